"""How far the reference arithmetic itself carries an input change of FP32 rounding size on
the config-3 volumes whose FP32 GPU masks differ most from the FP64 oracle's.

For each listed volume the FP64 oracle runs the config-3 loop (sigma 3, 30 outer iterations,
ls_steps 5, ls_tol 0) twice more: on the image rescaled to [0, 255] and then rounded to FP32
(the input the FP32 mode works from), and on the image with every 7th voxel raised by 1e-7
of its value.  The Dice of each run's mask against c3_batch.npz (the unperturbed oracle) is
written to c3_sensitivity.json.  Output is test data; the oracle is test infrastructure.

Usage: python tests/golden/c3_sensitivity.py 14 4 41 60
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import make_golden as mg  # noqa: E402

orc, tk = mg.orc, mg.tk


def main(vols):
    d = np.load(os.path.join(HERE, "c3_batch.npz"))
    dims, nell = d["dims"], d["nell"]
    p = orc.default_params(sigma=3.0, maxiter=30, tol=0.0, ls_steps=5, ls_tol=0.0)
    path = os.path.join(HERE, "c3_sensitivity.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    for v in vols:
        u0, phi0 = mg.c3_volume(v, dims[v], nell[v])
        ref = np.unpackbits(d[f"mask_{v}"])[:u0.size].astype(bool)
        u255 = orc.change_range(u0, 255.0, 0.0)
        rel = u0.astype(np.float64).copy()
        rel.reshape(-1)[::7] *= 1 + 1e-7
        row = {"dims": dims[v].tolist(), "n_ell": int(nell[v])}
        for name, img in (("input_rounded_to_fp32", u255.astype(np.float32).astype(np.float64)),
                          ("every_7th_voxel_times_1_plus_1e-7", orc.change_range(rel, 255.0, 0.0))):
            t0 = time.time()
            phi, _, _, _ = orc.rsfe_apply(img, phi0, p)
            row[name] = mg.dice((phi >= 0.5).reshape(-1), ref)
            print(f"volume {v} {name}: Dice vs unperturbed oracle {row[name]:.6f} "
                  f"({time.time() - t0:.0f} s)", flush=True)
        out[str(v)] = row
        json.dump(out, open(path, "w"), indent=1, sort_keys=True)


if __name__ == "__main__":
    main([int(x) for x in sys.argv[1:]])
