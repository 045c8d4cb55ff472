"""Write the golden fixtures of the RSFE Split-Bregman loop from the CPU FP64 oracle.

The reference itself cannot be built in this image (VTK, FFTW3 and GetPot are absent),
so the fixtures come from oracle/rsfe_oracle.cpp, whose pieces are pinned in
tests/test_oracle_pins.py.  Inputs come from the project's seeded generator.

  c0_64.npz   -- 64^3, two ellipsoids (image rescaled to [0,255] as the CLI does), sigma 3, 30 outer iterations, Gauss-Seidel
                 gated at a relative update of 1e-8 (ls_steps 5), Dirichlet 0,
                 cyclic convolution, reference default lambda/mu/beta.
  s5_64.npz   -- the same input with the reference default Poisson setting
                 (ls_steps 5, ls_tol 0), 30 iterations.

Usage: python tests/golden/make_golden.py [c0|s5|all]
"""
import importlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

import oracle_lib as orc  # noqa: E402

tk = importlib.import_module("3dimagetoolkit_amd")

C0_SEED = 20240601


def run(name, ls_tol, maxiter=30):
    u0, phi0 = tk.gen_standard((64, 64, 64), C0_SEED, 1, 2)
    # the reference's driver rescales the image to [0, 255] before apply
    # (examples/rsfe_splitbregman.cxx:86, image3d::change_range_of_intensity)
    u0 = orc.change_range(u0, 255.0, 0.0)
    p = orc.default_params(sigma=3.0, maxiter=maxiter, tol=0.0, ls_steps=5, ls_tol=ls_tol)
    t0 = time.time()
    phi, sweeps, tols, st = orc.rsfe_apply(u0, phi0, p)
    dt = time.time() - t0
    meta = dict(seed=C0_SEED, shape=[64, 64, 64], sigma=3.0, maxiter=maxiter, ls_steps=5,
                ls_tol=ls_tol, iters=int(st.iters), seconds=dt, threads=orc.lib().or_num_threads())
    np.savez_compressed(os.path.join(HERE, f"{name}_64.npz"), phi=phi, sweeps=sweeps, tols=tols,
                        meta=json.dumps(meta))
    print(name, meta, "sweeps", sweeps.tolist())


if __name__ == "__main__":
    which = sys.argv[1] if len(sys.argv) > 1 else "all"
    if which in ("s5", "all"):
        run("s5", 0.0)
    if which in ("c0", "all"):
        run("c0", 1e-8)
