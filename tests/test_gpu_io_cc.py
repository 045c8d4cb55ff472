"""The in-loop automatic connected component (rsfe_splitbregman_imp.hxx:447-491, called at
1808-1812) and the MetaImage dumps (imp:1939-1978; the format of interface<S>::write,
interface_imp.hxx:672-690, and of its reader, 33-47 + 125-150) on the device.

The reference holds no MetaImage file for this path, so the format is checked against the
published MetaImage layout read by a parser written here (key = value header lines up to
ElementDataFile; raw voxels x fastest) and against hand-written files of other element
types, byte orders and layouts; the loop against the CPU oracle in FP64."""
import os

import numpy as np
import pytest

from cases import two_balls

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def parse_mhd(path):
    hdr = {}
    with open(path) as f:
        for line in f:
            k, _, v = line.partition("=")
            hdr[k.strip()] = v.strip()
            if k.strip() == "ElementDataFile":
                break
    return hdr


# ---- MetaImage -------------------------------------------------------------------
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("shape", [(5, 6, 7), (33, 20, 65)])
def test_mhd_write_layout_and_round_trip(tk, ctx, rng, tmp_path, dtype, shape):
    say("mhd", shape, np.dtype(dtype).name)
    a = rng.standard_normal(shape).astype(dtype)
    h = (0.5, 1.25, 2.0)
    v = ctx.from_numpy(a, h=h)
    base = str(tmp_path / "vol")
    v.write_mhd(base)
    hdr = parse_mhd(base + ".mhd")
    assert hdr["ObjectType"] == "Image" and hdr["NDims"] == "3"
    assert hdr["ElementType"] == "MET_DOUBLE" and hdr["BinaryDataByteOrderMSB"] == "False"
    assert hdr["CompressedData"] == "False" and hdr["ElementDataFile"] == "vol.raw"
    assert [int(x) for x in hdr["DimSize"].split()] == list(shape)
    assert [float(x) for x in hdr["ElementSpacing"].split()] == list(h)
    raw = np.fromfile(base + ".raw", dtype="<f8")
    assert raw.size == a.size
    # x (i) fastest: Fortran order of the (i, j, k) array
    assert np.array_equal(raw.reshape(shape, order="F"), a.astype(np.float64))
    back = ctx.read_mhd(base + ".mhd", dtype)
    assert back.shape == shape and back.h_spacing == h
    assert np.array_equal(back.numpy(), a)


def write_mhd_by_hand(path, arr, etype, msb=False, header_size=0, local=False, channels=1,
                      ndims=3):
    """a MetaImage file of another layout, written independently of the library"""
    dims = arr.shape[:3] if ndims == 3 else arr.shape[:2]
    lines = ["ObjectType = Image", f"NDims = {ndims}",
             "DimSize = " + " ".join(str(d) for d in dims),
             "ElementSpacing = " + " ".join(["1.5"] * ndims),
             f"BinaryDataByteOrderMSB = {'True' if msb else 'False'}",
             f"ElementType = {etype}"]
    if channels > 1:
        lines.append(f"ElementNumberOfChannels = {channels}")
    if header_size:
        lines.append(f"HeaderSize = {header_size}")
    np_t = {"MET_SHORT": "i2", "MET_UCHAR": "u1", "MET_FLOAT": "f4", "MET_INT": "i4",
            "MET_USHORT": "u2", "MET_DOUBLE": "f8"}[etype]
    data = np.asfortranarray(arr).astype((">" if msb else "<") + np_t)
    payload = b"\0" * max(header_size, 0) + data.tobytes(order="F")
    if local:
        lines.append("ElementDataFile = LOCAL")
        with open(path, "wb") as f:
            f.write(("\n".join(lines) + "\n").encode() + payload)
    else:
        raw = os.path.splitext(path)[0] + ".dat"
        lines.append(f"ElementDataFile = {os.path.basename(raw)}")
        with open(path, "w") as f:
            f.write("\n".join(lines) + "\n")
        with open(raw, "wb") as f:
            f.write(payload)


def test_mhd_reader_variants(tk, ctx, rng, tmp_path):
    shape = (6, 5, 4)
    s16 = rng.integers(-3000, 3000, shape).astype(np.int16)
    p = str(tmp_path / "be.mhd")
    write_mhd_by_hand(p, s16, "MET_SHORT", msb=True, header_size=32)
    v = ctx.read_mhd(p, np.float64)
    assert v.h_spacing == (1.5, 1.5, 1.5)
    assert np.array_equal(v.numpy(), s16.astype(np.float64))
    u8 = rng.integers(0, 255, shape).astype(np.uint8)
    p = str(tmp_path / "local.mha.mhd")
    write_mhd_by_hand(p, u8, "MET_UCHAR", local=True)
    assert np.array_equal(ctx.read_mhd(p, np.float32).numpy(), u8.astype(np.float32))
    # two channels interleaved per voxel: component 0 (GetScalarComponentAsDouble(.., 0))
    two = rng.standard_normal(shape + (2,)).astype(np.float32)
    p = str(tmp_path / "ch2.mhd")
    dims = "DimSize = 6 5 4"
    payload = np.transpose(two, (2, 1, 0, 3)).astype("<f4").tobytes()
    with open(p, "w") as f:
        f.write(f"ObjectType = Image\nNDims = 3\n{dims}\nElementNumberOfChannels = 2\n"
                "ElementType = MET_FLOAT\nElementDataFile = ch2.raw\n")
    with open(str(tmp_path / "ch2.raw"), "wb") as f:
        f.write(payload)
    assert np.array_equal(ctx.read_mhd(p, np.float32).numpy(), two[..., 0])
    # a 2-D image (NDims 2) is one plane
    flat = rng.standard_normal((7, 9)).astype(np.float32)
    p = str(tmp_path / "flat.mhd")
    write_mhd_by_hand(p, flat, "MET_FLOAT", ndims=2)
    assert np.array_equal(ctx.read_mhd(p, np.float64).numpy()[:, :, 0], flat.astype(np.float64))
    # refusals: compressed data, missing files
    p = str(tmp_path / "z.mhd")
    with open(p, "w") as f:
        f.write("NDims = 3\nDimSize = 2 2 2\nElementType = MET_FLOAT\nCompressedData = True\n"
                "ElementDataFile = z.raw\n")
    with pytest.raises(tk.TkError) as e:
        ctx.read_mhd(p)
    assert e.value.code == tk.TK_ERR_UNSUPPORTED
    with pytest.raises(tk.TkError) as e:
        ctx.read_mhd(str(tmp_path / "missing.mhd"))
    assert e.value.code == tk.TK_ERR_IO


# ---- automatic connected component in the loop ------------------------------------
def run_gpu(tk, ctx, u0, phi0, n_iter, **kw):
    vu, vi = ctx.from_numpy(u0), ctx.from_numpy(phi0)
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(**kw))
    st = r.iterate(n_iter)
    phi = r.phi()
    r.close()
    return phi, st


@pytest.mark.parametrize("binary,init_vars", [(1, 1), (0, 1), (1, 0)])
def test_loop_auto_cc_f64_vs_oracle(tk, orc, ctx, binary, init_vars):
    shape = (20, 18, 16)
    u0, phi0, seed = two_balls(shape)
    pixel = tuple(x + 0.5 for x in seed)
    kw = dict(sigma=1.5, maxiter=5, tol=0.0, ls_steps=3, ls_tol=0.0)
    p = orc.default_params(auto_cc=1, auto_cc_freq=2, cc_pixel=pixel, cc_binary_output=binary,
                           cc_init_variables=init_vars, **kw)
    ref, sw, tl, st_or = orc.rsfe_apply(u0, phi0, p)
    phi, st = run_gpu(tk, ctx, u0, phi0, 5, gs_order=tk.TK_GS_LEX, auto_extract_conn_comp=1,
                      auto_extract_conn_comp_freq=2, cc_init_pixel=pixel,
                      cc_binary_output=binary, cc_init_variables=init_vars, **kw)
    err = np.max(np.abs(phi - ref))
    say(f"auto cc binary {binary} init {init_vars}: max|phi-ref| = {err:.3e}, extracted "
        f"{[s['cc_extracted'] for s in st]}")
    assert [s["cc_extracted"] for s in st] == [0, 1, 0, 1, 0]
    assert st_or.cc_extractions == 2
    assert err < 1e-9
    assert np.allclose([s["tollerance"] for s in st], tl, rtol=1e-9)
    # without the extraction the run keeps both balls
    phi_plain, _ = run_gpu(tk, ctx, u0, phi0, 5, gs_order=tk.TK_GS_LEX, **kw)
    assert np.max(np.abs(phi_plain - ref)) > 0.1


def test_loop_auto_cc_f32_and_seed_cases(tk, orc, ctx):
    """FP32: the extraction keeps the same component (mask Dice with the FP64 oracle); a
    seed outside the volume or on a black voxel leaves the run unchanged"""
    shape = (20, 18, 16)
    u0, phi0, seed = two_balls(shape)
    pixel = tuple(x + 0.5 for x in seed)
    kw = dict(sigma=1.5, maxiter=4, tol=0.0, ls_steps=3, ls_tol=0.0)
    ref = orc.rsfe_apply(u0, phi0, orc.default_params(auto_cc=1, auto_cc_freq=2, cc_pixel=pixel,
                                                      **kw))[0]
    vu, vi = ctx.from_numpy(u0.astype(np.float32)), ctx.from_numpy(phi0.astype(np.float32))
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(gs_order=tk.TK_GS_LEX, auto_extract_conn_comp=1,
                                               auto_extract_conn_comp_freq=2,
                                               cc_init_pixel=pixel, **kw))
    st = r.iterate(4)
    m = r.phi() >= 0.5
    r.close()
    a = ref >= 0.5
    dice = 2 * (a & m).sum() / (a.sum() + m.sum())
    say(f"auto cc FP32: Dice {dice:.6f}, extracted {[s['cc_extracted'] for s in st]}")
    assert dice > 0.99 and [s["cc_extracted"] for s in st] == [0, 1, 0, 1]
    plain, _ = run_gpu(tk, ctx, u0, phi0, 3, gs_order=tk.TK_GS_LEX, **kw)
    for px in ((-1.0, 3.0, 3.0), (3.0, 99.0, 3.0), (0.2, 0.2, 0.2)):
        phi, st = run_gpu(tk, ctx, u0, phi0, 3, gs_order=tk.TK_GS_LEX,
                          auto_extract_conn_comp=1, auto_extract_conn_comp_freq=1,
                          cc_init_pixel=px, **kw)
        assert all(s["cc_extracted"] == 0 for s in st)
        assert np.array_equal(phi, plain)


def test_dumps_every_dumpfrequency_and_final(tk, ctx, tmp_path):
    shape = (20, 18, 16)
    u0, phi0, _ = two_balls(shape)
    out = str(tmp_path / "result")
    vu, vi = ctx.from_numpy(u0), ctx.from_numpy(phi0)
    r = tk.Rsfe(ctx, vu, vi, tk.default_params(sigma=1.5, maxiter=5, tol=0.0, ls_steps=3,
                                               dumpfrequency=2, outputname=out))
    snaps = {}
    for t in range(1, 6):
        r.iterate(1)
        snaps[t] = r.phi()
    r.iterate(1)  # the run has ended: no second final dump, no error
    r.close()
    names = sorted(os.listdir(tmp_path))
    say("dumps:", names)
    assert names == ["result_2.mhd", "result_2.raw", "result_4.mhd", "result_4.raw",
                     "result_5_final.mhd", "result_5_final.raw"]
    for t, name in ((2, "result_2"), (4, "result_4"), (5, "result_5_final")):
        v = ctx.read_mhd(str(tmp_path / (name + ".mhd")))
        assert np.array_equal(v.numpy(), snaps[t])


def test_batch_refuses_per_run_options(tk, ctx):
    for kw in (dict(auto_extract_conn_comp=1), dict(dumpfrequency=1)):
        with pytest.raises(tk.TkError):
            tk.Batch(ctx, [(16, 16, 16)], tk.default_params(**kw))
