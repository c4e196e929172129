"""The CPU reconstruction path (the reference driver's --no-gpu mode:
ctbb_rebin_filter_cpu, ctbb_backproject_cpu, ctbb_finalize_image_stack_cpu)
against the reference's own CPU twins compiled from source (oracle/_ref,
where the reference tree is present) and the FP64 oracle.

Bars: backprojection bit-identical to the reference's backproject_cpu built
from source, finalize within 1e-6 of the maximum of the reference's
finalize_image_stack_cpu built from source (see its test); rebin + filter
within 2e-4 (max) and 1e-5 (RMS) of the filtered block's max |value| of the
FP64 oracle, the error of the FP32 interpolation (the HIP rebin's arithmetic)
after the ramp filter (the reference's rebin_filter_cpu needs FFTW3 and
cannot be built here); the whole CPU path
within 0.05 HU RMS of the FP64 oracle with CPU conventions.
"""
import ctypes as C

import numpy as np
import pytest

import ctbangbang_amd as cb
from ctbangbang_amd import scans


def _mr(tmp_path, tok, install, name="c.prm"):
    p = str(tmp_path / name)
    cb.write_prm(p, tok)
    return cb.metadata_from_prm(p, install_dir=install)


def _block_rebin(mr, raw):
    """ctbb_rebin_filter_cpu on a host pull; returns the filtered block."""
    s = cb.block_sizes(mr)
    reb = np.zeros(s.rebin, dtype=np.float32)
    r = np.ascontiguousarray(raw, dtype=np.float32)
    mr.ctd.raw = r.ctypes.data_as(C.POINTER(C.c_float))
    mr.ctd.rebin = reb.ctypes.data_as(C.POINTER(C.c_float))
    st = cb.lib().ctbb_rebin_filter_cpu(C.byref(mr), 4)
    mr.ctd.raw = None
    mr.ctd.rebin = None
    assert st == 0, st
    return reb.reshape(s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)


def _block_bp(mr, reb, threads=4):
    """ctbb_backproject_cpu of one block; returns [32][ny][nx]."""
    nblk = mr.ri.n_slices_block * mr.rp.nx * mr.rp.ny
    img = np.zeros(max(int(mr.ri.cb.block_idx), 1) * nblk, dtype=np.float32)
    reb = np.ascontiguousarray(reb, dtype=np.float32)
    mr.ctd.rebin = reb.ctypes.data_as(C.POINTER(C.c_float))
    mr.ctd.image = img.ctypes.data_as(C.POINTER(C.c_float))
    st = cb.lib().ctbb_backproject_cpu(C.byref(mr), threads)
    mr.ctd.rebin = None
    mr.ctd.image = None
    assert st == 0, st
    b = int(mr.ri.cb.block_idx) - 1
    return img[b * nblk:(b + 1) * nblk].reshape(mr.ri.n_slices_block, mr.rp.ny, mr.rp.nx)


@pytest.mark.parametrize("table_dir,ffs", [("out", "none"), ("in", "z")])
def test_backproject_cpu_equals_reference_build(tmp_path, mini_dir, table_dir, ffs):
    import oracle
    if oracle.ref_lib() is None:
        pytest.skip("reference sources not present")
    tok = scans.mini_tokens(ffs, nrows=8, nx=32, table_dir=table_dir)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(3)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    got = _block_bp(mr, reb)
    want = oracle.ref_backproject_cpu(mr, reb)
    print(f"backproject_cpu vs reference build ({table_dir}): max |diff| {np.abs(got - want).max():.3g}, "
          f"identical {np.array_equal(got, want)}")
    assert np.array_equal(got, want)
    # thread count does not change the bits
    assert np.array_equal(_block_bp(mr, reb, threads=1), got)
    cb.free_metadata(mr)


def test_finalize_cpu_equals_reference_build(tmp_path, mini_dir):
    """Within 1e-6 of the stack maximum (one or two float ulps): the
    reference source calls fabs on floats without including <math.h>, so
    its g++ build here takes the double overload (weights evaluated in
    double, rounded once), while an nvcc build gets CUDA's float overload;
    this path, like the GPU finalize, evaluates them in float.  For the same
    reason the slice range is not a whole number of thicknesses, where
    floor(|end - start| / thickness) would differ between the two."""
    import oracle
    if oracle.ref_lib() is None:
        pytest.skip("reference sources not present")
    for reverse in (False, True):
        tok = scans.mini_tokens("none", nrows=8, nx=32, n_slices=6)
        tok["EndPos"] = tok["StartPos"] + 5 * 2.4 + 0.5
        tok["NSlices"] = 0
        if reverse:
            tok["StartPos"], tok["EndPos"] = tok["EndPos"], tok["StartPos"]
        mr = _mr(tmp_path, tok, mini_dir, name=f"f{int(reverse)}.prm")
        s = cb.block_sizes(mr)
        rng = np.random.default_rng(7)
        img = (rng.standard_normal(s.image_stack) * 0.01 + 0.02).astype(np.float32)
        mr.ctd.image = img.ctypes.data_as(C.POINTER(C.c_float))
        assert cb.lib().ctbb_finalize_image_stack_cpu(C.byref(mr), 0.0, 3) == 0
        nf = cb.n_slices_final(mr)
        got = np.ctypeslib.as_array(mr.ctd.final_image_stack, shape=(nf * mr.rp.nx * mr.rp.ny,)).copy()
        mr.ctd.image = None
        want = oracle.ref_finalize_cpu(mr, img)
        d = np.abs(got.reshape(want.shape) - want).max()
        print(f"finalize_cpu vs reference build (reverse {reverse}): max |diff| {d:.3g} of {np.abs(want).max():.3g}")
        assert d <= 1e-6 * np.abs(want).max()
        cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "phi", "z", "both"])
def test_rebin_filter_cpu_vs_oracle(tmp_path, mini_dir, ffs):
    import oracle
    tok = scans.mini_tokens(ffs, nrows=8 if ffs in ("z", "both") else 4)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    raw = cb.synth_scan(mr, scans.phantom_for(tok, 3), cb.pull_start(mr), int(mr.ri.n_proj_pull))
    got = _block_rebin(mr, raw)
    want = oracle.filter_fft(oracle.rebin(mr, raw, double=True), cb.load_filter(mr))
    err = np.abs(got - want).max() / np.abs(want).max()
    rms = np.sqrt(np.mean((got - want) ** 2)) / np.abs(want).max()
    print(f"rebin_filter_cpu {ffs}: rel max err {err:.3g}, rms {rms:.3g}")
    assert err <= 2e-4 and rms <= 1e-5
    cb.free_metadata(mr)


def test_cpu_path_end_to_end_hu(tmp_path, mini_dir):
    """rebin_filter_cpu -> backproject_cpu per block -> finalize_cpu (HU)
    against the FP64 oracle with the CPU twin's conventions."""
    import oracle
    tok = scans.mini_tokens("both", nrows=8, nx=32)
    mr = _mr(tmp_path, tok, mini_dir)
    n_read = tok["Readings"]
    raw_all = cb.synth_scan(mr, scans.phantom_for(tok, 5), 0, n_read).reshape(n_read, -1)
    s = cb.block_sizes(mr)
    image = np.zeros(s.image_stack, dtype=np.float32)
    blocks_want = []
    taps = cb.load_filter(mr)
    nblk = mr.ri.n_slices_block * mr.rp.nx * mr.rp.ny
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
        win = np.zeros((n, raw_all.shape[1]), dtype=np.float32)
        lo, hi = max(s0, 0), min(s0 + n, n_read)
        win[lo - s0:hi - s0] = raw_all[lo:hi]
        reb = _block_rebin(mr, win)
        image[b * nblk:(b + 1) * nblk] = _block_bp(mr, reb).ravel()
        blocks_want.append(oracle.backproject(mr, oracle.rebin_filter(mr, win, taps, double=True), oracle.BP_CPU,
                                              double=True))
    mr.ctd.image = image.ctypes.data_as(C.POINTER(C.c_float))
    assert cb.lib().ctbb_finalize_image_stack_cpu(C.byref(mr), 0.02, 0) == 0
    nf = cb.n_slices_final(mr)
    got = np.ctypeslib.as_array(mr.ctd.final_image_stack, shape=(nf * mr.rp.nx * mr.rp.ny,)).copy()
    mr.ctd.image = None
    want = oracle.finalize(mr, np.concatenate(blocks_want), 0.02, double=True)
    d = got.reshape(want.shape).astype(np.float64) - want
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"cpu path end to end: HU rms {rms:.4f} max {mx:.4f}")
    assert rms <= 0.05 and mx <= 2.0
    cb.free_metadata(mr)
