"""Pins of the CPU oracle (CPU only).

Every oracle function is checked against something it does not restate:
the reference's own CPU code compiled from its sources (oracle/_ref), a
closed form, or an invariant a plausible mistake would break.
"""
import os

import numpy as np
import pytest

import ctbangbang_amd as cb
import oracle
from ctbangbang_amd import scans

PI_BP = np.float32(3.1415368979)


def _mr(tmp_path, tokens, install=None, name="o.prm"):
    p = str(tmp_path / name)
    cb.write_prm(p, tokens)
    return cb.metadata_from_prm(p, install_dir=install)


def need_ref():
    if oracle.ref_lib() is None:
        pytest.skip("reference sources not present (oracle/_ref not built)")


# ----------------------------------------------------------------------------
# backprojection


@pytest.mark.parametrize("table_dir", ["in", "out"])
def test_backproject_cpu_conventions_match_reference_build(tmp_path, mini_dir, table_dir):
    """FP32 restatement in the CPU twin's conventions == reference
    backproject_cpu (src/backproject_cpu.cu:49) run from its own source."""
    need_ref()
    mr = _mr(tmp_path, scans.mini_tokens("none", nrows=8, nx=32, table_dir=table_dir), mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(3)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    ref = oracle.ref_backproject_cpu(mr, reb)
    mine = oracle.backproject(mr, reb, oracle.BP_CPU, double=False)
    assert np.abs(mine - ref).max() <= 1e-5 * np.abs(ref).max()
    mine64 = oracle.backproject(mr, reb, oracle.BP_CPU, double=True)
    assert np.abs(mine64 - ref).max() <= 5e-4 * np.abs(ref).max()   # the reference's own FP32 error
    cb.free_metadata(mr)


def test_backproject_gpu_vs_cpu_conventions(tmp_path, mini_dir):
    """With TableDir in the GPU conventions differ from the CPU twin only by the
    output transpose, the FOV mask (inactive here) and the top-row border of
    the last half turn (reference backproject.cuh:60-63,84 vs
    backproject_cpu.cu:94,112; interp.cu:45-46)."""
    mr = _mr(tmp_path, scans.mini_tokens("none", nrows=8, nx=32, table_dir="in"), mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(4)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    # zero the last half turn so the border rules cannot differ
    half = mr.cg.n_proj_turn // 2
    reb[(s.n_half_turns - 1) * half:] = 0
    g = oracle.backproject(mr, reb, oracle.BP_GPU)
    c = oracle.backproject(mr, reb, oracle.BP_CPU)
    assert np.abs(g - c.transpose(0, 2, 1)).max() <= 1e-6 * np.abs(g).max()
    cb.free_metadata(mr)


def test_backproject_constant_sinogram_closed_form(tmp_path, mini_dir):
    """Constant filtered sinogram c: every angle contributes s_t/h_t = c, so
    interior voxels get 576 c 2pi'/1152 = pi' c; voxels outside AcqFOV/2 are 0."""
    tok = scans.mini_tokens("none", nrows=8, nx=32)
    tok["AcqFOV"] = 300.0     # FOV circle of radius 150 cuts the 250 mm square's corners
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    reb = np.full((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled), 1.5, dtype=np.float32)
    img = oracle.backproject(mr, reb, oracle.BP_GPU)
    nx = mr.rp.nx
    d = 250.0 / nx
    xi = (np.arange(nx) - (nx - 1) / 2) * d
    r = np.sqrt(xi[:, None] ** 2 + xi[None, :] ** 2)
    inside = r < 150.0
    assert np.allclose(img[:, inside], PI_BP * 1.5, rtol=1e-6)
    assert np.all(img[:, ~inside] == 0)
    cb.free_metadata(mr)


def test_backproject_linearity(tmp_path, mini_dir):
    mr = _mr(tmp_path, scans.mini_tokens("none", nrows=8, nx=32), mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(9)
    shp = (s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)
    a, b = rng.standard_normal(shp).astype(np.float32), rng.standard_normal(shp).astype(np.float32)
    cols = (np.array([3, 17, 30, 8]), np.array([5, 16, 2, 29]))
    fa = oracle.backproject(mr, a, cols=cols)
    fb = oracle.backproject(mr, b, cols=cols)
    fab = oracle.backproject(mr, (2 * a - 3 * b).astype(np.float32), cols=cols)
    assert np.allclose(fab, 2 * fa - 3 * fb, atol=1e-5 * np.abs(fab).max())
    assert np.all(oracle.backproject(mr, np.zeros(shp, np.float32), cols=cols) == 0)
    cb.free_metadata(mr)


# ----------------------------------------------------------------------------
# finalize


@pytest.mark.parametrize("table_dir,reverse", [("out", False), ("in", False), ("out", True)])
def test_finalize_matches_reference_build(tmp_path, mini_dir, table_dir, reverse):
    need_ref()
    tok = scans.mini_tokens("none", table_dir=table_dir)
    tok["NSlices"] = 5
    if reverse:      # EndPos before StartPos: recon direction -1
        tok["StartPos"], tok["EndPos"] = tok["StartPos"] + 4 * 2.4, tok["StartPos"]
        tok["NSlices"] = 0
    else:
        tok["EndPos"] = tok["StartPos"] + 4 * 2.4
    mr = _mr(tmp_path, tok, mini_dir)
    s = cb.block_sizes(mr)
    img = (np.random.default_rng(2).standard_normal(s.image_stack) * 0.01 + 0.02).astype(np.float32)
    ref = oracle.ref_finalize_cpu(mr, img)
    mine32 = oracle.finalize(mr, img, double=False)
    mine64 = oracle.finalize(mr, img, double=True)
    assert mine32.shape == ref.shape
    assert np.abs(mine32 - ref).max() <= 2e-7 * np.abs(ref).max()
    assert np.abs(mine64 - ref).max() <= 2e-7 * np.abs(ref).max()
    cb.free_metadata(mr)


def test_finalize_closed_forms(tmp_path, mini_dir):
    tok = scans.mini_tokens("none")
    tok["NSlices"] = 4
    tok["EndPos"] = tok["StartPos"] + 3 * 2.4
    mr = _mr(tmp_path, tok, mini_dir)
    s = cb.block_sizes(mr)
    nxy = mr.rp.nx * mr.rp.ny
    const = np.full(s.image_stack, 0.02, np.float32)
    out = oracle.finalize(mr, const, double=True)
    assert np.allclose(out, 0.02, rtol=1e-6)
    # a stack linear in z (raw slice i at recon_start + i*coll) is reproduced at
    # final slices whose triangle window lies inside the stack
    i = np.arange(mr.ri.n_slices_recon, dtype=np.float64)
    z = mr.ri.recon_start_pos + i * mr.rp.coll_slicewidth
    lin = np.repeat((0.001 * z).astype(np.float32), nxy)
    out = oracle.finalize(mr, lin, double=True)
    zf = mr.rp.start_pos + np.arange(out.shape[0]) * mr.rp.slice_thickness
    assert np.allclose(out[:, 0, 0], 0.001 * zf, rtol=1e-5)
    # HU of water is 0, of air -1000
    hu = oracle.finalize(mr, const, mu_water=0.02, double=True)
    assert np.allclose(hu, 0.0, atol=1e-3)
    hu = oracle.finalize(mr, np.zeros_like(const), mu_water=0.02, double=True)
    assert np.allclose(hu, -1000.0)
    cb.free_metadata(mr)


# ----------------------------------------------------------------------------
# ramp filter


def test_filter_direct_impulse_and_numpy():
    n = 96
    taps = cb.ramp_filter(n, 0.7, 0.65, 0.5)
    rows = np.zeros((3, n), np.float32)
    rows[0, 0] = 1
    rows[1, 50] = 1
    rows[2] = np.random.default_rng(1).standard_normal(n)
    out = oracle.filter_direct(rows, taps)
    assert np.allclose(out[0], taps[n:2 * n], atol=1e-7)
    assert np.allclose(out[1], taps[n - 50:2 * n - 50], atol=1e-7)
    full = np.convolve(rows[2].astype(np.float64), taps.astype(np.float64))   # full linear convolution
    assert np.allclose(out[2], full[n:2 * n], atol=1e-6)
    assert np.allclose(oracle.filter_fft(rows, taps), out, atol=1e-6)


# ----------------------------------------------------------------------------
# rebinning: analytic off-centre cylinder through the real FFS geometry


def _cylinder_raw(mr, x0, y0, R, mu, ffs):
    """In-plane chord length of rays from each focal spot to each detector
    channel through a vertical cylinder, for the pulled readings; rows are
    scaled by (1 + 0.1 j + 0.05 z-set) to expose the row interleave."""
    cg = mr.cg
    n_ffs = mr.ri.n_ffs
    da = cg.src_to_det * cg.r_f * cg.fan_angle_increment / (4.0 * (cg.src_to_det - cg.r_f))
    dr = cg.src_to_det * mr.rp.coll_slicewidth / (4.0 * (cg.src_to_det - cg.r_f) * np.tan(cg.anode_angle))
    table = {"none": [(0, 0, 0)], "phi": [(da, 0, 0), (-da, 0, 0)], "z": [(0, -dr, 0), (0, dr, 1)],
             "both": [(da, -dr, 0), (-da, -dr, 0), (da, dr, 1), (-da, dr, 1)]}[ffs]
    n0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
    idx = np.arange(n0, n0 + n)
    alpha = np.deg2rad(mr.rp.tube_start_angle + 360.0 * idx / cg.n_proj_ffs)
    ch = np.arange(cg.n_channels)
    b0 = (ch - cg.central_channel) * cg.fan_angle_increment
    raw = np.zeros((n, cg.n_rows_raw, cg.n_channels), np.float32)
    for fs in range(n_ffs):
        sel = (idx % n_ffs) == fs
        a = alpha[sel][:, None]
        fda, fdr, zset = table[fs]
        ca, sa = np.cos(a), np.sin(a)
        Sx, Sy = cg.r_f * ca, cg.r_f * sa
        a1x, a1y, a2x, a2y = -ca, -sa, sa, -ca
        Fx, Fy = Sx - fdr * a1x - fda * a2x, Sy - fdr * a1y - fda * a2y
        Dx = Sx + cg.src_to_det * (np.cos(b0) * a1x + np.sin(b0) * a2x)
        Dy = Sy + cg.src_to_det * (np.cos(b0) * a1y + np.sin(b0) * a2y)
        ux, uy = Dx - Fx, Dy - Fy
        L = np.hypot(ux, uy)
        ux, uy = ux / L, uy / L
        # distance from the cylinder axis to the line
        d = np.abs((x0 - Fx) * uy - (y0 - Fy) * ux)
        chord = 2 * mu * np.sqrt(np.clip(R * R - d * d, 0, None))
        for j in range(cg.n_rows_raw):
            raw[sel, j, :] = chord * (1 + 0.1 * j + 0.05 * zset)
    return raw


@pytest.mark.parametrize("ffs", ["none", "phi", "z", "both"])
def test_rebin_analytic_cylinder(tmp_path, ffs):
    tok = scans.c1_tokens()
    n_ffs = (2 if ffs in ("phi", "both") else 1) * (2 if ffs in ("z", "both") else 1)
    tok.update({"Phiffs": int(ffs in ("phi", "both")), "Zffs": int(ffs in ("z", "both")),
                "Nrows": 2 if ffs in ("z", "both") else 1, "Readings": 4 * 1152 * n_ffs, "StartPos": 30.0,
                "EndPos": 31.0, "NSlices": 1, "PitchValue": 19.2, "SliceThickness": 1.0})
    mr = _mr(tmp_path, tok)
    cb.plan_block(mr, 0)
    x0, y0, R, mu = 40.0, -25.0, 60.0, 0.02
    raw = _cylinder_raw(mr, x0, y0, R, mu, ffs)
    out = oracle.rebin(mr, raw, double=True)
    cg = mr.cg
    v = np.arange(out.shape[0])
    theta = np.deg2rad(mr.rp.tube_start_angle + 360.0 * cb.pull_start(mr) / cg.n_proj_ffs) + \
        2 * np.pi * (v + cg.add_projections) / cg.n_proj_turn
    c = np.arange(cg.n_channels_oversampled)
    s = cg.r_f * (c - 2 * cg.central_channel) * cg.fan_angle_increment / 2
    s0 = x0 * np.sin(theta) - y0 * np.cos(theta)
    dd = s[None, :] - s0[:, None]
    want = 2 * mu * np.sqrt(np.clip(R * R - dd ** 2, 0, None))
    ok = np.abs(dd) < 0.9 * R
    for r in range(cg.n_rows):
        zset = (r % 2) if ffs in ("z", "both") else 0
        j = r // 2 if ffs in ("z", "both") else r
        scale = 1 + 0.1 * j + 0.05 * zset
        err = np.abs(out[:, r, :] / scale - want)[ok].max()
        assert err <= 2e-3 * want.max(), (ffs, r, err)
    cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "phi", "z", "both"])
def test_rebin_constant_and_precision(tmp_path, mini_dir, ffs):
    mr = _mr(tmp_path, scans.mini_tokens(ffs, nrows=8 if ffs in ("z", "both") else 4), mini_dir)
    cb.plan_block(mr, 0)
    n = mr.ri.n_proj_pull * mr.cg.n_rows_raw * mr.cg.n_channels
    out = oracle.rebin(mr, np.full(n, 3.25, np.float32))
    assert np.allclose(out, 3.25, rtol=1e-6)
    pp = scans.phantom_for(scans.mini_tokens(ffs), 2)
    raw = cb.synth_scan(mr, pp, cb.pull_start(mr), mr.ri.n_proj_pull)
    a = oracle.rebin(mr, raw, double=True)
    b = oracle.rebin(mr, raw, double=False)
    assert np.abs(a - b).max() <= 1e-4 * np.abs(raw).max()
    cb.free_metadata(mr)


# ----------------------------------------------------------------------------
# adaptive filter (GPU-only in the reference: invariants only)


def test_adaptive_smoothing_kernel_quirk(tmp_path, mini_dir):
    """The reference's 5x5x5 weights (3-|i-3|)+(3-|j-3|)+(3-|k-3|)+1 peak at
    offset (+1,+1,+1) (preprocessing.cuh:64-67): a lone sample H above its
    threshold among zeros becomes 7 H / 725 in the interior."""
    mr = _mr(tmp_path, scans.mini_tokens("none", nrows=8), mini_dir)
    cb.plan_block(mr, 0)
    cg = mr.cg
    np_ = mr.ri.n_proj_pull
    raw = np.zeros((np_, cg.n_rows_raw, cg.n_channels), np.float32)
    raw[40, 3, 60] = 10.0
    thr = np.full(np_, 5.0, np.float32)
    out = oracle.adaptive_apply(mr, raw, thr)
    assert out[40, 3, 60] == np.float32(np.float32(1.0) / np.float32(725.0)) * np.float32(70.0)
    # everything at or below the threshold is copied
    mask = np.ones_like(raw, bool)
    mask[40, 3, 60] = False
    assert np.array_equal(out[mask], raw[mask])
    # a constant block above threshold stays constant (normalised weights)
    const = np.full_like(raw, 2.0)
    out = oracle.adaptive_apply(mr, const, np.full(np_, 1.0, np.float32))
    assert np.allclose(out, 2.0, rtol=1e-6)


@pytest.mark.parametrize("eccentric", [False, True])
def test_adaptive_threshold_property(tmp_path, mini_dir, eccentric):
    """At the returned threshold T, the fraction of smoothed view maxima above T
    in the 180-degree window is within 1% of 0.03 s e (reference
    preprocessing.cuh:167-206) for views where the bisection converged; T stays
    inside the window's [min, max]; samples <= T are copied."""
    tok = scans.mini_tokens("both", nrows=8)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    raw = cb.synth_scan(mr, scans.phantom_for(tok, 3), cb.pull_start(mr), mr.ri.n_proj_pull)
    if eccentric:
        # per-view maxima oscillating 1:5 with the tube angle -> eccentricity 1
        a = 4 * np.pi * (cb.pull_start(mr) + np.arange(mr.ri.n_proj_pull)) / mr.cg.n_proj_ffs
        rng = np.random.default_rng(0)
        raw = (raw * (3 + 2 * np.cos(a))[:, None, None] *
               (1 + 0.05 * rng.standard_normal(raw.shape[0]))[:, None, None]).astype(np.float32)
    out, T = oracle.adaptive_filter(mr, raw)
    np_ = mr.ri.n_proj_pull
    npf = mr.cg.n_proj_ffs
    sup = raw.reshape(np_, -1).max(axis=1).clip(min=0)
    n_avg = int(np.ceil(np.float32(18 / 360) * npf))
    n_avg += 1 - n_avg % 2
    idx = np.clip(np.arange(np_)[:, None] + np.arange(n_avg)[None, :] - n_avg // 2, 0, np_ - 1)
    sm = sup[idx].mean(axis=1)
    n_rng = int(np.ceil(np.float32(0.5) * npf))
    n_rng += 1 - n_rng % 2
    ok = 0
    for p in range(0, np_, 7):
        lo, hi = max(p - n_rng // 2, 0), min(p + n_rng // 2, np_ - 1)
        win = sm[lo:hi + 1]
        widx = np.clip(np.arange(p - n_rng // 2, p + n_rng // 2 + 1), 0, np_ - 1)
        wmin, wmax = sm[widx].min(), sm[widx].max()
        assert wmin - 1e-6 <= T[p] <= wmax + 1e-6
        e = np.clip((1 - wmin / wmax - 0.3) / 0.2, 0, 1)
        lhs = 0.03 * tok["AdaptiveFiltration"] * e
        rhs = np.count_nonzero(win > T[p]) / (np.pi * npf)
        if lhs > 0 and abs(lhs - rhs) / lhs <= 0.0101:
            ok += 1
        if lhs == 0:
            # e = 0: the bisection drives T to the top of the window
            assert T[p] >= wmax - (wmax - wmin) * 2.0 ** -19
    n_checked = len(range(0, np_, 7))
    if eccentric:
        assert ok >= 0.8 * n_checked, (ok, n_checked)
    # samples at or below their view's threshold are untouched
    below = raw <= T[:, None, None]
    assert np.array_equal(out[below], raw[below])


@pytest.mark.parametrize("ffs", ["none", "both"])
def test_nonfinite_raw_sample_reaches_rebinned_sinogram(mini_dir, tmp_path, ffs):
    """A NaN or infinite raw sample inside the pull shows up as non-finite
    values of the rebinned sinogram (every raw sample the reconstruction
    uses is an interpolation tap of some rebinned sample), which is where the
    HIP rebin checks its output; a finite pull gives a finite sinogram."""
    tok = scans.mini_tokens(ffs, nrows=8 if ffs == "both" else 4)
    p = str(tmp_path / "n.prm")
    cb.write_prm(p, tok)
    mr = cb.metadata_from_prm(p, install_dir=mini_dir)
    cb.plan_block(mr, 0)
    n = int(mr.ri.n_proj_pull)
    raw = np.ones((n, mr.cg.n_rows_raw, mr.cg.n_channels), dtype=np.float32)
    assert np.isfinite(oracle.rebin(mr, raw, double=False)).all()
    for bad in (np.nan, np.inf):
        r = raw.copy()
        r[n // 2, 1, 40] = bad
        out = oracle.rebin(mr, r, double=False)
        assert not np.isfinite(out).all()
        # only views near the bad reading are affected
        views = np.unique(np.nonzero(~np.isfinite(out))[0])
        assert views.size < 8 * int(mr.ri.n_ffs)
    cb.free_metadata(mr)
