"""HIP path vs the CPU oracle, stage by stage and end to end (needs a GPU).

Tolerances (written here, measured on MI355X):
  adaptive filter, finalize     bit-exact vs the FP32 restatement
  rebin                         |HIP - FP64 oracle| <= 2e-5 * max|raw|
  ramp filter F32               <= 2e-6 * max|FP64 result|
  ramp filter BF16X3 / F16X3    <= 3e-5 / 2e-6 * max|FP64 result|
  backprojection                <= 2e-4 * max|FP64 result| (random sinogram)
  end to end (mini scans)       RMS <= 1 HU, max <= 10 HU vs the FP64 oracle
"""
import os

import numpy as np
import pytest

import ctbangbang_amd as cb
from ctbangbang_amd import scans

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def _mr(tmp_path, tokens, install, name="s.prm"):
    p = str(tmp_path / name)
    cb.write_prm(p, tokens)
    return cb.metadata_from_prm(p, install_dir=install)


def _raw_pull(ctx, mr, pp):
    s = cb.block_sizes(mr)
    d = torch.empty(s.raw_pull, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, pp, cb.pull_start(mr), mr.ri.n_proj_pull, d)
    ctx.synchronize()
    return d


def test_synth_host_matches_device(ctx, tmp_path, mini_dir):
    mr = _mr(tmp_path, scans.mini_tokens("both"), mini_dir)
    pp = scans.phantom_for(scans.mini_tokens("both"), 9)
    d = torch.empty(40 * mr.cg.n_rows_raw * mr.cg.n_channels, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, pp, 100, 40, d)
    h = cb.synth_scan(mr, pp, 100, 40).ravel()
    g = host(d)
    assert np.abs(g - h).max() <= 1e-5 * max(1.0, np.abs(h).max())
    cb.free_metadata(mr)


@pytest.mark.parametrize("mode,tol", [(cb.FILTER_F32, 2e-6), (cb.FILTER_BF16X3, 3e-5), (cb.FILTER_F16X3, 2e-6),
                                      (cb.FILTER_FFT32, 2e-6)])
@pytest.mark.parametrize("n", [1472, 256, 1840])
def test_ramp_filter(ctx, mode, tol, n):
    import oracle
    rng = np.random.default_rng(n + 10 * mode)
    rows = 333
    x = rng.standard_normal((rows, n)).astype(np.float32) * 3.0 + 1.0
    taps = cb.ramp_filter(n, 0.35, 1.0, 1.0) if n != 256 else cb.ramp_filter(n, 2.0, 0.65, 0.5)
    want = oracle.filter_fft(x, taps)
    ctx.set_options(filter_mode=mode, proj_dtype=cb.PROJ_F32)
    out = torch.empty(rows * n, dtype=torch.float32, device="cuda")
    taps_d, x_d = dev(taps), dev(x)
    ctx.filter_rows_dev(taps_d, n, rows, x_d, out)
    got = host(out).reshape(rows, n)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"filter n={n} mode={mode}: rel max err {err:.3g}")
    t = tol if n % 64 == 0 else 2e-6   # non-multiple-of-64 widths use the FP32 VALU kernel
    assert err <= t
    ctx.set_options(filter_mode=cb.FILTER_F32)


def test_ramp_filter_fft_in_place(ctx):
    """FFT32 form filtering a buffer in place (each workgroup reads its rows
    before writing them): same result as out of place."""
    n, rows = 1472, 101
    rng = np.random.default_rng(5)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    taps = cb.ramp_filter(n, 0.35, 1.0, 1.0)
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    a = dev(x)
    out = torch.empty(rows * n, dtype=torch.float32, device="cuda")
    taps_d = dev(taps)
    ctx.filter_rows_dev(taps_d, n, rows, a, out)
    ctx.filter_rows_dev(taps_d, n, rows, a, a)
    assert np.array_equal(host(a).ravel(), host(out))
    # any other overlap is refused before a kernel runs: bfloat16 output in
    # place (its rows are half as wide), an offset overlap, and the tiled
    # forms in place
    # (one row of slack, so even the offset destination stays inside the buffer)
    buf = dev(np.concatenate([x.ravel(), np.zeros(n, np.float32)]))
    b = buf[:rows * n]
    for mode, proj, dst in ((cb.FILTER_FFT32, cb.PROJ_BF16, b), (cb.FILTER_FFT32, cb.PROJ_F32, buf[n:]),
                            (cb.FILTER_F32, cb.PROJ_F32, b), (cb.FILTER_BF16X3, cb.PROJ_F32, b)):
        ctx.set_options(filter_mode=mode, proj_dtype=proj)
        with pytest.raises(cb.CTBBError) as e:
            ctx.filter_rows_dev(taps_d, n, rows, b, dst)
        assert e.value.code == 1                                # CTBB_ERR_ARG
    ctx.synchronize()
    assert np.array_equal(host(b).reshape(rows, n), x)        # untouched
    ctx.set_options(filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32)


@pytest.mark.parametrize("mode", [cb.FILTER_F32, cb.FILTER_BF16X3, cb.FILTER_FFT32])
def test_ramp_filter_bf16_output(ctx, mode):
    """bfloat16 storage of the filtered rows: round-to-nearest-even of the
    FP32-accumulated result (<= 2^-8 relative per sample, plus the filter's
    own absolute error bound of 3e-5 * max)."""
    import oracle
    n, rows = 1472, 200
    rng = np.random.default_rng(77)
    x = rng.standard_normal((rows, n)).astype(np.float32)
    taps = cb.ramp_filter(n, 0.35, 1.0, 1.0)
    want = oracle.filter_fft(x, taps)
    ctx.set_options(filter_mode=mode, proj_dtype=cb.PROJ_BF16)
    out = torch.empty(rows * n, dtype=torch.int16, device="cuda")
    taps_d, x_d = dev(taps), dev(x)
    ctx.filter_rows_dev(taps_d, n, rows, x_d, out)
    ctx.set_options(filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32)
    bits = host(out).astype(np.uint16).astype(np.uint32) << 16
    got = bits.view(np.float32).reshape(rows, n)
    excess = np.abs(got - want) - 2.0 ** -8 * np.abs(want)
    print(f"filter bf16 out mode={mode}: max excess over 2^-8 rel {excess.max() / np.abs(want).max():.3g} of max")
    assert excess.max() <= 3e-5 * np.abs(want).max()


def test_filter_impulse_gives_taps(ctx):
    n = 1472
    taps = np.fromfile(os.path.join(cb.PKG_DIR, "resources", "filters", "f_736_sharp.txt"), dtype=np.float32)
    x = np.zeros((64, n), dtype=np.float32)
    for r in range(64):
        x[r, (r * 23) % n] = 1.0
    out = torch.empty(64 * n, dtype=torch.float32, device="cuda")
    ctx.set_options(filter_mode=cb.FILTER_F32)
    taps_d, x_d = dev(taps), dev(x)
    ctx.filter_rows_dev(taps_d, n, 64, x_d, out)
    got = host(out).reshape(64, n)
    for r in range(64):
        k = (r * 23) % n
        want = taps[n - k:2 * n - k]     # out[l] = taps[n + l - k]
        assert np.allclose(got[r], want, rtol=0, atol=1e-7 * np.abs(taps).max())


@pytest.mark.parametrize("ffs", ["none", "phi", "z", "both"])
def test_rebin(ctx, tmp_path, mini_dir, ffs):
    import oracle
    tok = scans.mini_tokens(ffs, nrows=8 if ffs in ("z", "both") else 4)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    raw = _raw_pull(ctx, mr, scans.phantom_for(tok, 3))
    s = cb.block_sizes(mr)
    out = torch.empty(s.rebin, dtype=torch.float32, device="cuda")
    ctx.rebin_dev(mr, raw, out)
    got = host(out).reshape(s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)
    r = host(raw)
    want = oracle.rebin(mr, r, double=True)
    err = np.abs(got - want).max() / np.abs(r).max()
    print(f"rebin {ffs}: rel max err {err:.3g}")
    assert err <= 2e-5
    cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "both"])
def test_adaptive_filter_bit_exact(ctx, tmp_path, mini_dir, ffs):
    import oracle
    tok = scans.mini_tokens(ffs, nrows=8 if ffs == "both" else 4)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    raw = _raw_pull(ctx, mr, scans.phantom_for(tok, 4))
    out = torch.empty_like(raw)
    ctx.adaptive_filter_dev(mr, raw, out)
    got = host(out)
    want, thr = oracle.adaptive_filter(mr, host(raw))
    changed = np.count_nonzero(want != host(raw))
    print(f"adaptive {ffs}: {changed} samples filtered, mismatches {np.count_nonzero(got != want)}")
    assert changed > 0
    assert np.array_equal(got, want)
    cb.free_metadata(mr)


@pytest.mark.parametrize("nz", [32, 16])
@pytest.mark.parametrize("table_dir", ["out", "in"])
def test_backproject_random_sinogram(ctx, tmp_path, mini_dir, table_dir, nz):
    import oracle
    tok = scans.mini_tokens("none", nrows=8, nx=32, table_dir=table_dir)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(5)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    ctx.set_options(proj_dtype=cb.PROJ_F32, bp_tile=nz)
    reb_d = dev(reb)
    ctx.backproject_dev(mr, reb_d, img)
    got = host(img).reshape(32, mr.rp.nx, mr.rp.ny)
    ctx.set_options(bp_tile=0)
    want = oracle.backproject(mr, reb, oracle.BP_GPU, double=True)
    err = np.abs(got - want).max() / np.abs(want).max()
    print(f"bp random ({table_dir}, nz={nz}): rel max err {err:.3g}")
    assert err <= 2e-4
    cb.free_metadata(mr)


@pytest.mark.parametrize("nz", [16, 32])
@pytest.mark.parametrize("nrows,fov", [(8, 50.0), (16, 50.0), (16, 120.0), (16, 250.0), (32, 60.0), (32, 120.0),
                                       (64, 40.0), (64, 100.0)])
def test_backproject_dispatch_shapes(ctx, tmp_path, mini_dir, nrows, fov, nz):
    """Detector heights 8-64 rows and voxel sizes 1.6-7.8 mm select different
    staging shapes (threads per row 8-64, channels per thread 1-14, with and
    without the LDS pad pair, compile-time and generic strides): random
    sinogram vs the FP64 oracle, all but the rare |q| = 1 steps within 2e-4
    of the image maximum."""
    import oracle
    tok = scans.mini_tokens("none", nrows=nrows, nx=32)
    tok["ReconFOV"] = fov
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(nrows + int(fov))
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    print(f"bp shapes: rows {nrows} fov {fov} nz {nz}", flush=True)
    ctx.set_options(proj_dtype=cb.PROJ_F32, bp_tile=nz)
    reb_d = dev(reb)
    ctx.backproject_dev(mr, reb_d, img)
    got = host(img).reshape(32, mr.rp.nx, mr.rp.ny)
    ctx.set_options(bp_tile=0)
    want = oracle.backproject(mr, reb, oracle.BP_GPU, double=True)
    err = np.abs(got - want) / np.abs(want).max()
    n_off = int((err > 2e-4).sum())
    print(f"bp shapes: rows {nrows} fov {fov} nz {nz}: rel max err {err.max():.3g}, {n_off} voxels beyond 2e-4")
    assert n_off <= max(2, 1e-4 * err.size)
    cb.free_metadata(mr)


@pytest.mark.parametrize("nz", [32, 16])
def test_backproject_tall_detector(ctx, tmp_path, mini_dir, nz):
    """128 rebinned rows (z-FFS over 64 detector rows) at a field of view
    whose footprint needs 8 channels per staging thread: with 16-slice
    z-groups the staged window takes 76 KB of LDS per workgroup, above the
    64 KB default, so the launch raises the limit first.  Random sinogram vs
    the FP64 oracle: all but at most 1e-4 of the voxels within 2e-4 of the
    image maximum.  The exceptions are the |q| = 1 steps of the weighting (a
    view that is an angle's only contributor, at |q| within FP32 rounding of
    1, adds its whole sample or nothing); with 64-128 rows a handful of the
    524 288 voxels hit one.  The FP32 build of the oracle shows the same
    steps against the FP64 one (8 voxels here, up to 0.079 of the maximum,
    two of them the voxels where the HIP result steps)."""
    import oracle
    tok = scans.mini_tokens("z", nrows=128, nx=128)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    assert mr.cg.n_rows == 128
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(6)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    ctx.set_options(proj_dtype=cb.PROJ_F32, bp_tile=nz)
    reb_d = dev(reb)
    ctx.backproject_dev(mr, reb_d, img)
    got = host(img).reshape(32, mr.rp.nx, mr.rp.ny)
    ctx.set_options(bp_tile=0)
    want = oracle.backproject(mr, reb, oracle.BP_GPU, double=True)
    err = np.abs(got - want) / np.abs(want).max()
    n_off = int((err > 2e-4).sum())
    print(f"bp tall detector (nz={nz}): rel max err {err.max():.3g}, {n_off} of {err.size} voxels beyond 2e-4")
    assert n_off <= 1e-4 * err.size
    cb.free_metadata(mr)


@pytest.mark.parametrize("nz", [32, 16])
def test_backproject_definitionas_block_sampled(ctx, tmp_path, nz):
    """Full-size phi+z FFS block (512^2, 64 rows): HIP vs FP64 oracle on 400
    sampled voxel columns plus the four 24x24 corner patches (voxels that
    the views of an angle see only at the detector edge, |q| -> 1, where W is
    tiny but h_t must stay non-zero) of a filtered synthetic scan, in HU."""
    import oracle
    tok = scans.c1_tokens()
    mr = _mr(tmp_path, tok, None)
    cb.plan_block(mr, 3)
    s = cb.block_sizes(mr)
    raw = _raw_pull(ctx, mr, scans.phantom_for(tok, 1))
    reb = torch.empty(s.rebin, dtype=torch.float32, device="cuda")
    ctx.set_options(filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32, bp_tile=nz)
    ctx.rebin_filter_dev(mr, raw, reb)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    ctx.backproject_dev(mr, reb, img)
    got = host(img).reshape(32, mr.rp.nx, mr.rp.ny)
    ctx.set_options(bp_tile=0)
    rng = np.random.default_rng(11)
    idx = rng.choice(mr.rp.nx * mr.rp.ny, 400, replace=False)
    cx, cy = list(idx // mr.rp.ny), list(idx % mr.rp.ny)
    for x0 in (0, mr.rp.nx - 24):
        for y0 in (0, mr.rp.ny - 24):
            for x in range(x0, x0 + 24):
                for y in range(y0, y0 + 24):
                    cx.append(x)
                    cy.append(y)
    cx, cy = np.array(cx), np.array(cy)
    want = oracle.backproject(mr, host(reb).reshape(s.n_proj_out, mr.cg.n_rows, -1), oracle.BP_GPU,
                              cols=(cx, cy), double=True)
    g = got[:, cx, cy].astype(np.float64)
    w = want[:, cx, cy].astype(np.float64)
    hu = 1000.0 * (g - w) / 0.02
    n_off = int((np.abs(hu) > 0.05).sum())
    print(f"bp c1 block (nz={nz}): HU max {np.abs(hu).max():.4f} rms {np.sqrt(np.mean(hu ** 2)):.4f}, "
          f"{n_off} of {hu.size} voxels beyond 0.05 HU")
    # W(q) jumps from ~8e-10 to 0 at |q| = 1 and h_t renormalises, so a voxel
    # that one view of an angle sees at |q| within rounding of 1 takes that
    # sample in full or not at all: rare single-voxel steps are inherent to
    # the weighting at FP32 geometry, anything systematic is not.
    assert np.sqrt(np.mean(hu ** 2)) <= 0.02
    assert np.abs(hu).max() <= 2.0
    assert n_off <= 1e-4 * hu.size
    cb.free_metadata(mr)


def test_backproject_constant_sinogram(ctx, tmp_path, mini_dir):
    """Closed form: a constant filtered sinogram c backprojects to pi'*c inside
    the acquisition FOV wherever every angle sees the voxel."""
    tok = scans.mini_tokens("none", nrows=8, nx=32)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    reb = np.full((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled), 2.5, dtype=np.float32)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    reb_d = dev(reb)
    ctx.backproject_dev(mr, reb_d, img)
    got = host(img).reshape(32, mr.rp.nx, mr.rp.ny)
    want = np.float32(3.1415368979) * 2.5
    assert np.abs(got - want).max() <= 2e-5 * want
    cb.free_metadata(mr)


def test_finalize_bit_exact(ctx, tmp_path, mini_dir):
    import oracle
    tok = scans.mini_tokens("none")
    tok["NSlices"] = 6
    tok["EndPos"] = tok["StartPos"] + 5 * 2.4
    mr = _mr(tmp_path, tok, mini_dir)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(1)
    img = (rng.standard_normal(s.image_stack) * 0.01 + 0.02).astype(np.float32)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    for mu in (0.0, 0.02):
        ctx.set_options(mu_water=mu)
        img_d = dev(img)
        ctx.finalize_dev(mr, img_d, out)
        got = host(out).reshape(-1, mr.rp.nx, mr.rp.ny)
        want = oracle.finalize(mr, img, mu, double=False)
        assert np.array_equal(got, want), mu
    ctx.set_options(mu_water=0.0)
    cb.free_metadata(mr)


def _oracle_volume(mr, raw_all, taps, adaptive, mu):
    """FP64 oracle pipeline over every block (FP32-exact adaptive filter)."""
    import oracle
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    n_read = raw_all.shape[0]
    blocks = []
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
        win = np.zeros((n, frame), dtype=np.float32)
        lo, hi = max(s0, 0), min(s0 + n, n_read)
        win[lo - s0:hi - s0] = raw_all[lo:hi]
        if adaptive:
            win, _ = oracle.adaptive_filter(mr, win)
        reb = oracle.rebin_filter(mr, win, taps, double=True)
        blocks.append(oracle.backproject(mr, reb, oracle.BP_GPU, double=True))
    img = np.concatenate(blocks, axis=0)
    return oracle.finalize(mr, img, mu, double=True)


@pytest.mark.parametrize("ffs,table_dir", [("none", "out"), ("both", "out"), ("z", "in"), ("phi", "out")])
def test_end_to_end_hu(ctx, tmp_path, mini_dir, ffs, table_dir):
    tok = scans.mini_tokens(ffs, nrows=8 if ffs in ("z", "both") else 4, nx=32, table_dir=table_dir)
    mr = _mr(tmp_path, tok, mini_dir)
    pp = scans.phantom_for(tok, 5)
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, pp, 0, n_read, raw)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32)
    ctx.reconstruct_dev(mr, raw, n_read, out)
    got = host(out).reshape(-1, mr.rp.nx, mr.rp.ny)
    taps = cb.load_filter(mr)
    want = _oracle_volume(mr, host(raw).reshape(n_read, frame), taps, True, 0.02)
    d = got.astype(np.float64) - want
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"end-to-end {ffs}/{table_dir}: HU rms {rms:.4f} max {mx:.4f}; image HU range {want.min():.0f}..{want.max():.0f}")
    assert rms <= 1.0 and mx <= 10.0
    cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "both"])
def test_end_to_end_pull_before_data(ctx, tmp_path, mini_dir, ffs):
    """TableDir in with the first slice at the highest position the planner
    accepts: the first block's pull starts before reading 0 (the reference
    keeps the start in a size_t, so it wraps), and the readings before the
    data are zeros.  Device-resident and host-streaming reconstructions
    agree bit for bit and match the FP64 oracle fed the same zero-padded
    window."""
    base = scans.mini_tokens(ffs, nrows=8 if ffs == "both" else 4, nx=32, table_dir="in")
    mr = None
    for k in range(199, -1, -1):
        tok = dict(base)
        tok["StartPos"] = 0.5 * k
        tok["EndPos"] = tok["StartPos"] + (tok["NSlices"] - 1) * 2.4
        try:
            mr = _mr(tmp_path, tok, mini_dir, name=f"edge_{k}.prm")
        except cb.CTBBError:
            continue
        break
    assert mr is not None
    cb.plan_block(mr, 0)
    assert cb.pull_start(mr) < 0
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 6), 0, n_read, raw)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    ctx.reconstruct_dev(mr, raw, n_read, out)
    got = host(out).reshape(-1, mr.rp.nx, mr.rp.ny)
    h_out = torch.empty(s.final_stack, dtype=torch.float32).pin_memory()
    ctx.reconstruct_host(mr, raw.cpu().pin_memory(), n_read, h_out)
    assert np.array_equal(h_out.numpy().reshape(got.shape), got)
    want = _oracle_volume(mr, host(raw).reshape(n_read, frame), cb.load_filter(mr), True, 0.02)
    d = got.astype(np.float64) - want
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"pull before data {ffs}: start {tok['StartPos']}, first pull {cb.pull_start(mr)}; HU rms {rms:.4f} max {mx:.4f}")
    assert rms <= 1.0 and mx <= 10.0
    cb.free_metadata(mr)


def test_back_to_back_calls(ctx, tmp_path, mini_dir):
    """Whole-volume calls queued back to back without synchronizing (the next
    call's first preparation runs beside the previous call's last
    backprojection, in workspaces of its own): every output equals the one
    of the same call made alone, also when the scans (and so the finalize
    tables, view tables and buffer sizes) change from call to call."""
    items = []
    for i, (ffs, nx, rot) in enumerate((("both", 32, 24), ("none", 48, 20), ("both", 32, 24))):
        tok = scans.mini_tokens(ffs, nrows=8, nx=nx, rotations=rot, n_slices=30)
        mr = _mr(tmp_path, tok, mini_dir, name=f"b2b_{i}.prm")
        frame = mr.cg.n_rows_raw * mr.cg.n_channels
        raw = torch.empty(tok["Readings"] * frame, dtype=torch.float32, device="cuda")
        ctx.synth_scan_dev(mr, scans.phantom_for(tok, 7 + i), 0, tok["Readings"], raw)
        items.append((mr, raw, tok["Readings"]))
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    alone = []
    for mr, raw, n in items:
        out = torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda")
        ctx.reconstruct_dev(mr, raw, n, out)
        ctx.synchronize()
        alone.append(host(out).copy())
    outs = [torch.zeros(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda") for mr, _, _ in items]
    torch.cuda.synchronize()
    for _ in range(2):
        for (mr, raw, n), out in zip(items, outs):
            ctx.reconstruct_dev(mr, raw, n, out)
    ctx.synchronize()
    for a, out in zip(alone, outs):
        assert np.array_equal(host(out), a)
    for mr, _, _ in items:
        cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "z", "phi", "both"])
def test_pipeline_matches_stage_chain(ctx, tmp_path, mini_dir, ffs):
    """The whole-scan pipeline (block b + 1 prepared on one stream while block
    b is backprojected on another, two filtered-projection buffers) gives
    exactly the result of the stage entry points called block by block
    (adaptive filter, rebin + ramp filter, backprojection, finalize)."""
    tok = scans.mini_tokens(ffs, nrows=8, nx=32, rotations=24, n_slices=40)
    mr = _mr(tmp_path, tok, mini_dir)
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    n_read = tok["Readings"]
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 4), 0, n_read, raw)
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_read, out)
    got = host(out)
    n_blocks = int(mr.ri.n_blocks)
    assert n_blocks >= 3
    stack = torch.empty(n_blocks * s.image_block, dtype=torch.float32, device="cuda")
    changed = 0
    for b in range(n_blocks):
        cb.plan_block(mr, b)
        sb = cb.block_sizes(mr)
        s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
        assert s0 >= 0 and s0 + n <= n_read
        win = raw[s0 * frame:(s0 + n) * frame]
        af = torch.empty_like(win)
        ctx.adaptive_filter_dev(mr, win, af)
        # the context's streams are not torch's: wait before torch reads af
        # or frees a buffer the queued work still uses
        ctx.synchronize()
        changed += int((af != win).sum())
        reb = torch.empty(sb.rebin, dtype=torch.float32, device="cuda")
        ctx.rebin_filter_dev(mr, af, reb)
        ctx.backproject_dev(mr, reb, stack[b * sb.image_block:(b + 1) * sb.image_block])
        ctx.synchronize()
    want = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.finalize_dev(mr, stack, want)
    want = host(want)
    print(f"pipeline vs stages {ffs}: {changed} samples smoothed, mismatches {np.count_nonzero(got != want)}")
    assert changed > 0
    assert np.array_equal(got, want)
    cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "both"])
def test_reconstruct_host_matches_dev(ctx, tmp_path, mini_dir, ffs):
    """Host-streaming reconstruction (pull double buffer, overlap reuse, zero
    padding past the end of the scan) gives exactly the device-resident
    result, from page-locked and from pageable host memory."""
    tok = scans.mini_tokens(ffs, nrows=8, nx=32, rotations=24, n_slices=40)
    mr = _mr(tmp_path, tok, mini_dir)
    assert mr.ri.n_blocks >= 3
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    n_read = tok["Readings"]
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 5), 0, n_read, raw)
    cb.plan_block(mr, int(mr.ri.n_blocks) - 1)
    n_use = cb.pull_start(mr) + int(mr.ri.n_proj_pull) - 100   # last pull runs past the data
    s = cb.block_sizes(mr)
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    out_d = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_use, out_d)
    want = host(out_d)
    h_raw = raw.cpu().pin_memory()
    h_out = torch.empty(s.final_stack, dtype=torch.float32).pin_memory()
    ctx.reconstruct_host(mr, h_raw, n_use, h_out)
    assert np.array_equal(h_out.numpy(), want)
    np_out = np.zeros(s.final_stack, dtype=np.float32)
    ctx.reconstruct_host(mr, host(raw), n_use, np_out)
    assert np.array_equal(np_out, want)
    cb.free_metadata(mr)


@pytest.mark.parametrize("mode,group_mb,filt,proj", [
    (cb.BATCH_PERSISTENT, 0, cb.FILTER_FFT32, cb.PROJ_F32), (cb.BATCH_PERSISTENT, 1, cb.FILTER_FFT32, cb.PROJ_F32),
    (cb.BATCH_PIPELINE, 0, cb.FILTER_FFT32, cb.PROJ_F32),
    # BF16 projections or the matrix-core filter: the preparation graph
    # captures the whole per-block preparation, and the persistent kernel
    # reads BF16 windows
    (cb.BATCH_PERSISTENT, 1, cb.FILTER_F16X3, cb.PROJ_BF16), (cb.BATCH_PERSISTENT, 0, cb.FILTER_FFT32, cb.PROJ_BF16)])
def test_reconstruct_batch_matches_single(ctx, tmp_path, mini_dir, mode, group_mb, filt, proj):
    """Batching form: scans of different length, pitch, FOV and matrix size
    give exactly the one-scan-at-a-time results, through the persistent form
    (scans grouped per kernel variant; with a 1 MiB group budget every scan
    is a group of its own, so the two arenas alternate) and through the
    per-block pipeline; a batch mixing focal-spot modes (other rebin maps)
    falls back and still agrees."""
    def scan(ffs, nx, rot, n_sl, fov, pitch, name):
        tok = scans.mini_tokens(ffs, nrows=8, nx=nx, rotations=rot, n_slices=n_sl)
        tok["ReconFOV"] = fov
        tok["PitchValue"] = round(pitch * 8 * 1.2, 4)
        length = rot * tok["PitchValue"]
        tok["StartPos"] = round(0.5 * length - 1.2 * n_sl, 3)
        tok["EndPos"] = tok["StartPos"] + (n_sl - 1) * 2.4
        mr = _mr(tmp_path, tok, mini_dir, name=name)
        frame = mr.cg.n_rows_raw * mr.cg.n_channels
        raw = torch.empty(tok["Readings"] * frame, dtype=torch.float32, device="cuda")
        ctx.synth_scan_dev(mr, scans.phantom_for(tok, 5), 0, tok["Readings"], raw)
        out = torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda")
        return mr, raw, tok["Readings"], out
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=filt, proj_dtype=proj)
    for combo in (("both", "both", "both"), ("both", "z", "both")):
        items = [scan(combo[0], 32, 24, 30, 250.0, 0.8, "a.prm"), scan(combo[1], 48, 30, 40, 200.0, 0.6, "b.prm"),
                 scan(combo[2], 16, 20, 12, 300.0, 1.1, "c.prm")]
        singles = []
        for mr, raw, n, out in items:
            ctx.reconstruct_dev(mr, raw, n, out)
            singles.append(host(out).copy())
            out.zero_()
        torch.cuda.synchronize()    # torch's zero fill before the context's streams write out
        ctx.set_options(batch_mode=mode, batch_group_mb=group_mb)
        ctx.reconstruct_batch_dev([i[0] for i in items], [i[1] for i in items], [i[2] for i in items],
                                  [i[3] for i in items])
        ctx.synchronize()
        ctx.set_options(batch_mode=cb.BATCH_PERSISTENT, batch_group_mb=0)
        for (mr, raw, n, out), want in zip(items, singles):
            assert np.array_equal(host(out), want)
            cb.free_metadata(mr)
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)


@pytest.mark.parametrize("mode,proj", [(cb.FILTER_BF16X3, cb.PROJ_F32), (cb.FILTER_F16X3, cb.PROJ_F32),
                                       (cb.FILTER_BF16X3, cb.PROJ_BF16), (cb.FILTER_FFT32, cb.PROJ_F32),
                                       (cb.FILTER_FFT32, cb.PROJ_BF16)])
def test_end_to_end_precision_form_hu(ctx, tmp_path, mini_dir, mode, proj):
    """Precision form (split MFMA filter, optionally BF16 projections with FP32
    accumulation) vs the FP64 oracle: the precision forms' bar is <= 1 HU RMS."""
    tok = scans.mini_tokens("both", nrows=8, nx=32)
    mr = _mr(tmp_path, tok, mini_dir)
    pp = scans.phantom_for(tok, 5)
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, pp, 0, n_read, raw)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=mode, proj_dtype=proj)
    ctx.reconstruct_dev(mr, raw, n_read, out)
    ctx.set_options(filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32)
    got = host(out).reshape(-1, mr.rp.nx, mr.rp.ny)
    want = _oracle_volume(mr, host(raw).reshape(n_read, frame), cb.load_filter(mr), True, 0.02)
    d = got.astype(np.float64) - want
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"precision form mode={mode} proj={proj}: HU rms {rms:.4f} max {mx:.4f}")
    assert rms <= 1.0
    cb.free_metadata(mr)


def test_reference_host_contract(ctx, tmp_path, mini_dir):
    """The reference's calling sequence (main.cu:207-256) on host buffers
    gives the same volume as the device-resident pipeline."""
    import ctypes as C
    tok = scans.mini_tokens("both", nrows=8, nx=32)
    tok["RawDataDir"] = str(tmp_path)
    mr = _mr(tmp_path, tok, mini_dir)
    pp = scans.phantom_for(tok, 6)
    n_read = tok["Readings"]
    raw_h = cb.synth_scan(mr, pp, 0, n_read)
    raw_h.tofile(tmp_path / "synthetic.raw")
    cb.free_metadata(mr)
    p = str(tmp_path / "host.prm")
    cb.write_prm(p, tok)
    mr = cb.metadata_from_prm(p, install_dir=mini_dir, plan=False)
    L = cb.lib()
    assert L.ctbb_configure_reconstruction(C.byref(mr)) == 0
    ctx.set_options(mu_water=0.0, adaptive_filter=1, filter_mode=cb.FILTER_F32, proj_dtype=cb.PROJ_F32)
    for _ in range(mr.ri.n_blocks):
        assert L.ctbb_update_block_info(C.byref(mr)) == 0
        assert L.ctbb_extract_projections(C.byref(mr)) == 0
        ctx.adaptive_filter_kk(mr)
        ctx.rebin_filter(mr)
        assert mr.ctd.d_rebin
        ctx.backproject(mr)
        assert not mr.ctd.d_rebin
    ctx.finalize_image_stack(mr)
    nf = cb.n_slices_final(mr)
    got = np.ctypeslib.as_array(mr.ctd.final_image_stack, shape=(nf * mr.rp.nx * mr.rp.ny,)).copy()
    raw = dev(raw_h.ravel())
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_read, out)
    want = host(out)
    assert np.array_equal(got, want)
    mr.ctd.final_image_stack = None
    cb.free_metadata(mr)


def test_driver_executable_matches_dev(ctx, tmp_path, mini_dir):
    """The reference-style driver (main.cu:69-271 over the compat stage
    functions) run as a program on a prm file writes the same stack as the
    device-resident pipeline with the driver's options."""
    import subprocess
    tok = scans.mini_tokens("phi", nrows=8, nx=32)
    tok["RawDataDir"] = str(tmp_path)
    tok["OutputDir"] = str(tmp_path)
    tok["OutputFile"] = "driver.img"
    mr = _mr(tmp_path, tok, mini_dir)
    pp = scans.phantom_for(tok, 4)
    n_read = tok["Readings"]
    raw_h = cb.synth_scan(mr, pp, 0, n_read)
    raw_h.tofile(tmp_path / "synthetic.raw")
    p = str(tmp_path / "driver.prm")
    cb.write_prm(p, tok)
    env = dict(os.environ, CTBB_INSTALL_DIR=mini_dir)
    exe = os.path.join(cb.PKG_DIR, "ctbangbang")
    print(f"driver: {exe} --device=0 {p}", flush=True)
    r = subprocess.run([exe, "--device=0", "--timing", p], capture_output=True, text=True, env=env, timeout=100)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ms for entire reconstruction" in r.stdout
    got = np.fromfile(tmp_path / "driver.img", dtype=np.float32)
    # the driver's options: F32 filter, FP32 projections, adaptive filter,
    # HU with the prm's MuWater
    ctx.set_options(mu_water=float(tok["MuWater"]), adaptive_filter=1, filter_mode=cb.FILTER_F32,
                    proj_dtype=cb.PROJ_F32)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, dev(raw_h.ravel()), n_read, out)
    want = host(out)
    ctx.set_options(mu_water=0.0, filter_mode=cb.FILTER_FFT32)
    assert got.shape == want.shape
    assert np.array_equal(got, want)
    cb.free_metadata(mr)


def _hu16_ref(x, slope, intercept):
    """int16 stored values in numpy float32: IEEE subtraction and division,
    round half to even, NaN -> 0, saturation."""
    v = np.rint((x.astype(np.float32) - np.float32(intercept)) / np.float32(slope))
    v = np.where(np.isnan(v), np.float32(0), v)
    return np.clip(v, -32768, 32767).astype(np.int16)


@pytest.mark.parametrize("n", [0, 1, 7, 8, 9, 4099, 1_000_003])
@pytest.mark.parametrize("offset", [0, 1])
@pytest.mark.parametrize("slope,intercept", [(1.0, 0.0), (0.5, -1024.0), (2.5, 3.0)])
def test_hu_to_int16_bit_exact(ctx, n, offset, slope, intercept):
    rng = np.random.default_rng(n + 7 * offset)
    x = (rng.standard_normal(n + 1) * 3000.0).astype(np.float32)
    special = np.array([0.5, 1.5, 2.5, -0.5, -1.5, np.nan, np.inf, -np.inf, 1e9, -1e9, 32767.5, -32768.5,
                        16383.75, -0.0], dtype=np.float32)
    x[1:1 + min(n, special.size)] = special[:min(n, special.size)]
    xd = dev(x)
    out = torch.full((n + 2,), 12345, dtype=torch.int16, device="cuda")
    # offset 1 element: neither buffer 16-byte aligned (single-sample path)
    ctx.hu_to_int16_dev(xd[offset:offset + n], n, out[offset:offset + n], slope=slope, intercept=intercept)
    got = host(out)
    np.testing.assert_array_equal(got[offset:offset + n], _hu16_ref(x[offset:offset + n], slope, intercept))
    assert (np.delete(got, np.arange(offset, offset + n)) == 12345).all()      # nothing written outside


def test_hu_to_int16_rejects_bad_rescale(ctx):
    x = torch.zeros(16, dtype=torch.float32, device="cuda")
    o = torch.zeros(16, dtype=torch.int16, device="cuda")
    for slope, intercept in [(0.0, 0.0), (-1.0, 0.0), (float("nan"), 0.0), (float("inf"), 0.0), (1.0, float("nan"))]:
        with pytest.raises(cb.CTBBError):
            ctx.hu_to_int16_dev(x, 16, o, slope=slope, intercept=intercept)


def test_dicom_series_from_reconstruction(ctx, tmp_path, mini_dir):
    """Reconstruction in HU -> int16 on the device -> DICOM series: the
    pixel data read back equals the device's int16 stack, which is the HU
    stack rounded."""
    import dicomfile as D
    tok = scans.mini_tokens("z", nrows=8, nx=32, n_slices=6)
    mr = _mr(tmp_path, tok, mini_dir)
    pp = scans.phantom_for(tok, 2)
    n_read = tok["Readings"]
    raw = dev(cb.synth_scan(mr, pp, 0, n_read).ravel())
    ctx.set_options(mu_water=float(tok["MuWater"]), adaptive_filter=1, filter_mode=cb.FILTER_FFT32,
                    proj_dtype=cb.PROJ_F32)
    s = cb.block_sizes(mr)
    hu = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_read, hu)
    i16 = torch.empty(s.final_stack, dtype=torch.int16, device="cuda")
    ctx.hu_to_int16_dev(hu, s.final_stack, i16)
    hu_h, st = host(hu), host(i16)
    ctx.set_options(mu_water=0.0)
    np.testing.assert_array_equal(st, _hu16_ref(hu_h, 1.0, 0.0))
    assert np.abs(st - hu_h).max() <= 0.5
    nf, nx, ny = cb.n_slices_final(mr), mr.rp.nx, mr.rp.ny
    cb.write_dicom_series(mr, st.reshape(nf, nx, ny), str(tmp_path), "recon")
    for k in range(nf):
        _, data, _ = D.read(str(tmp_path / f"recon_{k + 1:04d}.dcm"))
        np.testing.assert_array_equal(D.pixels(data[(0x7FE0, 0x10)], nx, ny), st.reshape(nf, nx, ny)[k])
    cb.free_metadata(mr)


def test_autotune_backprojection(ctx, tmp_path, mini_dir):
    """The tile autotuner picks 16 or 32, sets it, records it per shape in
    its cache and reuses the cached choice; reconstructions with either tile
    agree to FP32 rounding."""
    tok = scans.mini_tokens("z", nrows=8, nx=32)
    mr = _mr(tmp_path, tok, mini_dir)
    ctx.set_options(mu_water=0.02, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32,
                    bp_tile=0)
    cache = str(tmp_path / "tune.txt")
    cb.plan_block(mr, 0)
    before = (cb.pull_start(mr), int(mr.ri.n_proj_pull), int(mr.ri.cb.block_idx))
    t1 = ctx.autotune_backprojection(mr, cache)
    assert t1 in (16, 32) and ctx.options().bp_tile == t1
    assert (cb.pull_start(mr), int(mr.ri.n_proj_pull), int(mr.ri.cb.block_idx)) == before
    text = open(cache).read()
    assert text.count("\n") == 1 and text.split()[1] == str(t1)
    ctx.set_options(bp_tile=0)
    assert ctx.autotune_backprojection(mr, cache) == t1
    assert open(cache).read() == text                   # cached: not timed again
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 3), 0, n_read, raw)
    s = cb.block_sizes(mr)
    outs = []
    for tile in (16, 32):
        ctx.set_options(bp_tile=tile)
        out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
        ctx.reconstruct_dev(mr, raw, n_read, out)
        outs.append(host(out))
    ctx.set_options(bp_tile=0, mu_water=0.0)
    d = np.abs(outs[0].astype(np.float64) - outs[1])
    print(f"autotune: tile {t1}; tile 16 vs 32 max HU difference {d.max():.2e}")
    assert d.max() <= 1e-3       # HU
    cb.free_metadata(mr)


def test_context_options_validated(ctx):
    """Unsupported option values are rejected and leave the options unchanged."""
    before = ctx.options()
    for bad in (dict(bp_tile=7), dict(bp_views_per_pass=2), dict(filter_mode=9), dict(proj_dtype=5)):
        with pytest.raises(cb.CTBBError):
            ctx.set_options(**bad)
        now = ctx.options()
        assert (now.bp_tile, now.bp_views_per_pass, now.filter_mode, now.proj_dtype) == \
            (before.bp_tile, before.bp_views_per_pass, before.filter_mode, before.proj_dtype)
    ctx.set_options(bp_tile=32)
    ctx.set_options(bp_tile=0)


def test_nonfinite_raw_is_reported(ctx, tmp_path, mini_dir):
    """A NaN or infinite raw sample that enters the reconstruction makes the
    next synchronize report CTBB_ERR_NONFINITE (12) once; the host-streaming
    and host-contract entry points return it themselves.  A finite scan
    afterwards reconstructs cleanly and to the same bits as before."""
    tok = scans.mini_tokens("both", nrows=8, nx=32)
    mr = _mr(tmp_path, tok, mini_dir)
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 2), 0, n_read, raw)
    ctx.set_options(mu_water=0.02, adaptive_filter=1, adaptive_threshold=0.0, filter_mode=cb.FILTER_FFT32,
                    proj_dtype=cb.PROJ_F32, bp_tile=0)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.synchronize()
    ctx.reconstruct_dev(mr, raw, n_read, out)
    ctx.synchronize()
    want = host(out)
    cb.plan_block(mr, 0)
    mid = cb.pull_start(mr) + int(mr.ri.n_proj_pull) // 2
    for bad in (float("nan"), float("inf")):
        r = raw.clone()
        r[mid * frame + 3 * mr.cg.n_channels + 60] = bad
        torch.cuda.synchronize()
        ctx.reconstruct_dev(mr, r, n_read, out)
        with pytest.raises(cb.CTBBError) as e:
            ctx.synchronize()
        assert e.value.code == 12
        ctx.synchronize()                       # reported once
        h = r.cpu().pin_memory()
        h_out = torch.empty(s.final_stack, dtype=torch.float32).pin_memory()
        with pytest.raises(cb.CTBBError) as e:
            ctx.reconstruct_host(mr, h, n_read, h_out)
        assert e.value.code == 12
    ctx.reconstruct_dev(mr, raw, n_read, out)
    ctx.synchronize()
    assert np.array_equal(host(out), want)
    cb.free_metadata(mr)


@pytest.mark.parametrize("ffs", ["none", "both"])
def test_adaptive_fixed_threshold_bit_exact(ctx, tmp_path, mini_dir, ffs):
    """Fixed-threshold adaptive filter (every view smooths the samples above
    one line-integral threshold) bit for bit against the restatement's
    smoothing step with that threshold (stage entry point: the 16-byte apply
    kernel)."""
    import oracle
    tok = scans.mini_tokens(ffs, nrows=8 if ffs == "both" else 4)
    mr = _mr(tmp_path, tok, mini_dir)
    cb.plan_block(mr, 0)
    raw = _raw_pull(ctx, mr, scans.phantom_for(tok, 4))
    r = host(raw)
    T = float(np.float32(np.quantile(r, 0.7)))
    ctx.set_options(adaptive_threshold=T)
    out = torch.empty_like(raw)
    ctx.adaptive_filter_dev(mr, raw, out)
    ctx.synchronize()
    ctx.set_options(adaptive_threshold=0.0)
    n = int(mr.ri.n_proj_pull)
    want = oracle.adaptive_apply(mr, r.reshape(n, -1), np.full(n, np.float32(T), dtype=np.float32))
    got = host(out).reshape(want.shape)
    changed = int(np.count_nonzero(want != r.reshape(want.shape)))
    print(f"adaptive fixed T={T:.4f} ({ffs}): {changed} samples smoothed, mismatches {np.count_nonzero(got != want)}")
    assert changed > 0.2 * want.size
    assert np.array_equal(got, want)
    cb.free_metadata(mr)


def test_end_to_end_fixed_threshold(ctx, tmp_path, mini_dir):
    """Whole-scan pipeline with the fixed-threshold adaptive filter (blocks
    after the first use the one-sample apply kernel beside the running
    backprojection) against the FP64 oracle with the same threshold."""
    import oracle
    tok = scans.mini_tokens("both", nrows=8, nx=32, rotations=24, n_slices=40)
    mr = _mr(tmp_path, tok, mini_dir)
    assert int(mr.ri.n_blocks) >= 3
    n_read = tok["Readings"]
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, 6), 0, n_read, raw)
    raw_h = host(raw).reshape(n_read, frame)
    T = float(np.float32(np.log(1e5 / 3e4)))
    ctx.set_options(mu_water=0.02, adaptive_filter=1, adaptive_threshold=T, filter_mode=cb.FILTER_FFT32,
                    proj_dtype=cb.PROJ_F32, bp_tile=0)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_read, out)
    ctx.synchronize()
    ctx.set_options(adaptive_threshold=0.0)
    got = host(out).reshape(-1, mr.rp.nx, mr.rp.ny)
    taps = cb.load_filter(mr)
    blocks, smoothed = [], 0
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
        win = np.zeros((n, frame), dtype=np.float32)
        lo, hi = max(s0, 0), min(s0 + n, n_read)
        win[lo - s0:hi - s0] = raw_h[lo:hi]
        af = oracle.adaptive_apply(mr, win, np.full(n, np.float32(T), dtype=np.float32))
        smoothed += int(np.count_nonzero(af != win))
        blocks.append(oracle.backproject(mr, oracle.rebin_filter(mr, af, taps, double=True), oracle.BP_GPU,
                                         double=True))
    want = oracle.finalize(mr, np.concatenate(blocks), 0.02, double=True)
    d = got.astype(np.float64) - want
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"end-to-end fixed threshold {T:.4f}: {smoothed} samples smoothed; HU rms {rms:.4f} max {mx:.4f}")
    assert smoothed > 0
    assert rms <= 0.05 and mx <= 2.0
    cb.free_metadata(mr)


def test_device_planner_matches_host(ctx, tmp_path):
    """The batching form's work list is planned on the device: every block
    plan of the 32 seeded batch scans and of the c1 scan (pull range, slice
    range and its table-position indices, half turns) equals the host
    planner's (update_block_info arithmetic)."""
    mrs = []
    for i in range(32):
        mrs.append(_mr(tmp_path, scans.c3_tokens(i), None, name=f"c3_{i}.prm"))
    mrs.append(_mr(tmp_path, scans.c1_tokens(), None, name="c1.prm"))
    got = ctx.plan_batch_dev(mrs)
    j = 0
    for s, mr in enumerate(mrs):
        for b in range(mr.ri.n_blocks):
            cb.plan_block(mr, b)
            g = got[j]
            assert (g["scan"], g["block"]) == (s, b)
            # the host struct keeps the pull start as size_t (a negative start
            # wraps); the device plan keeps it as int32
            v = cb.pull_start(mr) & 0xffffffff
            assert g["idx_pull_start"] == (v - (1 << 32) if v >= (1 << 31) else v)
            assert g["n_proj_pull"] == int(mr.ri.n_proj_pull)
            assert g["n_half_turns"] == cb.block_sizes(mr).n_half_turns
            assert g["idx_block_slice_start"] == int(mr.ri.cb.idx_block_slice_start)
            assert g["idx_block_slice_end"] == int(mr.ri.cb.idx_block_slice_end)
            assert g["block_slice_start"] == np.float32(mr.ri.cb.block_slice_start)
            assert g["block_slice_end"] == np.float32(mr.ri.cb.block_slice_end)
            j += 1
    assert j == got.size
    print(f"device planner: {j} block plans of {len(mrs)} scans equal the host's")
    for mr in mrs:
        cb.free_metadata(mr)
