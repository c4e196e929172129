"""HIP path vs the CPU oracle at the benchmark configurations' own sizes
(needs a GPU; the oracle runs on the host's cores, OpenMP and threaded FFTs).

Workloads (ctbangbang_amd/scans.py):
  c0  736 ch x 16 rows, no flying focal spot, 256^2, one block, 16 slices of
      1 mm: the whole volume against the FP64 oracle.
  c1  phi+z flying focal spot, 32 -> 64 rows, 512^2, 7 blocks -> 40 slices of
      3 mm, adaptive filter (KK), FFT32 ramp filter (the benchmark's form):
      every block end to end, compared on sampled voxel columns (all slices).
      The adaptive filter is also checked bit for bit on a whole C1 pull
      (KK and fixed-threshold modes), and rebin + ramp filter on a whole C1
      block against the FP64 oracle's.
  c2  the precision forms of c1 (split MFMA filters, BF16 projections): HU
      error per dtype variant against the same FP64 reference.
  c3  two of the seeded batch scans (768^2 and 512^2, different pitch and
      FOV) through the batching entry point, against the oracle.
  c4  z flying focal spot, 64 rows of 0.5 mm, 1024^2 over 400 mm, the first
      two blocks of the scan (fewer slices), through the host-streaming entry
      point, against the oracle.

Bars (HU = 1000 (mu - mu_w) / mu_w, mu_w = 0.02 /mm):
  plain FP32 forms     RMS <= 0.05 HU, max <= 2 HU (the max allows the rare
                       |q| = 1 steps of the weighting, where one view's
                       sample enters in full or not at all)
  precision forms      RMS <= 1 HU (the precision forms' bar), max <= 10 HU
  adaptive filter      bit-exact against the FP32 restatement
  rebin + ramp filter  max |HIP - FP64| <= 1e-4 of the block's max |value|
"""
import os

import numpy as np
import pytest

import ctbangbang_amd as cb
from ctbangbang_amd import scans

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

torch = pytest.importorskip("torch")

MU_W = 0.02


def _mr(tmp_path, tok, name):
    p = str(tmp_path / name)
    cb.write_prm(p, tok)
    return cb.metadata_from_prm(p)


def _scan_dev(ctx, mr, tok, cfg):
    n_read = int(tok["Readings"])
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, cfg), 0, n_read, raw)
    ctx.synchronize()
    return raw, n_read, frame


def _columns(mr, n_random, seed, corner=8):
    """Sampled voxel columns: n_random random ones plus the four corner
    patches (corner x corner), where angles see the voxels only at the
    detector's edge rows."""
    nx, ny = int(mr.rp.nx), int(mr.rp.ny)
    rng = np.random.default_rng(seed)
    idx = rng.choice(nx * ny, n_random, replace=False)
    cx, cy = list(idx // ny), list(idx % ny)
    for x0 in (0, nx - corner):
        for y0 in (0, ny - corner):
            for x in range(x0, x0 + corner):
                for y in range(y0, y0 + corner):
                    cx.append(x)
                    cy.append(y)
    return np.array(cx, dtype=np.int32), np.array(cy, dtype=np.int32)


def _window(raw_h, s0, n, frame):
    n_read = raw_h.shape[0]
    win = np.zeros((n, frame), dtype=np.float32)
    lo, hi = max(s0, 0), min(s0 + n, n_read)
    if hi > lo:
        win[lo - s0:hi - s0] = raw_h[lo:hi]
    return win


def oracle_volume(mr, raw_h, cols, adaptive):
    """FP64 oracle reconstruction (adaptive filter in its FP32 restatement,
    which the HIP kernels match bit for bit): every block, backprojection on
    the voxel columns `cols` (None: all), then finalize in HU.
    adaptive: None, "kk" or a fixed threshold (float)."""
    import oracle
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    taps = cb.load_filter(mr)
    blocks = []
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        win = _window(raw_h, cb.pull_start(mr), int(mr.ri.n_proj_pull), frame)
        if adaptive == "kk":
            win, _ = oracle.adaptive_filter(mr, win)
        elif adaptive is not None:
            thr = np.full(int(mr.ri.n_proj_pull), np.float32(adaptive), dtype=np.float32)
            win = oracle.adaptive_apply(mr, win, thr)
        reb = oracle.rebin(mr, win, double=True)
        # the backprojection reads views i + k * half (i < half, k < n_ht)
        nv = int(mr.cg.n_proj_turn // 2) * int(cb.block_sizes(mr).n_half_turns)
        reb[:nv] = oracle.filter_fft(reb[:nv], taps).astype(np.float32)
        blocks.append(oracle.backproject(mr, reb, oracle.BP_GPU, cols=cols, double=True))
        del reb, win
    return oracle.finalize(mr, np.concatenate(blocks, axis=0), MU_W, double=True)


def hu_err(got, want, cols, label):
    if cols is not None:
        g, w = got[:, cols[0], cols[1]], want[:, cols[0], cols[1]]
    else:
        g, w = got, want
    d = g.astype(np.float64) - w
    rms, mx = float(np.sqrt(np.mean(d ** 2))), float(np.abs(d).max())
    print(f"{label}: HU rms {rms:.4f} max {mx:.4f} over {d.size} voxels; image HU {w.min():.0f}..{w.max():.0f}",
          flush=True)
    return rms, mx


# --------------------------------------------------------------------------
# c0: whole volume


def test_c0_whole_volume(ctx, tmp_path):
    tok = scans.c0_tokens()
    tok["MuWater"] = MU_W
    mr = _mr(tmp_path, tok, "c0.prm")
    assert int(mr.ri.n_blocks) == 1 and int(mr.rp.nx) == 256
    raw, n_read, frame = _scan_dev(ctx, mr, tok, 0)
    ctx.set_options(mu_water=MU_W, adaptive_filter=0, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32,
                    bp_tile=0)
    out = torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda")
    ctx.reconstruct_dev(mr, raw, n_read, out)
    ctx.synchronize()
    got = out.cpu().numpy().reshape(-1, mr.rp.nx, mr.rp.ny)
    want = oracle_volume(mr, raw.cpu().numpy().reshape(n_read, frame), None, None)
    rms, mx = hu_err(got, want, None, "c0 whole volume (fft32)")
    assert rms <= 0.05 and mx <= 2.0
    cb.free_metadata(mr)


# --------------------------------------------------------------------------
# c1 / c2: one oracle reference, several forms


@pytest.fixture(scope="module")
def c1_case(tmp_path_factory):
    from conftest import has_gpu
    if not has_gpu():
        pytest.skip("no GPU")
    tmp = tmp_path_factory.mktemp("c1")
    tok = scans.c1_tokens()
    mr = _mr(tmp, tok, "c1.prm")
    c = cb.Context(0)
    raw, n_read, frame = _scan_dev(c, mr, tok, 1)
    raw_h = raw.cpu().numpy().reshape(n_read, frame)
    cols = _columns(mr, 400, 11)
    want = oracle_volume(mr, raw_h, cols, "kk")
    yield {"tok": tok, "mr": mr, "raw": raw, "raw_h": raw_h, "n_read": n_read, "frame": frame, "cols": cols,
           "want": want}
    c.close()
    cb.free_metadata(mr)


def _recon(ctx, case, **opt):
    mr = case["mr"]
    ctx.set_options(mu_water=MU_W, adaptive_filter=1, bp_tile=0, **opt)
    out = torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.reconstruct_dev(mr, case["raw"], case["n_read"], out)
    ctx.synchronize()
    return out.cpu().numpy().reshape(-1, mr.rp.nx, mr.rp.ny)


def test_c1_end_to_end(ctx, c1_case):
    """All 7 blocks (adaptive KK filter, rebin, FFT32 filter, two-stream
    pipeline, finalize + HU) against the FP64 oracle on 656 voxel columns."""
    assert int(c1_case["mr"].ri.n_blocks) == 7
    got = _recon(ctx, c1_case, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    rms, mx = hu_err(got, c1_case["want"], c1_case["cols"], "c1 end to end (fft32)")
    assert rms <= 0.05 and mx <= 2.0


@pytest.mark.parametrize("mode,proj", [(cb.FILTER_F16X3, cb.PROJ_F32), (cb.FILTER_BF16X3, cb.PROJ_F32),
                                       (cb.FILTER_F16X3, cb.PROJ_BF16), (cb.FILTER_BF16X3, cb.PROJ_BF16),
                                       (cb.FILTER_FFT32, cb.PROJ_BF16), (cb.FILTER_F32, cb.PROJ_F32)])
def test_c2_precision_forms(ctx, c1_case, mode, proj):
    """Precision forms at C1 size: HU error per dtype variant against the
    FP64 oracle (precision forms' bar: RMS <= 1 HU)."""
    got = _recon(ctx, c1_case, filter_mode=mode, proj_dtype=proj)
    names = {cb.FILTER_F32: "f32", cb.FILTER_BF16X3: "bf16x3", cb.FILTER_F16X3: "f16x3", cb.FILTER_FFT32: "fft32"}
    label = f"c2 filter {names[mode]}, projections {'bf16' if proj == cb.PROJ_BF16 else 'f32'}"
    rms, mx = hu_err(got, c1_case["want"], c1_case["cols"], label)
    if proj == cb.PROJ_F32 and mode in (cb.FILTER_F32, cb.FILTER_F16X3):
        assert rms <= 0.05 and mx <= 2.0
    else:
        assert rms <= 1.0 and mx <= 10.0
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)


def test_c1_adaptive_filter_bit_exact(ctx, c1_case):
    """Adaptive filter (KK) of a whole C1 pull, bit for bit."""
    import oracle
    mr = c1_case["mr"]
    cb.plan_block(mr, 3)
    s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
    frame = c1_case["frame"]
    win = c1_case["raw"][s0 * frame:(s0 + n) * frame]
    out = torch.empty_like(win)
    ctx.adaptive_filter_dev(mr, win, out)
    ctx.synchronize()
    want, _ = oracle.adaptive_filter(mr, c1_case["raw_h"][s0:s0 + n])
    got = out.cpu().numpy().reshape(want.shape)
    changed = int(np.count_nonzero(want != c1_case["raw_h"][s0:s0 + n]))
    print(f"c1 adaptive (kk): {changed} of {want.size} samples smoothed, mismatches {np.count_nonzero(got != want)}")
    assert changed > 0
    assert np.array_equal(got, want)


def test_c1_rebin_filter_block(ctx, c1_case):
    """Rebin + FFT32 ramp filter of a whole C1 block (phi+z flying focal
    spot, 1934 views x 64 rows x 1472 channels) against the FP64 oracle's
    rebin and FFT filter of the same pull."""
    import oracle
    mr = c1_case["mr"]
    cb.plan_block(mr, 3)
    s = cb.block_sizes(mr)
    s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
    frame = c1_case["frame"]
    win = c1_case["raw"][s0 * frame:(s0 + n) * frame]
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    reb = torch.empty(s.rebin, dtype=torch.float32, device="cuda")
    ctx.rebin_filter_dev(mr, win, reb)
    ctx.synchronize()
    got = reb.cpu().numpy().reshape(s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)
    want = oracle.rebin(mr, c1_case["raw_h"][s0:s0 + n], double=True)
    want = oracle.filter_fft(want, cb.load_filter(mr))
    d = np.abs(got.astype(np.float64) - want)
    scale = np.abs(want).max()
    print(f"c1 rebin + fft32 filter block: max err {d.max() / scale:.3g}, rms {np.sqrt(np.mean(d ** 2)) / scale:.3g} "
          f"of max |value| {scale:.4g}")
    assert d.max() <= 1e-4 * scale


def test_c1_adaptive_fixed_threshold_bit_exact(ctx, c1_case):
    """Fixed-threshold mode (noise threshold 3e4 photons at 1e5 incident:
    T = ln(1e5 / 3e4) in line-integral units) on a whole C1 pull, bit for
    bit against the restatement's smoothing step with that threshold."""
    import oracle
    mr = c1_case["mr"]
    cb.plan_block(mr, 2)
    s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
    frame = c1_case["frame"]
    T = float(np.float32(np.log(1e5 / 3e4)))
    ctx.set_options(adaptive_threshold=T)
    win = c1_case["raw"][s0 * frame:(s0 + n) * frame]
    out = torch.empty_like(win)
    ctx.adaptive_filter_dev(mr, win, out)
    ctx.synchronize()
    ctx.set_options(adaptive_threshold=0.0)
    raw_w = c1_case["raw_h"][s0:s0 + n]
    want = oracle.adaptive_apply(mr, raw_w, np.full(n, np.float32(T), dtype=np.float32))
    got = out.cpu().numpy().reshape(want.shape)
    changed = int(np.count_nonzero(want != raw_w))
    print(f"c1 adaptive (fixed T {T:.4f}): {changed} of {want.size} samples smoothed, "
          f"mismatches {np.count_nonzero(got != want)}")
    assert changed > 0
    assert np.array_equal(got, want)


# --------------------------------------------------------------------------
# c3: batching form


def test_c3_batch_scans_vs_oracle(ctx, tmp_path):
    """Two seeded batch scans of different matrix, FOV and pitch (c3 scans 2
    and 30: 768^2 over 390 mm, 3 blocks; 512^2 over 426 mm, 4 blocks) through
    ctbb_reconstruct_batch_dev, each against the FP64 oracle."""
    items = []
    for i in (2, 30):
        tok = scans.c3_tokens(i)
        mr = _mr(tmp_path, tok, f"c3_{i}.prm")
        raw, n_read, frame = _scan_dev(ctx, mr, tok, 3)
        out = torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda")
        items.append((i, tok, mr, raw, n_read, frame, out))
    ctx.set_options(mu_water=MU_W, adaptive_filter=1, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32,
                    bp_tile=0)
    torch.cuda.synchronize()
    ctx.reconstruct_batch_dev([it[2] for it in items], [it[3] for it in items], [it[4] for it in items],
                              [it[6] for it in items])
    ctx.synchronize()
    for i, tok, mr, raw, n_read, frame, out in items:
        got = out.cpu().numpy().reshape(-1, mr.rp.nx, mr.rp.ny)
        cols = _columns(mr, 300, 30 + i)
        want = oracle_volume(mr, raw.cpu().numpy().reshape(n_read, frame), cols, "kk")
        rms, mx = hu_err(got, want, cols, f"c3 scan {i} ({tok['Nx']}^2, {int(mr.ri.n_blocks)} blocks) batched")
        assert rms <= 0.05 and mx <= 2.0
        cb.free_metadata(mr)


# --------------------------------------------------------------------------
# c4: 1024^2 through the host-streaming entry point


@pytest.mark.parametrize("form", ["plain", "precision"])
def test_c4_two_blocks_host_streaming(ctx, tmp_path, form):
    """The c4 shape through the host-streaming entry point (pulls uploaded
    through two device buffers on a copy stream), in the plain form and in
    the precision form c4 names (FP16 hi+lo matrix-core filter, bfloat16
    filtered projections, FP32 accumulation), against the FP64 oracle."""
    tok = scans.c4_tokens()
    # the first two blocks of the c4 scan: 40 raw slices of 0.5 mm -> 64 with
    # the block rounding, 2 mm thickening; enough rotations for their pulls
    tok["NSlices"] = 10
    tok["EndPos"] = tok["StartPos"] + 9 * 2.0
    tok["Readings"] = 20 * 2304
    mr = _mr(tmp_path, tok, "c4.prm")
    assert int(mr.ri.n_blocks) == 2 and int(mr.rp.nx) == 1024
    raw, n_read, frame = _scan_dev(ctx, mr, tok, 4)
    raw_h = raw.cpu().numpy().reshape(n_read, frame)
    del raw
    h_raw = torch.from_numpy(raw_h.ravel()).pin_memory()
    s = cb.block_sizes(mr)
    h_out = torch.empty(s.final_stack, dtype=torch.float32).pin_memory()
    precision = form == "precision"
    ctx.set_options(mu_water=MU_W, adaptive_filter=1, filter_mode=cb.FILTER_F16X3 if precision else cb.FILTER_FFT32,
                    proj_dtype=cb.PROJ_BF16 if precision else cb.PROJ_F32, bp_tile=0)
    ctx.reconstruct_host(mr, h_raw, n_read, h_out)
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32)
    got = h_out.numpy().reshape(-1, mr.rp.nx, mr.rp.ny)
    cols = _columns(mr, 400, 44)
    want = oracle_volume(mr, raw_h, cols, "kk")
    rms, mx = hu_err(got, want, cols, f"c4 two blocks (1024^2) host streaming, {form} form")
    if precision:
        assert rms <= 1.0 and mx <= 10.0
    else:
        assert rms <= 0.05 and mx <= 2.0
    cb.free_metadata(mr)
