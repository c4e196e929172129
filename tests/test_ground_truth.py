"""Ground truth for the geometry conventions: an off-centre analytic water
ellipsoid in air, projected in numpy independently of the library's scan
generator (tests/phantom_truth.py), reconstructed by the FP64 oracle with the
reference's GPU conventions (x flipped unless table_direction == -1, qhat
sign -table_direction, include/backproject.cuh:59-65,84).

Bars (phantom_truth.BARS): interior HU within 2 of water (RMS 3), exterior
within 3 HU of air, in-plane centroid within 0.5 mm of the ellipsoid centre
after mapping voxel indices to the world frame, z centroid of the final
slices within 0.5 mm (slice placement).  Measured: interior mean 0.02 HU,
RMS 0.35 HU; centroid off by 0.03 mm in x and y and 0.02-0.08 mm in z.

Planted mistakes in the oracle (dropping the x flip or the qhat sign) must
miss the bars: without the flip the object lands mirrored (centroid at -x),
without the sign the cone weighting reads the wrong rows (interior HU off by
tens).  Both conventions are no-ops for TableDir in (table_direction -1).
"""
import numpy as np
import pytest

import ctbangbang_amd as cb
from ctbangbang_amd import scans

import phantom_truth as PT


def _oracle_stack(mr, raw_all, conventions):
    import oracle
    taps = cb.load_filter(mr)
    n_read = raw_all.shape[0]
    blocks = []
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        s0, n = cb.pull_start(mr), int(mr.ri.n_proj_pull)
        win = np.zeros((n,) + raw_all.shape[1:], dtype=np.float32)
        lo, hi = max(s0, 0), min(s0 + n, n_read)
        win[lo - s0:hi - s0] = raw_all[lo:hi]
        reb = oracle.rebin_filter(mr, win, taps, double=True)
        blocks.append(oracle.backproject(mr, reb, conventions, double=True))
    return oracle.finalize(mr, np.concatenate(blocks), PT.MU_WATER, double=True)


@pytest.mark.parametrize("table_dir", ["out", "in"])
def test_oracle_reconstructs_the_phantom(tmp_path, mini_dir, table_dir):
    import oracle
    mr, e, raw = PT.make_case(cb, scans, mini_dir, tmp_path, table_dir)
    assert int(mr.cg.table_direction) == (1 if table_dir == "out" else -1)
    st = _oracle_stack(mr, raw, oracle.BP_GPU)
    res = PT.check_stack(mr, st, e, f"oracle {table_dir}")
    assert res["n_interior"] > 300
    assert PT.failures(res, e) == []
    cb.free_metadata(mr)


@pytest.mark.parametrize("plant,expect", [("no_flip", "centroid_xy"), ("no_qsign", "interior_rms")])
def test_planted_convention_mistakes_miss_the_bars(tmp_path, mini_dir, plant, expect):
    import oracle
    mr, e, raw = PT.make_case(cb, scans, mini_dir, tmp_path, "out")
    flag = {"no_flip": oracle.BP_PLANT_NO_FLIP, "no_qsign": oracle.BP_PLANT_NO_QSIGN}[plant]
    st = _oracle_stack(mr, raw, oracle.BP_GPU | flag)
    res = PT.check_stack(mr, st, e, f"oracle out, planted {plant}")
    assert expect in PT.failures(res, e)
    cb.free_metadata(mr)


@pytest.mark.gpu
@pytest.mark.parametrize("table_dir", ["out", "in"])
def test_hip_reconstructs_the_phantom(ctx, tmp_path, mini_dir, table_dir):
    """The HIP path end to end (rebin, FFT32 filter, backprojection,
    finalize + HU) on the same analytic scan: the same ground-truth bars,
    and within 0.05 HU RMS of the FP64 oracle's stack."""
    import oracle
    torch = pytest.importorskip("torch")
    mr, e, raw = PT.make_case(cb, scans, mini_dir, tmp_path, table_dir)
    n_read = raw.shape[0]
    ctx.set_options(mu_water=PT.MU_WATER, adaptive_filter=0, filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32,
                    bp_tile=0)
    s = cb.block_sizes(mr)
    out = torch.empty(s.final_stack, dtype=torch.float32, device="cuda")
    raw_d = torch.from_numpy(raw.ravel()).cuda()
    torch.cuda.synchronize()
    ctx.reconstruct_dev(mr, raw_d, n_read, out)
    ctx.synchronize()
    got = out.cpu().numpy().reshape(-1, mr.rp.nx, mr.rp.ny)
    res = PT.check_stack(mr, got, e, f"hip {table_dir}")
    assert PT.failures(res, e) == []
    want = _oracle_stack(mr, raw, oracle.BP_GPU)
    d = got.astype(np.float64) - want
    print(f"hip vs oracle ({table_dir}): HU rms {np.sqrt(np.mean(d ** 2)):.4f} max {np.abs(d).max():.4f}")
    assert np.sqrt(np.mean(d ** 2)) <= 0.05
    cb.free_metadata(mr)
