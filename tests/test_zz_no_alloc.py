"""ctbb_backproject_dev_noalloc (the backprojection with its view table in a
fixed per-GPU device array, for callers that cannot allocate) against
ctbb_backproject_dev through the C-ABI: bit-identical images on two calls in
a row, the second on a smaller block (a smaller filtered sinogram and image
matrix, other table direction; the array keeps the first call's contents,
and the second call's view table replaces them); inputs beyond the array's
capacity are rejected before anything is launched."""
import numpy as np
import pytest

import ctbangbang_amd as cb
from ctbangbang_amd import scans

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _block(tmp_path, mini_dir, name, nrows, nx, table_dir, seed):
    tok = scans.mini_tokens("none", nrows=nrows, nx=nx, table_dir=table_dir)
    p = str(tmp_path / name)
    cb.write_prm(p, tok)
    mr = cb.metadata_from_prm(p, install_dir=mini_dir)
    cb.plan_block(mr, 0)
    s = cb.block_sizes(mr)
    rng = np.random.default_rng(seed)
    reb = rng.standard_normal((s.n_proj_out, mr.cg.n_rows, mr.cg.n_channels_oversampled)).astype(np.float32)
    return mr, s, torch.from_numpy(reb).cuda()


def test_backproject_noalloc_matches_backproject_dev(ctx, tmp_path, mini_dir):
    ctx.set_options(proj_dtype=cb.PROJ_F32, bp_tile=0)
    sizes = []
    for name, nrows, nx, td, seed in (("a.prm", 32, 64, "out", 1), ("b.prm", 16, 32, "in", 2)):
        mr, s, reb = _block(tmp_path, mini_dir, name, nrows, nx, td, seed)
        views = int(mr.cg.n_proj_turn // 2) * int(s.n_half_turns)
        sizes.append((int(s.rebin), int(s.image_block)))
        print(f"noalloc: rows {nrows} nx {nx} table {td}: {views} views, sinogram {s.rebin}, image {s.image_block}",
              flush=True)
        want = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
        got = torch.full((s.image_block,), float("nan"), dtype=torch.float32, device="cuda")
        torch.cuda.synchronize()
        ctx.backproject_dev(mr, reb, want)
        ctx.backproject_dev_noalloc(mr, reb, got)
        ctx.synchronize()
        w, g = want.cpu().numpy(), got.cpu().numpy()
        assert np.isfinite(w).all() and np.abs(w).max() > 0
        assert np.array_equal(g, w)
        cb.free_metadata(mr)
    assert sizes[1][0] < sizes[0][0] and sizes[1][1] < sizes[0][1]


def test_backproject_noalloc_rejects_beyond_capacity(ctx, tmp_path, mini_dir):
    mr, s, reb = _block(tmp_path, mini_dir, "c.prm", 8, 32, "out", 3)
    half = int(mr.cg.n_proj_turn // 2)
    n_ht = cb.NOALLOC_MAX_VIEWS // half + 1
    # a pull long enough for n_ht half turns: only the view count is checked,
    # the call returns before any launch
    mr.ri.n_proj_pull = (n_ht * half + 2 * int(mr.cg.add_projections)) * int(mr.ri.n_ffs)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda")
    with pytest.raises(cb.CTBBError) as e:
        ctx.backproject_dev_noalloc(mr, reb, img)
    assert e.value.code == 11
    cb.free_metadata(mr)
