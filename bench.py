"""Benchmark: whole-volume reconstruction of the phi+z flying-focal-spot scan
(736 ch x 32->64 rows, 10 rotations, pitch 0.8, 512^2 x 224 raw slices in 7
blocks -> 40 slices of 3 mm in HU) with the scan resident in HBM.

One step = one complete reconstruction: per block adaptive filter, rebin,
ramp filter and backprojection, then reorder + thickening + HU.
Prints one JSON line (see keys below).  --form plain computes in FP32
throughout (ramp filter as an FP32 FFT convolution by default); --form mixed
selects the precision form (FP16 hi+lo MFMA Toeplitz filter, BF16 filtered
projections, FP32 accumulation).
"""
import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BP_BYTES_PER_UPDATE = 16.0   # 4 FP32 bilinear taps per voxel-update


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c3", "c4"],
                    help="c1 (default): one phi+z FFS scan; c3: the batching workload (many scans)")
    ap.add_argument("--c3-scans", type=int, default=32, help="number of scans in the c3 batch")
    ap.add_argument("--form", default="plain", choices=["plain", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-ingest", action="store_true",
                    help="skip the PCIe-inclusive run (scan in page-locked host memory, final stack to host)")
    ap.add_argument("--cpu-columns", type=int, default=1024, help="voxel columns in the CPU baseline sample")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "bp_traffic_c1.json"))
    ap.add_argument("--stage", default="all", choices=["all", "backproject", "filter", "hu16", "copy"],
                    help="all: whole volume (the benchmark); backproject / filter: only the backprojection / the "
                         "ramp filter (--filter form) of one block (for counter runs); hu16: the int16 output "
                         "conversion of the final stack; copy: a 1 GiB device copy (HBM counter calibration)")
    ap.add_argument("--bp-tile", type=int, default=16, choices=[16, 32])
    ap.add_argument("--filter", default=None, choices=["fft32", "f32", "bf16x3", "f16x3"],
                    help="ramp filter form (default: fft32 for --form plain, f16x3 for --form mixed)")
    return ap.parse_args()


def stage_copy(args):
    """1 GiB device-to-device copy by an elementwise kernel: the reference
    point that scales FETCH_SIZE / WRITE_SIZE to bytes."""
    import torch
    n = (1 << 30) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(args.warmup):
        b.copy_(a)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.copy_(a)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"stage": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "ms": dt * 1e3,
                      "GB_per_s": 8 * n / dt / 1e9}), flush=True)


def stage_backproject(args, cb, scans, mr, ctx, raw, frame):
    """Backprojection of the middle block only, K launches."""
    import torch
    cb.plan_block(mr, int(mr.ri.n_blocks) // 2)
    s = cb.block_sizes(mr)
    s0, n = int(mr.ri.idx_pull_start), int(mr.ri.n_proj_pull)
    win = raw[s0 * frame:(s0 + n) * frame]
    elem = 2 if ctx.options().proj_dtype == cb.PROJ_BF16 else 4
    reb = torch.empty(s.rebin * elem // 4 + 1, dtype=torch.float32, device="cuda:0")
    ctx.rebin_filter_dev(mr, win, reb)
    img = torch.empty(s.image_block, dtype=torch.float32, device="cuda:0")
    for _ in range(args.warmup):
        ctx.backproject_dev(mr, reb, img)
    ctx.synchronize()
    total, launches = 0.0, 0
    for _ in range(args.steps):
        ctx.backproject_dev(mr, reb, img)
        ms, k = ctx.kernel_time("backproject")
        total += ms
        launches += k
    avg = total / launches / 1e3
    print(json.dumps({"stage": "backproject", "voxel_updates": s.voxel_updates_block, "avg_launch_ms": avg * 1e3,
                      "updates_per_s": s.voxel_updates_block / avg,
                      "gather_GB_per_s": BP_BYTES_PER_UPDATE * s.voxel_updates_block / avg / 1e9,
                      "bp_tile": args.bp_tile, "form": args.form}), flush=True)


def stage_filter(args, cb, mr, ctx, raw, frame, filt):
    """Ramp filter of the middle block's rebinned rows only, K launches: the
    FFT form or a Toeplitz product on the matrix cores (--filter)."""
    import torch
    cb.plan_block(mr, int(mr.ri.n_blocks) // 2)
    s = cb.block_sizes(mr)
    s0, n = int(mr.ri.idx_pull_start), int(mr.ri.n_proj_pull)
    win = raw[s0 * frame:(s0 + n) * frame]
    reb = torch.empty(s.rebin, dtype=torch.float32, device="cuda:0")
    ctx.rebin_dev(mr, win, reb)
    nos = int(mr.cg.n_channels_oversampled)
    rows = s.rebin // nos
    taps = torch.from_numpy(cb.load_filter(mr)).to("cuda:0")
    out = torch.empty_like(reb)
    for _ in range(args.warmup):
        ctx.filter_rows_dev(taps, nos, rows, reb, out)
    ctx.synchronize()
    total = 0.0
    for _ in range(args.steps):
        ctx.filter_rows_dev(taps, nos, rows, reb, out)
        total += ctx.kernel_time("filter")[0]
    avg_ms = total / args.steps
    line = {"stage": "filter", "filter": filt, "rows": rows, "n": nos, "avg_launch_ms": avg_ms,
            "GB_per_s_in_out": 2.0 * 4 * rows * nos / avg_ms / 1e6}
    if filt != "fft32":
        products = 3 if filt in ("bf16x3", "f16x3") else 1
        line["matrix_TFLOP_per_s"] = products * 2.0 * rows * nos * nos / avg_ms / 1e9
    print(json.dumps(line), flush=True)


def stage_hu16(args, cb, mr, ctx):
    """Output conversion of the final stack (HU float32 -> int16 stored values
    for a DICOM series), K launches over a synthetic HU stack of the final
    stack's size: HBM-bound, 6 bytes per voxel."""
    import torch
    s = cb.block_sizes(mr)
    n = int(s.final_stack)
    g = torch.Generator(device="cuda:0").manual_seed(5)
    hu = torch.randn(n, generator=g, device="cuda:0") * 1000.0
    out = torch.empty(n, dtype=torch.int16, device="cuda:0")
    for _ in range(args.warmup):
        ctx.hu_to_int16_dev(hu, n, out)
    ctx.synchronize()
    total = 0.0
    for _ in range(args.steps):
        ctx.hu_to_int16_dev(hu, n, out)
        total += ctx.kernel_time("hu16")[0]
    avg_ms = total / args.steps
    gbs = 6.0 * n / avg_ms / 1e6
    print(json.dumps({"stage": "hu16", "voxels": n, "avg_launch_ms": avg_ms, "GB_per_s": gbs,
                      "hbm_frac": gbs / 8000.0}), flush=True)


def bench_c3(args):
    """Batching form: --c3-scans synthetic scans of uneven length, pitch, FOV
    and matrix size, reconstructed (a) one reconstruct_dev call per scan and
    (b) as one reconstruct_batch_dev pipeline; value = voxel-updates/s of (b)."""
    import numpy as np
    import torch

    import ctbangbang_amd as cb
    from ctbangbang_amd import scans
    d = tempfile.mkdtemp()
    ctx = cb.Context(0)
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32, mu_water=0.02, adaptive_filter=1,
                    bp_tile=args.bp_tile)
    mrs, raws, nreads, finals, vu = [], [], [], [], 0.0
    for i in range(args.c3_scans):
        tok = scans.c3_tokens(i)
        prm = os.path.join(d, f"c3_{i}.prm")
        cb.write_prm(prm, tok)
        mr = cb.metadata_from_prm(prm)
        for b in range(mr.ri.n_blocks):
            cb.plan_block(mr, b)
            vu += cb.block_sizes(mr).voxel_updates_block
        n_read = int(tok["Readings"])
        frame = mr.cg.n_rows_raw * mr.cg.n_channels
        raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda:0")
        ctx.synth_scan_dev(mr, scans.phantom_for(tok, 3), 0, n_read, raw)
        mrs.append(mr)
        raws.append(raw)
        nreads.append(n_read)
        finals.append(torch.empty(cb.block_sizes(mr).final_stack, dtype=torch.float32, device="cuda:0"))
    ctx.synchronize()
    # warm-up (allocations, tables) with the batch call
    ctx.reconstruct_batch_dev(mrs, raws, nreads, finals)
    ctx.synchronize()
    t0 = time.perf_counter()
    for mr, raw, n, out in zip(mrs, raws, nreads, finals):
        ctx.reconstruct_dev(mr, raw, n, out)
    ctx.synchronize()
    t_loop = time.perf_counter() - t0
    loop_out = [finals[0].cpu().numpy().copy(), finals[-1].cpu().numpy().copy()]
    t0 = time.perf_counter()
    ctx.reconstruct_batch_dev(mrs, raws, nreads, finals)
    ctx.synchronize()
    t_batch = time.perf_counter() - t0
    same = bool(np.array_equal(finals[0].cpu().numpy(), loop_out[0]) and
                np.array_equal(finals[-1].cpu().numpy(), loop_out[1]))
    n = args.c3_scans
    line = {
        "metric": "voxel-updates/s (batch of scans, one pipeline, scans resident in HBM)",
        "value": vu / t_batch, "unit": "voxel-updates/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded scans: pitch 0.5-1.5, FOV 200-500 mm, 512^2/768^2, 50-400 slices)",
        "config": {"workload": f"c3: {n} phi+z FFS scans, {vu:.4g} voxel-updates", "parallelism": "1 GPU"},
        "scans_per_s": n / t_batch,
        "loop": {"s": t_loop, "voxel_updates_per_s": vu / t_loop, "scans_per_s": n / t_loop},
        "batch_over_loop": t_loop / t_batch,
        "batch_equals_loop": same,
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    for mr in mrs:
        cb.free_metadata(mr)
    ctx.close()


def cpu_baseline(mr, raw_block, taps, n_cols):
    """FP32 CPU restatement (oracle/liboracle32.so, one thread) on one block:
    rebin + FFT ramp filter of the whole block and the backprojection of
    n_cols sampled voxel columns (all 32 slices), extrapolated to the block."""
    os.environ.setdefault("OMP_NUM_THREADS", "1")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np

    import ctbangbang_amd as cb
    import oracle
    s = cb.block_sizes(mr)
    t0 = time.perf_counter()
    reb = oracle.rebin(mr, raw_block, double=False)
    reb = oracle.filter_fft(reb, taps).astype(np.float32)
    t1 = time.perf_counter()
    rng = np.random.default_rng(0)
    idx = rng.choice(mr.rp.nx * mr.rp.ny, size=n_cols, replace=False)
    cx, cy = (idx // mr.rp.ny).astype(np.int32), (idx % mr.rp.ny).astype(np.int32)
    oracle.backproject(mr, reb, oracle.BP_GPU, cols=(cx, cy), double=False)
    t2 = time.perf_counter()
    t_block = (t1 - t0) + (t2 - t1) * (mr.rp.nx * mr.rp.ny) / n_cols
    return {
        "value": s.voxel_updates_block / t_block,
        "unit": "voxel-updates/s",
        "cores": 1,
        "kind": "port",
        "sample": (f"one block of the workload: FP32 rebin + FFT ramp filter of the full block "
                   f"({t1 - t0:.1f} s) and backprojection of {n_cols} of {mr.rp.nx * mr.rp.ny} voxel columns "
                   f"x 32 slices ({t2 - t1:.1f} s), extrapolated to the block"),
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(blocks shard with no exchange across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    if args.stage == "copy":
        stage_copy(args)
        return
    if args.config == "c3":
        bench_c3(args)
        return
    import torch

    import ctbangbang_amd as cb
    from ctbangbang_amd import scans

    tok = getattr(scans, args.config + "_tokens")()
    d = tempfile.mkdtemp()
    prm = os.path.join(d, "bench.prm")
    cb.write_prm(prm, tok)
    mr = cb.metadata_from_prm(prm)
    ctx = cb.Context(0)
    mixed = args.form == "mixed"
    modes = {"fft32": cb.FILTER_FFT32, "f32": cb.FILTER_F32, "bf16x3": cb.FILTER_BF16X3, "f16x3": cb.FILTER_F16X3}
    filt = args.filter or ("f16x3" if mixed else "fft32")
    ctx.set_options(filter_mode=modes[filt],
                    proj_dtype=cb.PROJ_BF16 if mixed else cb.PROJ_F32,
                    mu_water=float(tok.get("MuWater", 0.0)),
                    adaptive_filter=1 if float(tok["AdaptiveFiltration"]) > 0 else 0,
                    bp_tile=args.bp_tile)
    if args.stage == "hu16":
        stage_hu16(args, cb, mr, ctx)
        ctx.close()
        return
    n_read = int(tok["Readings"])
    frame = mr.cg.n_rows_raw * mr.cg.n_channels
    raw = torch.empty(n_read * frame, dtype=torch.float32, device="cuda:0")
    ctx.synth_scan_dev(mr, scans.phantom_for(tok, {"c0": 0, "c1": 1, "c4": 4}[args.config]), 0, n_read, raw)
    if args.stage == "backproject":
        stage_backproject(args, cb, scans, mr, ctx, raw, frame)
        ctx.close()
        return
    if args.stage == "filter":
        stage_filter(args, cb, mr, ctx, raw, frame, filt)
        ctx.close()
        return
    s = cb.block_sizes(mr)
    final = torch.empty(s.final_stack, dtype=torch.float32, device="cuda:0")
    image = torch.empty(s.image_stack, dtype=torch.float32, device="cuda:0")
    vu_blocks = []
    for b in range(mr.ri.n_blocks):
        cb.plan_block(mr, b)
        vu_blocks.append(cb.block_sizes(mr).voxel_updates_block)
    vu_total = float(sum(vu_blocks))

    for _ in range(args.warmup):
        ctx.reconstruct_dev(mr, raw, n_read, final, image)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.reconstruct_dev(mr, raw, n_read, final, image)
    ctx.synchronize()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps

    # per-kernel device time of the last step (HIP events on the ctx stream)
    fam = {}
    for f in ("adaptive", "rebin", "filter", "backproject", "finalize"):
        ms, n = ctx.kernel_time(f)
        fam[f] = {"ms_total": round(ms, 3), "launches": n}
    bp_ms, bp_n = ctx.kernel_time("backproject")
    bp_avg_s = bp_ms / max(bp_n, 1) / 1e3
    vu_launch = vu_blocks[-1]
    achieved = BP_BYTES_PER_UPDATE * vu_launch / bp_avg_s / 1e9
    traffic = None
    if os.path.exists(args.traffic_file):
        try:
            traffic = json.load(open(args.traffic_file)).get("bytes_per_launch")
        except Exception:
            traffic = None

    line = {
        "metric": "voxel-updates/s (end-to-end, whole volume, scan resident in HBM)",
        "value": vu_total / dt,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "s_per_volume": dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32" if mixed else "f32",
        "data": "synthetic (analytic Shepp-Logan-type phantom + 8 seeded ellipsoids, Poisson noise I0=1e5)",
        "config": {"workload": f"{args.config}: {int(mr.rp.nx)}^2 x {int(mr.ri.n_slices_recon)} raw slices, "
                               f"{int(mr.ri.n_blocks)} blocks, n_ffs={int(mr.ri.n_ffs)}, "
                               f"{int(mr.cg.n_rows)} rows, {n_read} readings",
                   "form": args.form, "filter": filt, "voxel_updates": vu_total, "parallelism": "1 GPU"},
        "backproject_updates_per_s": vu_launch / bp_avg_s,
        "kernels_last_step": fam,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "bp_kernel (backprojection)",
                     "per_unit": "16 B per voxel-update (4 FP32 bilinear taps)",
                     "avg_launch_ms": bp_avg_s * 1e3},
    }
    if not args.no_host_ingest:
        # same reconstruction from a page-locked host copy of the scan to a
        # host final stack: uploads overlap the compute (reconstruct_host)
        h_raw = torch.empty(n_read * frame, dtype=torch.float32, pin_memory=True)
        h_raw.copy_(raw)
        h_fin = torch.empty(s.final_stack, dtype=torch.float32, pin_memory=True)
        ctx.reconstruct_host(mr, h_raw, n_read, h_fin)
        t0 = time.perf_counter()
        n_h = max(1, min(args.steps, 3))
        for _ in range(n_h):
            ctx.reconstruct_host(mr, h_raw, n_read, h_fin)
        dt_h = (time.perf_counter() - t0) / n_h
        line["host_ingest"] = {"s_per_volume": dt_h, "voxel_updates_per_s": vu_total / dt_h, "runs": n_h,
                               "note": "scan in page-locked host memory to final stack in host memory "
                                       "(H2D of each pull overlapped with the previous block)"}
        del h_raw, h_fin
    if not args.no_cpu_baseline:
        cb.plan_block(mr, 0)
        s0, n = int(mr.ri.idx_pull_start), int(mr.ri.n_proj_pull)
        rb = raw[s0 * frame:(s0 + n) * frame].cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(mr, rb, cb.load_filter(mr), args.cpu_columns)
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
