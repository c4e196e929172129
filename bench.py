"""Benchmark: whole-volume reconstruction of the phi+z flying-focal-spot scan
(736 ch x 32->64 rows, 10 rotations, pitch 0.8, 512^2 x 224 raw slices in 7
blocks -> 40 slices of 3 mm in HU) with the scan resident in HBM.

One step = one complete reconstruction: per block adaptive filter, rebin,
ramp filter and backprojection, then reorder + thickening + HU.
Prints one JSON line.  `value` is voxel-updates/s with the scan resident in
HBM (plain FP32 form); `s_per_volume` is the host-to-host time (scan in
page-locked host memory to the final stack in host memory) and
`s_per_volume_hbm_resident` the time `value` is computed from.  The same
line carries the precision form (`forms.mixed`) and the batching form
(`forms.batched`, several seeded scans of uneven size through one
pipeline against one call per scan), measured in the same run.
--form mixed makes the precision form the headline instead.
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
FP32_PEAK_TFLOPS = 157.3     # MI355X published FP32 vector peak (packed; non-packed issue: half)
BP_BYTES_PER_UPDATE = 16.0   # 4 FP32 bilinear taps per voxel-update (served from LDS)
BP_FLOP_PER_UPDATE = 12.0    # q 2, row coordinate 2, bilinear 6, accumulate 2
N_SIMD = 1024                # 256 CUs x 4 SIMDs
CLOCK_HZ = 2.4e9             # peak engine clock


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="c1", choices=["c0", "c1", "c3", "c4"],
                    help="c1 (default): one phi+z FFS scan; c3: the batching workload (many scans)")
    ap.add_argument("--c3-scans", type=int, default=32, help="number of scans in the c3 batch")
    ap.add_argument("--c3-chunk", type=int, default=0,
                    help="c3 scans per reconstruct_batch_dev call (0: all in one; their readings must fit in HBM)")
    ap.add_argument("--form", default="plain", choices=["plain", "mixed"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-host-ingest", action="store_true",
                    help="skip the PCIe-inclusive run (scan in page-locked host memory, final stack to host)")
    ap.add_argument("--cpu-patch", type=int, default=64,
                    help="CPU baseline: side of each square sample of voxel columns (one sample per thread)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="CPU baseline threads (0: the usable cores, at most 16)")
    ap.add_argument("--no-forms", action="store_true", help="skip the mixed and batched figures of the default line")
    ap.add_argument("--line-scans", type=int, default=8, help="scans in the batched figure of the default line")
    ap.add_argument("--counters-file", default=os.path.join(ROOT, "profiles", "bp_kernel_c1.json"))
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


def batch_run(n_scans, bp_tile, first=0):
    """Batching form on n_scans seeded scans (c3 scans first .. first +
    n_scans - 1) of uneven length, pitch, FOV and matrix size: one
    reconstruct_dev call per scan (loop) against one reconstruct_batch_dev
    pipeline (batch), both after a warm-up batch.  Returns the timings,
    voxel-updates and whether both give the same bits."""
    import numpy as np
    import torch

    import ctbangbang_amd as cb
    from ctbangbang_amd import scans
    d = tempfile.mkdtemp()
    ctx = cb.Context(0)
    ctx.set_options(filter_mode=cb.FILTER_FFT32, proj_dtype=cb.PROJ_F32, mu_water=0.02, adaptive_filter=1,
                    bp_tile=bp_tile)
    mrs, raws, nreads, finals, vu = [], [], [], [], 0.0
    for i in range(first, first + n_scans):
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
    for mode in (cb.BATCH_PIPELINE, cb.BATCH_PERSISTENT):         # warm-up (allocations, tables)
        ctx.set_options(batch_mode=mode)
        ctx.reconstruct_batch_dev(mrs, raws, nreads, finals)
        ctx.synchronize()
    t0 = time.perf_counter()
    for mr, raw, n, out in zip(mrs, raws, nreads, finals):
        ctx.reconstruct_dev(mr, raw, n, out)
    ctx.synchronize()
    t_loop = time.perf_counter() - t0
    loop_out = [f.cpu().numpy().copy() for f in finals]
    t_modes, same = {}, True
    for mode in (cb.BATCH_PIPELINE, cb.BATCH_PERSISTENT):
        ctx.set_options(batch_mode=mode)
        t0 = time.perf_counter()
        ctx.reconstruct_batch_dev(mrs, raws, nreads, finals)
        ctx.synchronize()
        t_modes[mode] = time.perf_counter() - t0
        same = same and all(np.array_equal(f.cpu().numpy(), w) for f, w in zip(finals, loop_out))
    for mr in mrs:
        cb.free_metadata(mr)
    ctx.close()
    del raws, finals
    torch.cuda.empty_cache()
    return {"scans": n_scans, "voxel_updates": vu, "s_batch": t_modes[cb.BATCH_PERSISTENT],
            "s_pipeline": t_modes[cb.BATCH_PIPELINE], "s_loop": t_loop, "same_bits": bool(same)}


def bench_c3(args):
    """Batching workload alone: value = voxel-updates/s of the batch pipeline.
    With --c3-chunk N the scans run as consecutive batches of N (each one
    reconstruct_batch_dev call, its scans resident in HBM: the 32 scans'
    readings alone take 273 GB), and the times add up."""
    chunk = args.c3_chunk if args.c3_chunk > 0 else args.c3_scans
    r = {"scans": 0, "voxel_updates": 0.0, "s_batch": 0.0, "s_pipeline": 0.0, "s_loop": 0.0, "same_bits": True}
    for first in range(0, args.c3_scans, chunk):
        rc = batch_run(min(chunk, args.c3_scans - first), args.bp_tile, first)
        print(f"c3 scans {first}..{first + rc['scans'] - 1}: batch {rc['s_batch']:.3f} s, pipeline "
              f"{rc['s_pipeline']:.3f} s, loop {rc['s_loop']:.3f} s, same bits {rc['same_bits']}",
              file=sys.stderr, flush=True)
        for k in ("scans", "voxel_updates", "s_batch", "s_pipeline", "s_loop"):
            r[k] += rc[k]
        r["same_bits"] = r["same_bits"] and rc["same_bits"]
    n, vu, t_batch, t_loop = r["scans"], r["voxel_updates"], r["s_batch"], r["s_loop"]
    line = {
        "metric": "voxel-updates/s (batch of scans, one pipeline, scans resident in HBM)",
        "value": vu / t_batch, "unit": "voxel-updates/s", "n_gpus": 1, "steps": 1, "warmup": 1,
        "ms_per_step": t_batch * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded scans: pitch 0.5-1.5, FOV 200-500 mm, 512^2/768^2, 50-400 slices)",
        "config": {"workload": f"c3: {n} phi+z FFS scans, {vu:.4g} voxel-updates"
                               + (f", in batches of {chunk}" if chunk < n else ""), "parallelism": "1 GPU"},
        "scans_per_s": n / t_batch,
        "loop": {"s": t_loop, "voxel_updates_per_s": vu / t_loop, "scans_per_s": n / t_loop},
        "pipeline": {"s": r["s_pipeline"], "voxel_updates_per_s": vu / r["s_pipeline"]},
        "batch_over_loop": t_loop / t_batch,
        "batch_equals_loop": r["same_bits"],
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def _cpu_info():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"nproc": os.cpu_count(), "usable": usable, "model": model}


class _QuietStdout:
    """Send file descriptor 1 to /dev/null (the reference's backproject_cpu
    prints a progress line per angle); the JSON line is printed after."""

    def __enter__(self):
        sys.stdout.flush()
        self.saved = os.dup(1)
        nul = os.open(os.devnull, os.O_WRONLY)
        os.dup2(nul, 1)
        os.close(nul)

    def __exit__(self, *exc):
        import ctypes
        ctypes.CDLL(None).fflush(None)      # C stdio buffers go to /dev/null too
        os.dup2(self.saved, 1)
        os.close(self.saved)


def cpu_baseline(mr, raw_block, taps, patch, threads):
    """CPU path timed on this host.

    Backprojection: the reference's own backproject_cpu (src/backproject_cpu.cu,
    compiled from source into oracle/_ref) on samples of patch x patch voxel
    columns x 32 slices of the block, each a sub-grid of the same voxel size
    at its own place in the field of view, one sample per thread (threads
    share nothing: each has its own copy of the metadata and output); once on
    one thread, once on `threads` threads.  Voxel-updates per second are per
    update, so they carry over to the whole block.
    Rebin + ramp filter: the oracle's FP32 restatement (oracle/liboracle32.so
    and the FP64 numpy FFT) of the whole block, one thread.
    Falls back to the restatement's backprojection where the reference build
    is absent (kind "port")."""
    import ctypes as C
    import threading

    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import ctbangbang_amd as cb
    import oracle
    info = _cpu_info()
    if threads <= 0:
        threads = max(1, min(16, info["usable"]))
    t0 = time.perf_counter()
    reb = oracle.rebin(mr, raw_block, double=False)
    reb = oracle.filter_fft(reb, taps).astype(np.float32)
    t_prep = time.perf_counter() - t0
    nx, ny = int(mr.rp.nx), int(mr.rp.ny)
    dxy = float(mr.rp.recon_fov) / nx
    updates_per_col = cb.block_sizes(mr).voxel_updates_block / (nx * ny)
    ref = oracle.ref_lib()
    # sample centres: a ring of places inside the field of view
    centres = []
    for k in range(max(threads, 1)):
        ang = 2.0 * np.pi * k / max(threads, 1)
        rad = 0.25 * float(mr.rp.recon_fov) * (0.5 + 0.5 * (k % 2))
        centres.append((float(mr.rp.x_origin) + rad * np.cos(ang), float(mr.rp.y_origin) + rad * np.sin(ang)))

    def sample_mr(cx, cy):
        from ctbangbang_amd.structs import ReconMetadata
        m = ReconMetadata.from_buffer_copy(mr)
        m.rp.nx = m.rp.ny = patch
        m.rp.recon_fov = dxy * patch
        m.rp.x_origin, m.rp.y_origin = cx, cy
        return m

    if ref is not None:
        f = getattr(ref, "_Z15backproject_cpuP14recon_metadata")
        f.argtypes = [C.c_void_p]
        f.restype = C.c_int
        kind = "reference"

        def run(k, out):
            m = sample_mr(*centres[k])
            nblk = int(m.ri.n_slices_block) * patch * patch
            img = np.zeros(max(int(m.ri.cb.block_idx), 1) * nblk, dtype=np.float32)
            m.ctd.rebin = reb.ctypes.data_as(C.POINTER(C.c_float))
            m.ctd.image = img.ctypes.data_as(C.POINTER(C.c_float))
            out[k] = f(C.addressof(m))
    else:
        kind = "port"

        def run(k, out):
            m = sample_mr(*centres[k])
            oracle.backproject(m, reb, oracle.BP_CPU, double=False)
            out[k] = 0

    res = [None] * threads
    with _QuietStdout():
        t0 = time.perf_counter()
        run(0, res)
        t_one = time.perf_counter() - t0
        ths = [threading.Thread(target=run, args=(k, res)) for k in range(threads)]
        t0 = time.perf_counter()
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        t_all = time.perf_counter() - t0
    assert all(r == 0 for r in res)
    upd = updates_per_col * patch * patch
    one_core = upd / t_one
    all_cores = threads * upd / t_all
    return {
        "value": all_cores,
        "unit": "voxel-updates/s",
        "cores": threads,
        "kind": kind,
        "sample": (f"backprojection: {'reference backproject_cpu built from source' if kind == 'reference' else 'oracle FP32 restatement'} "
                   f"on {threads} samples of {patch}x{patch} voxel columns x 32 slices of one block (same voxel size, "
                   f"{upd:.3g} voxel-updates each), one per thread, {t_all:.1f} s; 1 thread: {t_one:.1f} s"),
        "one_core": one_core,
        "rebin_filter_s_per_block": t_prep,
        "rebin_filter_note": "oracle FP32 rebin + FP64 FFT filter of the whole block, one thread "
                             "(the reference's rebin_filter_cpu needs FFTW3, absent here)",
        "host_to_host_block_s_all_cores": t_prep + cb.block_sizes(mr).voxel_updates_block / all_cores,
        "nproc": info["nproc"], "usable_cpus": info["usable"], "cpu_model": info["model"],
    }


def run_volume(ctx, mr, raw, n_read, final, image, steps, warmup):
    """Time `steps` whole reconstructions (scan resident in HBM) after
    `warmup`; returns seconds per volume."""
    import torch
    for _ in range(warmup):
        ctx.reconstruct_dev(mr, raw, n_read, final, image)
    ctx.synchronize()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        ctx.reconstruct_dev(mr, raw, n_read, final, image)
    ctx.synchronize()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the figure is one GPU's throughput; blocks would shard across GPUs with no exchange)",
              file=sys.stderr)
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
    forms = {"plain": {"filter": "fft32", "proj": cb.PROJ_F32},
             "mixed": {"filter": "f16x3", "proj": cb.PROJ_BF16}}
    filt = args.filter or forms[args.form]["filter"]
    base_opts = dict(mu_water=float(tok.get("MuWater", 0.0)),
                     adaptive_filter=1 if float(tok["AdaptiveFiltration"]) > 0 else 0, bp_tile=args.bp_tile)
    ctx.set_options(filter_mode=modes[filt], proj_dtype=forms[args.form]["proj"], **base_opts)
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

    dt = run_volume(ctx, mr, raw, n_read, final, image, args.steps, args.warmup)

    # per-kernel device time of the last step (HIP events on the streams the
    # kernels run on)
    fam = {}
    for f in ("adaptive", "rebin", "filter", "backproject", "finalize"):
        ms, n = ctx.kernel_time(f)
        fam[f] = {"ms_total": round(ms, 3), "launches": n}
    bp_ms, bp_n = ctx.kernel_time("backproject")
    bp_avg_s = bp_ms / max(bp_n, 1) / 1e3
    vu_launch = vu_total / len(vu_blocks)          # mean updates per backprojection launch
    counters = {}
    if os.path.exists(args.counters_file):
        try:
            counters = json.load(open(args.counters_file))
        except (OSError, ValueError):
            counters = {}
    traffic = counters.get("bytes_per_launch")
    achieved = BP_FLOP_PER_UPDATE * vu_launch / bp_avg_s / 1e12
    roof = {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
            "kernel": "bp_kernel (backprojection)",
            "per_unit": f"{BP_FLOP_PER_UPDATE:.0f} FP32 FLOP per voxel-update (q 2, row coordinate 2, bilinear 6, "
                        f"accumulate 2) x {vu_launch:.4g} updates per launch",
            "avg_launch_ms": bp_avg_s * 1e3}
    if counters.get("valu_insts_per_launch"):
        # issue slots: a wave64 VALU instruction occupies its SIMD for 4 cycles
        roof["valu_issue_frac"] = 4.0 * counters["valu_insts_per_launch"] / (N_SIMD * CLOCK_HZ * bp_avg_s)
        roof["valu_issue_note"] = ("SQ_INSTS_VALU x 4 cycles / (1024 SIMDs x 2.4 GHz x launch time): the roof that "
                                   "binds; counters from " + os.path.relpath(args.counters_file, ROOT))
    if traffic:
        roof["hbm_GB_per_s"] = traffic / bp_avg_s / 1e9
        roof["hbm_frac"] = roof["hbm_GB_per_s"] / HBM_PEAK_GBS
    roof["lds_gather_GB_per_s"] = BP_BYTES_PER_UPDATE * vu_launch / bp_avg_s / 1e9
    roof["lds_gather_note"] = "16 B of bilinear taps per voxel-update, read from LDS (not HBM traffic)"

    line = {
        "metric": "voxel-updates/s (end-to-end, whole volume, scan resident in HBM)",
        "value": vu_total / dt,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3,
        "s_per_volume": None,
        "s_per_volume_hbm_resident": dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32" if mixed else "f32",
        "data": "synthetic (analytic Shepp-Logan-type phantom + 8 seeded ellipsoids, Poisson noise I0=1e5)",
        "config": {"workload": f"{args.config}: {int(mr.rp.nx)}^2 x {int(mr.ri.n_slices_recon)} raw slices, "
                               f"{int(mr.ri.n_blocks)} blocks, n_ffs={int(mr.ri.n_ffs)}, "
                               f"{int(mr.cg.n_rows)} rows, {n_read} readings, adaptive filter "
                               f"{'KK' if float(tok['AdaptiveFiltration']) > 0 else 'off'}",
                   "form": args.form, "filter": filt, "voxel_updates": vu_total, "parallelism": "1 GPU"},
        "backproject_updates_per_s": vu_launch / bp_avg_s,
        "kernels_last_step": fam,
        "roofline": roof,
    }
    if not args.no_host_ingest:
        # s/volume: the same reconstruction from a page-locked host copy of the
        # scan to a host final stack (uploads overlap the compute)
        h_raw = torch.empty(n_read * frame, dtype=torch.float32, pin_memory=True)
        h_raw.copy_(raw)
        h_fin = torch.empty(s.final_stack, dtype=torch.float32, pin_memory=True)
        ctx.reconstruct_host(mr, h_raw, n_read, h_fin)
        t0 = time.perf_counter()
        n_h = max(1, min(args.steps, 3))
        for _ in range(n_h):
            ctx.reconstruct_host(mr, h_raw, n_read, h_fin)
        dt_h = (time.perf_counter() - t0) / n_h
        line["s_per_volume"] = dt_h
        line["host_to_host"] = {"s_per_volume": dt_h, "voxel_updates_per_s": vu_total / dt_h, "runs": n_h,
                                "note": "scan in page-locked host memory to final stack in host memory "
                                        "(H2D of each pull overlapped with the previous block)"}
        del h_raw, h_fin
    if not args.no_forms and args.config == "c1":
        other = "plain" if mixed else "mixed"
        ctx.set_options(filter_mode=modes[forms[other]["filter"]], proj_dtype=forms[other]["proj"], **base_opts)
        dt_o = run_volume(ctx, mr, raw, n_read, final, image, max(1, min(args.steps, 5)), 1)
        ctx.set_options(filter_mode=modes[filt], proj_dtype=forms[args.form]["proj"], **base_opts)
        f_this = {"s_per_volume_hbm_resident": dt, "voxel_updates_per_s": vu_total / dt,
                  "filter": filt, "projections": "bf16" if mixed else "f32"}
        f_other = {"s_per_volume_hbm_resident": dt_o, "voxel_updates_per_s": vu_total / dt_o,
                   "filter": forms[other]["filter"], "projections": "bf16" if other == "mixed" else "f32"}
        plain_f, mixed_f = (f_other, f_this) if mixed else (f_this, f_other)
        line["forms"] = {"plain": plain_f, "mixed": mixed_f,
                         "mixed_over_plain": plain_f["s_per_volume_hbm_resident"] / mixed_f["s_per_volume_hbm_resident"]}
    if not args.no_cpu_baseline:
        cb.plan_block(mr, int(mr.ri.n_blocks) // 2)
        s0, n = int(mr.ri.idx_pull_start), int(mr.ri.n_proj_pull)
        rb = raw[s0 * frame:(s0 + n) * frame].cpu().numpy()
        line["cpu_baseline"] = cpu_baseline(mr, rb, cb.load_filter(mr), args.cpu_patch, args.cpu_threads)
    else:
        line["cpu_baseline"] = None
    ctx.close()
    del raw, final, image
    torch.cuda.empty_cache()
    if not args.no_forms and args.config == "c1" and args.line_scans > 0:
        r = batch_run(args.line_scans, args.bp_tile)
        line.setdefault("forms", {})["batched"] = {
            "scans": r["scans"], "voxel_updates": r["voxel_updates"],
            "batch_s": r["s_batch"], "pipeline_s": r["s_pipeline"], "loop_s": r["s_loop"],
            "batch_voxel_updates_per_s": r["voxel_updates"] / r["s_batch"],
            "loop_voxel_updates_per_s": r["voxel_updates"] / r["s_loop"],
            "batch_over_loop": r["s_loop"] / r["s_batch"], "same_bits": r["same_bits"],
            "workload": "the first scans of the c3 batch (seeds 3000+i): uneven rotations, pitch, FOV, matrix",
            "form": "batch: scans grouped per kernel variant, captured preparation graph + one persistent "
                    "backprojection launch per group; pipeline: one launch per block; loop: one call per scan"}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
