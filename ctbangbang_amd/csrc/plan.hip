// Work list of the batching form, built on the device: for every (scan,
// block) the reference's block plan (src/setup.cu:605-679 update_block_info:
// slice range, table-position search, pull range rounded to the focal-spot
// count and to 128 readings; src/backproject.cu:37-39: tube start angle and
// half turns), the backprojection view table of the block, the job
// descriptor the persistent backprojection reads, and the per-XCD queues
// of (block, tile) items.
//
// The arithmetic restates the host planner (setup.cpp) operation for
// operation, with contraction off, so the device plans equal the host ones
// bit for bit (the host plans size the preparation launches).
#include <hip/hip_runtime.h>

#include "kernels.h"

#pragma clang fp contract(off)

namespace ctbb {

namespace {

// reference src/setup.cu:806-825 (bounds tested before the element): the
// first idx in [0, n) at which the search stops, n if it never does.
__device__ __forceinline__ bool search_stops(float key, double v, int search_type) {
    switch (search_type) {
        case -1: return !(key < v);
        case 0: return !(key != v);
        default: return !(key > v);
    }
}

__device__ __forceinline__ int round_up_d(int v, int m) { return (v - 1) + (m - (v - 1) % m); }

// One workgroup per (scan, block) (grid.y = scan, grid.x = block): the
// block plan and its job's per-block fields.  The two table-position
// searches run over the whole workgroup (each thread takes every 256th
// entry; the smallest stopping index wins), so a long scan costs a few
// strided passes instead of a dependent chain of loads.
__global__ __launch_bounds__(256) void plan_blocks_kernel(const ScanPlanIn *scans, int n_scans, BlockPlan *plans,
                                                          BpArgs *jobs) {
    const int s = blockIdx.y, b = blockIdx.x;
    if (s >= n_scans) return;
    const ScanPlanIn &in = scans[s];
    if (b >= in.n_blocks) return;
    // setup.cu:611-617
    const float block_slice_start = in.recon_start_pos + in.recon_direction * b * in.coll_slicewidth *
                                                             static_cast<float>(in.n_slices_block);
    const float block_slice_end = block_slice_start + static_cast<float>(in.recon_direction) *
                                                          (static_cast<float>(in.n_slices_block) - 1.0f) *
                                                          in.coll_slicewidth;
    __shared__ int first_a, first_b;
    if (threadIdx.x == 0) {
        first_a = in.n_readings;
        first_b = in.n_readings;
    }
    __syncthreads();
    int fa = in.n_readings, fb = in.n_readings;
    for (int i = threadIdx.x; i < in.n_readings && (fa == in.n_readings || fb == in.n_readings); i += 256) {
        const double v = in.table_positions[i];
        if (fa == in.n_readings && search_stops(block_slice_start, v, in.array_direction)) fa = i;
        if (fb == in.n_readings && search_stops(block_slice_end, v, in.array_direction)) fb = i;
    }
    if (fa < in.n_readings) atomicMin(&first_a, fa);
    if (fb < in.n_readings) atomicMin(&first_b, fb);
    __syncthreads();
    if (threadIdx.x != 0) return;
    const int ia = first_a, ib = first_b;
    // setup.cu:637-661
    int st, en;
    if (ia > ib) {
        st = round_up_d(ib - in.ppb - in.add_ffs, in.n_ffs);
        en = round_up_d(ia + in.ppb + in.add_ffs, in.n_ffs);
    } else {
        st = round_up_d(ia - in.ppb - in.add_ffs, in.n_ffs);
        en = round_up_d(ib + in.ppb + in.add_ffs, in.n_ffs);
    }
    en += 256;
    const int n = round_up_d(en - st, 128);
    en = st + n;
    BlockPlan p;
    p.scan = s;
    p.block = b;
    p.slice_start = block_slice_start;
    p.slice_end = block_slice_end;
    p.idx_slice_start = ia;
    p.idx_slice_end = ib;
    p.pull_start = st;
    p.pull_end = en;
    p.n_pull = n;
    // backproject.cu:37 (float, pi' constant; the pull start as the
    // unsigned size_t of the reference's struct)
    const float t = in.tube_angle0 + (static_cast<float>(static_cast<size_t>(st) + in.add_projections_ffs) *
                                      360.0f / static_cast<float>(in.n_proj_ffs));
    p.tube_start = static_cast<float>(fmod(static_cast<double>(t), 360.0) * kPiBp / 180.0f);
    // backproject.cu:39
    p.n_ht = (n / in.n_ffs - 2 * in.add_projections) / (in.n_proj_turn / 2);
    plans[in.block_first + b] = p;
    BpArgs a = in.tmpl;
    a.n_ht = p.n_ht;
    a.out = in.image + static_cast<size_t>(b) * in.n_slices_block * static_cast<size_t>(a.nx) * a.ny;
    // the scan's filtered blocks lie slot_elems apart from its first one
    if (a.rebin_bf16) a.rebin_bf16 = in.tmpl.rebin_bf16 + static_cast<size_t>(b) * in.slot_elems;
    else a.rebin = in.tmpl.rebin + static_cast<size_t>(b) * in.slot_elems;
    jobs[in.block_first + b] = a;
}

// One thread: view-table offsets of all blocks (in block order), the jobs'
// table pointers, and the queues' item prefixes [8][n_jobs + 1] (band q of
// a job with n tiles is [floor(q n / 8), floor((q + 1) n / 8))).
__global__ void plan_lists_kernel(const BlockPlan *plans, const ScanPlanIn *scans, int n_blocks, BpView *vtab_base,
                                  size_t *vt_off, BpArgs *jobs, int *band_pre) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    size_t off = 0;
    for (int j = 0; j < n_blocks; j++) {
        const ScanPlanIn &in = scans[plans[j].scan];
        vt_off[j] = off;
        jobs[j].vtab = vtab_base + off;
        off += static_cast<size_t>(in.n_proj_turn / 2) * (plans[j].n_ht > 0 ? plans[j].n_ht : 0);
    }
    vt_off[n_blocks] = off;
    for (int q = 0; q < 8; q++) {
        int acc = 0;
        int *pre = band_pre + q * (n_blocks + 1);
        for (int j = 0; j < n_blocks; j++) {
            pre[j] = acc;
            const int n = jobs[j].n_tiles;
            acc += ((q + 1) * n) / 8 - (q * n) / 8;
        }
        pre[n_blocks] = acc;
    }
}

// View tables of all blocks from the device plans: grid.y = block; entry
// t = i * n_ht + k as view_table_kernel (bp.hip) computes it.
__global__ __launch_bounds__(256) void plan_view_tables_kernel(const BlockPlan *plans, const ScanPlanIn *scans,
                                                               const size_t *vt_off, BpView *vtab_base) {
    const BlockPlan p = plans[blockIdx.y];
    const ScanPlanIn &in = scans[p.scan];
    const int half = in.n_proj_turn / 2;
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (p.n_ht <= 0 || e >= half * p.n_ht) return;
    view_table_entry(vtab_base + vt_off[blockIdx.y] + e, e, static_cast<double>(p.tube_start), half, p.n_ht,
                     in.n_proj_turn, static_cast<double>(in.z_rot));
}

}  // namespace

int launch_plan_batch(const ScanPlanIn *d_scans, int n_scans, int n_blocks, int max_views_per_block,
                      BlockPlan *d_plans, BpArgs *d_jobs, BpView *d_vtab, size_t *d_vt_off, int *d_band_pre,
                      hipStream_t stream) {
    if (!d_scans || n_scans <= 0 || n_blocks <= 0 || !d_plans || !d_jobs || !d_vtab || !d_vt_off || !d_band_pre)
        return CTBB_ERR_ARG;
    // (n_blocks, the total over the scans, bounds each scan's block count)
    hipLaunchKernelGGL(plan_blocks_kernel, dim3(static_cast<unsigned>(n_blocks), static_cast<unsigned>(n_scans)),
                       dim3(256), 0, stream, d_scans, n_scans, d_plans, d_jobs);
    hipLaunchKernelGGL(plan_lists_kernel, dim3(1), dim3(64), 0, stream, d_plans, d_scans, n_blocks, d_vtab, d_vt_off,
                       d_jobs, d_band_pre);
    if (max_views_per_block > 0)
        hipLaunchKernelGGL(plan_view_tables_kernel,
                           dim3(static_cast<unsigned>((max_views_per_block + 255) / 256), static_cast<unsigned>(n_blocks)),
                           dim3(256), 0, stream, d_plans, d_scans, d_vt_off, d_vtab);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
