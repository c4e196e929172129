// Kernel launch interfaces (device pointers, HIP stream).  Each launcher
// validates its shapes on the host before launching and returns a
// ctbb_status.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "ctbb_internal.h"

namespace ctbb {

// ---------------------------------------------------------------------------
// Backprojection
struct BpArgs {
    const float *rebin;        // float32 [n_views][n_rows][nos]  (or null)
    const uint16_t *rebin_bf16;// bfloat16 [n_views][n_rows][nos] (or null)
    const BpView *vtab;        // [n_ht*half]
    float *out;                // [32][nx][ny]
    int nx, ny, nos, n_rows, half, n_ht;
    int tpr, tpr_log2;         // staging threads per detector row (set by the launcher)
    int asin_small_only;       // 1: every voxel has |p| <= 0.45 r_f (no asinf fallback needed)
    int nz;                    // slices per thread: 32 (one z chunk) or 16 (two)
    int tiles_y;               // ny / 16
    int n_tiles;
    float xs, xc, x0;          // x(xi) = xs * (xi - xc) + x0
    float ys, yc, y0;          // y(yi) = ys * (yi - yc) + y0
    float inv_ds, two_cc, inv_rf, rf2, k_asin, coll, tanc, neg_td, hr;
    float half_w;              // footprint half width of a 16x16 tile, channels
    float acq_half, out_scale;
};
// beside: the caller schedules other kernels to run beside this launch
// (then the grid is only the resident workgroups, which stride over the
// tiles; otherwise one workgroup per tile).
int launch_backproject(const BpArgs &a, hipStream_t stream, bool beside = false);
// Workgroups of launch_backproject(a) resident on the device at once (per
// slice range): *groups.
int backproject_resident(const BpArgs &a, long *groups);

// Persistent backprojection over the tiles of many blocks (jobs).  The jobs
// of one launch share the kernel variant (rows, staging shape, slices per
// thread, projection type): see backproject_variant.
struct BpQueues {
    const BpArgs *jobs;        // [n_jobs] (device memory)
    const int *band_pre;       // [8][n_jobs + 1]: items of queue q before job j (band q of each job)
    int n_jobs;
    int job_first, job_end;    // the launch covers jobs [job_first, job_end)
    int *counters;             // [8] item counters, one per queue, zeroed before the launch
    int slot_offset;           // set by the launcher: byte offset of the counter slots in dynamic LDS
};
// Fills the staging shape (tpr, tpr_log2) of a and returns an identifier of
// its kernel variant (equal identifiers: one persistent launch may hold
// both), or a negative ctbb_status.
int backproject_variant(BpArgs &a);
// Work-queue launch over the jobs of Q (all of variant `variant`, job
// arguments as the variant filled them); `sample` is any of those jobs
// (host copy).  max_groups bounds the grid (0: the resident workgroups).
int launch_backproject_persistent(const BpArgs &sample, int variant, const BpQueues &Q, int total_items,
                                  int max_groups, hipStream_t stream);

// ---------------------------------------------------------------------------
// Rebinning (fused single pass per FFS mode)
struct RebinArgs {
    const float *raw;          // [n_proj_pull][n_rows_raw][n_channels]
    float *out;                // [n_proj_out][n_rows][nos]
    const float *bidx;         // [n_sets][nos]
    const float *aoff;         // [n_sets][nos]
    int kind;                  // 0 nffs, 1 zffs, 2 pffs, 3 affs
    int n_ffs, nch, nos, n_rows_raw, n_rows, n_proj, n_proj_out, add;
    int reverse_rows;
    unsigned *nonfinite;       // set to 1 when an output sample is not finite (or null)
    // optional (null: every frame from raw): hot[f] != 0 reads frame f of the
    // pull from raw_af instead (the adaptive filter's output, written only
    // for the frames it changes)
    const unsigned char *hot;
    const float *raw_af;
    // > 0: at most this many workgroups, each looping over the work (the
    // rebin then cannot take the CUs a kernel launched beside it needs)
    int max_groups;
};
int launch_rebin(const RebinArgs &a, hipStream_t stream);

// ---------------------------------------------------------------------------
// Ramp filter as a Toeplitz product Y = X T, T[k][l] = taps[n + l - k]
struct FilterArgs {
    const float *in;           // [rows][n]
    void *out;                 // [rows][n] float32 or bfloat16
    const float *taps;         // [2n]
    const float2 *spec;        // FFT32: [H (fft_len)][twiddles (fft_len)]
    int fft_len;               // length of spec's transform: 3072 (2n <= 3072) or 4096
    size_t rows;
    int n;
    int mode;                  // ctbb_filter_mode
    int out_bf16;              // 1: write bfloat16
};
int launch_filter(const FilterArgs &a, hipStream_t stream);

// ---------------------------------------------------------------------------
// Adaptive raw-data filter
struct AdaptiveArgs {
    const float *raw;          // [n_proj][n_rows_raw][n_channels]
    float *out;
    float *work;               // >= 6 * n_proj floats
    int n_proj, n_rows_raw, nch;
    int n_proj_ffs;
    float strength;
    int vec4;                  // 1: 16-byte apply kernel where aligned (GPU otherwise idle)
    float fixed_threshold;     // > 0: this threshold for every view instead of the KK bisection
    // 1: write out only for the views the filter changes (those whose maximum
    // exceeds their threshold), flagged in adaptive_hot_flags(work, n_proj);
    // a reader takes the other views from raw (RebinArgs::hot)
    int hot_only;
    // > 0: the view-maximum kernel runs at most this many workgroups, each
    // looping over views (see RebinArgs::max_groups)
    int max_groups;
};
// the per-view flags of hot_only inside AdaptiveArgs::work (n_proj bytes)
inline unsigned char *adaptive_hot_flags(float *work, int n_proj) {
    return reinterpret_cast<unsigned char *>(work + 2 * static_cast<size_t>(n_proj));
}
int launch_adaptive(const AdaptiveArgs &a, hipStream_t stream);

// ---------------------------------------------------------------------------
// Finalize: reorder + thicken (+ HU)
struct FinalizeArgs {
    const float *image;        // [n_raw][nxy]
    float *out;                // [n_final][nxy]
    const int *src_slice;      // [n_final * max_taps] physical raw slice index, -1 = none
    const float *weight;       // [n_final * max_taps] normalised weight
    int max_taps, n_final;
    size_t nxy;
    float mu_water;            // > 0: HU
};
int launch_finalize(const FinalizeArgs &a, hipStream_t stream);

// Backprojection view table of one block (half * n_ht entries) on the device
int launch_view_table(BpView *tab, double tube_start, int half, int n_ht, int n_proj_turn, double z_rot,
                      hipStream_t stream);

// Entry e = i * n_ht + k (angle i < half, half turn k < n_ht) of a block's
// view table: theta = tube_start + (2 pi' / n_proj_turn) i + k pi'
// (reference include/backproject.cuh:75), its sine and cosine, the z term
// z_rot / 2 - z_rot (theta - tube_start) / (2 pi') and the rebinned
// projection i + k * half, in FP64 rounded to FP32 once.  Explicitly rounded
// operations keep the FP64 expression free of contractions.
__device__ inline void view_table_entry(BpView *out, int e, double tube_start, int half, int n_ht, int n_proj_turn,
                                        double z_rot) {
    const int i = e / n_ht, k = e - i * n_ht;
    const double pi = static_cast<double>(kPiBp);
    const double dtheta = __dadd_rn(__dmul_rn(__ddiv_rn(2.0 * pi, static_cast<double>(n_proj_turn)), i),
                                    __dmul_rn(static_cast<double>(k), pi));
    const double th = __dadd_rn(tube_start, dtheta);
    BpView v;
    v.sin_t = static_cast<float>(sin(th));
    v.cos_t = static_cast<float>(cos(th));
    v.c_v = static_cast<float>(__dsub_rn(z_rot / 2.0, __ddiv_rn(__dmul_rn(z_rot, dtheta), 2.0 * pi)));
    v.proj = i + k * half;
    *out = v;
}

// ---------------------------------------------------------------------------
// Batching form: the work list built on the device (plan.hip)
struct ScanPlanIn {
    const double *table_positions;  // device copy, n_readings entries
    float *image;                   // the scan's image stack [n_slices_recon][nx][ny]
    BpArgs tmpl;                    // geometry of the scan's jobs; rebin / rebin_bf16 = its first filtered block
    size_t slot_elems;              // elements between consecutive filtered blocks of the scan
    float recon_start_pos, coll_slicewidth, tube_angle0, z_rot;
    int n_readings, array_direction, recon_direction, n_blocks, n_slices_block, block_first;
    int n_ffs, ppb, add_ffs, n_proj_ffs, add_projections, n_proj_turn;
    size_t add_projections_ffs;
};
struct BlockPlan {
    int scan, block;
    float slice_start, slice_end, tube_start;
    int idx_slice_start, idx_slice_end, pull_start, pull_end, n_pull, n_ht;
};
// plans[n_blocks], jobs[n_blocks] (jobs' geometry from their scan's tmpl),
// view tables of every block at d_vtab + vt_off[block] (vt_off[n_blocks + 1]),
// queue prefixes band_pre[8][n_blocks + 1].
int launch_plan_batch(const ScanPlanIn *d_scans, int n_scans, int n_blocks, int max_views_per_block,
                      BlockPlan *d_plans, BpArgs *d_jobs, BpView *d_vtab, size_t *d_vt_off, int *d_band_pre,
                      hipStream_t stream);

// HU -> int16 stored pixel values (HU = slope * stored + intercept)
struct Hu16Args {
    const float *in;
    int16_t *out;
    size_t n;
    float slope, intercept;
};
int launch_hu16(const Hu16Args &a, hipStream_t stream);

// ---------------------------------------------------------------------------
// Synthetic scan generator
int launch_synth(const struct recon_metadata &mr, const ctbb_phantom_params &pp, long first, long count,
                 float *d_out, hipStream_t stream);

}  // namespace ctbb
