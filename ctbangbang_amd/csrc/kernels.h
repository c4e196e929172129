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
int launch_backproject(const BpArgs &a, hipStream_t stream);

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
};
int launch_rebin(const RebinArgs &a, hipStream_t stream);

// ---------------------------------------------------------------------------
// Ramp filter as a Toeplitz product Y = X T, T[k][l] = taps[n + l - k]
struct FilterArgs {
    const float *in;           // [rows][n]
    void *out;                 // [rows][n] float32 or bfloat16
    const float *taps;         // [2n]
    const float2 *spec;        // FFT32: [H (4096)][twiddles (4096)]
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
};
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
