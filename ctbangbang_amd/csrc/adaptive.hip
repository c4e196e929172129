// Adaptive raw-data filtering (Kachelriess-Kalender style), reference
// src/preprocessing.cu:31-156 and include/preprocessing.cuh:32-285.
//
// The reference runs one thread per view for every step (a serial loop over
// the 23 552 samples of a view in filter_projections).  Here each step is
// parallelised to its natural width while keeping the reference's float
// arithmetic in the same order (including the fused multiply-adds nvcc forms
// for "acc += a*b"), so results are bit-identical to a sequential FP32
// restatement:
//   1. view maximum         one workgroup per view, tree max (order-free)
//   2. 18-degree smoothing  one thread per view, sequential fmaf sum
//   3. 180-degree min/max + eccentricity, and
//   4. threshold bisection  one wavefront per view: the view's window is
//      copied once to LDS (min and max taken on the way, order-free), then
//      exact integer counts over it
//   5. 5x5x5 triangle smoothing of samples above threshold: one thread per
//      sample (coalesced along channels)
// A fixed-threshold mode (no counterpart in the reference) skips steps 1-4
// and smooths the samples above one given line-integral threshold, e.g.
// ln(I0 / N) for a noise threshold of N photons at I0 incident photons.
#include <hip/hip_runtime.h>

#include "kernels.h"

// Only the explicit fmaf calls may fuse: every other multiply-add rounds
// twice, as in the sequential FP32 statement of the algorithm.
#pragma clang fp contract(off)

namespace ctbb {

namespace {

__device__ __forceinline__ float wave_max(float v) {
    for (int s = 32; s > 0; s >>= 1) v = fmaxf(v, __shfl_xor(v, s, 64));
    return v;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int s = 32; s > 0; s >>= 1) v = fminf(v, __shfl_xor(v, s, 64));
    return v;
}

// View maximum, one workgroup per view (max is order-free, so any reduction
// order gives the reference's value), gridDim.x workgroups looping over the
// n_proj views.  frame % 4 == 0 is checked on the host.
__global__ __launch_bounds__(256) void af_sup(const float *raw, float *sup, size_t frame, int n_proj) {
    __shared__ float red[4];
    for (int p = blockIdx.x; p < n_proj; p += gridDim.x) {
        const float4 *f = reinterpret_cast<const float4 *>(raw + static_cast<size_t>(p) * frame);
        float m = 0.0f;   // reference starts from 0.0f (preprocessing.cuh:75)
        for (size_t i = threadIdx.x; i < frame / 4; i += 256) {
            const float4 v = f[i];
            m = fmaxf(fmaxf(m, fmaxf(v.x, v.y)), fmaxf(v.z, v.w));
        }
        m = wave_max(m);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
        __syncthreads();
        if (threadIdx.x == 0) sup[p] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
        __syncthreads();                        // red is rewritten for the next view
    }
}

// reference preprocessing.cuh:87-113
__global__ void af_smooth(const float *sup, float *sm, int n_proj, int n_avg) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_proj) return;
    const int off = n_avg / 2;
    const int max_idx = n_proj - 1;
    const float inv = 1.0f / static_cast<float>(n_avg);
    float acc = 0.0f;
    for (int i = 0; i < n_avg; i++) {
        int idx = p - off + i;
        idx = idx < 0 ? 0 : (idx > max_idx ? max_idx : idx);
        acc = fmaf(inv, sup[idx], acc);
    }
    sm[p] = acc;
}

// The counting window of reference preprocessing.cuh:32-62 (clipped to the
// pull), fixed for a view while the bisection moves T.
__device__ __forceinline__ void af_window(int &start_idx, int &numel_range, int max_idx) {
    const int diff = start_idx + (numel_range - 1);
    if (diff < numel_range - 1) {
        numel_range += start_idx;
        start_idx = 0;
    }
    if (diff > max_idx) numel_range -= (diff - max_idx);
}

// rhs of reference preprocessing.cuh:32-62 for a window of numel values:
// the exact count of values above T, then the reference's float products.
__device__ __forceinline__ float af_rhs_wave(const float *win, int numel, float T, float inv_pi_npf_a,
                                             float inv_npf) {
    int cnt = 0;
    for (int i = threadIdx.x & 63; i < numel; i += 64) cnt += win[i] > T ? 1 : 0;
    for (int s = 32; s > 0; s >>= 1) cnt += __shfl_xor(cnt, s, 64);
    return inv_pi_npf_a * static_cast<float>(cnt) * inv_npf;
}

// reference preprocessing.cuh:115-206; one wavefront per view.  The view's
// window is copied to LDS once (kLds, when it fits in the dynamic LDS the
// host gives the launch), its minimum and maximum taken on the way (the
// reference's 180-degree range, preprocessing.cuh:115-163: the same views as
// the counting window, its clamped edge indices repeat values inside it),
// and every bisection step counts from there; otherwise each step reads the
// window from memory.
template <bool kLds>
__global__ __launch_bounds__(64) void af_thresh(const float *sm, float *thr, int n_proj, int n_range, int n_proj_ffs,
                                                float strength) {
    extern __shared__ float win_lds[];
    const int p = blockIdx.x;
    const int max_idx = n_proj - 1;
    int start = p - n_range / 2, numel = n_range;
    af_window(start, numel, max_idx);
    const float *win = sm + start;
    float hi = 0.0f, lo = 10000.0f;     // the reference's starting values
    for (int i = threadIdx.x; i < numel; i += 64) {
        const float v = win[i];
        hi = fmaxf(hi, v);
        lo = fminf(lo, v);
        if (kLds) win_lds[i] = v;
    }
    hi = wave_max(hi);
    lo = wave_min(lo);
    if (kLds) {
        __syncthreads();
        win = win_lds;
    }
    float e = 1 - lo / hi;
    e = (e - 0.3f) / (0.5f - 0.3f);
    const float ecc = e > 1 ? 1.0f : (e < 0 ? 0.0f : e);
    const float inv_pi = 1.0f / 3.1415368979f;
    const float inv_npf = 1.0f / static_cast<float>(n_proj_ffs);
    float lower = lo, upper = hi;
    float T = (hi + lo) / 2.0f;
    const float lhs = 0.03f * strength * ecc;
    float rhs = af_rhs_wave(win, numel, T, inv_pi, inv_npf);
    int iter = 0;
    while (static_cast<double>(fabsf(lhs - rhs) / lhs) > 0.01) {
        if (iter >= 20) break;
        iter += 1;
        if (rhs > lhs) lower = T;
        else if (rhs < lhs) upper = T;
        else break;
        T = (lower + upper) / 2.0f;
        rhs = af_rhs_wave(win, numel, T, inv_pi, inv_npf);
    }
    if (threadIdx.x == 0) thr[p] = T;
}

// reference preprocessing.cuh:208-285 (3D branch, __FILTER_3D__ defined)
// Output sample (i, j, p): copied when x <= T, else the 5x5x5 weighted
// smoothing of its neighbourhood.
__device__ __forceinline__ float af_value(const float *__restrict__ raw, float x, float T, int i, int j, int p,
                                          int n_proj, int n_rows_raw, int nch, size_t ps) {
    if (!(x > T)) return x;
    float norm = 0.0f, val = 0.0f;
    // (outer loops rolled: few registers, so the kernels fit beside a
    // backprojection; the order of the sums is unchanged)
#pragma unroll 1
    for (int ii = 0; ii < 5; ii++) {
        const int fc = i + ii - 2;
#pragma unroll 1
        for (int jj = 0; jj < 5; jj++) {
            const int fr = j + jj - 2;
            for (int kk = 0; kk < 5; kk++) {
                const int fp = p + kk - 2;
                if (fr < 0 || fc < 0 || fp < 0 || fr >= n_rows_raw || fc >= nch || fp >= n_proj) continue;
                // triangle_3d(5, ii, jj, kk), ceil(5/2) = 3: peak at offset +1
                const float t = static_cast<float>((3 - abs(ii - 3)) + (3 - abs(jj - 3)) + (3 - abs(kk - 3)) + 1);
                norm += t;
                val = fmaf(raw[ps * fp + static_cast<size_t>(nch) * fr + fc], t, val);
            }
        }
    }
    return (1.0f / norm) * val;
}

// One thread per sample (low register use: runs beside a backprojection).
__global__ __launch_bounds__(256) void af_apply(const float *raw, const float *thr, float *out, int n_proj,
                                                int n_rows_raw, int nch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // channel
    const int j = blockIdx.y;                                // row
    const int p = blockIdx.z;                                // view
    if (i >= nch) return;
    const size_t ps = static_cast<size_t>(nch) * n_rows_raw;
    const size_t idx = ps * p + static_cast<size_t>(j) * nch + i;
    out[idx] = af_value(raw, raw[idx], thr[p], i, j, p, n_proj, n_rows_raw, nch, ps);
}

// Four consecutive samples of a view's frame per thread, 16-byte loads and
// stores (for a GPU that runs nothing else: the first block of a pipeline
// and the stage entry point).
__global__ __launch_bounds__(256) void af_apply4(const float *raw, const float *thr, float *out, int n_proj,
                                                 int n_rows_raw, int nch) {
    const int p = blockIdx.y;
    const size_t ps = static_cast<size_t>(nch) * n_rows_raw;
    const int e = (blockIdx.x * blockDim.x + threadIdx.x) * 4;   // sample in the frame
    if (static_cast<size_t>(e) >= ps) return;
    const size_t base = ps * p + e;
    const float T = thr[p];
    const float4 x = *reinterpret_cast<const float4 *>(raw + base);
    float4 y = x;
    if (x.x > T || x.y > T || x.z > T || x.w > T) {
        float *yv = reinterpret_cast<float *>(&y);
        const float xv[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int j = (e + q) / nch, i = e + q - j * nch;
            yv[q] = af_value(raw, xv[q], T, i, j, p, n_proj, n_rows_raw, nch, ps);
        }
    }
    *reinterpret_cast<float4 *>(out + base) = y;
}

// Only the views the filter changes: view p is unchanged where its maximum
// (af_sup, from 0) is at or below its threshold, since af_value keeps every
// sample x <= T.  One workgroup per view flags it (hot[p]) and, if changed,
// writes its filtered frame; a reader takes the other views from the raw
// pull.  (C1: about 6% of the views.)
__global__ __launch_bounds__(256) void af_apply_hot(const float *raw, const float *sup, const float *thr, float *out,
                                                    unsigned char *hot, int n_proj, int n_rows_raw, int nch) {
    const int p = blockIdx.x;
    const float T = thr[p];
    const bool h = sup[p] > T;
    if (threadIdx.x == 0) hot[p] = h ? 1 : 0;
    if (!h) return;
    const size_t ps = static_cast<size_t>(nch) * n_rows_raw;
    for (int j = 0; j < n_rows_raw; j++)
        for (int i = threadIdx.x; i < nch; i += 256) {
            const size_t idx = ps * p + static_cast<size_t>(j) * nch + i;
            out[idx] = af_value(raw, raw[idx], T, i, j, p, n_proj, n_rows_raw, nch, ps);
        }
}

// Fixed-threshold mode: every view's threshold is the same constant.
__global__ __launch_bounds__(256) void af_const(float *thr, int n_proj, float T) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n_proj) thr[p] = T;
}

int odd_window(float frac, int n_proj_ffs) {
    int n = static_cast<int>(ceilf(frac * static_cast<float>(n_proj_ffs)));
    return n + (1 - n % 2);
}

}  // namespace

int launch_adaptive(const AdaptiveArgs &a, hipStream_t stream) {
    if (!a.raw || !a.out || !a.work || a.raw == a.out) return CTBB_ERR_ARG;
    if (a.n_proj <= 0 || a.n_rows_raw <= 0 || a.nch <= 0 || a.n_proj_ffs <= 0) return CTBB_ERR_SHAPE;
    if (a.n_proj > 65535 || a.n_rows_raw > 65535) return CTBB_ERR_SHAPE;  // grid.z / grid.y limits
    if ((static_cast<size_t>(a.n_rows_raw) * a.nch) % 4 != 0) return CTBB_ERR_SHAPE;  // float4 frames
    const size_t np = static_cast<size_t>(a.n_proj);
    // work: [sup][sm][3 unused][thr], n_proj floats each
    float *sup = a.work, *sm = sup + np, *thr = sm + 4 * np;
    const size_t frame = static_cast<size_t>(a.n_rows_raw) * a.nch;
    const int n_avg = odd_window(18.0f / 360.0f, a.n_proj_ffs);
    const int n_range = odd_window(180.0f / 360.0f, a.n_proj_ffs);
    const unsigned g1 = static_cast<unsigned>((a.n_proj + 255) / 256);
    if (a.max_groups < 0) return CTBB_ERR_SHAPE;
    const unsigned gsup = static_cast<unsigned>(a.max_groups > 0 && a.max_groups < a.n_proj ? a.max_groups : a.n_proj);
    if (a.fixed_threshold > 0.0f) {
        if (!(a.fixed_threshold < 3.0e38f)) return CTBB_ERR_ARG;
        hipLaunchKernelGGL(af_const, dim3(g1), dim3(256), 0, stream, thr, a.n_proj, a.fixed_threshold);
        // (the view maxima decide which views hot_only writes)
        if (a.hot_only) hipLaunchKernelGGL(af_sup, dim3(gsup), dim3(256), 0, stream, a.raw, sup, frame, a.n_proj);
    } else {
        hipLaunchKernelGGL(af_sup, dim3(gsup), dim3(256), 0, stream, a.raw, sup, frame, a.n_proj);
        hipLaunchKernelGGL(af_smooth, dim3(g1), dim3(256), 0, stream, sup, sm, a.n_proj, n_avg);
        // the counting window (at most n_range values) in LDS when it takes <= 16 KB
        const size_t win_bytes = static_cast<size_t>(n_range) * sizeof(float);
        if (win_bytes <= 16384)
            hipLaunchKernelGGL(af_thresh<true>, dim3(a.n_proj), dim3(64), win_bytes, stream, sm, thr, a.n_proj,
                               n_range, a.n_proj_ffs, a.strength);
        else
            hipLaunchKernelGGL(af_thresh<false>, dim3(a.n_proj), dim3(64), 0, stream, sm, thr, a.n_proj, n_range,
                               a.n_proj_ffs, a.strength);
    }
    if (a.hot_only) {
        hipLaunchKernelGGL(af_apply_hot, dim3(a.n_proj), dim3(256), 0, stream, a.raw, sup, thr, a.out,
                           adaptive_hot_flags(a.work, a.n_proj), a.n_proj, a.n_rows_raw, a.nch);
        return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
    }
    const bool aligned = (reinterpret_cast<uintptr_t>(a.raw) & 15u) == 0 && (reinterpret_cast<uintptr_t>(a.out) & 15u) == 0;
    if (a.vec4 && aligned)
        hipLaunchKernelGGL(af_apply4, dim3(static_cast<unsigned>((frame / 4 + 255) / 256), a.n_proj), dim3(256), 0,
                           stream, a.raw, thr, a.out, a.n_proj, a.n_rows_raw, a.nch);
    else
        hipLaunchKernelGGL(af_apply, dim3((a.nch + 255) / 256, a.n_rows_raw, a.n_proj), dim3(256), 0, stream, a.raw,
                           thr, a.out, a.n_proj, a.n_rows_raw, a.nch);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
