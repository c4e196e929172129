// Fan-to-parallel rebinning with flying focal spot, fused into a single pass
// per FFS mode and writing the backprojection layout directly.
//
// Reference behaviour (GPU path, clamp-addressed linear texture fetches):
//   no FFS   n1_rebin/n2_rebin    include/rebin_filter.cuh:72-96
//   z FFS    z1_rebin/z2_rebin    include/rebin_filter.cuh:186-228
//   phi FFS  p1/p2_rebin_t + p1/p2_rebin   include/rebin_filter.cuh:111-169
//   phi+z    a1..a4_rebin_t + a1/a2_rebin_b include/rebin_filter.cuh:247-350
//   + reshape_out (view trimming, [proj][row][channel] layout)
//                                 include/rebin_filter.cuh:354-366
// The reference stages the raw data through focal-spot "sheets", texture
// arrays and a full-size intermediate per pass; here each output sample is
// computed directly from the raw pull with the same clamped bilinear
// arithmetic (2 taps along the azimuthal axis of each of the 2 channel taps),
// using FP64-derived per-channel coordinate maps (geometry.cpp).
//
// Output element (v, r, c) = rebinned projection v + add_projections, row r,
// oversampled channel c.  Threads run along c (coalesced writes, raw reads at
// channel ~c/2 of a handful of readings that stay in L1/L2).
//
// Every raw sample the reconstruction uses reaches some rebinned sample
// (non-finite values propagate through the interpolation), so the kernels
// also report non-finite input: a wave with a non-finite output sample
// stores 1 to *nonfinite (the reference instead drops such fetches in its
// backprojection and resets h_t, include/backproject.cuh:89-95).
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctbb {

namespace {

constexpr int kViewsPerThread = 8;

// Clamp a fractional index into [0, n-1] and split it into the two taps and
// the weight of the upper tap (texture clamp addressing).
__device__ __forceinline__ void taps(float x, int n, int &i0, int &i1, float &f) {
    x = fminf(fmaxf(x, 0.0f), static_cast<float>(n - 1));
    float fl = floorf(x);
    i0 = static_cast<int>(fl);
    i1 = min(i0 + 1, n - 1);
    f = x - fl;
}

// One plain store of 1 per wave that saw a non-finite output sample (every
// writer stores the same value, so no atomic is needed).
__device__ __forceinline__ void report_nonfinite(unsigned *flag, bool bad) {
    const unsigned long long any = __builtin_amdgcn_ballot_w64(bad);
    if (any && flag && (threadIdx.x & 63) == static_cast<unsigned>(__builtin_ctzll(any)))
        *reinterpret_cast<volatile unsigned *>(flag) = 1u;
}

// Base of frame f of the pull: with kHot the adaptive filter's output where
// it changed the frame, the raw pull otherwise.
template <bool kHot>
__device__ __forceinline__ const float *frame_src(const RebinArgs &a, size_t f) {
    if (!kHot) return a.raw;
    return a.hot[f] ? a.raw_af : a.raw;
}

// The (x, y, z) workgroups of the rebin's grid (x: 256 output channels, y:
// row, z: kViewsPerThread views), in dispatch order, taken in a stride of
// gridDim.x by the launched one-dimensional grid (fewer workgroups than the
// grid has: the rebin then leaves room on every CU for a kernel launched
// beside it).
template <class F>
__device__ __forceinline__ void for_each_block(const RebinArgs &a, F f) {
    const unsigned gx = static_cast<unsigned>(a.nos + 255) / 256u, gxy = gx * static_cast<unsigned>(a.n_rows);
    const unsigned total = gxy * static_cast<unsigned>((a.n_proj_out + kViewsPerThread - 1) / kViewsPerThread);
    for (unsigned vb = blockIdx.x; vb < total; vb += gridDim.x) {
        const unsigned bz = vb / gxy, rem = vb - bz * gxy, br = rem / gx;
        f(rem - br * gx, br, bz);
    }
}

template <bool kHot>
__device__ __forceinline__ void rebin_block(const RebinArgs &a, unsigned bx, unsigned br, unsigned bz, bool &bad) {
    const int c = static_cast<int>(bx) * 256 + static_cast<int>(threadIdx.x);
    const int r = static_cast<int>(br);
    if (c >= a.nos) return;
    const size_t frame = static_cast<size_t>(a.n_rows_raw) * a.nch;
    const float *__restrict__ aoff = a.aoff;
    float *__restrict__ outp = a.out;

    int set, j;
    if (a.kind == 0 || a.kind == 2) {
        set = 0;
        j = r;
    } else {
        // interleaved rows: set 0 (-dr focal spot(s)) -> row 2j, set 1 -> 2j+1,
        // swapped with ReverseRowInterleave
        set = (r & 1) ^ (a.reverse_rows ? 1 : 0);
        j = r >> 1;
    }
    const float by = a.bidx[set * a.nos + c];

    // per-thread constants of the output channel (independent of the view)
    int c0 = 0, c1 = 0, m0 = 0, m1 = 0;
    float fc = 0.0f, fm = 0.0f, ao0, ao1 = 0.0f;
    if (a.kind <= 1) {
        taps(by, a.nch, c0, c1, fc);
        ao0 = aoff[set * a.nos + c];
    } else {
        taps(by, a.nos, m0, m1, fm);
        ao0 = aoff[set * a.nos + m0];
        ao1 = aoff[set * a.nos + m1];
    }

    for (int vv = 0; vv < kViewsPerThread; vv++) {
        const int v = static_cast<int>(bz) * kViewsPerThread + vv;
        if (v >= a.n_proj_out) break;
        const float proj = static_cast<float>(v + a.add);
        float out;
        if (a.kind <= 1) {
            // single 2D bilinear fetch in (channel, sheet projection)
            const int fs = a.kind == 1 ? set : 0;
            int p0, p1;
            float fa;
            taps(proj - ao0, a.n_proj, p0, p1, fa);
            const size_t f0 = static_cast<size_t>(p0) * a.n_ffs + fs, f1 = static_cast<size_t>(p1) * a.n_ffs + fs;
            const float *s0 = frame_src<kHot>(a, f0) + static_cast<size_t>(j) * a.nch + f0 * frame;
            const float *s1 = frame_src<kHot>(a, f1) + static_cast<size_t>(j) * a.nch + f1 * frame;
            const float t0 = fmaf(fc, s0[c1] - s0[c0], s0[c0]);
            const float t1 = fmaf(fc, s1[c1] - s1[c0], s1[c0]);
            out = fmaf(fa, t1 - t0, t0);
        } else {
            // two-pass modes: radial tap pair over the interleaved azimuthal
            // rebin of the focal-spot pair `set`
            float t[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int m = h ? m1 : m0;
                const int fs = 2 * set + (m & 1);
                const int ch = m >> 1;
                int p0, p1;
                float fa;
                taps(proj - (h ? ao1 : ao0), a.n_proj, p0, p1, fa);
                const size_t f0 = static_cast<size_t>(p0) * a.n_ffs + fs, f1 = static_cast<size_t>(p1) * a.n_ffs + fs;
                const float u0 = frame_src<kHot>(a, f0)[static_cast<size_t>(j) * a.nch + ch + f0 * frame];
                const float u1 = frame_src<kHot>(a, f1)[static_cast<size_t>(j) * a.nch + ch + f1 * frame];
                t[h] = fmaf(fa, u1 - u0, u0);
            }
            out = fmaf(fm, t[1] - t[0], t[0]);
        }
        outp[(static_cast<size_t>(v) * a.n_rows + r) * a.nos + c] = out;
        bad |= !__builtin_isfinite(out);
    }
}

template <bool kHot>
__global__ __launch_bounds__(256) void rebin_kernel(RebinArgs a) {
    bool bad = false;
    rebin_block<kHot>(a, blockIdx.x, blockIdx.y, blockIdx.z, bad);
    report_nonfinite(a.nonfinite, bad);
}

// The same arithmetic with 32-bit byte offsets from the (uniform) raw and
// output bases, one template per channel-tap form (kinds 0/1: bilinear in
// raw channel and azimuth; 2/3: two oversampled-channel taps, each bilinear
// in azimuth), for pulls and outputs below 4 GiB.  Few registers (<= 32
// VGPRs), so its waves fit beside a running backprojection (4 waves per SIMD
// of 112 VGPRs leave 64 per lane) and the rebin of the next block overlaps
// it.
__device__ __forceinline__ float ld(const float *base, unsigned boff) {
    return *reinterpret_cast<const float *>(reinterpret_cast<const char *>(base) + boff);
}

// Sample at byte offset boff of frame f (kHot: from the adaptive filter's
// output where it changed frame f).
template <bool kHot>
__device__ __forceinline__ float ld_frame(const RebinArgs &a, unsigned f, unsigned boff) {
    if (!kHot) return ld(a.raw, boff);
    return ld(a.hot[f] ? a.raw_af : a.raw, boff);
}

template <bool kSplit, bool kHot>
__device__ __forceinline__ void rebin32_block(const RebinArgs &a, unsigned bx, unsigned br, unsigned bz, bool &bad) {
    const int c = bx * blockDim.x + threadIdx.x;
    const int r = br;
    if (c >= a.nos) return;
    const unsigned frame_b = static_cast<unsigned>(a.n_rows_raw * a.nch) * 4u;
    int set, j;
    if (a.kind == 0 || a.kind == 2) {
        set = 0;
        j = r;
    } else {
        set = (r & 1) ^ (a.reverse_rows ? 1 : 0);
        j = r >> 1;
    }
    const float by = a.bidx[set * a.nos + c];
    int c0, c1;
    float fc, ao0, ao1 = 0.0f;
    if (!kSplit) {
        taps(by, a.nch, c0, c1, fc);
        ao0 = a.aoff[set * a.nos + c];
    } else {
        taps(by, a.nos, c0, c1, fc);
        ao0 = a.aoff[set * a.nos + c0];
        ao1 = a.aoff[set * a.nos + c1];
    }
    const unsigned row_b = static_cast<unsigned>(j * a.nch) * 4u;
    const int v0 = bz * kViewsPerThread;
    unsigned ob = (static_cast<unsigned>(v0 * a.n_rows + r) * static_cast<unsigned>(a.nos) + c) * 4u;
    const unsigned ostep = static_cast<unsigned>(a.n_rows * a.nos) * 4u;
    for (int vv = 0; vv < kViewsPerThread; vv++) {
        const int v = v0 + vv;
        if (v >= a.n_proj_out) break;
        const float proj = static_cast<float>(v + a.add);
        float out;
        if (!kSplit) {
            const unsigned fs = a.kind == 1 ? static_cast<unsigned>(set) : 0u;
            int p0, p1;
            float fa;
            taps(proj - ao0, a.n_proj, p0, p1, fa);
            const unsigned f0 = static_cast<unsigned>(p0 * a.n_ffs) + fs, f1 = static_cast<unsigned>(p1 * a.n_ffs) + fs;
            const unsigned b0 = f0 * frame_b + row_b, b1 = f1 * frame_b + row_b;
            const float s00 = ld_frame<kHot>(a, f0, b0 + 4u * c0), s01 = ld_frame<kHot>(a, f0, b0 + 4u * c1);
            const float s10 = ld_frame<kHot>(a, f1, b1 + 4u * c0), s11 = ld_frame<kHot>(a, f1, b1 + 4u * c1);
            const float t0 = fmaf(fc, s01 - s00, s00);
            const float t1 = fmaf(fc, s11 - s10, s10);
            out = fmaf(fa, t1 - t0, t0);
        } else {
            float t[2];
#pragma unroll
            for (int h = 0; h < 2; h++) {
                const int m = h ? c1 : c0;
                const unsigned fs = static_cast<unsigned>(2 * set + (m & 1));
                const unsigned ch_b = row_b + 4u * static_cast<unsigned>(m >> 1);
                int p0, p1;
                float fa;
                taps(proj - (h ? ao1 : ao0), a.n_proj, p0, p1, fa);
                const unsigned f0 = static_cast<unsigned>(p0 * a.n_ffs) + fs, f1 = static_cast<unsigned>(p1 * a.n_ffs) + fs;
                const float u0 = ld_frame<kHot>(a, f0, f0 * frame_b + ch_b);
                const float u1 = ld_frame<kHot>(a, f1, f1 * frame_b + ch_b);
                t[h] = fmaf(fa, u1 - u0, u0);
            }
            out = fmaf(fc, t[1] - t[0], t[0]);
        }
        *reinterpret_cast<float *>(reinterpret_cast<char *>(a.out) + ob) = out;
        bad |= !__builtin_isfinite(out);
        ob += ostep;
    }
}

// One workgroup per (x, y, z) block of the grid.  At most 32 VGPRs (the
// attribute counts pairs here): two waves per SIMD fit in the 64 per lane a
// running backprojection leaves.
template <bool kSplit, bool kHot>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8), amdgpu_num_vgpr(16))) void rebin32_kernel(
    RebinArgs a) {
    bool bad = false;
    rebin32_block<kSplit, kHot>(a, blockIdx.x, blockIdx.y, blockIdx.z, bad);
    report_nonfinite(a.nonfinite, bad);
}

// A one-dimensional grid of RebinArgs::max_groups workgroups looping over the
// blocks, at most 32 VGPRs as above.
template <bool kSplit, bool kHot>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8), amdgpu_num_vgpr(16))) void rebin32_loop_kernel(
    RebinArgs a) {
    bool bad = false;
    for_each_block(a, [&](unsigned bx, unsigned br, unsigned bz) { rebin32_block<kSplit, kHot>(a, bx, br, bz, bad); });
    report_nonfinite(a.nonfinite, bad);
}

}  // namespace

int launch_rebin(const RebinArgs &a, hipStream_t stream) {
    if (!a.raw || !a.out || !a.bidx || !a.aoff || (a.hot && !a.raw_af)) return CTBB_ERR_ARG;
    if (a.kind < 0 || a.kind > 3 || a.nch <= 1 || a.nos != 2 * a.nch || a.n_proj < 2 || a.n_proj_out <= 0)
        return CTBB_ERR_SHAPE;
    if (a.add < 0 || a.add + a.n_proj_out > a.n_proj) return CTBB_ERR_SHAPE;
    const int rows_needed = (a.kind == 1 || a.kind == 3) ? 2 * a.n_rows_raw : a.n_rows_raw;
    if (a.n_rows != rows_needed) return CTBB_ERR_SHAPE;
    const dim3 grid((a.nos + 255) / 256, a.n_rows, (a.n_proj_out + kViewsPerThread - 1) / kViewsPerThread);
    if (grid.z > 65535u || grid.y > 65535u || a.max_groups < 0) return CTBB_ERR_SHAPE;
    const unsigned long long blocks = static_cast<unsigned long long>(grid.x) * grid.y * grid.z;
    const unsigned long long raw_bytes = static_cast<unsigned long long>(a.n_proj) * a.n_ffs * a.n_rows_raw * a.nch * 4ull;
    const unsigned long long out_bytes = static_cast<unsigned long long>(a.n_proj_out) * a.n_rows * a.nos * 4ull;
    const bool hot = a.hot != nullptr;
    if (raw_bytes < (1ull << 32) && out_bytes < (1ull << 32)) {
        if (a.max_groups > 0 && static_cast<unsigned long long>(a.max_groups) < blocks) {
            const dim3 g1(static_cast<unsigned>(a.max_groups));
            if (a.kind <= 1 && !hot)
                hipLaunchKernelGGL((rebin32_loop_kernel<false, false>), g1, dim3(256), 0, stream, a);
            else if (a.kind <= 1)
                hipLaunchKernelGGL((rebin32_loop_kernel<false, true>), g1, dim3(256), 0, stream, a);
            else if (!hot)
                hipLaunchKernelGGL((rebin32_loop_kernel<true, false>), g1, dim3(256), 0, stream, a);
            else
                hipLaunchKernelGGL((rebin32_loop_kernel<true, true>), g1, dim3(256), 0, stream, a);
        } else if (a.kind <= 1 && !hot) {
            hipLaunchKernelGGL((rebin32_kernel<false, false>), grid, dim3(256), 0, stream, a);
        } else if (a.kind <= 1) {
            hipLaunchKernelGGL((rebin32_kernel<false, true>), grid, dim3(256), 0, stream, a);
        } else if (!hot) {
            hipLaunchKernelGGL((rebin32_kernel<true, false>), grid, dim3(256), 0, stream, a);
        } else {
            hipLaunchKernelGGL((rebin32_kernel<true, true>), grid, dim3(256), 0, stream, a);
        }
    } else {
        // (more than 32 VGPRs: long-lived workgroups of it would keep a
        // backprojection workgroup off their CU, so its grid stays whole)
        if (hot)
            hipLaunchKernelGGL(rebin_kernel<true>, grid, dim3(256), 0, stream, a);
        else
            hipLaunchKernelGGL(rebin_kernel<false>, grid, dim3(256), 0, stream, a);
    }
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
