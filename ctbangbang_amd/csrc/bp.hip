// Weighted-FBP cone-beam backprojection (Stierstorfer weights) of one
// 32-slice block, written for CDNA4 (gfx950).
//
// Behaviour follows reference src/backproject.cu:31-108 and
// include/backproject.cuh:33-198 (GPU conventions):
//   x(xi) flipped unless table_direction == -1, q sign = -table_direction,
//   output layout [z][xi][yi], voxels with r >= AcqFOV/2 set to 0,
//   zero-border bilinear fetches, per angle i: s += s_t/h_t over half turns.
// The reference's texture fetches are replaced by explicit FP32 bilinear
// interpolation from an LDS-staged projection footprint.
//
// Mapping: one workgroup = 16x16 voxel columns x NZ slices; one thread = one
// column (xi, yi) with its NZ slice accumulators in registers.  Views are
// processed in the reference order (angle i outer, half turn k inner) through
// a two-stage pipeline: while the workgroup interpolates view t from one LDS
// buffer, the rows x W channel window of view t+1 is already in flight from
// global memory into registers and lands in the other buffer (one barrier
// per view).  The window holds (value, next-minus-value) pairs so both taps
// of a row come from one 64-bit LDS read and each row interpolation is one
// FMA; channels outside the detector are staged as zeros, which is exactly
// the reference's border addressing.
//
// W(q) = 0 for |q| >= 1 contributes exactly nothing; since q varies by less
// than 1% across a wavefront, slices no lane sees are skipped on a
// wave-uniform vote, and the cos^2 taper is evaluated only when some lane is
// in 0.6 <= |q| < 1.
//
// Workgroups are remapped so the tiles running on one XCD are spatially
// contiguous: their footprints on each view overlap and hit in that XCD's L2.
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctbb {

namespace {

// cos^2(pi'/2 * t) on t in [0, 1] (pi' = 3.1415368979f, the constant the
// reference's weight uses): degree-8 Chebyshev interpolant converted to a
// power series, max abs error 7e-7 in FP32 Horner form.
__device__ __forceinline__ float taper(float t) {
    float p = 3.0801288630755153e-06f;
    p = fmaf(p, t, 0.27724477648735046f);
    p = fmaf(p, t, -0.970396876335144f);
    p = fmaf(p, t, 0.18480874598026276f);
    p = fmaf(p, t, 1.9640635251998901f);
    p = fmaf(p, t, 0.012797481380403042f);
    p = fmaf(p, t, -2.468567371368408f);
    p = fmaf(p, t, 4.714697934105061e-05f);
    p = fmaf(p, t, 0.9999997019767761f);
    return p;
}

// Taper part of the Stierstorfer weight W(q), Q = 0.6 (reference
// include/backproject.cuh:33-46), for 0.6 <= |q| < 1.
__device__ __forceinline__ float stier_taper(float aq) {
    return taper(__builtin_amdgcn_fmed3f((aq - 0.6f) * 2.5f, 0.0f, 1.0f));
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// Raw bits of one filtered-projection sample (float32 bits, or bfloat16 bits
// in the low half).  Conversion happens when the prefetched window is written
// to LDS, so the loads stay in flight while the previous view is computed.
template <bool kBf16>
__device__ __forceinline__ uint32_t load_bits(const BpArgs &a, size_t i) {
    if (kBf16) return a.rebin_bf16[i];
    return __float_as_uint(a.rebin[i]);
}

template <bool kBf16>
__device__ __forceinline__ float bits_to_f32(uint32_t b) {
    return kBf16 ? bf16_to_f32(static_cast<uint16_t>(b)) : __uint_as_float(b);
}

// asin(x) for |x| <= 0.45 (the ray's fan angle seen from a parallel-beam
// coordinate phat/r_f): odd degree-11 polynomial fitted on Chebyshev nodes,
// max abs error 4.2e-8 in FP32; larger |x| (FOV radius > 0.45 r_f) use asinf.
__device__ __forceinline__ float asin_small(float x) {
    const float t = x * x;
    float p = 0.0372612439095974f;
    p = fmaf(p, t, 0.026611803099513054f);
    p = fmaf(p, t, 0.04506025090813637f);
    p = fmaf(p, t, 0.07497954368591309f);
    p = fmaf(p, t, 0.16666702926158905f);
    p = fmaf(p, t, 1.0f);
    return p * x;
}

// NZ slices per thread; CPT (even) staged channels per thread per view.  The
// staged window covers rows_pad x W channels, rows_pad = 256 / TPR a power of
// two >= n_rows, W = TPR * CPT: thread t stages row t / TPR, channels
// (t % TPR) * CPT ... + CPT - 1 from CPT + 1 consecutive samples of that row.
// In LDS every run of CPT pairs is followed by one pad pair, so channel c of
// row r sits at r * TPR * (CPT + 1) + c + c / CPT: the lanes of a wave then
// write at a stride of CPT + 1 pairs (odd), which spreads them over all banks.
template <int NZ, int CPT, bool kBf16>
__global__ __launch_bounds__(256) void bp_kernel(BpArgs a) {
    static_assert(CPT % 2 == 0, "CPT must be even");
    extern __shared__ float2 lds2[];   // 2 buffers x [rows_pad][TPR * (CPT + 1)] (value, delta)
    const int tid = threadIdx.x;

    // XCD-aware remap: hardware dispatches workgroup b to XCD b % 8; give each
    // XCD a contiguous range of tiles.
    int tile = blockIdx.x;
    if ((a.n_tiles & 7) == 0) {
        const int per = a.n_tiles >> 3;
        tile = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
    }
    const int tx = tile / a.tiles_y, ty = tile - tx * a.tiles_y;
    const int xi = tx * 16 + (tid >> 4);
    const int yi = ty * 16 + (tid & 15);
    const int z0 = blockIdx.y * NZ;

    const float x = a.xs * (static_cast<float>(xi) - a.xc) + a.x0;
    const float y = a.ys * (static_cast<float>(yi) - a.yc) + a.y0;
    // centre of the tile
    const float xm = a.xs * (static_cast<float>(tx * 16) + 7.5f - a.xc) + a.x0;
    const float ym = a.ys * (static_cast<float>(ty * 16) + 7.5f - a.yc) + a.y0;

    const int W = a.tpr * CPT;                          // staged channels
    const int row_pairs = a.tpr * (CPT + 1);            // LDS pairs per staged row
    const int n_stage = 256 * (CPT + 1);                // pairs per buffer
    const unsigned stride_b = static_cast<unsigned>(row_pairs) * 8u;
    const int n_views = a.half * a.n_ht;
    const int srow = tid >> a.tpr_log2;                 // staged row of this thread
    const int scol = (tid & (a.tpr - 1)) * CPT;         // first staged channel
    const bool srow_ok = srow < a.n_rows;

    // prefetch registers for the next view's window (raw sample bits)
    uint32_t pv[CPT + 1];

    auto window_lo = [&](const BpView &vw) {
        const float pm = (xm * vw.sin_t - ym * vw.cos_t) * a.inv_ds + a.two_cc;
        return static_cast<int>(floorf(pm - a.half_w)) - 1;
    };
    auto fetch = [&](int v, int p_lo) {
        const int g0 = p_lo + scol;
        const size_t rb = (static_cast<size_t>(v) * a.n_rows + srow) * a.nos;
        if (srow_ok && p_lo >= 0 && p_lo + W < a.nos) {
            // window inside the detector: CPT + 1 consecutive samples
#pragma unroll
            for (int e = 0; e <= CPT; e++) pv[e] = load_bits<kBf16>(a, rb + g0 + e);
        } else {
#pragma unroll
            for (int e = 0; e <= CPT; e++) {
                const int g = g0 + e;
                pv[e] = (srow_ok && g >= 0 && g < a.nos) ? load_bits<kBf16>(a, rb + g) : 0u;  // +0.0 in either format
            }
        }
    };

    float s[NZ], st[NZ], ht[NZ];
#pragma unroll
    for (int z = 0; z < NZ; z++) { s[z] = 0.0f; st[z] = 0.0f; ht[z] = 0.0f; }

    // pipeline prologue: view 0 (i = 0, k = 0)
    BpView vw_next = a.vtab[0];
    int plo_next = window_lo(vw_next);
    fetch(0, plo_next);

    for (int t = 0; t < n_views; t++) {
        const int i = t / a.n_ht, k = t - i * a.n_ht;
        const BpView vw = vw_next;
        const int p_lo = plo_next;
        float2 *buf = lds2 + (t & 1) * n_stage;
        {
            float2 *dst = buf + tid * (CPT + 1);   // = srow * row_pairs + scol + scol / CPT
            float prev = bits_to_f32<kBf16>(pv[0]);
#pragma unroll
            for (int e = 0; e < CPT; e++) {
                const float nxt = bits_to_f32<kBf16>(pv[e + 1]);
                dst[e] = make_float2(prev, nxt - prev);
                prev = nxt;
            }
        }
        __syncthreads();
        if (t + 1 < n_views) {
            const int i1 = (t + 1) / a.n_ht, k1 = (t + 1) - i1 * a.n_ht;
            vw_next = a.vtab[i1 + k1 * a.half];
            plo_next = window_lo(vw_next);
            fetch(i1 + k1 * a.half, plo_next);
        }

        // per-column geometry, shared by the NZ slices:
        // q(z) = A z + B is qhat of reference include/backproject.cuh:84
        const float phat = x * vw.sin_t - y * vw.cos_t;
        const float pl = phat * a.inv_ds + a.two_cc - static_cast<float>(p_lo);
        int c0 = static_cast<int>(floorf(pl));
        const float fp = pl - static_cast<float>(c0);
        c0 = min(max(c0, 0), W - 2);
        const float xr = phat * a.inv_rf;
        float asn = asin_small(xr);
        if (__any(fabsf(xr) > 0.45f)) asn = fabsf(xr) > 0.45f ? asinf(xr) : asn;
        const float dz0 = vw.c_v + a.k_asin * asn + static_cast<float>(z0) * a.coll;
        const float lhat = sqrtf(a.rf2 - phat * phat) - x * vw.cos_t - y * vw.sin_t;
        const float g = a.neg_td * __builtin_amdgcn_rcpf(lhat * a.tanc);
        const float A = g * a.coll;
        const float B = g * dz0;
        const float Ar = A * a.hr, Br = fmaf(B, a.hr, a.hr);   // row coordinate = Ar z + Br
        const unsigned cl = static_cast<unsigned>(c0) + static_cast<unsigned>(c0) / CPT;
        const char *colb = reinterpret_cast<const char *>(buf) + 8u * cl;

#pragma unroll
        for (int z = 0; z < NZ; z++) {
            const float q = fmaf(A, static_cast<float>(z), B);
            const float aq = fabsf(q);
            const bool zv = aq < 1.0f;
            if (__any(zv)) {
                float w = zv ? 1.0f : 0.0f;
                const bool tap = zv && aq >= 0.6f;
                if (__any(tap)) {
                    const float tw = stier_taper(aq);
                    w = tap ? tw : w;
                }
                const float rowf = __builtin_amdgcn_fmed3f(fmaf(Ar, static_cast<float>(z), Br), 0.0f, a.row_max);
                const unsigned r0 = static_cast<unsigned>(rowf);
                const float fr = __builtin_amdgcn_fractf(rowf);
                const char *p = colb + __umul24(r0, stride_b);
                const float2 e0 = *reinterpret_cast<const float2 *>(p);
                const float2 e1 = *reinterpret_cast<const float2 *>(p + stride_b);
                const float top = fmaf(fp, e0.y, e0.x);
                const float bot = fmaf(fp, e1.y, e1.x);
                const float val = fmaf(fr, bot - top, top);
                st[z] = fmaf(val, w, st[z]);
                ht[z] += w;
            }
        }
        if (k == a.n_ht - 1) {
            // s += (1/h_t) s_t where h_t != 0 (reference include/backproject.cuh:104-108)
#pragma unroll
            for (int z = 0; z < NZ; z++) {
                s[z] += ht[z] != 0.0f ? __builtin_amdgcn_rcpf(ht[z]) * st[z] : 0.0f;
                st[z] = 0.0f;
                ht[z] = 0.0f;
            }
        }
        (void)i;
    }

    const bool in_fov = sqrtf(x * x + y * y) < a.acq_half;
    const size_t nxy = static_cast<size_t>(a.nx) * a.ny;
    float *o = a.out + static_cast<size_t>(xi) * a.ny + yi + static_cast<size_t>(z0) * nxy;
#pragma unroll
    for (int z = 0; z < NZ; z++) o[z * nxy] = in_fov ? s[z] * a.out_scale : 0.0f;
}

template <int NZ, int CPT>
int launch_c(const BpArgs &a, hipStream_t stream) {
    const size_t lds_bytes = 2 * 256 * (CPT + 1) * sizeof(float2);   // <= 60 KiB
    dim3 grid(a.n_tiles, 32 / NZ);
    if (a.rebin_bf16) hipLaunchKernelGGL((bp_kernel<NZ, CPT, true>), grid, dim3(256), lds_bytes, stream, a);
    else hipLaunchKernelGGL((bp_kernel<NZ, CPT, false>), grid, dim3(256), lds_bytes, stream, a);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

// Smallest supported channels-per-thread whose window W = tpr * CPT covers
// the tile footprint (2 half_w + 3 channels).
template <int NZ>
int launch_nz(BpArgs a, hipStream_t stream) {
    const float need = 2.0f * a.half_w + 3.0f;
    int rows_pad = 1;
    while (rows_pad < a.n_rows) rows_pad <<= 1;
    if (rows_pad > 256) return CTBB_ERR_SHAPE;
    a.tpr = 256 / rows_pad;
    a.tpr_log2 = 0;
    while ((1 << a.tpr_log2) < a.tpr) a.tpr_log2++;
    const int cpt = static_cast<int>(ceilf(need / static_cast<float>(a.tpr)));
    static const int kCpt[] = {2, 4, 6, 8, 10, 12, 14};
    for (int c : kCpt) {
        if (c < cpt) continue;
        switch (c) {
            case 2: return launch_c<NZ, 2>(a, stream);
            case 4: return launch_c<NZ, 4>(a, stream);
            case 6: return launch_c<NZ, 6>(a, stream);
            case 8: return launch_c<NZ, 8>(a, stream);
            case 10: return launch_c<NZ, 10>(a, stream);
            case 12: return launch_c<NZ, 12>(a, stream);
            case 14: return launch_c<NZ, 14>(a, stream);
        }
    }
    return CTBB_ERR_SHAPE;   // footprint wider than 14 channels per thread
}

}  // namespace

int launch_backproject(const BpArgs &a, hipStream_t stream) {
    if ((!a.rebin && !a.rebin_bf16) || !a.vtab || !a.out) return CTBB_ERR_ARG;
    if (a.nx % 16 || a.ny % 16 || a.nx <= 0 || a.ny <= 0) return CTBB_ERR_SHAPE;
    if (a.n_rows < 2 || a.half <= 0 || a.n_ht <= 0) return CTBB_ERR_SHAPE;
    if (a.n_tiles != (a.nx / 16) * (a.ny / 16) || a.tiles_y != a.ny / 16) return CTBB_ERR_ARG;
    if (a.nz == 32) return launch_nz<32>(a, stream);
    if (a.nz == 16) return launch_nz<16>(a, stream);
    return CTBB_ERR_ARG;
}

}  // namespace ctbb
