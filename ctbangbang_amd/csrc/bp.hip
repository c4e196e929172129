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
// column (xi, yi) with its NZ slice accumulators in registers.  For every
// view the workgroup stages the rows x WCH channel window that covers the
// tile's footprint (channels outside the detector are staged as zeros, which
// is exactly the border rule), and every thread amortises the per-column
// geometry (phat, asin, lhat) over its NZ slices.  Views are processed in the
// reference order (angle i outer, half turn k inner).
//
// Workgroups are remapped so that the tiles running on one XCD are spatially
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

template <int NZ, bool kBf16>
__global__ __launch_bounds__(256) void bp_kernel(BpArgs a) {
    extern __shared__ float lds[];
    const int tid = threadIdx.x;
    const int wch = 1 << a.wch_log2;
    const int stride = wch + 1;

    // XCD-aware remap: hardware dispatches workgroup b to XCD b % 8; give each
    // XCD a contiguous range of tiles.
    int tile = blockIdx.x;
    if ((a.n_tiles & 7) == 0) {
        int per = a.n_tiles >> 3;
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

    float s[NZ], st[NZ], ht[NZ];
#pragma unroll
    for (int z = 0; z < NZ; z++) { s[z] = 0.0f; st[z] = 0.0f; ht[z] = 0.0f; }

    const int n_stage = a.n_rows << a.wch_log2;
    const unsigned stride_b = static_cast<unsigned>(stride) * 4u;

    for (int i = 0; i < a.half; i++) {
        for (int k = 0; k < a.n_ht; k++) {
            const int v = i + k * a.half;
            const BpView vw = a.vtab[v];
            // footprint window of the tile on this view
            const float pm = (xm * vw.sin_t - ym * vw.cos_t) * a.inv_ds + a.two_cc;
            const int p_lo = static_cast<int>(floorf(pm - a.half_w)) - 1;

            // per-column geometry, shared by the NZ slices:
            // q(z) = A z + B, q = qhat of reference include/backproject.cuh:84
            const float phat = x * vw.sin_t - y * vw.cos_t;
            const float pl = phat * a.inv_ds + a.two_cc - static_cast<float>(p_lo);
            int c0 = static_cast<int>(floorf(pl));
            const float fp = pl - static_cast<float>(c0);
            c0 = min(max(c0, 0), wch - 2);
            const float asn = asinf(phat * a.inv_rf);
            const float dz0 = vw.c_v + a.k_asin * asn + static_cast<float>(z0) * a.coll;
            const float lhat = sqrtf(a.rf2 - phat * phat) - x * vw.cos_t - y * vw.sin_t;
            const float g = a.neg_td * __builtin_amdgcn_rcpf(lhat * a.tanc);
            const float A = g * a.coll;
            const float B = g * dz0;
            // W(q) = 0 for |q| >= 1: slices outside (-1, 1) add exactly nothing
            const float qe = fmaf(A, static_cast<float>(NZ - 1), B);
            const bool vis = fminf(B, qe) < 1.0f && fmaxf(B, qe) > -1.0f;

            // also the barrier that retires the previous view's LDS reads
            if (!__syncthreads_or(vis)) continue;
            const size_t vbase = static_cast<size_t>(v) * a.n_rows * a.nos;
            for (int e = tid; e < n_stage; e += 256) {
                const int r = e >> a.wch_log2;
                const int c = e & (wch - 1);
                const int gch = p_lo + c;
                float val = 0.0f;
                if (gch >= 0 && gch < a.nos) {
                    const size_t gi = vbase + static_cast<size_t>(r) * a.nos + gch;
                    val = kBf16 ? bf16_to_f32(a.rebin_bf16[gi]) : a.rebin[gi];
                }
                lds[r * stride + c] = val;
            }
            __syncthreads();

            const char *colb = reinterpret_cast<const char *>(lds) + 4u * static_cast<unsigned>(c0);
#pragma unroll
            for (int z = 0; z < NZ; z++) {
                const float q = fmaf(A, static_cast<float>(z), B);
                const float aq = fabsf(q);
                // q varies by < 1% across a wavefront: the slice is skipped when
                // no lane of the wavefront sees it, and the taper is evaluated
                // only when some lane is in 0.6 <= |q| < 1 (wave-uniform votes)
                const bool zv = vis && aq < 1.0f;
                if (__any(zv)) {
                    float w = zv ? 1.0f : 0.0f;
                    const bool tap = zv && aq >= 0.6f;
                    if (__any(tap)) {
                        const float tw = stier_taper(aq);
                        w = tap ? tw : w;
                    }
                    const float rowf = __builtin_amdgcn_fmed3f(fmaf(q, a.hr, a.hr), 0.0f, a.row_max);
                    const unsigned r0 = static_cast<unsigned>(rowf);
                    const float fr = __builtin_amdgcn_fractf(rowf);
                    const float *p0 = reinterpret_cast<const float *>(colb + __umul24(r0, stride_b));
                    const float *p1 = reinterpret_cast<const float *>(colb + __umul24(r0, stride_b) + stride_b);
                    const float v00 = p0[0], v01 = p0[1];
                    const float v10 = p1[0], v11 = p1[1];
                    const float top = fmaf(fp, v01 - v00, v00);
                    const float bot = fmaf(fp, v11 - v10, v10);
                    const float val = fmaf(fr, bot - top, top);
                    st[z] = fmaf(val, w, st[z]);
                    ht[z] += w;
                }
            }
        }
        // s += (1/h_t) s_t where h_t != 0 (reference include/backproject.cuh:104-108)
#pragma unroll
        for (int z = 0; z < NZ; z++) {
            s[z] += ht[z] != 0.0f ? __builtin_amdgcn_rcpf(ht[z]) * st[z] : 0.0f;
            st[z] = 0.0f;
            ht[z] = 0.0f;
        }
    }

    const bool in_fov = sqrtf(x * x + y * y) < a.acq_half;
    const size_t nxy = static_cast<size_t>(a.nx) * a.ny;
    float *o = a.out + static_cast<size_t>(xi) * a.ny + yi + static_cast<size_t>(z0) * nxy;
#pragma unroll
    for (int z = 0; z < NZ; z++) o[z * nxy] = in_fov ? s[z] * a.out_scale : 0.0f;
}

}  // namespace

int launch_backproject(const BpArgs &a, hipStream_t stream) {
    if ((!a.rebin && !a.rebin_bf16) || !a.vtab || !a.out) return CTBB_ERR_ARG;
    if (a.nx % 16 || a.ny % 16 || a.nx <= 0 || a.ny <= 0) return CTBB_ERR_SHAPE;
    if (a.n_rows < 2 || a.wch_log2 < 2 || a.wch_log2 > 8 || a.half <= 0 || a.n_ht <= 0) return CTBB_ERR_SHAPE;
    if (a.n_tiles != (a.nx / 16) * (a.ny / 16) || a.tiles_y != a.ny / 16) return CTBB_ERR_ARG;
    const size_t lds_bytes = static_cast<size_t>(a.n_rows) * ((1 << a.wch_log2) + 1) * sizeof(float);
    if (lds_bytes > 64 * 1024) return CTBB_ERR_SHAPE;
    if (a.nz != 16 && a.nz != 32) return CTBB_ERR_ARG;
    dim3 grid(a.n_tiles, 32 / a.nz);
    if (a.nz == 32) {
        if (a.rebin_bf16) hipLaunchKernelGGL((bp_kernel<32, true>), grid, dim3(256), lds_bytes, stream, a);
        else hipLaunchKernelGGL((bp_kernel<32, false>), grid, dim3(256), lds_bytes, stream, a);
    } else {
        if (a.rebin_bf16) hipLaunchKernelGGL((bp_kernel<16, true>), grid, dim3(256), lds_bytes, stream, a);
        else hipLaunchKernelGGL((bp_kernel<16, false>), grid, dim3(256), lds_bytes, stream, a);
    }
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
