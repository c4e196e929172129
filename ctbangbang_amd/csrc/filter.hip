// Ramp / apodisation filter of rebinned rows as a Toeplitz matrix product on
// the CDNA4 matrix cores:
//     Y[m][l] = sum_k X[m][k] * taps[n + l - k]        (k, l in [0, n))
// which is the convolution of reference include/rebin_filter.cuh:369-409
// (and of the FFT form in src/rebin_filter_cpu.cu:647-709).
//
// The Toeplitz operand is never materialised: every workgroup keeps the 2n
// taps in LDS and forms the B fragments from them on the fly
// (B[k][l] = taps[n + l - k]).
//
// Variants (ctbb_filter_mode):
//   F32     v_mfma_f32_16x16x4_f32, FP32 operands, FP32 accumulation
//   BF16X3  v_mfma_f32_16x16x16_bf16 on hi/lo bfloat16 splits of both
//           operands: X*T ~= Xh*Th + Xh*Tl + Xl*Th, FP32 accumulation
//   F16X3   the same with float16 splits; taps are pre-scaled by 2^10 (and
//           the result by 2^-10) to keep the low parts out of the subnormals
// Tile: 64 rows x 64 columns per workgroup (4 waves, 2x2 of 32x32 per wave),
// K chunks of 32 staged in LDS.  Requires n % 64 == 0 (1472 and 1344 for the
// 736- and 672-channel scanners); other widths use a VALU kernel.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctbb {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int A_STRIDE = BK + 4;   // 36 floats: 16-byte aligned rows

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

template <int MODE>
__global__ __launch_bounds__(256) void toeplitz_kernel(FilterArgs a) {
    extern __shared__ float smem[];
    const int n = a.n;
    float *As = smem;                          // [BM][A_STRIDE]
    float *Tf = smem + BM * A_STRIDE;          // [2n] taps (F32 mode)
    uint16_t *Th = reinterpret_cast<uint16_t *>(Tf);         // [2n] hi (split modes)
    uint16_t *Tl = Th + 2 * n;                                // [2n] lo

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const size_t m0 = static_cast<size_t>(blockIdx.y) * BM;
    const int n0 = blockIdx.x * BN;
    const float tscale = MODE == CTBB_FILTER_F16X3 ? 1024.0f : 1.0f;

    // taps -> LDS (split into hi/lo for the split modes)
    for (int j = tid; j < 2 * n; j += 256) {
        float t = a.taps[j] * tscale;
        if (MODE == CTBB_FILTER_F32) {
            Tf[j] = t;
        } else if (MODE == CTBB_FILTER_BF16X3) {
            uint16_t h = f32_to_bf16_rne(t);
            Th[j] = h;
            Tl[j] = f32_to_bf16_rne(t - bf16_to_f32(h));
        } else {
            _Float16 h = static_cast<_Float16>(t);
            _Float16 l = static_cast<_Float16>(t - static_cast<float>(h));
            Th[j] = __builtin_bit_cast(uint16_t, h);
            Tl[j] = __builtin_bit_cast(uint16_t, l);
        }
    }

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int lr = lane & 15, lk = lane >> 4;
    const float *X = a.in;

    for (int k0 = 0; k0 < n; k0 += BK) {
        __syncthreads();
        // A tile: 64 rows x 32 floats, 8 per thread (two float4)
        {
            const int row = tid >> 2, seg = (tid & 3) * 8;
            const size_t gm = m0 + row;
            float4 v0 = make_float4(0.f, 0.f, 0.f, 0.f), v1 = v0;
            if (gm < a.rows) {
                const float4 *src = reinterpret_cast<const float4 *>(X + gm * n + k0 + seg);
                v0 = src[0];
                v1 = src[1];
            }
            float4 *dst = reinterpret_cast<float4 *>(As + row * A_STRIDE + seg);
            dst[0] = v0;
            dst[1] = v1;
        }
        __syncthreads();

        if (MODE == CTBB_FILTER_F32) {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 4) {
                float af[2], bfv[2];
#pragma unroll
                for (int rt = 0; rt < 2; rt++) af[rt] = As[(wr * 32 + rt * 16 + lr) * A_STRIDE + kk + lk];
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    const int col = n0 + wc * 32 + ct * 16 + lr;
                    bfv[ct] = Tf[n + col - (k0 + kk + lk)];
                }
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
#pragma unroll
                    for (int ct = 0; ct < 2; ct++)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(af[rt], bfv[ct], acc[rt][ct], 0, 0, 0);
            }
        } else {
#pragma unroll
            for (int kk = 0; kk < BK; kk += 16) {
                i16x4 ah[2], al[2], bh[2], bl[2];
#pragma unroll
                for (int rt = 0; rt < 2; rt++) {
                    const float4 x4 = *reinterpret_cast<const float4 *>(As + (wr * 32 + rt * 16 + lr) * A_STRIDE + kk + 4 * lk);
                    const float xs[4] = {x4.x, x4.y, x4.z, x4.w};
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        uint16_t h, l;
                        if (MODE == CTBB_FILTER_BF16X3) {
                            h = f32_to_bf16_rne(xs[e]);
                            l = f32_to_bf16_rne(xs[e] - bf16_to_f32(h));
                        } else {
                            _Float16 hh = static_cast<_Float16>(xs[e]);
                            _Float16 ll = static_cast<_Float16>(xs[e] - static_cast<float>(hh));
                            h = __builtin_bit_cast(uint16_t, hh);
                            l = __builtin_bit_cast(uint16_t, ll);
                        }
                        ah[rt][e] = static_cast<short>(h);
                        al[rt][e] = static_cast<short>(l);
                    }
                }
#pragma unroll
                for (int ct = 0; ct < 2; ct++) {
                    const int col = n0 + wc * 32 + ct * 16 + lr;
                    const int base = n + col - (k0 + kk + 4 * lk);
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        bh[ct][e] = static_cast<short>(Th[base - e]);
                        bl[ct][e] = static_cast<short>(Tl[base - e]);
                    }
                }
#pragma unroll
                for (int rt = 0; rt < 2; rt++)
#pragma unroll
                    for (int ct = 0; ct < 2; ct++) {
                        if (MODE == CTBB_FILTER_BF16X3) {
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(al[rt], bh[ct], acc[rt][ct], 0, 0, 0);
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[rt], bl[ct], acc[rt][ct], 0, 0, 0);
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(ah[rt], bh[ct], acc[rt][ct], 0, 0, 0);
                        } else {
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16f16(
                                __builtin_bit_cast(f16x4, al[rt]), __builtin_bit_cast(f16x4, bh[ct]), acc[rt][ct], 0, 0, 0);
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16f16(
                                __builtin_bit_cast(f16x4, ah[rt]), __builtin_bit_cast(f16x4, bl[ct]), acc[rt][ct], 0, 0, 0);
                            acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x16f16(
                                __builtin_bit_cast(f16x4, ah[rt]), __builtin_bit_cast(f16x4, bh[ct]), acc[rt][ct], 0, 0, 0);
                        }
                    }
            }
        }
    }

    // D layout of 16x16 MFMA tiles: lane l, register e -> row 4*(l/16)+e, column l%16
    const float oscale = 1.0f / tscale;
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const size_t gm = m0 + wr * 32 + rt * 16 + 4 * lk + e;
                const int col = n0 + wc * 32 + ct * 16 + lr;
                if (gm < a.rows) {
                    const float y = acc[rt][ct][e] * oscale;
                    if (a.out_bf16)
                        reinterpret_cast<uint16_t *>(a.out)[gm * n + col] = f32_to_bf16_rne(y);
                    else
                        reinterpret_cast<float *>(a.out)[gm * n + col] = y;
                }
            }
}

// Direct FP32 convolution on the vector ALUs for widths that are not a
// multiple of 64: one workgroup per row, taps and row staged in LDS.
__global__ __launch_bounds__(256) void conv_valu_kernel(FilterArgs a) {
    extern __shared__ float smem[];
    const int n = a.n;
    float *xr = smem;          // [n]
    float *tp = smem + n;      // [2n]
    const size_t row = blockIdx.x;
    for (int j = threadIdx.x; j < 2 * n; j += blockDim.x) tp[j] = a.taps[j];
    for (int j = threadIdx.x; j < n; j += blockDim.x) xr[j] = a.in[row * n + j];
    __syncthreads();
    for (int l = threadIdx.x; l < n; l += blockDim.x) {
        float acc = 0.0f;
        for (int k = 0; k < n; k++) acc = fmaf(tp[n + l - k], xr[k], acc);
        if (a.out_bf16)
            reinterpret_cast<uint16_t *>(a.out)[row * n + l] = f32_to_bf16_rne(acc);
        else
            reinterpret_cast<float *>(a.out)[row * n + l] = acc;
    }
}

}  // namespace

int launch_filter(const FilterArgs &a, hipStream_t stream) {
    if (!a.in || !a.out || !a.taps || a.n <= 0) return CTBB_ERR_ARG;
    if (a.rows == 0) return CTBB_OK;
    if (a.in == a.out && (a.out_bf16 || a.n % 64 == 0)) return CTBB_ERR_ARG;  // tiled kernel reads rows after writing others
    if (a.n % 64 != 0) {
        if (a.rows > 0x7fffffffu) return CTBB_ERR_SHAPE;
        size_t lds = 3 * static_cast<size_t>(a.n) * sizeof(float);
        if (lds > 64 * 1024) return CTBB_ERR_SHAPE;
        hipLaunchKernelGGL(conv_valu_kernel, dim3(static_cast<unsigned>(a.rows)), dim3(256), lds, stream, a);
        return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
    }
    const size_t gy = (a.rows + BM - 1) / BM;
    const size_t lds = BM * A_STRIDE * sizeof(float) +
                       (a.mode == CTBB_FILTER_F32 ? 2 * a.n * sizeof(float) : 4 * a.n * sizeof(uint16_t));
    if (lds > 64 * 1024) return CTBB_ERR_SHAPE;
    // grid.y may exceed 65535 for very tall inputs: split into launches
    const size_t max_y = 65535;
    for (size_t y0 = 0; y0 < gy; y0 += max_y) {
        FilterArgs s = a;
        const size_t ny = (gy - y0 < max_y) ? gy - y0 : max_y;
        s.in = a.in + y0 * BM * a.n;
        if (a.out_bf16) s.out = reinterpret_cast<uint16_t *>(a.out) + y0 * BM * a.n;
        else s.out = reinterpret_cast<float *>(a.out) + y0 * BM * a.n;
        s.rows = a.rows - y0 * BM < ny * BM ? a.rows - y0 * BM : ny * BM;
        dim3 grid(a.n / BN, static_cast<unsigned>(ny));
        switch (a.mode) {
            case CTBB_FILTER_F32: hipLaunchKernelGGL(toeplitz_kernel<CTBB_FILTER_F32>, grid, dim3(256), lds, stream, s); break;
            case CTBB_FILTER_BF16X3: hipLaunchKernelGGL(toeplitz_kernel<CTBB_FILTER_BF16X3>, grid, dim3(256), lds, stream, s); break;
            case CTBB_FILTER_F16X3: hipLaunchKernelGGL(toeplitz_kernel<CTBB_FILTER_F16X3>, grid, dim3(256), lds, stream, s); break;
            default: return CTBB_ERR_ARG;
        }
        if (hipGetLastError() != hipSuccess) return CTBB_ERR_HIP;
    }
    return CTBB_OK;
}

}  // namespace ctbb
