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
//   BF16X3  v_mfma_f32_32x32x16_bf16 on hi/lo bfloat16 splits of both
//           operands: X*T ~= Xh*Th + Xh*Tl + Xl*Th, FP32 accumulation
//   F16X3   the same with float16 splits; taps are pre-scaled by 2^10 (and
//           the result by 2^-10) to keep the low parts out of the subnormals
//   FFT32   FP32 4096-point FFT convolution (the form of the reference's CPU
//           path), for n <= 2048; wider rows fall back to F32
// Toeplitz tiles are 64 (F32) or 256 (split forms) rows x 64 columns; they
// require n % 64 == 0 (1472 and 1344 for the 736- and 672-channel scanners);
// other widths use a VALU kernel.
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>

#include "kernels.h"

namespace ctbb {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef short i16x4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 32;
constexpr int A_STRIDE = BK + 4;   // 36 floats: 16-byte aligned rows

__device__ __forceinline__ uint16_t f32_to_bf16_rne(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}

// FP32 Toeplitz product on v_mfma_f32_16x16x4_f32.
__global__ __launch_bounds__(256) void toeplitz_f32_kernel(FilterArgs a) {
    extern __shared__ float smem[];
    const int n = a.n;
    float *As = smem;                          // [BM][A_STRIDE]
    float *Tf = smem + BM * A_STRIDE;          // [2n] taps

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const size_t m0 = static_cast<size_t>(blockIdx.y) * BM;
    const int n0 = blockIdx.x * BN;

    for (int j = tid; j < 2 * n; j += 256) Tf[j] = a.taps[j];

    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    const int lr = lane & 15, lk = lane >> 4;
    const int arow = tid >> 2, aseg = (tid & 3) * 8;
    const size_t gm = m0 + arow;
    const bool row_ok = gm < a.rows;
    const float *xrow = a.in + gm * n + aseg;
    float4 x0 = make_float4(0.f, 0.f, 0.f, 0.f), x1 = x0;
    if (row_ok) {
        x0 = reinterpret_cast<const float4 *>(xrow)[0];
        x1 = reinterpret_cast<const float4 *>(xrow)[1];
    }

    for (int k0 = 0; k0 < n; k0 += BK) {
        __syncthreads();
        {
            float4 *dst = reinterpret_cast<float4 *>(As + arow * A_STRIDE + aseg);
            dst[0] = x0;
            dst[1] = x1;
        }
        __syncthreads();
        // prefetch the next K chunk of this thread's row while the MFMAs run
        if (k0 + BK < n && row_ok) {
            x0 = reinterpret_cast<const float4 *>(xrow + k0 + BK)[0];
            x1 = reinterpret_cast<const float4 *>(xrow + k0 + BK)[1];
        }
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
    }

    // D layout of 16x16 MFMA tiles: lane l, register e -> row 4*(l/16)+e, column l%16
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const size_t om = m0 + wr * 32 + rt * 16 + 4 * lk + e;
                const int col = n0 + wc * 32 + ct * 16 + lr;
                if (om < a.rows) {
                    const float y = acc[rt][ct][e];
                    if (a.out_bf16)
                        reinterpret_cast<uint16_t *>(a.out)[om * n + col] = f32_to_bf16_rne(y);
                    else
                        reinterpret_cast<float *>(a.out)[om * n + col] = y;
                }
            }
}

// Split-precision Toeplitz product: X = Xh + Xl and T = Th + Tl in bfloat16
// (MODE BF16X3) or float16 (MODE F16X3, taps scaled by 2^10 to keep the low
// parts out of the subnormals), accumulated in FP32 as Xl*Th + Xh*Tl + Xh*Th
// on gfx950's v_mfma_f32_32x32x16_{bf16,f16} (measured on MI355X at about
// 1.7x the rate of the 16x16x32 forms).  Workgroup = 256 rows x 64 columns;
// each wave owns 64 rows x 64 columns as 2 x 2 tiles of 32 x 32.
// The A operand goes straight from global memory into MFMA fragments: K is
// reduced in chunks of 32, and within a chunk lane l (half h = l / 32) holds
// k = 16 h + 8 s + e, e = 0..7, for k-step s, so its two A fragments are the
// 16 consecutive samples of its row (four float4 loads), split into hi/lo
// in registers and used for both column tiles; the next chunk's loads are
// issued before the current chunk's MFMAs.  Only the reversed taps live in
// LDS (four copies shifted by 0..3 so that every B fragment, 8 consecutive
// reversed taps, is two 8-byte reads), written once, so the K loop has no
// barriers.  D layout (32 x 32): lane l, register r -> row 8 (r / 4) + 4 h +
// r % 4, column l % 32.  Workgroups are ordered so that the n / 64 column
// tiles of one 256-row band run on the same XCD and share its L2 copy of the
// band.
template <int MODE>
__global__ __launch_bounds__(256) void toeplitz_split_kernel(FilterArgs a, int n_rt, int n_ct) {
    constexpr int TM = 256, TN = 64, TK = 32;
    typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
    typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
    typedef short i16x8 __attribute__((ext_vector_type(8)));
    typedef float f32x16 __attribute__((ext_vector_type(16)));
    extern __shared__ uint16_t sh[];
    const int n = a.n;
    const int tw = n + TN + 8;                 // window length per copy
    uint16_t *Gh = sh;                         // [4][tw]
    uint16_t *Gl = Gh + 4 * tw;                // [4][tw]

    int tile = blockIdx.x;
    const int n_tiles = n_rt * n_ct;
    if ((n_tiles & 7) == 0) tile = (blockIdx.x & 7) * (n_tiles >> 3) + (blockIdx.x >> 3);
    const int rt_i = tile / n_ct, ct_i = tile - rt_i * n_ct;
    const size_t m0 = static_cast<size_t>(rt_i) * TM;
    const int n0 = ct_i * TN;

    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int l32 = lane & 31, h = lane >> 5;
    const int jb = n - 1 - (n0 + TN - 1);      // first reversed-tap index this tile needs
    const float tscale = MODE == CTBB_FILTER_F16X3 ? 1024.0f : 1.0f;

    auto split = [&](float v, uint16_t &hi, uint16_t &lo) {
        if (MODE == CTBB_FILTER_BF16X3) {
            const __bf16 hb = static_cast<__bf16>(v);
            hi = __builtin_bit_cast(uint16_t, hb);
            lo = __builtin_bit_cast(uint16_t, static_cast<__bf16>(v - static_cast<float>(hb)));
        } else {
            const _Float16 hf = static_cast<_Float16>(v);
            hi = __builtin_bit_cast(uint16_t, hf);
            lo = __builtin_bit_cast(uint16_t, static_cast<_Float16>(v - static_cast<float>(hf)));
        }
    };

    for (int e = tid; e < 4 * tw; e += 256) {
        const int sft = e / tw, m = e - sft * tw;
        const int j = jb + m + sft;           // reversed-tap index
        const float t = (j >= 0 && j < 2 * n) ? a.taps[2 * n - 1 - j] * tscale : 0.0f;
        split(t, Gh[e], Gl[e]);
    }
    __syncthreads();

    f32x16 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++)
#pragma unroll
            for (int r = 0; r < 16; r++) acc[i][j][r] = 0.0f;

    // fragment rows of this lane; rows past the end read the last row (their
    // results are not stored)
    const float *xrow[2];
#pragma unroll
    for (int rt = 0; rt < 2; rt++) {
        size_t gm = m0 + wave * 64 + rt * 32 + l32;
        if (gm >= a.rows) gm = a.rows - 1;
        xrow[rt] = a.in + gm * n + 16 * h;
    }
    auto load = [&](float4 (&x)[2][4], int k0) {
#pragma unroll
        for (int rt = 0; rt < 2; rt++)
#pragma unroll
            for (int q = 0; q < 4; q++) x[rt][q] = *reinterpret_cast<const float4 *>(xrow[rt] + k0 + 4 * q);
    };
    auto mma = [&](const float4 (&x)[2][4], int k0) {
#pragma unroll
        for (int s = 0; s < 2; s++) {
            i16x8 ah[2], al[2];
#pragma unroll
            for (int rt = 0; rt < 2; rt++) {
                const float4 u = x[rt][2 * s], w = x[rt][2 * s + 1];
                const float v[8] = {u.x, u.y, u.z, u.w, w.x, w.y, w.z, w.w};
#pragma unroll
                for (int e = 0; e < 8; e++) {
                    uint16_t hi, lo;
                    split(v[e], hi, lo);
                    ah[rt][e] = static_cast<short>(hi);
                    al[rt][e] = static_cast<short>(lo);
                }
            }
#pragma unroll
            for (int ct = 0; ct < 2; ct++) {
                const int col = n0 + ct * 32 + l32;
                const int mm = (n - 1 - col + k0 + 16 * h + 8 * s) - jb;
                const int sft = mm & 3;
                const uint16_t *ph = Gh + sft * tw + (mm - sft);
                const uint16_t *pl = Gl + sft * tw + (mm - sft);
                const i16x4 h0 = *reinterpret_cast<const i16x4 *>(ph), h1 = *reinterpret_cast<const i16x4 *>(ph + 4);
                const i16x4 l0 = *reinterpret_cast<const i16x4 *>(pl), l1 = *reinterpret_cast<const i16x4 *>(pl + 4);
                const i16x8 bh = __builtin_shufflevector(h0, h1, 0, 1, 2, 3, 4, 5, 6, 7);
                const i16x8 bl = __builtin_shufflevector(l0, l1, 0, 1, 2, 3, 4, 5, 6, 7);
#pragma unroll
                for (int rt = 0; rt < 2; rt++) {
                    if (MODE == CTBB_FILTER_BF16X3) {
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, al[rt]), __builtin_bit_cast(bf16x8, bh), acc[rt][ct], 0, 0, 0);
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ah[rt]), __builtin_bit_cast(bf16x8, bl), acc[rt][ct], 0, 0, 0);
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            __builtin_bit_cast(bf16x8, ah[rt]), __builtin_bit_cast(bf16x8, bh), acc[rt][ct], 0, 0, 0);
                    } else {
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(f16x8, al[rt]), __builtin_bit_cast(f16x8, bh), acc[rt][ct], 0, 0, 0);
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(f16x8, ah[rt]), __builtin_bit_cast(f16x8, bl), acc[rt][ct], 0, 0, 0);
                        acc[rt][ct] = __builtin_amdgcn_mfma_f32_32x32x16_f16(
                            __builtin_bit_cast(f16x8, ah[rt]), __builtin_bit_cast(f16x8, bh), acc[rt][ct], 0, 0, 0);
                    }
                }
            }
        }
    };

    // n % 64 == 0, so the chunks come in pairs
    float4 xa[2][4], xb[2][4];
    load(xa, 0);
    for (int k0 = 0; k0 < n; k0 += 2 * TK) {
        load(xb, k0 + TK);
        mma(xa, k0);
        if (k0 + 2 * TK < n) load(xa, k0 + 2 * TK);
        mma(xb, k0 + TK);
    }

    const float oscale = 1.0f / tscale;
#pragma unroll
    for (int rt = 0; rt < 2; rt++)
#pragma unroll
        for (int ct = 0; ct < 2; ct++)
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const size_t om = m0 + wave * 64 + rt * 32 + 8 * (r / 4) + 4 * h + (r % 4);
                const int col = n0 + ct * 32 + l32;
                if (om < a.rows) {
                    const float y = acc[rt][ct][r] * oscale;
                    if (a.out_bf16)
                        reinterpret_cast<uint16_t *>(a.out)[om * n + col] = f32_to_bf16_rne(y);
                    else
                        reinterpret_cast<float *>(a.out)[om * n + col] = y;
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

// ---------------------------------------------------------------------------
// FP32 FFT form of the same convolution (CTBB_FILTER_FFT32).
// Embedding: with L = 4096 >= 2n, h[m] = taps[m] for m < 2n and 0 above, the
// circular convolution y = x (*) h of the zero-padded row gives
// y[n + l] = sum_k x[k] taps[n + l - k] for l in [0, n) without wrap-around
// (n + l - k lies in [1, 2n - 1]).  Two rows a, b travel as one complex
// signal a + i b: h is real, so y = (a (*) h) + i (b (*) h).
// The host supplies H = DFT_L(h) / L and the twiddles e^(-2 pi i t / L),
// both evaluated in FP64 and rounded once.  L = 3072 where 2n <= 3072 (the
// shipped scanners: 1472 and 1344 oversampled channels), else 4096.
// Transforms (Stockham, the passes in between through LDS with one pad
// entry per 16 to spread the banks; the inverse is conj(DFT(conj(Y)))):
//   L = 4096: radix 16, 16, 16, each pass 256 threads x 16 points;
//   L = 3072: forward radix 12, 16, 16 (256 threads x 12 points, then
//   192 x 16), the inverse in the reverse order 16, 16, 12, so each
//   transform starts from the layout the previous one ended in: 25% fewer
//   VALU instructions than L = 4096 for the same rows.
constexpr int FFT_L = 4096;
constexpr int FFT_LDS = FFT_L + FFT_L / 16;
constexpr int FFT_L3 = 3072;
constexpr int FFT_LDS3 = FFT_L3 + FFT_L3 / 8;

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// in-place 4-point DFT, forward sign: X1 = a0 - i a1 - a2 + i a3
__device__ __forceinline__ void dft4(float2 &a0, float2 &a1, float2 &a2, float2 &a3) {
    const float2 s02 = cadd(a0, a2), d02 = csub(a0, a2), s13 = cadd(a1, a3), d13 = csub(a1, a3);
    a0 = cadd(s02, s13);
    a2 = csub(s02, s13);
    a1 = make_float2(d02.x + d13.y, d02.y - d13.x);
    a3 = make_float2(d02.x - d13.y, d02.y + d13.x);
}

// 16-point DFT in natural order, as 4 x 4: inner DFTs over r2 (r = r1 + 4 r2),
// twiddles W16^(r1 k2), outer DFTs over r1; X[k2 + 4 k1] ends in slot k1 + 4 k2.
__device__ __forceinline__ void fft16(float2 (&v)[16]) {
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f, H2 = 0.70710678118654757f;
#pragma unroll
    for (int r1 = 0; r1 < 4; r1++) dft4(v[r1], v[r1 + 4], v[r1 + 8], v[r1 + 12]);
    // W16^m = (cos(pi m / 8), -sin(pi m / 8))
    v[5] = cmul(v[5], make_float2(C1, -S1));     // r1 1, k2 1: m 1
    v[9] = cmul(v[9], make_float2(H2, -H2));     // r1 1, k2 2: m 2
    v[13] = cmul(v[13], make_float2(S1, -C1));   // r1 1, k2 3: m 3
    v[6] = cmul(v[6], make_float2(H2, -H2));     // r1 2, k2 1: m 2
    v[10] = make_float2(v[10].y, -v[10].x);      // r1 2, k2 2: m 4 (-i)
    v[14] = cmul(v[14], make_float2(-H2, -H2));  // r1 2, k2 3: m 6
    v[7] = cmul(v[7], make_float2(S1, -C1));     // r1 3, k2 1: m 3
    v[11] = cmul(v[11], make_float2(-H2, -H2));  // r1 3, k2 2: m 6
    v[15] = cmul(v[15], make_float2(-C1, S1));   // r1 3, k2 3: m 9
#pragma unroll
    for (int k2 = 0; k2 < 4; k2++) dft4(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
    // slot k1 + 4 k2 holds X[k2 + 4 k1]: transpose to natural order
    float2 t;
    t = v[1]; v[1] = v[4]; v[4] = t;
    t = v[2]; v[2] = v[8]; v[8] = t;
    t = v[3]; v[3] = v[12]; v[12] = t;
    t = v[6]; v[6] = v[9]; v[9] = t;
    t = v[7]; v[7] = v[13]; v[13] = t;
    t = v[11]; v[11] = v[14]; v[14] = t;
}

// in-place 3-point DFT, forward sign: X1 = a0 + w a1 + w^2 a2, w = e^(-2 pi i / 3)
__device__ __forceinline__ void dft3(float2 &a0, float2 &a1, float2 &a2) {
    constexpr float S3 = 0.86602540378443865f;   // sin(pi / 3)
    const float2 t = cadd(a1, a2);
    const float2 d = make_float2(S3 * (a1.x - a2.x), S3 * (a1.y - a2.y));
    const float2 c = make_float2(fmaf(-0.5f, t.x, a0.x), fmaf(-0.5f, t.y, a0.y));
    a0 = cadd(a0, t);
    a1 = make_float2(c.x + d.y, c.y - d.x);     // c - i d
    a2 = make_float2(c.x - d.y, c.y + d.x);     // c + i d
}

// 12-point DFT of v[0 .. 11] in natural order, as 3 x 4: inner 3-point DFTs
// over r2 (r = r1 + 4 r2), twiddles W12^(r1 k2), outer 4-point DFTs over r1;
// X[k2 + 3 k1] ends in slot k1 + 4 k2.
__device__ __forceinline__ void fft12(float2 (&v)[16]) {
    constexpr float S3 = 0.86602540378443865f;
#pragma unroll
    for (int r1 = 0; r1 < 4; r1++) dft3(v[r1], v[r1 + 4], v[r1 + 8]);
    // W12^m = (cos(pi m / 6), -sin(pi m / 6))
    v[5] = cmul(v[5], make_float2(S3, -0.5f));     // r1 1, k2 1: m 1
    v[9] = cmul(v[9], make_float2(0.5f, -S3));     // r1 1, k2 2: m 2
    v[6] = cmul(v[6], make_float2(0.5f, -S3));     // r1 2, k2 1: m 2
    v[10] = cmul(v[10], make_float2(-0.5f, -S3));  // r1 2, k2 2: m 4
    v[7] = make_float2(v[7].y, -v[7].x);           // r1 3, k2 1: m 3 (-i)
    v[11] = make_float2(-v[11].x, -v[11].y);       // r1 3, k2 2: m 6 (-1)
#pragma unroll
    for (int k2 = 0; k2 < 3; k2++) dft4(v[4 * k2], v[4 * k2 + 1], v[4 * k2 + 2], v[4 * k2 + 3]);
    // natural order: X[k] sits in slot k / 3 + 4 (k % 3)
    float2 o[12];
#pragma unroll
    for (int k = 0; k < 12; k++) o[k] = v[k / 3 + 4 * (k % 3)];
#pragma unroll
    for (int k = 0; k < 12; k++) v[k] = o[k];
}

__device__ __forceinline__ int fpad(int i) { return i + (i >> 4); }
// (the 3072-point passes: one pad entry per 8, measured 2% faster than one
// per 16 for their stride-12 and stride-16 stores)
__device__ __forceinline__ int fpad3(int i) { return i + (i >> 3); }

// v[r] *= w^r for r = 1 .. 15, from b[m] = w^(2^m), m = 0..3: the other
// powers are products of at most three of them (each product rounds once,
// so the twiddles stay within a few FP32 ulps of the table values).
// (R = 12: r = 1 .. 11 only)
template <int R = 16>
__device__ __forceinline__ void twiddle16(float2 (&v)[16], const float2 (&b)[4]) {
    static_assert(R == 12 || R == 16, "radix 12 or 16");
    const float2 w3 = cmul(b[1], b[0]);
    const float2 w5 = cmul(b[2], b[0]);
    const float2 w6 = cmul(b[2], b[1]);
    const float2 w7 = cmul(b[2], w3);
    v[1] = cmul(v[1], b[0]);
    v[2] = cmul(v[2], b[1]);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], b[2]);
    v[5] = cmul(v[5], w5);
    v[6] = cmul(v[6], w6);
    v[7] = cmul(v[7], w7);
    v[8] = cmul(v[8], b[3]);
    v[9] = cmul(v[9], cmul(b[3], b[0]));
    v[10] = cmul(v[10], cmul(b[3], b[1]));
    v[11] = cmul(v[11], cmul(b[3], w3));
    if (R == 12) return;
    v[12] = cmul(v[12], cmul(b[3], b[2]));
    v[13] = cmul(v[13], cmul(b[3], w5));
    v[14] = cmul(v[14], cmul(b[3], w6));
    v[15] = cmul(v[15], cmul(b[3], w7));
}

// Forward 4096-point DFT.  On entry v[r] = x[j + 256 r]; on exit
// v[r] = X[j + 256 r].  tw: e^(-2 pi i t / 4096), t in [0, 4096).  Each
// pass's twiddles are w^r of one w per thread; its four power-of-two powers
// are requested one pass ahead (before the LDS exchange) and the other
// eleven are formed by products.
__device__ __forceinline__ void fft4096(float2 (&v)[16], float2 *sh, int j, const float2 *__restrict__ tw) {
    float2 b[4];
#pragma unroll
    for (int m = 0; m < 4; m++) b[m] = tw[(16 << m) * (j & 15)];
    fft16(v);                                           // pass Ns = 1: no twiddles
    __syncthreads();                                    // earlier readers of sh are done
#pragma unroll
    for (int r = 0; r < 16; r++) sh[fpad(16 * j + r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = sh[fpad(j + 256 * r)];
    twiddle16(v, b);
#pragma unroll
    for (int m = 0; m < 4; m++) b[m] = tw[j << m];
    fft16(v);                                           // pass Ns = 16
    __syncthreads();
    const int base = (j >> 4) * 256 + (j & 15);
#pragma unroll
    for (int r = 0; r < 16; r++) sh[fpad(base + 16 * r)] = v[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 16; r++) v[r] = sh[fpad(j + 256 * r)];
    twiddle16(v, b);
    fft16(v);                                           // pass Ns = 256
}

// Forward 3072-point DFT, radix 12, 16, 16.  On entry (every thread)
// v[r] = x[j + 256 r], r < 12; on exit (threads j < 192) v[r] = X[j + 192 r],
// r < 16.  tw: e^(-2 pi i t / 3072).  Stockham: the pass of radix R and
// stride Ns takes butterfly j from x[j + r N / R], turns it by
// w^r, w = e^(-2 pi i (j mod Ns) / (Ns R)), and leaves output r at
// (j / Ns) Ns R + j mod Ns + r Ns.
__device__ __forceinline__ void fft3072(float2 (&v)[16], float2 *sh, int j, const float2 *__restrict__ tw) {
    const bool act = j < 192;               // butterflies of the radix-16 passes
    float2 b[4];
#pragma unroll
    for (int m = 0; m < 4; m++) b[m] = tw[act ? (16 * (j % 12)) << m : 0];
    fft12(v);                                           // R 12, Ns 1
    __syncthreads();                                    // earlier readers of sh are done
#pragma unroll
    for (int r = 0; r < 12; r++) sh[fpad3(12 * j + r)] = v[r];
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = sh[fpad3(j + 192 * r)];
        twiddle16(v, b);
#pragma unroll
        for (int m = 0; m < 4; m++) b[m] = tw[j << m];
        fft16(v);                                       // R 16, Ns 12
    }
    __syncthreads();
    if (act) {
        const int base = (j / 12) * 192 + j % 12;
#pragma unroll
        for (int r = 0; r < 16; r++) sh[fpad3(base + 12 * r)] = v[r];
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = sh[fpad3(j + 192 * r)];
        twiddle16(v, b);
        fft16(v);                                       // R 16, Ns 192
    }
}

// The same DFT with the radices in reverse order (16, 16, 12): on entry
// (threads j < 192) v[r] = x[j + 192 r], r < 16; on exit (every thread)
// v[r] = X[j + 256 r], r < 12.
__device__ __forceinline__ void fft3072_rev(float2 (&v)[16], float2 *sh, int j, const float2 *__restrict__ tw) {
    const bool act = j < 192;
    float2 b[4];
#pragma unroll
    for (int m = 0; m < 4; m++) b[m] = tw[act ? (12 * (j & 15)) << m : 0];
    if (act) fft16(v);                                  // R 16, Ns 1
    __syncthreads();                                    // earlier readers of sh are done
    if (act) {
#pragma unroll
        for (int r = 0; r < 16; r++) sh[fpad3(16 * j + r)] = v[r];
    }
    __syncthreads();
    if (act) {
#pragma unroll
        for (int r = 0; r < 16; r++) v[r] = sh[fpad3(j + 192 * r)];
        twiddle16(v, b);
        fft16(v);                                       // R 16, Ns 16
    }
#pragma unroll
    for (int m = 0; m < 4; m++) b[m] = tw[j << m];
    __syncthreads();
    if (act) {
        const int base = (j >> 4) * 256 + (j & 15);
#pragma unroll
        for (int r = 0; r < 16; r++) sh[fpad3(base + 16 * r)] = v[r];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 12; r++) v[r] = sh[fpad3(j + 256 * r)];
    twiddle16<12>(v, b);
    fft12(v);                                           // R 12, Ns 256
}

// L = 3072 form of fft_filter_kernel (2n <= 3072); spec = [H (3072)]
// followed by [twiddles (3072)].
__global__ __launch_bounds__(256) void fft3072_filter_kernel(FilterArgs a) {
    __shared__ float2 sh[FFT_LDS3];
    const int j = threadIdx.x;
    const int n = a.n;
    const float2 *__restrict__ H = a.spec;
    const float2 *__restrict__ tw = a.spec + FFT_L3;
    const size_t ra = 2 * static_cast<size_t>(blockIdx.x);
    const bool has_b = ra + 1 < a.rows;
    const float *xa = a.in + ra * n;
    const float *xb = xa + n;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 12; r++) {
        const int m = j + 256 * r;
        v[r] = make_float2(m < n ? xa[m] : 0.0f, (m < n && has_b) ? xb[m] : 0.0f);
    }
    fft3072(v, sh, j, tw);
    if (j < 192) {
#pragma unroll
        for (int r = 0; r < 16; r++) {
            const float2 y = cmul(v[r], H[j + 192 * r]);
            v[r] = make_float2(y.x, -y.y);
        }
    }
    fft3072_rev(v, sh, j, tw);
    // y = conj(v); keep y[n .. 2n)
#pragma unroll
    for (int r = 0; r < 12; r++) {
        const int m = j + 256 * r;
        if (m >= n && m < 2 * n) {
            const size_t o = ra * n + (m - n);
            if (a.out_bf16) {
                uint16_t *ob = reinterpret_cast<uint16_t *>(a.out);
                ob[o] = f32_to_bf16_rne(v[r].x);
                if (has_b) ob[o + n] = f32_to_bf16_rne(-v[r].y);
            } else {
                float *of = reinterpret_cast<float *>(a.out);
                of[o] = v[r].x;
                if (has_b) of[o + n] = -v[r].y;
            }
        }
    }
}

// One workgroup per pair of rows; spec = [H (L)] followed by [twiddles (L)]
// as float2.
__global__ __launch_bounds__(256) void fft_filter_kernel(FilterArgs a) {
    __shared__ float2 sh[FFT_LDS];
    const int j = threadIdx.x;
    const int n = a.n;
    const float2 *__restrict__ H = a.spec;
    const float2 *__restrict__ tw = a.spec + FFT_L;
    const size_t ra = 2 * static_cast<size_t>(blockIdx.x);
    const bool has_b = ra + 1 < a.rows;
    const float *xa = a.in + ra * n;
    const float *xb = xa + n;
    float2 v[16];
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = j + 256 * r;
        v[r] = make_float2(m < n ? xa[m] : 0.0f, (m < n && has_b) ? xb[m] : 0.0f);
    }
    fft4096(v, sh, j, tw);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const float2 y = cmul(v[r], H[j + 256 * r]);
        v[r] = make_float2(y.x, -y.y);
    }
    fft4096(v, sh, j, tw);
    // y = conj(v); keep y[n .. 2n)
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int m = j + 256 * r;
        if (m >= n && m < 2 * n) {
            const size_t o = ra * n + (m - n);
            if (a.out_bf16) {
                uint16_t *ob = reinterpret_cast<uint16_t *>(a.out);
                ob[o] = f32_to_bf16_rne(v[r].x);
                if (has_b) ob[o + n] = f32_to_bf16_rne(-v[r].y);
            } else {
                float *of = reinterpret_cast<float *>(a.out);
                of[o] = v[r].x;
                if (has_b) of[o + n] = -v[r].y;
            }
        }
    }
}

}  // namespace

int launch_filter(const FilterArgs &a, hipStream_t stream) {
    if (!a.in || !a.out || !a.taps || a.n <= 0) return CTBB_ERR_ARG;
    if (a.rows == 0) return CTBB_OK;
    // (the spectrum's length a.fft_len: 3072 or 4096, 0 without one)
    const bool fft = a.mode == CTBB_FILTER_FFT32 && 2 * a.n <= FFT_L;
    const bool fft3 = fft && a.fft_len == FFT_L3 && 2 * a.n <= FFT_L3;
    if (fft && !fft3 && a.spec && a.fft_len != FFT_L) return CTBB_ERR_ARG;
    {
        // In place (out == in) only for the FFT kernel with FP32 output: each
        // of its workgroups reads its two rows completely before writing the
        // same two rows.  Narrower (bfloat16) output rows, an offset overlap,
        // or the tiled kernels (which read rows other workgroups write) would
        // race, so any other overlap is refused.
        const uintptr_t ib = reinterpret_cast<uintptr_t>(a.in), ob = reinterpret_cast<uintptr_t>(a.out);
        const uintptr_t ie = ib + a.rows * static_cast<size_t>(a.n) * sizeof(float);
        const uintptr_t oe = ob + a.rows * static_cast<size_t>(a.n) * (a.out_bf16 ? sizeof(uint16_t) : sizeof(float));
        if (ib < oe && ob < ie && !(fft && !a.out_bf16 && ib == ob)) return CTBB_ERR_ARG;
    }
    if (fft) {
        // grid.x is limited to 2^31 - 1
        if (!a.spec) return CTBB_ERR_ARG;
        const size_t n_pairs = (a.rows + 1) / 2;
        const size_t max_pairs = 0x7fffffffu;
        for (size_t p0 = 0; p0 < n_pairs; p0 += max_pairs) {
            FilterArgs s = a;
            const size_t np = n_pairs - p0 < max_pairs ? n_pairs - p0 : max_pairs;
            s.in = a.in + 2 * p0 * a.n;
            if (a.out_bf16) s.out = reinterpret_cast<uint16_t *>(a.out) + 2 * p0 * a.n;
            else s.out = reinterpret_cast<float *>(a.out) + 2 * p0 * a.n;
            s.rows = a.rows - 2 * p0 < 2 * np ? a.rows - 2 * p0 : 2 * np;
            if (fft3)
                hipLaunchKernelGGL(fft3072_filter_kernel, dim3(static_cast<unsigned>(np)), dim3(256), 0, stream, s);
            else
                hipLaunchKernelGGL(fft_filter_kernel, dim3(static_cast<unsigned>(np)), dim3(256), 0, stream, s);
            if (hipGetLastError() != hipSuccess) return CTBB_ERR_HIP;
        }
        return CTBB_OK;
    }
    // FFT32 with 2n > 4096 uses the FP32 Toeplitz product below
    if (a.n % 64 != 0) {
        if (a.rows > 0x7fffffffu) return CTBB_ERR_SHAPE;
        size_t lds = 3 * static_cast<size_t>(a.n) * sizeof(float);
        if (lds > 64 * 1024) return CTBB_ERR_SHAPE;
        hipLaunchKernelGGL(conv_valu_kernel, dim3(static_cast<unsigned>(a.rows)), dim3(256), lds, stream, a);
        return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
    }
    if (a.mode == CTBB_FILTER_BF16X3 || a.mode == CTBB_FILTER_F16X3) {
        // 256 x 64 tiles on the double-K MFMAs; > 64 KB of LDS per workgroup
        constexpr int TM = 256;
        const size_t n_rt = (a.rows + TM - 1) / TM, n_ct = static_cast<size_t>(a.n) / 64;
        if (n_rt * n_ct > 0x7fffffffu) return CTBB_ERR_SHAPE;
        const size_t lds2 = 8 * (static_cast<size_t>(a.n) + 72) * sizeof(uint16_t);
        const void *kf = a.mode == CTBB_FILTER_BF16X3
                             ? reinterpret_cast<const void *>(&toeplitz_split_kernel<CTBB_FILTER_BF16X3>)
                             : reinterpret_cast<const void *>(&toeplitz_split_kernel<CTBB_FILTER_F16X3>);
        if (lds2 > 65536) {
            int dev = 0, max_lds = 0;
            if (hipGetDevice(&dev) != hipSuccess ||
                hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
                return CTBB_ERR_HIP;
            if (lds2 > static_cast<size_t>(max_lds)) return CTBB_ERR_SHAPE;
            if (hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds2)) !=
                hipSuccess)
                return CTBB_ERR_SHAPE;
        }
        const dim3 grid(static_cast<unsigned>(n_rt * n_ct));
        if (a.mode == CTBB_FILTER_BF16X3)
            hipLaunchKernelGGL((toeplitz_split_kernel<CTBB_FILTER_BF16X3>), grid, dim3(256), lds2, stream, a,
                               static_cast<int>(n_rt), static_cast<int>(n_ct));
        else
            hipLaunchKernelGGL((toeplitz_split_kernel<CTBB_FILTER_F16X3>), grid, dim3(256), lds2, stream, a,
                               static_cast<int>(n_rt), static_cast<int>(n_ct));
        return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
    }
    const size_t gy = (a.rows + BM - 1) / BM;
    // FP32 Toeplitz product (CTBB_FILTER_F32, and FFT32 rows wider than 2048)
    const size_t lds = BM * A_STRIDE * sizeof(float) + 2 * a.n * sizeof(float);
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
        if (a.mode != CTBB_FILTER_F32 && a.mode != CTBB_FILTER_FFT32) return CTBB_ERR_ARG;
        hipLaunchKernelGGL(toeplitz_f32_kernel, grid, dim3(256), lds, stream, s);
        if (hipGetLastError() != hipSuccess) return CTBB_ERR_HIP;
    }
    return CTBB_OK;
}

}  // namespace ctbb
