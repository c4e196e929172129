// Final image stack: block reorder + triangle slice thickening (+ HU).
//
// Reference behaviour: src/finalize_image_stack.cu:26-120 and
// include/finalize_image_stack.cuh:26-76 (and the identical CPU twin
// src/finalize_image_stack_cpu.cu:24-117).  The reference reverses each
// 32-slice block into a temporary stack and launches one kernel plus a device
// synchronisation per final slice.  Here the host folds the block reversal
// into a per-final-slice list of (physical raw slice, normalised weight) taps
// in the reference's summation order, and a single kernel walks every voxel
// column once, accumulating with fmaf exactly as nvcc contracts the
// reference's "acc += w/sum * v".  The HU conversion HU = 1000 (mu - mu_w)/mu_w
// (new; the reference stops at mu) is fused into the store.  A second kernel
// turns a stack in HU into int16 stored values for DICOM output (the
// reference writes float32 only).
#include <hip/hip_runtime.h>

#include "kernels.h"

// Only the explicit fmaf calls may fuse: every other multiply-add rounds
// twice, as in the sequential FP32 statement of the algorithm.
#pragma clang fp contract(off)

namespace ctbb {

namespace {

__global__ __launch_bounds__(256) void finalize_kernel(FinalizeArgs a) {
    const size_t v = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (v >= a.nxy) return;
    const int k = blockIdx.y;
    const int *src = a.src_slice + static_cast<size_t>(k) * a.max_taps;
    const float *w = a.weight + static_cast<size_t>(k) * a.max_taps;
    float acc = 0.0f;
    for (int t = 0; t < a.max_taps; t++) {
        const int s = src[t];
        if (s < 0) break;
        acc = fmaf(w[t], a.image[static_cast<size_t>(s) * a.nxy + v], acc);
    }
    if (a.mu_water > 0.0f) acc = 1000.0f * (acc - a.mu_water) / a.mu_water;
    a.out[static_cast<size_t>(k) * a.nxy + v] = acc;
}

// Stored value of one HU sample: (hu - intercept) / slope (IEEE division),
// rounded half to even, NaN -> 0, saturated to the int16 range.
__device__ __forceinline__ int hu16_one(float hu, float slope, float intercept) {
    float v = rintf((hu - intercept) / slope);
    if (!(v == v)) v = 0.0f;
    v = fminf(fmaxf(v, -32768.0f), 32767.0f);
    return static_cast<int>(v);
}

// 8 samples per thread: two 16-byte loads, one 16-byte store where both
// buffers are 16-byte aligned and the 8 samples are inside the array
// (vec), single samples otherwise.  HBM-bound: 6 bytes per sample.
template <bool kVec>
__global__ __launch_bounds__(256) void hu16_kernel(Hu16Args a) {
    const size_t i0 = (static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x) * 8;
    if (i0 >= a.n) return;
    if (kVec && i0 + 8 <= a.n) {
        const float4 p = *reinterpret_cast<const float4 *>(a.in + i0);
        const float4 q = *reinterpret_cast<const float4 *>(a.in + i0 + 4);
        const float f[8] = {p.x, p.y, p.z, p.w, q.x, q.y, q.z, q.w};
        unsigned w[4];
#pragma unroll
        for (int e = 0; e < 4; e++)
            w[e] = (static_cast<unsigned>(hu16_one(f[2 * e], a.slope, a.intercept)) & 0xffffu) |
                   (static_cast<unsigned>(hu16_one(f[2 * e + 1], a.slope, a.intercept)) << 16);
        *reinterpret_cast<uint4 *>(a.out + i0) = make_uint4(w[0], w[1], w[2], w[3]);
        return;
    }
    for (size_t i = i0; i < i0 + 8 && i < a.n; i++)
        a.out[i] = static_cast<int16_t>(hu16_one(a.in[i], a.slope, a.intercept));
}

}  // namespace

int launch_hu16(const Hu16Args &a, hipStream_t stream) {
    if (a.n == 0) return CTBB_OK;
    if (!a.in || !a.out) return CTBB_ERR_ARG;
    if (!(a.slope > 0.0f) || !(a.slope < 3.0e38f) || !(a.intercept == a.intercept) ||
        !(a.intercept < 3.0e38f && a.intercept > -3.0e38f))
        return CTBB_ERR_ARG;
    const size_t n8 = (a.n + 7) / 8;
    const size_t blocks = (n8 + 255) / 256;
    if (blocks > 0x7fffffffu) return CTBB_ERR_SHAPE;
    const bool vec = (reinterpret_cast<uintptr_t>(a.in) & 15u) == 0 && (reinterpret_cast<uintptr_t>(a.out) & 15u) == 0;
    if (vec)
        hipLaunchKernelGGL(hu16_kernel<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, a);
    else
        hipLaunchKernelGGL(hu16_kernel<false>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, a);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

int launch_finalize(const FinalizeArgs &a, hipStream_t stream) {
    if (!a.image || !a.out || !a.src_slice || !a.weight) return CTBB_ERR_ARG;
    if (a.n_final <= 0 || a.n_final > 65535 || a.max_taps <= 0 || a.nxy == 0) return CTBB_ERR_SHAPE;
    dim3 grid(static_cast<unsigned>((a.nxy + 255) / 256), a.n_final);
    hipLaunchKernelGGL(finalize_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
