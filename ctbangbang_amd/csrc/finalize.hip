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
// (new; the reference stops at mu) is fused into the store.
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

}  // namespace

int launch_finalize(const FinalizeArgs &a, hipStream_t stream) {
    if (!a.image || !a.out || !a.src_slice || !a.weight) return CTBB_ERR_ARG;
    if (a.n_final <= 0 || a.n_final > 65535 || a.max_taps <= 0 || a.nxy == 0) return CTBB_ERR_SHAPE;
    dim3 grid(static_cast<unsigned>((a.nxy + 255) / 256), a.n_final);
    hipLaunchKernelGGL(finalize_kernel, grid, dim3(256), 0, stream, a);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
