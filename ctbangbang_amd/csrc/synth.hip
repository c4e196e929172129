// HIP evaluation of the synthetic helical scan (one thread per sample, FP64).
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels.h"
#include "synth_common.h"
#include "synth_host.h"

namespace ctbb {

namespace {

constexpr int kMaxEllipsoids = 64;

__global__ __launch_bounds__(256) void synth_kernel(SynthGeom g, const Ellipsoid *E, int n_e, long first,
                                                    long count, float *out) {
    __shared__ Ellipsoid se[kMaxEllipsoids];
    for (int i = threadIdx.x; i < n_e; i += blockDim.x) se[i] = E[i];
    __syncthreads();
    const long per = static_cast<long>(g.n_rows_raw) * g.nch;
    const long total = count * per;
    for (long i = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x; i < total;
         i += static_cast<long>(gridDim.x) * blockDim.x) {
        const long n = first + i / per;
        const int j = static_cast<int>((i % per) / g.nch), ch = static_cast<int>(i % g.nch);
        out[i] = synth_sample(g, se, n_e, n, j, ch);
    }
}

}  // namespace

int launch_synth(const recon_metadata &mr, const ctbb_phantom_params &pp, long first, long count, float *d_out,
                 hipStream_t stream) {
    if (!d_out || count < 0) return CTBB_ERR_ARG;
    const std::vector<Ellipsoid> E = make_phantom(mr, pp);
    if (E.size() > static_cast<size_t>(kMaxEllipsoids)) return CTBB_ERR_ARG;
    const SynthGeom g = make_synth_geom(mr, pp);
    Ellipsoid *dE = nullptr;
    if (hipMalloc(&dE, E.size() * sizeof(Ellipsoid)) != hipSuccess) return CTBB_ERR_NOMEM;
    int st = CTBB_OK;
    if (hipMemcpyAsync(dE, E.data(), E.size() * sizeof(Ellipsoid), hipMemcpyHostToDevice, stream) != hipSuccess)
        st = CTBB_ERR_HIP;
    if (st == CTBB_OK && count > 0) {
        const long total = count * g.n_rows_raw * static_cast<long>(g.nch);
        long blocks = (total + 255) / 256;
        if (blocks > 1048576) blocks = 1048576;
        hipLaunchKernelGGL(synth_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, stream, g, dE,
                           static_cast<int>(E.size()), first, count, d_out);
        if (hipGetLastError() != hipSuccess) st = CTBB_ERR_HIP;
    }
    if (hipStreamSynchronize(stream) != hipSuccess) st = CTBB_ERR_HIP;
    (void)hipFree(dE);
    return st;
}

}  // namespace ctbb
