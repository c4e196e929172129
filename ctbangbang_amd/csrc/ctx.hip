// Execution context and the C-ABI of the four GPU stages.
//
// The context owns one non-blocking HIP stream (default priority), grow-only
// device workspaces, cached coordinate maps / filter taps, and HIP events
// recorded around every launch so callers can read per-kernel device time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "ctbb_hip.h"
#include "geometry.h"
#include "kernels.h"

#ifndef CTBB_PIPE_DEPTH
#define CTBB_PIPE_DEPTH 2
#endif

// View table of ctbb_backproject_dev_noalloc: a fixed device array, one copy
// per GPU at its own address.  Host code never passes the array's name to a
// kernel or a copy; it uses the per-device address hipGetSymbolAddress gives.
static __device__ __align__(256) ctbb::BpView scratch[CTBB_NOALLOC_MAX_VIEWS];

namespace {

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t n) {
        if (n <= bytes) return CTBB_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, n) != hipSuccess) return CTBB_ERR_NOMEM;
        bytes = n;
        return CTBB_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(p); }
};

struct EventPair {
    hipEvent_t a, b;
    hipStream_t s;
};

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// Address of `scratch` on `device` (which must be current), looked up once
// per device.
int scratch_address(int device, ctbb::BpView **out) {
    if (device < 0 || device >= kMaxDevices) return CTBB_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CTBB_ERR_HIP;
        g_scratch_ptr[device] = p;
    }
    *out = static_cast<ctbb::BpView *>(g_scratch_ptr[device]);
    return CTBB_OK;
}

}  // namespace

struct ctbb_ctx {
    int device = 0;
    // workgroups of a preparation kernel launched beside a per-block
    // backprojection: 2 per CU (set at creation)
    int prep_groups = 0;
    hipStream_t stream = nullptr;
    ctbb_options opt{};
    ctbb::BpView *noalloc_vtab = nullptr;   // this device's `scratch` array
    DevBuf raw_in, raw_af, rebin_tmp, image_block, image_stack, af_work;
    DevBuf vtab, maps, taps, spec, fin_tab;
    int spec_len = 0;          // transform length of `spec` (3072 or 4096; 0: none)
    // 1 once a rebin met a non-finite output sample (read and cleared by
    // ctbb_ctx_synchronize)
    DevBuf nonfinite;
    // host-streaming reconstruction: two raw pull buffers filled on a copy
    // stream (default priority) while the compute stream works on the other
    DevBuf pull[2];
    hipStream_t copy_stream = nullptr;
    hipEvent_t ev_ready[2] = {nullptr, nullptr}, ev_free[2] = {nullptr, nullptr};
    // whole-scan pipelines: the next blocks are prepared (adaptive filter,
    // rebin, ramp filter) on `stream` while block b is backprojected on
    // `stream_bp`, with kPipeDepth filtered-projection buffers
    static constexpr int kPipeDepth = CTBB_PIPE_DEPTH;
    hipStream_t stream_bp = nullptr;
    DevBuf filt[kPipeDepth], image_stack2;
    hipEvent_t ev_filt[kPipeDepth] = {}, ev_bp[kPipeDepth] = {}, ev_sync = nullptr;
    // Head of a whole-volume call: the first block's adaptive filter and
    // rebin run on `stream_head`, into workspaces of their own, after
    // ev_prep_done (recorded on `stream` after the previous call's last
    // preparation), so in back-to-back calls they run beside the previous
    // call's last backprojection; ev_head hands their result to the ramp
    // filter on `stream`.
    hipStream_t stream_head = nullptr;
    hipEvent_t ev_prep_done = nullptr, ev_head = nullptr;
    // the head's own workspaces (adaptive-filter output and work, rebinned
    // rows, zero-padded pull), so it touches nothing the previous call or
    // other entry points use
    DevBuf head_af, head_work, head_rebin, head_pad;
    // finalize tap lists last uploaded to fin_tab (unchanged lists are not
    // uploaded again: the upload waits for the compute stream)
    std::vector<char> fin_blob;
    // cached host-side keys
    std::string taps_key, maps_key;
    int map_kind = 0, map_sets = 0;
    std::map<std::string, std::vector<EventPair>> events;
    std::vector<EventPair> free_events;
    bool timing_failed = false;
    // batching form: filtered-projection and image arenas per group parity,
    // the device work list (plans, jobs, view tables, queues), and the
    // executable graphs of launched group preparations (kept until the
    // context's work has finished), the item counters of the persistent
    // backprojection launches (8 per launch)
    DevBuf arena[2], img_arena[2], plan_buf, tpos_buf, bp_counters;
    hipEvent_t ev_gdone[2] = {nullptr, nullptr}, ev_gprep = nullptr;
    std::vector<hipGraphExec_t> live_graphs;
    bool capturing = false;      // a graph is being captured: no event timing

    EventPair begin(const char *family, hipStream_t on = nullptr) {
        EventPair e;
        if (capturing) return EventPair{nullptr, nullptr, nullptr};
        if (!free_events.empty()) {
            e = free_events.back();
            free_events.pop_back();
        } else if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) {
            // no timing for this launch (ctbb_ctx_kernel_time reports an error)
            timing_failed = true;
            return EventPair{nullptr, nullptr, nullptr};
        }
        e.s = on ? on : stream;
        if (hipEventRecord(e.a, e.s) != hipSuccess) timing_failed = true;
        events[family].push_back(e);
        return e;
    }
    void end(const EventPair &e) {
        if (e.b && hipEventRecord(e.b, e.s) != hipSuccess) timing_failed = true;
    }
    void reset_events() {
        for (auto &kv : events)
            for (auto &e : kv.second) free_events.push_back(e);
        events.clear();
        timing_failed = false;
    }
};

namespace {

using namespace ctbb;

int hip_ok(hipError_t e) { return e == hipSuccess ? CTBB_OK : CTBB_ERR_HIP; }

int upload(ctbb_ctx *ctx, DevBuf &dst, const void *src, size_t bytes) {
    int st = dst.ensure(bytes);
    if (st) return st;
    if (hipMemcpyAsync(dst.p, src, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return CTBB_ERR_HIP;
    // the host staging buffer may be reused as soon as this returns
    return hip_ok(hipStreamSynchronize(ctx->stream));
}

// Backprojection view tables are built on the device, per block, into
// ctx->vtab: a job holds one block's parameters and its entry offset.
struct VtabJob {
    size_t off;
    double tube_start, z_rot;
    int half, n_ht, n_proj_turn;
};

VtabJob vtab_job(const recon_metadata &mr, size_t off) {
    VtabJob j;
    j.off = off;
    j.tube_start = static_cast<double>(bp_tube_start(mr));
    j.z_rot = static_cast<double>(mr.cg.z_rot);
    j.half = static_cast<int>(mr.cg.n_proj_turn / 2);
    j.n_ht = bp_n_half_turns(mr);
    j.n_proj_turn = static_cast<int>(mr.cg.n_proj_turn);
    return j;
}

size_t vtab_len(const VtabJob &j) { return static_cast<size_t>(j.half) * static_cast<size_t>(j.n_ht); }

// Queue the tables of all jobs on the context stream (total entries in all).
int build_view_tables(ctbb_ctx *ctx, const std::vector<VtabJob> &jobs, size_t total) {
    int st = ctx->vtab.ensure(total * sizeof(BpView));
    if (st) return st;
    for (const VtabJob &j : jobs) {
        st = launch_view_table(ctx->vtab.as<BpView>() + j.off, j.tube_start, j.half, j.n_ht, j.n_proj_turn, j.z_rot,
                               ctx->stream);
        if (st) return st;
    }
    return CTBB_OK;
}

// Zero fill on a stream (a kernel rather than hipMemsetAsync, so the
// sequence stays valid inside a captured graph).
__global__ void fill_zero_kernel(float *p, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = 0.0f;
}

int fill_zero(float *p, size_t n, hipStream_t stream) {
    if (n == 0) return CTBB_OK;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 16384));
    hipLaunchKernelGGL(fill_zero_kernel, dim3(grid), dim3(256), 0, stream, p, n);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

int check_geometry(const recon_metadata *mr) {
    if (!mr) return CTBB_ERR_ARG;
    const ct_geom &cg = mr->cg;
    if (mr->ri.n_ffs == 0 || cg.n_channels == 0 || cg.n_rows == 0 || cg.n_proj_turn < 2) return CTBB_ERR_ARG;
    if (cg.n_channels_oversampled != 2 * cg.n_channels) return CTBB_ERR_SHAPE;
    if (mr->ri.n_proj_pull % mr->ri.n_ffs) return CTBB_ERR_SHAPE;
    if (mr->ri.n_proj_pull / mr->ri.n_ffs <= 2 * cg.add_projections) return CTBB_ERR_SHAPE;
    return CTBB_OK;
}

// Filter taps of the prm's ReconKernel, cached per (install dir, channels, kernel).
// In-place radix-2 DFT (forward sign) of a power-of-two length, FP64.
void dft_pow2(std::vector<std::complex<double>> &x) {
    const size_t n = x.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(x[i], x[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        const double ang = -2.0 * M_PI / static_cast<double>(len);
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                const std::complex<double> w(std::cos(ang * k), std::sin(ang * k));
                const std::complex<double> u = x[i + k], v = x[i + k + len / 2] * w;
                x[i + k] = u + v;
                x[i + k + len / 2] = u - v;
            }
    }
}

// DFT of length L (forward sign) of the first `count` values of x zero-padded,
// FP64, directly from a table of the L roots of unity.
std::vector<std::complex<double>> dft_direct(const float *x, size_t count, size_t L) {
    std::vector<std::complex<double>> w(L), X(L);
    for (size_t t = 0; t < L; t++) {
        const double a = -2.0 * M_PI * static_cast<double>(t) / static_cast<double>(L);
        w[t] = std::complex<double>(std::cos(a), std::sin(a));
    }
    for (size_t k = 0; k < L; k++) {
        std::complex<double> acc = 0.0;
        size_t t = 0;                                   // m k mod L
        for (size_t m = 0; m < count; m++) {
            acc += static_cast<double>(x[m]) * w[t];
            t += k;
            if (t >= L) t -= L;
        }
        X[k] = acc;
    }
    return X;
}

// FFT32 form of the filter: H = DFT_L(taps zero-padded) / L and the
// twiddles e^(-2 pi i t / L), evaluated in FP64 and rounded once, into
// ctx->spec, with L = 3072 for 2n <= 3072 and 4096 for 2n <= 4096 (wider
// filters use the FP32 product).
int make_spec(ctbb_ctx *ctx, const float *taps, size_t count) {
    if (count > 4096) return CTBB_OK;
    const size_t L = count <= 3072 ? 3072 : 4096;
    std::vector<std::complex<double>> h;
    if (L == 4096) {
        h.assign(L, 0.0);
        for (size_t m = 0; m < count; m++) h[m] = taps[m];
        dft_pow2(h);
    } else {
        h = dft_direct(taps, count, L);
    }
    std::vector<float> sp(4 * L);
    for (size_t k = 0; k < L; k++) {
        sp[2 * k] = static_cast<float>(h[k].real() / L);
        sp[2 * k + 1] = static_cast<float>(h[k].imag() / L);
        const double a = -2.0 * M_PI * static_cast<double>(k) / L;
        sp[2 * L + 2 * k] = static_cast<float>(std::cos(a));
        sp[2 * L + 2 * k + 1] = static_cast<float>(std::sin(a));
    }
    ctx->spec_len = 0;
    const int st = upload(ctx, ctx->spec, sp.data(), sp.size() * sizeof(float));
    if (st) return st;
    ctx->spec_len = static_cast<int>(L);
    return CTBB_OK;
}

int ensure_taps(ctbb_ctx *ctx, const recon_metadata *mr) {
    std::string key = std::string(mr->install_dir) + "|" + std::to_string(mr->cg.n_channels) + "|" +
                      std::to_string(mr->rp.recon_kernel);
    if (key == ctx->taps_key && ctx->taps.p) return CTBB_OK;
    std::vector<float> t(2 * mr->cg.n_channels_oversampled);
    int st = ctbb_load_filter(mr, t.data());
    if (st) return st;
    st = upload(ctx, ctx->taps, t.data(), t.size() * sizeof(float));
    if (st) return st;
    st = make_spec(ctx, t.data(), t.size());
    if (st) return st;
    ctx->taps_key = key;
    return CTBB_OK;
}

int ensure_maps(ctbb_ctx *ctx, const recon_metadata *mr) {
    const ct_geom &cg = mr->cg;
    char key[512];
    std::snprintf(key, sizeof(key), "%zu|%zu|%zu|%a|%a|%a|%a|%a|%a|%zu|%d|%d|%d|%a", cg.n_channels, cg.n_rows,
                  cg.n_proj_ffs, cg.r_f, cg.src_to_det, cg.fan_angle_increment, cg.central_channel, cg.anode_angle,
                  mr->rp.coll_slicewidth, mr->ri.n_ffs, mr->rp.z_ffs, mr->rp.phi_ffs, cg.reverse_row_interleave,
                  static_cast<double>(cg.n_proj_turn));
    if (ctx->maps_key == key && ctx->maps.p) return CTBB_OK;
    std::vector<float> bidx, aoff;
    int kind = 0, sets = 0;
    int st = rebin_maps(*mr, bidx, aoff, &kind, &sets);
    if (st) return st;
    std::vector<float> both(bidx.size() + aoff.size());
    std::memcpy(both.data(), bidx.data(), bidx.size() * sizeof(float));
    std::memcpy(both.data() + bidx.size(), aoff.data(), aoff.size() * sizeof(float));
    st = upload(ctx, ctx->maps, both.data(), both.size() * sizeof(float));
    if (st) return st;
    ctx->maps_key = key;
    ctx->map_kind = kind;
    ctx->map_sets = sets;
    return CTBB_OK;
}

// n_views_max < 0: every rebinned view of the block; otherwise at most the
// first n_views_max views.
// hot (with d_af): frames whose flag is set come from d_af, the adaptive
// filter's hot_only output, the others from d_raw.
int do_rebin(ctbb_ctx *ctx, const recon_metadata *mr, const float *d_raw, float *d_out, int n_views_max = -1,
             hipStream_t on = nullptr, const unsigned char *hot = nullptr, const float *d_af = nullptr,
             int max_groups = 0) {
    int st = ensure_maps(ctx, mr);
    if (st) return st;
    const ct_geom &cg = mr->cg;
    RebinArgs a{};
    a.raw = d_raw;
    a.out = d_out;
    const size_t nos = cg.n_channels_oversampled;
    a.bidx = ctx->maps.as<float>();
    a.aoff = ctx->maps.as<float>() + static_cast<size_t>(ctx->map_sets) * nos;
    a.kind = ctx->map_kind;
    a.n_ffs = static_cast<int>(mr->ri.n_ffs);
    a.nch = static_cast<int>(cg.n_channels);
    a.nos = static_cast<int>(nos);
    a.n_rows_raw = static_cast<int>(cg.n_rows_raw);
    a.n_rows = static_cast<int>(cg.n_rows);
    a.n_proj = static_cast<int>(mr->ri.n_proj_pull / mr->ri.n_ffs);
    a.add = static_cast<int>(cg.add_projections);
    a.n_proj_out = a.n_proj - 2 * a.add;
    if (n_views_max >= 0 && n_views_max < a.n_proj_out) a.n_proj_out = n_views_max;
    a.reverse_rows = cg.reverse_row_interleave;
    a.nonfinite = ctx->nonfinite.as<unsigned>();
    a.hot = hot;
    a.raw_af = d_af;
    a.max_groups = max_groups;
    hipStream_t s = on ? on : ctx->stream;
    EventPair e = ctx->begin("rebin", s);
    st = launch_rebin(a, s);
    ctx->end(e);
    return st;
}

int do_filter(ctbb_ctx *ctx, const float *d_taps, int n, size_t rows, const float *d_in, void *d_out, int out_bf16) {
    FilterArgs f{};
    f.in = d_in;
    f.out = d_out;
    f.taps = d_taps;
    f.spec = ctx->spec.p ? ctx->spec.as<float2>() : nullptr;
    f.fft_len = ctx->spec.p ? ctx->spec_len : 0;
    f.rows = rows;
    f.n = n;
    f.mode = ctx->opt.filter_mode;
    f.out_bf16 = out_bf16;
    EventPair e = ctx->begin("filter");
    int st = launch_filter(f, ctx->stream);
    ctx->end(e);
    return st;
}

int fill_bp_args(const recon_metadata *mr, BpArgs &a) {
    const ct_geom &cg = mr->cg;
    const recon_params &rp = mr->rp;
    if (rp.nx % 16 || rp.ny % 16 || rp.nx == 0 || rp.ny == 0) return CTBB_ERR_SHAPE;
    if (mr->ri.n_slices_block != 32) return CTBB_ERR_SHAPE;
    a.nx = static_cast<int>(rp.nx);
    a.ny = static_cast<int>(rp.ny);
    a.nos = static_cast<int>(cg.n_channels_oversampled);
    a.n_rows = static_cast<int>(cg.n_rows);
    a.half = static_cast<int>(cg.n_proj_turn / 2);
    a.n_ht = bp_n_half_turns(*mr);
    if (a.n_ht <= 0) return CTBB_ERR_SHAPE;
    a.tiles_y = a.ny / 16;
    a.n_tiles = (a.nx / 16) * a.tiles_y;
    // reference include/backproject.cuh:59-64
    const double dx = static_cast<double>(rp.recon_fov) / rp.nx, dy = static_cast<double>(rp.recon_fov) / rp.ny;
    a.xs = static_cast<float>(cg.table_direction == -1 ? dx : -dx);
    a.xc = (rp.nx - 1) / 2.0f;
    a.x0 = rp.x_origin;
    a.ys = static_cast<float>(dy);
    a.yc = (rp.ny - 1) / 2.0f;
    a.y0 = rp.y_origin;
    const double ds = static_cast<double>(cg.r_f) * cg.fan_angle_increment / 2.0;
    a.inv_ds = static_cast<float>(1.0 / ds);
    a.two_cc = 2.0f * cg.central_channel;
    a.inv_rf = static_cast<float>(1.0 / cg.r_f);
    a.rf2 = cg.r_f * cg.r_f;
    a.k_asin = static_cast<float>(cg.z_rot / (2.0 * static_cast<double>(kPiBp)));
    a.coll = rp.coll_slicewidth;
    a.tanc = static_cast<float>(std::tan(static_cast<double>(cg.theta_cone) / 2.0));
    a.neg_td = static_cast<float>(-cg.table_direction);
    a.hr = (static_cast<float>(cg.n_rows) - 1.0f) / 2.0f;
    const double dmax = dx > dy ? dx : dy;
    a.half_w = static_cast<float>(7.5 * std::sqrt(2.0) * dmax / ds + 1.0);
    a.acq_half = rp.acq_fov / 2.0f;
    // largest |p| of a voxel centre: the grid corner farthest from the isocentre
    const double hx = 0.5 * rp.recon_fov + std::fabs(static_cast<double>(rp.x_origin));
    const double hy = 0.5 * rp.recon_fov + std::fabs(static_cast<double>(rp.y_origin));
    a.asin_small_only = std::sqrt(hx * hx + hy * hy) <= 0.449 * cg.r_f ? 1 : 0;
    a.out_scale = static_cast<float>(2.0 * static_cast<double>(kPiBp) / cg.n_proj_turn);
    return CTBB_OK;
}

int do_backproject(ctbb_ctx *ctx, const recon_metadata *mr, const void *d_rebin, const BpView *d_vtab,
                   float *d_out, hipStream_t on = nullptr, bool timed = true, bool beside = false) {
    if (!on) on = ctx->stream;
    BpArgs a{};
    int st = fill_bp_args(mr, a);
    if (st) return st;
    a.nz = ctx->opt.bp_tile == 32 ? 32 : 16;
    if (ctx->opt.proj_dtype == CTBB_PROJ_BF16) a.rebin_bf16 = static_cast<const uint16_t *>(d_rebin);
    else a.rebin = static_cast<const float *>(d_rebin);
    a.vtab = d_vtab;
    a.out = d_out;
    if (!timed) return launch_backproject(a, on, beside);
    EventPair e = ctx->begin("backproject", on);
    st = launch_backproject(a, on, beside);
    ctx->end(e);
    return st;
}

// Final-slice tap lists: reference src/finalize_image_stack.cu:45-105 with the
// block reversal of :56-63 folded into the source slice index.
int finalize_table(const recon_metadata *mr, std::vector<int> &src, std::vector<float> &w, int *max_taps,
                   int *n_final) {
    const recon_params &rp = mr->rp;
    const recon_info &ri = mr->ri;
    const int nb = static_cast<int>(ri.n_slices_block);
    const int n_raw = nb * static_cast<int>(ri.n_blocks);
    if (n_raw <= 0 || !mr->table_positions || rp.n_readings <= 1000) return CTBB_ERR_ARG;
    float d = rp.end_pos - rp.start_pos;
    int recon_direction = d == 0.0f ? 1 : static_cast<int>(std::fabs(d) / d);
    if (recon_direction != 1 && recon_direction != -1) recon_direction = 1;
    double tdd = mr->table_positions[1000] - mr->table_positions[0];
    int table_direction = tdd == 0.0 ? 1 : static_cast<int>(std::fabs(tdd) / tdd);
    const int nf = ctbb_n_slices_final(mr);
    std::vector<float> raw_loc(n_raw), wt(n_raw);
    for (int i = 0; i < n_raw; i++) raw_loc[i] = ri.recon_start_pos + recon_direction * i * rp.coll_slicewidth;
    std::vector<std::vector<std::pair<int, float>>> lists(nf);
    int mt = 1;
    for (int k = 0; k < nf; k++) {
        const float loc = rp.start_pos + recon_direction * k * rp.slice_thickness;
        float sum = 0.0f;
        for (int i = 0; i < n_raw; i++) {
            wt[i] = std::fmax(0.0f, 1.0f - std::fabs(raw_loc[i] - loc) / rp.slice_thickness);
            sum += wt[i];
        }
        for (int i = 0; i < n_raw; i++) {
            if (wt[i] == 0.0f) continue;
            const int b = i / nb, z = i % nb;
            const int s = recon_direction != table_direction ? b * nb + (nb - 1 - z) : i;
            lists[k].push_back({s, wt[i] / sum});
        }
        if (static_cast<int>(lists[k].size()) > mt) mt = static_cast<int>(lists[k].size());
    }
    src.assign(static_cast<size_t>(nf) * mt, -1);
    w.assign(static_cast<size_t>(nf) * mt, 0.0f);
    for (int k = 0; k < nf; k++)
        for (size_t t = 0; t < lists[k].size(); t++) {
            src[static_cast<size_t>(k) * mt + t] = lists[k][t].first;
            w[static_cast<size_t>(k) * mt + t] = lists[k][t].second;
        }
    *max_taps = mt;
    *n_final = nf;
    return CTBB_OK;
}

// Finalize tap lists of one scan as one blob: [src (nf * mt ints)][w (nf * mt floats)]
int finalize_blob(const recon_metadata *mr, std::vector<char> &blob, int *mt, int *nf) {
    std::vector<int> src;
    std::vector<float> w;
    int st = finalize_table(mr, src, w, mt, nf);
    if (st) return st;
    blob.resize(src.size() * sizeof(int) + w.size() * sizeof(float));
    std::memcpy(blob.data(), src.data(), src.size() * sizeof(int));
    std::memcpy(blob.data() + src.size() * sizeof(int), w.data(), w.size() * sizeof(float));
    return CTBB_OK;
}

// Launch the finalize kernel with tap lists already on the device at d_blob.
int launch_finalize_blob(ctbb_ctx *ctx, const recon_metadata *mr, const float *d_image, float *d_final,
                         const char *d_blob, int mt, int nf, hipStream_t on) {
    FinalizeArgs f{};
    f.image = d_image;
    f.out = d_final;
    f.src_slice = reinterpret_cast<const int *>(d_blob);
    f.weight = reinterpret_cast<const float *>(d_blob + static_cast<size_t>(nf) * mt * sizeof(int));
    f.max_taps = mt;
    f.n_final = nf;
    f.nxy = static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    f.mu_water = ctx->opt.mu_water;
    EventPair e = ctx->begin("finalize", on);
    int st = launch_finalize(f, on);
    ctx->end(e);
    return st;
}

int do_finalize(ctbb_ctx *ctx, const recon_metadata *mr, const float *d_image, float *d_final) {
    std::vector<char> blob;
    int mt = 0, nf = 0;
    int st = finalize_blob(mr, blob, &mt, &nf);
    if (st) return st;
    if (blob != ctx->fin_blob || !ctx->fin_tab.p) {
        st = upload(ctx, ctx->fin_tab, blob.data(), blob.size());
        if (st) return st;
        ctx->fin_blob = blob;
    }
    return launch_finalize_blob(ctx, mr, d_image, d_final, ctx->fin_tab.as<char>(), mt, nf, ctx->stream);
}

// vec4: the 16-byte apply kernel (for a GPU that runs nothing else; the
// one-sample kernel needs few registers and runs beside a backprojection).
// hot: write d_out only for the views the filter changes and return their
// flags in *hot (inside the work buffer, valid until its next use).
int do_adaptive(ctbb_ctx *ctx, const recon_metadata *mr, const float *d_raw, float *d_out, bool vec4 = true,
                hipStream_t on = nullptr, DevBuf *work = nullptr, const unsigned char **hot = nullptr,
                int max_groups = 0) {
    const ct_geom &cg = mr->cg;
    AdaptiveArgs a{};
    a.raw = d_raw;
    a.out = d_out;
    a.n_proj = static_cast<int>(mr->ri.n_proj_pull);
    a.n_rows_raw = static_cast<int>(cg.n_rows_raw);
    a.nch = static_cast<int>(cg.n_channels);
    a.n_proj_ffs = static_cast<int>(cg.n_proj_ffs);
    a.strength = mr->rp.adaptive_filtration_s;
    a.vec4 = vec4 ? 1 : 0;
    a.fixed_threshold = ctx->opt.adaptive_threshold;
    a.hot_only = hot ? 1 : 0;
    a.max_groups = max_groups;
    DevBuf &wb = work ? *work : ctx->af_work;
    int st = wb.ensure(6 * static_cast<size_t>(a.n_proj) * sizeof(float));
    if (st) return st;
    a.work = wb.as<float>();
    if (hot) *hot = adaptive_hot_flags(a.work, a.n_proj);
    hipStream_t s = on ? on : ctx->stream;
    EventPair e = ctx->begin("adaptive", s);
    st = launch_adaptive(a, s);
    ctx->end(e);
    return st;
}

size_t proj_elem_bytes(const ctbb_ctx *ctx) { return ctx->opt.proj_dtype == CTBB_PROJ_BF16 ? 2 : 4; }

int rebin_filter_dev(ctbb_ctx *ctx, const recon_metadata *mr, const float *d_raw, void *d_rebin) {
    int st = check_geometry(mr);
    if (st) return st;
    ctbb_sizes sz;
    ctbb_block_sizes(mr, &sz);
    st = ensure_taps(ctx, mr);
    if (st) return st;
    st = ctx->rebin_tmp.ensure(sz.rebin * sizeof(float));
    if (st) return st;
    st = do_rebin(ctx, mr, d_raw, ctx->rebin_tmp.as<float>());
    if (st) return st;
    const int n = static_cast<int>(mr->cg.n_channels_oversampled);
    return do_filter(ctx, ctx->taps.as<float>(), n, sz.n_proj_out * mr->cg.n_rows, ctx->rebin_tmp.as<float>(),
                     d_rebin, ctx->opt.proj_dtype == CTBB_PROJ_BF16);
}

int backproject_dev(ctbb_ctx *ctx, const recon_metadata *mr, const void *d_rebin, float *d_img) {
    int st = check_geometry(mr);
    if (st) return st;
    if (!mr->tube_angles) return CTBB_ERR_ARG;
    const VtabJob j = vtab_job(*mr, 0);
    if (j.n_ht <= 0) return CTBB_ERR_SHAPE;
    st = build_view_tables(ctx, {j}, vtab_len(j));
    if (st) return st;
    return do_backproject(ctx, mr, d_rebin, ctx->vtab.as<BpView>(), d_img);
}

// Adaptive filter -> rebin -> ramp filter of the current block (mr->ri.cb)
// on ctx->stream, raw pull at `win`, filtered projections into d_out.
// head: run the adaptive filter and the rebin on this stream (after
// ev_prep_done; the caller has made the zero-padded pull, if any, there
// too) and hand over to ctx->stream for the filter.
// filter_after: the ramp filter waits for this event (the backprojection
// running beside the preparation: the FFT kernel beside it would slow both).
int prep_block(ctbb_ctx *c, recon_metadata *mr, const float *win, void *d_out, hipStream_t head = nullptr,
               hipEvent_t filter_after = nullptr) {
    const ct_geom &cg = mr->cg;
    const size_t frame = cg.n_rows_raw * cg.n_channels;
    const size_t n = mr->ri.n_proj_pull;
    ctbb_sizes sz;
    ctbb_block_sizes(mr, &sz);
    int st;
    DevBuf &af = head ? c->head_af : c->raw_af;
    DevBuf &rb = head ? c->head_rebin : c->rebin_tmp;
    // The adaptive filter writes only the views it changes (C1: about 6%);
    // the rebin reads those from its output and the rest from the pull.
    // The first kernel of the chain (the view maxima, or the rebin without
    // the adaptive filter) is released together with a backprojection
    // launch: with its full grid it took the CUs before every
    // backprojection workgroup was placed, and those finished late (C1: 0.3
    // ms per volume with the filter, 3 ms without it).  That kernel runs at
    // most prep_groups workgroups (2 per CU, looping over its work).  The
    // rebin after the filter starts once the backprojection is placed and
    // keeps its full grid (looping workgroups are slower there).
    const unsigned char *hot = nullptr;
    if (c->opt.adaptive_filter) {
        st = af.ensure(n * frame * sizeof(float));
        if (st) return st;
        st = do_adaptive(c, mr, win, af.as<float>(), false, head, head ? &c->head_work : nullptr, &hot,
                         c->prep_groups);
        if (st) return st;
    }
    st = rb.ensure(sz.rebin * sizeof(float));
    if (st) return st;
    // The backprojection reads views i + k * half (i < half, k < n_ht) only
    // (view table, reference src/backproject.cu:39); the pull's remaining
    // rebinned views (C1: 206 of 1934) are neither rebinned nor filtered.
    // The FFT filter transforms rows in pairs (2m, 2m + 1) as one complex
    // signal, so a used row's result depends on its partner: when the used
    // rows end on an even row, its partner (the next view's first row) is
    // rebinned and filtered too, and every used value equals the one of the
    // whole-block filter.
    const size_t n_out = sz.n_proj_out;
    size_t views = std::min(n_out, static_cast<size_t>(cg.n_proj_turn / 2) * static_cast<size_t>(bp_n_half_turns(*mr)));
    size_t rows = views * cg.n_rows;
    if ((rows & 1) && views < n_out) {
        views += 1;
        rows += 1;
    }
    st = do_rebin(c, mr, win, rb.as<float>(), static_cast<int>(views), head, hot, hot ? af.as<float>() : nullptr,
                  c->opt.adaptive_filter ? 0 : c->prep_groups);
    if (st) return st;
    if (head && (hipEventRecord(c->ev_head, head) != hipSuccess ||
                 hipStreamWaitEvent(c->stream, c->ev_head, 0) != hipSuccess))
        return CTBB_ERR_HIP;
    if (filter_after && hipStreamWaitEvent(c->stream, filter_after, 0) != hipSuccess) return CTBB_ERR_HIP;
    return do_filter(c, c->taps.as<float>(), static_cast<int>(cg.n_channels_oversampled), rows, rb.as<float>(), d_out,
                     c->opt.proj_dtype == CTBB_PROJ_BF16);
}

int ensure_pipeline(ctbb_ctx *c, size_t rebin_bytes) {
    if (!c->stream_bp) {
        if (hipStreamCreateWithFlags(&c->stream_bp, hipStreamNonBlocking) != hipSuccess) return CTBB_ERR_HIP;
        for (int i = 0; i < ctbb_ctx::kPipeDepth; i++)
            if (hipEventCreateWithFlags(&c->ev_filt[i], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_bp[i], hipEventDisableTiming) != hipSuccess)
                return CTBB_ERR_HIP;
        if (hipEventCreateWithFlags(&c->ev_sync, hipEventDisableTiming) != hipSuccess) return CTBB_ERR_HIP;
    }
    if (!c->stream_head) {
        // ev_prep_done starts out recorded after everything queued so far
        if (hipStreamCreateWithFlags(&c->stream_head, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_prep_done, hipEventDisableTiming) != hipSuccess ||
            hipEventCreateWithFlags(&c->ev_head, hipEventDisableTiming) != hipSuccess ||
            hipEventRecord(c->ev_prep_done, c->stream) != hipSuccess)
            return CTBB_ERR_HIP;
    }
    for (DevBuf &b : c->filt) {
        const int st = b.ensure(rebin_bytes);
        if (st) return st;
    }
    // the backprojection stream starts after everything queued so far on ctx->stream
    if (hipEventRecord(c->ev_sync, c->stream) != hipSuccess ||
        hipStreamWaitEvent(c->stream_bp, c->ev_sync, 0) != hipSuccess)
        return CTBB_ERR_HIP;
    return CTBB_OK;
}

// Whether a block's backprojection has at least 8 rounds of tiles for the
// resident workgroups.  Such a block launches one workgroup per tile, and
// the next block's filter does not wait for it: the filter's workgroups take
// the slots of its retiring workgroups between rounds.  A block of fewer
// rounds keeps the resident grid, and the filter runs after it (beside a
// resident grid it would only get the tail).  Measured per volume: C4
// (1024^2, 8 rounds) 482 against 494 ms; C1 (512^2, 2 rounds) and the
// 768^2 c3 scans (4.5 rounds) were faster with the resident grid (c3, one
// call per scan: 1.19 against 1.30 s with the threshold at 4).
bool bp_many_rounds(ctbb_ctx *c, const recon_metadata *mr) {
    BpArgs a{};
    if (fill_bp_args(mr, a)) return false;
    a.nz = c->opt.bp_tile == 32 ? 32 : 16;
    if (c->opt.proj_dtype == CTBB_PROJ_BF16) a.rebin_bf16 = reinterpret_cast<const uint16_t *>(16);
    else a.rebin = reinterpret_cast<const float *>(16);      // placeholders for the variant choice
    a.vtab = reinterpret_cast<const BpView *>(16);
    a.out = reinterpret_cast<float *>(16);
    long groups = 0;
    return backproject_resident(a, &groups) == CTBB_OK && groups > 0 && a.n_tiles >= 8 * groups;
}

// Item `item` of a pipeline (block b of its scan): prepare on ctx->stream into
// filtered-projection buffer item % kPipeDepth (after the backprojection of
// item - kPipeDepth has read it), then backproject block b into d_image on
// stream_bp once the preparation is done.
int pipeline_block(ctbb_ctx *c, recon_metadata *mr, int item, int b, const float *win, const BpView *d_vtab,
                   float *d_image, hipStream_t head = nullptr) {
    const int k = item % ctbb_ctx::kPipeDepth;
    DevBuf &out = c->filt[k];
    if (item >= ctbb_ctx::kPipeDepth && hipStreamWaitEvent(c->stream, c->ev_bp[k], 0) != hipSuccess)
        return CTBB_ERR_HIP;
    // the previous item's backprojection (launched just before) is the one
    // this preparation runs beside
    const bool many_rounds = bp_many_rounds(c, mr);
    hipEvent_t running = item > 0 && !many_rounds ? c->ev_bp[(item - 1) % ctbb_ctx::kPipeDepth] : nullptr;
    int st = prep_block(c, mr, win, out.p, head, running);
    if (st) return st;
    if (hipEventRecord(c->ev_filt[k], c->stream) != hipSuccess ||
        hipStreamWaitEvent(c->stream_bp, c->ev_filt[k], 0) != hipSuccess)
        return CTBB_ERR_HIP;
    const size_t nxy = static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    st = do_backproject(c, mr, out.p, d_vtab, d_image + static_cast<size_t>(b) * mr->ri.n_slices_block * nxy,
                        c->stream_bp, true, /*beside: the next block's preparation*/ !many_rounds);
    if (st) return st;
    return hip_ok(hipEventRecord(c->ev_bp[k], c->stream_bp));
}

// After the context's work has finished: CTBB_ERR_NONFINITE (and the flag
// cleared) when a rebin met a non-finite sample since the last check.
int take_nonfinite(ctbb_ctx *c) {
    unsigned flag = 0;
    if (hipMemcpy(&flag, c->nonfinite.p, sizeof(flag), hipMemcpyDeviceToHost) != hipSuccess) return CTBB_ERR_HIP;
    if (!flag) return CTBB_OK;
    if (hipMemset(c->nonfinite.p, 0, sizeof(flag)) != hipSuccess) return CTBB_ERR_HIP;
    return CTBB_ERR_NONFINITE;
}

// ctx->stream waits for every backprojection queued on stream_bp
int join_pipeline(ctbb_ctx *c) {
    if (hipEventRecord(c->ev_sync, c->stream_bp) != hipSuccess) return CTBB_ERR_HIP;
    return hip_ok(hipStreamWaitEvent(c->stream, c->ev_sync, 0));
}

}  // namespace

extern "C" {

int ctbb_ctx_create(int device, ctbb_ctx **out) {
    if (!out) return CTBB_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return CTBB_ERR_HIP;
    if (device >= 0) {
        if (device >= n) return CTBB_ERR_ARG;
        if (hipSetDevice(device) != hipSuccess) return CTBB_ERR_HIP;
    }
    ctbb_ctx *c = new ctbb_ctx();
    (void)hipGetDevice(&c->device);
    int n_cu = 0;
    if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c->device) == hipSuccess && n_cu > 0)
        c->prep_groups = 2 * n_cu;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CTBB_ERR_HIP;
    }
    // the no-allocation entry point's array on this device (current here)
    int sst = scratch_address(c->device, &c->noalloc_vtab);
    if (!sst) sst = c->nonfinite.ensure(sizeof(unsigned));
    if (!sst && hipMemsetAsync(c->nonfinite.p, 0, sizeof(unsigned), c->stream) != hipSuccess) sst = CTBB_ERR_HIP;
    if (sst) {
        c->nonfinite.release();
        (void)hipStreamDestroy(c->stream);
        delete c;
        return sst;
    }
    c->opt.filter_mode = CTBB_FILTER_FFT32;
    c->opt.proj_dtype = CTBB_PROJ_F32;
    c->opt.mu_water = 0.0f;
    c->opt.bp_tile = 0;
    c->opt.bp_views_per_pass = 0;
    c->opt.adaptive_filter = 1;
    *out = c;
    return CTBB_OK;
}

void ctbb_ctx_destroy(ctbb_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    if (c->stream_bp) (void)hipStreamSynchronize(c->stream_bp);
    if (c->stream_head) (void)hipStreamSynchronize(c->stream_head);
    for (DevBuf *b : {&c->raw_in, &c->raw_af, &c->rebin_tmp, &c->image_block, &c->image_stack, &c->af_work,
                      &c->vtab, &c->maps, &c->taps, &c->spec, &c->fin_tab, &c->pull[0], &c->pull[1],
                      &c->image_stack2, &c->nonfinite, &c->head_af, &c->head_work, &c->head_rebin, &c->head_pad})
        b->release();
    for (DevBuf &b : c->filt) b.release();
    for (int i = 0; i < 2; i++) {
        if (c->ev_ready[i]) (void)hipEventDestroy(c->ev_ready[i]);
        if (c->ev_free[i]) (void)hipEventDestroy(c->ev_free[i]);
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    for (int i = 0; i < ctbb_ctx::kPipeDepth; i++) {
        if (c->ev_filt[i]) (void)hipEventDestroy(c->ev_filt[i]);
        if (c->ev_bp[i]) (void)hipEventDestroy(c->ev_bp[i]);
    }
    if (c->ev_sync) (void)hipEventDestroy(c->ev_sync);
    if (c->stream_bp) (void)hipStreamDestroy(c->stream_bp);
    for (hipEvent_t e : {c->ev_prep_done, c->ev_head})
        if (e) (void)hipEventDestroy(e);
    if (c->stream_head) (void)hipStreamDestroy(c->stream_head);
    for (hipGraphExec_t g : c->live_graphs) (void)hipGraphExecDestroy(g);
    c->live_graphs.clear();
    for (DevBuf *b : {&c->arena[0], &c->arena[1], &c->img_arena[0], &c->img_arena[1], &c->plan_buf, &c->tpos_buf,
                      &c->bp_counters})
        b->release();
    for (hipEvent_t e : {c->ev_gdone[0], c->ev_gdone[1], c->ev_gprep})
        if (e) (void)hipEventDestroy(e);
    c->reset_events();
    for (auto &e : c->free_events) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    (void)hipStreamDestroy(c->stream);
    delete c;
}

int ctbb_ctx_get_options(const ctbb_ctx *c, ctbb_options *o) {
    if (!c || !o) return CTBB_ERR_ARG;
    *o = c->opt;
    return CTBB_OK;
}

int ctbb_ctx_set_options(ctbb_ctx *c, const ctbb_options *o) {
    if (!c || !o) return CTBB_ERR_ARG;
    if (o->filter_mode < CTBB_FILTER_F32 || o->filter_mode > CTBB_FILTER_FFT32) return CTBB_ERR_ARG;
    if (o->proj_dtype != CTBB_PROJ_F32 && o->proj_dtype != CTBB_PROJ_BF16) return CTBB_ERR_ARG;
    if (o->bp_tile != 0 && o->bp_tile != 16 && o->bp_tile != 32) return CTBB_ERR_ARG;
    if (o->bp_views_per_pass != 0 && o->bp_views_per_pass != 1) return CTBB_ERR_ARG;
    if (!(o->adaptive_threshold >= 0.0f && o->adaptive_threshold < 3.0e38f)) return CTBB_ERR_ARG;
    if (o->batch_mode != CTBB_BATCH_PERSISTENT && o->batch_mode != CTBB_BATCH_PIPELINE) return CTBB_ERR_ARG;
    if (o->batch_group_mb < 0) return CTBB_ERR_ARG;
    c->opt = *o;
    return CTBB_OK;
}

void *ctbb_ctx_stream(ctbb_ctx *c) { return c ? static_cast<void *>(c->stream) : nullptr; }

int ctbb_ctx_synchronize(ctbb_ctx *c) {
    if (!c) return CTBB_ERR_ARG;
    if (c->stream_bp && hipStreamSynchronize(c->stream_bp) != hipSuccess) return CTBB_ERR_HIP;
    if (c->stream_head && hipStreamSynchronize(c->stream_head) != hipSuccess) return CTBB_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return CTBB_ERR_HIP;
    for (hipGraphExec_t g : c->live_graphs) (void)hipGraphExecDestroy(g);
    c->live_graphs.clear();
    return take_nonfinite(c);
}

int ctbb_ctx_kernel_time(ctbb_ctx *c, const char *family, float *total_ms, int *launches) {
    if (!c || !family || !total_ms || !launches) return CTBB_ERR_ARG;
    // (the streams only: a pending non-finite report stays for synchronize)
    if ((c->stream_bp && hipStreamSynchronize(c->stream_bp) != hipSuccess) ||
        (c->stream_head && hipStreamSynchronize(c->stream_head) != hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    *total_ms = 0.0f;
    *launches = 0;
    if (c->timing_failed) return CTBB_ERR_HIP;
    auto it = c->events.find(family);
    if (it == c->events.end()) return CTBB_OK;
    for (auto &e : it->second) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, e.a, e.b) != hipSuccess) return CTBB_ERR_HIP;
        *total_ms += ms;
        *launches += 1;
    }
    return CTBB_OK;
}

int ctbb_adaptive_filter_dev(ctbb_ctx *c, const struct recon_metadata *mr, const float *d_raw, float *d_out) {
    if (!c || !mr || !d_raw || !d_out) return CTBB_ERR_ARG;
    int st = check_geometry(mr);
    if (st) return st;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    return do_adaptive(c, mr, d_raw, d_out);
}

int ctbb_rebin_dev(ctbb_ctx *c, const struct recon_metadata *mr, const float *d_raw, float *d_rebin) {
    if (!c || !mr || !d_raw || !d_rebin) return CTBB_ERR_ARG;
    int st = check_geometry(mr);
    if (st) return st;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    return do_rebin(c, mr, d_raw, d_rebin);
}

int ctbb_filter_rows_dev(ctbb_ctx *c, const float *d_taps, int n, size_t n_rows_total, const float *d_in,
                         void *d_out) {
    if (!c || !d_taps || !d_in || !d_out || n <= 0) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    if (c->opt.filter_mode == CTBB_FILTER_FFT32) {
        // the spectrum of these taps (the cached one belongs to a prm's filter)
        std::vector<float> t(2 * static_cast<size_t>(n));
        if (hipStreamSynchronize(c->stream) != hipSuccess ||
            hipMemcpy(t.data(), d_taps, t.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
            return CTBB_ERR_HIP;
        int st = make_spec(c, t.data(), t.size());
        if (st) return st;
        c->taps_key.clear();
    }
    return do_filter(c, d_taps, n, n_rows_total, d_in, d_out, c->opt.proj_dtype == CTBB_PROJ_BF16);
}

int ctbb_rebin_filter_dev(ctbb_ctx *c, const struct recon_metadata *mr, const float *d_raw, void *d_rebin) {
    if (!c || !mr || !d_raw || !d_rebin) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    return rebin_filter_dev(c, mr, d_raw, d_rebin);
}

int ctbb_backproject_dev(ctbb_ctx *c, const struct recon_metadata *mr, const void *d_rebin, float *d_image_block) {
    if (!c || !mr || !d_rebin || !d_image_block) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    return backproject_dev(c, mr, d_rebin, d_image_block);
}

// Same as ctbb_backproject_dev with the view table in the device array
// `scratch` (through this device's address, looked up at context creation):
// no allocation, no event timing.
int ctbb_backproject_dev_noalloc(ctbb_ctx *c, const struct recon_metadata *mr, const void *d_rebin,
                                 float *d_image_block) {
    if (!c || !mr || !d_rebin || !d_image_block || !mr->tube_angles || !c->noalloc_vtab) return CTBB_ERR_ARG;
    int st = check_geometry(mr);
    if (st) return st;
    const VtabJob j = vtab_job(*mr, 0);
    if (j.n_ht <= 0 || j.half <= 0) return CTBB_ERR_SHAPE;
    if (vtab_len(j) > static_cast<size_t>(CTBB_NOALLOC_MAX_VIEWS)) return CTBB_ERR_LIMIT;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    st = launch_view_table(c->noalloc_vtab, j.tube_start, j.half, j.n_ht, j.n_proj_turn, j.z_rot, c->stream);
    if (st) return st;
    return do_backproject(c, mr, d_rebin, c->noalloc_vtab, d_image_block, c->stream, false);
}

// Backprojection autotuner: bp_tile 16 (two z-groups) or 32 (one group).
// The shape that decides it is the staged window (rows, oversampled channels,
// voxel size through the footprint), the matrix, the number of views and the
// projection type; results are cached per shape in a text file of lines
// "<key> <tile>".
int ctbb_autotune_backprojection(ctbb_ctx *c, struct recon_metadata *mr, const char *cache_path, int *bp_tile_out) {
    if (!c || !mr || !mr->tube_angles || !bp_tile_out) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    if (mr->ri.n_blocks <= 0) return CTBB_ERR_ARG;
    // plan the middle block for the measurement; the caller's block state
    // (mr->ri) is restored on return
    const recon_info saved_ri = mr->ri;
    struct Restore {
        recon_metadata *m;
        const recon_info &ri;
        ~Restore() { m->ri = ri; }
    } restore{mr, saved_ri};
    ctbb::plan_block(mr, static_cast<int>(mr->ri.n_blocks / 2));
    int st = check_geometry(mr);
    if (st) return st;
    char key[256];
    std::snprintf(key, sizeof(key), "rows%zu_ch%zu_nx%u_ny%u_fov%.4f_ht%d_half%d_%s", mr->cg.n_rows,
                  mr->cg.n_channels_oversampled, mr->rp.nx, mr->rp.ny, static_cast<double>(mr->rp.recon_fov),
                  bp_n_half_turns(*mr), static_cast<int>(mr->cg.n_proj_turn / 2),
                  c->opt.proj_dtype == CTBB_PROJ_BF16 ? "bf16" : "f32");
    int best = 0;
    if (cache_path) {
        if (FILE *f = std::fopen(cache_path, "r")) {
            char k[256];
            int t = 0;
            while (std::fscanf(f, "%255s %d", k, &t) == 2)
                if (std::strcmp(k, key) == 0 && (t == 16 || t == 32)) best = t;
            std::fclose(f);
        }
    }
    if (!best) {
        ctbb_sizes sz;
        ctbb_block_sizes(mr, &sz);
        const size_t reb_floats = (sz.rebin * proj_elem_bytes(c) + 3) / 4;
        DevBuf reb, img;
        const int keep = c->opt.bp_tile;
        float best_ms = 0.0f;
        // the work does not depend on the values: a zero block
        if (!(st = reb.ensure(reb_floats * sizeof(float))) && !(st = img.ensure(sz.image_block * sizeof(float))))
            st = fill_zero(reb.as<float>(), reb_floats, c->stream);
        for (int tile : {16, 32}) {
            if (st) break;
            c->opt.bp_tile = tile;
            st = backproject_dev(c, mr, reb.p, img.as<float>());          // warm-up
            if (!st) st = hip_ok(hipStreamSynchronize(c->stream));
            c->reset_events();
            for (int r = 0; r < 3 && !st; r++) st = backproject_dev(c, mr, reb.p, img.as<float>());
            float ms = 0.0f;
            int n = 0;
            if (!st) st = ctbb_ctx_kernel_time(c, "backproject", &ms, &n);
            if (!st && n > 0 && (!best || ms / n < best_ms)) {
                best = tile;
                best_ms = ms / n;
            }
        }
        c->opt.bp_tile = keep;
        if (!st) st = hip_ok(hipStreamSynchronize(c->stream));
        reb.release();
        img.release();
        if (st) return st;
        if (!best) return CTBB_ERR_HIP;
        if (cache_path) {
            FILE *f = std::fopen(cache_path, "a");
            if (!f) return CTBB_ERR_IO;
            std::fprintf(f, "%s %d\n", key, best);
            if (std::fclose(f) != 0) return CTBB_ERR_IO;
        }
    }
    c->opt.bp_tile = best;
    *bp_tile_out = best;
    return CTBB_OK;
}

int ctbb_finalize_dev(ctbb_ctx *c, const struct recon_metadata *mr, const float *d_image, float *d_final) {
    if (!c || !mr || !d_image || !d_final) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    return do_finalize(c, mr, d_image, d_final);
}

int ctbb_hu_to_int16_dev(ctbb_ctx *c, const float *d_hu, size_t n, float slope, float intercept, int16_t *d_out) {
    if (!c || (n && (!d_hu || !d_out))) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    Hu16Args a{d_hu, d_out, n, slope, intercept};
    EventPair e = c->begin("hu16");
    const int st = launch_hu16(a, c->stream);
    c->end(e);
    return st;
}

int ctbb_reconstruct_dev(ctbb_ctx *c, struct recon_metadata *mr, const float *d_raw, size_t n_raw_readings,
                         float *d_image, float *d_final) {
    if (!c || !mr || !d_raw || !d_final || !mr->table_positions || !mr->tube_angles) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const ct_geom &cg = mr->cg;
    const int n_blocks = static_cast<int>(mr->ri.n_blocks);
    if (n_blocks <= 0) return CTBB_ERR_ARG;
    const size_t nxy = static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    const size_t frame = cg.n_rows_raw * cg.n_channels;
    int st;
    if (!d_image) {
        st = c->image_stack.ensure(mr->ri.n_slices_recon * nxy * sizeof(float));
        if (st) return st;
        d_image = c->image_stack.as<float>();
    }
    // Per-block view tables, built on the device up front.
    std::vector<VtabJob> jobs;
    std::vector<size_t> off(n_blocks);
    size_t total = 0;
    for (int b = 0; b < n_blocks; b++) {
        ctbb::plan_block(mr, b);
        st = check_geometry(mr);
        if (st) return st;
        jobs.push_back(vtab_job(*mr, total));
        if (jobs.back().n_ht <= 0) return CTBB_ERR_SHAPE;
        off[b] = total;
        total += vtab_len(jobs.back());
    }
    st = build_view_tables(c, jobs, total);
    if (st) return st;
    st = ensure_taps(c, mr);
    if (st) return st;
    st = ensure_maps(c, mr);
    if (st) return st;
    size_t max_rebin = 0;
    for (int b = 0; b < n_blocks; b++) {
        ctbb::plan_block(mr, b);
        ctbb_sizes sz;
        ctbb_block_sizes(mr, &sz);
        max_rebin = std::max(max_rebin, sz.rebin);
    }
    st = ensure_pipeline(c, max_rebin * proj_elem_bytes(c));
    if (st) return st;
    for (int b = 0; b < n_blocks; b++) {
        ctbb::plan_block(mr, b);
        ctbb_sizes sz;
        ctbb_block_sizes(mr, &sz);
        const long start = static_cast<long>(static_cast<int>(mr->ri.idx_pull_start));
        const long n = static_cast<long>(mr->ri.n_proj_pull);
        // the first block's adaptive filter and rebin on the head stream
        // (beside a previous call's last backprojection)
        hipStream_t head = nullptr;
        if (b == 0) {
            head = c->stream_head;
            if (hipStreamWaitEvent(head, c->ev_prep_done, 0) != hipSuccess) return CTBB_ERR_HIP;
        }
        hipStream_t ps = head ? head : c->stream;
        const float *win;
        if (start >= 0 && start + n <= static_cast<long>(n_raw_readings)) {
            win = d_raw + static_cast<size_t>(start) * frame;
        } else {
            // pull reaches outside the scan: zero-padded copy
            DevBuf &pad = head ? c->head_pad : c->raw_in;
            st = pad.ensure(static_cast<size_t>(n) * frame * sizeof(float));
            if (st) return st;
            float *dst = pad.as<float>();
            st = fill_zero(dst, static_cast<size_t>(n) * frame, ps);
            if (st) return st;
            long lo = start < 0 ? 0 : start, hi = start + n;
            if (hi > static_cast<long>(n_raw_readings)) hi = static_cast<long>(n_raw_readings);
            if (hi > lo &&
                hipMemcpyAsync(dst + static_cast<size_t>(lo - start) * frame, d_raw + static_cast<size_t>(lo) * frame,
                               static_cast<size_t>(hi - lo) * frame * sizeof(float), hipMemcpyDeviceToDevice,
                               ps) != hipSuccess)
                return CTBB_ERR_HIP;
            win = dst;
        }
        st = pipeline_block(c, mr, b, b, win, c->vtab.as<BpView>() + off[b], d_image, head);
        if (st) return st;
    }
    // the next call's head starts once every preparation of this call is done
    // (so it runs beside this call's last backprojection, not beside a
    // preparation) and the head's workspaces have been read
    if (hipEventRecord(c->ev_prep_done, c->stream) != hipSuccess) return CTBB_ERR_HIP;
    st = join_pipeline(c);
    if (st) return st;
    return do_finalize(c, mr, d_image, d_final);
}

int ctbb_reconstruct_host(ctbb_ctx *c, struct recon_metadata *mr, const float *h_raw, size_t n_raw_readings,
                          float *h_final) {
    if (!c || !mr || !h_raw || !h_final || !mr->table_positions || !mr->tube_angles) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const ct_geom &cg = mr->cg;
    const int n_blocks = static_cast<int>(mr->ri.n_blocks);
    if (n_blocks <= 0) return CTBB_ERR_ARG;
    const size_t nxy = static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    const size_t frame = cg.n_rows_raw * cg.n_channels;
    int st = c->image_stack.ensure(mr->ri.n_slices_recon * nxy * sizeof(float));
    if (st) return st;
    float *d_image = c->image_stack.as<float>();
    std::vector<VtabJob> jobs;
    std::vector<size_t> off(n_blocks);
    size_t total = 0;
    std::vector<long> start(n_blocks), count(n_blocks);
    size_t max_pull = 0;
    for (int b = 0; b < n_blocks; b++) {
        ctbb::plan_block(mr, b);
        st = check_geometry(mr);
        if (st) return st;
        jobs.push_back(vtab_job(*mr, total));
        if (jobs.back().n_ht <= 0) return CTBB_ERR_SHAPE;
        off[b] = total;
        total += vtab_len(jobs.back());
        start[b] = static_cast<long>(static_cast<int>(mr->ri.idx_pull_start));
        count[b] = static_cast<long>(mr->ri.n_proj_pull);
        max_pull = std::max(max_pull, static_cast<size_t>(count[b]));
    }
    st = build_view_tables(c, jobs, total);
    if (st) return st;
    st = ensure_taps(c, mr);
    if (st) return st;
    st = ensure_maps(c, mr);
    if (st) return st;
    if (!c->copy_stream) {
        if (hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking) != hipSuccess) return CTBB_ERR_HIP;
        for (int i = 0; i < 2; i++)
            if (hipEventCreateWithFlags(&c->ev_ready[i], hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&c->ev_free[i], hipEventDisableTiming) != hipSuccess)
                return CTBB_ERR_HIP;
    }
    for (int i = 0; i < 2; i++) {
        st = c->pull[i].ensure(max_pull * frame * sizeof(float));
        if (st) return st;
    }
    size_t max_rebin = 0;
    for (int b = 0; b < n_blocks; b++) {
        ctbb::plan_block(mr, b);
        ctbb_sizes sz;
        ctbb_block_sizes(mr, &sz);
        max_rebin = std::max(max_rebin, sz.rebin);
    }
    st = ensure_pipeline(c, max_rebin * proj_elem_bytes(c));
    if (st) return st;
    // the copy stream must not overwrite a buffer before earlier work on the
    // compute stream (a previous call) is done with it
    if (hipEventRecord(c->ev_free[0], c->stream) != hipSuccess ||
        hipEventRecord(c->ev_free[1], c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    const size_t fb = frame * sizeof(float);
    for (int b = 0; b < n_blocks; b++) {
        const int k = b & 1;
        float *dst = c->pull[k].as<float>();
        if (hipStreamWaitEvent(c->copy_stream, c->ev_free[k], 0) != hipSuccess) return CTBB_ERR_HIP;
        // readings [start, start + count): the part shared with the previous
        // pull comes from the other buffer (device to device), the rest from
        // the host scan, and readings outside the scan are zeros
        long lo_prev = 0, hi_prev = 0;
        if (b > 0) {
            lo_prev = std::max(start[b], start[b - 1]);
            hi_prev = std::min(start[b] + count[b], start[b - 1] + count[b - 1]);
        }
        if (hi_prev > lo_prev &&
            hipMemcpyAsync(dst + static_cast<size_t>(lo_prev - start[b]) * frame,
                           c->pull[k ^ 1].as<float>() + static_cast<size_t>(lo_prev - start[b - 1]) * frame,
                           static_cast<size_t>(hi_prev - lo_prev) * fb, hipMemcpyDeviceToDevice,
                           c->copy_stream) != hipSuccess)
            return CTBB_ERR_HIP;
        // the rest: [start, lo_prev) and [max(hi_prev, start), end) in reading space
        const long seg[2][2] = {{start[b], hi_prev > lo_prev ? lo_prev : start[b] + count[b]},
                                {hi_prev > lo_prev ? hi_prev : start[b] + count[b], start[b] + count[b]}};
        for (const auto &sg : seg) {
            const long a0 = sg[0], a1 = sg[1];
            if (a1 <= a0) continue;
            const long in0 = std::max<long>(a0, 0), in1 = std::min<long>(a1, static_cast<long>(n_raw_readings));
            if (in0 > a0) {
                st = fill_zero(dst + static_cast<size_t>(a0 - start[b]) * frame,
                               static_cast<size_t>(std::min(in0, a1) - a0) * frame, c->copy_stream);
                if (st) return st;
            }
            if (in1 > in0 &&
                hipMemcpyAsync(dst + static_cast<size_t>(in0 - start[b]) * frame, h_raw + static_cast<size_t>(in0) * frame,
                               static_cast<size_t>(in1 - in0) * fb, hipMemcpyHostToDevice, c->copy_stream) != hipSuccess)
                return CTBB_ERR_HIP;
            const long z0 = std::max(in1, a0);
            if (a1 > z0) {
                st = fill_zero(dst + static_cast<size_t>(z0 - start[b]) * frame, static_cast<size_t>(a1 - z0) * frame,
                               c->copy_stream);
                if (st) return st;
            }
        }
        if (hipEventRecord(c->ev_ready[k], c->copy_stream) != hipSuccess ||
            hipStreamWaitEvent(c->stream, c->ev_ready[k], 0) != hipSuccess)
            return CTBB_ERR_HIP;
        ctbb::plan_block(mr, b);
        st = pipeline_block(c, mr, b, b, dst, c->vtab.as<BpView>() + off[b], d_image);
        if (st) return st;
        // the pull in buffer k has been read once this block's preparation
        // on ctx->stream is done (the backprojection reads the filtered
        // projections); the next block copies its overlap out of it, so that
        // copy is also ordered before buffer k is refilled (copy_stream)
        if (hipEventRecord(c->ev_free[k], c->stream) != hipSuccess) return CTBB_ERR_HIP;
    }
    st = join_pipeline(c);
    if (st) return st;
    const size_t final_bytes = static_cast<size_t>(ctbb_n_slices_final(mr)) * nxy * sizeof(float);
    st = c->raw_in.ensure(final_bytes);
    if (st) return st;
    st = do_finalize(c, mr, d_image, c->raw_in.as<float>());
    if (st) return st;
    if (hipMemcpyAsync(h_final, c->raw_in.p, final_bytes, hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return CTBB_ERR_HIP;
    return take_nonfinite(c);
}

}  // extern "C"

namespace {

int batch_pipeline(ctbb_ctx *c, int n_scans, struct recon_metadata *const *mrs, const float *const *d_raws,
                   const size_t *n_raw_readings, float *const *d_finals);

// Copy host bytes to a device address and wait (the host buffer may be
// reused when this returns).
int upload_to(ctbb_ctx *ctx, void *dst, const void *src, size_t bytes) {
    if (!bytes) return CTBB_OK;
    if (hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) return CTBB_ERR_HIP;
    return hip_ok(hipStreamSynchronize(ctx->stream));
}

// Raw pull of the current block of scan mr (mr->ri), padded with zeros where
// it reaches outside the scan's readings [0, n_raw) (into ctx->raw_in).
int block_window(ctbb_ctx *c, recon_metadata *m, const float *d_raw, size_t n_raw, const float **win) {
    const size_t frame = m->cg.n_rows_raw * m->cg.n_channels;
    const long start = static_cast<long>(static_cast<int>(m->ri.idx_pull_start));
    const long n = static_cast<long>(m->ri.n_proj_pull);
    if (start >= 0 && start + n <= static_cast<long>(n_raw)) {
        *win = d_raw + static_cast<size_t>(start) * frame;
        return CTBB_OK;
    }
    int st = c->raw_in.ensure(static_cast<size_t>(n) * frame * sizeof(float));
    if (st) return st;
    float *dst = c->raw_in.as<float>();
    st = fill_zero(dst, static_cast<size_t>(n) * frame, c->stream);
    if (st) return st;
    long lo = start < 0 ? 0 : start, hi = start + n;
    if (hi > static_cast<long>(n_raw)) hi = static_cast<long>(n_raw);
    if (hi > lo && hipMemcpyAsync(dst + static_cast<size_t>(lo - start) * frame, d_raw + static_cast<size_t>(lo) * frame,
                                  static_cast<size_t>(hi - lo) * frame * sizeof(float), hipMemcpyDeviceToDevice,
                                  c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    *win = dst;
    return CTBB_OK;
}

// Views the preparation rebins and filters for the current block (see
// prep_block): the backprojection's views plus, for an odd row count, the
// FFT row-pair partner.
size_t prep_views(const recon_metadata *m) {
    ctbb_sizes sz;
    ctbb_block_sizes(m, &sz);
    size_t views = std::min(sz.n_proj_out, static_cast<size_t>(m->cg.n_proj_turn / 2) *
                                               static_cast<size_t>(std::max(bp_n_half_turns(*m), 0)));
    if (((views * m->cg.n_rows) & 1) && views < sz.n_proj_out) views += 1;
    return views;
}

// Planner input of scan m (table positions at d_tpos, its blocks numbered
// from block_first in the planner's output); image and filtered-block
// pointers are set by the caller.
ScanPlanIn scan_plan_in(const recon_metadata *m, const BpArgs &tmpl, const double *d_tpos, int block_first) {
    ScanPlanIn in{};
    in.table_positions = d_tpos;
    in.tmpl = tmpl;
    in.recon_start_pos = m->ri.recon_start_pos;
    in.coll_slicewidth = m->rp.coll_slicewidth;
    in.tube_angle0 = m->tube_angles[0];
    in.z_rot = m->cg.z_rot;
    in.n_readings = static_cast<int>(m->rp.n_readings);
    // direction_of / recon_direction_of of the host planner (setup.cpp)
    const double dt = m->table_positions[100] - m->table_positions[0];
    in.array_direction = dt == 0.0 ? 1 : (dt > 0.0 ? 1 : -1);
    const float dpos = m->rp.end_pos - m->rp.start_pos;
    in.recon_direction = dpos == 0.0f ? 1 : (dpos > 0.0f ? 1 : -1);
    in.n_blocks = static_cast<int>(m->ri.n_blocks);
    in.n_slices_block = static_cast<int>(m->ri.n_slices_block);
    in.block_first = block_first;
    in.n_ffs = static_cast<int>(std::pow(2, m->rp.z_ffs) * std::pow(2, m->rp.phi_ffs));
    in.ppb = static_cast<int>(m->cg.n_proj_ffs / 2);
    in.add_ffs = static_cast<int>(m->cg.add_projections_ffs);
    in.n_proj_ffs = static_cast<int>(m->cg.n_proj_ffs);
    in.add_projections = static_cast<int>(m->cg.add_projections);
    in.n_proj_turn = static_cast<int>(m->cg.n_proj_turn);
    in.add_projections_ffs = m->cg.add_projections_ffs;
    return in;
}

struct BatchScan {
    int s = 0, variant = 0, n_blocks = 0, views_max = 0;
    size_t slot_elems = 0, img_elems = 0, pull_max = 0, rebin_max = 0, views_total = 0;
    BpArgs tmpl{};
};

struct BatchGroup {
    std::vector<int> scans;       // indices into the batch's scan list, one kernel variant
    size_t slot_elems = 0, arena_elems = 0, img_elems = 0;
    int n_blocks = 0, views_max = 0;
    size_t views_total = 0;
};

// Persistent batching form (see ctbb_reconstruct_batch_dev).
int batch_persistent(ctbb_ctx *c, int n_scans, recon_metadata *const *mrs, const float *const *d_raws,
                     const size_t *n_raw, float *const *d_finals) {
    const size_t eb = proj_elem_bytes(c);
    std::vector<BatchScan> info(n_scans);
    size_t pull_max = 0, rebin_max = 0;
    for (int s = 0; s < n_scans; s++) {
        recon_metadata *m = mrs[s];
        BatchScan &bs = info[s];
        bs.s = s;
        bs.n_blocks = static_cast<int>(m->ri.n_blocks);
        for (int b = 0; b < bs.n_blocks; b++) {
            ctbb::plan_block(m, b);
            int st = check_geometry(m);
            if (st) return st;
            if (bp_n_half_turns(*m) <= 0) return CTBB_ERR_SHAPE;
            ctbb_sizes sz;
            ctbb_block_sizes(m, &sz);
            const size_t views = prep_views(m);
            bs.slot_elems = std::max(bs.slot_elems, views * m->cg.n_rows * m->cg.n_channels_oversampled);
            bs.views_max = std::max(bs.views_max, static_cast<int>(m->cg.n_proj_turn / 2) * bp_n_half_turns(*m));
            bs.views_total += static_cast<size_t>(m->cg.n_proj_turn / 2) * bp_n_half_turns(*m);
            pull_max = std::max(pull_max, static_cast<size_t>(m->ri.n_proj_pull) * m->cg.n_rows_raw * m->cg.n_channels);
            rebin_max = std::max(rebin_max, sz.rebin);
        }
        bs.slot_elems = (bs.slot_elems + 63) & ~static_cast<size_t>(63);      // 256-byte aligned slots
        bs.img_elems = static_cast<size_t>(m->ri.n_slices_recon) * m->rp.nx * m->rp.ny;
        int st = fill_bp_args(m, bs.tmpl);
        if (st) return st;
        bs.tmpl.nz = c->opt.bp_tile == 32 ? 32 : 16;
        if (c->opt.proj_dtype == CTBB_PROJ_BF16) bs.tmpl.rebin_bf16 = reinterpret_cast<const uint16_t *>(16);
        else bs.tmpl.rebin = reinterpret_cast<const float *>(16);   // placeholders for the variant choice
        bs.tmpl.vtab = reinterpret_cast<const BpView *>(16);
        bs.tmpl.out = reinterpret_cast<float *>(16);
        bs.variant = backproject_variant(bs.tmpl);
        if (bs.variant < 0) return -bs.variant;
    }
    // groups: scans of one kernel variant (in batch order within it) whose
    // filtered blocks and images fit the budget (a larger scan is a group of
    // its own)
    const size_t budget = (c->opt.batch_group_mb > 0 ? static_cast<size_t>(c->opt.batch_group_mb) : 8192) << 20;
    std::vector<int> order(n_scans);
    for (int s = 0; s < n_scans; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return info[a].variant < info[b].variant; });
    std::vector<BatchGroup> groups;
    for (int s : order) {
        const BatchScan &bs = info[s];
        const size_t bytes = bs.slot_elems * bs.n_blocks * eb + bs.img_elems * sizeof(float);
        BatchGroup *g = groups.empty() ? nullptr : &groups.back();
        if (!g || info[g->scans.front()].variant != bs.variant ||
            (g->arena_elems * eb + g->img_elems * sizeof(float) + bytes > budget)) {
            groups.emplace_back();
            g = &groups.back();
        }
        g->scans.push_back(s);
        g->arena_elems += bs.slot_elems * bs.n_blocks;
        g->img_elems += bs.img_elems;
        g->n_blocks += bs.n_blocks;
        g->views_max = std::max(g->views_max, bs.views_max);
        g->views_total += bs.views_total;
    }
    // the first group's preparation overlaps nothing: start with the smallest
    if (groups.size() > 1) {
        auto smallest = std::min_element(groups.begin(), groups.end(), [](const BatchGroup &x, const BatchGroup &y) {
            return x.n_blocks < y.n_blocks;
        });
        std::rotate(groups.begin(), smallest, smallest + 1);
    }
    size_t arena_max = 0, img_max = 0, views_all = 0;
    int blocks_all = 0;
    for (const BatchGroup &g : groups) {
        arena_max = std::max(arena_max, g.arena_elems);
        img_max = std::max(img_max, g.img_elems);
        views_all += g.views_total;
        blocks_all += g.n_blocks;
    }
    // device memory: arenas, preparation workspaces, view tables
    int st;
    for (int k = 0; k < 2 && k < static_cast<int>(groups.size()); k++) {
        if ((st = c->arena[k].ensure(arena_max * eb)) || (st = c->img_arena[k].ensure(img_max * sizeof(float))))
            return st;
    }
    if ((st = c->raw_af.ensure(pull_max * sizeof(float))) || (st = c->raw_in.ensure(pull_max * sizeof(float))) ||
        (st = c->rebin_tmp.ensure(rebin_max * sizeof(float))))
        return st;
    size_t np_max = 0;
    for (int s = 0; s < n_scans; s++)
        for (int b = 0; b < info[s].n_blocks; b++) {
            ctbb::plan_block(mrs[s], b);
            np_max = std::max(np_max, static_cast<size_t>(mrs[s]->ri.n_proj_pull));
        }
    if ((st = c->af_work.ensure(6 * np_max * sizeof(float)))) return st;
    if ((st = c->vtab.ensure(views_all * sizeof(BpView)))) return st;
    if ((st = ensure_pipeline(c, 0))) return st;
    for (hipEvent_t *e : {&c->ev_gdone[0], &c->ev_gdone[1], &c->ev_gprep})
        if (!*e && hipEventCreateWithFlags(e, hipEventDisableTiming) != hipSuccess) return CTBB_ERR_HIP;
    // host-built inputs of the device planner, one upload: per scan its
    // ScanPlanIn (in group order) and table positions; per group the plans,
    // jobs, view-table offsets and queue prefixes
    size_t tpos_total = 0;
    std::vector<size_t> tpos_off(n_scans);
    for (int s = 0; s < n_scans; s++) {
        tpos_off[s] = tpos_total;
        tpos_total += static_cast<size_t>(mrs[s]->rp.n_readings);
    }
    if ((st = c->tpos_buf.ensure(tpos_total * sizeof(double)))) return st;
    auto align = [](size_t v) { return (v + 255) & ~static_cast<size_t>(255); };
    struct GroupLayout {
        size_t scans, plans, jobs, vt_off, band_pre, vtab;
    };
    std::vector<GroupLayout> lay(groups.size());
    size_t bytes = 0, vtab_at = 0;
    for (size_t g = 0; g < groups.size(); g++) {
        const int nb = groups[g].n_blocks;
        lay[g].scans = bytes;
        bytes = align(bytes + groups[g].scans.size() * sizeof(ScanPlanIn));
        lay[g].plans = bytes;
        bytes = align(bytes + nb * sizeof(BlockPlan));
        lay[g].jobs = bytes;
        bytes = align(bytes + nb * sizeof(BpArgs));
        lay[g].vt_off = bytes;
        bytes = align(bytes + (nb + 1) * sizeof(size_t));
        lay[g].band_pre = bytes;
        bytes = align(bytes + 8 * (nb + 1) * sizeof(int));
        lay[g].vtab = vtab_at;
        vtab_at += groups[g].views_total;
    }
    if ((st = c->plan_buf.ensure(bytes))) return st;
    std::vector<char> host(bytes, 0);
    char *dplan = c->plan_buf.as<char>();
    for (size_t g = 0; g < groups.size(); g++) {
        const int k = static_cast<int>(g & 1);
        size_t slot_at = 0, img_at = 0;
        int block_first = 0;
        for (size_t i = 0; i < groups[g].scans.size(); i++) {
            const int s = groups[g].scans[i];
            const recon_metadata *m = mrs[s];
            const BatchScan &bs = info[s];
            ScanPlanIn in = scan_plan_in(m, bs.tmpl, c->tpos_buf.as<double>() + tpos_off[s], block_first);
            in.image = c->img_arena[k].as<float>() + img_at;
            if (c->opt.proj_dtype == CTBB_PROJ_BF16) {
                in.tmpl.rebin_bf16 = c->arena[k].as<uint16_t>() + slot_at;
                in.tmpl.rebin = nullptr;
            } else {
                in.tmpl.rebin = c->arena[k].as<float>() + slot_at;
                in.tmpl.rebin_bf16 = nullptr;
            }
            in.slot_elems = bs.slot_elems;
            std::memcpy(host.data() + lay[g].scans + i * sizeof(ScanPlanIn), &in, sizeof(in));
            slot_at += bs.slot_elems * bs.n_blocks;
            img_at += bs.img_elems;
            block_first += bs.n_blocks;
        }
    }
    if ((st = upload_to(c, c->plan_buf.p, host.data(), bytes))) return st;
    for (int s = 0; s < n_scans; s++)
        if ((st = upload_to(c, c->tpos_buf.as<double>() + tpos_off[s], mrs[s]->table_positions,
                            static_cast<size_t>(mrs[s]->rp.n_readings) * sizeof(double))))
            return st;
    // finalize tap lists of every scan
    std::vector<char> fin;
    std::vector<size_t> fin_off(n_scans);
    std::vector<int> fin_mt(n_scans), fin_nf(n_scans);
    for (int s = 0; s < n_scans; s++) {
        std::vector<char> blob;
        if ((st = finalize_blob(mrs[s], blob, &fin_mt[s], &fin_nf[s]))) return st;
        fin_off[s] = align(fin.size());
        fin.resize(fin_off[s] + blob.size());
        std::memcpy(fin.data() + fin_off[s], blob.data(), blob.size());
    }
    if ((st = upload(c, c->fin_tab, fin.data(), fin.size()))) return st;
    // the device work list of every group: plans, jobs, view tables, queues
    for (size_t g = 0; g < groups.size(); g++) {
        st = launch_plan_batch(reinterpret_cast<const ScanPlanIn *>(dplan + lay[g].scans),
                               static_cast<int>(groups[g].scans.size()), groups[g].n_blocks, groups[g].views_max,
                               reinterpret_cast<BlockPlan *>(dplan + lay[g].plans),
                               reinterpret_cast<BpArgs *>(dplan + lay[g].jobs), c->vtab.as<BpView>() + lay[g].vtab,
                               reinterpret_cast<size_t *>(dplan + lay[g].vt_off),
                               reinterpret_cast<int *>(dplan + lay[g].band_pre), c->stream);
        if (st) return st;
    }
    // the item counters of every persistent launch (the first group runs one
    // launch per block, every later group one), zeroed behind the work list
    const size_t n_launch = static_cast<size_t>(groups[0].n_blocks) + groups.size() - 1;
    if ((st = c->bp_counters.ensure(8 * n_launch * sizeof(int))) ||
        (st = fill_zero(c->bp_counters.as<float>(), 8 * n_launch, c->stream)))
        return st;
    // the backprojection stream starts after the work list
    if (hipEventRecord(c->ev_sync, c->stream) != hipSuccess || hipStreamWaitEvent(c->stream_bp, c->ev_sync, 0) != hipSuccess)
        return CTBB_ERR_HIP;
    const bool split = c->opt.proj_dtype != CTBB_PROJ_BF16 && c->opt.filter_mode == CTBB_FILTER_FFT32;
    size_t launch = 0;
    struct JobRef {
        int s, b;
        size_t slot;        // element offset of the block's filtered projections in the arena
    };
    for (size_t g = 0; g < groups.size(); g++) {
        const int k = static_cast<int>(g & 1);
        // preparation into arena k once the group that used it before (g - 2)
        // is backprojected and finalized
        if (g >= 2 && hipStreamWaitEvent(c->stream, c->ev_gdone[k], 0) != hipSuccess) return CTBB_ERR_HIP;
        std::vector<JobRef> jobs;
        size_t slot_at = 0;
        for (int s : groups[g].scans) {
            for (int b = 0; b < info[s].n_blocks; b++) jobs.push_back({s, b, slot_at + b * info[s].slot_elems});
            slot_at += info[s].slot_elems * info[s].n_blocks;
        }
        // Job ranges, each prepared (one captured graph) and then
        // backprojected (one persistent launch).  A group is one range,
        // prepared while the previous group is backprojected, except the
        // first group: its preparation has nothing to overlap, so it runs
        // one block per range, each block prepared beside the previous
        // block's backprojection (beside the backprojection the adaptive
        // filter and the rebin of a block take about 0.9 of the block's
        // backprojection time, C1: 5.4 of 6 ms, so larger ranges would not
        // be ready in time).
        std::vector<std::pair<int, int>> ranges;
        const int nj = static_cast<int>(jobs.size());
        if (g == 0) {
            for (int j0 = 0; j0 < nj; j0++) ranges.push_back({j0, j0 + 1});
        } else {
            ranges.push_back({0, nj});
        }
        const BatchScan &s0 = info[groups[g].scans.front()];
        for (const auto &rg : ranges) {
            if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return CTBB_ERR_HIP;
            c->capturing = true;
            auto slot_ptr = [&](const JobRef &j) { return c->arena[k].as<char>() + j.slot * eb; };
            if (split) {
                // FP32 projections, FFT filter: first the adaptive filter and
                // the rebinning of every block, straight into its slot
                // (low-register kernels that run beside the backprojection
                // before; the first kernel of each block capped as in
                // prep_block), then the filter of every slot in place (the
                // FFT kernel needs the registers the backprojection holds)
                for (int j = rg.first; j < rg.second && !st; j++) {
                    recon_metadata *m = mrs[jobs[j].s];
                    ctbb::plan_block(m, jobs[j].b);
                    const float *win = nullptr;
                    st = block_window(c, m, d_raws[jobs[j].s], n_raw[jobs[j].s], &win);
                    const unsigned char *hot = nullptr;
                    if (!st && c->opt.adaptive_filter)
                        st = do_adaptive(c, m, win, c->raw_af.as<float>(), false, nullptr, nullptr, &hot,
                                         c->prep_groups);
                    if (!st)
                        st = do_rebin(c, m, win, reinterpret_cast<float *>(slot_ptr(jobs[j])),
                                      static_cast<int>(prep_views(m)), nullptr, hot,
                                      hot ? c->raw_af.as<float>() : nullptr,
                                      c->opt.adaptive_filter ? 0 : c->prep_groups);
                }
                for (int j = rg.first; j < rg.second && !st; j++) {
                    recon_metadata *m = mrs[jobs[j].s];
                    ctbb::plan_block(m, jobs[j].b);
                    float *slot = reinterpret_cast<float *>(slot_ptr(jobs[j]));
                    st = do_filter(c, c->taps.as<float>(), static_cast<int>(m->cg.n_channels_oversampled),
                                   prep_views(m) * m->cg.n_rows, slot, slot, 0);
                }
            } else {
                for (int j = rg.first; j < rg.second && !st; j++) {
                    recon_metadata *m = mrs[jobs[j].s];
                    ctbb::plan_block(m, jobs[j].b);
                    const float *win = nullptr;
                    st = block_window(c, m, d_raws[jobs[j].s], n_raw[jobs[j].s], &win);
                    if (!st) st = prep_block(c, m, win, slot_ptr(jobs[j]));
                }
            }
            c->capturing = false;
            hipGraph_t graph = nullptr;
            const hipError_t ec = hipStreamEndCapture(c->stream, &graph);
            if (st || ec != hipSuccess || !graph) {
                if (graph) (void)hipGraphDestroy(graph);
                return st ? st : CTBB_ERR_HIP;
            }
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess || !exec) return CTBB_ERR_HIP;
            c->live_graphs.push_back(exec);
            EventPair e = c->begin("prepare", c->stream);
            if (hipGraphLaunch(exec, c->stream) != hipSuccess) return CTBB_ERR_HIP;
            c->end(e);
            if (hipEventRecord(c->ev_gprep, c->stream) != hipSuccess ||
                hipStreamWaitEvent(c->stream_bp, c->ev_gprep, 0) != hipSuccess)
                return CTBB_ERR_HIP;
            // one persistent backprojection over every (block, tile) of the range
            BpQueues Q{};
            Q.jobs = reinterpret_cast<const BpArgs *>(dplan + lay[g].jobs);
            Q.band_pre = reinterpret_cast<const int *>(dplan + lay[g].band_pre);
            Q.n_jobs = groups[g].n_blocks;
            Q.job_first = rg.first;
            Q.job_end = rg.second;
            Q.counters = c->bp_counters.as<int>() + 8 * launch++;
            int total = 0;
            for (int j = rg.first; j < rg.second; j++) total += info[jobs[j].s].tmpl.n_tiles;
            EventPair eb2 = c->begin("backproject", c->stream_bp);
            st = launch_backproject_persistent(s0.tmpl, s0.variant, Q, total, 0, c->stream_bp);
            c->end(eb2);
            if (st) return st;
        }
        // finalize every scan of the group behind it
        size_t img_at = 0;
        for (int s : groups[g].scans) {
            st = launch_finalize_blob(c, mrs[s], c->img_arena[k].as<float>() + img_at, d_finals[s],
                                      c->fin_tab.as<char>() + fin_off[s], fin_mt[s], fin_nf[s], c->stream_bp);
            if (st) return st;
            img_at += info[s].img_elems;
        }
        if (hipEventRecord(c->ev_gdone[k], c->stream_bp) != hipSuccess) return CTBB_ERR_HIP;
    }
    return join_pipeline(c);
}

}  // namespace

extern "C" {

int ctbb_plan_batch_dev(ctbb_ctx *c, int n_scans, struct recon_metadata *const *mrs, ctbb_block_plan *out,
                        int max_out) {
    if (!c || n_scans <= 0 || !mrs || !out) return CTBB_ERR_ARG;
    int n_blocks = 0;
    for (int s = 0; s < n_scans; s++) {
        const recon_metadata *m = mrs[s];
        if (!m || !m->table_positions || !m->tube_angles || m->ri.n_blocks <= 0 || m->rp.n_readings <= 100)
            return CTBB_ERR_ARG;
        n_blocks += static_cast<int>(m->ri.n_blocks);
    }
    if (n_blocks > max_out) return CTBB_ERR_LIMIT;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    // planner inputs (geometry templates are not needed for the plans)
    size_t tpos_total = 0;
    std::vector<size_t> tpos_off(n_scans);
    for (int s = 0; s < n_scans; s++) {
        tpos_off[s] = tpos_total;
        tpos_total += static_cast<size_t>(mrs[s]->rp.n_readings);
    }
    int st;
    DevBuf tpos, work;
    const size_t b_scans = n_scans * sizeof(ScanPlanIn), b_plans = n_blocks * sizeof(BlockPlan),
                 b_jobs = n_blocks * sizeof(BpArgs), b_off = (n_blocks + 1) * sizeof(size_t),
                 b_pre = 8 * (n_blocks + 1) * sizeof(int);
    if ((st = tpos.ensure(tpos_total * sizeof(double))) ||
        (st = work.ensure(b_scans + b_plans + b_jobs + b_off + b_pre + 4 * 256)))
        return st;
    char *w = work.as<char>();
    std::vector<ScanPlanIn> ins;
    int first = 0;
    for (int s = 0; s < n_scans; s++) {
        BpArgs t{};
        t.n_tiles = 0;
        ins.push_back(scan_plan_in(mrs[s], t, tpos.as<double>() + tpos_off[s], first));
        first += static_cast<int>(mrs[s]->ri.n_blocks);
    }
    st = upload_to(c, w, ins.data(), b_scans);
    for (int s = 0; s < n_scans && !st; s++)
        st = upload_to(c, tpos.as<double>() + tpos_off[s], mrs[s]->table_positions,
                       static_cast<size_t>(mrs[s]->rp.n_readings) * sizeof(double));
    if (st) return st;
    char *p_plans = w + ((b_scans + 255) & ~size_t(255));
    char *p_jobs = p_plans + ((b_plans + 255) & ~size_t(255));
    char *p_off = p_jobs + ((b_jobs + 255) & ~size_t(255));
    char *p_pre = p_off + ((b_off + 255) & ~size_t(255));
    // (no view tables: max_views_per_block 0; the table base only offsets
    // the jobs' pointers, which are not read back)
    st = launch_plan_batch(reinterpret_cast<const ScanPlanIn *>(w), n_scans, n_blocks, 0,
                           reinterpret_cast<BlockPlan *>(p_plans), reinterpret_cast<BpArgs *>(p_jobs),
                           reinterpret_cast<BpView *>(p_jobs), reinterpret_cast<size_t *>(p_off),
                           reinterpret_cast<int *>(p_pre), c->stream);
    if (st) return st;
    std::vector<BlockPlan> plans(n_blocks);
    if (hipMemcpyAsync(plans.data(), p_plans, b_plans, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    for (int j = 0; j < n_blocks; j++) {
        ctbb_block_plan &o = out[j];
        o.scan = plans[j].scan;
        o.block = plans[j].block;
        o.idx_pull_start = plans[j].pull_start;
        o.n_proj_pull = plans[j].n_pull;
        o.n_half_turns = plans[j].n_ht;
        o.idx_block_slice_start = plans[j].idx_slice_start;
        o.idx_block_slice_end = plans[j].idx_slice_end;
        o.block_slice_start = plans[j].slice_start;
        o.block_slice_end = plans[j].slice_end;
        o.tube_start = plans[j].tube_start;
    }
    return CTBB_OK;
}

int ctbb_reconstruct_batch_dev(ctbb_ctx *c, int n_scans, struct recon_metadata *const *mrs,
                               const float *const *d_raws, const size_t *n_raw_readings, float *const *d_finals) {
    if (!c || n_scans <= 0 || !mrs || !d_raws || !n_raw_readings || !d_finals) return CTBB_ERR_ARG;
    for (int s = 0; s < n_scans; s++) {
        const recon_metadata *m = mrs[s];
        if (!m || !d_raws[s] || !d_finals[s] || !m->table_positions || !m->tube_angles || m->ri.n_blocks <= 0 ||
            m->rp.n_readings <= 100)
            return CTBB_ERR_ARG;
    }
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    // The persistent form gains by running the blocks of many scans per
    // launch; a batch of one scan runs the per-block pipeline (C1: 48.7 ms
    // against 51.1 ms persistent, which adds the device planning and the
    // per-block graphs and launches of its first group).
    if (c->opt.batch_mode == CTBB_BATCH_PIPELINE || n_scans == 1)
        return batch_pipeline(c, n_scans, mrs, d_raws, n_raw_readings, d_finals);
    int st = ensure_taps(c, mrs[0]);
    if (st) return st;
    st = ensure_maps(c, mrs[0]);
    if (st) return st;
    const std::string taps0 = c->taps_key, maps0 = c->maps_key;
    for (int s = 1; s < n_scans; s++) {
        if ((st = ensure_taps(c, mrs[s])) || (st = ensure_maps(c, mrs[s]))) return st;
        if (c->taps_key != taps0 || c->maps_key != maps0) {
            // different scanners, kernels or focal-spot modes: one scan at a time
            for (int t = 0; t < n_scans; t++) {
                st = ctbb_reconstruct_dev(c, mrs[t], d_raws[t], n_raw_readings[t], nullptr, d_finals[t]);
                if (st) return st;
            }
            return CTBB_OK;
        }
    }
    c->reset_events();
    return batch_persistent(c, n_scans, mrs, d_raws, n_raw_readings, d_finals);
}

}  // extern "C"

namespace {

// Per-block pipeline form of the batch (batch_mode CTBB_BATCH_PIPELINE): one
// launch per block, block preparation overlapping the previous block's
// backprojection.
int batch_pipeline(ctbb_ctx *c, int n_scans, struct recon_metadata *const *mrs, const float *const *d_raws,
                   const size_t *n_raw_readings, float *const *d_finals) {
    for (int s = 0; s < n_scans; s++) {
        const recon_metadata *m = mrs[s];
        if (!m || !d_raws[s] || !d_finals[s] || !m->table_positions || !m->tube_angles || m->ri.n_blocks <= 0)
            return CTBB_ERR_ARG;
    }
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    // every scan must use the filter table and rebin maps of the first one
    int st = ensure_taps(c, mrs[0]);
    if (st) return st;
    st = ensure_maps(c, mrs[0]);
    if (st) return st;
    const std::string taps0 = c->taps_key, maps0 = c->maps_key;
    for (int s = 1; s < n_scans; s++) {
        st = ensure_taps(c, mrs[s]);
        if (st) return st;
        st = ensure_maps(c, mrs[s]);
        if (st) return st;
        if (c->taps_key != taps0 || c->maps_key != maps0) {
            // different scanners, kernels or focal-spot modes: one scan at a time
            for (int t = 0; t < n_scans; t++) {
                st = ctbb_reconstruct_dev(c, mrs[t], d_raws[t], n_raw_readings[t], nullptr, d_finals[t]);
                if (st) return st;
            }
            return CTBB_OK;
        }
    }
    c->reset_events();
    // view tables of every (scan, block), built on the device, and finalize
    // tap lists of every scan, uploaded once
    std::vector<VtabJob> jobs;
    size_t total = 0;
    std::vector<std::vector<size_t>> off(n_scans);
    std::vector<char> fin;
    std::vector<size_t> fin_off(n_scans);
    std::vector<int> fin_mt(n_scans), fin_nf(n_scans);
    size_t max_rebin = 0, max_image = 0;
    for (int s = 0; s < n_scans; s++) {
        recon_metadata *m = mrs[s];
        const int nb = static_cast<int>(m->ri.n_blocks);
        off[s].resize(nb);
        for (int b = 0; b < nb; b++) {
            ctbb::plan_block(m, b);
            st = check_geometry(m);
            if (st) return st;
            jobs.push_back(vtab_job(*m, total));
            if (jobs.back().n_ht <= 0) return CTBB_ERR_SHAPE;
            off[s][b] = total;
            total += vtab_len(jobs.back());
            ctbb_sizes sz;
            ctbb_block_sizes(m, &sz);
            max_rebin = std::max(max_rebin, sz.rebin);
        }
        max_image = std::max(max_image, static_cast<size_t>(m->ri.n_slices_recon) * m->rp.nx * m->rp.ny);
        std::vector<char> blob;
        st = finalize_blob(m, blob, &fin_mt[s], &fin_nf[s]);
        if (st) return st;
        fin_off[s] = (fin.size() + 255) & ~static_cast<size_t>(255);
        fin.resize(fin_off[s] + blob.size());
        std::memcpy(fin.data() + fin_off[s], blob.data(), blob.size());
    }
    st = build_view_tables(c, jobs, total);
    if (st) return st;
    st = upload(c, c->fin_tab, fin.data(), fin.size());
    if (st) return st;
    st = c->image_stack.ensure(max_image * sizeof(float));
    if (st) return st;
    st = c->image_stack2.ensure(max_image * sizeof(float));
    if (st) return st;
    st = ensure_pipeline(c, max_rebin * proj_elem_bytes(c));
    if (st) return st;
    // one pipeline over the (scan, block) items: preparation on ctx->stream,
    // backprojection and then each scan's finalize on stream_bp; scans
    // alternate between two image stacks, so finalize(s) and the
    // backprojections of scan s + 1 never touch the same stack
    int item = 0;
    for (int s = 0; s < n_scans; s++) {
        recon_metadata *m = mrs[s];
        const int nb = static_cast<int>(m->ri.n_blocks);
        const size_t frame = m->cg.n_rows_raw * m->cg.n_channels;
        float *d_image = (s & 1) ? c->image_stack2.as<float>() : c->image_stack.as<float>();
        for (int b = 0; b < nb; b++, item++) {
            ctbb::plan_block(m, b);
            const long start = static_cast<long>(static_cast<int>(m->ri.idx_pull_start));
            const long n = static_cast<long>(m->ri.n_proj_pull);
            const float *win;
            if (start >= 0 && start + n <= static_cast<long>(n_raw_readings[s])) {
                win = d_raws[s] + static_cast<size_t>(start) * frame;
            } else {
                // pull reaches outside the scan: zero-padded copy on ctx->stream
                st = c->raw_in.ensure(static_cast<size_t>(n) * frame * sizeof(float));
                if (st) return st;
                float *dst = c->raw_in.as<float>();
                st = fill_zero(dst, static_cast<size_t>(n) * frame, c->stream);
                if (st) return st;
                long lo = start < 0 ? 0 : start, hi = start + n;
                if (hi > static_cast<long>(n_raw_readings[s])) hi = static_cast<long>(n_raw_readings[s]);
                if (hi > lo && hipMemcpyAsync(dst + static_cast<size_t>(lo - start) * frame,
                                              d_raws[s] + static_cast<size_t>(lo) * frame,
                                              static_cast<size_t>(hi - lo) * frame * sizeof(float),
                                              hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                    return CTBB_ERR_HIP;
                win = dst;
            }
            st = pipeline_block(c, m, item, b, win, c->vtab.as<BpView>() + off[s][b], d_image);
            if (st) return st;
        }
        st = launch_finalize_blob(c, m, d_image, d_finals[s], c->fin_tab.as<char>() + fin_off[s], fin_mt[s],
                                  fin_nf[s], c->stream_bp);
        if (st) return st;
    }
    return join_pipeline(c);
}

}  // namespace

extern "C" {

// --- reference host-memory contract ---------------------------------------

int ctbb_adaptive_filter(ctbb_ctx *c, struct recon_metadata *mr) {
    if (!c || !mr || !mr->ctd.raw) return CTBB_ERR_ARG;
    int st = check_geometry(mr);
    if (st) return st;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const size_t n = mr->ri.n_proj_pull * mr->cg.n_rows_raw * mr->cg.n_channels;
    if ((st = c->raw_in.ensure(n * sizeof(float))) || (st = c->raw_af.ensure(n * sizeof(float)))) return st;
    if (hipMemcpyAsync(c->raw_in.p, mr->ctd.raw, n * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    st = do_adaptive(c, mr, c->raw_in.as<float>(), c->raw_af.as<float>());
    if (st) return st;
    if (hipMemcpyAsync(mr->ctd.raw, c->raw_af.p, n * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    return hip_ok(hipStreamSynchronize(c->stream));
}

int ctbb_rebin_filter(ctbb_ctx *c, struct recon_metadata *mr) {
    if (!c || !mr || !mr->ctd.raw) return CTBB_ERR_ARG;
    int st = check_geometry(mr);
    if (st) return st;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const size_t n = mr->ri.n_proj_pull * mr->cg.n_rows_raw * mr->cg.n_channels;
    if ((st = c->raw_in.ensure(n * sizeof(float)))) return st;
    if (hipMemcpyAsync(c->raw_in.p, mr->ctd.raw, n * sizeof(float), hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    ctbb_sizes sz;
    ctbb_block_sizes(mr, &sz);
    void *d = nullptr;
    if (hipMalloc(&d, sz.rebin * proj_elem_bytes(c)) != hipSuccess) return CTBB_ERR_NOMEM;
    st = rebin_filter_dev(c, mr, c->raw_in.as<float>(), d);
    if (st) {
        (void)hipFree(d);
        return st;
    }
    if (mr->flags.testing && mr->ctd.rebin && c->opt.proj_dtype == CTBB_PROJ_F32) {
        if (hipMemcpyAsync(mr->ctd.rebin, d, sz.rebin * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            return CTBB_ERR_HIP;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        (void)hipFree(d);
        return CTBB_ERR_HIP;
    }
    if ((st = take_nonfinite(c))) {
        (void)hipFree(d);
        return st;
    }
    mr->ctd.d_rebin = static_cast<float *>(d);
    return CTBB_OK;
}

int ctbb_backproject(ctbb_ctx *c, struct recon_metadata *mr) {
    if (!c || !mr || !mr->ctd.d_rebin || !mr->ctd.image || mr->ri.cb.block_idx < 1) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const size_t nblk = mr->ri.n_slices_block * static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    int st = c->image_block.ensure(nblk * sizeof(float));
    if (st) return st;
    st = backproject_dev(c, mr, mr->ctd.d_rebin, c->image_block.as<float>());
    if (st) return st;
    float *dst = mr->ctd.image + static_cast<size_t>(mr->ri.cb.block_idx - 1) * nblk;
    if (hipMemcpyAsync(dst, c->image_block.p, nblk * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        return CTBB_ERR_HIP;
    st = hip_ok(hipStreamSynchronize(c->stream));
    (void)hipFree(mr->ctd.d_rebin);
    mr->ctd.d_rebin = nullptr;
    return st;
}

int ctbb_finalize_image_stack(ctbb_ctx *c, struct recon_metadata *mr) {
    if (!c || !mr || !mr->ctd.image) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    c->reset_events();
    const size_t nxy = static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    const size_t n_img = mr->ri.n_slices_recon * nxy;
    const int nf = ctbb_n_slices_final(mr);
    int st = c->image_stack.ensure(n_img * sizeof(float));
    if (st) return st;
    DevBuf fin;
    if ((st = fin.ensure(static_cast<size_t>(nf) * nxy * sizeof(float)))) return st;
    if (hipMemcpyAsync(c->image_stack.p, mr->ctd.image, n_img * sizeof(float), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess) {
        fin.release();
        return CTBB_ERR_HIP;
    }
    st = do_finalize(c, mr, c->image_stack.as<float>(), fin.as<float>());
    if (st == CTBB_OK) {
        std::free(mr->ctd.final_image_stack);
        mr->ctd.final_image_stack = static_cast<float *>(std::calloc(static_cast<size_t>(nf) * nxy, sizeof(float)));
        if (!mr->ctd.final_image_stack) st = CTBB_ERR_NOMEM;
        else if (hipMemcpyAsync(mr->ctd.final_image_stack, fin.p, static_cast<size_t>(nf) * nxy * sizeof(float),
                                hipMemcpyDeviceToHost, c->stream) != hipSuccess)
            st = CTBB_ERR_HIP;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess && st == CTBB_OK) st = CTBB_ERR_HIP;
    fin.release();
    return st;
}

int ctbb_synth_scan_dev(ctbb_ctx *c, const struct recon_metadata *mr, const ctbb_phantom_params *pp, long first,
                        long count, float *d_out) {
    if (!c || !mr || !pp || !d_out) return CTBB_ERR_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return CTBB_ERR_HIP;
    return launch_synth(*mr, *pp, first, count, d_out, c->stream);
}

}  // extern "C"
