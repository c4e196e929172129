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
// Mapping: one workgroup = 16x16 voxel columns x the block's 32 slices, as
// two z-groups of NZ = 16 slices (interleaved by chunks of 4) or one of 32;
// one thread = one column (xi, yi) with its NZ slice accumulators in
// registers.  Views are
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
// than 1% across a wavefront, chunks of 4 slices that no lane sees are
// skipped on a wave-uniform vote (about 40% of the (slice, view) pairs are
// seen at the C1 geometry), chunks with |q| < 0.6 in every lane take a path
// with W = 1, and the cos^2 taper is evaluated only for slices where some
// lane is in 0.6 <= |q| < 1.  View-table entries reach the
// workgroup through a small LDS ring refilled every 64 views.
//
// Workgroups are remapped so the tiles running on one XCD are spatially
// contiguous: their footprints on each view overlap and hit in that XCD's L2.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "kernels.h"

namespace ctbb {

namespace {

// Taper part of the Stierstorfer weight W(q), Q = 0.6 (reference
// include/backproject.cuh:33-46), for 0.6 <= |q| < 1:
//   cos^2(pi'/2 t) = sin^2(pi'/2 (1 - t) + (pi - pi')/2),  t = (|q| - 0.6) / 0.4,
// pi' = 3.1415368979f.  The sine is an odd degree-7 polynomial v + v^3 P(v^2)
// fitted for relative error (1.2e-6 on [0, pi/2]), so the weight stays
// strictly positive up to |q| -> 1 (7.8e-10 there, as the reference's): a
// view that is the only one to see a voxel at an angle still gets
// s_t / h_t = its sample.  Error of W < 2.5e-6 of W.
__device__ __forceinline__ float stier_taper(float aq) {
    const float u = __builtin_amdgcn_fmed3f(fmaf(aq, -2.5f, 2.5f), 0.0f, 1.0f);   // 1 - t
    const float v = fmaf(u, 1.5707684f, 2.7877844e-05f);
    const float v2 = v * v;
    float p = -0.0001854245929280296f;
    p = fmaf(p, v2, 0.008314281702041626f);
    p = fmaf(p, v2, -0.16665853559970856f);
    const float sn = fmaf(v * v2, p, v);
    return sn * sn;
}

__device__ __forceinline__ float bf16_to_f32(uint16_t h) {
    return __uint_as_float(static_cast<uint32_t>(h) << 16);
}

// Raw bits of one filtered-projection sample (float32 bits, or bfloat16 bits
// in the low half).  Conversion happens when the prefetched window is written
// to LDS, so the loads stay in flight while the previous view is computed.
template <bool kBf16, class Args>
__device__ __forceinline__ uint32_t load_bits(Args &a, size_t i) {
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

// NZ slices per thread; CPT staged channels per thread per view; TPR staging
// threads per detector row (0: taken from the arguments at run time).  The
// staged window covers rows_pad x W channels, rows_pad = 256 / TPR a power of
// two >= n_rows, W = TPR * CPT: thread t stages row t / TPR, channels
// (t % TPR) * CPT ... + CPT - 1 from CPT + 1 consecutive samples of that row.
// LDS holds (value, delta) pairs row-major.  For even CPT every run of CPT
// pairs is followed by one pad pair (PAD = 1), for odd CPT there is none, so
// channel c of row r sits at r * RS + c + PAD * (c / CPT) with
// RS = TPR * (CPT + PAD): a wave's staging writes land at an odd stride of
// CPT + PAD pairs, 16 consecutive lanes on 16 different 8-byte bank pairs.
// Row rows_pad is kept zero so a row coordinate that rounds to exactly
// n_rows - 1 reads a finite neighbour (its weight is 0).
static_assert(sizeof(BpView) == 16 && alignof(BpView) <= 16, "view table entries are read as float4");

template <int CPT> constexpr int kPad = (CPT & 1) ? 0 : 1;

constexpr int ilog2(int v) { return v <= 1 ? 0 : 1 + ilog2(v >> 1); }

// ZS z-groups of NZ slices share one workgroup (256 * ZS threads): thread t
// takes column t % 256 and slices (t / 256) * NZ ... + NZ - 1 of the block.
// NZ = 32, ZS = 1 keeps 96 accumulators per thread (150 VGPRs, 3 waves per
// SIMD); NZ = 16, ZS = 2 halves them (120 VGPRs, 4 waves per SIMD) and
// shares one staged window between the two halves.
// One 16x16-column tile (all 32 slices) of the block described by `a`, by
// the whole workgroup; zblk selects the slice range when a launch splits
// the 32 slices over several workgroups (NZ * ZS < 32).  lds2: the dynamic
// LDS (2 window buffers x [rows_pad + 1][RS] pairs, a ring of 128
// view-table entries, with two z-groups 2 x 256 column geometries, and the
// ring's 128 window bases).
// The workgroup leaves LDS free for reuse once every thread has returned.
// Args: BpArgs, either a kernel argument or a job read through the scalar
// cache (constant address space), so its fields stay in scalar registers.
template <int NZ, int CPT, int TPR, int ZS, bool kBf16, class Args>
__device__ __forceinline__ void bp_tile(Args &a, int tile, int zblk, float2 *lds2) {
    constexpr int THREADS = 256 * ZS;
    const int tid = threadIdx.x;
    const int tx = tile / a.tiles_y, ty = tile - tx * a.tiles_y;
    // Column of this thread in the 16x16 tile.  Each wave covers 4 x-rows of
    // 16; within it each group of 16 lanes (one LDS cycle of a 64-bit read)
    // takes a 4x4 patch, whose projections spread over fewer window channels
    // than a 1x16 line, so their LDS reads meet fewer bank conflicts.
    const int col = tid & 255;
    const int xi = tx * 16 + 4 * (col >> 6) + ((col >> 2) & 3);
    const int yi = ty * 16 + 4 * ((col >> 4) & 3) + (col & 3);
    // Slices of this thread.  With two z-groups the groups interleave by
    // chunks of 4 (group g: slices 8m + 4g .. 8m + 4g + 3 of the block's 32),
    // so both see about the same number of visible chunks per view and
    // neither waits long at the view's barrier; chunk zc of the thread is
    // slice z0 + ZSTEP * zc + c.
    static_assert(ZS == 1 || NZ * ZS == 32, "two z-groups split the block's 32 slices");
    constexpr int ZSTEP = ZS == 2 ? 2 : 1;
    const int z0 = ZS == 2 ? zblk * 32 + 4 * (tid >> 8) : zblk * NZ;

    const float x = a.xs * (static_cast<float>(xi) - a.xc) + a.x0;
    const float y = a.ys * (static_cast<float>(yi) - a.yc) + a.y0;
    // centre of the tile
    const float xm = a.xs * (static_cast<float>(tx * 16) + 7.5f - a.xc) + a.x0;
    const float ym = a.ys * (static_cast<float>(ty * 16) + 7.5f - a.yc) + a.y0;

    const int tpr = TPR ? TPR : a.tpr;
    const int tpr_log2 = TPR ? ilog2(TPR) : a.tpr_log2;
    const int rows_pad = THREADS >> tpr_log2;
    const int W = tpr * CPT;                            // staged channels
    constexpr int PAD = kPad<CPT>;
    const int RS = tpr * (CPT + PAD);                   // pairs per staged row
    const int n_stage = (rows_pad + 1) * RS;            // pairs per buffer
    const int srow = tid >> tpr_log2;                   // staged row of this thread
    const int scol = (tid & (tpr - 1)) * CPT;           // first staged channel
    const bool srow_ok = srow < a.n_rows;

    // the zero row (row rows_pad) of both buffers
    for (int c = tid; c < RS; c += THREADS) {
        lds2[rows_pad * RS + c] = make_float2(0.0f, 0.0f);
        lds2[n_stage + rows_pad * RS + c] = make_float2(0.0f, 0.0f);
    }

    // prefetch registers for the next view's window (raw sample bits)
    uint32_t pv[CPT + 1];

    auto window_lo = [&](const float4 vw) {
        const float pm = (xm * vw.x - ym * vw.y) * a.inv_ds + a.two_cc;
        return static_cast<int>(floorf(pm - a.half_w)) - 1;
    };
    // byte offset of this thread's first sample inside a projection window
    // (row srow, channel scol)
    constexpr unsigned kEs = kBf16 ? 2u : 4u;
    const unsigned toff_b = static_cast<unsigned>(srow * a.nos + scol) * kEs;
    auto fetch = [&](int v, int p_lo) {
        // v and p_lo are the same in every lane: the window base stays in
        // scalar registers and a lane adds its 32-bit byte offset
        v = __builtin_amdgcn_readfirstlane(v);
        p_lo = __builtin_amdgcn_readfirstlane(p_lo);
        const int g0 = p_lo + scol;
        const size_t rb = (static_cast<size_t>(v) * a.n_rows + srow) * a.nos;
        // (bfloat16 reads whole dwords, up to 3 samples past the window)
        if (srow_ok && p_lo >= 0 && p_lo + W + (kBf16 ? 3 : 0) < a.nos) {
            // window inside the detector: CPT + 1 consecutive samples, read
            // through a buffer descriptor of projection v (raw buffer, gfx9
            // dword 3): per-lane offset toff_b, scalar offset p_lo, and the
            // sample index as the instruction's immediate offset
            const char *pb = (kBf16 ? reinterpret_cast<const char *>(a.rebin_bf16)
                                    : reinterpret_cast<const char *>(a.rebin)) +
                             static_cast<size_t>(v) * a.n_rows * a.nos * kEs;
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<char *>(pb), static_cast<short>(0), a.n_rows * a.nos * static_cast<int>(kEs), 0x00020000);
            const int soff = p_lo * static_cast<int>(kEs);
            if (kBf16) {
                // bfloat16: whole dwords from the dword-aligned start of the
                // lane's samples (vector and scalar offsets both even in
                // samples), then each sample pair re-aligned by 0 or 16 bits;
                // sub-dword loads are slower
                constexpr int ND = CPT / 2 + 2;
                const int lo_el = static_cast<int>(toff_b >> 1) + (p_lo & 1);   // past the even part of p_lo
                const int par = lo_el & 1;
                const int dbase = 2 * (lo_el - par);
                const int soff2 = 2 * (p_lo & ~1);
                uint32_t d[ND];
#pragma unroll
                for (int q = 0; q < ND; q++) {
                    d[q] = __builtin_amdgcn_raw_buffer_load_b32(rs, dbase + 4 * q, soff2, 0);
                    asm volatile("" ::: "memory");
                }
#pragma unroll
                for (int e = 0; e <= CPT; e += 2) {
                    const uint32_t pr = __builtin_amdgcn_alignbit(d[e / 2 + 1], d[e / 2], 16u * par);
                    pv[e] = pr & 0xffffu;
                    if (e + 1 <= CPT) pv[e + 1] = pr >> 16;
                }
            } else {
#pragma unroll
                for (int e = 0; e <= CPT; e++) {
                    pv[e] = __builtin_amdgcn_raw_buffer_load_b32(rs, static_cast<int>(toff_b) + e * 4, soff, 0);
                    // keep the loads separate: merged multi-dword loads land in
                    // register tuples that the (value, delta) pairs cannot use,
                    // and the copies out of them would wait for the loads here
                    asm volatile("" ::: "memory");
                }
            }
        } else {
#pragma unroll
            for (int e = 0; e <= CPT; e++) {
                const int g = g0 + e;
                pv[e] = (srow_ok && g >= 0 && g < a.nos) ? load_bits<kBf16>(a, rb + g) : 0u;  // +0.0 in either format
            }
        }
    };

    // s: the image; acc = (s_t, h_t + 1e-30) of the current angle.  A
    // nonzero weight is at least 7.8e-10, so starting h_t at 1e-30 instead
    // of adding it before the reciprocal gives the same bits, and an angle
    // no view sees ends with s_t = 0, h_t = 1e-30 and adds 0.
    float s[NZ];
    float2 acc[NZ];
#pragma unroll
    for (int z = 0; z < NZ; z++) { s[z] = 0.0f; acc[z] = make_float2(0.0f, 1e-30f); }

    // Views run in reference order t = i * n_ht + k (angle i, half turn k);
    // the table holds them in that order.  Pipeline: while view t is
    // computed, the window of view t + 1 is in flight.  Table entries pass
    // through an LDS ring of 128: chunk c (views 64c .. 64c + 63) is loaded
    // during view 64(c - 1), published by the next barrier and first read
    // 63 views later; its slot last held chunk c - 2.
    const int n_views = a.half * a.n_ht;
    const float4 *vtab = reinterpret_cast<const float4 *>(a.vtab);
    float4 *ring = reinterpret_cast<float4 *>(lds2 + 2 * n_stage);
    // Each entry's window base (the staged window's first channel for this
    // tile, the same in every lane) is computed once, by the thread that
    // loads the entry, into a parallel ring of ints.
    int *plo_ring = reinterpret_cast<int *>(ring + 128 + (ZS == 2 ? 512 : 0));
    auto load_chunk = [&](int c) {
        if (tid < 64) {
            const float4 e = vtab[min(64 * c + tid, n_views - 1)];
            ring[(c & 1) * 64 + tid] = e;
            plo_ring[(c & 1) * 64 + tid] = window_lo(e);
        }
    };

    // Per-column geometry of a view (window base plo), shared by the slices:
    // channel fraction fp, LDS column index ci of the bilinear tap (within
    // a window buffer), g = -td / (lhat tan(theta_cone/2)) and the z offset
    // dz of slice 0 of the block, so that q(z) = g (dz + z0 coll + z coll) is
    // qhat of reference include/backproject.cuh:84.
    auto geometry = [&](const float4 v, int plo, float &fp, unsigned &ci, float &g, float &dz) {
        const float phat = x * v.x - y * v.y;
        const float pl = phat * a.inv_ds + a.two_cc - static_cast<float>(plo);
        int c0 = static_cast<int>(floorf(pl));
        fp = pl - static_cast<float>(c0);
        c0 = min(max(c0, 0), W - 2);
        ci = static_cast<unsigned>(c0) + (PAD ? static_cast<unsigned>(c0) / CPT : 0u);
        const float xr = phat * a.inv_rf;
        float asn = asin_small(xr);
        if (!a.asin_small_only && __builtin_amdgcn_ballot_w64(fabsf(xr) > 0.45f))
            asn = fabsf(xr) > 0.45f ? asinf(xr) : asn;
        dz = fmaf(a.k_asin, asn, v.z);
        // hardware square root (1 ulp): no denormal scaling or correction steps
        const float lhat = __builtin_amdgcn_sqrtf(a.rf2 - phat * phat) - x * v.y - y * v.x;
        g = a.neg_td * __builtin_amdgcn_rcpf(lhat * a.tanc);
    };
    // With two z-groups the groups take turns: during view t, group
    // (t + 1) % 2 evaluates view t + 1 for its 256 columns and leaves it in
    // LDS slot (t + 1) % 2 (published by the next barrier), so each
    // evaluation serves both groups.
    float4 *geo = ring + 128;
    const int zg = tid >> 8;

    load_chunk(0);
    __syncthreads();
    int plo_next;
    {
        const float4 v0 = ring[0];
        plo_next = plo_ring[0];
        fetch(__float_as_int(v0.w), plo_next);
        if (ZS == 2 && zg == 0) {
            float fp, g, dz;
            unsigned ci;
            geometry(v0, plo_next, fp, ci, g, dz);
            geo[col] = make_float4(fp, __uint_as_float(ci), g, dz);
        }
    }

    int k = 0;                                          // half turn of view t
    for (int t = 0; t < n_views; t++) {
        const bool last_k = k == a.n_ht - 1;            // view t closes its angle
        if (++k == a.n_ht) k = 0;
        const int p_lo = plo_next;
        float2 *buf = lds2 + (t & 1) * n_stage;
        __builtin_amdgcn_s_setprio(1);     // staging writes, ahead of other waves' arithmetic
        {
            float2 *dst = buf + tid * (CPT + PAD);      // = srow * RS + scol + PAD * (scol / CPT)
            float prev = bits_to_f32<kBf16>(pv[0]);
#pragma unroll
            for (int e = 0; e < CPT; e++) {
                const float nxt = bits_to_f32<kBf16>(pv[e + 1]);
                dst[e] = make_float2(prev, nxt - prev);
                prev = nxt;
            }
        }
        __builtin_amdgcn_s_setprio(0);
        __syncthreads();
        if ((t & 63) == 0 && t + 64 < n_views) load_chunk((t >> 6) + 1);
        float fp, g, dz;
        unsigned ci;
        if (ZS == 2) {
            const float4 gv = geo[(t & 1) * 256 + col];
            fp = gv.x;
            ci = __float_as_uint(gv.y);
            g = gv.z;
            dz = gv.w;
        }
        const float4 vw = ring[t & 127];
        if (t + 1 < n_views) {
            const float4 vn = ring[(t + 1) & 127];
            plo_next = plo_ring[(t + 1) & 127];
            __builtin_amdgcn_s_setprio(1);
            fetch(__float_as_int(vn.w), plo_next);
            __builtin_amdgcn_s_setprio(0);
            if (ZS == 2 && zg == ((t + 1) & 1)) {
                float fpn, gn, dzn;
                unsigned cin;
                geometry(vn, plo_next, fpn, cin, gn, dzn);
                geo[((t + 1) & 1) * 256 + col] = make_float4(fpn, __uint_as_float(cin), gn, dzn);
            }
        }
        if (ZS == 1) geometry(vw, p_lo, fp, ci, g, dz);

        const float A = g * a.coll;
        const float B = g * fmaf(static_cast<float>(z0), a.coll, dz);
        // row coordinate (q + 1) hr = Ar z + Br, used directly where |q| < 0.6
        const float Ar = A * a.hr;
        const float Br = fmaf(B, a.hr, a.hr);
        // LDS index of (row 0, channel c0) in this view's buffer
        const unsigned colb = static_cast<unsigned>((t & 1) * n_stage) + ci;

        // Slices in chunks of 4.  q is linear in z, so over a chunk it stays
        // within hw = 1.5 |A| of its value qm at the chunk centre: a chunk can
        // be seen only if |qm| < 1 + hw, and W = 1 for all of it if
        // |qm| < 0.6 - hw.  Both thresholds carry a margin of 1e-4, far above
        // the rounding of q, so a skipped chunk has |q| > 1 in every lane and
        // an interior chunk |q| < 0.6.  One wave vote skips chunks no lane
        // sees; a second sends chunks interior for every lane (the common
        // case) through a path with W = 1 and no clamping.  Elsewhere W is
        // evaluated per slice, the taper only for slices where some lane has
        // 0.6 <= |q| < 1 (a third, per-slice vote).  The 4 slices of a
        // visible chunk then interpolate without branches (W = 0 for lanes
        // that do not see a slice, row coordinate clamped into the window),
        // so their LDS reads are in flight together.
        const float hw = 1.5f * fabsf(A);
        const float t_vis = 1.0001f + hw;
        const float t_int = 0.5999f - hw;
        const float two_hr = 2.0f * a.hr;
#pragma unroll
        for (int zc = 0; zc < NZ; zc += 4) {
            const float qm = fmaf(A, static_cast<float>(ZSTEP * zc) + 1.5f, B);
            if (__builtin_amdgcn_ballot_w64(fabsf(qm) < t_vis) == 0) continue;
            float rowf[4], w[4];
            if (__builtin_amdgcn_ballot_w64(!(fabsf(qm) < t_int)) == 0) {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    rowf[c] = fmaf(Ar, static_cast<float>(ZSTEP * zc + c), Br);
                    w[c] = 1.0f;
                }
            } else {
#pragma unroll
                for (int c = 0; c < 4; c++) {
                    const float q = fmaf(A, static_cast<float>(ZSTEP * zc + c), B);
                    const float aq = fabsf(q);
                    const bool vis = aq < 1.0f;
                    const bool band = vis && aq >= 0.6f;
                    float wc = vis ? 1.0f : 0.0f;
                    if (__builtin_amdgcn_ballot_w64(band)) {
                        const float tp = stier_taper(aq);
                        wc = band ? tp : wc;
                    }
                    w[c] = wc;
                    // row coordinate (q + 1) hr, in [0, n_rows - 1] where W != 0
                    rowf[c] = __builtin_amdgcn_fmed3f(fmaf(q, a.hr, a.hr), 0.0f, two_hr);
                }
            }
            float2 e0[4], e1[4];
            // the chunk's LDS reads (and, below, the next view's window
            // loads) are issued at raised wave priority, so waves start their
            // long-latency requests ahead of other waves' arithmetic
            __builtin_amdgcn_s_setprio(1);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const unsigned idx = colb + __umul24(static_cast<unsigned>(rowf[c]), static_cast<unsigned>(RS));
                e0[c] = lds2[idx];
                e1[c] = lds2[idx + RS];
            }
            // all 8 reads of the chunk are issued before the first is used
            __builtin_amdgcn_sched_barrier(0);
            __builtin_amdgcn_s_setprio(0);
#pragma unroll
            for (int c = 0; c < 4; c++) {
                const float fr = __builtin_amdgcn_fractf(rowf[c]);
                const float top = fmaf(fp, e0[c].y, e0[c].x);
                const float bot = fmaf(fp, e1[c].y, e1[c].x);
                const float val = fmaf(fr, bot - top, top);
                acc[zc + c].x = fmaf(val, w[c], acc[zc + c].x);
                acc[zc + c].y += w[c];
            }
        }
        if (last_k) {
            // s += (1/h_t) s_t where h_t != 0 (reference include/backproject.cuh:104-108);
            // h_t = 0 (then s_t = 0) is 0 * 1e30 = 0 without a select.
#pragma unroll
            for (int z = 0; z < NZ; z++) {
                s[z] = fmaf(acc[z].x, __builtin_amdgcn_rcpf(acc[z].y), s[z]);
                acc[z] = make_float2(0.0f, 1e-30f);
            }
        }
    }

    const bool in_fov = sqrtf(x * x + y * y) < a.acq_half;
    // slice bases in scalar registers (z0 is the same across a wave), the
    // column's 32-bit offset per lane
    const size_t nxy = static_cast<size_t>(a.nx) * a.ny;
    const int zw = __builtin_amdgcn_readfirstlane(z0);
    const unsigned lane_off = static_cast<unsigned>(xi * a.ny + yi);
#pragma unroll
    for (int z = 0; z < NZ; z++) {
        float *os = a.out + static_cast<size_t>(zw + ZSTEP * (z & ~3) + (z & 3)) * nxy;
        os[lane_off] = in_fov ? s[z] * a.out_scale : 0.0f;
    }
}

// One launch per block: slice range blockIdx.y; a grid of at most the
// resident workgroups (a multiple of 8 where n_tiles is), workgroup b taking
// tiles b, b + gridDim.x, ... — all of them placed at launch, so kernels
// started beside the backprojection cannot take the registers a later
// workgroup of it would need.  XCD-aware remap: hardware dispatches workgroup
// b to XCD b % 8, and item i maps to tile (i % 8) per + i / 8, so each XCD
// works through a contiguous band of tiles.
template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__device__ __forceinline__ void bp_block_body(const BpArgs &a, float2 *lds2) {
    const int per = (a.n_tiles & 7) == 0 ? a.n_tiles >> 3 : 0;
    for (int item = blockIdx.x; item < a.n_tiles; item += gridDim.x) {
        const int tile = per ? (item & 7) * per + (item >> 3) : item;
        if (item != static_cast<int>(blockIdx.x)) __syncthreads();   // LDS reuse by the next tile
        bp_tile<NZ, CPT, TPR, ZS, kBf16>(a, tile, blockIdx.y, lds2);
    }
}

// Two z-groups (512 threads, 4 waves per SIMD).  VGPR limits (the
// attribute counts register pairs on gfx950): 112 registers for windows of
// up to 8 channels per staging thread (bp_kernel2), so the 4 waves per SIMD
// leave 64 per lane to the next block's preparation kernels, which then run
// beside the backprojection (at 120 its 4 waves hold the whole register
// file and the preparation only gets the registers of retiring workgroups;
// the values the cap spills are reloaded once per tile; alone it runs as
// fast as at 120); 120 for the wider windows (bp_kernel2w), whose prefetch
// registers would spill in the view loop.
template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4), amdgpu_num_vgpr(56))) void bp_kernel2(BpArgs a) {
    extern __shared__ float2 lds2[];
    bp_block_body<NZ, CPT, TPR, ZS, kBf16>(a, lds2);
}

template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4), amdgpu_num_vgpr(60))) void bp_kernel2w(BpArgs a) {
    extern __shared__ float2 lds2[];
    bp_block_body<NZ, CPT, TPR, ZS, kBf16>(a, lds2);
}

// One z-group of 32 slices (256 threads, 3 waves per SIMD).
template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bp_kernel1(BpArgs a) {
    extern __shared__ float2 lds2[];
    bp_block_body<NZ, CPT, TPR, ZS, kBf16>(a, lds2);
}

template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
constexpr auto bp_kernel_of() {
    if constexpr (ZS == 1) return &bp_kernel1<NZ, CPT, TPR, ZS, kBf16>;
    else if constexpr (CPT <= 8) return &bp_kernel2<NZ, CPT, TPR, ZS, kBf16>;
    else return &bp_kernel2w<NZ, CPT, TPR, ZS, kBf16>;
}

// Persistent form: a grid of at most the resident workgroups works through
// the (block, tile) items of many blocks (jobs).  Eight queues, one per XCD
// (workgroup b starts on queue b % 8, the XCD the hardware dispatches it
// to): queue q holds band q of every job, the contiguous tiles
// [floor(q n / 8), floor((q + 1) n / 8)) of its n tiles, in job order; a
// launch covers the items of jobs [job_first, job_end).  A workgroup takes
// items one at a time from its queue's counter, so the tiles an XCD runs at
// one time lie side by side and share its L2, as in one launch per block;
// when its queue is empty it takes what is left of the other queues in turn,
// and returns when all eight are empty.  Handing items out this way keeps
// the workgroups level: the preparation kernels that run beside a long
// launch slow the workgroups of some CUs more than others, and with fixed
// shares the slowest ran alone for the last 10-15% of a launch over 10-20
// blocks.  (The counter's answer is passed through dynamic LDS: a static LDS
// variable would move the dynamic base off the 16-byte alignment its 128-bit
// accesses need, which cost 12% here.)
template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__device__ __forceinline__ void bp_persistent_tile(const BpQueues &Q, const int *pre, int q, int item,
                                                   float2 *lds2) {
    // the job holding this item: the last j with pre[j] <= item
    int lo = Q.job_first, hi = Q.job_end - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (__builtin_amdgcn_readfirstlane(pre[mid]) <= item) lo = mid;
        else hi = mid - 1;
    }
    // the job through the scalar cache: the plan kernels wrote the jobs
    // before this launch and nothing writes them during it
    const __attribute__((address_space(4))) BpArgs &a =
        *(const __attribute__((address_space(4))) BpArgs *)(Q.jobs + lo);
    const int tile = (q * a.n_tiles) / 8 + (item - __builtin_amdgcn_readfirstlane(pre[lo]));
    bp_tile<NZ, CPT, TPR, ZS, kBf16>(a, tile, 0, lds2);
}

template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__device__ __forceinline__ void bp_persistent_body(const BpQueues &Q, float2 *lds2) {
    int *slot = reinterpret_cast<int *>(reinterpret_cast<char *>(lds2) + Q.slot_offset);   // double-buffered
    const int q0 = blockIdx.x & 7;
    int par = 0;
    for (int k = 0; k < 8; k++) {
        const int q = (q0 + k) & 7;
        const int *pre = Q.band_pre + q * (Q.n_jobs + 1);
        const int first = __builtin_amdgcn_readfirstlane(pre[Q.job_first]);
        const int len = __builtin_amdgcn_readfirstlane(pre[Q.job_end]);
        for (;;) {
            if (threadIdx.x == 0) slot[par] = first + atomicAdd(Q.counters + q, 1);
            __syncthreads();                    // slot written; every thread has left the previous tile
            const int item = __builtin_amdgcn_readfirstlane(slot[par]);
            par ^= 1;
            if (item >= len) break;
            bp_persistent_tile<NZ, CPT, TPR, ZS, kBf16>(Q, pre, q, item, lds2);
        }
    }
}

// Two z-groups (512 threads), the VGPR limits of bp_kernel2 / bp_kernel2w
// (the next group's low-register preparation kernels fit beside it).
template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) __attribute__((amdgpu_num_vgpr(56)))
void bp_persistent2_kernel(BpQueues Q) {
    extern __shared__ float2 lds2[];
    bp_persistent_body<NZ, CPT, TPR, ZS, kBf16>(Q, lds2);
}

template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(4))) __attribute__((amdgpu_num_vgpr(60)))
void bp_persistent2w_kernel(BpQueues Q) {
    extern __shared__ float2 lds2[];
    bp_persistent_body<NZ, CPT, TPR, ZS, kBf16>(Q, lds2);
}


template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void bp_persistent1_kernel(BpQueues Q) {
    extern __shared__ float2 lds2[];
    bp_persistent_body<NZ, CPT, TPR, ZS, kBf16>(Q, lds2);
}

template <int NZ, int CPT, int TPR, int ZS, bool kBf16>
constexpr auto bp_persistent_of() {
    if constexpr (ZS == 1) return &bp_persistent1_kernel<NZ, CPT, TPR, ZS, kBf16>;
    else if constexpr (CPT <= 8) return &bp_persistent2_kernel<NZ, CPT, TPR, ZS, kBf16>;
    else return &bp_persistent2w_kernel<NZ, CPT, TPR, ZS, kBf16>;
}

// View table of one block, built on the device (the backprojection's work
// list): entry t = i * n_ht + k (angle i < half, half turn k < n_ht) is
// theta = tube_start + (2 pi' / n_proj_turn) i + k pi' (reference
// include/backproject.cuh:75), its sine and cosine, the z term
// z_rot / 2 - z_rot (theta - tube_start) / (2 pi') and the rebinned
// projection i + k * half, in FP64 rounded to FP32 once.  Explicitly
// rounded operations keep the FP64 expression free of contractions.
__global__ __launch_bounds__(256) void view_table_kernel(BpView *tab, double tube_start, int half, int n_ht,
                                                          int n_proj_turn, double z_rot) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= half * n_ht) return;
    view_table_entry(tab + e, e, tube_start, half, n_ht, n_proj_turn, z_rot);
}

template <int NZ, int CPT, int TPR, int ZS>
size_t lds_bytes_of(const BpArgs &a) {
    const int rows_pad = (256 * ZS) >> a.tpr_log2;
    return 2 * static_cast<size_t>(rows_pad + 1) * a.tpr * (CPT + kPad<CPT>) * sizeof(float2) +
           128 * sizeof(float4) + (ZS == 2 ? 2 * 256 * sizeof(float4) : 0) + 128 * sizeof(int);
}

// Windows above 64 KB (tall detectors): raise the kernel's LDS limit, up to
// the device's per-workgroup maximum.
int allow_lds(const void *kf, size_t lds_bytes) {
    if (lds_bytes <= 65536) return CTBB_OK;
    int dev = 0, max_lds = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess)
        return CTBB_ERR_HIP;
    if (lds_bytes > static_cast<size_t>(max_lds)) return CTBB_ERR_SHAPE;
    if (hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds_bytes)) != hipSuccess)
        return CTBB_ERR_SHAPE;
    return CTBB_OK;
}

template <int NZ, int CPT, int TPR, int ZS>
int launch_c(const BpArgs &a, hipStream_t stream, bool beside) {
    const size_t lds_bytes = lds_bytes_of<NZ, CPT, TPR, ZS>(a);
    auto kf = a.rebin_bf16 ? reinterpret_cast<const void *>(bp_kernel_of<NZ, CPT, TPR, ZS, true>())
                           : reinterpret_cast<const void *>(bp_kernel_of<NZ, CPT, TPR, ZS, false>());
    const int st = allow_lds(kf, lds_bytes);
    if (st) return st;
    // Beside other kernels: the resident workgroups only (per slice range),
    // a multiple of 8 where the XCD remap applies, each striding over tiles,
    // so nothing of the launch waits for registers after its start.  Alone:
    // one workgroup per tile, the hardware starting each as one finishes
    // (C1 block: 5.82 against 5.87 ms for the resident grid, medians of 15
    // launches).
    const int ny = 32 / (NZ * ZS);
    long gx = a.n_tiles;
    if (beside) {
        int dev = 0, n_cu = 0, per_cu = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, 256 * ZS, lds_bytes) != hipSuccess)
            return CTBB_ERR_HIP;
        gx = std::max(1, per_cu) * static_cast<long>(std::max(1, n_cu)) / ny;
        if ((a.n_tiles & 7) == 0) gx = std::max<long>(8, gx / 8 * 8);
        gx = std::min<long>(gx, a.n_tiles);
    }
    dim3 grid(static_cast<unsigned>(gx), ny);
    void *args[] = {const_cast<BpArgs *>(&a)};
    return hipLaunchKernel(kf, grid, dim3(256 * ZS), args, lds_bytes, stream) == hipSuccess ? CTBB_OK
                                                                                           : CTBB_ERR_HIP;
}

template <int NZ, int CPT, int TPR, int ZS>
int launch_p(const BpArgs &a, const BpQueues &Q, int total_items, int max_groups, hipStream_t stream) {
    const size_t lds_bytes = lds_bytes_of<NZ, CPT, TPR, ZS>(a);
    const void *kf = a.rebin_bf16 ? reinterpret_cast<const void *>(bp_persistent_of<NZ, CPT, TPR, ZS, true>())
                                  : reinterpret_cast<const void *>(bp_persistent_of<NZ, CPT, TPR, ZS, false>());
    BpQueues QL = Q;
    QL.slot_offset = static_cast<int>((lds_bytes + 15) & ~static_cast<size_t>(15));
    const size_t lds_total = QL.slot_offset + 16;               // + the counter slots
    int st = allow_lds(kf, lds_total);
    if (st) return st;
    int dev = 0, n_cu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, 256 * ZS, lds_total) != hipSuccess)
        return CTBB_ERR_HIP;
    long groups = static_cast<long>(std::max(per_cu, 1)) * std::max(n_cu, 1);
    if (max_groups > 0) groups = std::min<long>(groups, max_groups);
    groups = std::min<long>(groups, std::max(total_items, 1));
    groups = std::max<long>(8, (groups + 7) / 8 * 8);     // every queue has workgroups starting on it
    void *args[] = {&QL};
    return hipLaunchKernel(kf, dim3(static_cast<unsigned>(groups)), dim3(256 * ZS), args, lds_total, stream) ==
                   hipSuccess
               ? CTBB_OK
               : CTBB_ERR_HIP;
}

// Kernel variant of a block: slices per thread (NZ, ZS), channels per
// staging thread (CPT: the smallest supported value whose window
// W = tpr * CPT covers the tile footprint, 2 half_w + 3 channels) and the
// staging threads per row (TPR: compile-time for the (TPR, CPT) shapes of
// 16-, 32- and 64-row detectors at the usual fields of view, 0 = taken from
// the arguments).  Encoded as one int; negative: a ctbb_status.
constexpr int kCptList[] = {1, 2, 3, 4, 5, 6, 8, 10, 12, 14};

bool has_static_tpr(int zs, int tpr, int cpt) {
    const int key = tpr * 100 + cpt;
    if (zs == 1) {
        switch (key) {
            case 406: case 408: case 410: case 412: case 414: case 804: case 806: case 808: case 1602: case 1604:
                return true;
        }
        return false;
    }
    switch (key) {
        case 804: case 805: case 806: case 1602: case 1603: case 3201: case 3202: return true;
    }
    return false;
}

int encode_variant(int nz, int cpt, int tpr_ct, int tpr, int bf16) {
    return ((((bf16 * 64 + nz) * 32 + cpt) * 128 + tpr_ct) * 512 + tpr);
}

template <class Op>
int with_static_tpr(int zs, int tpr, int cpt, Op op) {
    if (zs == 1) {
        switch (tpr * 100 + cpt) {
            case 406: return op(std::integral_constant<int, 6>{}, std::integral_constant<int, 4>{});
            case 408: return op(std::integral_constant<int, 8>{}, std::integral_constant<int, 4>{});
            case 410: return op(std::integral_constant<int, 10>{}, std::integral_constant<int, 4>{});
            case 412: return op(std::integral_constant<int, 12>{}, std::integral_constant<int, 4>{});
            case 414: return op(std::integral_constant<int, 14>{}, std::integral_constant<int, 4>{});
            case 804: return op(std::integral_constant<int, 4>{}, std::integral_constant<int, 8>{});
            case 806: return op(std::integral_constant<int, 6>{}, std::integral_constant<int, 8>{});
            case 808: return op(std::integral_constant<int, 8>{}, std::integral_constant<int, 8>{});
            case 1602: return op(std::integral_constant<int, 2>{}, std::integral_constant<int, 16>{});
            case 1604: return op(std::integral_constant<int, 4>{}, std::integral_constant<int, 16>{});
        }
        return CTBB_ERR_SHAPE;
    }
    switch (tpr * 100 + cpt) {
        case 804: return op(std::integral_constant<int, 4>{}, std::integral_constant<int, 8>{});
        case 805: return op(std::integral_constant<int, 5>{}, std::integral_constant<int, 8>{});
        case 806: return op(std::integral_constant<int, 6>{}, std::integral_constant<int, 8>{});
        case 1602: return op(std::integral_constant<int, 2>{}, std::integral_constant<int, 16>{});
        case 1603: return op(std::integral_constant<int, 3>{}, std::integral_constant<int, 16>{});
        case 3201: return op(std::integral_constant<int, 1>{}, std::integral_constant<int, 32>{});
        case 3202: return op(std::integral_constant<int, 2>{}, std::integral_constant<int, 32>{});
    }
    return CTBB_ERR_SHAPE;
}

template <class Op>
int with_generic(int cpt, Op op) {
    switch (cpt) {
        case 1: return op(std::integral_constant<int, 1>{}, std::integral_constant<int, 0>{});
        case 2: return op(std::integral_constant<int, 2>{}, std::integral_constant<int, 0>{});
        case 3: return op(std::integral_constant<int, 3>{}, std::integral_constant<int, 0>{});
        case 4: return op(std::integral_constant<int, 4>{}, std::integral_constant<int, 0>{});
        case 5: return op(std::integral_constant<int, 5>{}, std::integral_constant<int, 0>{});
        case 6: return op(std::integral_constant<int, 6>{}, std::integral_constant<int, 0>{});
        case 8: return op(std::integral_constant<int, 8>{}, std::integral_constant<int, 0>{});
        case 10: return op(std::integral_constant<int, 10>{}, std::integral_constant<int, 0>{});
        case 12: return op(std::integral_constant<int, 12>{}, std::integral_constant<int, 0>{});
        case 14: return op(std::integral_constant<int, 14>{}, std::integral_constant<int, 0>{});
    }
    return CTBB_ERR_SHAPE;
}

// op(nz, zs, cpt, tpr) with compile-time constants for the variant id
template <class Op>
int with_variant(int id, Op op) {
    const int tpr_ct = (id / 512) % 128, cpt = (id / (512 * 128)) % 32, nz = (id / (512 * 128 * 32)) % 64;
    auto inner = [&](auto nzc, auto zsc) {
        auto call = [&](auto cptc, auto tprc) { return op(nzc, zsc, cptc, tprc); };
        return tpr_ct ? with_static_tpr(decltype(zsc)::value, tpr_ct, cpt, call) : with_generic(cpt, call);
    };
    if (nz == 32) return inner(std::integral_constant<int, 32>{}, std::integral_constant<int, 1>{});
    if (nz == 16) return inner(std::integral_constant<int, 16>{}, std::integral_constant<int, 2>{});
    return CTBB_ERR_ARG;
}

}  // namespace

int backproject_variant(BpArgs &a) {
    if (a.nx % 16 || a.ny % 16 || a.nx <= 0 || a.ny <= 0) return -CTBB_ERR_SHAPE;
    if (a.n_rows < 2 || a.half <= 0 || a.n_ht <= 0) return -CTBB_ERR_SHAPE;
    if (a.n_tiles != (a.nx / 16) * (a.ny / 16) || a.tiles_y != a.ny / 16) return -CTBB_ERR_ARG;
    if (a.nz != 16 && a.nz != 32) return -CTBB_ERR_ARG;
    const int zs = a.nz == 16 ? 2 : 1;
    const int threads = 256 * zs;
    int rows_pad = 2;
    while (rows_pad < a.n_rows) rows_pad <<= 1;
    if (rows_pad > threads) return -CTBB_ERR_SHAPE;
    a.tpr = threads / rows_pad;
    a.tpr_log2 = 0;
    while ((1 << a.tpr_log2) < a.tpr) a.tpr_log2++;
    const int need = static_cast<int>(ceilf((2.0f * a.half_w + 3.0f) / static_cast<float>(a.tpr)));
    for (int c : kCptList) {
        if (c < need) continue;
        const int tpr_ct = has_static_tpr(zs, a.tpr, c) ? a.tpr : 0;
        return encode_variant(a.nz, c, tpr_ct, a.tpr, a.rebin_bf16 ? 1 : 0);
    }
    return -CTBB_ERR_SHAPE;   // footprint wider than 14 channels per thread
}

int launch_backproject(const BpArgs &a0, hipStream_t stream, bool beside) {
    if ((!a0.rebin && !a0.rebin_bf16) || !a0.vtab || !a0.out) return CTBB_ERR_ARG;
    BpArgs a = a0;
    const int id = backproject_variant(a);
    if (id < 0) return -id;
    return with_variant(id, [&](auto nz, auto zs, auto cpt, auto tpr) {
        return launch_c<decltype(nz)::value, decltype(cpt)::value, decltype(tpr)::value, decltype(zs)::value>(a, stream,
                                                                                                           beside);
    });
}

template <int NZ, int CPT, int TPR, int ZS>
int resident_c(const BpArgs &a, long *groups) {
    const size_t lds_bytes = lds_bytes_of<NZ, CPT, TPR, ZS>(a);
    auto kf = a.rebin_bf16 ? reinterpret_cast<const void *>(bp_kernel_of<NZ, CPT, TPR, ZS, true>())
                           : reinterpret_cast<const void *>(bp_kernel_of<NZ, CPT, TPR, ZS, false>());
    const int st = allow_lds(kf, lds_bytes);
    if (st) return st;
    int dev = 0, n_cu = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kf, 256 * ZS, lds_bytes) != hipSuccess)
        return CTBB_ERR_HIP;
    *groups = std::max(1, per_cu) * static_cast<long>(std::max(1, n_cu)) / (32 / (NZ * ZS));
    return CTBB_OK;
}

int backproject_resident(const BpArgs &a0, long *groups) {
    if (!groups) return CTBB_ERR_ARG;
    BpArgs a = a0;
    const int id = backproject_variant(a);
    if (id < 0) return -id;
    return with_variant(id, [&](auto nz, auto zs, auto cpt, auto tpr) {
        return resident_c<decltype(nz)::value, decltype(cpt)::value, decltype(tpr)::value, decltype(zs)::value>(
            a, groups);
    });
}

int launch_backproject_persistent(const BpArgs &sample, int variant, const BpQueues &Q, int total_items,
                                  int max_groups, hipStream_t stream) {
    if (!Q.jobs || !Q.band_pre || !Q.counters || Q.n_jobs <= 0 || Q.job_first < 0 || Q.job_end <= Q.job_first ||
        Q.job_end > Q.n_jobs || variant < 0)
        return CTBB_ERR_ARG;
    if (total_items <= 0) return CTBB_OK;
    return with_variant(variant, [&](auto nz, auto zs, auto cpt, auto tpr) {
        return launch_p<decltype(nz)::value, decltype(cpt)::value, decltype(tpr)::value, decltype(zs)::value>(
            sample, Q, total_items, max_groups, stream);
    });
}

int launch_view_table(BpView *tab, double tube_start, int half, int n_ht, int n_proj_turn, double z_rot,
                      hipStream_t stream) {
    if (!tab || half <= 0 || n_ht <= 0 || n_proj_turn <= 0) return CTBB_ERR_ARG;
    const long n = static_cast<long>(half) * n_ht;
    if (n > 0x7fffffffL - 255) return CTBB_ERR_SHAPE;
    hipLaunchKernelGGL(view_table_kernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, tab,
                       tube_start, half, n_ht, n_proj_turn, z_rot);
    return hipGetLastError() == hipSuccess ? CTBB_OK : CTBB_ERR_HIP;
}

}  // namespace ctbb
