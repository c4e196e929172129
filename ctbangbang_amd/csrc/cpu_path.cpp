// CPU reconstruction path (the reference driver's --no-gpu mode) on host
// threads: rebin + ramp filter, weighted backprojection with the reference
// CPU twin's conventions, and slice thickening.
//
//   ctbb_rebin_filter_cpu          reference src/rebin_filter_cpu.cu:75 (+ filter_cpu :647-709)
//   ctbb_backproject_cpu           reference src/backproject_cpu.cu:49-154
//   ctbb_finalize_image_stack_cpu  reference src/finalize_image_stack_cpu.cu:24-117
//
// Rebinning uses the same FP64-derived coordinate maps and clamped bilinear
// taps as the HIP kernels (rebin.hip), the filter is an FP64 FFT convolution
// (the reference CPU twin uses FFTW in double), the backprojection follows
// backproject_cpu's arithmetic statement by statement (float and double
// exactly where the reference has them), with the per-(voxel column, view)
// terms evaluated once instead of per slice (the same values), over threads
// that each own whole voxel columns.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <thread>
#include <vector>

#include "ctbb_hip.h"
#include "ctbb_internal.h"
#include "geometry.h"

namespace {

int thread_count(int requested) {
    if (requested > 0) return requested;
    const unsigned hc = std::thread::hardware_concurrency();
    return hc ? static_cast<int>(hc) : 1;
}

// f(begin, end) over [0, n) in contiguous ranges, one per thread.
template <class F>
void parallel_ranges(long n, int threads, F f) {
    threads = static_cast<int>(std::min<long>(std::max(threads, 1), std::max<long>(n, 1)));
    if (threads == 1) {
        f(0L, n);
        return;
    }
    std::vector<std::thread> pool;
    const long per = (n + threads - 1) / threads;
    for (int t = 0; t < threads; t++) {
        const long b = t * per, e = std::min(n, b + per);
        if (b >= e) break;
        pool.emplace_back([=, &f] { f(b, e); });
    }
    for (auto &th : pool) th.join();
}

// ---------------------------------------------------------------------------
// rebin: the HIP kernels' arithmetic (clamped bilinear taps of the raw pull
// at FP64-derived coordinates), written per output sample

void clamp_taps(float x, int n, int &i0, int &i1, float &f) {
    x = std::fmin(std::fmax(x, 0.0f), static_cast<float>(n - 1));
    const float fl = std::floor(x);
    i0 = static_cast<int>(fl);
    i1 = std::min(i0 + 1, n - 1);
    f = x - fl;
}

int rebin_block(const recon_metadata &mr, const float *raw, float *out, int threads) {
    const ct_geom &cg = mr.cg;
    std::vector<float> bidx, aoff;
    int kind = 0, sets = 0;
    int st = ctbb::rebin_maps(mr, bidx, aoff, &kind, &sets);
    if (st) return st;
    const int nch = static_cast<int>(cg.n_channels), nos = static_cast<int>(cg.n_channels_oversampled);
    const int nrr = static_cast<int>(cg.n_rows_raw), nrows = static_cast<int>(cg.n_rows);
    const int n_ffs = static_cast<int>(mr.ri.n_ffs);
    const int n_proj = static_cast<int>(mr.ri.n_proj_pull / mr.ri.n_ffs);
    const int add = static_cast<int>(cg.add_projections);
    const int n_out = n_proj - 2 * add;
    if (n_out <= 0) return CTBB_ERR_SHAPE;
    const size_t frame = static_cast<size_t>(nrr) * nch;
    parallel_ranges(static_cast<long>(nrows) * nos, threads, [&](long b, long e) {
        for (long rc = b; rc < e; rc++) {
            const int r = static_cast<int>(rc / nos), c = static_cast<int>(rc % nos);
            int set = 0, j = r;
            if (kind == 1 || kind == 3) {
                set = (r & 1) ^ (cg.reverse_row_interleave ? 1 : 0);
                j = r >> 1;
            }
            const float by = bidx[static_cast<size_t>(set) * nos + c];
            int c0, c1;
            float fc, ao0, ao1 = 0.0f;
            if (kind <= 1) {
                clamp_taps(by, nch, c0, c1, fc);
                ao0 = aoff[static_cast<size_t>(set) * nos + c];
            } else {
                clamp_taps(by, nos, c0, c1, fc);
                ao0 = aoff[static_cast<size_t>(set) * nos + c0];
                ao1 = aoff[static_cast<size_t>(set) * nos + c1];
            }
            const float *row = raw + static_cast<size_t>(j) * nch;
            for (int v = 0; v < n_out; v++) {
                const float proj = static_cast<float>(v + add);
                float o;
                if (kind <= 1) {
                    const int fs = kind == 1 ? set : 0;
                    int p0, p1;
                    float fa;
                    clamp_taps(proj - ao0, n_proj, p0, p1, fa);
                    const float *s0 = row + (static_cast<size_t>(p0) * n_ffs + fs) * frame;
                    const float *s1 = row + (static_cast<size_t>(p1) * n_ffs + fs) * frame;
                    const float t0 = std::fma(fc, s0[c1] - s0[c0], s0[c0]);
                    const float t1 = std::fma(fc, s1[c1] - s1[c0], s1[c0]);
                    o = std::fma(fa, t1 - t0, t0);
                } else {
                    float t[2];
                    for (int h = 0; h < 2; h++) {
                        const int m = h ? c1 : c0;
                        const int fs = 2 * set + (m & 1);
                        int p0, p1;
                        float fa;
                        clamp_taps(proj - (h ? ao1 : ao0), n_proj, p0, p1, fa);
                        const float u0 = row[(static_cast<size_t>(p0) * n_ffs + fs) * frame + (m >> 1)];
                        const float u1 = row[(static_cast<size_t>(p1) * n_ffs + fs) * frame + (m >> 1)];
                        t[h] = std::fma(fa, u1 - u0, u0);
                    }
                    o = std::fma(fc, t[1] - t[0], t[0]);
                }
                out[(static_cast<size_t>(v) * nrows + r) * nos + c] = o;
            }
        }
    });
    return CTBB_OK;
}

// ---------------------------------------------------------------------------
// ramp filter: FP64 FFT convolution, two rows per complex transform

struct Fft {
    int L = 0;
    std::vector<std::complex<double>> w;    // e^(-2 pi i t / L), t < L / 2
    std::vector<int> rev;
    explicit Fft(int len) : L(len), w(len / 2), rev(len) {
        for (int t = 0; t < L / 2; t++) w[t] = std::polar(1.0, -2.0 * M_PI * t / L);
        int bits = 0;
        while ((1 << bits) < L) bits++;
        for (int i = 0; i < L; i++) {
            int r = 0;
            for (int b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
            rev[i] = r;
        }
    }
    // in place; inverse: conjugate twiddles, no scaling
    void run(std::complex<double> *x, bool inverse) const {
        for (int i = 0; i < L; i++)
            if (i < rev[i]) std::swap(x[i], x[rev[i]]);
        for (int len = 2; len <= L; len <<= 1) {
            const int step = L / len;
            for (int i = 0; i < L; i += len)
                for (int k = 0; k < len / 2; k++) {
                    const double wr = w[k * step].real(), wi = inverse ? -w[k * step].imag() : w[k * step].imag();
                    const std::complex<double> u = x[i + k], b = x[i + k + len / 2];
                    const std::complex<double> v(b.real() * wr - b.imag() * wi, b.real() * wi + b.imag() * wr);
                    x[i + k] = u + v;
                    x[i + k + len / 2] = u - v;
                }
        }
    }
};

// rows[r][l] <- sum_k taps[n + l - k] rows[r][k]: with h[m] = taps[m]
// (m < 2n) zero-padded to L >= 2n, the circular convolution of the padded
// row gives y[n + l] without wrap-around.
void filter_rows(float *rows, size_t n_rows, int n, const float *taps, int threads) {
    int L = 1;
    while (L < 2 * n) L <<= 1;
    const Fft fft(L);
    std::vector<std::complex<double>> H(L, 0.0);
    for (int m = 0; m < 2 * n; m++) H[m] = taps[m];
    fft.run(H.data(), false);
    for (auto &h : H) h /= static_cast<double>(L);
    const long pairs = static_cast<long>((n_rows + 1) / 2);
    parallel_ranges(pairs, threads, [&](long b, long e) {
        std::vector<std::complex<double>> x(L);
        for (long p = b; p < e; p++) {
            float *ra = rows + static_cast<size_t>(2 * p) * n;
            const bool has_b = static_cast<size_t>(2 * p + 1) < n_rows;
            float *rb = ra + n;
            std::fill(x.begin(), x.end(), std::complex<double>(0.0, 0.0));
            for (int k = 0; k < n; k++) x[k] = std::complex<double>(ra[k], has_b ? rb[k] : 0.0f);
            fft.run(x.data(), false);
            for (int i = 0; i < L; i++)
                x[i] = std::complex<double>(x[i].real() * H[i].real() - x[i].imag() * H[i].imag(),
                                            x[i].real() * H[i].imag() + x[i].imag() * H[i].real());
            fft.run(x.data(), true);
            for (int l = 0; l < n; l++) {
                ra[l] = static_cast<float>(x[n + l].real());
                if (has_b) rb[l] = static_cast<float>(x[n + l].imag());
            }
        }
    });
}

// ---------------------------------------------------------------------------
// backprojection, reference CPU twin conventions (no x flip, qhat sign +1,
// output [z][yi][xi], interp2's zero border)

constexpr float kPi = 3.1415368979f;     // reference src/backproject_cpu.cu:29

float weight_w(float q) {                  // reference src/backproject_cpu.cu:34-47
    float out;
    const float Q = 0.6f;
    if (std::fabs(q) < Q) out = 1.0f;
    else if ((std::fabs(q) >= Q) && (std::fabs(q) < 1.0f))
        out = std::pow(std::cos((kPi / 2.0f) * (std::fabs(q) - Q) / (1.0f - Q)), 2.0f);
    else out = 0.0f;
    return out;
}

// reference src/interp.cu:29-64 (border mode: zero outside [0, dim - 2])
float interp2_zero(const float *array, int dim1, int dim2, float idx1, float idx2) {
    if (idx1 > (dim1 - 2) || (idx1 < 0) || idx2 > (dim2 - 2) || (idx2 < 0)) return 0.0f;
    const float v = idx1 - std::floor(idx1);
    const float w = idx2 - std::floor(idx2);
    const int i1 = static_cast<int>(std::floor(idx1)), i2 = static_cast<int>(std::floor(idx2));
    const int a1 = i2 * dim1 + i1, a2 = i2 * dim1 + i1 + 1, a3 = (i2 + 1) * dim1 + i1, a4 = (i2 + 1) * dim1 + i1 + 1;
    return array[a1] * (1.0f - v) * (1.0f - w) + array[a2] * v * (1.0f - w) + array[a3] * (1.0f - v) * w +
           array[a4] * v * w;
}

int backproject_block(const recon_metadata *mr, const float *rebin, float *out, int threads) {
    const ct_geom cg = mr->cg;
    const recon_params rp = mr->rp;
    const float tube_start =
        std::fmod(static_cast<double>(mr->tube_angles[0] + ((mr->ri.idx_pull_start + cg.add_projections_ffs) *
                                                            360.0f / cg.n_proj_ffs)),
                  360.0) *
        kPi / 180.0f;
    const int n_half_turns = static_cast<int>((mr->ri.n_proj_pull / mr->ri.n_ffs - 2 * cg.add_projections) /
                                              (cg.n_proj_turn / 2));
    if (n_half_turns <= 0) return CTBB_ERR_SHAPE;
    const int nx = static_cast<int>(rp.nx), ny = static_cast<int>(rp.ny);
    const int nz = static_cast<int>(mr->ri.n_slices_block);
    const int nos = static_cast<int>(cg.n_channels_oversampled), nrows = static_cast<int>(cg.n_rows);
    const int dim1 = nos, dim2 = nrows * n_half_turns;
    const size_t view = static_cast<size_t>(nos) * nrows;
    const int half = static_cast<int>(cg.n_proj_turn / 2);
    const size_t nxy = static_cast<size_t>(nx) * ny;
    std::fill(out, out + nxy * nz, 0.0f);
    // per slice: z of the reference's voxel loop
    std::vector<float> zpos(nz);
    for (int zi = 0; zi < nz; zi++)
        zpos[zi] = zi * rp.coll_slicewidth + cg.z_rot / 2.0f + cg.z_rot * tube_start / (2.0f * kPi);
    parallel_ranges(nx, threads, [&](long xb, long xe) {
        std::vector<float> proj(view * n_half_turns);
        std::vector<float> st(static_cast<size_t>(nz) * ny), ht(static_cast<size_t>(nz) * ny);
        for (int i = 0; i < half; i++) {
            for (int kk = 0; kk < n_half_turns; kk++)
                std::memcpy(&proj[view * kk], rebin + (static_cast<size_t>(i) + static_cast<size_t>(kk) * half) * view,
                            view * sizeof(float));
            for (long xi = xb; xi < xe; xi++) {
                const float x = (rp.recon_fov / rp.nx) * (static_cast<float>(xi) - (rp.nx - 1) / 2.0f) + rp.x_origin;
                std::fill(st.begin(), st.end(), 0.0f);
                std::fill(ht.begin(), ht.end(), 0.0f);
                for (int yi = 0; yi < ny; yi++) {
                    const float y = (rp.recon_fov / rp.ny) * (static_cast<float>(yi) - (rp.ny - 1) / 2.0f) + rp.y_origin;
                    for (int k = 0; k < n_half_turns; k++) {
                        // (float overloads of the math functions, as the
                        // reference's C++ build resolves them)
                        const float theta = tube_start + (2.0f * kPi / cg.n_proj_turn) * i + k * kPi;
                        const float phat = x * std::sin(theta) - y * std::cos(theta);
                        const float p_idx = phat / (cg.r_f * cg.fan_angle_increment / 2.0f) + 2.0f * cg.central_channel;
                        // terms of the reference's z loop that do not depend on z
                        const float ray_pos = (cg.z_rot * (theta - std::asin(phat / cg.r_f)) / (2.0f * kPi));
                        const float lhat = std::sqrt(std::pow(cg.r_f, 2.0f) - std::pow(phat, 2.0f)) -
                                           x * std::cos(theta) - y * std::sin(theta);
                        const float tanc = std::tan(cg.theta_cone / 2.0f);
                        for (int zi = 0; zi < nz; zi++) {
                            const float qhat = (zpos[zi] - ray_pos) / (lhat * tanc);
                            const float q_idx = ((qhat + 1.0f) / 2.0f) * (cg.n_rows - 1.0f) + k * cg.n_rows;
                            const float wq = weight_w(qhat);
                            const size_t vi = static_cast<size_t>(zi) * ny + yi;
                            st[vi] += interp2_zero(proj.data(), dim1, dim2, p_idx, q_idx) * wq;
                            ht[vi] += wq;
                        }
                    }
                }
                for (int yi = 0; yi < ny; yi++)
                    for (int zi = 0; zi < nz; zi++) {
                        const size_t vi = static_cast<size_t>(zi) * ny + yi;
                        if (ht[vi] != 0)
                            out[zi * nxy + static_cast<size_t>(yi) * nx + xi] +=
                                (1.0f / ht[vi]) * st[vi] * 2 * kPi / cg.n_proj_turn;
                    }
            }
        }
    });
    return CTBB_OK;
}

}  // namespace

extern "C" {

int ctbb_rebin_filter_cpu(struct recon_metadata *mr, int n_threads) {
    if (!mr || !mr->ctd.raw || !mr->ctd.rebin || mr->ri.n_ffs == 0) return CTBB_ERR_ARG;
    if (mr->cg.n_channels_oversampled != 2 * mr->cg.n_channels) return CTBB_ERR_SHAPE;
    const int threads = thread_count(n_threads);
    int st = rebin_block(*mr, mr->ctd.raw, mr->ctd.rebin, threads);
    if (st) return st;
    const int n = static_cast<int>(mr->cg.n_channels_oversampled);
    std::vector<float> taps(2 * static_cast<size_t>(n));
    st = ctbb_load_filter(mr, taps.data());
    if (st) return st;
    const size_t n_out = mr->ri.n_proj_pull / mr->ri.n_ffs - 2 * mr->cg.add_projections;
    filter_rows(mr->ctd.rebin, n_out * mr->cg.n_rows, n, taps.data(), threads);
    return CTBB_OK;
}

int ctbb_backproject_cpu(struct recon_metadata *mr, int n_threads) {
    if (!mr || !mr->ctd.rebin || !mr->ctd.image || !mr->tube_angles || mr->ri.cb.block_idx < 1) return CTBB_ERR_ARG;
    const size_t nblk = mr->ri.n_slices_block * static_cast<size_t>(mr->rp.nx) * mr->rp.ny;
    return backproject_block(mr, mr->ctd.rebin, mr->ctd.image + static_cast<size_t>(mr->ri.cb.block_idx - 1) * nblk,
                             thread_count(n_threads));
}

int ctbb_finalize_image_stack_cpu(struct recon_metadata *mr, float mu_water, int n_threads) {
    if (!mr || !mr->ctd.image || !mr->table_positions || mr->rp.n_readings <= 1000) return CTBB_ERR_ARG;
    const recon_params &rp = mr->rp;
    const recon_info &ri = mr->ri;
    const size_t nxy = static_cast<size_t>(rp.nx) * rp.ny;
    const float dpos = rp.end_pos - rp.start_pos;
    const int recon_direction = dpos < 0.0f ? -1 : 1;      // 1 for a single slice (end == start)
    const double tdd = mr->table_positions[1000] - mr->table_positions[0];
    const int table_direction = tdd == 0.0 ? 1 : static_cast<int>(tdd / std::fabs(tdd));
    const int nb = static_cast<int>(ri.n_slices_block);
    const int n_raw = nb * static_cast<int>(ri.n_blocks);
    const int nf = ctbb_n_slices_final(mr);
    // block reorder (reference :43-61) folded into the source slice index
    auto src_slice = [&](int s) {
        if (recon_direction == table_direction) return s;
        const int b = s / nb, z = s % nb;
        return b * nb + (nb - 1 - z);
    };
    std::vector<float> raw_loc(n_raw), w(n_raw);
    for (int i = 0; i < n_raw; i++) raw_loc[i] = ri.recon_start_pos + recon_direction * i * rp.coll_slicewidth;
    float *fin = static_cast<float *>(std::calloc(static_cast<size_t>(nf) * nxy, sizeof(float)));
    if (!fin) return CTBB_ERR_NOMEM;
    for (int k = 0; k < nf; k++) {
        const float loc = rp.start_pos + recon_direction * k * rp.slice_thickness;
        float sum = 0;
        for (int s = 0; s < n_raw; s++) {
            w[s] = std::fmax(0.0f, 1.0f - std::fabs(raw_loc[s] - loc) / rp.slice_thickness);
            sum += w[s];
        }
        float *o = fin + static_cast<size_t>(k) * nxy;
        parallel_ranges(static_cast<long>(nxy), thread_count(n_threads), [&](long b, long e) {
            for (int s = 0; s < n_raw; s++) {
                if (w[s] == 0) continue;
                const float ws = w[s] / sum;
                const float *img = mr->ctd.image + static_cast<size_t>(src_slice(s)) * nxy;
                for (long p = b; p < e; p++) o[p] += ws * img[p];
            }
            if (mu_water > 0.0f)
                for (long p = b; p < e; p++) o[p] = 1000.0f * (o[p] - mu_water) / mu_water;
        });
    }
    std::free(mr->ctd.final_image_stack);
    mr->ctd.final_image_stack = fin;
    return CTBB_OK;
}

}  // extern "C"
