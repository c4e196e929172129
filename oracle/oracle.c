/*
 * oracle.c — CPU restatement of the reconstruction hot path (TEST
 * INFRASTRUCTURE ONLY).
 *
 * This file is the checker the parity tests compare the HIP path against.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product library never links or calls it.
 *
 * It is compiled twice (see oracle/Makefile):
 *   liboracle64.so  REAL=double  the FP64 oracle
 *   liboracle32.so  REAL=float   the FP32 restatement of the reference's CPU
 *                                path (used as the timed CPU baseline)
 *
 * Every function follows the reference's discrete algorithm (same index
 * quantisation, same pi constants, same clamping/border rules) and cites the
 * reference lines it restates.  Where the reference GPU and CPU twins
 * disagree, the GPU conventions are the default and the CPU conventions are
 * selectable (ORC_BP_CPU_CONVENTIONS) so this restatement can be pinned
 * against the reference's own CPU code compiled from source (oracle/_ref).
 *
 * Pinning status of each function (checks in tests/test_oracle.py):
 *   orc_backproject      pinned: CPU-convention mode vs reference
 *                        backproject_cpu built from source; constant-sinogram
 *                        closed form; linearity.
 *   orc_finalize         pinned: vs reference finalize_image_stack_cpu built
 *                        from source; constant / linear-in-z closed forms.
 *   orc_filter_direct    pinned: impulse -> taps; numpy.convolve; FFT path.
 *   orc_rebin            pinned: closed-form coordinate maps (constant
 *                        sinogram -> constant; analytic disc -> analytic
 *                        parallel projection); FP32 vs FP64.
 *   orc_adaptive_filter  pinned by invariants only (samples <= T unchanged,
 *                        constant input unchanged, normalised weights,
 *                        brute-force threshold property); the reference
 *                        stage is GPU-only and cannot be built here.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "recon_structs.h"

#ifndef REAL
#define REAL double
#endif

#define ORC_IS_DOUBLE (sizeof(REAL) == sizeof(double))
#define RSIN(x) (ORC_IS_DOUBLE ? sin((double)(x)) : (double)sinf((float)(x)))
#define RCOS(x) (ORC_IS_DOUBLE ? cos((double)(x)) : (double)cosf((float)(x)))
#define RASIN(x) (ORC_IS_DOUBLE ? asin((double)(x)) : (double)asinf((float)(x)))
#define RSQRT(x) (ORC_IS_DOUBLE ? sqrt((double)(x)) : (double)sqrtf((float)(x)))
#define RTAN(x) (ORC_IS_DOUBLE ? tan((double)(x)) : (double)tanf((float)(x)))
#define RFLOOR(x) (ORC_IS_DOUBLE ? floor((double)(x)) : (double)floorf((float)(x)))

/* pi constants as float literals, as the reference uses them:
 * backprojection and preprocessing use 3.1415368979f (backproject.cuh:23,
 * preprocessing.cuh:25); rebinning and setup use 3.1415926535897f
 * (rebin_filter.cuh:23). */
static const float PI_BP = 3.1415368979f;
static const float PI_TRUE = 3.1415926535897f;

#define ORC_BP_GPU_CONVENTIONS 0
#define ORC_BP_CPU_CONVENTIONS 1
/* Planted mistakes for the ground-truth tests (bits added to the GPU
 * conventions): drop the x flip, drop the -table_direction sign of qhat.
 * A test that plants one must see its checks fail. */
#define ORC_BP_PLANT_NO_FLIP 2
#define ORC_BP_PLANT_NO_QSIGN 4

/* ------------------------------------------------------------------------ */
/* Interpolation helpers                                                    */
/* ------------------------------------------------------------------------ */

/* Clamped bilinear sample of a 2D array (x fastest, width w, height h) at
 * fractional index (x, y): the texture "clamp" mode of reference
 * rebin_filter.cu:91-101 and the CPU interp3 clamping of interp.cu:71-119
 * (both clamp the coordinate into [0, dim-1] before interpolating). */
static REAL bilerp_clamp(const float *a, long w, long h, REAL x, REAL y) {
    if (x > (REAL)(w - 1)) x = (REAL)(w - 1);
    if (x < 0) x = 0;
    if (y > (REAL)(h - 1)) y = (REAL)(h - 1);
    if (y < 0) y = 0;
    long x0 = (long)RFLOOR(x), y0 = (long)RFLOOR(y);
    REAL fx = x - (REAL)x0, fy = y - (REAL)y0;
    long x1 = x0 + 1 < w ? x0 + 1 : w - 1;
    long y1 = y0 + 1 < h ? y0 + 1 : h - 1;
    REAL v00 = a[y0 * w + x0], v01 = a[y0 * w + x1];
    REAL v10 = a[y1 * w + x0], v11 = a[y1 * w + x1];
    return v00 * (1 - fx) * (1 - fy) + v01 * fx * (1 - fy) + v10 * (1 - fx) * fy + v11 * fx * fy;
}

/* ------------------------------------------------------------------------ */
/* Rebinning geometry helpers: reference rebin_filter.cuh:36-64 and
 * rebin_filter_cpu.cu:37-65.                                               */
/* ------------------------------------------------------------------------ */

static REAL orc_angle(REAL x1, REAL x2, REAL y1, REAL y2) {
    return RASIN((x1 * y2 - x2 * y1) / (RSQRT(x1 * x1 + x2 * x2) * RSQRT(y1 * y1 + y2 * y2)));
}

static REAL orc_beta_rk(const struct ct_geom *cg, REAL da, REAL dr, REAL channel) {
    REAL b0 = (channel - (REAL)cg->central_channel) * (REAL)cg->fan_angle_increment;
    return orc_angle(-((REAL)cg->r_f + dr), -da, -((REAL)cg->src_to_det * RCOS(b0) + dr),
                     -((REAL)cg->src_to_det * RSIN(b0) + da));
}

static REAL orc_d_alpha_r(const struct ct_geom *cg, REAL da, REAL dr) {
    return orc_angle((REAL)cg->r_f, 0, (REAL)cg->r_f + dr, da);
}

static REAL orc_r_fr(const struct ct_geom *cg, REAL da, REAL dr) {
    return RSQRT(((REAL)cg->r_f + dr) * ((REAL)cg->r_f + dr) + da * da);
}

/* Linear search in an increasing lookup table with linear extrapolation past
 * its ends: reference rebin_filter.cuh:53-64 / rebin_filter_cpu.cu:54-65. */
static REAL orc_get_beta_idx(REAL beta, const REAL *lut, int n) {
    int lo = 0;
    while (beta > lut[lo] && lo < n - 1) lo++;
    if (lo == 0) lo++;
    return (REAL)lo - 1 + (beta - lut[lo - 1]) / (lut[lo] - lut[lo - 1]);
}

/* FFS focal-spot deflections: reference rebin_filter.cu:194,397-398,593-594. */
static void orc_ffs_offsets(const struct recon_metadata *mr, double *da, double *dr) {
    const struct ct_geom *cg = &mr->cg;
    *da = cg->src_to_det * cg->r_f * cg->fan_angle_increment / (4.0 * (cg->src_to_det - cg->r_f));
    *dr = cg->src_to_det * mr->rp.coll_slicewidth / (4.0 * (cg->src_to_det - cg->r_f) * tan(cg->anode_angle));
}

/* ------------------------------------------------------------------------ */
/* Rebinning (all four FFS modes)                                           */
/* ------------------------------------------------------------------------ */

/*
 * orc_rebin: fan-to-parallel rebinning of one raw pull into the reference's
 * post-reshape layout out[v][row][ch_os] for v in [0, n_proj_out), i.e. the
 * rebinned projection (v + add_projections) of
 *   rebin_nffs  reference rebin_filter.cu:58-186, rebin_filter.cuh:72-96
 *   rebin_pffs  reference rebin_filter.cu:188-385, rebin_filter.cuh:99-169
 *   rebin_zffs  reference rebin_filter.cu:386-585, rebin_filter.cuh:173-228
 *   rebin_affs  reference rebin_filter.cu:586-912, rebin_filter.cuh:231-350
 * without the ramp filter (which acts per row and commutes with the view
 * trimming of reshape_out, rebin_filter.cuh:354-366).
 * raw: [n_proj_pull][n_rows_raw][n_channels].
 */
int orc_rebin(const struct recon_metadata *mr, const float *raw, float *out) {
    const struct ct_geom *cg = &mr->cg;
    const long nch = (long)cg->n_channels, nos = (long)cg->n_channels_oversampled;
    const long nrr = (long)cg->n_rows_raw, nrows = (long)cg->n_rows;
    const long n_ffs = (long)mr->ri.n_ffs;
    const long n_proj = (long)mr->ri.n_proj_pull / n_ffs;
    const long add = (long)cg->add_projections;
    const long n_out = n_proj - 2 * add;
    const REAL K = (REAL)cg->n_proj_ffs / (2 * (REAL)PI_TRUE);
    const REAL cc = (REAL)cg->central_channel, dbeta = (REAL)cg->fan_angle_increment;
    double da_d, dr_d;
    orc_ffs_offsets(mr, &da_d, &dr_d);
    const REAL da = (REAL)da_d, dr = (REAL)dr_d;
    if (n_out <= 0) return 1;

    /* focal-spot sheets: sheet[fs][row][ch][p] = raw[n_ffs*p + fs][row][ch]
     * (p_reshape / z_reshape / a_reshape, rebin_filter.cuh:99-109,173-183,231-245) */
    float *sheet = (float *)malloc(sizeof(float) * n_ffs * nrr * nch * n_proj);
    if (!sheet) return 2;
    /* (the loops below write disjoint outputs: any thread order gives the
     * same bits) */
#pragma omp parallel for collapse(2) schedule(static)
    for (long fs = 0; fs < n_ffs; fs++)
        for (long r = 0; r < nrr; r++)
            for (long c = 0; c < nch; c++)
                for (long p = 0; p < n_proj; p++)
                    sheet[((fs * nrr + r) * nch + c) * n_proj + p] = raw[((n_ffs * p + fs) * nrr + r) * nch + c];

    if (n_ffs == 1) {
        /* n1_rebin/n2_rebin: texture (x = channel, y = projection) of the
         * row sheet [proj][ch] */
        float *rs = (float *)malloc(sizeof(float) * n_proj * nch);
        if (!rs) { free(sheet); return 2; }
        for (long r = 0; r < nrows; r++) {
            for (long p = 0; p < n_proj; p++)
                for (long c = 0; c < nch; c++) rs[p * nch + c] = raw[(p * nrr + r) * nch + c];
#pragma omp parallel for schedule(static)
            for (long ch = 0; ch < nos; ch++) {
                REAL beta = RASIN(((REAL)ch - 2 * cc) * (dbeta / 2));
                REAL beta_idx = beta / dbeta + cc;
                for (long v = 0; v < n_out; v++) {
                    REAL alpha_idx = (REAL)(v + add) - beta * (REAL)cg->n_proj_turn / (2 * (REAL)PI_TRUE);
                    out[(v * nrows + r) * nos + ch] = (float)bilerp_clamp(rs, nch, n_proj, beta_idx, alpha_idx);
                }
            }
        }
        free(rs);
        free(sheet);
        return 0;
    }

    if (n_ffs == 2 && mr->rp.z_ffs == 1) {
        /* z-FFS: lookup tables over raw channels (beta_lookup kernel,
         * rebin_filter.cu:493-498) and one bilinear pass per focal spot with
         * texture (x = projection, y = channel) of the sheet [ch][proj]
         * (z1_rebin/z2_rebin, rebin_filter.cuh:186-228). */
        REAL *lut = (REAL *)malloc(sizeof(REAL) * 2 * nch);
        if (!lut) { free(sheet); return 2; }
        for (long c = 0; c < nch; c++) {
            lut[c] = orc_beta_rk(cg, 0, -dr, (REAL)c);
            lut[nch + c] = orc_beta_rk(cg, 0, dr, (REAL)c);
        }
        for (int fs = 0; fs < 2; fs++) {
            REAL drs = fs == 0 ? -dr : dr;
            const REAL *l = lut + fs * nch;
            REAL rfr = orc_r_fr(cg, 0, drs), dal = orc_d_alpha_r(cg, 0, drs);
            for (long r = 0; r < nrr; r++) {
                long orow = (fs == 0) ^ (cg->reverse_row_interleave != 0) ? 2 * r : 2 * r + 1;
                const float *sh = sheet + (fs * nrr + r) * nch * n_proj;
#pragma omp parallel for schedule(static)
                for (long ch = 0; ch < nos; ch++) {
                    REAL beta = RASIN(((REAL)ch - 2 * cc) * (dbeta / 2) * (REAL)cg->r_f / rfr);
                    REAL beta_idx = orc_get_beta_idx(beta, l, (int)nch);
                    for (long v = 0; v < n_out; v++) {
                        REAL a = (REAL)n_ffs * (REAL)(v + add) - beta * K - dal * K;
                        a = fs == 0 ? a / 2 : (a - 1) / 2;
                        out[(v * nrows + orow) * nos + ch] = (float)bilerp_clamp(sh, n_proj, nch, a, beta_idx);
                    }
                }
            }
        }
        free(lut);
        free(sheet);
        return 0;
    }

    /* phi-FFS (n_ffs 2) and phi+z FFS (n_ffs 4): two passes.
     * Pass 1 (p1/p2_rebin_t, a1..a4_rebin_t): azimuthal rebin of each focal
     * spot into interleaved oversampled channels, texture (x = projection,
     * y = channel exact).  Pass 2 (p1/p2_rebin, a1/a2_rebin_b): radial
     * rebin through the interleaved beta table, texture (x = projection
     * exact, y = beta index). */
    {
        const int npairs = n_ffs == 4 ? 2 : 1;
        /* focal spot parameters: (da, dr, alpha offset) per fs index */
        REAL fda[4], fdr[4];
        if (n_ffs == 2) {
            fda[0] = da; fdr[0] = 0;
            fda[1] = -da; fdr[1] = 0;
        } else {
            fda[0] = da; fdr[0] = -dr;
            fda[1] = -da; fdr[1] = -dr;
            fda[2] = da; fdr[2] = dr;
            fda[3] = -da; fdr[3] = dr;
        }
        REAL *lut = (REAL *)malloc(sizeof(REAL) * npairs * nos);
        REAL *aoff = (REAL *)malloc(sizeof(REAL) * npairs * nos);
        float *tsheet = (float *)malloc(sizeof(float) * nos * n_proj);
        if (!lut || !aoff || !tsheet) { free(lut); free(aoff); free(tsheet); free(sheet); return 2; }
        for (int pr = 0; pr < npairs; pr++)
            for (long c = 0; c < nch; c++)
                for (int h = 0; h < 2; h++) {
                    int fs = 2 * pr + h;
                    REAL beta = orc_beta_rk(cg, fda[fs], fdr[fs], (REAL)c);
                    lut[pr * nos + 2 * c + h] = beta;
                    aoff[pr * nos + 2 * c + h] = beta * K + orc_d_alpha_r(cg, fda[fs], fdr[fs]) * K;
                }
        for (int pr = 0; pr < npairs; pr++) {
            const REAL *l = lut + pr * nos;
            REAL rfr = n_ffs == 4 ? orc_r_fr(cg, 0, pr == 0 ? -dr : dr) : (REAL)cg->r_f;
            for (long r = 0; r < nrr; r++) {
                /* pass 1 for this raw row: tsheet[m][p] */
#pragma omp parallel for schedule(static)
                for (long m = 0; m < nos; m++) {
                    int h = (int)(m & 1);
                    int fs = 2 * pr + h;
                    long c = m >> 1;
                    const float *sh = sheet + ((long)fs * nrr + r) * nch * n_proj;
                    for (long p = 0; p < n_proj; p++) {
                        REAL a = ((REAL)n_ffs * (REAL)p - aoff[pr * nos + m] - (REAL)fs) / (REAL)n_ffs;
                        tsheet[m * n_proj + p] = (float)bilerp_clamp(sh, n_proj, nch, a, (REAL)c);
                    }
                }
                long orow;
                if (n_ffs == 4) orow = (pr == 0) ^ (cg->reverse_row_interleave != 0) ? 2 * r : 2 * r + 1;
                else orow = r;
#pragma omp parallel for schedule(static)
                for (long ch = 0; ch < nos; ch++) {
                    REAL beta = RASIN(((REAL)ch - 2 * cc) * (dbeta / 2) * (REAL)cg->r_f / rfr);
                    REAL beta_idx = orc_get_beta_idx(beta, l, (int)nos);
                    for (long v = 0; v < n_out; v++)
                        out[(v * nrows + orow) * nos + ch] =
                            (float)bilerp_clamp(tsheet, n_proj, nos, (REAL)(v + add), beta_idx);
                }
            }
        }
        free(lut);
        free(aoff);
        free(tsheet);
    }
    free(sheet);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Ramp filter                                                              */
/* ------------------------------------------------------------------------ */

/* rows[r][l] <- sum_k taps[n + l - k] * rows[r][k] (direct sum in REAL):
 * reference rebin_filter.cuh:369-409 (GPU) and the FFT form of
 * rebin_filter_cpu.cu:647-709, which computes the same linear convolution. */
void orc_filter_direct(int n, const float *taps, long n_rows_total, float *rows) {
    REAL *in = (REAL *)malloc(sizeof(REAL) * n);
    for (long r = 0; r < n_rows_total; r++) {
        float *row = rows + r * n;
        for (int k = 0; k < n; k++) in[k] = row[k];
        for (int l = 0; l < n; l++) {
            REAL acc = 0;
            for (int k = 0; k < n; k++) acc += (REAL)taps[n + l - k] * in[k];
            row[l] = (float)acc;
        }
    }
    free(in);
}

/* ------------------------------------------------------------------------ */
/* Backprojection                                                           */
/* ------------------------------------------------------------------------ */

/* Stierstorfer weight, Q = 0.6: reference backproject.cuh:33-46 */
static REAL orc_W(REAL q) {
    REAL a = q < 0 ? -q : q;
    const REAL Q = (REAL)0.6f;
    if (a < Q) return 1;
    if (a < 1) {
        REAL c = RCOS(((REAL)PI_BP / 2) * (a - Q) / (1 - Q));
        return c * c;
    }
    return 0;
}

/* Zero-border bilinear sample of the pass-stacked texture the reference GPU
 * builds for one 16-angle pass (backproject.cu:71-86, addressing
 * backproject.cu:55-65, fetch backproject.cuh:85-87): texture row R of the
 * pass starting at angle i_pass belongs to rebinned view
 * i_pass + (R / n_rows) % 16 + ((R / n_rows) / 16) * half, detector row
 * R % n_rows.  Texels outside the array are zero. */
static REAL bp_fetch_gpu(const float *rebin, long nos, long nrows, long half, long n_ht, long i_pass,
                         REAL p, REAL R) {
    long h = 16 * nrows * n_ht;
    long x0 = (long)RFLOOR(p), y0 = (long)RFLOOR(R);
    REAL fx = p - (REAL)x0, fy = R - (REAL)y0;
    REAL v[2][2];
    for (int dy = 0; dy < 2; dy++)
        for (int dx = 0; dx < 2; dx++) {
            long x = x0 + dx, y = y0 + dy;
            if (x < 0 || x >= nos || y < 0 || y >= h) { v[dy][dx] = 0; continue; }
            long blk = y / nrows, row = y % nrows;
            long view = i_pass + blk % 16 + (blk / 16) * half;
            v[dy][dx] = rebin[(view * nrows + row) * nos + x];
        }
    return v[0][0] * (1 - fx) * (1 - fy) + v[0][1] * fx * (1 - fy) + v[1][0] * (1 - fx) * fy + v[1][1] * fx * fy;
}

/* Border rule of the CPU twin's interp2 (interp.cu:29-64): zero when the
 * index leaves [0, dim-2]. */
static REAL bp_fetch_cpu(const float *proj, long w, long h, REAL x, REAL y) {
    if (x > (REAL)(w - 2) || x < 0 || y > (REAL)(h - 2) || y < 0) return 0;
    long x0 = (long)RFLOOR(x), y0 = (long)RFLOOR(y);
    REAL v = x - (REAL)x0, ww = y - (REAL)y0;
    return proj[y0 * w + x0] * (1 - v) * (1 - ww) + proj[y0 * w + x0 + 1] * v * (1 - ww) +
           proj[(y0 + 1) * w + x0] * (1 - v) * ww + proj[(y0 + 1) * w + x0 + 1] * v * ww;
}

/*
 * orc_backproject: weighted FBP of one 32-slice block.
 *
 * GPU conventions (default; reference backproject.cu:31-108,
 * backproject.cuh:48-198): x flipped unless table_direction == -1, q sign
 * -table_direction, output layout [z][xi][yi], voxels outside AcqFOV/2 set to
 * zero, NaN fetch resets h_t, zero-border bilinear on the pass-stacked texture.
 * CPU conventions (reference backproject_cpu.cu:49-154): no flip, q sign +1,
 * output layout [z][yi][xi], no FOV mask, interp2 border rule, one view per
 * pass.
 *
 * Only the voxel columns listed in (cols_x, cols_y) are computed when n_cols
 * > 0 (others are left untouched); n_cols == 0 computes the whole block.
 * rebin: [n_proj_out][n_rows][n_channels_oversampled]; image: 32*nx*ny.
 */
int orc_backproject(const struct recon_metadata *mr, const float *rebin, float *image, int conventions,
                    const int *cols_x, const int *cols_y, long n_cols) {
    const struct ct_geom *cg = &mr->cg;
    const struct recon_params *rp = &mr->rp;
    const long nx = rp->nx, ny = rp->ny, nz = (long)mr->ri.n_slices_block;
    const long nos = (long)cg->n_channels_oversampled, nrows = (long)cg->n_rows;
    const long half = (long)cg->n_proj_turn / 2;
    const long n_ht = (long)((mr->ri.n_proj_pull / mr->ri.n_ffs - 2 * cg->add_projections) / (cg->n_proj_turn / 2));
    const int cpu = (conventions & 1) == ORC_BP_CPU_CONVENTIONS;
    const int no_flip = (conventions & ORC_BP_PLANT_NO_FLIP) != 0;
    const int no_qsign = (conventions & ORC_BP_PLANT_NO_QSIGN) != 0;
    /* reference backproject.cu:37 (float tube_start, pi' constant) */
    float tube_start_f = (float)(fmod((double)(mr->tube_angles[0] +
                                                ((float)(mr->ri.idx_pull_start + cg->add_projections_ffs) * 360.0f /
                                                 (float)cg->n_proj_ffs)),
                                      360.0) *
                                 PI_BP / 180.0f);
    const REAL tube_start = (REAL)tube_start_f;
    const REAL pi = (REAL)PI_BP;
    const REAL td = (REAL)cg->table_direction;
    const REAL rf = (REAL)cg->r_f, ds = rf * (REAL)cg->fan_angle_increment / 2;
    const REAL tanc = RTAN((REAL)cg->theta_cone / 2);
    const REAL scale = 2 * pi / (REAL)cg->n_proj_turn;
    long ncol = n_cols > 0 ? n_cols : nx * ny;
    int fail = 0;
    /* voxel columns are independent: one per iteration, any thread (each
     * column's arithmetic is the same sequential statement) */
#pragma omp parallel reduction(| : fail)
    {
    REAL *s = (REAL *)malloc(sizeof(REAL) * nz);
    REAL *st = (REAL *)malloc(sizeof(REAL) * nz);
    REAL *ht = (REAL *)malloc(sizeof(REAL) * nz);
    if (!s || !st || !ht) fail = 1;
#pragma omp for schedule(dynamic, 16)
    for (long ci = 0; ci < ncol; ci++) {
        if (fail) continue;
        long xi = n_cols > 0 ? cols_x[ci] : ci / ny;
        long yi = n_cols > 0 ? cols_y[ci] : ci % ny;
        REAL xc = (REAL)xi - (REAL)(nx - 1) / 2;
        REAL x = ((REAL)rp->recon_fov / (REAL)nx) * ((cpu || no_flip || cg->table_direction == -1) ? xc : -xc) +
                 (REAL)rp->x_origin;
        REAL y = ((REAL)rp->recon_fov / (REAL)ny) * ((REAL)yi - (REAL)(ny - 1) / 2) + (REAL)rp->y_origin;
        for (long z = 0; z < nz; z++) s[z] = 0;
        for (long i = 0; i < half; i++) {
            for (long z = 0; z < nz; z++) { st[z] = 0; ht[z] = 0; }
            long i_pass = (i / 16) * 16;
            for (long k = 0; k < n_ht; k++) {
                REAL theta = tube_start + (2 * pi / (REAL)cg->n_proj_turn) * (REAL)i + (REAL)k * pi;
                REAL sn = RSIN(theta), cs = RCOS(theta);
                REAL phat = x * sn - y * cs;
                REAL p_idx = phat / ds + 2 * (REAL)cg->central_channel;
                REAL ray = (REAL)cg->z_rot * (theta - RASIN(phat / rf)) / (2 * pi);
                REAL lhat = RSQRT(rf * rf - phat * phat) - x * cs - y * sn;
                for (long z = 0; z < nz; z++) {
                    REAL zz = (REAL)z * (REAL)rp->coll_slicewidth + (REAL)cg->z_rot / 2 +
                              (REAL)cg->z_rot * tube_start / (2 * pi);
                    REAL qhat = ((cpu || no_qsign) ? 1 : -td) * (zz - ray) / (lhat * tanc);
                    REAL w = orc_W(qhat);
                    REAL val;
                    if (cpu) {
                        REAL q_idx = ((qhat + 1) / 2) * (REAL)(nrows - 1) + (REAL)(k * nrows);
                        /* proj array of the CPU twin: [k][row][ch] of views i + k*half */
                        long hgt = nrows * n_ht;
                        if (p_idx > (REAL)(nos - 2) || p_idx < 0 || q_idx > (REAL)(hgt - 2) || q_idx < 0) val = 0;
                        else {
                            long y0 = (long)RFLOOR(q_idx);
                            long kk = y0 / nrows, rr = y0 % nrows;
                            /* rows y0 and y0+1 may straddle two stacked views */
                            long kk1 = (y0 + 1) / nrows, rr1 = (y0 + 1) % nrows;
                            long x0 = (long)RFLOOR(p_idx);
                            REAL fx = p_idx - (REAL)x0, fy = q_idx - (REAL)y0;
                            const float *r0 = rebin + ((i + kk * half) * nrows + rr) * nos;
                            const float *r1 = rebin + ((i + kk1 * half) * nrows + rr1) * nos;
                            val = r0[x0] * (1 - fx) * (1 - fy) + r0[x0 + 1] * fx * (1 - fy) +
                                  r1[x0] * (1 - fx) * fy + r1[x0 + 1] * fx * fy;
                        }
                        st[z] += val * w;
                        ht[z] += w;
                    } else {
                        REAL R = ((qhat + 1) / 2) * (REAL)(nrows - 1) + (REAL)(nrows * (i - i_pass)) +
                                 (REAL)(k * 16 * nrows);
                        val = bp_fetch_gpu(rebin, nos, nrows, half, n_ht, i_pass, p_idx, R) * w;
                        if (val == val) { st[z] += val; ht[z] += w; }
                        else ht[z] = 0;
                    }
                }
            }
            for (long z = 0; z < nz; z++)
                if (ht[z] != 0) s[z] += (1 / ht[z]) * st[z];
        }
        for (long z = 0; z < nz; z++) {
            if (cpu) {
                image[z * nx * ny + yi * nx + xi] = (float)(s[z] * scale);
            } else {
                int in_fov = RSQRT(x * x + y * y) < (REAL)rp->acq_fov / 2;
                image[z * nx * ny + xi * ny + yi] = in_fov ? (float)(s[z] * scale) : 0.0f;
            }
        }
    }
    free(s);
    free(st);
    free(ht);
    }
    (void)bp_fetch_cpu;
    return fail ? 2 : 0;
}

/* ------------------------------------------------------------------------ */
/* Finalize: block reorder, slice thickening, HU                            */
/* ------------------------------------------------------------------------ */

/*
 * orc_finalize: reference finalize_image_stack.cu:26-120 /
 * finalize_image_stack.cuh:26-76 (and the identical CPU twin
 * finalize_image_stack_cpu.cu:24-117): reverse each 32-slice block when the
 * recon direction differs from the table direction, then the normalised
 * triangle average with FWHM = SliceThickness at every final slice.  The
 * FP32 build accumulates with fmaf in the reference's order (nvcc contracts
 * "acc += w*v"); mu_water > 0 appends HU = 1000 (mu - mu_w) / mu_w.
 */
int orc_finalize(const struct recon_metadata *mr, const float *image, float *final_stack, float mu_water) {
    const struct recon_params *rp = &mr->rp;
    const struct recon_info *ri = &mr->ri;
    const long nxy = (long)rp->nx * rp->ny;
    const long nb = (long)ri->n_slices_block;
    const long n_raw = nb * (long)ri->n_blocks;
    float d = rp->end_pos - rp->start_pos;
    int recon_direction = d == 0.0f ? 1 : (int)(fabsf(d) / d);
    if (recon_direction != 1 && recon_direction != -1) recon_direction = 1;
    double td = mr->table_positions[1000] - mr->table_positions[0];
    int table_direction = td == 0 ? 1 : (int)(fabs(td) / td);
    int n_final = (int)floorf(fabsf(rp->end_pos - rp->start_pos) / rp->slice_thickness) + 1;
    float *w = (float *)malloc(sizeof(float) * n_raw);
    float *raw_loc = (float *)malloc(sizeof(float) * n_raw);
    if (!w || !raw_loc) { free(w); free(raw_loc); return 2; }
    for (long i = 0; i < n_raw; i++) raw_loc[i] = ri->recon_start_pos + recon_direction * (int)i * rp->coll_slicewidth;
    memset(final_stack, 0, sizeof(float) * n_final * nxy);
    for (int k = 0; k < n_final; k++) {
        float slice_location = rp->start_pos + recon_direction * k * rp->slice_thickness;
        float sum_w = 0.0f;
        for (long i = 0; i < n_raw; i++) {
            w[i] = fmaxf(0.0f, 1.0f - fabsf(raw_loc[i] - slice_location) / rp->slice_thickness);
            sum_w += w[i];
        }
        for (long v = 0; v < nxy; v++) {
            REAL acc = 0;
            for (long i = 0; i < n_raw; i++) {
                if (w[i] == 0) continue;
                long b = i / nb, z = i % nb;
                long src = recon_direction != table_direction ? b * nb + (nb - 1 - z) : i;
                float wn = w[i] / sum_w;
                if (ORC_IS_DOUBLE) acc += (REAL)wn * (REAL)image[src * nxy + v];
                else acc = (REAL)fmaf(wn, image[src * nxy + v], (float)acc);
            }
            float out = (float)acc;
            if (mu_water > 0) {
                if (ORC_IS_DOUBLE) out = (float)(1000.0 * ((double)acc - (double)mu_water) / (double)mu_water);
                else out = 1000.0f * (out - mu_water) / mu_water;
            }
            final_stack[(long)k * nxy + v] = out;
        }
    }
    free(w);
    free(raw_loc);
    return n_final;
}

/* ------------------------------------------------------------------------ */
/* Adaptive raw-data filter (FP32 only: the reference stage is FP32 and its  */
/* thresholding is discontinuous, so the restatement keeps float arithmetic */
/* and the fused multiply-adds nvcc forms by default).                      */
/* ------------------------------------------------------------------------ */

static int odd_window(float frac, long n_proj_ffs) {
    int n = (int)ceilf(frac * (float)n_proj_ffs);
    return n + (1 - n % 2);
}

/* reference preprocessing.cuh:32-62 */
static float orc_rhs(const float *array, int start_idx, int max_idx, int numel_range, float T, long n_proj_ffs) {
    int sum = 0;
    int diff = start_idx + (numel_range - 1);
    if (diff < (numel_range - 1)) {
        numel_range += start_idx;
        start_idx = 0;
    }
    if (diff > max_idx) numel_range -= (diff - max_idx);
    for (int i = 0; i < numel_range; i++)
        if (array[start_idx + i] > T) sum += 1;
    return (1.0f / PI_BP) * (float)sum * (1.0f / (float)n_proj_ffs);
}

/* Per-view thresholds: extract_sup, smooth_sup, eccentricity and
 * find_thresholds (reference preprocessing.cuh:69-206, driver
 * preprocessing.cu:54-120).  thresholds has n_proj_pull entries. */
void orc_adaptive_thresholds(const struct recon_metadata *mr, const float *raw, float *thresholds) {
    const long nch = (long)mr->cg.n_channels, nrr = (long)mr->cg.n_rows_raw;
    const long np = (long)mr->ri.n_proj_pull;
    const long npf = (long)mr->cg.n_proj_ffs;
    float *sup = (float *)malloc(sizeof(float) * np);
    float *sm = (float *)malloc(sizeof(float) * np);
    float *mx = (float *)malloc(sizeof(float) * np);
    float *mn = (float *)malloc(sizeof(float) * np);
    float *ecc = (float *)malloc(sizeof(float) * np);
    for (long p = 0; p < np; p++) {
        float m = 0.0f;
        for (long i = 0; i < nch * nrr; i++) m = fmaxf(m, raw[p * nch * nrr + i]);
        sup[p] = m;
    }
    int n_avg = odd_window(18.0f / 360.0f, npf);
    int off = n_avg / 2;
    int max_idx = (int)np - 1;
    float inv = 1.0f / (float)n_avg;
    for (long p = 0; p < np; p++) {
        float acc = 0.0f;
        for (int i = 0; i < n_avg; i++) {
            int idx = (int)p - off + i;
            if (idx < 0) idx = 0;
            else if (idx > max_idx) idx = max_idx;
            acc = fmaf(inv, sup[idx], acc);
        }
        sm[p] = acc;
    }
    int n_range = odd_window(180.0f / 360.0f, npf);
    int roff = n_range / 2;
    for (long p = 0; p < np; p++) {
        float hi = 0.0f, lo = 10000.0f;
        for (int i = 0; i < n_range; i++) {
            int idx = (int)p - roff + i;
            if (idx < 0) idx = 0;
            else if (idx > max_idx) idx = max_idx;
            float v = sm[idx];
            if (v > hi) hi = v;
            if (v < lo) lo = v;
        }
        mx[p] = hi;
        mn[p] = lo;
        float e = 1 - lo / hi;
        e = (e - 0.3f) / (0.5f - 0.3f);
        ecc[p] = e > 1 ? 1.0f : (e < 0 ? 0.0f : e);
    }
    float strength = mr->rp.adaptive_filtration_s;
    for (long p = 0; p < np; p++) {
        float lower = mn[p], upper = mx[p];
        float T = (mx[p] + mn[p]) / 2.0f;
        float lhs = 0.03f * strength * ecc[p];
        float rhs = orc_rhs(sm, (int)p - roff, max_idx, n_range, T, npf);
        int iter = 0;
        while ((double)(fabsf(lhs - rhs) / lhs) > 0.01) {
            if (iter >= 20) break;
            iter += 1;
            if (rhs > lhs) lower = T;
            else if (rhs < lhs) upper = T;
            else break;
            T = (lower + upper) / 2.0f;
            rhs = orc_rhs(sm, (int)p - roff, max_idx, n_range, T, npf);
        }
        thresholds[p] = T;
    }
    free(sup); free(sm); free(mx); free(mn); free(ecc);
}

/* 5x5x5 triangle weight with ceil(5/2) = 3 (peak at offset +1):
 * reference preprocessing.cuh:64-67 */
static float orc_triangle(int i, int j, int k) {
    int c = 3;
    return (float)((c - abs(i - c)) + (c - abs(j - c)) + (c - abs(k - c)) + 1);
}

/* reference preprocessing.cuh:208-285 (3D branch) */
void orc_adaptive_apply(const struct recon_metadata *mr, const float *raw, const float *thresholds, float *out) {
    const int nch = (int)mr->cg.n_channels, nrr = (int)mr->cg.n_rows_raw;
    const int np = (int)mr->ri.n_proj_pull;
    const long ps = (long)nch * nrr;
    /* every output sample is its own sequential statement: any thread order
     * gives the same bits */
#pragma omp parallel for schedule(dynamic, 4)
    for (long p = 0; p < np; p++) {
        float T = thresholds[p];
        for (int i = 0; i < nch; i++)
            for (int j = 0; j < nrr; j++) {
                long idx = ps * p + (long)j * nch + i;
                if (raw[idx] > T) {
                    float norm = 0.0f, val = 0.0f;
                    for (int ii = 0; ii < 5; ii++)
                        for (int jj = 0; jj < 5; jj++)
                            for (int kk = 0; kk < 5; kk++) {
                                int fr = j + jj - 2, fc = i + ii - 2;
                                long fp = p + kk - 2;
                                if (fr < 0 || fc < 0 || fp < 0 || fr >= nrr || fc >= nch || fp >= np) continue;
                                float t = orc_triangle(ii, jj, kk);
                                norm += t;
                                val = fmaf(raw[ps * fp + (long)nch * fr + fc], t, val);
                            }
                    out[idx] = (1.0f / norm) * val;
                } else {
                    out[idx] = raw[idx];
                }
            }
    }
}

void orc_adaptive_filter(const struct recon_metadata *mr, const float *raw, float *out, float *thresholds_out) {
    long np = (long)mr->ri.n_proj_pull;
    float *T = thresholds_out ? thresholds_out : (float *)malloc(sizeof(float) * np);
    orc_adaptive_thresholds(mr, raw, T);
    orc_adaptive_apply(mr, raw, T, out);
    if (!thresholds_out) free(T);
}

/* Number of bytes of REAL in this build (8 = FP64 oracle, 4 = FP32). */
int orc_real_size(void) { return (int)sizeof(REAL); }
