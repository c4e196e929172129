// FP64 geometry on the host: the per-view backprojection table and the
// flying-focal-spot rebinning coordinate maps.  Both are evaluated in double
// from the reference formulas and rounded to float once, so the kernels only
// do the per-sample interpolation arithmetic in FP32.
#include <cmath>
#include <vector>

#include "ctbb_internal.h"
#include "geometry.h"

namespace ctbb {

namespace {

double angle2(double x1, double x2, double y1, double y2) {
    // reference include/rebin_filter.cuh:36-38
    return std::asin((x1 * y2 - x2 * y1) / (std::sqrt(x1 * x1 + x2 * x2) * std::sqrt(y1 * y1 + y2 * y2)));
}

double beta_rk(const ct_geom &cg, double da, double dr, double channel) {
    // reference include/rebin_filter.cuh:40-43 (os_flag = 0)
    double b0 = (channel - static_cast<double>(cg.central_channel)) * static_cast<double>(cg.fan_angle_increment);
    double sdd = cg.src_to_det, rf = cg.r_f;
    return angle2(-(rf + dr), -da, -(sdd * std::cos(b0) + dr), -(sdd * std::sin(b0) + da));
}

double d_alpha_r(const ct_geom &cg, double da, double dr) {
    // reference include/rebin_filter.cuh:45-47
    return angle2(cg.r_f, 0.0, cg.r_f + dr, da);
}

double r_fr(const ct_geom &cg, double da, double dr) {
    // reference include/rebin_filter.cuh:49-51
    return std::sqrt((cg.r_f + dr) * (cg.r_f + dr) + da * da);
}

double get_beta_idx(double beta, const std::vector<double> &lut) {
    // reference include/rebin_filter.cuh:53-64: first entry >= beta by linear
    // search, linear extrapolation beyond the table ends.
    int n = static_cast<int>(lut.size());
    int lo = 0;
    while (beta > lut[lo] && lo < n - 1) lo++;
    if (lo == 0) lo++;
    return static_cast<double>(lo) - 1.0 + (beta - lut[lo - 1]) / (lut[lo] - lut[lo - 1]);
}

}  // namespace

void ffs_offsets(const recon_metadata &mr, double *da, double *dr) {
    // reference src/rebin_filter.cu:194,397-398,593-594 (double on the host)
    const ct_geom &cg = mr.cg;
    *da = cg.src_to_det * cg.r_f * cg.fan_angle_increment / (4.0f * (cg.src_to_det - cg.r_f));
    *dr = cg.src_to_det * mr.rp.coll_slicewidth / (4.0 * (cg.src_to_det - cg.r_f) * std::tan(cg.anode_angle));
}

float bp_tube_start(const recon_metadata &mr) {
    // reference src/backproject.cu:37 (float result, pi' constant)
    const ct_geom &cg = mr.cg;
    float t = mr.tube_angles[0] + (static_cast<float>(mr.ri.idx_pull_start + cg.add_projections_ffs) * 360.0f /
                                   static_cast<float>(cg.n_proj_ffs));
    return static_cast<float>(std::fmod(static_cast<double>(t), 360.0) * kPiBp / 180.0f);
}

int bp_n_half_turns(const recon_metadata &mr) {
    // reference src/backproject.cu:39
    const ct_geom &cg = mr.cg;
    return static_cast<int>((mr.ri.n_proj_pull / mr.ri.n_ffs - 2 * cg.add_projections) / (cg.n_proj_turn / 2));
}

int rebin_maps(const recon_metadata &mr, std::vector<float> &bidx, std::vector<float> &aoff, int *kind,
               int *n_sets) {
    const ct_geom &cg = mr.cg;
    const int nch = static_cast<int>(cg.n_channels), nos = static_cast<int>(cg.n_channels_oversampled);
    const int n_ffs = static_cast<int>(mr.ri.n_ffs);
    const double cc = cg.central_channel, dbeta = cg.fan_angle_increment;
    const double K = static_cast<double>(cg.n_proj_ffs) / (2.0 * static_cast<double>(kPiTrue));
    double da, dr;
    ffs_offsets(mr, &da, &dr);
    if (n_ffs == 1) {
        // n1_rebin / n2_rebin, reference include/rebin_filter.cuh:72-96
        *kind = 0;
        *n_sets = 1;
        bidx.resize(nos);
        aoff.resize(nos);
        for (int c = 0; c < nos; c++) {
            double beta = std::asin((static_cast<double>(c) - 2.0 * cc) * (dbeta / 2.0));
            bidx[c] = static_cast<float>(beta / dbeta + cc);
            aoff[c] = static_cast<float>(beta * static_cast<double>(cg.n_proj_turn) / (2.0 * static_cast<double>(kPiTrue)));
        }
        return CTBB_OK;
    }
    if (n_ffs == 2 && mr.rp.z_ffs == 1) {
        // z1_rebin / z2_rebin with beta_lookup tables, reference
        // src/rebin_filter.cu:493-498, include/rebin_filter.cuh:186-228
        *kind = 1;
        *n_sets = 2;
        bidx.resize(2 * nos);
        aoff.resize(2 * nos);
        for (int fs = 0; fs < 2; fs++) {
            double drs = fs == 0 ? -dr : dr;
            std::vector<double> lut(nch);
            for (int c = 0; c < nch; c++) lut[c] = beta_rk(cg, 0.0, drs, c);
            double rfr = r_fr(cg, 0.0, drs), dal = d_alpha_r(cg, 0.0, drs);
            for (int c = 0; c < nos; c++) {
                double beta = std::asin((c - 2.0 * cc) * (dbeta / 2.0) * cg.r_f / rfr);
                bidx[fs * nos + c] = static_cast<float>(get_beta_idx(beta, lut));
                aoff[fs * nos + c] = static_cast<float>((beta * K + dal * K + fs) / 2.0);
            }
        }
        return CTBB_OK;
    }
    if (n_ffs == 2 || n_ffs == 4) {
        // Two passes: p1/p2_rebin_t + p1/p2_rebin (reference
        // include/rebin_filter.cuh:111-169) or a1..a4_rebin_t + a1/a2_rebin_b
        // (include/rebin_filter.cuh:247-350).
        *kind = n_ffs == 2 ? 2 : 3;
        const int npairs = n_ffs == 4 ? 2 : 1;
        *n_sets = npairs;
        double fda[4], fdr[4];
        if (n_ffs == 2) {
            fda[0] = da; fdr[0] = 0.0;
            fda[1] = -da; fdr[1] = 0.0;
        } else {
            fda[0] = da; fdr[0] = -dr;
            fda[1] = -da; fdr[1] = -dr;
            fda[2] = da; fdr[2] = dr;
            fda[3] = -da; fdr[3] = dr;
        }
        bidx.resize(static_cast<size_t>(npairs) * nos);
        aoff.resize(static_cast<size_t>(npairs) * nos);
        for (int pr = 0; pr < npairs; pr++) {
            std::vector<double> lut(nos);
            for (int c = 0; c < nch; c++)
                for (int h = 0; h < 2; h++) {
                    int fs = 2 * pr + h;
                    double beta = beta_rk(cg, fda[fs], fdr[fs], c);
                    lut[2 * c + h] = beta;
                    aoff[pr * nos + 2 * c + h] =
                        static_cast<float>((beta * K + d_alpha_r(cg, fda[fs], fdr[fs]) * K + fs) / n_ffs);
                }
            double rfr = n_ffs == 4 ? r_fr(cg, 0.0, pr == 0 ? -dr : dr) : static_cast<double>(cg.r_f);
            for (int c = 0; c < nos; c++) {
                double beta = std::asin((c - 2.0 * cc) * (dbeta / 2.0) * cg.r_f / rfr);
                bidx[pr * nos + c] = static_cast<float>(get_beta_idx(beta, lut));
            }
        }
        return CTBB_OK;
    }
    return CTBB_ERR_SHAPE;
}

}  // namespace ctbb
