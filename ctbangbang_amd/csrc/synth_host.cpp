// Host side of the synthetic scan generator: phantom construction (seeded
// std::mt19937_64) and the CPU evaluation of the same samples the HIP kernel
// produces.
#include <cmath>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "geometry.h"
#include "synth_common.h"
#include "synth_host.h"

namespace ctbb {

namespace {

// Rotation about z by phi, then the Euler z-x-z angles for the random
// ellipsoids; returns world->body rows scaled by 1/semi-axes.
void make_rinv(double ax, double ay, double az, double phi, double theta, double psi, double *R) {
    const double c1 = std::cos(phi), s1 = std::sin(phi), c2 = std::cos(theta), s2 = std::sin(theta);
    const double c3 = std::cos(psi), s3 = std::sin(psi);
    // body axes expressed in world coordinates (rows of the world->body matrix)
    double M[9] = {c1 * c3 - s1 * c2 * s3, s1 * c3 + c1 * c2 * s3, s2 * s3,
                   -c1 * s3 - s1 * c2 * c3, -s1 * s3 + c1 * c2 * c3, s2 * c3,
                   s1 * s2, -c1 * s2, c2};
    const double inv[3] = {1.0 / ax, 1.0 / ay, 1.0 / az};
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) R[3 * r + c] = M[3 * r + c] * inv[r];
}

}  // namespace

std::vector<Ellipsoid> make_phantom(const recon_metadata &mr, const ctbb_phantom_params &pp) {
    // Shepp-Logan-type 3D head (Kak-Slaney layout): x0, y0, z0, a, b, c,
    // rotation about z (deg), attenuation relative to mu_water.
    static const double head[10][8] = {
        {0.00, 0.00, 0.00, 0.690, 0.920, 0.900, 0, 1.80},
        {0.00, -0.0184, 0.00, 0.6624, 0.874, 0.880, 0, -0.80},
        {0.22, 0.00, -0.25, 0.11, 0.31, 0.22, -18, -0.20},
        {-0.22, 0.00, -0.25, 0.16, 0.41, 0.28, 18, -0.20},
        {0.00, 0.35, -0.25, 0.21, 0.25, 0.41, 0, 0.10},
        {0.00, 0.10, -0.25, 0.046, 0.046, 0.05, 0, 0.10},
        {0.00, -0.10, -0.25, 0.046, 0.046, 0.05, 0, 0.10},
        {-0.08, -0.605, -0.25, 0.046, 0.023, 0.05, 0, 0.10},
        {0.00, -0.605, -0.25, 0.023, 0.023, 0.05, 0, 0.10},
        {0.06, -0.605, -0.25, 0.023, 0.046, 0.05, 0, 0.10},
    };
    const double s = pp.scale_mm > 0 ? pp.scale_mm : 100.0;
    const double muw = pp.mu_water > 0 ? pp.mu_water : 0.02;
    std::vector<Ellipsoid> E;
    for (const auto &h : head) {
        Ellipsoid e;
        e.c[0] = h[0] * s;
        e.c[1] = h[1] * s;
        e.c[2] = h[2] * s + pp.z_center;
        make_rinv(h[3] * s, h[4] * s, h[5] * s, h[6] * 3.14159265358979323846 / 180.0, 0.0, 0.0, e.rinv);
        e.mu = h[7] * muw;
        E.push_back(e);
    }
    std::mt19937_64 rng(pp.phantom_seed);
    std::uniform_real_distribution<double> U(0.0, 1.0);
    const double rmax = 0.7 * mr.rp.recon_fov / 2.0;
    const double zext = 0.9 * s;
    for (int i = 0; i < pp.n_random; i++) {
        Ellipsoid e;
        const double rr = rmax * std::sqrt(U(rng)), th = 2.0 * 3.14159265358979323846 * U(rng);
        e.c[0] = rr * std::cos(th);
        e.c[1] = rr * std::sin(th);
        e.c[2] = pp.z_center + (2.0 * U(rng) - 1.0) * zext;
        const double ax = 5.0 + 25.0 * U(rng), ay = 5.0 + 25.0 * U(rng), az = 5.0 + 25.0 * U(rng);
        const double phi = 2.0 * 3.14159265358979323846 * U(rng);
        const double theta = std::acos(2.0 * U(rng) - 1.0);
        const double psi = 2.0 * 3.14159265358979323846 * U(rng);
        make_rinv(ax, ay, az, phi, theta, psi, e.rinv);
        e.mu = 0.01 + 0.02 * U(rng);
        E.push_back(e);
    }
    return E;
}

SynthGeom make_synth_geom(const recon_metadata &mr, const ctbb_phantom_params &pp) {
    const ct_geom &cg = mr.cg;
    SynthGeom g;
    std::memset(&g, 0, sizeof(g));
    g.r_f = cg.r_f;
    g.sdd = cg.src_to_det;
    g.dbeta = cg.fan_angle_increment;
    g.cc = cg.central_channel;
    g.tanc = std::tan(static_cast<double>(cg.theta_cone) / 2.0);
    g.zrot = cg.z_rot;
    g.n_proj_ffs = static_cast<double>(cg.n_proj_ffs);
    g.start_deg = mr.rp.tube_start_angle;
    g.n_ffs = static_cast<int>(cg.n_proj_ffs / cg.n_proj_turn);
    g.n_rows = static_cast<int>(cg.n_rows);
    g.n_rows_raw = static_cast<int>(cg.n_rows_raw);
    g.nch = static_cast<int>(cg.n_channels);
    g.reverse_rows = cg.reverse_row_interleave ? 1 : 0;
    g.table_dir = cg.table_direction == -1 ? -1 : 1;
    g.n_readings = mr.rp.n_readings;
    g.noise_seed = pp.noise_seed;
    g.i0 = pp.i0;
    double da, dr;
    ffs_offsets(mr, &da, &dr);
    // focal-spot order of the reference's reshape kernels
    // (include/rebin_filter.cuh:99-109,173-183,231-245)
    if (mr.rp.phi_ffs && mr.rp.z_ffs) {
        const double tda[4] = {da, -da, da, -da}, tdr[4] = {-dr, -dr, dr, dr};
        const int set[4] = {0, 0, 1, 1};
        for (int i = 0; i < 4; i++) { g.fs_da[i] = tda[i]; g.fs_dr[i] = tdr[i]; g.fs_set[i] = set[i]; }
        g.interleave = 1;
    } else if (mr.rp.z_ffs) {
        g.fs_dr[0] = -dr; g.fs_dr[1] = dr; g.fs_set[0] = 0; g.fs_set[1] = 1;
        g.interleave = 1;
    } else if (mr.rp.phi_ffs) {
        g.fs_da[0] = da; g.fs_da[1] = -da;
    }
    return g;
}

}  // namespace ctbb

extern "C" int ctbb_synth_scan(const struct recon_metadata *mr, const ctbb_phantom_params *pp, long first, long count,
                               float *out) {
    if (!mr || !pp || !out || count < 0) return CTBB_ERR_ARG;
    if (mr->cg.n_proj_turn == 0 || mr->cg.n_channels == 0 || mr->cg.n_rows_raw == 0) return CTBB_ERR_ARG;
    const std::vector<ctbb::Ellipsoid> E = ctbb::make_phantom(*mr, *pp);
    const ctbb::SynthGeom g = ctbb::make_synth_geom(*mr, *pp);
    const long per = static_cast<long>(g.n_rows_raw) * g.nch;
    const long total = count * per;
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 16) nt = 16;
    std::vector<std::thread> pool;
    for (unsigned t = 0; t < nt; t++)
        pool.emplace_back([&, t] {
            for (long i = t; i < total; i += nt) {
                const long n = first + i / per;
                const int j = static_cast<int>((i % per) / g.nch), ch = static_cast<int>(i % g.nch);
                out[i] = ctbb::synth_sample(g, E.data(), static_cast<int>(E.size()), n, j, ch);
            }
        });
    for (auto &th : pool) th.join();
    return CTBB_OK;
}
