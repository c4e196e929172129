// Synthetic helical scan: analytic line integrals of an ellipsoid phantom
// through the flying-focal-spot geometry the rebinning assumes, plus Poisson
// noise.  Shared by the host generator (synth_host.cpp) and the HIP kernel
// (synth.hip); everything is evaluated in double.
//
// Geometry (per reading n, focal spot fs = n mod n_ffs):
//   tube angle   alpha = (TubeStartAngle + 360 n / n_proj_ffs) degrees
//   nominal source S = r_f (cos a, sin a); towards-iso axis a1 = -(cos a, sin a);
//   lateral axis a2 = (sin a, -cos a)
//   focal spot   F = S - dr_fs a1 - da_fs a2  (the frame in which the
//                reference's beta_rk / d_alpha_r / r_fr are written,
//                include/rebin_filter.cuh:36-51)
//   channel ch   D = S + SDD (cos b0 a1 + sin b0 a2), b0 = (ch - cc) dbeta
//   row          rebinned row r of raw row j (z-FFS interleave), cone slope
//                dz/dl = -q_r tan(theta_cone/2), q_r = 2 r/(n_rows-1) - 1
//                (the row model of the weighted backprojection,
//                include/backproject.cuh:82-85)
//   source z     the table position of reading n.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define CTBB_HD __host__ __device__
#else
#include <cmath>
#define CTBB_HD
#endif

namespace ctbb {

struct Ellipsoid {
    double c[3];      // centre (mm)
    double rinv[9];   // world -> body rotation, rows scaled by 1/semi-axis
    double mu;        // attenuation added inside (1/mm)
};

struct SynthGeom {
    double r_f, sdd, dbeta, cc, tanc, zrot, n_proj_ffs, start_deg;
    double fs_da[4], fs_dr[4];
    int fs_set[4];     // rebinned-row parity of each focal spot (z-FFS)
    int n_ffs, n_rows, n_rows_raw, nch, interleave, reverse_rows;
    int table_dir;     // +1: z = n*zrot/n_proj_ffs, -1: z = (n_readings - n)*zrot/n_proj_ffs
    long n_readings;
    uint64_t noise_seed;
    double i0;
};

CTBB_HD inline uint64_t splitmix64(uint64_t x) {
    x += 0x9e3779b97f4a7c15ull;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

CTBB_HD inline double u01(uint64_t x) { return (static_cast<double>(x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

CTBB_HD inline double ellipsoid_sum(const Ellipsoid *E, int n_e, const double p0[3], const double d[3]) {
    double sum = 0.0;
    for (int e = 0; e < n_e; e++) {
        const double px = p0[0] - E[e].c[0], py = p0[1] - E[e].c[1], pz = p0[2] - E[e].c[2];
        const double *R = E[e].rinv;
        const double u0 = R[0] * px + R[1] * py + R[2] * pz;
        const double u1 = R[3] * px + R[4] * py + R[5] * pz;
        const double u2 = R[6] * px + R[7] * py + R[8] * pz;
        const double v0 = R[0] * d[0] + R[1] * d[1] + R[2] * d[2];
        const double v1 = R[3] * d[0] + R[4] * d[1] + R[5] * d[2];
        const double v2 = R[6] * d[0] + R[7] * d[1] + R[8] * d[2];
        const double A = v0 * v0 + v1 * v1 + v2 * v2;
        const double B = u0 * v0 + u1 * v1 + u2 * v2;
        const double C = u0 * u0 + u1 * u1 + u2 * u2 - 1.0;
        const double disc = B * B - A * C;
        if (disc > 0.0) sum += E[e].mu * 2.0 * sqrt(disc) / A;
    }
    return sum;
}

// Poisson(lambda) draw from two counter-based uniforms: inversion below
// lambda = 64, rounded normal approximation above.
CTBB_HD inline double poisson_draw(double lambda, uint64_t key) {
    const uint64_t h1 = splitmix64(key), h2 = splitmix64(h1 ^ 0x5851f42d4c957f2dull);
    if (lambda < 64.0) {
        double u = u01(h1);
        double p = exp(-lambda), cdf = p;
        double k = 0.0;
        while (u > cdf && k < 1000.0) {
            k += 1.0;
            p *= lambda / k;
            cdf += p;
        }
        return k;
    }
    const double r = sqrt(-2.0 * log(u01(h1)));
    const double z = r * cos(6.283185307179586 * u01(h2));
    const double k = floor(lambda + sqrt(lambda) * z + 0.5);
    return k < 0.0 ? 0.0 : k;
}

CTBB_HD inline float synth_sample(const SynthGeom &g, const Ellipsoid *E, int n_e, long n, int j, int ch) {
    const int fs = static_cast<int>(n % g.n_ffs);
    const double a = (g.start_deg + 360.0 * static_cast<double>(n) / g.n_proj_ffs) * (3.14159265358979323846 / 180.0);
    const double ca = cos(a), sa = sin(a);
    const double Sx = g.r_f * ca, Sy = g.r_f * sa;
    const double a1x = -ca, a1y = -sa, a2x = sa, a2y = -ca;
    const double Fx = Sx - g.fs_dr[fs] * a1x - g.fs_da[fs] * a2x;
    const double Fy = Sy - g.fs_dr[fs] * a1y - g.fs_da[fs] * a2y;
    const double b0 = (static_cast<double>(ch) - g.cc) * g.dbeta;
    const double cb = cos(b0), sb = sin(b0);
    const double Dx = Sx + g.sdd * (cb * a1x + sb * a2x);
    const double Dy = Sy + g.sdd * (cb * a1y + sb * a2y);
    double ux = Dx - Fx, uy = Dy - Fy;
    const double L = sqrt(ux * ux + uy * uy);
    ux /= L;
    uy /= L;
    int r = j;
    if (g.interleave) r = 2 * j + (g.fs_set[fs] ^ g.reverse_rows);
    const double q = 2.0 * r / (g.n_rows - 1) - 1.0;
    const double slope = -q * g.tanc;
    const double zs = g.table_dir == -1 ? (static_cast<double>(g.n_readings) - static_cast<double>(n)) * g.zrot / g.n_proj_ffs
                                        : static_cast<double>(n) * g.zrot / g.n_proj_ffs;
    const double nrm = sqrt(1.0 + slope * slope);
    const double p0[3] = {Fx, Fy, zs};
    const double d[3] = {ux / nrm, uy / nrm, slope / nrm};
    double p = ellipsoid_sum(E, n_e, p0, d);
    if (g.i0 > 0.0) {
        const uint64_t key = g.noise_seed * 0x9e3779b97f4a7c15ull +
                             static_cast<uint64_t>((n * g.n_rows_raw + j) * static_cast<long>(g.nch) + ch);
        const double counts = poisson_draw(g.i0 * exp(-p), key);
        p = -log((counts > 1.0 ? counts : 1.0) / g.i0);
    }
    return static_cast<float>(p);
}

}  // namespace ctbb
