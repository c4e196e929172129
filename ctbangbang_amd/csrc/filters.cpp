// Ramp / apodisation filter taps: file loader and closed-form generator.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

#include "ctbb_hip.h"

namespace {

// r(t) = sin(t)/t + (cos(t)-1)/t^2 with r(0) = 1/2
// (reference resources/ramp.m:13-16).
double ramp_r(double t) {
    if (t == 0.0) return 0.5;
    return std::sin(t) / t + (std::cos(t) - 1.0) / (t * t);
}

const char *kernel_name(int recon_kernel, char *buf, size_t n) {
    // reference src/rebin_filter.cu:929-948
    switch (recon_kernel) {
        case -100: return "ramp";
        case -1: return "exp";
        case 1: return "smooth";
        case 2: return "medium";
        case 3: return "sharp";
        default: std::snprintf(buf, n, "b%i", recon_kernel); return buf;
    }
}

}  // namespace

extern "C" {

// f(k) = c^2/(2 ds) * (a r(c pi k) + (1-a)/2 r(pi c k + pi) + (1-a)/2 r(pi c k - pi))
// for k = j - n, j in [0, 2n): reference resources/ramp.m:10.  Evaluated in
// double and rounded to float, which reproduces the reference's 736-channel
// tables bit for bit.
int ctbb_ramp_filter(int n, double ds, double c, double a, float *taps) {
    if (n <= 0 || !taps || ds <= 0.0) return CTBB_ERR_ARG;
    const double pi = 3.14159265358979323846;
    for (int j = 0; j < 2 * n; j++) {
        double k = static_cast<double>(j - n);
        double t = c * pi * k;
        double v = a * ramp_r(t) + ((1.0 - a) / 2.0) * ramp_r(pi * c * k + pi) +
                   ((1.0 - a) / 2.0) * ramp_r(pi * c * k - pi);
        taps[j] = static_cast<float>((c * c / (2.0 * ds)) * v);
    }
    return CTBB_OK;
}

// reference src/rebin_filter.cu:922-959
int ctbb_load_filter(const struct recon_metadata *mr, float *taps) {
    if (!mr || !taps) return CTBB_ERR_ARG;
    char nb[32];
    const char *name = kernel_name(mr->rp.recon_kernel, nb, sizeof(nb));
    std::string path = std::string(mr->install_dir) + "/resources/filters/f_" +
                       std::to_string(mr->cg.n_channels) + "_" + name + ".txt";
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return CTBB_ERR_FILTER;
    size_t n = 2 * mr->cg.n_channels_oversampled;
    size_t got = std::fread(taps, sizeof(float), n, f);
    std::fclose(f);
    if (got != n) return CTBB_ERR_FILTER;
    return CTBB_OK;
}

}  // extern "C"
