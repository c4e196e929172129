// Reference-named entry points (C++ linkage) over the ctbb_* C-ABI, with the
// reference's exit() behaviour.  See include/ctbb_compat.h.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <mutex>

#include "ctbb_compat.h"
#include "ctbb_hip.h"

namespace {

ctbb_ctx *g_ctx = nullptr;
std::once_flag g_once;
ctbb_options g_opt{};      // last options given to ctbb_compat_set_options (the CPU path reads mu_water)

ctbb_ctx *default_ctx() {
    std::call_once(g_once, [] {
        if (ctbb_ctx_create(-1, &g_ctx) != CTBB_OK) g_ctx = nullptr;
    });
    if (!g_ctx) {
        std::fprintf(stderr, "CTBangBang: no HIP device available\n");
        std::exit(1);
    }
    return g_ctx;
}

int check(int st, const char *what) {
    if (st == CTBB_OK) return 0;
    std::fprintf(stderr, "%s: %s\n", what, ctbb_strerror(st));
    if (st == CTBB_ERR_HIP) std::exit(1);
    std::exit(1);
}


}  // namespace

int adaptive_filter_kk(struct recon_metadata *mr) { return check(ctbb_adaptive_filter(default_ctx(), mr), "adaptive_filter_kk"); }

int rebin_filter(struct recon_metadata *mr) {
    int st = ctbb_rebin_filter(default_ctx(), mr);
    if (st == CTBB_ERR_FILTER) {
        // reference src/rebin_filter.cu:952-955
        std::perror("Filter file not found.  You may need to create one.  We hope to add on-the-fly creation in a future update.");
        std::exit(1);
    }
    return check(st, "rebin_filter");
}

int backproject(struct recon_metadata *mr) { return check(ctbb_backproject(default_ctx(), mr), "backproject"); }

int finalize_image_stack(struct recon_metadata *mr) {
    return check(ctbb_finalize_image_stack(default_ctx(), mr), "finalize_image_stack");
}

// CPU twins (the reference driver's --no-gpu mode), on every host core
int rebin_filter_cpu(struct recon_metadata *mr) {
    int st = ctbb_rebin_filter_cpu(mr, 0);
    if (st == CTBB_ERR_FILTER) {
        std::perror("Filter file not found.  You may need to create one.  We hope to add on-the-fly creation in a future update.");
        std::exit(1);
    }
    return check(st, "rebin_filter_cpu");
}
int backproject_cpu(struct recon_metadata *mr) { return check(ctbb_backproject_cpu(mr, 0), "backproject_cpu"); }
int finalize_image_stack_cpu(struct recon_metadata *mr) {
    return check(ctbb_finalize_image_stack_cpu(mr, g_opt.mu_water, 0), "finalize_image_stack_cpu");
}

int configure_paths(struct recon_metadata *mr) { return ctbb_configure_paths(mr); }

struct recon_params configure_recon_params(char *filename) {
    struct recon_params p;
    int st = ctbb_configure_recon_params(filename, &p);
    if (st == CTBB_ERR_PRM) {
        std::printf("An essential parameter was not properly set in configuration.  Check parameter file.\n");
        std::exit(1);  // reference src/setup.cu:171-173
    }
    if (st) {
        std::perror("Parameter file not found. Check path and try again. Exiting:");
        std::exit(13);
    }
    return p;
}

struct ct_geom configure_ct_geom(struct recon_metadata *mr) {
    struct ct_geom cg;
    if (ctbb_configure_ct_geom(mr, &cg) != CTBB_OK) {
        std::perror("Could not parse selected scanner");
        std::exit(1);  // reference src/setup.cu:204-207
    }
    return cg;
}

void configure_reconstruction(struct recon_metadata *mr) {
    int st = ctbb_configure_reconstruction(mr);
    if (st == CTBB_ERR_RANGE) {
        std::printf("Requested reconstruction is outside of allowed data range: %.2f to %.2f\n", mr->ri.allowed_begin,
                    mr->ri.allowed_end);
        std::exit(1);  // reference src/setup.cu:537-545
    }
    if (st == CTBB_ERR_IO) {
        std::perror("Raw data file not found.");
        std::exit(1);
    }
    if (st) check(st, "configure_reconstruction");
}

void update_block_info(recon_metadata *mr) { check(ctbb_update_block_info(mr), "update_block_info"); }
void extract_projections(struct recon_metadata *mr) { check(ctbb_extract_projections(mr), "extract_projections"); }
void finish_and_cleanup(struct recon_metadata *mr) { check(ctbb_finish_and_cleanup(mr), "finish_and_cleanup"); }

void parse_config(char *config_file, struct recon_params *structure) {
    if (ctbb_parse_config(config_file, structure) != CTBB_OK) {
        std::perror("Parameter file not found. Check path and try again. Exiting:");
        std::exit(13);
    }
}

extern "C" int ctbb_compat_set_options(const struct ctbb_options_s *opt) {
    if (!opt) return CTBB_ERR_ARG;
    g_opt = *opt;
    return ctbb_ctx_set_options(default_ctx(), opt);
}
