// ctbangbang — command-line driver with the reference program's interface
// (reference src/main.cu:44-271): same flags, same prm input, same block loop
// and stage order, same float32 output stack.  Built on the reference-named
// wrappers of ctbb_compat.h, so this file is also the proof that a driver
// written against the reference's API links against libctbb_hip.so.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "ctbb_compat.h"
#include "ctbb_hip.h"

namespace {

int g_verbose = 0;

void logv(const char *fmt, ...) {
    if (!g_verbose) return;
    va_list ap;
    va_start(ap, fmt);
    std::vprintf(fmt, ap);
    va_end(ap);
}

void usage() {
    std::printf("\nusage: ctbangbang [options] input_prm_file\n\n");
    std::printf("    Options:\n");
    std::printf("          -v: verbose.\n");
    std::printf("          -t: testing (keep the rebinned sinogram of each block in host memory).\n");
    std::printf("    --no-gpu: rebin/filter, backproject and finalize on the host CPU cores\n"
                "              (the adaptive filter still runs on the GPU, as in the reference).\n");
    std::printf("  --device=i: run on GPU device number 'i'\n");
    std::printf("    --timing: Display timing information for each step of the recon process\n");
    std::printf(" --benchmark: accepted for compatibility (no effect)\n");
    std::printf(" --empty-prm: write an empty parameter file to the given path\n\n");
    std::exit(0);
}

double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

int main(int argc, char **argv) {
    static struct recon_metadata mr;
    std::memset(&mr, 0, sizeof(mr));
    if (argc < 2) usage();
    for (int i = 1; i < argc - 1; i++) {
        if (!std::strcmp(argv[i], "-t")) mr.flags.testing = 1;
        else if (!std::strcmp(argv[i], "-v")) mr.flags.verbose = 1;
        else if (!std::strcmp(argv[i], "--no-gpu")) mr.flags.no_gpu = 1;
        else if (!std::strncmp(argv[i], "--device=", 9)) {
            mr.flags.set_device = 1;
            mr.flags.device_number = std::atoi(argv[i] + 9);
        } else if (!std::strcmp(argv[i], "--timing")) mr.flags.timing = 1;
        else if (!std::strcmp(argv[i], "--benchmark")) mr.flags.benchmark = 1;
        else if (!std::strcmp(argv[i], "--empty-prm")) std::exit(ctbb_empty_config(argv[argc - 1]) ? 1 : 0);
        else usage();
    }
    g_verbose = mr.flags.verbose;
    // device selection: last device unless --device (reference main.cu:125-152)
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count == 0) {
        std::fprintf(stderr, "no HIP device found\n");
        return 1;
    }
    int dev = mr.flags.set_device ? mr.flags.device_number : count - 1;
    if (hipSetDevice(dev) != hipSuccess) std::printf("There was a problem setting device.\n");
    logv("Working on GPU %i\n", dev);

    const char *prm = argv[argc - 1];
    logv("Reading PRM file...\n");
    mr.rp = configure_recon_params(const_cast<char *>(prm));
    int err = configure_paths(&mr);
    if (err) {
        if (err == 1) std::perror("Cannot read specified raw file.  Ensure the path is correct and you have permission to read the file");
        if (err == 2) std::perror("Cannot write to specified output file/directory.  Ensure the path is correct and you have permission to write to the directory");
        return 13;
    }
    ctbb_options opt;
    std::memset(&opt, 0, sizeof(opt));
    opt.filter_mode = CTBB_FILTER_F32;
    opt.proj_dtype = CTBB_PROJ_F32;
    opt.adaptive_filter = 1;
    float mu_w = 0.0f, thr = 0.0f;
    ctbb_prm_mu_water(prm, &mu_w);
    opt.mu_water = mu_w;
    // fixed adaptive-filter threshold (line integral), a token of this build
    ctbb_prm_float(prm, "AdaptiveThreshold:", &thr);
    opt.adaptive_threshold = thr > 0.0f ? thr : 0.0f;
    ctbb_compat_set_options(&opt);

    logv("Configuring scanner geometry...\n");
    mr.cg = configure_ct_geom(&mr);
    logv("Configuring final reconstruction parameters...\n");
    configure_reconstruction(&mr);
    logv("Allowed recon range: %.2f to %.2f\n", mr.ri.allowed_begin, mr.ri.allowed_end);

    double t_all = now_ms(), t_af = 0, t_rb = 0, t_bp = 0;
    for (size_t b = 0; b < mr.ri.n_blocks; b++) {
        update_block_info(&mr);
        logv("Working on block %zu of %zu\n", b + 1, mr.ri.n_blocks);
        extract_projections(&mr);
        double t0 = now_ms();
        adaptive_filter_kk(&mr);
        double t1 = now_ms();
        if (mr.flags.no_gpu) rebin_filter_cpu(&mr);
        else rebin_filter(&mr);
        double t2 = now_ms();
        if (mr.flags.no_gpu) backproject_cpu(&mr);
        else backproject(&mr);
        double t3 = now_ms();
        t_af += t1 - t0;
        t_rb += t2 - t1;
        t_bp += t3 - t2;
    }
    double t4 = now_ms();
    if (mr.flags.no_gpu) finalize_image_stack_cpu(&mr);
    else finalize_image_stack(&mr);
    double t5 = now_ms();
    logv("Writing image data to %s/%s\n", mr.rp.output_dir, mr.rp.output_file);
    finish_and_cleanup(&mr);
    if (mr.flags.timing) {
        std::printf("%0.2f ms for adaptive_filtration\n", t_af);
        std::printf("%0.2f ms for rebinning and filtering\n", t_rb);
        std::printf("%0.2f ms for backprojections\n", t_bp);
        std::printf("%0.2f ms for reordering and thickening slices\n", t5 - t4);
        std::printf("%0.2f ms for entire reconstruction\n", now_ms() - t_all);
    }
    logv("\nDone.\n");
    return 0;
}
