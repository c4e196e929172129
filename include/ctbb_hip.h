/*
 * ctbb_hip.h — C-ABI of the MI355X reconstruction hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns a ctbb_status (0 = success) instead of calling exit() the way the
 * reference program does.  The drop-in wrappers with the reference's own
 * C++ names (ctbb_compat.h) map these codes back to the reference's exits.
 *
 * Three groups of entry points:
 *   1. planning / I/O on `struct recon_metadata` (host only, no GPU needed);
 *   2. the four hot-path stages with the reference's host-memory contract
 *      (raw in host memory, d_rebin on the device between rebin and
 *      backprojection, image and final stack in host memory);
 *   3. the same stages on device-resident buffers, for callers that keep the
 *      scan in HBM (whole-volume pipelines, benchmarks, batching).
 *
 * Data layouts (float32 unless noted):
 *   raw pull        [n_proj_pull][n_rows_raw][n_channels]        channel fastest
 *   rebinned block  [n_proj_out][n_rows][n_channels_oversampled] (n_proj_out =
 *                   n_proj_pull/n_ffs - 2*add_projections), ramp-filtered
 *   image block     [32][nx][ny]  (slice, xi, yi; yi fastest)
 *   image stack     [n_slices_recon][nx][ny]
 *   final stack     [n_slices_final][nx][ny]
 */
#ifndef CTBB_HIP_H
#define CTBB_HIP_H

#include <stddef.h>
#include <stdint.h>
#include "recon_structs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    CTBB_OK = 0,
    CTBB_ERR_ARG = 1,       /* invalid argument / null pointer */
    CTBB_ERR_IO = 2,        /* file missing, unreadable or unwritable */
    CTBB_ERR_PRM = 3,       /* an essential prm value is missing or zero */
    CTBB_ERR_RANGE = 4,     /* requested slices outside the allowed data range */
    CTBB_ERR_SCANNER = 5,   /* scanner cannot be parsed */
    CTBB_ERR_FILTER = 6,    /* filter file missing or unreadable */
    CTBB_ERR_SHAPE = 7,     /* geometry outside what the kernels support */
    CTBB_ERR_HIP = 8,       /* HIP runtime error (no device, launch failure) */
    CTBB_ERR_NOMEM = 9,     /* host or device allocation failed */
    CTBB_ERR_UNSUPPORTED = 10, /* raw file type without a reader in this build */
    CTBB_ERR_LIMIT = 11,      /* input beyond the fixed capacity of a no-allocation entry point */
    CTBB_ERR_NONFINITE = 12   /* the raw data reaching the reconstruction held NaN or infinite samples */
} ctbb_status;

const char *ctbb_strerror(int status);

/* ------------------------------------------------------------------------ */
/* 1. Planning and I/O (host only).                                          */
/* ------------------------------------------------------------------------ */

/* Parse a prm file, convert TableDir to table_dir and check the essential
 * fields.  Replaces configure_recon_params (reference src/setup.cu:108-176);
 * CTBB_ERR_PRM where the reference exits with code 1, CTBB_ERR_IO where the
 * prm is missing (reference main.cu:165-168 exits 13). */
int ctbb_configure_recon_params(const char *prm_file, struct recon_params *out);

/* Linear attenuation of water (1/mm) given by the "MuWater:" token of a prm
 * file, or 0 when absent.  The token is unknown to the reference parser. */
int ctbb_prm_mu_water(const char *prm_file, float *mu_water);

/* Value of a numeric prm token this build adds (e.g. "AdaptiveThreshold:"),
 * or 0 when absent; the reference parser ignores such tokens. */
int ctbb_prm_float(const char *prm_file, const char *token, float *value);

/* Resolve cwd / install_dir / output path and check the raw file is readable
 * (1) and the output writable (2): reference src/setup.cu:40-106.
 * install_dir is the directory holding resources/ (the package directory of
 * this library unless the environment variable CTBB_INSTALL_DIR is set). */
int ctbb_configure_paths(struct recon_metadata *mr);

/* Scanner geometry from a scanner file (path, or file name under
 * install_dir/resources/scanners) or a hard-coded scanner 0/1/2:
 * reference src/setup.cu:178-383. */
int ctbb_configure_ct_geom(struct recon_metadata *mr, struct ct_geom *out);

/* Tube angles, table positions, slice/block plan, first pull and the host
 * arrays raw/rebin/image: reference src/setup.cu:385-603.  Reads only the raw
 * file's existence for binary (FileType 0) data. */
int ctbb_configure_reconstruction(struct recon_metadata *mr);

/* Same as ctbb_configure_reconstruction but never allocates ctd.raw,
 * ctd.rebin or ctd.image and does not open the raw file (tube angles and
 * table positions are analytic for FileType 0).  For device-resident use. */
int ctbb_plan_reconstruction(struct recon_metadata *mr);

/* Pull range of the next block; post-increments ri.cb.block_idx and
 * reallocates ctd.raw / ctd.rebin: reference src/setup.cu:605-679. */
int ctbb_update_block_info(struct recon_metadata *mr);

/* Same arithmetic as ctbb_update_block_info for block `block_idx` (0-based),
 * without touching any allocation; ri.cb.block_idx becomes block_idx+1. */
int ctbb_plan_block(struct recon_metadata *mr, int block_idx);

/* Read the current pull into ctd.raw (FileType 0 binary only; reference
 * src/setup.cu:681-767).  Frames past the end of the file read as zeros. */
int ctbb_extract_projections(struct recon_metadata *mr);

/* Write final_image_stack to OutputDir/OutputFile and free every host
 * allocation of mr: reference src/setup.cu:769-787. */
int ctbb_finish_and_cleanup(struct recon_metadata *mr);

/* reference src/setup.cu:806-825 */
int ctbb_array_search(float key, const double *array, int numel_array, int search_type);

/* Number of final (thickened) slices: reference src/finalize_image_stack.cu:76 */
int ctbb_n_slices_final(const struct recon_metadata *mr);

/* Load the 2*n_channels_oversampled ramp-filter taps for rp.recon_kernel from
 * install_dir/resources/filters (reference src/rebin_filter.cu:922-959). */
int ctbb_load_filter(const struct recon_metadata *mr, float *taps);

/* Closed-form apodised ramp filter (reference resources/ramp.m:1-16):
 * taps[j] = ramp(j - n, ds, c, a) for j in [0, 2n), n = n_channels_oversampled. */
int ctbb_ramp_filter(int n, double ds, double c, double a, float *taps);

/* ------------------------------------------------------------------------ */
/* 2./3. GPU execution.                                                      */
/* ------------------------------------------------------------------------ */

typedef struct ctbb_ctx ctbb_ctx;

typedef enum {
    CTBB_FILTER_F32 = 0,        /* FP32 Toeplitz product, FP32 accumulation */
    CTBB_FILTER_BF16X3 = 1,     /* BF16 hi+lo split operands on MFMA, FP32 accumulation */
    CTBB_FILTER_F16X3 = 2,      /* FP16 hi+lo split operands on MFMA, FP32 accumulation */
    CTBB_FILTER_FFT32 = 3       /* FP32 4096-point FFT convolution (rows up to 2048 channels; wider: F32) */
} ctbb_filter_mode;

typedef enum {
    CTBB_PROJ_F32 = 0,          /* filtered projections stored as float32 */
    CTBB_PROJ_BF16 = 1          /* filtered projections stored as bfloat16 */
} ctbb_proj_dtype;

typedef enum {
    CTBB_BATCH_PERSISTENT = 0,  /* groups of scans: captured preparation graph + one persistent
                                   backprojection launch per group (default; a batch of one
                                   scan runs the pipeline form) */
    CTBB_BATCH_PIPELINE = 1     /* one launch per block, preparation beside the previous block */
} ctbb_batch_mode;

typedef struct ctbb_options_s {
    int filter_mode;            /* ctbb_filter_mode; default CTBB_FILTER_FFT32 */
    int proj_dtype;             /* ctbb_proj_dtype */
    float mu_water;             /* >0: convert the final stack to HU */
    int bp_tile;                /* backprojection slices per thread: 16 (two z-groups per
                                   workgroup; 0 selects it) or 32; other values are rejected */
    int bp_views_per_pass;      /* reserved: the kernels stage one view per LDS pass (0 or 1) */
    int adaptive_filter;        /* 1 = run the adaptive raw-data filter (default) */
    float adaptive_threshold;   /* 0: the reference's per-view threshold (Kachelriess-Kalender
                                   bisection); > 0: smooth every sample above this fixed
                                   line-integral threshold, e.g. ln(I0 / N) for a noise threshold
                                   of N photons at I0 incident photons (no counterpart in the
                                   reference) */
    int batch_mode;             /* ctbb_batch_mode of ctbb_reconstruct_batch_dev */
    int batch_group_mb;         /* device memory of one batch group (filtered projections and
                                   images), MiB; 0: 8192 */
} ctbb_options;

/* Create a context on HIP device `device` (-1: current device).  The context
 * owns a non-blocking stream, device workspaces and per-device tables. */
int ctbb_ctx_create(int device, ctbb_ctx **out);
void ctbb_ctx_destroy(ctbb_ctx *ctx);
int ctbb_ctx_get_options(const ctbb_ctx *ctx, ctbb_options *opt);
int ctbb_ctx_set_options(ctbb_ctx *ctx, const ctbb_options *opt);
/* The HIP stream (hipStream_t) the context launches on, as an opaque pointer. */
void *ctbb_ctx_stream(ctbb_ctx *ctx);
/* Block until all work queued on the context's streams has finished.
 * Returns CTBB_ERR_NONFINITE (once) when a rebin since the previous call
 * met a non-finite raw sample: the results of that work are undefined.  The
 * reference's backprojection drops such samples and resets h_t
 * (include/backproject.cuh:89-95); this build reports them instead. */
int ctbb_ctx_synchronize(ctbb_ctx *ctx);
/* Device time of a kernel family during the most recent top-level call on
 * the context, from HIP events recorded on the context stream around each
 * launch.  family: "adaptive", "rebin", "filter", "backproject", "finalize",
 * "hu16".
 * Synchronises the stream; *total_ms = summed duration, *launches = count. */
int ctbb_ctx_kernel_time(ctbb_ctx *ctx, const char *family, float *total_ms, int *launches);

/* --- stages with the reference's host-memory contract -------------------- */

/* Adaptive raw-data filter, in place on mr->ctd.raw.
 * Replaces adaptive_filter_kk (reference include/preprocessing.h:26). */
int ctbb_adaptive_filter(ctbb_ctx *ctx, struct recon_metadata *mr);

/* Rebin (n/p/z/a flying focal spot) + ramp filter of mr->ctd.raw; allocates
 * mr->ctd.d_rebin on the device.  Replaces rebin_filter
 * (reference include/rebin_filter.h:27). */
int ctbb_rebin_filter(ctbb_ctx *ctx, struct recon_metadata *mr);

/* Weighted-FBP backprojection of mr->ctd.d_rebin into block
 * (ri.cb.block_idx-1) of mr->ctd.image; frees d_rebin.  Replaces backproject
 * (reference include/backproject.h:26). */
int ctbb_backproject(ctbb_ctx *ctx, struct recon_metadata *mr);

/* Block reorder + slice thickening (+ HU when options.mu_water > 0) of
 * mr->ctd.image into a newly allocated mr->ctd.final_image_stack.  Replaces
 * finalize_image_stack (reference include/finalize_image_stack.h:26). */
int ctbb_finalize_image_stack(ctbb_ctx *ctx, struct recon_metadata *mr);

/* --- device-resident stages ---------------------------------------------- */

/* Element counts of the current block's buffers (from mr->ri / mr->cg / mr->rp). */
typedef struct {
    size_t raw_pull;        /* n_proj_pull * n_rows_raw * n_channels */
    size_t rebin;           /* n_proj_out * n_rows * n_channels_oversampled */
    size_t image_block;     /* 32 * nx * ny */
    size_t image_stack;     /* n_slices_recon * nx * ny */
    size_t final_stack;     /* n_slices_final * nx * ny */
    size_t n_proj_out;
    int n_half_turns;
    int n_slices_final;
    double voxel_updates_block; /* nx*ny*32*(n_proj_turn/2)*n_half_turns */
} ctbb_sizes;

int ctbb_block_sizes(const struct recon_metadata *mr, ctbb_sizes *out);

/* d_out[pull] = adaptive filter of d_raw[pull] (out of place, both device). */
int ctbb_adaptive_filter_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                             const float *d_raw, float *d_out);

/* d_rebin = ramp-filtered rebinned block from the device raw pull.  d_rebin is
 * float* for CTBB_PROJ_F32 and uint16_t* (bfloat16) for CTBB_PROJ_BF16. */
int ctbb_rebin_filter_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                          const float *d_raw, void *d_rebin);

/* Rebin only (no ramp filter), float32 output.  Exposes the rebinned
 * sinogram the reference writes as its rebin test dump before filtering. */
int ctbb_rebin_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                   const float *d_raw, float *d_rebin);

/* Ramp filter rows of length n_channels_oversampled:
 * rows[r][l] <- sum_k taps[n + l - k] * rows[r][k].  d_out == d_in (in place)
 * is allowed only for the CTBB_FILTER_FFT32 filter (rows of at most 2048
 * channels) with CTBB_PROJ_F32 output; any other overlap of d_in and d_out
 * returns CTBB_ERR_ARG. */
int ctbb_filter_rows_dev(ctbb_ctx *ctx, const float *d_taps, int n, size_t n_rows_total,
                         const float *d_in, void *d_out);

/* Backproject a filtered block into d_image_block [32][nx][ny] (overwritten). */
int ctbb_backproject_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                         const void *d_rebin, float *d_image_block);

/* ctbb_backproject_dev without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * block's view table (its only scratch) goes to a fixed device array of
 * CTBB_NOALLOC_MAX_VIEWS entries (256 KiB per GPU, held until the process
 * ends) instead of the context's workspace; views = (n_proj_turn / 2) *
 * half turns of the block beyond that limit return CTBB_ERR_LIMIT before
 * anything is launched.  The array is shared by every call on a GPU: calls
 * that overlap on one GPU (other contexts or streams) race.  No event
 * timing is recorded.  Create the context (ctbb_ctx_create) beforehand. */
#define CTBB_NOALLOC_MAX_VIEWS 16384
int ctbb_backproject_dev_noalloc(ctbb_ctx *ctx, const struct recon_metadata *mr,
                                 const void *d_rebin, float *d_image_block);

/* Reorder + thicken (+HU) d_image [n_slices_recon][nx][ny] into d_final. */
int ctbb_finalize_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                      const float *d_image, float *d_final);

/* Backprojection autotuner: times one launch per candidate bp_tile (16: two
 * z-groups of 16 slices per workgroup; 32: one group of 32) for the middle
 * block of the plan in mr (on a zero-filled block: the work does not depend
 * on the values), sets the faster in the context's options and returns it in
 * *bp_tile_out.  With cache_path, results are kept per shape (rows, channels,
 * matrix, field of view, views, projection type) as lines "<key> <tile>" in
 * that text file, and a cached shape is used without timing.  mr->ri is
 * left as it was.  Both tiles
 * give the same images to FP32 rounding (the interior-chunk shortcut is
 * decided per wavefront, whose voxels and slices differ between the two).
 * No counterpart in the reference, whose tile is a compile-time constant
 * (include/backproject.cuh:24-25). */
int ctbb_autotune_backprojection(ctbb_ctx *ctx, struct recon_metadata *mr, const char *cache_path,
                                 int *bp_tile_out);

/* Output format: int16 stored values of a stack in HU, for a DICOM series
 * (HU = slope * stored + intercept, the Rescale Slope / Intercept of the
 * files): d_out[i] = (d_hu[i] - intercept) / slope in IEEE FP32, rounded half
 * to even, saturated to [-32768, 32767], NaN -> 0.  slope > 0 and finite,
 * intercept finite, else CTBB_ERR_ARG; n = 0 does nothing.  No counterpart in
 * the reference, which writes the float32 stack only (src/setup.cu:769-787). */
int ctbb_hu_to_int16_dev(ctbb_ctx *ctx, const float *d_hu, size_t n, float slope, float intercept,
                         int16_t *d_out);

/* Write a final stack of int16 stored values ([n_slices_final][nx][ny] host
 * memory, as ctbb_hu_to_int16_dev produces from the HU stack) as a DICOM CT
 * series: one Part-10 file per slice, explicit VR little endian, named
 * <dir>/<prefix>_NNNN.dcm (NNNN = slice number from 1).  Geometry in the
 * scanner frame of the reconstruction: slice k at z = start_pos +
 * direction * k * slice_thickness, pixel (row i, column j) of a slice is
 * voxel (x index i, y index j) of the stack, with the voxel centres of
 * reference include/backproject.cuh:59-64.  Study, series and instance UIDs
 * are new per call (2.25 root).  CTBB_ERR_IO if a file cannot be written.
 * No counterpart in the reference. */
int ctbb_write_dicom_series(const struct recon_metadata *mr, const int16_t *h_stack, float slope,
                            float intercept, const char *dir, const char *prefix);

/* Whole reconstruction from a device-resident scan: every block of the plan
 * in mr (ctbb_plan_reconstruction), then finalize, into d_final.  d_raw holds
 * readings [0, n_raw_readings) of the scan; pulls outside that range read
 * zeros.  d_image (n_slices_recon*nx*ny floats) may be NULL (context
 * workspace is used).  Leaves mr->ri.cb at the last block. */
int ctbb_reconstruct_dev(ctbb_ctx *ctx, struct recon_metadata *mr,
                         const float *d_raw, size_t n_raw_readings,
                         float *d_image, float *d_final);

/* Whole reconstruction from a scan in host memory: the same pipeline as
 * ctbb_reconstruct_dev, with each block's raw pull assembled in one of two
 * device buffers on a copy stream (readings shared with the previous pull
 * copied device to device, the new ones host to device, readings outside
 * [0, n_raw_readings) zero) while the previous block is reconstructed.  The
 * final stack (n_slices_final*nx*ny floats) is copied to h_final and the call
 * returns when it is there.  Pass page-locked h_raw (hipHostMalloc) for the
 * copies to overlap the compute.  No counterpart in the reference, whose
 * driver reads and uploads every pull whole (src/setup.cu:605-767). */
int ctbb_reconstruct_host(ctbb_ctx *ctx, struct recon_metadata *mr,
                          const float *h_raw, size_t n_raw_readings, float *h_final);

/* Batching form: n_scans device-resident scans (each with its own plan in
 * mrs[s], readings d_raws[s][0, n_raw_readings[s]) and output d_finals[s])
 * reconstructed as one pipeline over all (scan, block) items: block
 * preparation on the context stream overlaps the backprojection of the
 * previous item, and each scan's finalize runs behind its last
 * backprojection.  Scans may differ in length, pitch, FOV and matrix size;
 * scans that do not share the first scan's filter table and rebin maps
 * (scanner, kernel, focal-spot mode, collimation) make the call fall back to
 * one scan at a time.  Returns when the work is queued; ctbb_ctx_synchronize
 * waits for it.  No counterpart in the reference (one scan per process). */
int ctbb_reconstruct_batch_dev(ctbb_ctx *ctx, int n_scans, struct recon_metadata *const *mrs,
                               const float *const *d_raws, const size_t *n_raw_readings,
                               float *const *d_finals);

/* Block plans of a batch, computed on the device the way the batching form
 * builds its work list (the arithmetic of update_block_info, reference
 * src/setup.cu:605-679, and of the backprojection's tube start and half
 * turns, src/backproject.cu:37-39): one entry per (scan, block) in batch
 * order, into out[0 .. total blocks).  CTBB_ERR_LIMIT if they exceed
 * max_out.  For checking the device planner against ctbb_plan_block. */
typedef struct {
    int scan, block;
    int idx_pull_start, n_proj_pull, n_half_turns;
    int idx_block_slice_start, idx_block_slice_end;
    float block_slice_start, block_slice_end, tube_start;
} ctbb_block_plan;
int ctbb_plan_batch_dev(ctbb_ctx *ctx, int n_scans, struct recon_metadata *const *mrs, ctbb_block_plan *out,
                        int max_out);

/* ------------------------------------------------------------------------ */
/* CPU reconstruction path (the reference driver's --no-gpu mode).           */
/* ------------------------------------------------------------------------ */

/* Rebin + ramp filter of mr->ctd.raw into the host array mr->ctd.rebin
 * ([n_proj_out][n_rows][n_channels_oversampled]) on n_threads host threads
 * (0: every core): the same coordinate maps and interpolation as the HIP
 * rebin, an FP64 FFT filter.  Replaces rebin_filter_cpu (reference
 * include/rebin_filter_cpu.h:27, src/rebin_filter_cpu.cu:75). */
int ctbb_rebin_filter_cpu(struct recon_metadata *mr, int n_threads);

/* Backprojection of mr->ctd.rebin into block (ri.cb.block_idx - 1) of
 * mr->ctd.image with the reference CPU twin's conventions (no x flip, qhat
 * sign +1, layout [z][yi][xi], zero border outside [0, dim - 2]).  Replaces
 * backproject_cpu (reference include/backproject_cpu.h:26,
 * src/backproject_cpu.cu:49-154). */
int ctbb_backproject_cpu(struct recon_metadata *mr, int n_threads);

/* Block reorder + slice thickening (+ HU when mu_water > 0) of mr->ctd.image
 * into a newly allocated mr->ctd.final_image_stack.  Replaces
 * finalize_image_stack_cpu (reference include/finalize_image_stack_cpu.h:26,
 * src/finalize_image_stack_cpu.cu:24-117). */
int ctbb_finalize_image_stack_cpu(struct recon_metadata *mr, float mu_water, int n_threads);

/* ------------------------------------------------------------------------ */
/* Synthetic helical scans (test data generator).                           */
/* ------------------------------------------------------------------------ */

/* Analytic line integrals of a 3D Shepp-Logan-type head phantom plus
 * `n_random` seeded ellipsoids for every reading in [first, first+count) of
 * the scan planned in mr (geometry, FFS focal positions, tube angles, table
 * positions), with Poisson noise at i0 incident photons (i0 <= 0: no noise).
 * Output layout [count][n_rows_raw][n_channels].  `_dev` writes device memory,
 * the host variant host memory; both evaluate the same FP64 expressions (they
 * can differ in the last bits where the host and device math libraries do). */
typedef struct {
    uint64_t phantom_seed;
    uint64_t noise_seed;
    int n_random;           /* extra random ellipsoids (8 in the standard scans) */
    double i0;              /* incident photons per sample; <=0 disables noise */
    double scale_mm;        /* phantom unit length in mm (head ~ 0.69 x 0.92 units) */
    double z_center;        /* phantom centre in table coordinates (mm) */
    double mu_water;        /* attenuation of the phantom's "brain" level (1/mm) */
} ctbb_phantom_params;

int ctbb_synth_scan(const struct recon_metadata *mr, const ctbb_phantom_params *pp,
                    long first, long count, float *out);
int ctbb_synth_scan_dev(ctbb_ctx *ctx, const struct recon_metadata *mr,
                        const ctbb_phantom_params *pp, long first, long count, float *d_out);

#ifdef __cplusplus
}
#endif

#endif
