/*
 * recon_structs.h — reconstruction metadata (drop-in contract).
 *
 * `struct recon_metadata` is handed, by pointer, to every stage of the
 * reconstruction.  Its layout is an ABI shared with drivers written against the
 * reference program (reference include/recon_structs.h:26-113), so field order,
 * types and array sizes are kept identical and pinned below with static
 * assertions (x86-64 / LP64).
 *
 * Ownership rules of the pointers in `struct ct_data` follow the reference
 * (src/setup.cu:600-602,676-677; src/rebin_filter.cu:160; src/backproject.cu:99):
 *   raw, rebin, image      host, owned by the caller/driver
 *   d_rebin                device, allocated by the rebin/filter stage, released
 *                          by the backprojection stage
 *   final_image_stack      host, allocated by the finalize stage, freed by the
 *                          driver after writing the output file
 */
#ifndef CTBB_RECON_STRUCTS_H
#define CTBB_RECON_STRUCTS_H

#include <stddef.h>
#include "parse_config.h"

struct block_info {
    int block_idx;                /* post-incremented by update_block_info */
    float block_slice_start;
    float block_slice_end;
    int idx_block_slice_start;
    int idx_block_slice_end;
};

struct recon_info {
    size_t n_ffs;
    float data_begin_pos;
    float data_end_pos;
    float allowed_begin;
    float allowed_end;
    size_t n_slices_requested;
    size_t n_slices_recon;
    size_t n_slices_block;
    size_t n_blocks;
    size_t idx_slice_start;
    size_t idx_slice_end;
    float recon_start_pos;
    float recon_end_pos;
    size_t idx_pull_start;
    size_t idx_pull_end;
    size_t n_proj_pull;
    struct block_info cb;
};

struct ct_data {
    float *raw;
    float *rebin;
    float *image;
    float *d_raw;
    float *d_rebin;
    float *d_image;
    float *d_final_image_stack;
    float *final_image_stack;
};

struct ct_geom {
    size_t n_proj_turn;
    size_t n_proj_ffs;
    size_t n_channels;
    size_t n_channels_oversampled;
    size_t n_rows;
    size_t n_rows_raw;
    float r_f;
    float z_rot;
    float theta_cone;
    float fan_angle_increment;
    float src_to_det;
    float anode_angle;
    float central_channel;
    float acq_fov;
    size_t projection_offset;
    size_t add_projections;
    size_t add_projections_ffs;
    int reverse_row_interleave;
    int reverse_channel_interleave;
    /* -1: table positions decrease with reading index ("in"),
     *  1: table positions increase ("out"). */
    int table_direction;
};

struct flags {
    int testing;
    int verbose;
    int no_gpu;
    int set_device;
    int device_number;
    int timing;
    int benchmark;
};

struct recon_metadata {
    char home_dir[4096];
    char cwd[4096];
    char install_dir[4096];
    struct flags flags;
    struct recon_params rp;
    struct recon_info ri;
    struct ct_geom cg;
    struct ct_data ctd;
    float *tube_angles;          /* degrees, one per reading */
    double *table_positions;     /* mm, one per reading */
};

#if defined(__cplusplus)
#define CTBB_STATIC_ASSERT(c, m) static_assert(c, m)
#else
#define CTBB_STATIC_ASSERT(c, m) _Static_assert(c, m)
#endif

#if defined(__x86_64__) || defined(__aarch64__)
CTBB_STATIC_ASSERT(sizeof(struct recon_params) == 14172, "recon_params layout");
CTBB_STATIC_ASSERT(sizeof(struct recon_info) == 128, "recon_info layout");
CTBB_STATIC_ASSERT(sizeof(struct ct_geom) == 120, "ct_geom layout");
CTBB_STATIC_ASSERT(sizeof(struct ct_data) == 64, "ct_data layout");
CTBB_STATIC_ASSERT(sizeof(struct recon_metadata) == 26816, "recon_metadata layout");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, rp) == 12316, "rp offset");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, ri) == 26488, "ri offset");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, cg) == 26616, "cg offset");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, ctd) == 26736, "ctd offset");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, tube_angles) == 26800, "tube_angles offset");
CTBB_STATIC_ASSERT(offsetof(struct recon_metadata, table_positions) == 26808, "table_positions offset");
#endif

#endif
