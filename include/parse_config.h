/*
 * parse_config.h — prm parameter block of the reconstruction (drop-in contract).
 *
 * The field order, types and array sizes of `struct recon_params` are an ABI:
 * a driver built against the reference program passes this struct by value and
 * by pointer, so it must be byte-identical to the reference's definition
 * (reference include/parse_config.h:36-67).  Everything else in this header is
 * written for this project.
 *
 * Token grammar accepted by parse_config() (reference src/parse_config.cu:40-216,
 * resources/config_dict.txt:4-33): one "Token: value" per line, '%' starts a
 * comment, unknown tokens are ignored, "TableFeed:" is an alias of "PitchValue:".
 * This implementation additionally understands "MuWater:" (linear attenuation of
 * water in 1/mm, used for the HU conversion).  Because the reference ignores
 * unknown tokens, a prm carrying "MuWater:" still parses there unchanged; the
 * value is stored outside the ABI struct (see ctbb_prm_mu_water in ctbb_hip.h).
 */
#ifndef CTBB_PARSE_CONFIG_H
#define CTBB_PARSE_CONFIG_H

#include <stddef.h>

#define COMMENT_DELIM '%'

struct recon_params {
    char raw_data_dir[4096];
    char raw_data_file[255];
    char output_dir[4096];
    char output_file[255];
    int n_rows;
    float coll_slicewidth;
    float start_pos;
    float end_pos;
    float pitch_value;          /* table feed per rotation in mm ("pitch" in mm) */
    float slice_thickness;
    float acq_fov;
    float recon_fov;
    int recon_kernel;
    int n_readings;
    float x_origin;
    float y_origin;
    int z_ffs;
    int phi_ffs;
    char scanner[4351];
    int file_type;
    int file_subtype;
    int raw_data_offset;
    unsigned int nx;
    unsigned int ny;
    float tube_start_angle;
    float adaptive_filtration_s;
    int n_slices;
    char table_dir_str[1024];
    int table_dir;
};

#ifdef __cplusplus
extern "C" {
#endif

/* Parse a prm file into *structure (fields not present keep their value).
 * Returns 0 on success, nonzero if the file cannot be opened.
 * Replaces reference src/parse_config.cu:40 (which returns void and crashes on
 * a missing file). */
int ctbb_parse_config(const char *config_file, struct recon_params *structure);

/* Write an empty prm template listing every token (reference
 * include/parse_config.h:69-113).  Returns 0 on success. */
int ctbb_empty_config(const char *filepath);

#ifdef __cplusplus
}
#endif

#endif
