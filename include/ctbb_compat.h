/*
 * ctbb_compat.h — the reference program's own entry points, with C++
 * linkage and unchanged signatures, so that a driver written against the
 * reference (its main.cu) links against libctbb_hip.so without changes.
 *
 * Each function wraps the ctbb_* C-ABI (ctbb_hip.h) on a process-wide default
 * context created on the current HIP device, and reproduces the reference's
 * error behaviour: where the reference calls exit(), so does the wrapper,
 * with the same exit code.
 *
 *   reference declaration                                   replaced by
 *   int adaptive_filter_kk(recon_metadata*)  preprocessing.h:26      ctbb_adaptive_filter
 *   int rebin_filter(recon_metadata*)        rebin_filter.h:27       ctbb_rebin_filter
 *   int rebin_filter_cpu(recon_metadata*)    rebin_filter_cpu.h:27   (no CPU path, see below)
 *   int backproject(recon_metadata*)         backproject.h:26        ctbb_backproject
 *   int backproject_cpu(recon_metadata*)     backproject_cpu.h:26    (no CPU path)
 *   int finalize_image_stack(recon_metadata*) finalize_image_stack.h:26     ctbb_finalize_image_stack
 *   int finalize_image_stack_cpu(recon_metadata*) finalize_image_stack_cpu.h:26 (no CPU path)
 *   setup.h:28-34 and parse_config.h:116    ctbb_configure_* / ctbb_update_block_info / ...
 *
 * This build has no CPU reconstruction path: the *_cpu symbols exist so a
 * reference driver links, print a diagnostic and exit(1) when called.
 */
#ifndef CTBB_COMPAT_H
#define CTBB_COMPAT_H

#include "recon_structs.h"

#ifdef __cplusplus
int adaptive_filter_kk(struct recon_metadata *mr);
int rebin_filter(struct recon_metadata *mr);
int rebin_filter_cpu(struct recon_metadata *mr);
int backproject(struct recon_metadata *mr);
int backproject_cpu(struct recon_metadata *mr);
int finalize_image_stack(struct recon_metadata *mr);
int finalize_image_stack_cpu(struct recon_metadata *mr);

int configure_paths(struct recon_metadata *mr);
struct recon_params configure_recon_params(char *filename);
struct ct_geom configure_ct_geom(struct recon_metadata *mr);
void configure_reconstruction(struct recon_metadata *mr);
void update_block_info(recon_metadata *mr);
void extract_projections(struct recon_metadata *mr);
void finish_and_cleanup(struct recon_metadata *mr);
void parse_config(char *config_file, struct recon_params *structure);

/* Options of the default context the wrappers above run on (HU conversion,
 * filter / projection precision).  No counterpart in the reference. */
extern "C" int ctbb_compat_set_options(const struct ctbb_options_s *opt);
#endif

#endif
