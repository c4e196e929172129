// FP64 host geometry for the kernels (see geometry.cpp).
#pragma once

#include <vector>

#include "ctbb_internal.h"

namespace ctbb {

void ffs_offsets(const recon_metadata &mr, double *da, double *dr);
float bp_tube_start(const recon_metadata &mr);
int bp_n_half_turns(const recon_metadata &mr);
// kind: 0 nffs, 1 zffs, 2 pffs, 3 affs; maps laid out [set][n_channels_oversampled]
int rebin_maps(const recon_metadata &mr, std::vector<float> &bidx, std::vector<float> &aoff, int *kind,
               int *n_sets);

}  // namespace ctbb
