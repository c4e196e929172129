// Internal declarations shared by the host-side planning code and the HIP
// stage implementations.  Not part of the public C-ABI.
#pragma once

#include <cstddef>
#include <cstdint>

#include "ctbb_hip.h"

namespace ctbb {

// Planning (setup.cpp)
void binary_angles_positions(const recon_params &rp, const ct_geom &cg, float *tube_angles,
                             double *table_positions);
int plan_reconstruction(recon_metadata *mr, bool allocate);
void plan_block(recon_metadata *mr, int block_idx);

// pi constants as the reference spells them (float literals):
// backprojection / preprocessing use 3.1415368979f (reference
// include/backproject.cuh:23, include/preprocessing.cuh:25); rebinning and
// setup use 3.1415926535897f (include/rebin_filter.cuh:23, src/setup.cu:33).
constexpr float kPiBp = 3.1415368979f;
constexpr float kPiTrue = 3.1415926535897f;

// Per-view table entry of the backprojection (FP64-derived, stored FP32),
// in the order the views are processed: entry t = i * n_half_turns + k for
// angle i and half turn k (reference include/backproject.cuh:70-109 loops).
struct BpView {
    float sin_t;    // sin(theta_v)
    float cos_t;    // cos(theta_v)
    float c_v;      // z_rot/2 + z_rot*(tube_start - theta_v)/(2 pi'): z - ray_pos at slice 0, column term excluded
    int32_t proj;   // rebinned projection index of the view, i + k * (n_proj_turn / 2)
};

// Host-computed rebinning maps of one block (FP64 geometry rounded to FP32).
struct RebinMaps {
    int kind;          // 0 nffs, 1 zffs (2D single pass), 2 pffs, 3 affs (two pass)
    int n_sets;        // 2D: focal spots (1 or 2); two pass: pairs (1 or 2)
    int map_len;       // entries per set in aoff (nos for two-pass, nos for 2D)
    // bidx[set][c]  beta index of output channel c (raw channels for 2D kinds,
    //               oversampled channels for two-pass kinds)
    // aoff[set][m]  alpha offset: sheet projection index = proj - aoff
    //               (2D kinds: indexed by output channel c; two-pass kinds:
    //               indexed by oversampled channel m of the azimuthal pass)
    float *bidx;
    float *aoff;
};

}  // namespace ctbb
