#pragma once

#include <vector>

#include "ctbb_hip.h"
#include "synth_common.h"

namespace ctbb {

std::vector<Ellipsoid> make_phantom(const recon_metadata &mr, const ctbb_phantom_params &pp);
SynthGeom make_synth_geom(const recon_metadata &mr, const ctbb_phantom_params &pp);

}  // namespace ctbb
