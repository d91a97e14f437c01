// Internal declarations of the warm-start network shared by its kernels (natray_net.hip) and
// the simulation library (natray_sim.cpp). The C ABI is in include/natray.h.
#pragma once
#include <stdint.h>

#include "natray.h"

namespace natray {
namespace net {
// Runs the network forward on the tiles already in the net's device input buffer
// ([tiles][4096][8] BF16, channels 0..1 used): min(tiles, max_tiles) of them, or fewer when
// `tiles_dev` (device memory, read by the kernels) holds a smaller count. Channel 0 of the last
// layer's FP32 output ([tiles][4096][16]) is the prediction. Returns a NATRAY_E_* code.
int forward_device(natray_net* n, uint32_t tiles, const uint32_t* tiles_dev, void* stream);
// The net's device input buffer and prediction buffer, and its tile capacity.
void* input_buffer(natray_net* n);
const float* output_buffer(natray_net* n);
uint32_t max_tiles(natray_net* n);
}  // namespace net
}  // namespace natray
