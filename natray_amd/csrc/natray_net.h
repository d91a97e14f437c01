// Internal declarations of the warm-start network shared by its kernels (natray_net.hip) and
// the simulation library (natray_sim.cpp). The C ABI is in include/natray.h.
#pragma once
#include <stdint.h>

#include "natray.h"

namespace natray {
namespace net {
// Runs the network forward on `tiles` tiles already in the net's device input buffer
// ([tiles][4096][8] BF16, channels 0..1 used); channel 0 of the last layer's FP32 output
// ([tiles][4096][16]) is the prediction. Returns a hipError_t as int.
int forward_device(natray_net* n, uint32_t tiles, void* stream);
// The net's device input buffer and prediction buffer, and its tile capacity.
void* input_buffer(natray_net* n);
const float* output_buffer(natray_net* n);
uint32_t max_tiles(natray_net* n);
}  // namespace net
}  // namespace natray
