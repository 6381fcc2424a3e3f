// What the encoder needs of a learned MV-proposal network handle (mvnet_api.cpp).
#pragma once
#include "mvnet.hpp"

struct riv_mvnet;

namespace rivnet {
Weights net_weights(const riv_mvnet* n);
// Records, on `st`, that an inference queued there reads the network's weights: every later write
// of the weights (new parameters, a training step, the fragment repack) waits for it on the GPU.
hipError_t note_use(const riv_mvnet* n, hipStream_t st);
int net_device(const riv_mvnet* n);
// logit threshold of a confidence in (0, 1): +inf at >= 1 (never confident), -inf at <= 0
float conf_logit(float confidence);
}  // namespace rivnet
