// Entry point of the warm-start CNN shared by the library's translation units.
#ifndef ARTEFACT_CNN_INTERNAL_H
#define ARTEFACT_CNN_INTERNAL_H

#include <hip/hip_runtime.h>

#include "../../include/artefact_hip.h"

// Prediction of `nplanes` device planes of h x w (image units) into device planes `out`, with
// each plane's quantisation table (host values, 64 per plane, natural order), on stream s.
int cnn_predict_planes(artefact_cnn *net, const float *const *planes, const float *quant, uint32_t nplanes,
                       uint32_t h, uint32_t w, float *const *out, hipStream_t s);

#endif
