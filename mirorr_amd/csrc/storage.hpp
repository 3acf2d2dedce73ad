// Host-side encoding of float images into binary16 storage.
#pragma once

#include "common.hpp"

namespace mirorr {

// v -> (v - offset) * scale, computed in float and rounded to binary16 to
// nearest even: offset the nearest integer to the middle of the finite range,
// scale the power of two that brings the largest of the half-range and
// |offset| into (1024, 2048]. Integer intensities of up to 12 bits (and
// their offset) are then exact in binary16, and so is the encoding of
// intensity 0, which resampling writes outside the field of view.
struct HalfEncoding {
  double offset = 0.0;
  double scale = 1.0;
  float zero() const { return (float)(half_t)(float)(-offset * scale); }
};

HalfEncoding half_encoding(const float* data, size_t n);
void encode_half(const float* data, size_t n, const HalfEncoding& e, half_t* out);

}  // namespace mirorr
