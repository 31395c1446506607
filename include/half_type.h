// half_type.h -- the 16-bit IEEE float used by the host classes.
//
// The reference's classes hold `half` from cuda_fp16.h (Matrix.cuh:3,14).  Here
// `half` is the compiler's native binary16 (_Float16 under hipcc/amdclang);
// float<->half conversion rounds to nearest even, as __float2half does.
#pragma once

typedef _Float16 half;

inline half float2half_rn(float f) { return static_cast<half>(f); }
inline float half2float(half h) { return static_cast<float>(h); }
