// Drop-in for the reference's collision_integral.h (collision_integral.h:
// 11-24).  All arrays are device memory laid out [velocity][cell] (npx cells
// fastest), FP32, as in the reference.
#ifndef GPU_DETERM_COLLISION_INTEGRAL_H
#define GPU_DETERM_COLLISION_INTEGRAL_H

#include "options.hpp"

// direct_integral <- corrected loss frequency nu, inverse_integral <- gain.
void evolve_colisions(float* f, float* direct_integral, float* inverse_integral, float* b, float* a,
                      float* correction_array, const Options& opts);

// f <- f (1 - nu dt) + gain dt with dt = min courant / nu; *time += dt.
void relax_f(float* f, float* frequency, float* inverse_integral, float courant, float* time, const Options& opts);

#endif
