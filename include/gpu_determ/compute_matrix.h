// Drop-in for the reference's compute_matrix.h (compute_matrix.h:13).
#ifndef GPU_DETERM_COMPUTE_MATRIX_H
#define GPU_DETERM_COMPUTE_MATRIX_H

#include "options.hpp"
#include "velocity_grid.hpp"

// Allocates and fills *b (nxyz floats) and *a (nxyz*nxyz/8 floats) in device
// memory, reference layout; the caller frees them with gdt_free.
void init_matrices(const DeviceVelocityGrid&, float** b, float** a, const Options& opts);

#endif
