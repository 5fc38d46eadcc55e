// Drop-in for the reference's velocity.h (velocity.h:22-41): the moment-basis
// plane accessors and init_correction_array.  The reference's compile-time
// grid sizes (N_X.., velocity.h:15-20) are gone: sizes come from Options.
#ifndef GPU_DETERM_VELOCITY_H
#define GPU_DETERM_VELOCITY_H

#include "common.hpp"
#include "options.hpp"
#include "velocity_grid.hpp"

#define XI_X(v) (v)
#define XI_Y(v) (&((v)[opts.nxyz]))
#define XI_Z(v) (&((v)[2 * opts.nxyz]))
#define XI_2(v) (&((v)[3 * opts.nxyz]))
#define XI_X_X(v) (&((v)[4 * opts.nxyz]))
#define XI_X_Y(v) (&((v)[5 * opts.nxyz]))
#define XI_X_Z(v) (&((v)[6 * opts.nxyz]))
#define XI_X_2(v) (&((v)[7 * opts.nxyz]))
#define XI_Y_Y(v) (&((v)[8 * opts.nxyz]))
#define XI_Y_Z(v) (&((v)[9 * opts.nxyz]))
#define XI_Y_2(v) (&((v)[10 * opts.nxyz]))
#define XI_Z_Z(v) (&((v)[11 * opts.nxyz]))
#define XI_Z_2(v) (&((v)[12 * opts.nxyz]))
#define XI_2_2(v) (&((v)[13 * opts.nxyz]))

#define ERR 1.0e-10

// Allocates 14*nxyz floats in device memory and fills the planes.
void init_correction_array(float** correction_array, const DeviceVelocityGrid&, const Options& opts);

#endif
