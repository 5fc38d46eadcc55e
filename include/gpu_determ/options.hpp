// Drop-in for the reference's options.hpp (options.hpp:6-17): the same
// Options POD and load_options / print_options, with additive keys for the
// grid, the method and the run (all optional; defaults reproduce the
// reference's hard-coded setup, determ.cu:47-53).
#ifndef GPU_DETERM_OPTIONS_HPP
#define GPU_DETERM_OPTIONS_HPP

#include <string>

struct Options
{
    int nx;
    int ny;
    int nz;
    int nyz;
    int nxyz;
    int npx;
    // additive keys (gpu.conf), not in the reference
    double vmin = -5.0;          // box [vmin, vmax]^3 (determ.cu:48-49)
    double vmax = 5.0;
    double R = 16.5;             // sphere cut radius (determ.cu:50); <= 0 keeps all
    int method = 0;              // 0 analytic matrix (the reference's), 1 Korobov projection
    int tables = 1;              // 0 FP64 tables, 1 the reference's FP32 rounding
    int correction = 1;          // conservation correction of the loss frequency
    unsigned korobov_M = 1u << 20;
    unsigned long long seed = 1;
    double s0 = 1.0;             // differential cross-section per steradian
    int steps = 1;               // driver: explicit steps (determ.cu:69)
    double courant = 0.5;        // driver: courant number (determ.cu:77)
};

void load_options(Options&, const std::string&);
void print_options(const Options&);

#endif
