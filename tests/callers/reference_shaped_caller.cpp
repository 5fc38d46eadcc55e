// A host program shaped like the reference's driver (determ.cu:7-169), built
// against the drop-in headers (include/gpu_determ) and libraries only.  It
// keeps the reference's program structure and call sequence: file-scope
// Options with a const reference, host/device arrays as raw pointers, a
// VelocityGrid built on [-5,5]^3 with R = 16.5, init_f / init_integrals
// helpers, init_correction_array then init_matrices on vgrid.device(), a
// GpuTimer around one step of evolve_colisions + relax_f (courant 0.5), the
// 9-column results.out over i, j, k < n/2 and the moment line.
//
// What differs from determ.cu, and nothing else:
//   * <cuda.h> -> "gdt.h"; cudaMalloc / cudaMemcpy / cudaFree ->
//     gdt_malloc / gdt_memcpy_htod / gdt_memcpy_dtoh / gdt_free
//     (determ.cu:9,65-66,71,74,79,88-89,119-120,167-168);
//   * cudaThreadSynchronize -> gdt_device_synchronize and
//     cudaGetErrorString(cudaGetLastError()) -> gdt_last_error()
//     (determ.cu:82,86,98);
//   * the compile-time N_Z of the reference's velocity.h -> opts.nz
//     (determ.cu:153-154): grid sizes come from gpu.conf;
//   * gpu.conf is read from argv[1] when given (default "gpu.conf", as the
//     reference) and results.out written to argv[2] when given.
#include <cassert>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>

#include "gdt.h"
#include "velocity.h"
#include "compute_matrix.h"
#include "collision_integral.h"
#include "options.hpp"
#include "velocity_grid.hpp"
#include "gputimer.hpp"

static void release_all();
static void fill_initial_f();
static void alloc_integrals();
static void write_results(const std::string& path);
static void report_moments();

Options options;
const Options& opts = options;

float* h_f;
float* d_f;
float* h_gain;
float* d_gain;
float* h_freq;
float* d_freq;
float* a;
float* b;
float* correction_array;

VelocityGrid vgrid;

int main(int argc, char** argv)
{
    const std::string conf = argc > 1 ? argv[1] : "gpu.conf";
    const std::string out = argc > 2 ? argv[2] : "results.out";
    load_options(options, conf);
    print_options(options);

    const Vector3i npts(opts.nx, opts.ny, opts.nz);
    const Vector3f lo(-5.0, -5.0, -5.0);
    const Vector3f hi(5.0, 5.0, 5.0);
    const float radius = 16.5;  // the reference's cut sphere (keeps the whole box)

    vgrid.init(npts, lo, hi, radius);
    fill_initial_f();
    alloc_integrals();
    init_correction_array(&correction_array, vgrid.device(), opts);
    init_matrices(vgrid.device(), &b, &a, opts);

    GpuTimer timer;
    timer.start();

    float h_time = 0;
    float* d_time;
    gdt_malloc((void**)&d_time, sizeof(float));
    gdt_memcpy_htod(d_time, &h_time, sizeof(float));
    const int every = 5;
    for (int step = 0; step <= 0; ++step) {
        gdt_memcpy_dtoh(&h_time, d_time, sizeof(float));
        std::cout << "STEP=" << step << ",  TIME=" << h_time << std::endl;
        if (step % every == 0) gdt_memcpy_dtoh(h_f, d_f, (size_t)opts.nxyz * opts.npx * sizeof(float));
        evolve_colisions(d_f, d_freq, d_gain, b, a, correction_array, opts);
        relax_f(d_f, d_freq, d_gain, 0.5, d_time, opts);
    }
    gdt_free(d_time);

    gdt_device_synchronize();
    timer.stop();
    std::cout << "COLLISIONS EVOLVED: " << gdt_last_error() << std::endl;

    gdt_memcpy_dtoh(h_freq, d_freq, (size_t)opts.nxyz * sizeof(float));
    gdt_memcpy_dtoh(h_gain, d_gain, (size_t)opts.nxyz * sizeof(float));

    write_results(out);
    report_moments();
    std::cout << "COLLISION INTEGRAL EVALUATION TOOK " << timer.elapsed() << " MS." << std::endl;

    release_all();
    std::cout << "END_PROGRAM: " << gdt_last_error() << std::endl;
    return 0;
}

// determ.cu:101-121: one shifted Maxwellian, the same in every cell.
static void fill_initial_f()
{
    const size_t count = (size_t)opts.nxyz * opts.npx;
    h_f = new float[count];
    const float ux = 1.75, uy = 0.25;
    int slot = 0;
    for (int i = 0; i < opts.nx; i++)
        for (int j = 0; j < opts.ny; j++)
            for (int k = 0; k < opts.nz; k++, slot++) {
                const float du = vgrid.u[i] - ux, dv = vgrid.v[j] - uy, w = vgrid.w[k];
                for (int c = 0; c < opts.npx; c++) h_f[(size_t)slot * opts.npx + c] = exp(-du * du - dv * dv - w * w);
            }
    gdt_malloc((void**)&d_f, count * sizeof(float));
    gdt_memcpy_htod(d_f, h_f, count * sizeof(float));
}

// determ.cu:123-144: moments of the first nxyz entries of the integral, as
// the reference sums them.
static void report_moments()
{
    float m0 = 0, mu = 0, mv = 0, mw = 0, m2 = 0;
    int slot = 0;
    for (int i = 0; i < opts.nx; i++)
        for (int j = 0; j < opts.ny; j++)
            for (int k = 0; k < opts.nz; k++, slot++) {
                const float q = -h_f[slot] * h_freq[slot] + h_gain[slot];
                const float u = vgrid.u[i], v = vgrid.v[j], w = vgrid.w[k];
                m0 += q * vgrid.d3v;
                mu += q * u * vgrid.d3v;
                mv += q * v * vgrid.d3v;
                mw += q * w * vgrid.d3v;
                m2 += q * (u * u + v * v + w * w) * vgrid.d3v;
            }
    printf("M0=%e, MU=%e, MV=%e, MW=%e, M2=%e\n", m0, mu, mv, mw, m2);
}

// determ.cu:146-160: the 9 columns for every element index < nxyz whose
// i, j, k indices lie in the lower half of each axis.
static void write_results(const std::string& path)
{
    std::ofstream fr(path.c_str());
    for (int e = 0; e < opts.nxyz; e++)
        if (vgrid.u_index[e] < vgrid.n_u / 2 && vgrid.v_index[e] < vgrid.n_v / 2 && vgrid.w_index[e] < vgrid.n_w / 2)
            fr << e << ' ' << e / opts.nyz << ' ' << (e % opts.nyz) / opts.nz << ' ' << (e % opts.nyz) % opts.nz
               << ' ' << h_f[e] << ' ' << h_freq[e] << ' ' << h_gain[e] << ' ' << -h_f[e] * h_freq[e] + h_gain[e]
               << ' ' << -h_f[e] * h_freq[e] / h_gain[e] + 1.0 << std::endl;
}

// determ.cu:162-169
static void alloc_integrals()
{
    h_gain = new float[(size_t)opts.nxyz * opts.npx];
    h_freq = new float[(size_t)opts.nxyz * opts.npx];
    gdt_malloc((void**)&d_freq, (size_t)opts.nxyz * opts.npx * sizeof(float));
    gdt_malloc((void**)&d_gain, (size_t)opts.nxyz * opts.npx * sizeof(float));
}

static void release_all()
{
    gdt_free(d_f);
    gdt_free(d_freq);
    gdt_free(d_gain);
    gdt_free(a);
    gdt_free(b);
    gdt_free(correction_array);
    delete[] h_f;
    delete[] h_freq;
    delete[] h_gain;
}
