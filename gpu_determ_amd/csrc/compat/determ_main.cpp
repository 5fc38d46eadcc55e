// The reference's driver (determ.cu:41-186) on the drop-in API: gpu.conf ->
// grid -> initial f -> steps of evolve_colisions + relax_f -> results.out.
// results.out keeps the reference's 9 columns and row selection so that its
// comparator (test.py) applies unchanged: element `index` (< nxyz) of the
// [velocity][cell] arrays for every index whose grid node has i, j, k < n/2
// (determ.cu:146-160) -- i.e. velocity slot index / npx at cell index % npx.
// The moments printed at the end are those of cell 0's collision integral
// over its velocity nodes (the reference's print_moments, determ.cu:123-144,
// mixes cells and velocities), followed by cell 0's entropy H = sum f ln f d3v
// and its rate sum Q ln f d3v (non-positive by the H theorem, up to the
// discretisation), both for the f the last collision integral was evaluated
// on.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "../../../include/gdt.h"
#include "collision_integral.h"
#include "compute_matrix.h"
#include "gputimer.hpp"
#include "options.hpp"
#include "velocity.h"
#include "velocity_grid.hpp"

static int run(int argc, char** argv)
{
    Options opts{};
    opts.nx = opts.ny = opts.nz = 16;
    opts.npx = 256;
    if (argc > 2 && std::string(argv[1]) == "--options") {
        // parse only (no device): nx ny nz nyz nxyz npx and the additive keys
        load_options(opts, argv[2]);
        std::cout << opts.nx << ' ' << opts.ny << ' ' << opts.nz << ' ' << opts.nyz << ' ' << opts.nxyz << ' '
                  << opts.npx << ' ' << opts.vmin << ' ' << opts.vmax << ' ' << opts.R << ' ' << opts.method << ' '
                  << opts.tables << ' ' << opts.correction << ' ' << opts.korobov_M << ' ' << opts.seed << ' '
                  << opts.s0 << ' ' << opts.steps << ' ' << opts.courant << std::endl;
        return 0;
    }
    const std::string conf = argc > 1 ? argv[1] : "gpu.conf";
    const std::string out = argc > 2 ? argv[2] : "results.out";
    load_options(opts, conf);
    print_options(opts);

    VelocityGrid vgrid;
    vgrid.init(Vector3i(opts.nx, opts.ny, opts.nz), Vector3f(opts.vmin, opts.vmin, opts.vmin),
               Vector3f(opts.vmax, opts.vmax, opts.vmax), opts.R);
    const size_t count = (size_t)opts.nxyz * opts.npx;

    // determ.cu:101-121: the same Maxwellian in every cell, [velocity][cell]
    std::vector<float> h_f(count), h_dc(count), h_ic(count), h_flast(count);
    const float u1 = 1.75f, v1 = 0.25f;
    for (int i = 0, index = 0; i < opts.nx; i++)
        for (int j = 0; j < opts.ny; j++)
            for (int k = 0; k < opts.nz; k++, index++) {
                const float u = vgrid.u[i], v = vgrid.v[j], w = vgrid.w[k];
                const float du = u - u1, dv = v - v1;
                const float val = std::exp(-du * du - dv * dv - w * w);
                for (int c = 0; c < opts.npx; c++) h_f[(size_t)index * opts.npx + c] = val;
            }
    void *d_f = nullptr, *d_dc = nullptr, *d_ic = nullptr, *d_time = nullptr;
    gdt_malloc(&d_f, count * sizeof(float));
    gdt_malloc(&d_dc, count * sizeof(float));
    gdt_malloc(&d_ic, count * sizeof(float));
    gdt_malloc(&d_time, sizeof(float));
    gdt_memcpy_htod(d_f, h_f.data(), count * sizeof(float));
    float h_time = 0.0f;
    gdt_memcpy_htod(d_time, &h_time, sizeof(float));

    float *corr = nullptr, *a = nullptr, *b = nullptr;
    init_correction_array(&corr, vgrid.device(), opts);
    init_matrices(vgrid.device(), &b, &a, opts);

    GpuTimer timer;
    timer.start();
    const int out_step = 5;
    for (int step = 0; step < opts.steps; ++step) {
        gdt_memcpy_dtoh(&h_time, d_time, sizeof(float));
        std::cout << "STEP=" << step << ",  TIME=" << h_time << std::endl;
        if (step % out_step == 0) gdt_memcpy_dtoh(h_f.data(), d_f, count * sizeof(float));
        if (step == opts.steps - 1) gdt_memcpy_dtoh(h_flast.data(), d_f, count * sizeof(float));
        evolve_colisions((float*)d_f, (float*)d_dc, (float*)d_ic, b, a, corr, opts);
        relax_f((float*)d_f, (float*)d_dc, (float*)d_ic, (float)opts.courant, (float*)d_time, opts);
    }
    timer.stop();
    std::cout << "COLLISIONS EVOLVED: " << gdt_last_error() << std::endl;
    gdt_memcpy_dtoh(h_dc.data(), d_dc, count * sizeof(float));
    gdt_memcpy_dtoh(h_ic.data(), d_ic, count * sizeof(float));

    std::ofstream fr(out.c_str());
    for (int index = 0; index < opts.nxyz; index++) {
        const int i = index / opts.nyz, j = (index % opts.nyz) / opts.nz, k = index % opts.nz;
        if (i < vgrid.n_u / 2 && j < vgrid.n_v / 2 && k < vgrid.n_w / 2)
            fr << index << ' ' << i << ' ' << j << ' ' << k << ' ' << h_f[index] << ' ' << h_dc[index] << ' '
               << h_ic[index] << ' ' << -h_f[index] * h_dc[index] + h_ic[index] << ' '
               << -h_f[index] * h_dc[index] / h_ic[index] + 1.0 << std::endl;
    }

    double mom[5] = {0, 0, 0, 0, 0};
    for (int i = 0, index = 0; i < opts.nx; i++)
        for (int j = 0; j < opts.ny; j++)
            for (int k = 0; k < opts.nz; k++, index++) {
                const size_t e = (size_t)index * opts.npx;  // cell 0
                const double ci = -(double)h_f[e] * h_dc[e] + h_ic[e];
                const double u = vgrid.u[i], v = vgrid.v[j], w = vgrid.w[k];
                mom[0] += ci * vgrid.d3v;
                mom[1] += ci * u * vgrid.d3v;
                mom[2] += ci * v * vgrid.d3v;
                mom[3] += ci * w * vgrid.d3v;
                mom[4] += ci * (u * u + v * v + w * w) * vgrid.d3v;
            }
    std::printf("CELL 0: M0=%e, MU=%e, MV=%e, MW=%e, M2=%e\n", mom[0], mom[1], mom[2], mom[3], mom[4]);
    if (opts.steps > 0) {
        double H = 0.0, dH = 0.0;
        for (int index = 0; index < opts.nxyz; index++) {
            const size_t e = (size_t)index * opts.npx;  // cell 0
            const double f = h_flast[e];
            if (f > 0.0) {
                const double q = -f * h_dc[e] + h_ic[e];
                H += f * std::log(f) * vgrid.d3v;
                dH += q * std::log(f) * vgrid.d3v;
            }
        }
        std::printf("CELL 0: H=%e, DHDT=%e\n", H, dH);
    }
    std::cout << "COLLISION INTEGRAL EVALUATION TOOK " << timer.elapsed() << " MS." << std::endl;

    gdt_free(d_f);
    gdt_free(d_dc);
    gdt_free(d_ic);
    gdt_free(d_time);
    gdt_free(corr);
    gdt_free(b);
    gdt_free(a);
    std::cout << "END_PROGRAM" << std::endl;
    return 0;
}

int main(int argc, char** argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::cerr << "determ: " << e.what() << std::endl;
        return 1;
    }
}
