// Spatial transport of f over a box of cells (no reference counterpart: the
// step a kinetic solver takes beside the collision operator).
//
// Cells c = (ix * ny + iy) * nz + iz of an nx x ny x nz box, f[c][v] FP64 on
// the handle's velocity grid (cell-major).  One explicit first-order upwind
// step of df/dt + xi . grad_x f = 0, all three axes in one update:
//   f'[c][v] = f[c][v] - sum_a k_a(v) * D_a,   k_a(v) = (dt / h_a) xi_a(v),
//   D_a = f[c] - f[c - e_a]  if k_a > 0,   f[c + e_a] - f[c]  otherwise,
// the axes taken in the order x, y, z.  Boundaries per axis: periodic, or
// zero-gradient (a missing neighbour takes the cell's own value, so D_a = 0
// there).  Stable and positivity-preserving while max_v sum_a |k_a(v)| <= 1,
// which the entry point checks.  With periodic boundaries the flux
// differences telescope: the cell sum of every velocity's f is conserved.
//
// A thread per velocity walks lines of cells along z: the reads of a wave
// (the cell ahead in z, the upwind neighbours in x and y) are contiguous runs
// of f, the z neighbours come from registers, and an x or y neighbour's run
// is read again by the lines beside it, so the HBM traffic stays near f in +
// f out.
#include <algorithm>
#include <cmath>

#include "gdt_internal.h"

#define TR_THREADS 256

// A block takes one work unit (a velocity block of TR_THREADS velocities, a
// line of cells (ix, iy)) and walks the line along z.  A thread keeps its
// velocity's k_a in registers and the z neighbours in a sliding window, so
// per cell it reads f of the cell ahead in z and of its upwind x and y
// neighbours.
__global__ __launch_bounds__(TR_THREADS) void k_transport(int nvb, int n, int nx, int ny, int nz, double cx, double cy,
                                                          double cz, int bcx, int bcy, int bcz,
                                                          const double *__restrict__ xi,
                                                          const double *__restrict__ f, double *__restrict__ out)
{
#pragma clang fp contract(off)
    // work unit u = (velocity block, ix, iy), iy fastest then ix then the
    // velocity block; hardware block b runs on XCD b % 8, so it takes unit
    // (b % 8) * per + b / 8: each XCD walks one contiguous eighth of the unit
    // order, and a line's x and y neighbours were read recently on the same
    // XCD (3.36 -> 3.48 TB/s on 16^3 cells of the 24^3 grid, with the z loop
    // unrolled by 4)
    const long units = (long)nvb * nx * ny, per = (units + 7) / 8;
    const long u = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (u >= units) return;
    const int vb = (int)(u / ((long)nx * ny)), line = (int)(u % ((long)nx * ny));
    const int v = vb * TR_THREADS + threadIdx.x;
    if (v >= n) return;  // no barrier below: threads past the last velocity just leave
    const double kx = cx * xi[3 * v], ky = cy * xi[3 * v + 1], kz = cz * xi[3 * v + 2];
    const size_t N = (size_t)n;
    {
        const int ix = line / ny, iy = line % ny;
        // upwind neighbour lines in x and y (the line itself where a
        // zero-gradient boundary has none)
        int px = kx > 0.0 ? ix - 1 : ix + 1, py = ky > 0.0 ? iy - 1 : iy + 1;
        if (px < 0 || px >= nx) px = bcx == GDT_BC_PERIODIC ? (px < 0 ? nx - 1 : 0) : ix;
        if (py < 0 || py >= ny) py = bcy == GDT_BC_PERIODIC ? (py < 0 ? ny - 1 : 0) : iy;
        const size_t base = (size_t)line * nz, bx = ((size_t)px * ny + iy) * nz, by = ((size_t)ix * ny + py) * nz;
        double cur = f[base * N + v];
        double prev = bcz == GDT_BC_PERIODIC ? f[(base + nz - 1) * N + v] : cur;
#pragma unroll 4
        for (int iz = 0; iz < nz; iz++) {
            const double next = iz + 1 < nz ? f[(base + iz + 1) * N + v]
                                            : (bcz == GDT_BC_PERIODIC ? f[base * N + v] : cur);
            const double fx = f[(bx + iz) * N + v], fy = f[(by + iz) * N + v];
            double acc = cur;
            acc = acc - kx * (kx > 0.0 ? cur - fx : fx - cur);
            acc = acc - ky * (ky > 0.0 ? cur - fy : fy - cur);
            const double fz = kz > 0.0 ? prev : next;
            acc = acc - kz * (kz > 0.0 ? cur - fz : fz - cur);
            out[(base + iz) * N + v] = acc;
            prev = cur;
            cur = next;
        }
    }
}

extern "C" {

int gdt_transport(gdt_handle h, const double *f, double *f_out, int nx, int ny, int nz, const double *spacing,
                  double dt, const int *bc)
{
    GDT_CHECK(h, GDT_E_ARG, "null gdt handle");
    GDT_CHECK(f && f_out && spacing && bc && nx > 0 && ny > 0 && nz > 0 && dt >= 0.0, GDT_E_ARG,
              "gdt_transport: bad arguments");

    double c[3];
    for (int a = 0; a < 3; a++) {
        GDT_CHECK(spacing[a] > 0.0, GDT_E_ARG, "gdt_transport: spacing[%d] = %g", a, spacing[a]);
        GDT_CHECK(bc[a] == GDT_BC_PERIODIC || bc[a] == GDT_BC_ZERO_GRADIENT, GDT_E_ARG,
                  "gdt_transport: boundary %d of axis %d", bc[a], a);
        c[a] = dt / spacing[a];
    }
    // Courant condition: max over velocities of sum_a |k_a| <= 1
    const int n = h->n_pnt;
    double cfl = 0.0;
    for (int v = 0; v < n; v++) {
        double s = 0.0;
        for (int a = 0; a < 3; a++) s += std::fabs(c[a] * h->h_xi[3 * (size_t)v + a]);
        cfl = std::max(cfl, s);
    }
    GDT_CHECK(cfl <= 1.0, GDT_E_ARG, "gdt_transport: Courant number %.6g > 1 (dt too large for the spacing)", cfl);
    const long cells = (long)nx * ny * nz;
    GDT_CHECK(cells <= (1L << 30), GDT_E_ARG, "gdt_transport: %ld cells", cells);
    {
        const uintptr_t a0 = (uintptr_t)f, b0 = (uintptr_t)f_out, len = (uintptr_t)cells * n * sizeof(double);
        GDT_CHECK(a0 + len <= b0 || b0 + len <= a0, GDT_E_ARG,
                  "gdt_transport: f_out overlaps f (the neighbours of every cell are read)");
    }
    GDT_HIP(hipSetDevice(h->device));
    const int vblocks = (n + TR_THREADS - 1) / TR_THREADS;
    const long units = (long)vblocks * nx * ny, grid = (units + 7) / 8 * 8;
    GDT_CHECK(grid <= 0x7FFFFFFFL, GDT_E_ARG, "gdt_transport: %ld work units", units);
    hipLaunchKernelGGL(k_transport, dim3((unsigned)grid), dim3(TR_THREADS), 0, h->stream, vblocks, n, nx, ny, nz, c[0], c[1],
                       c[2], bc[0], bc[1], bc[2], h->xi.as<double>(), f, f_out);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

}  // extern "C"
