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
// A thread per pair of velocities walks lines of cells along z: the reads of
// a wave (the cell ahead in z, the upwind neighbours in x and y) are
// contiguous runs of f, the z neighbours come from registers, and an x or y
// neighbour's run is read again by the lines beside it, so the HBM traffic
// stays near f in + f out.  The output is written with non-temporal stores:
// nothing in the step reads it back, so it does not displace f from L2.
// Measured on 16^3 cells of the 24^3 grid (7,208 velocities): one velocity
// per thread in 256-thread blocks 3.45 TB/s; two velocities per thread
// (16-B accesses) 3.60; non-temporal stores 3.90; 64-thread blocks 4.03.
#include <algorithm>
#include <cmath>

#include "gdt_internal.h"

#define TR_THREADS 64

// A block takes one work unit (a velocity block of TR_THREADS * VEC
// velocities, a line of cells (ix, iy)) and walks the line along z.  A thread
// keeps its VEC velocities' k_a in registers and the z neighbours in a sliding
// window, so per cell it reads f of the cell ahead in z and of its upwind x
// and y neighbours.  VEC = 2 reads and writes f in 16-B pieces (n even and f,
// out 16-B aligned, checked on the host).
template <int VEC>
__device__ __forceinline__ void tr_load(const double *__restrict__ p, double (&a)[VEC])
{
    if constexpr (VEC == 2) {
        const double2 t = *reinterpret_cast<const double2 *>(p);
        a[0] = t.x;
        a[1] = t.y;
    } else {
        a[0] = p[0];
    }
}

template <int VEC>
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
    // XCD (3.36 -> 3.48 TB/s at one velocity per thread)
    const long units = (long)nvb * nx * ny, per = (units + 7) / 8;
    const long u = (long)(blockIdx.x % 8) * per + blockIdx.x / 8;
    if (u >= units) return;
    const int vb = (int)(u / ((long)nx * ny)), line = (int)(u % ((long)nx * ny));
    const int v = (vb * TR_THREADS + threadIdx.x) * VEC;
    if (v >= n) return;  // no barrier below: threads past the last velocity just leave
    const size_t N = (size_t)n;
    const int ix = line / ny, iy = line % ny;
    double kx[VEC], ky[VEC], kz[VEC];
    size_t bx[VEC], by[VEC];
    const size_t base = (size_t)line * nz;
#pragma unroll
    for (int j = 0; j < VEC; j++) {
        kx[j] = cx * xi[3 * (v + j)];
        ky[j] = cy * xi[3 * (v + j) + 1];
        kz[j] = cz * xi[3 * (v + j) + 2];
        // upwind neighbour lines in x and y (the line itself where a
        // zero-gradient boundary has none)
        int px = kx[j] > 0.0 ? ix - 1 : ix + 1, py = ky[j] > 0.0 ? iy - 1 : iy + 1;
        if (px < 0 || px >= nx) px = bcx == GDT_BC_PERIODIC ? (px < 0 ? nx - 1 : 0) : ix;
        if (py < 0 || py >= ny) py = bcy == GDT_BC_PERIODIC ? (py < 0 ? ny - 1 : 0) : iy;
        bx[j] = ((size_t)px * ny + iy) * nz;
        by[j] = ((size_t)ix * ny + py) * nz;
    }
    // the VEC velocities of a thread share their x (y) neighbour line unless
    // a sign of xi_x (xi_y) changes between them
    bool xs = true, ys = true;
#pragma unroll
    for (int j = 1; j < VEC; j++) {
        xs = xs && bx[j] == bx[0];
        ys = ys && by[j] == by[0];
    }
    double cur[VEC], prev[VEC];
    tr_load<VEC>(f + base * N + v, cur);
    if (bcz == GDT_BC_PERIODIC)
        tr_load<VEC>(f + (base + nz - 1) * N + v, prev);
    else
#pragma unroll
        for (int j = 0; j < VEC; j++) prev[j] = cur[j];
#pragma unroll 4
    for (int iz = 0; iz < nz; iz++) {
        double next[VEC], fx[VEC], fy[VEC];
        if (iz + 1 < nz)
            tr_load<VEC>(f + (base + iz + 1) * N + v, next);
        else if (bcz == GDT_BC_PERIODIC)
            tr_load<VEC>(f + base * N + v, next);
        else
#pragma unroll
            for (int j = 0; j < VEC; j++) next[j] = cur[j];
        if (xs)
            tr_load<VEC>(f + (bx[0] + iz) * N + v, fx);
        else
#pragma unroll
            for (int j = 0; j < VEC; j++) fx[j] = f[(bx[j] + iz) * N + v + j];
        if (ys)
            tr_load<VEC>(f + (by[0] + iz) * N + v, fy);
        else
#pragma unroll
            for (int j = 0; j < VEC; j++) fy[j] = f[(by[j] + iz) * N + v + j];
        double acc[VEC];
#pragma unroll
        for (int j = 0; j < VEC; j++) {
            const double c = cur[j];
            double a = c;
            a = a - kx[j] * (kx[j] > 0.0 ? c - fx[j] : fx[j] - c);
            a = a - ky[j] * (ky[j] > 0.0 ? c - fy[j] : fy[j] - c);
            const double fz = kz[j] > 0.0 ? prev[j] : next[j];
            a = a - kz[j] * (kz[j] > 0.0 ? c - fz : fz - c);
            acc[j] = a;
            prev[j] = c;
            cur[j] = next[j];
        }
        double *o = out + (base + iz) * N + v;
        if constexpr (VEC == 2) {
            typedef double d2v __attribute__((ext_vector_type(2)));
            d2v t;
            t.x = acc[0];
            t.y = acc[1];
            __builtin_nontemporal_store(t, reinterpret_cast<d2v *>(o));
        } else
            __builtin_nontemporal_store(acc[0], o);
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
    // 16-B pieces of f when every row starts 16-B aligned
    const bool v2 = n % 2 == 0 && (uintptr_t)f % 16 == 0 && (uintptr_t)f_out % 16 == 0;
    const int per_block = v2 ? 2 * TR_THREADS : TR_THREADS;
    const int vblocks = (n + per_block - 1) / per_block;
    const long units = (long)vblocks * nx * ny, grid = (units + 7) / 8 * 8;
    GDT_CHECK(grid <= 0x7FFFFFFFL, GDT_E_ARG, "gdt_transport: %ld work units", units);
    if (v2)
        hipLaunchKernelGGL(k_transport<2>, dim3((unsigned)grid), dim3(TR_THREADS), 0, h->stream, vblocks, n, nx, ny, nz,
                           c[0], c[1], c[2], bc[0], bc[1], bc[2], h->xi.as<double>(), f, f_out);
    else
        hipLaunchKernelGGL(k_transport<1>, dim3((unsigned)grid), dim3(TR_THREADS), 0, h->stream, vblocks, n, nx, ny, nz,
                           c[0], c[1], c[2], bc[0], bc[1], bc[2], h->xi.as<double>(), f, f_out);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

}  // extern "C"
