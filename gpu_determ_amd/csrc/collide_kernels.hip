// Collision-operator kernels for gfx950 (wave64).
//
// Layouts: velocity arrays are [cell][node]; "box" arrays cover the whole
// n_u x n_v x n_w box with zeros outside the sphere cut, "cmp" arrays hold the
// kept nodes only.  All arithmetic is FP64.
#include <algorithm>
#include <cstdio>
#include <vector>

#include "gdt_internal.h"
#include "kernel_common.h"

// ---------------------------------------------------------------------------
// layout helpers
// ---------------------------------------------------------------------------

__global__ void k_cmp_to_box(int n_box, int n_pnt, int cells, int layout, const int *__restrict__ box2c,
                             const double *__restrict__ src, double *__restrict__ dst)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_box * cells) return;
    const int c = (int)(t / n_box), b = (int)(t % n_box);
    const int v = box2c[b];
    double x = 0.0;
    if (v >= 0) x = layout == GDT_LAYOUT_NODE_MAJOR ? src[(long)v * cells + c] : src[(long)c * n_pnt + v];
    dst[t] = x;
}

__global__ void k_box_to_cmp(int n_box, int n_pnt, int cells, int layout, const int *__restrict__ cmp2box,
                             const double *__restrict__ src, double *__restrict__ dst)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)n_pnt * cells) return;
    const int c = (int)(t / n_pnt), v = (int)(t % n_pnt);
    const double x = src[(long)c * n_box + cmp2box[v]];
    if (layout == GDT_LAYOUT_NODE_MAJOR)
        dst[(long)v * cells + c] = x;
    else
        dst[t] = x;
}

int launch_cmp_to_box(gdt_context *h, const double *f_cmp, double *f_box, int cells, int layout)
{
    const long tot = (long)h->n_box * cells;
    hipLaunchKernelGGL(k_cmp_to_box, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->n_box, h->n_pnt,
                       cells, layout, h->box2c.as<int>(), f_cmp, f_box);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

int launch_box_to_cmp(gdt_context *h, const double *f_box, double *f_cmp, int cells, int layout)
{
    const long tot = (long)h->n_pnt * cells;
    hipLaunchKernelGGL(k_box_to_cmp, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, h->stream, h->n_box, h->n_pnt,
                       cells, layout, h->cmp2box.as<int>(), f_box, f_cmp);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// loss frequency: 3-D Toeplitz correlation with the b table
// ---------------------------------------------------------------------------
// nu(i,j,k) = sum_{i1,j1,k1} b(|i-i1|,|j-j1|,|k-k1|) f(i1,j1,k1)
// (evolve_direct_colisions, collision_integral.cu:32-72).  One lane owns a
// destination row (i, j, 0..NW-1) of one cell in registers; per partner row
// (i1, j1) it loads the b row of (|i-i1|, |j-j1|) and the (wave-uniform) f
// row and performs NW*NW FMAs, so the loop is FMA bound, not load bound.
template <int NW>
__global__ __launch_bounds__(64) void k_loss(int nu, int nv, const double *__restrict__ b,
                                             const double *__restrict__ f, double *__restrict__ out)
{
    const int c = blockIdx.y;
    const int rows = nu * nv;
    const int row = blockIdx.x * 64 + threadIdx.x;
    const bool act = row < rows;
    const int i = act ? row / nv : 0, j = act ? row % nv : 0;
    const double *fc = f + (size_t)c * rows * NW;
    double acc[NW];
#pragma unroll
    for (int k = 0; k < NW; k++) acc[k] = 0.0;
    for (int i1 = 0; i1 < nu; i1++) {
        const int gi = i > i1 ? i - i1 : i1 - i;
        for (int j1 = 0; j1 < nv; j1++) {
            const int gj = j > j1 ? j - j1 : j1 - j;
            const double *brow = b + (gi * nv + gj) * NW;
            const double *frow = fc + (i1 * nv + j1) * NW;
            double bb[NW], ff[NW];
#pragma unroll
            for (int k = 0; k < NW; k++) {
                bb[k] = brow[k];
                ff[k] = frow[k];
            }
#pragma unroll
            for (int k = 0; k < NW; k++)
#pragma unroll
                for (int k1 = 0; k1 < NW; k1++) acc[k] += bb[k > k1 ? k - k1 : k1 - k] * ff[k1];
        }
    }
    if (act) {
        double *o = out + (size_t)c * rows * NW + (size_t)row * NW;
#pragma unroll
        for (int k = 0; k < NW; k++) o[k] = acc[k];
    }
}

// Generic-width fallback (any n_w): the same sum, one destination per lane.
__global__ __launch_bounds__(256) void k_loss_generic(int nu, int nv, int nw, const double *__restrict__ b,
                                                      const double *__restrict__ f, double *__restrict__ out)
{
    const int c = blockIdx.y;
    const int nb = nu * nv * nw;
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nb) return;
    const int i = v / (nv * nw), j = (v / nw) % nv, k = v % nw;
    const double *fc = f + (size_t)c * nb;
    double s = 0.0;
    for (int i1 = 0; i1 < nu; i1++)
        for (int j1 = 0; j1 < nv; j1++) {
            const double *brow = b + ((i > i1 ? i - i1 : i1 - i) * nv + (j > j1 ? j - j1 : j1 - j)) * nw;
            const double *frow = fc + (i1 * nv + j1) * nw;
            for (int k1 = 0; k1 < nw; k1++) s += brow[k > k1 ? k - k1 : k1 - k] * frow[k1];
        }
    out[(size_t)c * nb + v] = s;
}

int launch_loss(gdt_context *h, const double *f_box, double *nu_box, int cells)
{
    GDT_CHECK_GRID_Y(cells, "loss: cells");
    const int nu = h->n[0], nv = h->n[1], nw = h->n[2];
    const dim3 grid((nu * nv + 63) / 64, cells);
    const double *b = h->tb.as<double>();
    KernelTimer kt(h, "loss");
    switch (nw) {
    case 8: hipLaunchKernelGGL(k_loss<8>, grid, dim3(64), 0, h->stream, nu, nv, b, f_box, nu_box); break;
    case 16: hipLaunchKernelGGL(k_loss<16>, grid, dim3(64), 0, h->stream, nu, nv, b, f_box, nu_box); break;
    case 20: hipLaunchKernelGGL(k_loss<20>, grid, dim3(64), 0, h->stream, nu, nv, b, f_box, nu_box); break;
    case 24: hipLaunchKernelGGL(k_loss<24>, grid, dim3(64), 0, h->stream, nu, nv, b, f_box, nu_box); break;
    default:
        hipLaunchKernelGGL(k_loss_generic, dim3((h->n_box + 255) / 256, cells), dim3(256), 0, h->stream, nu, nv, nw, b,
                           f_box, nu_box);
    }
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// gain of the analytic matrix method
// ---------------------------------------------------------------------------
// gain(v) = sum_{v1} sum_{boxes e of g = |v - v1|} a_e t4(v, v1, e), with the
// four sign terms t4 of evolve_inverse_colisions (collision_integral.cu:
// 113-162) and the intended output slot.  f of one cell sits in LDS.  A wave
// owns a 4x4x4 block of destinations; it walks the relative-velocity triples
// g (wave-uniform), the 8 partner signs v1 = v +- g per axis, and the g's
// nonzero boxes (scalar loads).  Sign choices with no partner inside the grid
// for any lane of the block are skipped as a whole.  Blocks of 8 waves share
// one LDS copy of f: f takes 32-64 KB at 16^3-20^3, so 4-wave blocks left 2-4
// waves per SIMD to hide each entry's scalar loads (8 waves: 414 -> 509
// evaluations per second at 16^3, 49 -> 70 at 20^3; 16 waves: 496, 70).
// Rows of f in LDS are padded to gain_row(nw) = nw rounded up to 4 mod 8
// doubles: the 16 lanes of a wave that share i read 4 rows j of 4 k each, and
// with a row stride of 4 mod 8 doubles (8 or 24 banks of 4 bytes mod 32) the
// four rows fall on disjoint banks.
__host__ __device__ constexpr int gain_row(int nw) { return nw + ((12 - nw % 8) % 8); }

__global__ __launch_bounds__(512) void k_gain(int nu, int nv, int nw, const int *__restrict__ gptr,
                                              const int *__restrict__ gbox, const double *__restrict__ gval,
                                              const double *__restrict__ f, double *__restrict__ out)
{
    extern __shared__ double fs[];
    const int nb = nu * nv * nw, rw = gain_row(nw), nyz = nv * rw;
    const int c = blockIdx.y;
    const double *fc = f + (size_t)c * nb;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) fs[(t / nw) * rw + t % nw] = fc[t];
    __syncthreads();

    const int bu = (nu + 3) / 4, bv = (nv + 3) / 4, bw = (nw + 3) / 4;
    const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
    const int blk = blockIdx.x * 8 + wave;
    if (blk >= bu * bv * bw) return;  // whole wave exits; no barrier follows
    const int bi = blk / (bv * bw), bj = (blk / bw) % bv, bk = blk % bw;
    const int i = bi * 4 + lane / 16, j = bj * 4 + (lane / 4) % 4, k = bk * 4 + lane % 4;
    const bool act = i < nu && j < nv && k < nw;
    double acc = 0.0;
    for (int gi = 0; gi < nu; gi++)
        for (int gj = 0; gj < nv; gj++)
            for (int gk = 0; gk < nw; gk++) {
                const int g = (gi * nv + gj) * nw + gk;
                const int e0 = gptr[g], e1 = gptr[g + 1];
                if (e0 == e1) continue;
                for (int sgn = 0; sgn < 8; sgn++) {
                    if (((sgn & 1) && gi == 0) || ((sgn & 2) && gj == 0) || ((sgn & 4) && gk == 0)) continue;
                    const int i1 = (sgn & 1) ? i + gi : i - gi;
                    const int j1 = (sgn & 2) ? j + gj : j - gj;
                    const int k1 = (sgn & 4) ? k + gk : k - gk;
                    const bool ok = act && i1 >= 0 && i1 < nu && j1 >= 0 && j1 < nv && k1 >= 0 && k1 < nw;
                    if (!__any(ok)) continue;
                    // The pair of a box offset (in, jn, kn) is ih = (SI + in) / 2,
                    // il = (SI - in) / 2 per axis: SI = 2 i +- gi and in carry
                    // the parity of gi, so both halves are exact.  Hence
                    //   ih < nu and il >= 0  <=>  in <= min(SI, 2 nu - 1 - SI)
                    // (one bound per lane and axis; -1 folds in a partner off
                    // the grid), and the LDS index of (ih or il, jh or jl,
                    // kh or kl) is (P +- in nyz +- jn rw +- kn) / 2 with
                    // P = SI nyz + SJ rw + SK: a lane value plus a
                    // wave-uniform offset.
                    const int SI = i + i1, SJ = j + j1, SK = k + k1;
                    const int LI = ok ? min(SI, 2 * nu - 1 - SI) : -1;
                    const int LJ = min(SJ, 2 * nv - 1 - SJ), LK = min(SK, 2 * nw - 1 - SK);
                    const int P = SI * nyz + SJ * rw + SK;
                    // unrolled by 4: the scalar loads of four entries go out
                    // together (508 -> 535 evaluations per second at 16^3,
                    // 70 -> 77 at 20^3; by 2: 529, 73)
#pragma unroll 4
                    for (int e = e0; e < e1; e++) {
                        const int box = gbox[e];
                        const double val = gval[e];
                        const int in = box & 1023, jn = (box >> 10) & 1023, kn = box >> 20;
                        if (in <= LI && jn <= LJ && kn <= LK) {
                            const int qi = in * nyz, qj = jn * rw;
                            const int hhh = (P + qi + qj + kn) >> 1, lll = (P - qi - qj - kn) >> 1;
                            const int hll = (P + qi - qj - kn) >> 1, lhh = (P - qi + qj + kn) >> 1;
                            const int hhl = (P + qi + qj - kn) >> 1, llh = (P - qi - qj + kn) >> 1;
                            const int hlh = (P + qi - qj + kn) >> 1, lhl = (P - qi + qj - kn) >> 1;
                            const double t4 = fs[hhh] * fs[lll] + fs[hll] * fs[lhh] + fs[hhl] * fs[llh] +
                                              fs[hlh] * fs[lhl];
                            acc += val * t4;
                        }
                    }
                }
            }
    if (act) out[(size_t)c * nb + (i * nv + j) * nw + k] = acc;
}

// Enables k_gain for the LDS copy of f of the context's box (above 64 KiB
// the kernel must be enabled for its dynamic LDS).  Called when the tables
// are built, outside any stream capture; launch_gain checks it again.
int gain_prepare(gdt_context *h)
{
    const size_t lds = sizeof(double) * h->n[0] * h->n[1] * gain_row(h->n[2]);
    if (lds > 160 * 1024 || h->gain_lds_attr >= (int)lds) return GDT_OK;  // launch_gain rejects the former
    GDT_HIP(hipFuncSetAttribute((const void *)k_gain, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    h->gain_lds_attr = (int)lds;
    return GDT_OK;
}

int launch_gain(gdt_context *h, const double *f_box, double *gain_box, int cells)
{
    GDT_CHECK_GRID_Y(cells, "gain: cells");
    const int nu = h->n[0], nv = h->n[1], nw = h->n[2];
    const int blocks = ((nu + 3) / 4) * ((nv + 3) / 4) * ((nw + 3) / 4);
    const size_t lds = sizeof(double) * nu * nv * gain_row(nw);
    GDT_CHECK(lds <= 160 * 1024, GDT_E_ARG, "gain: grid of %d box nodes exceeds the LDS budget", h->n_box);
    int rc = gain_prepare(h);
    if (rc) return rc;
    KernelTimer kt(h, "gain");
    hipLaunchKernelGGL(k_gain, dim3((blocks + 7) / 8, cells), dim3(512), lds, h->stream, nu, nv, nw,
                       h->gptr.as<int>(), h->gbox.as<int>(), h->gval.as<double>(), f_box, gain_box);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// conservation correction (+ optional explicit step), one block per cell
// ---------------------------------------------------------------------------
// With psi = (1, xi_x, xi_y, xi_z, |xi|^2) and L = f nu (MATRIX) or the loss
// term itself (KOROBOV):
//   B_m = sum (Q0) psi_m  with Q0 = gain - L (MATRIX) or the raw Q (KOROBOV)
//   A_mn = sum L psi_m psi_n,  A X = B (partial pivoting),
//   MATRIX:  nu_c = nu (1 + X.psi),  Q = gain - f nu_c
//   KOROBOV: Q = Q0 - L X.psi,       nu_c = (L/f)(1 + X.psi)
// (before_correction, linear_solver, correct_frequency,
// collision_integral.cu:217-526, solved for every cell).  With dt != 0 the
// block also applies f <- f + dt Q (evolve_f, :809-833).
// MODE 0: matrix, corrected; 1: Korobov, corrected; 2: matrix, X = 0;
// 3: Korobov, X = 0.
template <int MODE>
__global__ __launch_bounds__(256) void k_correct(int npts, const double *__restrict__ xi, const double *f,
                                                 const double *__restrict__ nu, const double *__restrict__ gain,
                                                 double *__restrict__ nu_out, double *__restrict__ Q,
                                                 double *__restrict__ Xout, double dt, double *f_upd)
{
    constexpr bool KOR = (MODE == 1 || MODE == 3);
    constexpr bool SOLVE = (MODE == 0 || MODE == 1);
    __shared__ double part[4][20];
    __shared__ double Xs[5];
    const int c = blockIdx.x;
    const size_t o = (size_t)c * npts;
    if (SOLVE) {
        double m[20];
#pragma unroll
        for (int q = 0; q < 20; q++) m[q] = 0.0;
        for (int v = threadIdx.x; v < npts; v += blockDim.x) {
            const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
            const double e = x * x + y * y + z * z;
            const double psi[5] = {1.0, x, y, z, e};
            double L, R;
            if (KOR) {
                L = nu[o + v];
                R = gain[o + v];
            } else {
                L = nu[o + v] * f[o + v];
                R = gain[o + v] - L;
            }
#pragma unroll
            for (int a = 0; a < 5; a++) m[a] += R * psi[a];
            int q = 5;
#pragma unroll
            for (int a = 0; a < 5; a++)
#pragma unroll
                for (int b = a; b < 5; b++) m[q++] += L * psi[a] * psi[b];
        }
        const int wave = threadIdx.x / WAVE, lane = threadIdx.x % WAVE;
        wave_sum_n<20>(m);
        if (lane == 0)
#pragma unroll
            for (int q = 0; q < 20; q++) part[wave][q] = m[q];
        __syncthreads();
        if (threadIdx.x == 0) {
            double sm[20], X[5];
            for (int q = 0; q < 20; q++) sm[q] = part[0][q] + part[1][q] + part[2][q] + part[3][q];
            moments_solve(sm, X);
            for (int a = 0; a < 5; a++) Xs[a] = X[a];
        }
        __syncthreads();
    } else if (threadIdx.x < 5) {
        Xs[threadIdx.x] = 0.0;
    }
    __syncthreads();
    if (Xout && threadIdx.x < 5) Xout[5 * c + threadIdx.x] = Xs[threadIdx.x];
    const double X0 = Xs[0], X1 = Xs[1], X2 = Xs[2], X3 = Xs[3], X4 = Xs[4];
    for (int v = threadIdx.x; v < npts; v += blockDim.x) {
        const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
        const double e = x * x + y * y + z * z;
        const double fv = f[o + v];
        double q;
        if (KOR) {
            const double s = X0 + X1 * x + X2 * y + X3 * z + X4 * e;
            const double L = nu[o + v];
            q = gain[o + v] - L * s;
            if (nu_out) nu_out[o + v] = fv > 0.0 ? (L / fv) * (1.0 + s) : 0.0;
        } else {
            const double s = 1.0 + X0 + X1 * x + X2 * y + X3 * z + X4 * e;
            const double nc = nu[o + v] * s;
            q = gain[o + v] - fv * nc;
            if (nu_out) nu_out[o + v] = nc;
        }
        Q[o + v] = q;
        if (dt != 0.0) f_upd[o + v] = fv + dt * q;
    }
}

// gain = Q + f nu (the gain of the Korobov form after correction).
__global__ void k_compose_gain(long count, const double *__restrict__ f, const double *__restrict__ nu,
                               const double *__restrict__ Q, double *__restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) out[t] = Q[t] + f[t] * nu[t];
}

// method: GDT_METHOD_MATRIX / GDT_METHOD_KOROBOV = corrected; -1 / -2 = the
// same without correction (X = 0); -3 = gain composition out = Q + f nu with
// (nu, gain) = (nu, Q) and out = Q argument.
int launch_correct(gdt_context *h, int method, const double *f, const double *nu, const double *gain, double *nu_out,
                   double *Q, double *X, int cells, double dt, double *f_upd)
{
    if (method == -3) {
        const long count = (long)cells * h->n_pnt;
        hipLaunchKernelGGL(k_compose_gain, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, count, f, nu,
                           gain, Q);
        GDT_HIP(hipGetLastError());
        return GDT_OK;
    }
    KernelTimer kt(h, "correct");
    const dim3 g(cells), b(256);
    const double *xi = h->xi.as<double>();
    const int np = h->n_pnt;
    switch (method) {
    case GDT_METHOD_MATRIX:
        hipLaunchKernelGGL(k_correct<0>, g, b, 0, h->stream, np, xi, f, nu, gain, nu_out, Q, X, dt, f_upd);
        break;
    case GDT_METHOD_KOROBOV:
        hipLaunchKernelGGL(k_correct<1>, g, b, 0, h->stream, np, xi, f, nu, gain, nu_out, Q, X, dt, f_upd);
        break;
    case -1:
        hipLaunchKernelGGL(k_correct<2>, g, b, 0, h->stream, np, xi, f, nu, gain, nu_out, Q, X, dt, f_upd);
        break;
    case -2:
        hipLaunchKernelGGL(k_correct<3>, g, b, 0, h->stream, np, xi, f, nu, gain, nu_out, Q, X, dt, f_upd);
        break;
    default: gdt_set_error(GDT_E_ARG, "launch_correct: method %d", method); return GDT_E_ARG;
    }
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// explicit step helpers (relax_f, collision_integral.cu:776-880)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(1024) void k_courant_dt(const double *__restrict__ nu, long count, double courant,
                                                     double *__restrict__ dt)
{
    __shared__ double red[1024 / WAVE];
    double m = 1.0e10;
    for (long t = threadIdx.x; t < count; t += blockDim.x) {
        const double v = nu[t];
        if (v > 0.0) {
            const double d = courant / v;
            m = d < m ? d : m;
        }
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        const double o = __shfl_xor(m, off, WAVE);
        m = o < m ? o : m;
    }
    if (threadIdx.x % WAVE == 0) red[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = red[0];
        for (int w = 1; w < (int)(blockDim.x / WAVE); w++) r = red[w] < r ? red[w] : r;
        dt[0] = r;
    }
}

int launch_courant_dt(gdt_context *h, const double *nu, long count, double courant, double *dt_out)
{
    hipLaunchKernelGGL(k_courant_dt, dim3(1), dim3(1024), 0, h->stream, nu, count, courant, dt_out);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// f <- f (1 - nu dt) + gain dt (evolve_f, collision_integral.cu:809-833),
// dt from the device (courant) or the argument; time += dt.
__global__ void k_euler(double *__restrict__ f, const double *__restrict__ nu, const double *__restrict__ gain,
                        long count, double dt, const double *__restrict__ dt_dev, double *__restrict__ time)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const double d = dt_dev ? dt_dev[0] : dt;
    if (t < count) f[t] = f[t] * (1.0 - nu[t] * d) + gain[t] * d;
    if (t == 0 && time) time[0] += d;
}

int launch_euler(gdt_context *h, double *f, const double *nu, const double *gain, long count, double dt,
                 const double *dt_dev, double *time)
{
    hipLaunchKernelGGL(k_euler, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, f, nu, gain, count, dt,
                       dt_dev, time);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

__global__ void k_convert(const float *__restrict__ s32, double *__restrict__ d64, long count,
                          const double *__restrict__ s64, float *__restrict__ d32)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    if (s32)
        d64[t] = (double)s32[t];
    else
        d32[t] = (float)s64[t];
}

int launch_convert(const float *s32, double *d64, size_t count, const double *s64, float *d32)
{
    if (count == 0) return GDT_OK;
    hipLaunchKernelGGL(k_convert, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, 0, s32, d64, (long)count, s64,
                       d32);
    GDT_HIP(hipGetLastError());
    GDT_HIP(hipDeviceSynchronize());
    return GDT_OK;
}

// [rows][cols] -> [cols][rows] through a padded LDS tile (transpose_float,
// collision_integral.cu:528-547).
__global__ __launch_bounds__(256) void k_transpose(const double *__restrict__ src, double *__restrict__ dst, int rows,
                                                   int cols)
{
    __shared__ double tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
    for (int r = ty; r < 32; r += 8) {
        const int row = by + r, col = bx + tx;
        if (row < rows && col < cols) tile[r][tx] = src[(size_t)row * cols + col];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int col = bx + r, row = by + tx;
        if (row < rows && col < cols) dst[(size_t)col * rows + row] = tile[tx][r];
    }
}

int launch_transpose(gdt_context *h, const double *src, double *dst, int rows, int cols)
{
    hipLaunchKernelGGL(k_transpose, dim3((cols + 31) / 32, (rows + 31) / 32), dim3(256), 0, h->stream, src, dst, rows,
                       cols);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// The 14 moment-basis planes of fill_correction_array (velocity.cu:17-50)
// over the box: xi_x, xi_y, xi_z, xi^2, xi_x xi_x, xi_x xi_y, xi_x xi_z,
// xi_x xi^2, xi_y xi_y, xi_y xi_z, xi_y xi^2, xi_z xi_z, xi_z xi^2, xi^4.
__global__ void k_basis(int nu, int nv, int nw, double u0, double v0, double w0, double du, double dv, double dw,
                        double *__restrict__ dst)
{
    const int nb = nu * nv * nw;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    const int i = t / (nv * nw), j = (t / nw) % nv, k = t % nw;
    const double x = u0 + du * (2.0 * i + 1.0) / 2.0;
    const double y = v0 + dv * (2.0 * j + 1.0) / 2.0;
    const double z = w0 + dw * (2.0 * k + 1.0) / 2.0;
    const double e = x * x + y * y + z * z;
    const double p[14] = {x, y, z, e, x * x, x * y, x * z, x * e, y * y, y * z, y * e, z * z, z * e, e * e};
#pragma unroll
    for (int q = 0; q < 14; q++) dst[(size_t)q * nb + t] = p[q];
}

int launch_basis(gdt_context *h, double *dst)
{
    hipLaunchKernelGGL(k_basis, dim3((h->n_box + 255) / 256), dim3(256), 0, h->stream, h->n[0], h->n[1], h->n[2],
                       h->vmin[0], h->vmin[1], h->vmin[2], h->d[0], h->d[1], h->d[2], dst);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

__global__ void k_fill(double *dst, long count, double value)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < count) dst[t] = value;
}

int launch_fill(gdt_context *h, double *dst, long count, double value)
{
    hipLaunchKernelGGL(k_fill, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, dst, count, value);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}


// f[c][v] = sum_k n_k (2 pi T_k)^(-3/2) exp(-|xi_v - u_k|^2 / (2 T_k)),
// params[c][k] = (n, T, ux, uy, uz); components with n = 0 contribute nothing.
__global__ void k_init_mixture(int npts, int comps, long cells, const double *__restrict__ xi,
                               const double *__restrict__ params, double *__restrict__ f)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= cells * npts) return;
    const long c = t / npts;
    const int v = (int)(t % npts);
    const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
    double s = 0.0;
    for (int k = 0; k < comps; k++) {
        const double *p = params + (c * comps + k) * 5;
        if (p[0] == 0.0) continue;
        const double dx = x - p[2], dy = y - p[3], dz = z - p[4];
        s += p[0] * pow(2.0 * M_PI * p[1], -1.5) * exp(-(dx * dx + dy * dy + dz * dz) / (2.0 * p[1]));
    }
    f[t] = s;
}

int launch_init_mixture(gdt_context *h, const double *params, int comps, long cells, double *f)
{
    const long count = cells * h->n_pnt;
    hipLaunchKernelGGL(k_init_mixture, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, h->stream, h->n_pnt,
                       comps, cells, h->xi.as<double>(), params, f);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}
