// Setup kernels: analytic collision-matrix tables and the Korobov lattice
// stencil.  This file is compiled with -ffp-contract=off: every floating-point
// step that feeds an integer decision (table selection, nearest node, energy
// pair) is one IEEE operation, so the selections are reproducible bit for bit.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <thread>
#include <vector>

#include "gdt_internal.h"

#define GDT_ERR 1.0e-10 /* velocity.h:37 */
#define HALF_PI_ 1.5707963267948966
#define PI_ 3.141592653589793

// ---------------------------------------------------------------------------
// analytic matrix element
// ---------------------------------------------------------------------------

// Area antiderivatives f1..f4 (compute_matrix.cu:21-108).  With
// q = sqrt(1 - A^2 - p^2): H = p atan(q/A) + A atan(q/p) + atan(A p / q);
// f1, f2 -> H and f3, f4 -> pi/2 p - H, with the ERR special cases.
__device__ __forceinline__ double area_H(double p, double A, double s)
{
    const double q = sqrt(s);
    return p * atan(q / A) + A * atan(q / p) + atan(A * p / q);
}

__device__ double area_f(int kind, double p, double A)
{
    const double s = 1.0 - A * A - p * p;
    const bool pz = (p < GDT_ERR) && (p > -GDT_ERR);
    const bool sz = (s < GDT_ERR) && (s > -GDT_ERR);
    const bool reg = (s >= GDT_ERR) && (p >= GDT_ERR);
    if (kind == 1) {
        if ((A < GDT_ERR) && (A > -GDT_ERR)) return HALF_PI_ * p;
        if (pz) return HALF_PI_ * A;
        if (sz) return HALF_PI_;
        if (reg && A >= GDT_ERR) return area_H(p, A, s);
        return 0.0;
    }
    if (kind == 2) {
        if (A > 1.0 - GDT_ERR) return 0.0;
        if (pz) return HALF_PI_ * A;
        if (sz) return HALF_PI_;
        if (reg) return area_H(p, A, s);
        return 0.0;
    }
    if (kind == 3) {
        if (A < GDT_ERR) return 0.0;
        if (pz) return -HALF_PI_ * A;
        if (sz) return HALF_PI_ * (p - 1.0);
        if (reg) return HALF_PI_ * p - area_H(p, A, s);
        return 0.0;
    }
    if (A > 1.0 - GDT_ERR) return HALF_PI_ * p;
    if (pz) return -HALF_PI_ * A;
    if (sz) return HALF_PI_ * (p - 1.0);
    if (reg) return HALF_PI_ * p - area_H(p, A, s);
    return 0.0;
}

// Solid angle of the sphere |x| = g inside the lattice box of (in, jn, kn)
// (calc_matrix_element, compute_matrix.cu:110-257) as the integral over the
// height z of the arc inside the box's xy rectangle: the corner heights
// p0 <= {p1, p3} <= p2 cut [p0, p2] into three pieces with fixed bounding
// sides (f4-f2, then f1-f2 or f4-f3, then f1-f3).
__device__ double box_area(double du, double dv, double dw, int in, int jn, int kn, double g)
{
    const double u1 = in > 0 ? du * (in - 1) : 0.0, u2 = du * (in + 1);
    const double v1 = jn > 0 ? dv * (jn - 1) : 0.0, v2 = dv * (jn + 1);
    const double w1 = kn > 0 ? dw * (kn - 1) : 0.0, w2 = dw * (kn + 1);
    const double g2 = g * g;
    const double uu1 = u1 * u1, uu2 = u2 * u2, vv1 = v1 * v1, vv2 = v2 * v2;
    const double ww1 = w1 * w1, ww2 = w2 * w2;
    if (!(uu1 + vv1 + ww1 < g2 && uu2 + vv2 + ww2 > g2)) return 0.0;
    if (!(uu1 + vv1 < g2)) return 0.0;
    const double s0 = 1.0 - (uu2 + vv2) / g2, s1 = 1.0 - (uu1 + vv2) / g2;
    const double s2 = 1.0 - (uu1 + vv1) / g2, s3 = 1.0 - (uu2 + vv1) / g2;
    const double p0 = s0 < 0.0 ? 0.0 : sqrt(s0), p1 = s1 < 0.0 ? 0.0 : sqrt(s1);
    const double p2 = s2 < 0.0 ? 0.0 : sqrt(s2), p3 = s3 < 0.0 ? 0.0 : sqrt(s3);
    const double z0 = w1 / g;
    const double z1 = (w2 / g > 1.0) ? 1.0 : w2 / g;
    const double A1 = u1 / g, A2 = u2 / g, A3 = v1 / g, A4 = v2 / g;
    const bool m13 = p1 < p3;
    const double bl = m13 ? p1 : p3, bh = m13 ? p3 : p1;

    double area = 0.0;
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const double lo = s == 0 ? p0 : (s == 1 ? bl : bh);
        const double hi = s == 0 ? bl : (s == 1 ? bh : p2);
        const int ka = s == 0 ? 4 : (s == 1 ? (m13 ? 1 : 4) : 1);
        const int kb = s == 0 ? 2 : (s == 1 ? (m13 ? 2 : 3) : 3);
        const double Aa = s == 0 ? A4 : (s == 1 ? (m13 ? A1 : A4) : A1);
        const double Ab = s == 0 ? A2 : (s == 1 ? (m13 ? A2 : A3) : A3);
        const double za = z0 > lo ? z0 : lo;
        const double zb = z1 < hi ? z1 : hi;
        if (!(zb > za)) continue;
        const double Fb = area_f(ka, zb, Aa) - area_f(kb, zb, Ab);
        const double Fa = area_f(ka, za, Aa) - area_f(kb, za, Ab);
        area += Fb - Fa;
    }
    return area;
}

struct TableParams {
    int nu, nv, nw;
    int mode;
    double du, dv, dw, d3v;
    float fdu, fdv, fdw;
};

__device__ __forceinline__ double rel_speed(const TableParams &P, int gi, int gj, int gk)
{
    if (P.mode == GDT_TABLES_REFERENCE) {
        const float gu = P.fdu * (float)gi, gv = P.fdv * (float)gj, gw = P.fdw * (float)gk;
        return (double)sqrtf(gu * gu + gv * gv + gw * gw);
    }
    const double gu = P.du * gi, gv = P.dv * gj, gw = P.dw * gk;
    return sqrt(gu * gu + gv * gv + gw * gw);
}

// b[gi,gj,gk] = 4 pi g d3v (compute_matrix.cu:519-530).
__global__ void k_table_b(TableParams P, double *__restrict__ b)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = P.nu * P.nv * P.nw;
    if (t >= nb) return;
    const int gk = t % P.nw, gj = (t / P.nw) % P.nv, gi = t / (P.nv * P.nw);
    const double g = rel_speed(P, gi, gj, gk);
    const double v = 4.0 * PI_ * g * P.d3v;
    b[t] = (P.mode == GDT_TABLES_REFERENCE) ? (double)(float)v : v;
}

// a[in1,gi,jn1,kn1,gj,gk] = area < ERR ? 0 : area 2 g d3v, reference index
// order (compute_matrix.cu:532-562).  One thread per entry.
__global__ void k_table_a(TableParams P, double *__restrict__ a, long n_a)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_a) return;
    const int hv = P.nv / 2, hw = P.nw / 2;
    long r = t;
    const int gk = (int)(r % P.nw); r /= P.nw;
    const int gj = (int)(r % P.nv); r /= P.nv;
    const int kn1 = (int)(r % hw); r /= hw;
    const int jn1 = (int)(r % hv); r /= hv;
    const int gi = (int)(r % P.nu); r /= P.nu;
    const int in1 = (int)r;
    const int in = (in1 << 1) | (gi & 1), jn = (jn1 << 1) | (gj & 1), kn = (kn1 << 1) | (gk & 1);
    const double g = rel_speed(P, gi, gj, gk);
    const double A = box_area(P.du, P.dv, P.dw, in, jn, kn, g);
    double v = (A < GDT_ERR) ? 0.0 : A * 2.0 * g * P.d3v;
    if (P.mode == GDT_TABLES_REFERENCE) v = (double)(float)v;
    a[t] = v;
}

// Per relative-velocity triple: count / list the nonzero boxes in
// (in1, jn1, kn1) order, each as glist_pack(in, jn, kn).
__global__ void k_glist(int nu, int nv, int nw, const double *__restrict__ a, const int *__restrict__ gptr,
                        int *__restrict__ gcount, int *__restrict__ gbox, double *__restrict__ gval)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    const int nb = nu * nv * nw;
    if (g >= nb) return;
    const int hu = nu / 2, hv = nv / 2, hw = nw / 2, nyz = nv * nw;
    const int gk = g % nw, gj = (g / nw) % nv, gi = g / nyz;
    int cnt = 0;
    const int base = gptr ? gptr[g] : 0;
    for (int in1 = 0; in1 < hu; in1++)
        for (int jn1 = 0; jn1 < hv; jn1++)
            for (int kn1 = 0; kn1 < hw; kn1++) {
                const long ai = ((long)((in1 * nu + gi) * (hv * hw) + jn1 * hw + kn1)) * nyz + gj * nw + gk;
                const double v = a[ai];
                if (v != 0.0) {
                    if (gptr) {
                        gbox[base + cnt] = glist_pack(in1 << 1 | (gi & 1), jn1 << 1 | (gj & 1), kn1 << 1 | (gk & 1));
                        gval[base + cnt] = v;
                    }
                    cnt++;
                }
            }
    if (!gptr) gcount[g] = cnt;
}

int launch_build_tables(gdt_context *h, int mode)
{
    TableParams P;
    P.nu = h->n[0];
    P.nv = h->n[1];
    P.nw = h->n[2];
    P.mode = mode;
    if (mode == GDT_TABLES_REFERENCE) {
        P.fdu = ((float)h->vmax[0] - (float)h->vmin[0]) / (float)P.nu;
        P.fdv = ((float)h->vmax[1] - (float)h->vmin[1]) / (float)P.nv;
        P.fdw = ((float)h->vmax[2] - (float)h->vmin[2]) / (float)P.nw;
        const float fd3v = P.fdu * P.fdv * P.fdw;
        P.du = P.fdu;
        P.dv = P.fdv;
        P.dw = P.fdw;
        P.d3v = fd3v;
    } else {
        P.du = (h->vmax[0] - h->vmin[0]) / P.nu;
        P.dv = (h->vmax[1] - h->vmin[1]) / P.nv;
        P.dw = (h->vmax[2] - h->vmin[2]) / P.nw;
        P.d3v = P.du * P.dv * P.dw;
        P.fdu = P.fdv = P.fdw = 0.0f;
    }
    const int nb = h->n_box;
    const long n_a = (long)nb * nb / 8;
    GDT_CHECK(h->tb.ensure(sizeof(double) * nb) == 0, GDT_E_NOMEM, "table b allocation failed");
    GDT_CHECK(h->ta.ensure(sizeof(double) * n_a) == 0, GDT_E_NOMEM, "table a allocation failed (%ld entries)", n_a);
    hipLaunchKernelGGL(k_table_b, dim3((nb + 255) / 256), dim3(256), 0, h->stream, P, h->tb.as<double>());
    GDT_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_table_a, dim3((unsigned)((n_a + 255) / 256)), dim3(256), 0, h->stream, P,
                       h->ta.as<double>(), n_a);
    GDT_HIP(hipGetLastError());

    // nonzero-box lists: count, host scan, fill
    DevBuf cnt;
    GDT_CHECK(cnt.ensure(sizeof(int) * nb) == 0, GDT_E_NOMEM, "glist count allocation failed");
    hipLaunchKernelGGL(k_glist, dim3((nb + 127) / 128), dim3(128), 0, h->stream, P.nu, P.nv, P.nw,
                       h->ta.as<double>(), (const int *)nullptr, cnt.as<int>(), (int *)nullptr, (double *)nullptr);
    GDT_HIP(hipGetLastError());
    std::vector<int> hc(nb), hp(nb + 1);
    GDT_HIP(hipMemcpyAsync(hc.data(), cnt.p, sizeof(int) * nb, hipMemcpyDeviceToHost, h->stream));
    GDT_HIP(hipStreamSynchronize(h->stream));
    cnt.release();
    long tot = 0;
    for (int g = 0; g < nb; g++) {
        hp[g] = (int)tot;
        tot += hc[g];
    }
    hp[nb] = (int)tot;
    GDT_CHECK(h->gptr.ensure(sizeof(int) * (nb + 1)) == 0, GDT_E_NOMEM, "glist ptr allocation failed");
    GDT_CHECK(h->gbox.ensure(sizeof(int) * (tot > 0 ? tot : 1)) == 0, GDT_E_NOMEM, "glist allocation failed");
    GDT_CHECK(h->gval.ensure(sizeof(double) * (tot > 0 ? tot : 1)) == 0, GDT_E_NOMEM, "glist allocation failed");
    GDT_HIP(hipMemcpyAsync(h->gptr.p, hp.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_glist, dim3((nb + 127) / 128), dim3(128), 0, h->stream, P.nu, P.nv, P.nw,
                       h->ta.as<double>(), h->gptr.as<int>(), (int *)nullptr, h->gbox.as<int>(),
                       h->gval.as<double>());
    GDT_HIP(hipGetLastError());
    GDT_HIP(hipStreamSynchronize(h->stream));
    h->n_a = n_a;
    h->n_gent = tot;
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// Korobov lattice stencil
// ---------------------------------------------------------------------------

__constant__ double c_sin[12] = {
    1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
    2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
    -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18,
    1.9572941063391263e-20, -3.8681701706306835e-23};
__constant__ double c_cos[12] = {
    1.0, -0.5, 0.041666666666666664, -0.001388888888888889,
    2.48015873015873e-05, -2.755731922398589e-07, 2.08767569878681e-09,
    -1.1470745597729725e-11, 4.779477332387385e-14, -1.5619206968586225e-16,
    4.110317623312165e-19, -8.896791392450574e-22};

// sin, cos of 2 pi t, t in [0,1): quadrant from 4t, Taylor polynomials in
// theta^2 on [0, pi/2).  Only +,-,* so the result is the same on every
// IEEE machine that evaluates it in this order.
__device__ __forceinline__ void sincos_2pi(double t, double &s, double &c)
{
    const double t4 = 4.0 * t;
    const int q = (int)floor(t4);
    const double th = (t4 - (double)q) * HALF_PI_;
    const double th2 = th * th;
    double ps = c_sin[11], pc = c_cos[11];
#pragma unroll
    for (int k = 10; k >= 0; k--) {
        ps = ps * th2 + c_sin[k];
        pc = pc * th2 + c_cos[k];
    }
    const double sv = th * ps, cv = pc;
    switch (q & 3) {
    case 0: s = sv; c = cv; break;
    case 1: s = cv; c = -sv; break;
    case 2: s = -sv; c = -cv; break;
    default: s = -cv; c = sv; break;
    }
}

struct KorParams {
    int n0, n1, n2;
    uint32_t M, mask;
    uint32_t z[8];
    double K;
};

__device__ __forceinline__ int dom(const KorParams &P, const int *__restrict__ box2c, int x, int y, int z)
{
    if (x < 0 || x >= P.n0 || y < 0 || y >= P.n1 || z < 0 || z >= P.n2) return -1;
    return box2c[(x * P.n1 + y) * P.n2 + z];
}

// One thread per lattice node nu in [0, M): pre-collision pair, scattering
// direction, nearest-node projection and energy-bracketing second pair.
// Writes the record and a validity byte; the host keeps valid nodes in order.
__global__ void k_korobov_nodes(KorParams P, const int *__restrict__ box2c, int32_t *__restrict__ idx6,
                                double *__restrict__ rr, double *__restrict__ ww, int32_t *__restrict__ ints,
                                uint8_t *__restrict__ valid)
{
    const uint32_t nu = blockIdx.x * blockDim.x + threadIdx.x;
    if (nu >= P.M) return;
    uint32_t x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = (uint32_t)(((uint64_t)nu * P.z[j]) & P.mask);
    const int nax[3] = {P.n0, P.n1, P.n2};
    int ia[3], ib[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        ia[k] = (int)(((uint64_t)x[k] * (uint64_t)nax[k]) / P.M);
        ib[k] = (int)(((uint64_t)x[3 + k] * (uint64_t)nax[k]) / P.M);
    }
    valid[nu] = 0;
    const int ca = dom(P, box2c, ia[0], ia[1], ia[2]);
    const int cb = dom(P, box2c, ib[0], ib[1], ib[2]);
    if (ca < 0 || cb < 0) return;
    const int D2 = (ia[0] - ib[0]) * (ia[0] - ib[0]) + (ia[1] - ib[1]) * (ia[1] - ib[1]) +
                   (ia[2] - ib[2]) * (ia[2] - ib[2]);
    if (D2 == 0) return;

    const double t7 = ((double)x[6] + 0.5) / (double)P.M;
    const double t8 = ((double)x[7] + 0.5) / (double)P.M;
    double sn, cs;
    sincos_2pi(t8, sn, cs);
    const double rho = 2.0 * sqrt(t7 * (1.0 - t7));
    const double om[3] = {rho * cs, rho * sn, 1.0 - 2.0 * t7};
    const double hh = 0.5 * sqrt((double)D2);

    int lam[3], mu[3];
    double del[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double xp = 0.5 * (double)(ia[k] + ib[k]) + hh * om[k];
        lam[k] = (int)floor(xp + 0.5);
        del[k] = xp - (double)lam[k];
        mu[k] = ia[k] + ib[k] - lam[k];
    }
    const int cl = dom(P, box2c, lam[0], lam[1], lam[2]);
    const int cm = dom(P, box2c, mu[0], mu[1], mu[2]);
    if (cl < 0 || cm < 0) return;

    int E0 = 0, El = 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const int pa = 2 * ia[k] + 1 - nax[k], pb = 2 * ib[k] + 1 - nax[k];
        const int pl = 2 * lam[k] + 1 - nax[k], pm = 2 * mu[k] + 1 - nax[k];
        E0 += pa * pa + pb * pb;
        El += pl * pl + pm * pm;
    }
    int cls = cl, cms = cm, den = 0;
    double r = 0.0;
    if (El != E0) {
        int best = -1, bestE = 0, bl = -1, bm = -1;
        double bestd = DBL_MAX;
        for (int S = 1; S < 8; S++) {
            int ls[3], ms[3];
            double d2 = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int sk = ((S >> k) & 1) ? (del[k] >= 0.0 ? 1 : -1) : 0;
                ls[k] = lam[k] + sk;
                ms[k] = mu[k] - sk;
                const double dk = sk ? 1.0 - fabs(del[k]) : del[k];
                d2 = d2 + dk * dk;
            }
            const int c1 = dom(P, box2c, ls[0], ls[1], ls[2]);
            const int c2 = dom(P, box2c, ms[0], ms[1], ms[2]);
            if (c1 < 0 || c2 < 0) continue;
            int Es = 0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const int p1 = 2 * ls[k] + 1 - nax[k], p2 = 2 * ms[k] + 1 - nax[k];
                Es += p1 * p1 + p2 * p2;
            }
            const bool other = (El < E0) ? (Es >= E0) : (Es <= E0);
            if (other && d2 < bestd) {
                bestd = d2;
                best = S;
                bestE = Es;
                bl = c1;
                bm = c2;
            }
        }
        if (best < 0) return;
        cls = bl;
        cms = bm;
        den = bestE - El;
        r = (double)(E0 - El) / (double)den;
    }
    int32_t *o = idx6 + 6 * (size_t)nu;
    o[0] = ca;
    o[1] = cb;
    o[2] = cl;
    o[3] = cm;
    o[4] = cls;
    o[5] = cms;
    rr[nu] = r;
    ww[nu] = P.K * sqrt((double)D2);
    // integer facts that fix r and w: E0 - E_lm, E_s - E_lm (0 when r = 0), D2
    int32_t *q = ints + 3 * (size_t)nu;
    q[0] = E0 - El;
    q[1] = den;
    q[2] = D2;
    valid[nu] = 1;
}

int launch_korobov_nodes(gdt_context *h, uint32_t M, uint32_t a, double K, int32_t *d_idx6, double *d_r,
                         double *d_w, int32_t *d_ints, uint8_t *d_valid)
{
    KorParams P;
    P.n0 = h->n[0];
    P.n1 = h->n[1];
    P.n2 = h->n[2];
    P.M = M;
    P.mask = M - 1;
    uint64_t z = 1;
    for (int j = 0; j < 8; j++) {
        P.z[j] = (uint32_t)z;
        z = (z * a) & (uint64_t)(M - 1);
    }
    P.K = K;
    hipLaunchKernelGGL(k_korobov_nodes, dim3((M + 255) / 256), dim3(256), 0, h->stream, P, h->box2c.as<int>(),
                       d_idx6, d_r, d_w, d_ints, d_valid);
    GDT_HIP(hipGetLastError());
    return GDT_OK;
}

// ---------------------------------------------------------------------------
// Korobov generator choice (host)
// ---------------------------------------------------------------------------

// Korobov's P_2 figure of merit of the 8-D lattice with generator a, summed
// in node order so the value is reproducible.
static double korobov_p2(uint32_t M, uint32_t a)
{
    uint64_t z[8];
    z[0] = 1;
    for (int j = 1; j < 8; j++) z[j] = (z[j - 1] * a) % M;
    double sum = 0.0;
    for (uint64_t k = 0; k < M; k++) {
        double prod = 1.0;
        for (int j = 0; j < 8; j++) {
            const double x = (double)((k * z[j]) % M) / (double)M;
            const double b2 = x * x - x + 1.0 / 6.0;
            prod = prod * (1.0 + 19.739208802178716 * b2);
        }
        sum = sum + prod;
    }
    return sum / (double)M - 1.0;
}

static uint64_t splitmix(uint64_t &st)
{
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// ncand odd candidates from splitmix64(seed); smallest P_2 wins (ties: the
// earlier candidate).  Candidates are scored on host threads.
uint32_t host_korobov_pick(uint32_t M, uint64_t seed, int ncand, double *p2)
{
    uint64_t st = seed;
    std::vector<uint32_t> cand(ncand);
    for (int c = 0; c < ncand; c++) {
        uint32_t a = (uint32_t)((splitmix(st) % (M / 2)) * 2 + 1);
        cand[c] = a < 3 ? 3 : a;
    }
    std::vector<double> score(ncand);
    std::vector<std::thread> th;
    for (int c = 0; c < ncand; c++) th.emplace_back([&, c] { score[c] = korobov_p2(M, cand[c]); });
    for (auto &t : th) t.join();
    int best = 0;
    for (int c = 1; c < ncand; c++)
        if (score[c] < score[best]) best = c;
    if (p2) *p2 = score[best];
    return cand[best];
}
