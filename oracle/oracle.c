/*
 * oracle.c -- FP64 CPU restatement of the gpu_determ collision-integral path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library in
 * gpu_determ_amd/csrc.  It is loaded by tests/, by __graft_entry__.smoke()
 * and by the cpu_baseline leg of bench.py, and by nothing else.  The product
 * path never links or calls it.
 *
 * Every function names the reference file:line it restates.  The reference
 * runs in FP32 with several latent defects; where this restatement follows the
 * intended behaviour instead of the literal one, the comment says so.
 *
 * Pins (what checks this file independently of the HIP code):
 *   grid / cut ........ closed-form node counts (tests/test_oracle.py)
 *   orc_init_f ........ golden F column of true_results.out
 *   orc_loss .......... golden DC column (cells != 0) and the closed-form
 *                       hard-sphere collision frequency of a Maxwellian
 *   orc_box_area ...... independent 1-D quadrature of the sphere/box area
 *                       (tests/golden/box_area.json) and the octant-sum
 *                       invariant sum_boxes area = pi/2
 *   orc_gain .......... golden IC column (slots 0..7) and brute force over a
 *                       dense (v, v1, n) enumeration on a 6^3 grid
 *   orc_correct ....... residual of the 5x5 system, conserved moments
 *   Korobov ........... lattice permutation identity, exact per-node mass /
 *                       momentum / energy balance, Q(Maxwellian) = 0, loss ->
 *                       closed-form collision frequency, gain -> orc_gain.
 *                       The method has no reference counterpart, so these
 *                       invariants are its only pins: bitwise agreement with
 *                       an independent implementation is parity unpinned.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).  Contraction is
 * off so that every floating-point step of the integer-selecting code (grid
 * cut, Korobov stencil) is a single IEEE operation, as on the GPU side.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_HALF_PI 1.5707963267948966
#define ORC_PI 3.141592653589793
#define ORC_TWO_PI_SQ 19.739208802178716 /* 2*pi^2 */

static inline int iabs_(int x) { return x < 0 ? -x : x; }

/* ======================================================================== */
/* velocity grid                                                            */
/* ======================================================================== */

/* Membership test and compact ordering: velocity_grid.cu:6-14 (centre
 * measured from the middle of the box, ignoring v_min's offset) and
 * velocity_grid.cu:114-126 (compact list in i-j-k order).  Evaluated in FP64;
 * the reference's FP32 evaluation cannot differ at the configurations used
 * (sums of three odd squares never equal 4 R^2 / d^2 there). */
int orc_grid_nodes(const int n[3], const double vmin[3], const double vmax[3],
                   double R, int *box2c, int *cidx)
{
    int cnt = 0, idx = 0;
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++, idx++) {
                double x = (vmax[0] - vmin[0]) * (2 * i + 1 - n[0]) / 2.0 / n[0];
                double y = (vmax[1] - vmin[1]) * (2 * j + 1 - n[1]) / 2.0 / n[1];
                double z = (vmax[2] - vmin[2]) * (2 * k + 1 - n[2]) / 2.0 / n[2];
                int keep = (R <= 0.0) || (x * x + y * y + z * z < R * R);
                box2c[idx] = keep ? cnt : -1;
                if (keep) {
                    if (cidx) {
                        cidx[3 * cnt + 0] = i;
                        cidx[3 * cnt + 1] = j;
                        cidx[3 * cnt + 2] = k;
                    }
                    cnt++;
                }
            }
    return cnt;
}

/* Node centres, velocity_grid.cu:291-299: u_i = min + d (2i+1)/2.  With
 * ref_float the reference's FP32 members are reproduced (d = (max-min)/n in
 * float, the centre formed in double and stored as float). */
void orc_grid_centres(const int n[3], const double vmin[3],
                      const double vmax[3], int ref_float, double *u,
                      double *v, double *w)
{
    double *out[3] = {u, v, w};
    for (int a = 0; a < 3; a++) {
        if (ref_float) {
            float lo = (float)vmin[a], hi = (float)vmax[a];
            float d = (hi - lo) / (float)n[a];
            for (int i = 0; i < n[a]; i++)
                out[a][i] = (double)(float)((double)lo + (double)d * (2.0 * i + 1.0) / 2.0);
        } else {
            double d = (vmax[a] - vmin[a]) / n[a];
            for (int i = 0; i < n[a]; i++)
                out[a][i] = vmin[a] + d * (2.0 * i + 1.0) / 2.0;
        }
    }
}

/* The reference's initial distribution, determ.cu:101-121: the same FP32
 * exp(-(u-1.75)^2 - (v-0.25)^2 - w^2) in every cell.  One cell is written. */
void orc_init_f_reference(const int n[3], const double vmin[3],
                          const double vmax[3], double *f_box)
{
    double *u = malloc(sizeof(double) * (n[0] + n[1] + n[2]));
    double *v = u + n[0], *w = v + n[1];
    orc_grid_centres(n, vmin, vmax, 1, u, v, w);
    const float u1 = 1.75f, v1 = 0.25f;
    int idx = 0;
    for (int i = 0; i < n[0]; i++)
        for (int j = 0; j < n[1]; j++)
            for (int k = 0; k < n[2]; k++, idx++) {
                float fu = (float)u[i], fv = (float)v[j], fw = (float)w[k];
                float du = fu - u1, dv = fv - v1;
                float e = -du * du - dv * dv - fw * fw;
                f_box[idx] = (double)expf(e);
            }
    free(u);
}

/* ======================================================================== */
/* analytic collision matrix                                                */
/* ======================================================================== */

/* The four area antiderivatives f1..f4 of compute_matrix.cu:21-108.  With
 * q = sqrt(1 - A^2 - p^2) the common part is
 *     H(p, A) = p atan(q/A) + A atan(q/p) + atan(A p / q);
 * f1 and f2 return H, f3 and f4 return pi/2 p - H, each with the reference's
 * ERR special cases for A ~ 0, p ~ 0, q ~ 0 and A ~ 1. */
static double orc_H(double p, double A, double s)
{
    double q = sqrt(s);
    return p * atan(q / A) + A * atan(q / p) + atan(A * p / q);
}

double orc_fterm(int kind, double p, double A)
{
    const double s = 1.0 - A * A - p * p;
    const int p0 = (p < ORC_ERR) && (p > -ORC_ERR);
    const int s0 = (s < ORC_ERR) && (s > -ORC_ERR);
    const int regular = (s >= ORC_ERR) && (p >= ORC_ERR);
    switch (kind) {
    case 1: /* compute_matrix.cu:21-42 */
        if ((A < ORC_ERR) && (A > -ORC_ERR)) return ORC_HALF_PI * p;
        if (p0) return ORC_HALF_PI * A;
        if (s0) return ORC_HALF_PI;
        if (regular && A >= ORC_ERR) return orc_H(p, A, s);
        return 0.0;
    case 2: /* compute_matrix.cu:45-64 */
        if (A > 1.0 - ORC_ERR) return 0.0;
        if (p0) return ORC_HALF_PI * A;
        if (s0) return ORC_HALF_PI;
        if (regular) return orc_H(p, A, s);
        return 0.0;
    case 3: /* compute_matrix.cu:67-86 */
        if (A < ORC_ERR) return 0.0;
        if (p0) return -ORC_HALF_PI * A;
        if (s0) return ORC_HALF_PI * (p - 1.0);
        if (regular) return ORC_HALF_PI * p - orc_H(p, A, s);
        return 0.0;
    default: /* 4: compute_matrix.cu:89-108 */
        if (A > 1.0 - ORC_ERR) return ORC_HALF_PI * p;
        if (p0) return -ORC_HALF_PI * A;
        if (s0) return ORC_HALF_PI * (p - 1.0);
        if (regular) return ORC_HALF_PI * p - orc_H(p, A, s);
        return 0.0;
    }
}

/* Area on the unit sphere of {w : g w in box}, box = [u1,u2]x[v1,v2]x[w1,w2]
 * with u1 = du (in-1)^+, u2 = du (in+1) (same for v, w):
 * calc_matrix_element, compute_matrix.cu:110-257.
 *
 * The reference selects one of 26 closed forms.  Each of them is the integral
 * over the height z in [w1/g, min(1, w2/g)] of the arc of the circle
 * x^2+y^2 = 1-z^2 inside [A1,A2]x[A3,A4] (A = u1/g, u2/g, v1/g, v2/g),
 * telescoped.  Here the same integral is formed directly: the corner heights
 * p0 <= {p1,p3} <= p2 split [p0, p2] into three pieces, on each of which the
 * arc is bounded by a fixed pair of rectangle sides, so the integrand has a
 * fixed antiderivative F = f_a - f_b:
 *     [p0, min(p1,p3)]            f4(.,A4) - f2(.,A2)
 *     [min(p1,p3), max(p1,p3)]    f1(.,A1) - f2(.,A2)   if p1 < p3
 *                                 f4(.,A4) - f3(.,A3)   otherwise
 *     [max(p1,p3), p2]            f1(.,A1) - f3(.,A3)
 * and the area is the sum over the pieces of F(min(z1,hi)) - F(max(z0,lo)).
 * The reference's entry test (nearest corner inside the sphere, farthest
 * outside, :135) is kept; its eps1 > eps3 test (:136-151) is the condition
 * u1^2 + v1^2 < g^2, which this form meets by p2 collapsing to 0. */
double orc_box_area(double du, double dv, double dw, int in, int jn, int kn,
                    double g)
{
    const double u1 = in > 0 ? du * (in - 1) : 0.0, u2 = du * (in + 1);
    const double v1 = jn > 0 ? dv * (jn - 1) : 0.0, v2 = dv * (jn + 1);
    const double w1 = kn > 0 ? dw * (kn - 1) : 0.0, w2 = dw * (kn + 1);
    const double g2 = g * g;
    const double uu1 = u1 * u1, uu2 = u2 * u2, vv1 = v1 * v1, vv2 = v2 * v2;
    const double ww1 = w1 * w1, ww2 = w2 * w2;
    if (!(uu1 + vv1 + ww1 < g2 && uu2 + vv2 + ww2 > g2)) return 0.0;
    if (!(uu1 + vv1 < g2)) return 0.0;

    double sq[4] = {1.0 - (uu2 + vv2) / g2, 1.0 - (uu1 + vv2) / g2,
                    1.0 - (uu1 + vv1) / g2, 1.0 - (uu2 + vv1) / g2};
    double p[4];
    for (int c = 0; c < 4; c++) p[c] = sq[c] < 0.0 ? 0.0 : sqrt(sq[c]);
    const double z0 = w1 / g;
    const double z1 = (w2 / g > 1.0) ? 1.0 : w2 / g;
    const double A1 = u1 / g, A2 = u2 / g, A3 = v1 / g, A4 = v2 / g;

    const int mid13 = p[1] < p[3];
    const double bl = mid13 ? p[1] : p[3], bh = mid13 ? p[3] : p[1];
    const double lo[3] = {p[0], bl, bh}, hi[3] = {bl, bh, p[2]};
    /* (kind_a, A_a, kind_b, A_b) of each piece */
    const int ka[3] = {4, mid13 ? 1 : 4, 1}, kb[3] = {2, mid13 ? 2 : 3, 3};
    const double Aa[3] = {A4, mid13 ? A1 : A4, A1};
    const double Ab[3] = {A2, mid13 ? A2 : A3, A3};

    double area = 0.0;
    for (int s = 0; s < 3; s++) {
        double za = z0 > lo[s] ? z0 : lo[s];
        double zb = z1 < hi[s] ? z1 : hi[s];
        if (!(zb > za)) continue;
        double Fb = orc_fterm(ka[s], zb, Aa[s]) - orc_fterm(kb[s], zb, Ab[s]);
        double Fa = orc_fterm(ka[s], za, Aa[s]) - orc_fterm(kb[s], za, Ab[s]);
        area += Fb - Fa;
    }
    return area;
}

/* b and a tables, store_matrix_elements (compute_matrix.cu:505-562).
 *   b[gi,gj,gk]                   = 4 pi g d3v
 *   a[in1,gi,jn1,kn1,gj,gk]       = area < ERR ? 0 : area * 2 g d3v
 * with in = (in1 << 1) | (gi & 1) (same for jn, kn) and the reference's index
 * ((in1*nu + gi)*(nv/2*nw/2) + jn1*(nw/2) + kn1)*(nv*nw) + gj*nw + gk.
 * mode 0: FP64 grid spacing and g = sqrt(du^2 gi^2 + ...) in FP64.
 * mode 1: the reference's FP32 spacing, FP32 g (:511,517,553) and FP32
 *         stored values (the golden file was produced this way). */
void orc_build_tables(const int n[3], const double vmin[3],
                      const double vmax[3], int mode, double *b, double *a)
{
    const int nu = n[0], nv = n[1], nw = n[2];
    double du, dv, dw, d3v;
    float fdu = 0, fdv = 0, fdw = 0, fd3v = 0;
    if (mode == 1) {
        fdu = ((float)vmax[0] - (float)vmin[0]) / (float)nu;
        fdv = ((float)vmax[1] - (float)vmin[1]) / (float)nv;
        fdw = ((float)vmax[2] - (float)vmin[2]) / (float)nw;
        fd3v = fdu * fdv * fdw;
        du = fdu; dv = fdv; dw = fdw; d3v = fd3v;
    } else {
        du = (vmax[0] - vmin[0]) / nu;
        dv = (vmax[1] - vmin[1]) / nv;
        dw = (vmax[2] - vmin[2]) / nw;
        d3v = du * dv * dw;
    }
    const int nyz = nv * nw, hv = nv / 2, hw = nw / 2, hu = nu / 2;

    double *gtab = malloc(sizeof(double) * nu * nyz);
    for (int gi = 0; gi < nu; gi++)
        for (int gj = 0; gj < nv; gj++)
            for (int gk = 0; gk < nw; gk++) {
                double g;
                if (mode == 1) {
                    float gu = fdu * (float)gi, gv = fdv * (float)gj, gw = fdw * (float)gk;
                    float s = gu * gu + gv * gv + gw * gw;
                    g = (double)sqrtf(s);
                } else {
                    double gu = du * gi, gv = dv * gj, gw = dw * gk;
                    g = sqrt(gu * gu + gv * gv + gw * gw);
                }
                gtab[(gi * nv + gj) * nw + gk] = g;
                double bv = 4.0 * ORC_PI * g * d3v;
                b[(gi * nv + gj) * nw + gk] = (mode == 1) ? (double)(float)bv : bv;
            }

#pragma omp parallel for collapse(2) schedule(dynamic)
    for (int in1 = 0; in1 < hu; in1++)
        for (int gi = 0; gi < nu; gi++) {
            const int in = (in1 << 1) | (gi & 1);
            for (int jn1 = 0; jn1 < hv; jn1++)
                for (int kn1 = 0; kn1 < hw; kn1++) {
                    long base = ((long)((in1 * nu + gi) * (hv * hw) + jn1 * hw + kn1)) * nyz;
                    for (int gj = 0; gj < nv; gj++)
                        for (int gk = 0; gk < nw; gk++) {
                            const int jn = (gj & 1) | (jn1 << 1);
                            const int kn = (gk & 1) | (kn1 << 1);
                            const double g = gtab[(gi * nv + gj) * nw + gk];
                            double A = orc_box_area(du, dv, dw, in, jn, kn, g);
                            double val = (A < ORC_ERR) ? 0.0 : A * 2.0 * g * d3v;
                            if (mode == 1) val = (double)(float)val;
                            a[base + gj * nw + gk] = val;
                        }
                }
        }
    free(gtab);
}

/* Nonzero boxes of each relative-velocity triple (gi,gj,gk): the sparse view
 * of the a table that both the gain below and the GPU kernels walk.  gptr has
 * n_box+1 entries; an entry's box is packed as (in1*(nv/2) + jn1)*(nw/2)+kn1.
 * Returns the entry count (or the needed count if it exceeds cap). */
long orc_glist_build(const int n[3], const double *a, int *gptr, int *gbox,
                     double *gval, long cap)
{
    const int nu = n[0], nv = n[1], nw = n[2], nyz = nv * nw;
    const int hu = nu / 2, hv = nv / 2, hw = nw / 2;
    long cnt = 0;
    for (int gi = 0; gi < nu; gi++)
        for (int gj = 0; gj < nv; gj++)
            for (int gk = 0; gk < nw; gk++) {
                int gidx = (gi * nv + gj) * nw + gk;
                if (gptr) gptr[gidx] = (int)cnt;
                for (int in1 = 0; in1 < hu; in1++)
                    for (int jn1 = 0; jn1 < hv; jn1++)
                        for (int kn1 = 0; kn1 < hw; kn1++) {
                            long ai = ((long)((in1 * nu + gi) * (hv * hw) + jn1 * hw + kn1)) * nyz + gj * nw + gk;
                            double val = a[ai];
                            if (val != 0.0) {
                                if (cnt < cap) {
                                    gbox[cnt] = (in1 * hv + jn1) * hw + kn1;
                                    gval[cnt] = val;
                                }
                                cnt++;
                            }
                        }
            }
    if (gptr) gptr[nu * nyz] = (int)cnt;
    return cnt;
}

/* ======================================================================== */
/* loss frequency and gain                                                  */
/* ======================================================================== */

/* nu(v) = sum_v1 b(|i-i1|,|j-j1|,|k-k1|) f(v1): evolve_direct_colisions,
 * collision_integral.cu:32-72, with one cell per call slice. */
void orc_loss(const int n[3], const double *b, const double *f_box,
              double *nu_box, int cells)
{
    const int nu = n[0], nv = n[1], nw = n[2], nb = nu * nv * nw;
#pragma omp parallel for collapse(2) schedule(static)
    for (int c = 0; c < cells; c++)
        for (int v = 0; v < nb; v++) {
            const double *f = f_box + (long)c * nb;
            const int i = v / (nv * nw), j = (v / nw) % nv, k = v % nw;
            double s = 0.0;
            for (int i1 = 0; i1 < nu; i1++)
                for (int j1 = 0; j1 < nv; j1++) {
                    const double *brow = b + (iabs_(i - i1) * nv + iabs_(j - j1)) * nw;
                    const double *frow = f + (i1 * nv + j1) * nw;
                    for (int k1 = 0; k1 < nw; k1++) s += brow[iabs_(k - k1)] * frow[k1];
                }
            nu_box[(long)c * nb + v] = s;
        }
}

/* Gain term, evolve_inverse_colisions (collision_integral.cu:75-188) summed
 * over all 512 host-loop boxes (:637-648) and written to the intended slot
 * i*nyz + (j*nz+k) (the reference writes i*nx + ..., a defect).  For
 * destination v, partner v1 and box (in1,jn1,kn1):
 *   in = (in1<<1) | ((i+i1)&1),  il = (i+i1+in)/2 <= nu-1,  i+i1-in >= 0,
 *   il1 = (i+i1-in)/2; jn, jl, jl1 and kn, kl, kl1 alike with jl < nv,
 *   j+j1-jn >= 0, kl < nw, k+k1-kn >= 0;
 *   gain += a * ( f(il,jl,kl) f(il1,jl1,kl1) + f(il,jl1,kl1) f(il1,jl,kl)
 *               + f(il,jl,kl1) f(il1,jl1,kl) + f(il,jl1,kl) f(il1,jl,kl1) ).
 * Boxes with a == 0 contribute nothing and are skipped via the g-list.
 * dest (ndest entries, box indices) restricts the destinations; NULL = all. */
void orc_gain(const int n[3], const double *a, const double *f_box,
              double *gain_box, int cells, const int *dest, int ndest)
{
    const int nu = n[0], nv = n[1], nw = n[2], nb = nu * nv * nw;
    const int hv = nv / 2, hw = nw / 2;
    long ne = orc_glist_build(n, a, NULL, NULL, NULL, 0);
    int *gptr = malloc(sizeof(int) * (nb + 1));
    int *gbox = malloc(sizeof(int) * (ne > 0 ? ne : 1));
    double *gval = malloc(sizeof(double) * (ne > 0 ? ne : 1));
    orc_glist_build(n, a, gptr, gbox, gval, ne);
    const int nd = dest ? ndest : nb;

#pragma omp parallel for collapse(2) schedule(dynamic, 16)
    for (int c = 0; c < cells; c++)
        for (int t = 0; t < nd; t++) {
            const int v = dest ? dest[t] : t;
            const double *f = f_box + (long)c * nb;
            const int i = v / (nv * nw), j = (v / nw) % nv, k = v % nw;
            double s = 0.0;
            for (int i1 = 0; i1 < nu; i1++)
                for (int j1 = 0; j1 < nv; j1++)
                    for (int k1 = 0; k1 < nw; k1++) {
                        const int gidx = (iabs_(i - i1) * nv + iabs_(j - j1)) * nw + iabs_(k - k1);
                        for (int e = gptr[gidx]; e < gptr[gidx + 1]; e++) {
                            const int box = gbox[e];
                            const int kn1 = box % hw, jn1 = (box / hw) % hv, in1 = box / (hw * hv);
                            const int in = (in1 << 1) | ((i + i1) & 1);
                            const int jn = (jn1 << 1) | ((j + j1) & 1);
                            const int kn = (kn1 << 1) | ((k + k1) & 1);
                            const int il = (i + i1 + in) >> 1, im = i + i1 - in;
                            const int jl = (j + j1 + jn) >> 1, jm = j + j1 - jn;
                            const int kl = (k + k1 + kn) >> 1, km = k + k1 - kn;
                            if (il > nu - 1 || im < 0 || jl > nv - 1 || jm < 0 || kl > nw - 1 || km < 0)
                                continue;
                            const int il1 = im >> 1, jl1 = jm >> 1, kl1 = km >> 1;
#define F_(x, y, z) f[((x) * nv + (y)) * nw + (z)]
                            const double t4 = F_(il, jl, kl) * F_(il1, jl1, kl1) + F_(il, jl1, kl1) * F_(il1, jl, kl) +
                                              F_(il, jl, kl1) * F_(il1, jl1, kl) + F_(il, jl1, kl) * F_(il1, jl, kl1);
#undef F_
                            s += gval[e] * t4;
                        }
                    }
            gain_box[(long)c * nb + (dest ? t : v)] = s;
        }
    free(gptr);
    free(gbox);
    free(gval);
}

/* ======================================================================== */
/* conservation correction                                                  */
/* ======================================================================== */

/* 5x5 Gaussian elimination with partial pivoting.  Replaces linear_solver
 * (collision_integral.cu:464-503), which pivots not at all, solves only cell 0
 * and races on its shared arrays.  Returns 0, or -1 for a singular system. */
int orc_solve5(const double A[25], const double B[5], double X[5])
{
    double M[5][6];
    for (int r = 0; r < 5; r++) {
        for (int c = 0; c < 5; c++) M[r][c] = A[r * 5 + c];
        M[r][5] = B[r];
    }
    for (int col = 0; col < 5; col++) {
        int piv = col;
        for (int r = col + 1; r < 5; r++)
            if (fabs(M[r][col]) > fabs(M[piv][col])) piv = r;
        if (M[piv][col] == 0.0) return -1;
        if (piv != col)
            for (int c = 0; c < 6; c++) {
                double t = M[col][c];
                M[col][c] = M[piv][c];
                M[piv][c] = t;
            }
        for (int r = col + 1; r < 5; r++) {
            double m = M[r][col] / M[col][col];
            for (int c = col; c < 6; c++) M[r][c] -= m * M[col][c];
        }
    }
    for (int r = 4; r >= 0; r--) {
        double s = M[r][5];
        for (int c = r + 1; c < 5; c++) s -= M[r][c] * X[c];
        X[r] = s / M[r][r];
    }
    return 0;
}

/* Moments and scaling of the loss frequency so that the discrete mass,
 * momentum and energy moments of Q vanish: before_correction
 * (collision_integral.cu:217-461), linear_solver (:464-503) and
 * correct_frequency (:506-526), for every cell.  With psi = (1, xi_x, xi_y,
 * xi_z, |xi|^2):
 *   B_m = sum_v (gain - f nu) psi_m,   A_mn = sum_v nu f psi_m psi_n,
 *   A X = B,   nu_corr = nu (1 + X.psi),   Q = gain - f nu_corr.
 * Sums are accumulated in long double.  xi is [npts][3]. */
int orc_correct(int npts, const double *xi, const double *f, const double *nu,
                const double *gain, double *nu_corr, double *Q, double *X,
                int cells)
{
    int bad = 0;
    for (int c = 0; c < cells; c++) {
        const long o = (long)c * npts;
        long double A[25] = {0}, B[5] = {0};
        for (int v = 0; v < npts; v++) {
            const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
            const double psi[5] = {1.0, x, y, z, x * x + y * y + z * z};
            const double L = nu[o + v] * f[o + v];
            const double R = gain[o + v] - L;
            for (int m = 0; m < 5; m++) {
                B[m] += (long double)R * psi[m];
                for (int q = 0; q < 5; q++) A[m * 5 + q] += (long double)L * psi[m] * psi[q];
            }
        }
        double Ad[25], Bd[5], Xc[5];
        for (int m = 0; m < 25; m++) Ad[m] = (double)A[m];
        for (int m = 0; m < 5; m++) Bd[m] = (double)B[m];
        if (orc_solve5(Ad, Bd, Xc) != 0) {
            bad++;
            for (int m = 0; m < 5; m++) Xc[m] = 0.0;
        }
        if (X)
            for (int m = 0; m < 5; m++) X[5 * c + m] = Xc[m];
        for (int v = 0; v < npts; v++) {
            const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
            const double s = 1.0 + Xc[0] + Xc[1] * x + Xc[2] * y + Xc[3] * z + Xc[4] * (x * x + y * y + z * z);
            const double nc = nu[o + v] * s;
            if (nu_corr) nu_corr[o + v] = nc;
            if (Q) Q[o + v] = gain[o + v] - f[o + v] * nc;
        }
    }
    return bad;
}

/* The same correction for an operator that yields Q and the loss term
 * L = f nu directly (the Korobov form): A_mn = sum L psi_m psi_n,
 * B_m = sum Q psi_m, Qout = Q - L (X.psi). */
int orc_correct_loss(int npts, const double *xi, const double *loss,
                     const double *Qin, double *Qout, double *X, int cells)
{
    int bad = 0;
    for (int c = 0; c < cells; c++) {
        const long o = (long)c * npts;
        long double A[25] = {0}, B[5] = {0};
        for (int v = 0; v < npts; v++) {
            const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
            const double psi[5] = {1.0, x, y, z, x * x + y * y + z * z};
            for (int m = 0; m < 5; m++) {
                B[m] += (long double)Qin[o + v] * psi[m];
                for (int q = 0; q < 5; q++) A[m * 5 + q] += (long double)loss[o + v] * psi[m] * psi[q];
            }
        }
        double Ad[25], Bd[5], Xc[5];
        for (int m = 0; m < 25; m++) Ad[m] = (double)A[m];
        for (int m = 0; m < 5; m++) Bd[m] = (double)B[m];
        if (orc_solve5(Ad, Bd, Xc) != 0) {
            bad++;
            for (int m = 0; m < 5; m++) Xc[m] = 0.0;
        }
        if (X)
            for (int m = 0; m < 5; m++) X[5 * c + m] = Xc[m];
        for (int v = 0; v < npts; v++) {
            const double x = xi[3 * v], y = xi[3 * v + 1], z = xi[3 * v + 2];
            const double s = Xc[0] + Xc[1] * x + Xc[2] * y + Xc[3] * z + Xc[4] * (x * x + y * y + z * z);
            Qout[o + v] = Qin[o + v] - loss[o + v] * s;
        }
    }
    return bad;
}

/* ======================================================================== */
/* explicit step                                                            */
/* ======================================================================== */

/* get_delta_t (collision_integral.cu:776-807): dt = min over all values of
 * courant / nu (entries with nu <= 0 are skipped). */
double orc_courant_dt(const double *nu, long count, double courant)
{
    double dt = 1.0e10;
    for (long t = 0; t < count; t++)
        if (nu[t] > 0.0 && courant / nu[t] < dt) dt = courant / nu[t];
    return dt;
}

/* evolve_f (collision_integral.cu:809-833): f <- f (1 - nu dt) + gain dt,
 * written as f <- f + dt Q with Q = gain - f nu. */
void orc_euler(double *f, const double *Q, long count, double dt)
{
    for (long t = 0; t < count; t++) f[t] = f[t] + dt * Q[t];
}

/* ======================================================================== */
/* Korobov lattice / two-point projection method                            */
/* ======================================================================== */
/*
 * No reference counterpart: the reference evaluates the gain by dense
 * enumeration.  This is the conservative projection method of Tcheremissine
 * (Comput. Math. Math. Phys. 46 (2006) 315-329): an 8-D rank-1 Korobov
 * lattice x_nu = frac(nu (1, a, a^2, ..., a^7) / M) gives, per node nu, a
 * pre-collision pair (alpha, beta) of grid nodes (3 + 3 coordinates) and a
 * unit scattering direction w (2 coordinates; hard spheres scatter
 * isotropically in the centre-of-mass frame).  The post-collision velocity
 * xi'_alpha = (xi_a + xi_b)/2 + |g|/2 w is projected onto the grid by the
 * nearest node lambda and mu = alpha + beta - lambda (momentum exact) and a
 * second pair (lambda+s, mu-s) whose energy lies on the other side of the
 * exact energy E0; r = (E0 - E_lm)/(E_s - E_lm) splits the node between the
 * two pairs, so mass, momentum and energy balance exactly per node.
 *
 * Per node and cell: P = f_a f_b,  R = (f_l f_m)^(1-r) (f_ls f_ms)^r,
 * Delta = P - R, and with the node weight w_nu = K sqrt(D2):
 *   Q[a] -= w Delta, Q[b] -= w Delta, Q[l],Q[m] += (1-r) w Delta,
 *   Q[ls],Q[ms] += r w Delta;   loss[a], loss[b] += 2 w P.
 * Normalisation from the weak form: for a test function phi,
 *   int Q phi = 1/2 int g f f1 (phi' + phi1' - phi - phi1)
 *             = 1/4 int g (f f1 - f' f1') (phi' + phi1' - phi - phi1),
 * the symmetric form counting every node's direct and inverse collision.  A
 * rank-1 lattice of M nodes over (v, v1, w) in box x box x S^2 has volume
 * weight V_box^2 4 pi s0 / M, so with V_box = n_box d^3 and g = d sqrt(D2)
 * K = pi s0 n_box^2 d^4 / M for a differential cross-section s0 per
 * steradian (s0 = 1: the reference's 4 pi g total; s0 = 1/4: unit-diameter
 * spheres); the loss f nu = 1/2 int g f f1 (...phi) takes twice that weight.
 * D2 = |i_a - i_b|^2 in grid units.
 *
 * Integer geometry: with p = 2 i + 1 - n per axis the energy of a pair is the
 * integer sum of p^2 (spacing equal on all axes), so E0, E_lm, E_s and r are
 * exact integer facts; the direction uses only +,-,*,/,sqrt and a fixed
 * polynomial for sin/cos, so the stencil is reproducible bit for bit.
 */

static const double orc_sin_c[12] = {
    1.0, -0.16666666666666666, 0.008333333333333333, -0.0001984126984126984,
    2.7557319223985893e-06, -2.505210838544172e-08, 1.6059043836821613e-10,
    -7.647163731819816e-13, 2.8114572543455206e-15, -8.22063524662433e-18,
    1.9572941063391263e-20, -3.8681701706306835e-23};
static const double orc_cos_c[12] = {
    1.0, -0.5, 0.041666666666666664, -0.001388888888888889,
    2.48015873015873e-05, -2.755731922398589e-07, 2.08767569878681e-09,
    -1.1470745597729725e-11, 4.779477332387385e-14, -1.5619206968586225e-16,
    4.110317623312165e-19, -8.896791392450574e-22};

/* sin and cos of 2 pi t for t in [0,1): quadrant from 4t (exact for dyadic
 * t), then Taylor polynomials (Horner in theta^2) on [0, pi/2). */
void orc_sincos_2pi(double t, double *s, double *c)
{
    const double t4 = 4.0 * t;
    const int q = (int)floor(t4);
    const double th = (t4 - (double)q) * ORC_HALF_PI;
    const double th2 = th * th;
    double ps = orc_sin_c[11], pc = orc_cos_c[11];
    for (int k = 10; k >= 0; k--) {
        ps = ps * th2 + orc_sin_c[k];
        pc = pc * th2 + orc_cos_c[k];
    }
    const double sv = th * ps, cv = pc;
    switch (q & 3) {
    case 0: *s = sv; *c = cv; break;
    case 1: *s = cv; *c = -sv; break;
    case 2: *s = -sv; *c = -cv; break;
    default: *s = -cv; *c = sv; break;
    }
}

/* Korobov's P_2 figure of merit of the lattice with generator a. */
double orc_korobov_p2(uint32_t M, uint32_t a, int dims)
{
    uint64_t z[8];
    z[0] = 1;
    for (int j = 1; j < dims; j++) z[j] = (z[j - 1] * a) % M;
    double sum = 0.0;
    for (uint64_t k = 0; k < M; k++) {
        double prod = 1.0;
        for (int j = 0; j < dims; j++) {
            double x = (double)((k * z[j]) % M) / (double)M;
            double b2 = x * x - x + 1.0 / 6.0;
            prod = prod * (1.0 + ORC_TWO_PI_SQ * b2);
        }
        sum = sum + prod;
    }
    return sum / (double)M - 1.0;
}

static uint64_t orc_splitmix(uint64_t *st)
{
    uint64_t z = (*st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Generator choice: ncand odd candidates drawn from splitmix64(seed); the
 * one with the smallest P_2 (ties: earliest) is kept. */
uint32_t orc_korobov_pick(uint32_t M, uint64_t seed, int ncand, double *best_p2)
{
    uint64_t st = seed;
    uint32_t best = 0;
    double bp = DBL_MAX;
    for (int c = 0; c < ncand; c++) {
        uint32_t a = (uint32_t)((orc_splitmix(&st) % (M / 2)) * 2 + 1);
        if (a < 3) a = 3;
        double p2 = orc_korobov_p2(M, a, 8);
        if (p2 < bp) {
            bp = p2;
            best = a;
        }
    }
    if (best_p2) *best_p2 = bp;
    return best;
}

static inline int orc_in_domain(const int n[3], const int *box2c, const int x[3])
{
    if (x[0] < 0 || x[0] >= n[0] || x[1] < 0 || x[1] >= n[1] || x[2] < 0 || x[2] >= n[2]) return -1;
    return box2c[(x[0] * n[1] + x[1]) * n[2] + x[2]];
}

/* Stencil of every lattice node that survives the domain tests, in node
 * order: idx6 = (alpha, beta, lambda, mu, lambda+s, mu-s) compact indices,
 * r, w = K sqrt(D2) and the lattice node number.  Returns the count (or the
 * needed count when it exceeds cap; then nothing past cap is written). */
long orc_korobov_stencil(const int n[3], const int *box2c, int npts,
                         uint32_t M, uint32_t a, double K, int32_t *idx6,
                         double *r, double *w, uint32_t *node, long cap)
{
    (void)npts;
    uint64_t z[8];
    z[0] = 1;
    for (int j = 1; j < 8; j++) z[j] = (z[j - 1] * a) % M;
    long cnt = 0;
    for (uint64_t nu = 0; nu < M; nu++) {
        uint64_t x[8];
        for (int j = 0; j < 8; j++) x[j] = (nu * z[j]) % M;
        int ia[3], ib[3];
        for (int k = 0; k < 3; k++) {
            ia[k] = (int)((x[k] * (uint64_t)n[k]) / M);
            ib[k] = (int)((x[3 + k] * (uint64_t)n[k]) / M);
        }
        const int ca = orc_in_domain(n, box2c, ia), cb = orc_in_domain(n, box2c, ib);
        if (ca < 0 || cb < 0) continue;
        const int D2 = (ia[0] - ib[0]) * (ia[0] - ib[0]) + (ia[1] - ib[1]) * (ia[1] - ib[1]) +
                       (ia[2] - ib[2]) * (ia[2] - ib[2]);
        if (D2 == 0) continue;

        const double t7 = ((double)x[6] + 0.5) / (double)M;
        const double t8 = ((double)x[7] + 0.5) / (double)M;
        double sn, cs;
        orc_sincos_2pi(t8, &sn, &cs);
        const double rho = 2.0 * sqrt(t7 * (1.0 - t7));
        const double om[3] = {rho * cs, rho * sn, 1.0 - 2.0 * t7};
        const double h = 0.5 * sqrt((double)D2);

        int lam[3], mu[3];
        double del[3];
        for (int k = 0; k < 3; k++) {
            const double xp = 0.5 * (double)(ia[k] + ib[k]) + h * om[k];
            lam[k] = (int)floor(xp + 0.5);
            del[k] = xp - (double)lam[k];
            mu[k] = ia[k] + ib[k] - lam[k];
        }
        const int cl = orc_in_domain(n, box2c, lam), cm = orc_in_domain(n, box2c, mu);
        if (cl < 0 || cm < 0) continue;

        long E0 = 0, El = 0;
        for (int k = 0; k < 3; k++) {
            const long pa = 2 * ia[k] + 1 - n[k], pb = 2 * ib[k] + 1 - n[k];
            const long pl = 2 * lam[k] + 1 - n[k], pm = 2 * mu[k] + 1 - n[k];
            E0 += pa * pa + pb * pb;
            El += pl * pl + pm * pm;
        }
        int cls = cl, cms = cm;
        double rr = 0.0;
        if (El != E0) {
            int best = -1;
            double bestd = DBL_MAX;
            long bestE = 0;
            int bl = -1, bm = -1;
            for (int S = 1; S < 8; S++) {
                int ls[3], ms[3];
                double d2 = 0.0;
                for (int k = 0; k < 3; k++) {
                    const int sk = ((S >> k) & 1) ? (del[k] >= 0.0 ? 1 : -1) : 0;
                    ls[k] = lam[k] + sk;
                    ms[k] = mu[k] - sk;
                    const double dk = sk ? 1.0 - fabs(del[k]) : del[k];
                    d2 = d2 + dk * dk;
                }
                const int c1 = orc_in_domain(n, box2c, ls), c2 = orc_in_domain(n, box2c, ms);
                if (c1 < 0 || c2 < 0) continue;
                long Es = 0;
                for (int k = 0; k < 3; k++) {
                    const long p1 = 2 * ls[k] + 1 - n[k], p2 = 2 * ms[k] + 1 - n[k];
                    Es += p1 * p1 + p2 * p2;
                }
                const int other = (El < E0) ? (Es >= E0) : (Es <= E0);
                if (other && d2 < bestd) {
                    bestd = d2;
                    best = S;
                    bestE = Es;
                    bl = c1;
                    bm = c2;
                }
            }
            if (best < 0) continue;
            cls = bl;
            cms = bm;
            rr = (double)(E0 - El) / (double)(bestE - El);
        }
        if (cnt < cap) {
            int32_t *o = idx6 + 6 * cnt;
            o[0] = ca; o[1] = cb; o[2] = cl; o[3] = cm; o[4] = cls; o[5] = cms;
            r[cnt] = rr;
            w[cnt] = K * sqrt((double)D2);
            if (node) node[cnt] = (uint32_t)nu;
        }
        cnt++;
    }
    return cnt;
}

/* Loss and Q of the projection method for `cells` cells of compact f. */
void orc_korobov_collide(int npts, long nn, const int32_t *idx6,
                         const double *r, const double *w,
                         const double *f_cmp, double *loss, double *Q,
                         int cells)
{
    double *val = malloc(sizeof(double) * 2 * (nn > 0 ? nn : 1));
    for (int c = 0; c < cells; c++) {
        const double *f = f_cmp + (long)c * npts;
        double *L = loss + (long)c * npts, *q = Q + (long)c * npts;
#pragma omp parallel for schedule(static)
        for (long t = 0; t < nn; t++) {
            const int32_t *o = idx6 + 6 * t;
            const double P = f[o[0]] * f[o[1]];
            const double Pl = f[o[2]] * f[o[3]];
            const double rr = r[t];
            double R;
            if (rr == 0.0) {
                R = Pl;
            } else {
                const double Ps = f[o[4]] * f[o[5]];
                if (Pl > 0.0 && Ps > 0.0)
                    R = Pl * exp(rr * log(Ps / Pl));
                else
                    R = (1.0 - rr) * Pl + rr * Ps;
            }
            val[2 * t] = w[t] * (P - R);
            val[2 * t + 1] = 2.0 * (w[t] * P);
        }
        for (int v = 0; v < npts; v++) L[v] = q[v] = 0.0;
        for (long t = 0; t < nn; t++) {
            const int32_t *o = idx6 + 6 * t;
            const double wd = val[2 * t], wp = val[2 * t + 1], rr = r[t];
            q[o[0]] -= wd;
            q[o[1]] -= wd;
            q[o[2]] += (1.0 - rr) * wd;
            q[o[3]] += (1.0 - rr) * wd;
            q[o[4]] += rr * wd;
            q[o[5]] += rr * wd;
            L[o[0]] += wp;
            L[o[1]] += wp;
        }
    }
    free(val);
}

/* Destination lists of a stencil, for the CPU baseline's parallel form of
 * orc_korobov_collide.  For every compact velocity v, the writes of the
 * serial scatter above to q[v] and to loss[v], in the order it makes them
 * (node t ascending, slot k = 0..5 within a node): qnode[e] = t with
 * qcoef[e] = -1 (k = 0, 1), 1 - r[t] (k = 2, 3) or r[t] (k = 4, 5), for q;
 * lnode[e] = t (k = 0, 1) for the loss.  qrow and lrow have npts + 1
 * entries; qnode/qcoef hold 6 nn entries, lnode 2 nn.  Built once per
 * stencil. */
void orc_korobov_dest_lists(int npts, long nn, const int32_t *idx6, const double *r,
                            long *qrow, int32_t *qnode, double *qcoef, long *lrow, int32_t *lnode)
{
    for (int v = 0; v <= npts; v++) qrow[v] = lrow[v] = 0;
    for (long e = 0; e < 6 * nn; e++) {
        qrow[idx6[e] + 1]++;
        if (e % 6 < 2) lrow[idx6[e] + 1]++;
    }
    for (int v = 0; v < npts; v++) {
        qrow[v + 1] += qrow[v];
        lrow[v + 1] += lrow[v];
    }
    long *fq = malloc(sizeof(long) * 2 * (npts > 0 ? npts : 1)), *fl = fq + npts;
    for (int v = 0; v < npts; v++) {
        fq[v] = qrow[v];
        fl[v] = lrow[v];
    }
    for (long t = 0; t < nn; t++) {
        for (int k = 0; k < 6; k++) {
            const int v = idx6[6 * t + k];
            const long e = fq[v]++;
            qnode[e] = (int32_t)t;
            qcoef[e] = k < 2 ? -1.0 : (k < 4 ? 1.0 - r[t] : r[t]);
            if (k < 2) lnode[fl[v]++] = (int32_t)t;
        }
    }
    free(fq);
}

/* orc_korobov_collide with the scatter turned into a gather over
 * destinations (OpenMP over v) through orc_korobov_dest_lists.  Each
 * destination adds its writes in the serial scatter's order with the same
 * operations (q - wd is q + (-1) wd exactly; the (1 - r) and r factors are
 * the same products), so loss and Q are bitwise those of
 * orc_korobov_collide. */
void orc_korobov_collide_dest(int npts, long nn, const int32_t *idx6,
                              const double *r, const double *w,
                              const long *qrow, const int32_t *qnode, const double *qcoef,
                              const long *lrow, const int32_t *lnode,
                              const double *f_cmp, double *loss, double *Q,
                              int cells)
{
    double *val = malloc(sizeof(double) * 2 * (nn > 0 ? nn : 1));
    for (int c = 0; c < cells; c++) {
        const double *f = f_cmp + (long)c * npts;
        double *L = loss + (long)c * npts, *q = Q + (long)c * npts;
#pragma omp parallel for schedule(static)
        for (long t = 0; t < nn; t++) {
            const int32_t *o = idx6 + 6 * t;
            const double P = f[o[0]] * f[o[1]];
            const double Pl = f[o[2]] * f[o[3]];
            const double rr = r[t];
            double R;
            if (rr == 0.0) {
                R = Pl;
            } else {
                const double Ps = f[o[4]] * f[o[5]];
                if (Pl > 0.0 && Ps > 0.0)
                    R = Pl * exp(rr * log(Ps / Pl));
                else
                    R = (1.0 - rr) * Pl + rr * Ps;
            }
            val[2 * t] = w[t] * (P - R);
            val[2 * t + 1] = 2.0 * (w[t] * P);
        }
#pragma omp parallel for schedule(dynamic, 64)
        for (int v = 0; v < npts; v++) {
            double qa = 0.0, la = 0.0;
            for (long e = qrow[v]; e < qrow[v + 1]; e++) qa += qcoef[e] * val[2 * (long)qnode[e]];
            for (long e = lrow[v]; e < lrow[v + 1]; e++) la += val[2 * (long)lnode[e] + 1];
            q[v] = qa;
            L[v] = la;
        }
    }
    free(val);
}
