// Device helpers shared by the collision kernels (wave64, FP64).
#pragma once

#include <hip/hip_runtime.h>

#define WAVE 64

// Lane exchange within rows of 16 lanes through DPP (a VALU source modifier:
// no LDS round trip, unlike ds_bpermute).  Doubles move as two 32-bit halves.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// DPP exchange that leaves `old` in lanes without a source lane (row_shr at
// the row start, wave_shl at lane 63) and in rows outside ROWMASK.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ int dpp_i32_or(int v, int old)
{
    return __builtin_amdgcn_update_dpp(old, v, CTRL, ROWMASK, 0xF, false);
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_f64_or(double v, double old)
{
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROWMASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

#define DPP_ROW_SHR(k) (0x110 + (k))  // lane i of a row reads lane i - k of the row
#define DPP_WAVE_SHL1 0x130           // lane i reads lane i + 1
#define DPP_ROW_BCAST15 0x142         // lane 15 of each row feeds the next row
#define DPP_ROW_BCAST31 0x143         // lane 31 feeds rows 2 and 3
#define DPP_QUAD_XOR1 0xB1       // quad_perm [1,0,3,2]
#define DPP_QUAD_XOR2 0x4E       // quad_perm [2,3,0,1]
#define DPP_ROW_HALF_MIRROR 0x141
#define DPP_ROW_MIRROR 0x140

// Sum over each row of 16 lanes; every lane of the row gets the row's sum.
// The pairing order is fixed, so the result is reproducible.
__device__ __forceinline__ double row16_sum(double v)
{
    v += dpp_f64<DPP_QUAD_XOR1>(v);
    v += dpp_f64<DPP_QUAD_XOR2>(v);
    v += dpp_f64<DPP_ROW_HALF_MIRROR>(v);
    v += dpp_f64<DPP_ROW_MIRROR>(v);
    return v;
}

template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_step(double &q, double &l, int key)
{
    const int ks = dpp_i32_or<CTRL, ROWMASK>(key, -1);
    const double qs = dpp_f64_or<CTRL, ROWMASK>(q, 0.0);
    const double ls = dpp_f64_or<CTRL, ROWMASK>(l, 0.0);
    if (ks == key) {
        q += qs;
        l += ls;
    }
}

// Inclusive segmented scan of (q, l) over the 64 lanes: afterwards every lane
// holds the sum over the lanes up to it that share its key.  Keys of valid
// lanes are >= 0 and non-decreasing with the lane (runs are contiguous);
// other lanes carry distinct keys < -1.  Row steps (shift 1, 2, 4, 8) and the
// two cross-row broadcasts are DPP modifiers (no LDS traffic); the pairing
// order is fixed, so sums are reproducible.
__device__ __forceinline__ void seg_scan_dpp(double &q, double &l, int key)
{
    seg_step<DPP_ROW_SHR(1), 0xF>(q, l, key);
    seg_step<DPP_ROW_SHR(2), 0xF>(q, l, key);
    seg_step<DPP_ROW_SHR(4), 0xF>(q, l, key);
    seg_step<DPP_ROW_SHR(8), 0xF>(q, l, key);
    seg_step<DPP_ROW_BCAST15, 0xA>(q, l, key);
    seg_step<DPP_ROW_BCAST31, 0xC>(q, l, key);
}

// The same scan for one value.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ void seg_step1(double &u, int key)
{
    const int ks = dpp_i32_or<CTRL, ROWMASK>(key, -1);
    const double us = dpp_f64_or<CTRL, ROWMASK>(u, 0.0);
    if (ks == key) u += us;
}

__device__ __forceinline__ void seg_scan1_dpp(double &u, int key)
{
    seg_step1<DPP_ROW_SHR(1), 0xF>(u, key);
    seg_step1<DPP_ROW_SHR(2), 0xF>(u, key);
    seg_step1<DPP_ROW_SHR(4), 0xF>(u, key);
    seg_step1<DPP_ROW_SHR(8), 0xF>(u, key);
    seg_step1<DPP_ROW_BCAST15, 0xA>(u, key);
    seg_step1<DPP_ROW_BCAST31, 0xC>(u, key);
}

__device__ __forceinline__ double readlane_f64(double v, int lane)
{
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes of the wave (every lane gets it): rows by DPP, the
// four row sums by readlane, added as (r0 + r1) + (r2 + r3).
__device__ __forceinline__ double wave_sum(double v)
{
    v = row16_sum(v);
    return (readlane_f64(v, 0) + readlane_f64(v, 16)) + (readlane_f64(v, 32) + readlane_f64(v, 48));
}

// Wave sums of N independent values, stage by stage so the chains overlap.
template <int N>
__device__ __forceinline__ void wave_sum_n(double (&m)[N])
{
#pragma unroll
    for (int q = 0; q < N; q++) m[q] += dpp_f64<DPP_QUAD_XOR1>(m[q]);
#pragma unroll
    for (int q = 0; q < N; q++) m[q] += dpp_f64<DPP_QUAD_XOR2>(m[q]);
#pragma unroll
    for (int q = 0; q < N; q++) m[q] += dpp_f64<DPP_ROW_HALF_MIRROR>(m[q]);
#pragma unroll
    for (int q = 0; q < N; q++) m[q] += dpp_f64<DPP_ROW_MIRROR>(m[q]);
#pragma unroll
    for (int q = 0; q < N; q++)
        m[q] = (readlane_f64(m[q], 0) + readlane_f64(m[q], 16)) + (readlane_f64(m[q], 32) + readlane_f64(m[q], 48));
}

// 5x5 Gaussian elimination with partial pivoting, fully unrolled so that the
// matrix lives in registers (row swaps are selects).  Same pivoting and
// elimination steps as the oracle's orc_solve5.  A zero pivot gives X = 0
// and false.
__device__ __forceinline__ bool solve5(double (&M)[5][6], double (&X)[5])
{
    bool ok = true;
#pragma unroll
    for (int col = 0; col < 5; col++) {
        int piv = col;
        double best = fabs(M[col][col]);
#pragma unroll
        for (int r = col + 1; r < 5; r++) {
            const double a = fabs(M[r][col]);
            if (a > best) {
                best = a;
                piv = r;
            }
        }
        if (best == 0.0) ok = false;
#pragma unroll
        for (int r = col + 1; r < 5; r++) {
            const bool sw = (piv == r);
#pragma unroll
            for (int cc = 0; cc < 6; cc++) {
                const double a = M[col][cc], b = M[r][cc];
                M[col][cc] = sw ? b : a;
                M[r][cc] = sw ? a : b;
            }
        }
#pragma unroll
        for (int r = col + 1; r < 5; r++) {
            const double mm = M[r][col] / M[col][col];
#pragma unroll
            for (int cc = col; cc < 6; cc++) M[r][cc] -= mm * M[col][cc];
        }
    }
#pragma unroll
    for (int r = 4; r >= 0; r--) {
        double sm = M[r][5];
#pragma unroll
        for (int cc = r + 1; cc < 5; cc++) sm -= M[r][cc] * X[cc];
        X[r] = sm / M[r][r];
    }
    if (!ok)
#pragma unroll
        for (int a = 0; a < 5; a++) X[a] = 0.0;
    return ok;
}

// 1 / sqrt(d) for d > 0 to FP64 accuracy: the hardware estimate and two
// Newton steps y <- y (3/2 - d y^2 / 2).
__device__ __forceinline__ double rsqrt_nr(double d)
{
    const double h = 0.5 * d;
    double y = __builtin_amdgcn_rsq(d);
    y = y * fma(-h * y, y, 1.5);
    y = y * fma(-h * y, y, 1.5);
    return y;
}

// Cholesky solve of the symmetric positive definite 5x5 system (moment
// matrices sum L psi psi^T with L >= 0): short dependency chains, registers
// only, the pivots' inverse square roots by rsqrt_nr (no division or square
// root on the chain).  A non-positive pivot (not positive definite, e.g.
// negative f) falls back to the pivoted elimination.
__device__ __forceinline__ void solve5_spd(const double (&A)[5][5], const double (&B)[5], double (&X)[5])
{
    double Lm[5][5], inv[5];
    bool spd = true;
#pragma unroll
    for (int j = 0; j < 5; j++) {
        double d = A[j][j];
#pragma unroll
        for (int k = 0; k < j; k++) d -= Lm[j][k] * Lm[j][k];
        spd = spd && (d > 0.0);
        inv[j] = rsqrt_nr(d > 0.0 ? d : 1.0);
#pragma unroll
        for (int i = j + 1; i < 5; i++) {
            double v = A[i][j];
#pragma unroll
            for (int k = 0; k < j; k++) v -= Lm[i][k] * Lm[j][k];
            Lm[i][j] = v * inv[j];
        }
    }
    if (spd) {
        double y[5];
#pragma unroll
        for (int i = 0; i < 5; i++) {
            double v = B[i];
#pragma unroll
            for (int k = 0; k < i; k++) v -= Lm[i][k] * y[k];
            y[i] = v * inv[i];
        }
#pragma unroll
        for (int i = 4; i >= 0; i--) {
            double v = y[i];
#pragma unroll
            for (int k = i + 1; k < 5; k++) v -= Lm[k][i] * X[k];
            X[i] = v * inv[i];
        }
        return;
    }
    double M[5][6];
#pragma unroll
    for (int i = 0; i < 5; i++) {
#pragma unroll
        for (int k = 0; k < 5; k++) M[i][k] = A[i][k];
        M[i][5] = B[i];
    }
    solve5(M, X);
}

// The 20 moment sums of a cell: s[0..4] = sum Q psi_m (right-hand side),
// s[5..19] = sum L psi_m psi_n for m <= n row by row; builds the symmetric
// system and solves for X.
__device__ __forceinline__ void moments_solve(const double (&s)[20], double (&X)[5])
{
    double A[5][5], B[5];
    int q = 5;
#pragma unroll
    for (int a = 0; a < 5; a++)
#pragma unroll
        for (int b = a; b < 5; b++) {
            A[a][b] = s[q];
            A[b][a] = s[q];
            q++;
        }
#pragma unroll
    for (int a = 0; a < 5; a++) B[a] = s[a];
    solve5_spd(A, B, X);
}

// Adds the moment contributions of a destination at velocity (x0, y0, z0)
// (Q value q, loss value l).
__device__ __forceinline__ void moments_add_psi(double (&m)[20], double x0, double y0, double z0, double q, double l)
{
    const double psi[5] = {1.0, x0, y0, z0, x0 * x0 + y0 * y0 + z0 * z0};
#pragma unroll
    for (int a = 0; a < 5; a++) m[a] += q * psi[a];
    int k = 5;
#pragma unroll
    for (int a = 0; a < 5; a++)
#pragma unroll
        for (int b = a; b < 5; b++) m[k++] += l * psi[a] * psi[b];
}

// moments_add_psi for destination g of the node list xi.
__device__ __forceinline__ void moments_add(double (&m)[20], const double *xi, int g, double q, double l)
{
    moments_add_psi(m, xi[3 * g], xi[3 * g + 1], xi[3 * g + 2], q, l);
}
