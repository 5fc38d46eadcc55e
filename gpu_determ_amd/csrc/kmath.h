// Table-driven FP64 log and exp for the projection weight R = Pl (Ps/Pl)^r of
// the Korobov node terms, after P. T. P. Tang, "Table-driven implementation
// of the exponential function in IEEE floating-point arithmetic", ACM TOMS 15
// (1989) 144-157, and "Table-driven implementation of the logarithm function
// in IEEE floating-point arithmetic", ACM TOMS 16 (1990) 378-400.
//
// The tables (KMT_SIZE doubles) are built once on the host in extended
// precision (kmath_tables) and read from LDS by the kernels:
//   log x = (e - 1) ln 2 - log c_j + log1p(t),  t = 2m c_j - 1,  |t| <= 2^-8,
//           x = m 2^e (m in [1/2, 1)), j = the top 7 bits of 2m's mantissa,
//           c_j = RN(1 / (1 + (j + 1/2) / 128)), log1p by its Taylor series
//           to t^7 (truncation below 2^-59 relative);
//   exp y = 2^(k >> 6) 2^((k & 63) / 64) (1 + expm1(R)),  k = rint(64 y / ln 2),
//           R = y - k ln2 / 64 (two-part constant), |R| <= ln2 / 128, expm1
//           by its Taylor series to R^5 (truncation below 2^-54).
// Both are within about one ulp (exp) and two ulps (log) over the normal
// range; the accuracy hook gdt_kmath_pow is checked against numpy's pow.
#pragma once

#define KMT_LOG_N 128
#define KMT_EXP_N 64
#define KMT_C 0      // c_j, j < 128
#define KMT_LHI 128  // -log(c_j), high part
#define KMT_LLO 256  // -log(c_j), low part
#define KMT_SHI 384  // 2^(j/64), high part, j < 64
#define KMT_SLO 448  // 2^(j/64), low part
#define KMT_K 512    // ln2 (33-bit high part, low part), 64 / ln2, ln2 / 64 (33-bit high part, low part)
#define KMT_SIZE 520

__device__ __forceinline__ double kmt_log(double x, const double *T)
{
    int e;
    const double m2 = 2.0 * frexp(x, &e);        // [1, 2), exact
    int j = (int)((m2 - 1.0) * 128.0);           // exact product, truncated: 0..127 for
    j = j < 0 ? 0 : (j > 127 ? 127 : j);         // finite x > 0 (clamped: table reads stay in bounds)
    const double t = fma(m2, T[KMT_C + j], -1.0);
    const double q = fma(t, fma(t, fma(t, fma(t, fma(t, 1.0 / 7.0, -1.0 / 6.0), 0.2), -0.25), 1.0 / 3.0), -0.5);
    const double lp = fma(t * t, q, t);          // log1p(t)
    const double ef = (double)(e - 1);
    return fma(ef, T[KMT_K], T[KMT_LHI + j]) + (fma(ef, T[KMT_K + 1], T[KMT_LLO + j]) + lp);
}

__device__ __forceinline__ double kmt_exp(double y, const double *T)
{
    const double kf = rint(y * T[KMT_K + 2]);
    const int k = (int)kf;
    const double R = fma(-kf, T[KMT_K + 4], fma(-kf, T[KMT_K + 3], y));
    const double p = fma(R * R, fma(R, fma(R, fma(R, 1.0 / 120.0, 1.0 / 24.0), 1.0 / 6.0), 0.5), R);  // expm1(R)
    const int j = k & 63;
    const double s = T[KMT_SHI + j];
    return ldexp(s + fma(s, p, T[KMT_SLO + j]), k >> 6);
}
