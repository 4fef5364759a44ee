// dvc_math.h -- transcendentals that return, for every input, the same bits as
// the host C library the reference decoder links (glibc 2.35 on x86-64, whose
// expf/logf/exp are dispatched to their FMA variants on FMA-capable CPUs, and
// whose tanhf is the fdlibm routine built on expm1f).
//
// The reference evaluates these at:
//   DvcDec/src/LdpcaDecoder.cpp:513,519   -log(tanh((x + 1e-9f)/2))  (float overloads)
//   DvcDec/src/LdpcaDecoder.cpp:666,135   log(LR), log(Pco)          (logf)
//   DvcDec/src/noiseModel.cpp:93,126      exp(-alpha*|k-SI|)          (expf)
//   DvcDec/src/motionLearning.cpp:111     exp(-MU*SSE)                (expf)
//   DvcDec/src/reconstruction.cpp:211-277 exp(alpha*step), exp(-alpha*r) (double exp)
//
// Algorithms (published methods):
//   logf  : ARM optimized-routines logf (table of 16 (1/c, log c) pairs, c the
//           centre of each sub-interval of [0x1.66p-1, 0x1.66p0); degree-4
//           polynomial for log1p(r)), evaluated in double.
//   expf  : ARM optimized-routines expf (2^(i/32) table, cubic for 2^(r/32)),
//           evaluated in double with the fused multiply-adds of the FMA build.
//   tanhf : fdlibm s_tanhf.c via fdlibm s_expm1f.c, pure single precision.
// The logf/expf tables were regenerated from their definitions with 200-bit
// arithmetic; every function was compared against the host libm over all 2^32
// float inputs with zero mismatches (tests/test_math_exhaustive.py repeats the
// comparison on a strided subset and on the full BP domain).
//
// Device code using this header must be compiled with -ffp-contract=off and
// IEEE-correct float division (hipcc's default), because the single-precision
// steps of expm1f/tanhf must round exactly as the scalar x86 code does.
#pragma once
#include <stdint.h>
#include <string.h>
#include <math.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define DVC_MATH_FN __device__ __forceinline__
#define DVC_MATH_TABLE __constant__
#else
#define DVC_MATH_FN static inline
#define DVC_MATH_TABLE static const
#endif

DVC_MATH_FN uint32_t dvcm_asu(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
DVC_MATH_FN float dvcm_asf(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
DVC_MATH_FN uint64_t dvcm_asu64(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
DVC_MATH_FN double dvcm_asd(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

// ---------------------------------------------------------------- logf
// Sub-interval i (0..15) holds the floats whose bits minus 0x3f330000 have
// bits 22..19 equal to i.  invc = 1/c rounded to double, c the midpoint of the
// first and last float of the sub-interval; the interval around 1.0 (i = 9)
// uses c = 1 and the one above it (i = 10) a shifted c; logc = -log(invc).
DVC_MATH_TABLE double dvcm_logf_invc[16] = {
    0x1.661ec79f8f3bep+0, 0x1.571ed4aaf883dp+0, 0x1.49539f0f010b0p+0, 0x1.3c995b0b80385p+0,
    0x1.30d190c8864a5p+0, 0x1.25e227b0b8ea0p+0, 0x1.1bb4a4a1a343fp+0, 0x1.12358f08ae5bap+0,
    0x1.0953f419900a7p+0, 0x1p+0,               0x1.e608ce2f1958dp-1, 0x1.ca4b31f026aa0p-1,
    0x1.b2036576afce6p-1, 0x1.9c2d163a1aa2dp-1, 0x1.886e6037841edp-1, 0x1.767dcf5534862p-1};
DVC_MATH_TABLE double dvcm_logf_logc[16] = {
    -0x1.57bf7808caadep-2, -0x1.2bef0a7c06ddbp-2, -0x1.01eae7f513a67p-2, -0x1.b31d8a68224e9p-3,
    -0x1.6574f0ac07758p-3, -0x1.1aa2bc79c8100p-3, -0x1.a4e76ce8c0e5ep-4, -0x1.1973c5a611cccp-4,
    -0x1.252f438e10c1ep-5, 0x0p+0,                0x1.aa5ac1f470d98p-5,  0x1.c5e53aa362eb4p-4,
    0x1.526e57720db08p-3,  0x1.bc2860d224770p-3,  0x1.1058bc8a07ee1p-2,  0x1.4043057b6ee09p-2};

DVC_MATH_FN float dvc_logf(float x) {
  const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
  const double Ln2 = 0x1.62e42fefa39efp-1;
  uint32_t ix = dvcm_asu(x);
  if (ix == 0x3f800000u) return 0.0f;
  if (ix - 0x00800000u >= 0x7f800000u - 0x00800000u) {
    if (ix * 2u == 0u) return -INFINITY;               // log(+-0) = -inf
    if (ix == 0x7f800000u) return x;                   // log(inf) = inf
    if ((ix & 0x80000000u) || ix * 2u >= 0xff000000u)  // negative or NaN
      return (x - x) / (x - x);
    ix = dvcm_asu(x * 0x1p23f);                        // subnormal: normalise
    ix -= 23u << 23;
  }
  uint32_t tmp = ix - 0x3f330000u;
  int i = (int)((tmp >> 19) % 16u);
  int k = (int32_t)tmp >> 23;
  uint32_t iz = ix - (tmp & 0xff800000u);
  double invc = dvcm_logf_invc[i], logc = dvcm_logf_logc[i];
  double z = (double)dvcm_asf(iz);
  double r = fma(z, invc, -1.0);
  double y0 = fma((double)k, Ln2, logc);
  double r2 = r * r;
  double y = fma(A1, r, A2);
  y = fma(A0, r2, y);
  y = fma(y, r2, y0 + r);
  return (float)y;
}

// ---------------------------------------------------------------- expf
// tab[i] = bits(2^(i/32)) - (i << 52)/32
DVC_MATH_TABLE uint64_t dvcm_exp2f_tab[32] = {
    0x3ff0000000000000ull, 0x3fefd9b0d3158574ull, 0x3fefb5586cf9890full, 0x3fef9301d0125b51ull,
    0x3fef72b83c7d517bull, 0x3fef54873168b9aaull, 0x3fef387a6e756238ull, 0x3fef1e9df51fdee1ull,
    0x3fef06fe0a31b715ull, 0x3feef1a7373aa9cbull, 0x3feedea64c123422ull, 0x3feece086061892dull,
    0x3feebfdad5362a27ull, 0x3feeb42b569d4f82ull, 0x3feeab07dd485429ull, 0x3feea47eb03a5585ull,
    0x3feea09e667f3bcdull, 0x3fee9f75e8ec5f74ull, 0x3feea11473eb0187ull, 0x3feea589994cce13ull,
    0x3feeace5422aa0dbull, 0x3feeb737b0cdc5e5ull, 0x3feec49182a3f090ull, 0x3feed503b23e255dull,
    0x3feee89f995ad3adull, 0x3feeff76f2fb5e47ull, 0x3fef199bdd85529cull, 0x3fef3720dcef9069ull,
    0x3fef5818dcfba487ull, 0x3fef7c97337b9b5full, 0x3fefa4afa2a490daull, 0x3fefd0765b6e4540ull};

DVC_MATH_FN float dvc_expf(float x) {
  const double InvLn2N = 0x1.71547652b82fep+0 * 32;
  const double C0 = 0x1.c6af84b912394p-5 / 32 / 32 / 32, C1 = 0x1.ebfce50fac4f3p-3 / 32 / 32,
               C2 = 0x1.62e42ff0c52d6p-1 / 32;
  const double SHIFT = 0x1.8p+52;
  uint32_t abstop = (dvcm_asu(x) >> 20) & 0x7ffu;
  if (abstop >= 0x42bu) {  // |x| >= 88 or NaN (top12(88.0f) = 0x42b)
    if (dvcm_asu(x) == 0xff800000u) return 0.0f;   // exp(-inf) = 0
    if (abstop >= 0x7f8u) return x + x;            // inf or NaN
    if (x > 0x1.62e42ep6f) return INFINITY;        // overflow
    if (x < -0x1.9fe368p6f) return 0.0f;           // underflow
  }
  double xd = (double)x;
  double kd = fma(InvLn2N, xd, SHIFT);
  uint64_t ki = dvcm_asu64(kd);
  kd -= SHIFT;
  double r = fma(InvLn2N, xd, -kd);
  uint64_t t = dvcm_exp2f_tab[ki % 32];
  t += ki << (52 - 5);
  double s = dvcm_asd(t);
  double z = fma(C0, r, C1);
  double r2 = r * r;
  double y = fma(C2, r, 1.0);
  y = fma(z, r2, y);
  y = y * s;
  return (float)y;
}

// ---------------------------------------------------------------- tanhf
#if defined(__clang__)
#define DVCM_NOCONTRACT _Pragma("clang fp contract(off)")
#else
#define DVCM_NOCONTRACT
#endif

DVC_MATH_FN float dvc_expm1f(float x) {
  DVCM_NOCONTRACT
  const float one = 1.0f, huge = 1.0e+30f, tiny = 1.0e-30f, o_threshold = 8.8721679688e+01f,
              ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f, invln2 = 1.4426950216e+00f,
              Q1 = -3.3333335072e-02f, Q2 = 1.5873016091e-03f, Q3 = -7.9365076090e-05f,
              Q4 = 4.0082177293e-06f, Q5 = -2.0109921195e-07f;
  float y, hi, lo, c = 0.0f, t, e, hxs, hfx, r1;
  int32_t k;
  uint32_t hx = dvcm_asu(x);
  uint32_t xsb = hx & 0x80000000u;
  hx &= 0x7fffffffu;
  if (hx >= 0x4195b844u) {            // |x| >= 27*ln2
    if (hx >= 0x42b17218u) {          // |x| >= 88.72...
      if (hx > 0x7f800000u) return x + x;
      if (hx == 0x7f800000u) return (xsb == 0) ? x : -1.0f;
      if (x > o_threshold) return huge * huge;
    }
    if (xsb != 0) return tiny - one;  // x < -27*ln2
  }
  if (hx > 0x3eb17218u) {             // |x| > 0.5*ln2
    if (hx < 0x3F851592u) {           // and |x| < 1.5*ln2
      if (xsb == 0) { hi = x - ln2_hi; lo = ln2_lo; k = 1; }
      else { hi = x + ln2_hi; lo = -ln2_lo; k = -1; }
    } else {
      k = (int32_t)(invln2 * x + ((xsb == 0) ? 0.5f : -0.5f));
      t = (float)k;
      hi = x - t * ln2_hi;
      lo = t * ln2_lo;
    }
    x = hi - lo;
    c = (hi - x) - lo;
  } else if (hx < 0x33000000u) {      // |x| < 2^-25
    t = huge + x;
    return x - (t - (huge + x));
  } else {
    k = 0;
  }
  hfx = 0.5f * x;
  hxs = x * hfx;
  r1 = one + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * (Q4 + hxs * Q5))));
  t = 3.0f - r1 * hfx;
  e = hxs * ((r1 - t) / (6.0f - x * t));
  if (k == 0) return x - (x * e - hxs);
  e = (x * (e - c) - c);
  e -= hxs;
  if (k == -1) return 0.5f * (x - e) - 0.5f;
  if (k == 1) {
    if (x < -0.25f) return -2.0f * (e - (x + 0.5f));
    return one + 2.0f * (x - e);
  }
  if (k <= -2 || k > 56) {
    y = one - (e - x);
    y = dvcm_asf(dvcm_asu(y) + ((uint32_t)k << 23));
    return y - one;
  }
  if (k < 23) {
    t = dvcm_asf(0x3f800000u - (0x1000000u >> k));
    y = t - (e - x);
    y = dvcm_asf(dvcm_asu(y) + ((uint32_t)k << 23));
  } else {
    t = dvcm_asf((uint32_t)(0x7f - k) << 23);
    y = x - (e + t);
    y += one;
    y = dvcm_asf(dvcm_asu(y) + ((uint32_t)k << 23));
  }
  return y;
}

DVC_MATH_FN float dvc_tanhf(float x) {
  DVCM_NOCONTRACT
  const float one = 1.0f, tiny = 1.0e-30f;
  float t, z;
  int32_t jx = (int32_t)dvcm_asu(x);
  int32_t ix = jx & 0x7fffffff;
  if (ix >= 0x7f800000) return (jx >= 0) ? one / x + one : one / x - one;
  if (ix < 0x41b00000) {           // |x| < 22
    if (ix == 0) return x;
    if (ix < 0x24000000) return x * (one + x);   // |x| < 2^-55
    if (ix >= 0x3f800000) {        // |x| >= 1
      t = dvc_expm1f(2.0f * fabsf(x));
      z = one - 2.0f / (t + 2.0f);
    } else {
      t = dvc_expm1f(-2.0f * fabsf(x));
      z = -t / (t + 2.0f);
    }
  } else {
    z = one - tiny;
  }
  return (jx >= 0) ? z : -z;
}

// The check-node kernel of the reference's sum-product decoder,
//   phi(x) = -log(tanh((x + 1e-9f) / 2))       (DvcDec/src/LdpcaDecoder.cpp:513,519)
// with the same single-precision steps.
DVC_MATH_FN float dvc_bp_phi(float x) {
  DVCM_NOCONTRACT
  return -dvc_logf(dvc_tanhf((x + (float)0.000000001) / 2.0f));
}

// The general path of dvc_logf for positive normal finite x, reading the
// (invc, logc) table through `tab` (tab[i] = invc[i], tab[16 + i] = logc[i])
// so that a kernel can keep it in LDS.  x = 1.0f gives +0 here as well.
DVC_MATH_FN float dvc_logf_normal(float x, const double *tab) {
  const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
  const double Ln2 = 0x1.62e42fefa39efp-1;
  const uint32_t ix = dvcm_asu(x);
  const uint32_t tmp = ix - 0x3f330000u;
  const int i = (int)((tmp >> 19) % 16u);
  const int k = (int32_t)tmp >> 23;
  const uint32_t iz = ix - (tmp & 0xff800000u);
  const double invc = tab[i], logc = tab[16 + i];
  const double z = (double)dvcm_asf(iz);
  const double r = fma(z, invc, -1.0);
  const double y0 = fma((double)k, Ln2, logc);
  const double r2 = r * r;
  double y = fma(A1, r, A2);
  y = fma(A0, r2, y);
  y = fma(y, r2, y0 + r);
  return (float)y;
}

// dvc_logf_normal for the arguments phi reaches, z = tanhf(y) with y >=
// 5e-10, so z >= 2^-31 and k = (bits(z) - 0x3f330000) >> 23 lies in [-31, 0].
// Everything of the general path that depends on (k, i) only comes from one
// 16-byte table entry e = (k + 31) * 16 + i, which a kernel keeps in LDS:
//   * invc[i] * 2^-k: the general path's reduced argument is z * 2^-k (z with
//     its exponent moved, exact for normal z), so fma(z, invc * 2^-k, -1) has
//     the same exact value before its one rounding as fma(z * 2^-k, invc, -1),
//     and the power-of-two scaling of invc is exact;
//   * fma((double)k, Ln2, logc[i]), the same expression, so the same bits.
// With t = bits(z) - 0x2fb30000 (= bits(z) - 0x3f330000 + 31 * 2^23), e is
// t >> 19, masked to the table, so an argument outside that range (a NaN lane
// whose value is discarded) still reads inside it.
#define DVC_PHI_LOGTAB 512  // entries of two doubles each (9 index bits)
DVC_MATH_FN void dvc_phi_logtab_entry(int e, double *tab) {
  const double Ln2 = 0x1.62e42fefa39efp-1;
  const int i = e & 15, k = (e >> 4) - 31;
  tab[2 * e] = ldexp(dvcm_logf_invc[i], -k);
  tab[2 * e + 1] = fma((double)k, Ln2, dvcm_logf_logc[i]);
}
DVC_MATH_FN float dvc_logf_phi(float x, const double *tab) {
  const double A0 = -0x1.00ea348b88334p-2, A1 = 0x1.5575b0be00b6ap-2, A2 = -0x1.ffffef20a4123p-2;
  const uint32_t t = dvcm_asu(x) - 0x2fb30000u;
#if defined(__HIP_DEVICE_COMPILE__)
  // one v_bfe_u32 (kept as written: the compiler would otherwise fold the
  // field extraction and the address scaling into a shift, a mask and an add)
  uint32_t e;
  __asm__("v_bfe_u32 %0, %1, 19, 9" : "=v"(e) : "v"(t));
#else
  const uint32_t e = (t >> 19) & (DVC_PHI_LOGTAB - 1);
#endif
  const double invcs = tab[2 * e], y0 = tab[2 * e + 1];
  const double r = fma((double)x, invcs, -1.0);
  const double r2 = r * r;
  double y = fma(A1, r, A2);
  y = fma(A0, r2, y);
  y = fma(y, r2, y0 + r);
  return (float)y;
}

// dvc_bp_phi for x >= 0 (the magnitudes BP applies it to), bit-identical to
// it, without data-dependent branches, so the lanes of a wave never diverge.
// With y = (x + 1e-9f)/2 >= 5e-10:
//   * tanhf: y >= 22 gives 1; otherwise its two expm1f calls (argument 2y for
//     y >= 1, -2y below) become one call on the selected argument, and its
//     two quotients one division with a selected numerator;
//   * expm1f on a in (-2, -1e-9] or [2, 44): every reduction case and every
//     result case is computed and the reached one selected (|a| < 2^-25
//     returns a; k = 0, -1, 1, <= -2 or > 56, 2..22, 23..56);
//   * logf of z in (0, 1] is its general path (dvc_logf_normal).
// Arguments outside these ranges only reach lanes whose value the final
// selects discard.
DVC_MATH_FN float dvc_bp_phi_nonneg(float x, const double *logtab) {
  DVCM_NOCONTRACT
  const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f, invln2 = 1.4426950216e+00f,
              Q1 = -3.3333335072e-02f, Q2 = 1.5873016091e-03f, Q3 = -7.9365076090e-05f,
              Q4 = 4.0082177293e-06f, Q5 = -2.0109921195e-07f;
  const float y = (x + (float)0.000000001) / 2.0f;
  const uint32_t iy = dvcm_asu(y) & 0x7fffffffu;
  const int big = iy >= 0x3f800000u;
  const float a = big ? 2.0f * y : -2.0f * y;
  // ---- expm1f(a)
  const uint32_t ha = dvcm_asu(a);
  const int neg = (ha >> 31) != 0;
  const uint32_t hx = ha & 0x7fffffffu;
  const int red = hx > 0x3eb17218u;
  const int near1 = hx < 0x3F851592u;
  const float hi1 = neg ? a + ln2_hi : a - ln2_hi;
  const float lo1 = neg ? -ln2_lo : ln2_lo;
  float kf = invln2 * a + (neg ? -0.5f : 0.5f);
  kf = kf > 256.0f ? 256.0f : (kf < -256.0f ? -256.0f : kf);
  const int kg = (int)kf;
  const float tg = (float)kg;
  const float hig = a - tg * ln2_hi;
  const float log_ = tg * ln2_lo;
  const float hi = near1 ? hi1 : hig, lo = near1 ? lo1 : log_;
  const int kr = near1 ? (neg ? -1 : 1) : kg;
  const float xr = hi - lo;
  const float c = (hi - xr) - lo;
  const float xx = red ? xr : a;
  const int k = red ? kr : 0;
  const float hfx = 0.5f * xx;
  const float hxs = xx * hfx;
  const float r1 = 1.0f + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * (Q4 + hxs * Q5))));
  const float t3 = 3.0f - r1 * hfx;
  const float e = hxs * ((r1 - t3) / (6.0f - xx * t3));
  const float r_k0 = xx - (xx * e - hxs);
  float e2 = (xx * (e - c) - c);
  e2 -= hxs;
  const float r_km1 = 0.5f * (xx - e2) - 0.5f;
  const float r_k1 = (xx < -0.25f) ? -2.0f * (e2 - (xx + 0.5f)) : 1.0f + 2.0f * (xx - e2);
  const uint32_t ksh = (uint32_t)k << 23;
  const float r_big = dvcm_asf(dvcm_asu(1.0f - (e2 - xx)) + ksh) - 1.0f;
  const int kc = k < 0 ? 0 : (k > 31 ? 31 : k);
  const float t23 = dvcm_asf(0x3f800000u - (0x1000000u >> kc));
  const float r_lt23 = dvcm_asf(dvcm_asu(t23 - (e2 - xx)) + ksh);
  const float tge = dvcm_asf((uint32_t)(0x7f - k) << 23);
  float yge = xx - (e2 + tge);
  yge += 1.0f;
  const float r_ge23 = dvcm_asf(dvcm_asu(yge) + ksh);
  float em1 = (k == 0) ? r_k0
            : (k == -1) ? r_km1
            : (k == 1) ? r_k1
            : (k <= -2 || k > 56) ? r_big
            : (k < 23) ? r_lt23 : r_ge23;
  em1 = (!red && hx < 0x33000000u) ? a : em1;
  // ---- tanhf(y)
  const float q = (big ? 2.0f : -em1) / (em1 + 2.0f);
  float z = big ? 1.0f - q : q;
  z = iy >= 0x41b00000u ? 1.0f : z;
  return -dvc_logf_normal(z, logtab);
}

// Float quotient a/b for dvc_bp_phi_fast: on the device the hardware
// reciprocal and one fused correction of the quotient (q = a*y, then
// q + (a - b*q)*y), without the range scaling and reciprocal refinement of
// the general IEEE sequence.  This is not a correctly rounded division in
// general; it is used only for phi's two quotients, and phi's result with it
// equals the reference expression's on every non-negative float (the device
// self-test sweeps all 2^31 of them).  Host builds divide.
DVC_MATH_FN float dvcm_div_phi(float a, float b) {
#if defined(__HIP_DEVICE_COMPILE__)
  const float y = __builtin_amdgcn_rcpf(b);
  const float q = a * y;
  const float r = fmaf(-b, q, a);
  return fmaf(r, y, q);
#else
  return a / b;
#endif
}

// dvc_bp_phi_nonneg with fewer instructions, same value for every x >= 0:
//   * y is clamped to 9.75: tanhf(y) is 1.0f from y = 9.02 on (2/(e^2y+1)
//     is below half an ulp of 1), so every larger argument gives the same
//     result, every intermediate is finite, and expm1f's reduction index k
//     stays in [-3, 28];
//   * with the reduction index k taken on every argument, the reduction
//     a - k*ln2_hi - k*ln2_lo reproduces expm1f's special cases by itself:
//     k = 0 leaves a unchanged (the unreduced case, c = 0) and k = -1 is
//     exactly the hx < 1.5 ln2 case's a + ln2_hi, -ln2_lo; k is rint(a/ln2)
//     where expm1f has (int)(a/ln2 +- 0.5): they differ on 14 floats a of
//     this range, on which phi's result is the same;
//   * k takes the values -3..0 and 3..28; the result cases k = 0, k = -1,
//     k <= -2 and k >= 3 are all computed and one selected (no branches);
//     for k = 23..28 (y >= 7.8) the k <= 22 form stands in for expm1f's
//     k > 22 form, which gives the same tanhf there; |a| < 2^-25 needs no
//     case of its own (x - (x*e - hxs) is x there);
//   * expm1f's polynomial stops at Q4: on the reduced arguments here
//     (|r| <= 0.35) the Q5 term never changes the final result;
//   * both quotients use dvcm_div_phi;
//   * a - k*ln2_hi is one fused multiply-add: k*ln2_hi is exact (|k| <= 28
//     and ln2_hi has 17 significant bits), so the single rounding is the
//     subtraction's; likewise 0.5*(xx - e2) - 0.5, whose product is exact
//     (and xx - e2 is the negated e2 - xx, which the k <= -2 and k >= 3
//     cases use as well);
//   * the k = 0 result is -u (u = e2 - xx, below), and one expression gives
//     the k <= -2 and k >= 3 cases' Y;
//   * logf is dvc_logf_phi with its (k, i) table in `phitab`.
// Pins a value in a register at this point, so the compiler cannot sink its
// computation into a branch (device builds; a no-op on the host).
DVC_MATH_FN float dvcm_pin(float v) {
#if defined(__HIP_DEVICE_COMPILE__)
  __asm__("" : "+v"(v));
#endif
  return v;
}

DVC_MATH_FN float dvc_bp_phi_fast(float x, const double *phitab) {
  DVCM_NOCONTRACT
  const float ln2_hi = 6.9313812256e-01f, ln2_lo = 9.0580006145e-06f, invln2 = 1.4426950216e+00f,
              Q1 = -3.3333335072e-02f, Q2 = 1.5873016091e-03f, Q3 = -7.9365076090e-05f,
              Q4 = 4.0082177293e-06f;
  // y0 = (x + 1e-9f)/2 and a = +-2*min(y0, 9.75): halving and doubling s =
  // x + 1e-9f (a normal float) are exact, so a = min(s, 19.5) when y0 >= 1
  // (s >= 2) and -s below, with the same bits as the halved-then-doubled form
  const float s = x + (float)0.000000001;
  const int big = s >= 2.0f;
  const float a = big ? fminf(s, 19.5f) : -s;
  // ---- expm1f(a), a in (-2, -1e-9] or [2, 19.5]
  // k = rint(invln2*a) in place of trunc(invln2*a +- 0.5): the two differ on
  // 14 floats a of this range, and phi's result is the same on those (every
  // argument is checked exhaustively); rint leaves tk a float and k one
  // conversion away
  const float tk = rintf(invln2 * a);
  const int k = (int)tk;
  const float hi = fmaf(-tk, ln2_hi, a);
  const float lo = tk * ln2_lo;
  const float xx = hi - lo;
  const float c = (hi - xx) - lo;
  const float hfx = 0.5f * xx;
  const float hxs = xx * hfx;
  const float r1 = 1.0f + hxs * (Q1 + hxs * (Q2 + hxs * (Q3 + hxs * Q4)));
  const float t3 = 3.0f - r1 * hfx;
  const float e = hxs * dvcm_div_phi(r1 - t3, 6.0f - xx * t3);
  float e2 = (xx * (e - c) - c);
  e2 -= hxs;
  const float u = e2 - xx;
  // k = 0: c is +0 (hi = a, lo = +0), so e2 is xx*e - hxs with the same two
  // roundings as the k = 0 result's inner term, and xx - (xx*e - hxs) is -u
  const float r_k0 = -u;
  const float r_km1 = fmaf(-0.5f, u, -0.5f);  // 0.5f * (xx - e2) - 0.5f: xx - e2 is -u exactly
  // k >= 3: Y = (1 - 2^-k) - u; k <= -2: Y = 1 - u, which the same expression
  // gives, since k & 31 is then 29 or 30 and 2^24 >> 29 (or 30) is 0
  const float Y = dvcm_pin(dvcm_asf(0x3f800000u - (0x1000000u >> (k & 31))) - u);
  const float R = dvcm_asf(dvcm_asu(Y) + ((uint32_t)k << 23));
  // k >= 3 exactly on the lanes with a >= 2 (big); the others have k in -3..0
  float em1 = (k == -1) ? r_km1 : r_k0;
  em1 = (k <= -2) ? R - 1.0f : em1;
  em1 = big ? R : em1;
  // ---- tanhf(y)
  const float q = dvcm_div_phi(big ? 2.0f : -em1, em1 + 2.0f);
  const float z = big ? 1.0f - q : q;
  return -dvc_logf_phi(z, phitab);
}

// ---------------------------------------------------------------- exp (double)
// glibc 2.35's exp (the ARM optimized-routines algorithm), as its x86-64 FMA
// build evaluates it.  The reference calls it in reconstruction
// (DvcDec/src/reconstruction.cpp:211-222, 266-277).  exp(x) = 2^(k/128) *
// exp(r), x = k*ln2/128 + r, |r| <= ln2/256; 2^(k/128) = H*(1 + T) from the
// table; exp(r) - 1 by a degree-5 polynomial.  The FMA build contracts every
// a + b*c whose product has no other use (so the k < 0 special case, which
// reuses scale*tmp, stays unfused).
// Table entry k: (bits of T, bits of H - (k << 45)) with H the double nearest
// 2^(k/128) and T the double nearest (2^(k/128) - H)/H (regenerated from that
// definition with 80-digit decimal arithmetic; tests/test_math.py re-derives
// it).
DVC_MATH_TABLE uint64_t dvcm_exp_tab[256] = {
    0x0000000000000000ull, 0x3ff0000000000000ull, 0x3c9b3b4f1a88bf6eull, 0x3feff63da9fb3335ull,
    0xbc7160139cd8dc5dull, 0x3fefec9a3e778061ull, 0xbc905e7a108766d1ull, 0x3fefe315e86e7f85ull,
    0x3c8cd2523567f613ull, 0x3fefd9b0d3158574ull, 0xbc8bce8023f98efaull, 0x3fefd06b29ddf6deull,
    0x3c60f74e61e6c861ull, 0x3fefc74518759bc8ull, 0x3c90a3e45b33d399ull, 0x3fefbe3ecac6f383ull,
    0x3c979aa65d837b6dull, 0x3fefb5586cf9890full, 0x3c8eb51a92fdeffcull, 0x3fefac922b7247f7ull,
    0x3c3ebe3d702f9cd1ull, 0x3fefa3ec32d3d1a2ull, 0xbc6a033489906e0bull, 0x3fef9b66affed31bull,
    0xbc9556522a2fbd0eull, 0x3fef9301d0125b51ull, 0xbc5080ef8c4eea55ull, 0x3fef8abdc06c31ccull,
    0xbc91c923b9d5f416ull, 0x3fef829aaea92de0ull, 0x3c80d3e3e95c55afull, 0x3fef7a98c8a58e51ull,
    0xbc801b15eaa59348ull, 0x3fef72b83c7d517bull, 0xbc8f1ff055de323dull, 0x3fef6af9388c8deaull,
    0x3c8b898c3f1353bfull, 0x3fef635beb6fcb75ull, 0xbc96d99c7611eb26ull, 0x3fef5be084045cd4ull,
    0x3c9aecf73e3a2f60ull, 0x3fef54873168b9aaull, 0xbc8fe782cb86389dull, 0x3fef4d5022fcd91dull,
    0x3c8a6f4144a6c38dull, 0x3fef463b88628cd6ull, 0x3c807a05b0e4047dull, 0x3fef3f49917ddc96ull,
    0x3c968efde3a8a894ull, 0x3fef387a6e756238ull, 0x3c875e18f274487dull, 0x3fef31ce4fb2a63full,
    0x3c80472b981fe7f2ull, 0x3fef2b4565e27cddull, 0xbc96b87b3f71085eull, 0x3fef24dfe1f56381ull,
    0x3c82f7e16d09ab31ull, 0x3fef1e9df51fdee1ull, 0xbc3d219b1a6fbffaull, 0x3fef187fd0dad990ull,
    0x3c8b3782720c0ab4ull, 0x3fef1285a6e4030bull, 0x3c6e149289cecb8full, 0x3fef0cafa93e2f56ull,
    0x3c834d754db0abb6ull, 0x3fef06fe0a31b715ull, 0x3c864201e2ac744cull, 0x3fef0170fc4cd831ull,
    0x3c8fdd395dd3f84aull, 0x3feefc08b26416ffull, 0xbc86a3803b8e5b04ull, 0x3feef6c55f929ff1ull,
    0xbc924aedcc4b5068ull, 0x3feef1a7373aa9cbull, 0xbc9907f81b512d8eull, 0x3feeecae6d05d866ull,
    0xbc71d1e83e9436d2ull, 0x3feee7db34e59ff7ull, 0xbc991919b3ce1b15ull, 0x3feee32dc313a8e5ull,
    0x3c859f48a72a4c6dull, 0x3feedea64c123422ull, 0xbc9312607a28698aull, 0x3feeda4504ac801cull,
    0xbc58a78f4817895bull, 0x3feed60a21f72e2aull, 0xbc7c2c9b67499a1bull, 0x3feed1f5d950a897ull,
    0x3c4363ed60c2ac11ull, 0x3feece086061892dull, 0x3c9666093b0664efull, 0x3feeca41ed1d0057ull,
    0x3c6ecce1daa10379ull, 0x3feec6a2b5c13cd0ull, 0x3c93ff8e3f0f1230ull, 0x3feec32af0d7d3deull,
    0x3c7690cebb7aafb0ull, 0x3feebfdad5362a27ull, 0x3c931dbdeb54e077ull, 0x3feebcb299fddd0dull,
    0xbc8f94340071a38eull, 0x3feeb9b2769d2ca7ull, 0xbc87deccdc93a349ull, 0x3feeb6daa2cf6642ull,
    0xbc78dec6bd0f385full, 0x3feeb42b569d4f82ull, 0xbc861246ec7b5cf6ull, 0x3feeb1a4ca5d920full,
    0x3c93350518fdd78eull, 0x3feeaf4736b527daull, 0x3c7b98b72f8a9b05ull, 0x3feead12d497c7fdull,
    0x3c9063e1e21c5409ull, 0x3feeab07dd485429ull, 0x3c34c7855019c6eaull, 0x3feea9268a5946b7ull,
    0x3c9432e62b64c035ull, 0x3feea76f15ad2148ull, 0xbc8ce44a6199769full, 0x3feea5e1b976dc09ull,
    0xbc8c33c53bef4da8ull, 0x3feea47eb03a5585ull, 0xbc845378892be9aeull, 0x3feea34634ccc320ull,
    0xbc93cedd78565858ull, 0x3feea23882552225ull, 0x3c5710aa807e1964ull, 0x3feea155d44ca973ull,
    0xbc93b3efbf5e2228ull, 0x3feea09e667f3bcdull, 0xbc6a12ad8734b982ull, 0x3feea012750bdabfull,
    0xbc6367efb86da9eeull, 0x3fee9fb23c651a2full, 0xbc80dc3d54e08851ull, 0x3fee9f7df9519484ull,
    0xbc781f647e5a3ecfull, 0x3fee9f75e8ec5f74ull, 0xbc86ee4ac08b7db0ull, 0x3fee9f9a48a58174ull,
    0xbc8619321e55e68aull, 0x3fee9feb564267c9ull, 0x3c909ccb5e09d4d3ull, 0x3feea0694fde5d3full,
    0xbc7b32dcb94da51dull, 0x3feea11473eb0187ull, 0x3c94ecfd5467c06bull, 0x3feea1ed0130c132ull,
    0x3c65ebe1abd66c55ull, 0x3feea2f336cf4e62ull, 0xbc88a1c52fb3cf42ull, 0x3feea427543e1a12ull,
    0xbc9369b6f13b3734ull, 0x3feea589994cce13ull, 0xbc805e843a19ff1eull, 0x3feea71a4623c7adull,
    0xbc94d450d872576eull, 0x3feea8d99b4492edull, 0x3c90ad675b0e8a00ull, 0x3feeaac7d98a6699ull,
    0x3c8db72fc1f0eab4ull, 0x3feeace5422aa0dbull, 0xbc65b6609cc5e7ffull, 0x3feeaf3216b5448cull,
    0x3c7bf68359f35f44ull, 0x3feeb1ae99157736ull, 0xbc93091fa71e3d83ull, 0x3feeb45b0b91ffc6ull,
    0xbc5da9b88b6c1e29ull, 0x3feeb737b0cdc5e5ull, 0xbc6c23f97c90b959ull, 0x3feeba44cbc8520full,
    0xbc92434322f4f9aaull, 0x3feebd829fde4e50ull, 0xbc85ca6cd7668e4bull, 0x3feec0f170ca07baull,
    0x3c71affc2b91ce27ull, 0x3feec49182a3f090ull, 0x3c6dd235e10a73bbull, 0x3feec86319e32323ull,
    0xbc87c50422622263ull, 0x3feecc667b5de565ull, 0x3c8b1c86e3e231d5ull, 0x3feed09bec4a2d33ull,
    0xbc91bbd1d3bcbb15ull, 0x3feed503b23e255dull, 0x3c90cc319cee31d2ull, 0x3feed99e1330b358ull,
    0x3c8469846e735ab3ull, 0x3feede6b5579fdbfull, 0xbc82dfcd978e9db4ull, 0x3feee36bbfd3f37aull,
    0x3c8c1a7792cb3387ull, 0x3feee89f995ad3adull, 0xbc907b8f4ad1d9faull, 0x3feeee07298db666ull,
    0xbc55c3d956dcaebaull, 0x3feef3a2b84f15fbull, 0xbc90a40e3da6f640ull, 0x3feef9728de5593aull,
    0xbc68d6f438ad9334ull, 0x3feeff76f2fb5e47ull, 0xbc91eee26b588a35ull, 0x3fef05b030a1064aull,
    0x3c74ffd70a5fddcdull, 0x3fef0c1e904bc1d2ull, 0xbc91bdfbfa9298acull, 0x3fef12c25bd71e09ull,
    0x3c736eae30af0cb3ull, 0x3fef199bdd85529cull, 0x3c8ee3325c9ffd94ull, 0x3fef20ab5fffd07aull,
    0x3c84e08fd10959acull, 0x3fef27f12e57d14bull, 0x3c63cdaf384e1a67ull, 0x3fef2f6d9406e7b5ull,
    0x3c676b2c6c921968ull, 0x3fef3720dcef9069ull, 0xbc808a1883ccb5d2ull, 0x3fef3f0b555dc3faull,
    0xbc8fad5d3ffffa6full, 0x3fef472d4a07897cull, 0xbc900dae3875a949ull, 0x3fef4f87080d89f2ull,
    0x3c74a385a63d07a7ull, 0x3fef5818dcfba487ull, 0xbc82919e2040220full, 0x3fef60e316c98398ull,
    0x3c8e5a50d5c192acull, 0x3fef69e603db3285ull, 0x3c843a59ac016b4bull, 0x3fef7321f301b460ull,
    0xbc82d52107b43e1full, 0x3fef7c97337b9b5full, 0xbc892ab93b470dc9ull, 0x3fef864614f5a129ull,
    0x3c74b604603a88d3ull, 0x3fef902ee78b3ff6ull, 0x3c83c5ec519d7271ull, 0x3fef9a51fbc74c83ull,
    0xbc8ff7128fd391f0ull, 0x3fefa4afa2a490daull, 0xbc8dae98e223747dull, 0x3fefaf482d8e67f1ull,
    0x3c8ec3bc41aa2008ull, 0x3fefba1bee615a27ull, 0x3c842b94c3a9eb32ull, 0x3fefc52b376bba97ull,
    0x3c8a64a931d185eeull, 0x3fefd0765b6e4540ull, 0xbc8e37bae43be3edull, 0x3fefdbfdad9cbe14ull,
    0x3c77893b4d91cd9dull, 0x3fefe7c1819e90d8ull, 0x3c5305c14160cc89ull, 0x3feff3c22b8f71f1ull};

DVC_MATH_FN double dvcm_exp_special(double tmp, uint64_t sbits, uint64_t ki) {
  if ((ki & 0x80000000u) == 0) {  // k > 0: the scale's exponent may have overflowed by <= 460
    sbits -= 1009ull << 52;
    const double scale = dvcm_asd(sbits);
    return 0x1p1009 * fma(scale, tmp, scale);
  }
  sbits += 1022ull << 52;  // k < 0: care in the subnormal range
  const double scale = dvcm_asd(sbits);
  const double st = scale * tmp;
  double y = scale + st;
  if (y < 1.0) {
    // round y to the final precision before it is scaled into the subnormals
    double lo = scale - y + st;
    const double hi = 1.0 + y;
    lo = 1.0 - hi + y + lo;
    y = (hi + lo) - 1.0;
    if (y == 0.0) y = 0.0;
  }
  return 0x1p-1022 * y;
}

DVC_MATH_FN double dvc_exp(double x) {
  const double InvLn2N = 0x1.71547652b82fep0 * 128, NegLn2hiN = -0x1.62e42fefa0000p-8,
               NegLn2loN = -0x1.cf79abc9e3b3ap-47, Shift = 0x1.8p52;
  const double C2 = 0x1.ffffffffffdbdp-2, C3 = 0x1.555555555543cp-3, C4 = 0x1.55555cf172b91p-5,
               C5 = 0x1.1111167a4d017p-7;
  uint32_t abstop = (uint32_t)(dvcm_asu64(x) >> 52) & 0x7ffu;
  if (abstop - 0x3c9u >= 0x408u - 0x3c9u) {  // |x| < 2^-54 or |x| >= 512 (or inf, NaN)
    if (abstop - 0x3c9u >= 0x80000000u) return 1.0 + x;  // tiny: 1 + x
    if (abstop >= 0x409u) {                               // |x| >= 1024
      if (dvcm_asu64(x) == 0xfff0000000000000ull) return 0.0;
      if (abstop >= 0x7ffu) return 1.0 + x;
      return (dvcm_asu64(x) >> 63) ? 0.0 : INFINITY;
    }
    abstop = 0;  // 512 <= |x| < 1024: the special case below
  }
  double kd = fma(InvLn2N, x, Shift);
  const uint64_t ki = dvcm_asu64(kd);
  kd -= Shift;
  const double r = fma(kd, NegLn2loN, fma(kd, NegLn2hiN, x));
  const uint64_t idx = 2 * (ki % 128), top = ki << 45;
  const double tail = dvcm_asd(dvcm_exp_tab[idx]);
  const uint64_t sbits = dvcm_exp_tab[idx + 1] + top;
  const double r2 = r * r;
  const double tmp = fma(r2 * r2, fma(r, C5, C4), fma(r2, fma(r, C3, C2), tail + r));
  if (abstop == 0) return dvcm_exp_special(tmp, sbits, ki);
  const double scale = dvcm_asd(sbits);
  return fma(scale, tmp, scale);
}
