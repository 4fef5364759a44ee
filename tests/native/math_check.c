/* Compares the single-precision routines of dvc_amd/csrc/dvc_math.h with the
 * host C library over a range of float bit patterns.
 *   math_check <func> <first_bits_hex> <last_bits_hex> <stride>
 * func: logf | expf | tanhf | expm1f | phi | phi_nonneg | phi_fast
 *   math_check exp <count> <seed_hex> 0   (double exp on pseudo-random arguments)
 * Prints "<func> checked=<n> mismatches=<m>" and the first few mismatches.
 * phi is -logf(tanhf((x + (float)1e-9)/2)), the check-node kernel of
 * DvcDec/src/LdpcaDecoder.cpp:513.  phi_nonneg is the branch-free form the BP
 * kernel uses for x >= 0, compared with the same C library value; phi_fast is
 * its shortened form (on the host its quotients are plain IEEE divisions; the
 * device form is compared with phi_nonneg on the GPU). */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../dvc_amd/csrc/dvc_math.h"

static float ref_phi(float x) { return -logf(tanhf((x + (float)0.000000001) / 2)); }

/* Double exp (reconstruction.cpp:211-277) on n pseudo-random arguments
 * (xorshift64, seed): a third uniform in [-700, 700], a third in [-20, 20], a
 * third the arguments reconstruction forms, +-alpha*k with alpha a float in
 * (1e-4, 64) and k an integer in [0, 4096]. */
static int check_exp(long long n, unsigned long long seed) {
  long long mism = 0;
#pragma omp parallel for reduction(+:mism) schedule(static)
  for (long long i = 0; i < n; i++) {
    unsigned long long s = seed ^ (0x9e3779b97f4a7c15ull * (unsigned long long)(i + 1));
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    s ^= s << 13; s ^= s >> 7; s ^= s << 17;
    const double u = (double)(s >> 11) / 9007199254740992.0;
    double x;
    if (i % 3 == 0) x = u * 1400.0 - 700.0;
    else if (i % 3 == 1) x = u * 40.0 - 20.0;
    else {
      const float a = (float)(1e-4 + u * 64.0);
      const int k = (int)((s >> 3) % 4097);
      x = ((s & 1) ? -(double)a : (double)a) * (double)k;
    }
    const double g = exp(x), m = dvc_exp(x);
    if (dvcm_asu64(g) != dvcm_asu64(m)) {
      mism++;
#pragma omp critical
      { static int shown = 0; if (shown++ < 5) fprintf(stderr, "mismatch exp x=%a libm=%a ours=%a\n", x, g, m); }
    }
  }
  printf("exp checked=%lld mismatches=%lld\n", n, mism);
  return mism ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc != 5) { fprintf(stderr, "usage: %s func first last stride\n", argv[0]); return 2; }
  const char *fn = argv[1];
  if (!strcmp(fn, "exp")) return check_exp(strtoll(argv[2], 0, 10), strtoull(argv[3], 0, 16));
  unsigned long long first = strtoull(argv[2], 0, 16), last = strtoull(argv[3], 0, 16);
  unsigned long long stride = strtoull(argv[4], 0, 10);
  if (stride == 0) stride = 1;
  int which = !strcmp(fn, "logf") ? 0 : !strcmp(fn, "expf") ? 1 : !strcmp(fn, "tanhf") ? 2
            : !strcmp(fn, "expm1f") ? 3 : !strcmp(fn, "phi") ? 4 : !strcmp(fn, "phi_nonneg") ? 5 : !strcmp(fn, "phi_fast") ? 6 : -1;
  if (which < 0) { fprintf(stderr, "unknown func %s\n", fn); return 2; }
  double logtab[32], phitab[2 * DVC_PHI_LOGTAB];
  for (int i = 0; i < 16; i++) { logtab[i] = dvcm_logf_invc[i]; logtab[16 + i] = dvcm_logf_logc[i]; }
  for (int e = 0; e < DVC_PHI_LOGTAB; e++) dvc_phi_logtab_entry(e, phitab);
  long long checked = 0, mism = 0;
#pragma omp parallel for reduction(+:checked,mism) schedule(static)
  for (long long b = (long long)first; b <= (long long)last; b += (long long)stride) {
    float x = dvcm_asf((uint32_t)b);
    if (x != x) continue;  /* NaN payloads are not compared */
    float g, m;
    switch (which) {
      case 0: g = logf(x); m = dvc_logf(x); break;
      case 1: g = expf(x); m = dvc_expf(x); break;
      case 2: g = tanhf(x); m = dvc_tanhf(x); break;
      case 3: g = expm1f(x); m = dvc_expm1f(x); break;
      case 4: g = ref_phi(x); m = dvc_bp_phi(x); break;
      case 5: g = ref_phi(x); m = dvc_bp_phi_nonneg(x, logtab); break;
      default: g = ref_phi(x); m = dvc_bp_phi_fast(x, phitab); break;
    }
    checked++;
    if (dvcm_asu(g) != dvcm_asu(m) && !(g != g && m != m)) {
      mism++;
#pragma omp critical
      { static int shown = 0; if (shown++ < 5) fprintf(stderr, "mismatch %s x=%a libm=%a ours=%a\n", fn, x, g, m); }
    }
  }
  printf("%s checked=%lld mismatches=%lld\n", fn, checked, mism);
  return mism ? 1 : 0;
}
