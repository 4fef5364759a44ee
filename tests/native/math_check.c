/* Compares the single-precision routines of dvc_amd/csrc/dvc_math.h with the
 * host C library over a range of float bit patterns.
 *   math_check <func> <first_bits_hex> <last_bits_hex> <stride>
 * func: logf | expf | tanhf | expm1f | phi | phi_nonneg | phi_fast
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

int main(int argc, char **argv) {
  if (argc != 5) { fprintf(stderr, "usage: %s func first last stride\n", argv[0]); return 2; }
  const char *fn = argv[1];
  unsigned long long first = strtoull(argv[2], 0, 16), last = strtoull(argv[3], 0, 16);
  unsigned long long stride = strtoull(argv[4], 0, 10);
  if (stride == 0) stride = 1;
  int which = !strcmp(fn, "logf") ? 0 : !strcmp(fn, "expf") ? 1 : !strcmp(fn, "tanhf") ? 2
            : !strcmp(fn, "expm1f") ? 3 : !strcmp(fn, "phi") ? 4 : !strcmp(fn, "phi_nonneg") ? 5 : !strcmp(fn, "phi_fast") ? 6 : -1;
  if (which < 0) { fprintf(stderr, "unknown func %s\n", fn); return 2; }
  double logtab[32];
  for (int i = 0; i < 16; i++) { logtab[i] = dvcm_logf_invc[i]; logtab[16 + i] = dvcm_logf_logc[i]; }
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
      default: g = ref_phi(x); m = dvc_bp_phi_fast(x, logtab); break;
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
