// ldpca.hip -- LDPCA sum-product decoding on gfx950.
//
// Replaces beliefPropagation_CUDA / beliefPropagation_CUDA_earlyJump
// (DvcDec/src/LDPCA_cuda.cu:872-1031) and the rate search of LdpcaDecoder
// (DvcDec/src/LdpcaDecoder.cpp:646-771).  Arithmetic follows the reference's
// ParallelMode 0 decoder (beliefPropagation, LdpcaDecoder.cpp:462-580) step
// for step, so every message, decision and iteration count equals the CPU
// path's:
//   * phi(x) = -logf(tanhf((x + 1e-9f)/2)) with the glibc-identical routines
//     of dvc_math.h, in their branch-free form for x >= 0 (every argument
//     here is a magnitude: a row's sum of |phi| is at least each of its
//     terms, so |sum| - |phi_e| >= 0), dvc_bp_phi_fast;
//   * check-node sums accumulate in ascending edge order (= ascending column),
//     the order of `rowTotal[ir[k]] += ...` (:514-515);
//   * variable-node sums accumulate LLR + c2v0 + c2v1 + c2v2 (:527-539);
//   * decisions are the sign bit of the overall LLR (:548); the syndrome test
//     is the row parity (:551-568).
//
// Layout.  One workgroup owns one (bitplane, code) pair.  Messages live in LDS:
// phi[E] holds |phi| of every variable-to-check message with the message's
// sign in the sign bit, stored by check row (row r's edges at rowPtr[r] ..
// rowPtr[r+1]-1, ascending edge order), so a check sums a contiguous run;
// rowTot[m] holds each check's sum of |phi| with the check's sign product
// (including the syndrome bit) in the sign bit.  Thread t owns columns t,
// t+T, ...; a column's three edges come as (row, message slot) words, the
// slot precomputed per code, so the variable-node update, the next check-node input and
// the decision are computed in registers without further exchange.  Two
// barriers per iteration: after the check (row) pass and after the variable
// (column) pass.  The syndrome test's decision parities are XORed into
// per-check bits by the variable pass (LDS atomics, double-buffered across
// iterations), so the check pass reads no column data besides phi.
//
// Batched form.  A persistent grid walks a queue of (bitplane, code offset)
// items in offset-major order.  An item whose code lies above the lowest code
// that already converged with a matching CRC for its bitplane is skipped, and
// a running item notices that at its next iteration and stops.  Every code
// below the final choice always runs to completion, so the kept code is the
// one the sequential search returns.
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <vector>
#include "dvc_internal.h"
#include "dvc_math.h"
#include "ldpca_kernels.h"

#pragma clang fp contract(off)

namespace dvc {

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }

// Pointers the kernels read from descriptors in memory are generic, and a
// generic (flat) load is counted by the LDS wait counter as well: a wait for
// an LDS read would also wait for it.  These views make such loads global.
typedef const float __attribute__((address_space(1))) gfloat_t;
typedef const uint8_t __attribute__((address_space(1))) gbyte_t;
typedef const uint32_t __attribute__((address_space(1))) gword_t;
__device__ __forceinline__ gfloat_t *as_global(const float *p) { return (gfloat_t *)p; }
__device__ __forceinline__ gbyte_t *as_global(const uint8_t *p) { return (gbyte_t *)p; }
__device__ __forceinline__ gword_t *as_global(const uint32_t *p) { return (gword_t *)p; }

// A column's three edges as the ladder upload packs them: row | slot << 15,
// slot = the edge's message position in row order; off = slot * 4, the
// message's byte offset in the phi array, is the word shifted right by 13.
__device__ __forceinline__ void col_edges(const ColEdges &c, int row[3], uint32_t off[3]) {
#pragma unroll
  for (int q = 0; q < 3; q++) {
    row[q] = (int)(c.e[q] & 0x1fffu);
    off[q] = c.e[q] >> 13;
  }
}
__device__ __forceinline__ float &at_byte(float *base, uint32_t off) {
  return *(float *)((uint8_t *)base + off);
}
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// Accumulated-syndrome bit i (MSB-first bytes, ByteToBit LdpcaDecoder.cpp:81-107).
__device__ __forceinline__ uint32_t acc_bit(gbyte_t *acc, int i) {
  return (acc[i >> 3] >> (7 - (i & 7))) & 1u;
}

// ------------------------------------------------------------------ CE + LLR
// ComputeCE (LdpcaDecoder.cpp:127-143) and the LR -> LLR conversion (:663-667)
// for one bitplane per workgroup.  The entropy is summed sequentially in
// column order by lane 0, as the reference does.
#define DVC_LOG2 (0.693147181f)
__global__ __launch_bounds__(256) void ce_llr_kernel(const PlaneDesc *planes, int nplanes, int fixed_start,
                                                     const float *codeRate1584, const float *codeRate6336,
                                                     int numCodes1584, int numCodes6336) {
  __shared__ float terms[6336];
  const int p = blockIdx.x;
  if (p >= nplanes) return;
  PlaneDesc d = planes[p];
  const int N = d.N;
  int nan = 0;
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    float lr = d.LR[k];
    float Pco = lr / (1.0f + lr);
    float tmp = dvc_logf(1 - Pco);
    terms[k] = (Pco * (-dvc_logf(Pco) + tmp) - tmp) / DVC_LOG2;
    const float llr = dvc_logf(lr);
    d.LLR[k] = llr;
    nan |= llr != llr;
  }
  nan = __syncthreads_or(nan);
  if (threadIdx.x == 0) {
    float rlb = 0.0f;
    for (int k = 0; k < N; k++) rlb += terms[k];
    rlb /= (float)N;
    // A NaN likelihood ratio makes the entropy NaN; the x86 path's is its
    // default NaN (sign set), which the GPU's NaN bits need not be.  The
    // start-code loop below then stops at 0, as :657-659 does.
    if (rlb != rlb) rlb = __uint_as_float(0xffc00000u);
    const float *rate = N == 1584 ? codeRate1584 : codeRate6336;
    const int numCodes = N == 1584 ? numCodes1584 : numCodes6336;
    int s = 0;
    while (s < numCodes - 1 && rate[s] < rlb) s++;
    PlaneState *st = d.state;
    st->entropy = rlb;
    if (st->start < 0 || !fixed_start) st->start = s;
    st->run = st->start;
    st->taint = nan;
  }
}
// ------------------------------------------------------------------ BP
template <int N, int T>
struct BpShape {
  static constexpr int E = 3 * N;
  static constexpr int CPT = (N + T - 1) / T;      // columns per thread
  static constexpr int DEC_WORDS = (CPT * T) / 32;  // ballot words of every column slot
  static constexpr int SYN_WORDS = (N + 31) / 32;
};

// MODE1: beliefPropagation_OpenMP's semantics (ParallelMode 1), else
// beliefPropagation's (ParallelMode 0).
//
// NaN likelihood ratios (PlaneState.taint).  On the x86 path a NaN LLR makes
// every message that depends on it NaN.  With ParallelMode 0's sign-bit
// decisions the outcome then hangs on NaN sign bits, and every code of such a
// bitplane fails after MAX_ITERATIONS (the reference run over the golden GOP
// frames and over seeded bitplanes with injected NaNs never converges), so a
// tainted code is recorded as failed after max_iter iterations without being
// run.  ParallelMode 1 decides LLR < 0, false for NaN, so only which values
// are NaN matters: the kernel then keeps NaN through phi (whose branch-free
// form would turn it into a finite value) and IEEE arithmetic carries it like
// the CPU path.
template <int N, int T, bool MODE1>
__global__ __launch_bounds__(T) void bp_persistent_kernel(BpArgs a) {
  using S = BpShape<N, T>;
  constexpr int E = S::E;
  // phi's logf table: (invc * 2^-k, k*ln2 + logc) for each (k, i) that phi's
  // arguments reach, read with per-lane indices
  __shared__ __align__(16) double s_phitab[2 * DVC_PHI_LOGTAB];
  // by check row; 3 spare slots at the end, so that the check pass may read
  // a row's last group of four unconditionally
  __shared__ float phi[E + 3];
  __shared__ float rowTot[N];
  __shared__ uint16_t rowPtr[N + 2];
  __shared__ uint32_t decBits[S::DEC_WORDS];
  __shared__ uint32_t synBits[S::SYN_WORDS];
  __shared__ int s_item, s_cancel;
  __shared__ uint32_t s_crc[T / 64];
  __shared__ uint32_t parBits[2][S::SYN_WORDS];  // decision parity per check, written by alternate iterations
  __shared__ int s_changed[2];  // ParallelMode 1: some decision changed in the variable pass (alternate iterations)

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const LadderDev L = a.lad;
  for (int e = tid; e < DVC_PHI_LOGTAB; e += T) dvc_phi_logtab_entry(e, s_phitab);

  for (;;) {
    // ---- grab the next item (lane 0 of the workgroup), broadcast through LDS
    if (tid == 0) {
      int item = -1;
      for (;;) {
        int w = atomicAdd(a.next, 1);
        if (w >= a.total) break;
        int b = w % a.nplanes, off = w / a.nplanes;
        PlaneState *st = a.planes[b].state;
        int level = st->run + off;
        if (off >= a.max_offset || level >= L.numCodes - 1) continue;
        if (a.rate_step && !dynstep_tries(off)) continue;
        if (__hip_atomic_load(&st->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < level) continue;
        item = w;
        break;
      }
      s_item = item;
    }
    __syncthreads();
    const int w = s_item;
    if (w < 0) return;
    const int b = w % a.nplanes;
    const PlaneDesc pd = a.planes[b];
    PlaneState *st = pd.state;
    const int run = st->run;
    const int level = run + w / a.nplanes;
    const bool taint = st->taint != 0;
    if (!MODE1 && taint) {
      if (tid == 0) pd.rec[level] = make_int4(a.max_iter, REC_FAILED, -1, 0);
      __syncthreads();  // every thread has read s_item before the next grab rewrites it
      continue;
    }
    // phi of the kernel, NaN-preserving on a tainted ParallelMode 1 bitplane
    auto phi_of = [&](float x) {
      const float v = dvc_bp_phi_fast(x, s_phitab);
      return (MODE1 && taint && x != x) ? x : v;
    };
    // ParallelMode 1: the decisions carried into this code are known only for
    // the queue's first code; a later code runs as if its first decisions
    // differ from them and records the one case where that matters
    const bool mode1 = MODE1, exact = level == run;
    const int m = L.syndLength[level];
    const ColEdges *gcol = L.colRows + (size_t)level * N;

    // ---- stage the code's row starts and the syndrome (LdpcaDecoder.cpp:684-693)
    const uint16_t *growPtr = L.rowPtr + (size_t)level * (N + 1);
    for (int i = tid; i <= m; i += T) rowPtr[i] = growPtr[i];
    for (int wd = tid; wd < S::SYN_WORDS; wd += T) {
      uint32_t bits = 0;
      for (int q = 0; q < 32; q++) {
        int k = wd * 32 + q;
        if (k < m) {
          int cur = L.tx[(size_t)level * N + k];
          uint32_t s = acc_bit(as_global(pd.acc), cur);
          if (k > 0) s ^= acc_bit(as_global(pd.acc), L.tx[(size_t)level * N + k - 1]);
          bits |= s << q;
        }
      }
      synBits[wd] = bits;
    }
    // ---- variable-to-check messages start at the channel LLR (:478-481)
    for (int j = 0; j < S::CPT; j++) {
      int k = tid + j * T;
      if (k < N) {
        float v = as_global(pd.LLR)[k];
        uint32_t s = f2u(v) >> 31;
        float mag = (s ? -1.0f : 1.0f) * v;
        uint32_t f = f2u(copysignf(phi_of(mag), v));  // |phi| with the LLR's sign bit
        int row[3];
        uint32_t off[3];
        col_edges(gcol[k], row, off);
#pragma unroll
        for (int q = 0; q < 3; q++) at_byte(phi, off[q]) = u2f(f);
      }
    }
    __syncthreads();

    int status = 0, iters = 0;
    int sameCount = 0, bcaseSlot = 0;
    for (int t = 0;; t++) {
      // ---- check-node pass: |phi| sums and sign products; parity of the
      // previous decisions (:495-515, :551-568)
      // The previous variable pass XORed each column's decision into its
      // three checks' bits of parBits[(t - 1) & 1]; this pass reads them and
      // clears parBits[t & 1] for the coming variable pass.
      int fail = 0;
      const uint32_t *parPrev = parBits[(t + 1) & 1];
      for (int r = tid; r < m; r += T) {
        const int e0 = rowPtr[r], e1 = rowPtr[r + 1];
        const uint32_t syn = (synBits[r >> 5] >> (r & 31)) & 1u;
        uint32_t sx = 0u;  // XOR of the row's message words: the sign product in bit 31
        float sum = 0.0f;
        // groups of four loads without conditions, then the last 1..3 slots
        // read unconditionally too (the rows are contiguous and phi has spare
        // slots at its end) with those past the row's end taken as +0: sum +
        // 0.0f is sum bit for bit (sum is never -0)
        int j = e0;
        for (; j + 4 <= e1; j += 4) {
          uint32_t f[4];
#pragma unroll
          for (int i = 0; i < 4; i++) f[i] = f2u(phi[j + i]);
#pragma unroll
          for (int i = 0; i < 4; i++) sum += u2f(f[i] & 0x7fffffffu);
          sx ^= f[0] ^ f[1] ^ f[2] ^ f[3];
        }
        if (j < e1) {
          uint32_t f[3];
#pragma unroll
          for (int i = 0; i < 3; i++) f[i] = f2u(phi[j + i]);
          f[1] = j + 1 < e1 ? f[1] : 0u;
          f[2] = j + 2 < e1 ? f[2] : 0u;
#pragma unroll
          for (int i = 0; i < 3; i++) sum += u2f(f[i] & 0x7fffffffu);
          sx ^= f[0] ^ f[1] ^ f[2];
        }
        const uint32_t sgn = syn ^ (sx >> 31);
        rowTot[r] = u2f(f2u(sum) | (sgn << 31));
        if (t > 0 && (syn ^ ((parPrev[r >> 5] >> (r & 31)) & 1u))) fail = 1;
      }
      for (int wd = tid; wd < S::SYN_WORDS; wd += T) parBits[t & 1][wd] = 0u;
      // ParallelMode 1 (beliefPropagation_OpenMP, LdpcaDecoder.cpp:254-264):
      // count consecutive iterations whose decisions all repeat the previous
      // ones; the fourth such iteration ends the code as a failure, before
      // the parity test of the same iteration
      if (mode1 && t > 0) sameCount = s_changed[(t - 1) & 1] ? 0 : sameCount + 1;
      if (tid == 0) {
        s_changed[t & 1] = 0;
        s_cancel = (t % a.poll == 0) &&
                   __hip_atomic_load(&st->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < level;
      }
      const int anyFail = __syncthreads_or(fail);
      if (mode1 && !exact && t == 4 && sameCount == 3) {
        // Decisions of iterations 0-3 all equal: had the carried decisions
        // equalled them too, the code would have failed here after 4
        // iterations.  Keep these decisions for the resolution.
        if (tid == 0) s_item = atomicAdd(a.slot_count, 1);
        __syncthreads();
        const int bs = s_item;
        if (bs < a.slot_capacity) {
          uint8_t *dst = a.slots + (size_t)bs * N;
          for (int j = 0; j < S::CPT; j++) {
            const int k = tid + j * T;
            if (k < N) dst[k] = (uint8_t)((decBits[k >> 5] >> (k & 31)) & 1u);
          }
        }
        bcaseSlot = bs + 1;
        __syncthreads();
      }
      if (mode1 && sameCount == 4) { status = 0; iters = t; break; }
      if (t > 0 && !anyFail) { status = 1; iters = t; break; }
      if (s_cancel) { status = 3; iters = t; break; }
      if (t == a.max_iter) { status = 0; iters = t; break; }

      // ---- variable-node pass: c2v, overall LLR, decision, next v2c (:512-548)
      // The next column's rows and channel LLR are fetched while this
      // column's six phi evaluations run.
      bool changed = false;
      ColEdges rowsNext = tid < N ? gcol[tid] : ColEdges{{0u, 0u, 0u}};
      gfloat_t *gLLR = as_global(pd.LLR);
      float llrNext = tid < N ? gLLR[tid] : 0.0f;
      for (int j = 0; j < S::CPT; j++) {
        const int k = tid + j * T;
        const ColEdges rows = rowsNext;
        const float llr = llrNext;
        if (j + 1 < S::CPT && k + T < N) {
          rowsNext = gcol[k + T];
          llrNext = gLLR[k + T];
        }
        uint32_t dec = 0;
        if (k < N) {
          float c2v[3];
          uint32_t ph[3];
          int row[3];
          uint32_t pos[3];
          col_edges(rows, row, pos);
          uint32_t R[3];
          float x[3], v[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            R[q] = f2u(rowTot[row[q]]);
            ph[q] = f2u(at_byte(phi, pos[q]));
          }
#pragma unroll
          for (int q = 0; q < 3; q++) x[q] = u2f(R[q] & 0x7fffffffu) - u2f(ph[q] & 0x7fffffffu);
#pragma unroll
          for (int q = 0; q < 3; q++) v[q] = phi_of(x[q]);
#pragma unroll
          for (int q = 0; q < 3; q++)
            // times -1 or 1 by the sign product; with no NaN (ParallelMode 0
            // never runs a NaN bitplane) that is the sign bit flipped
            c2v[q] = MODE1 ? (((R[q] ^ ph[q]) >> 31) ? v[q] * -1.0f : v[q] * 1.0f)
                           : u2f((f2u(v[q]) & 0x7fffffffu) | ((f2u(v[q]) ^ R[q] ^ ph[q]) & 0x80000000u));
          float tmp = llr;
          tmp += c2v[0];
          tmp += c2v[1];
          tmp += c2v[2];
          if (!mode1) {
            dec = f2u(tmp) >> 31;
          } else {
            dec = tmp < 0.0f;  // LLR_overall < 0 (LdpcaDecoder.cpp:257-260): -0 decides 0
            if (t == 0) changed |= !exact || as_global(pd.carry)[k] != (uint8_t)dec;
            else changed |= ((decBits[k >> 5] >> (k & 31)) & 1u) != dec;
          }
          if (dec) {
#pragma unroll
            for (int q = 0; q < 3; q++) {
              // word row >> 5 = bits 5..12 of the column word, extracted by one
              // v_bfe_u32 (kept as written, so that the word's address is one
              // shift-and-add); the bit's shift takes the row's low 5 bits
              uint32_t w;
              __asm__("v_bfe_u32 %0, %1, 5, 8" : "=v"(w) : "v"(rows.e[q]));
              atomicXor(&parBits[t & 1][w], 1u << (row[q] & 31));
            }
          }
          uint32_t sg[3];
#pragma unroll
          for (int q = 0; q < 3; q++) {
            const float v2c = tmp - c2v[q];
            sg[q] = f2u(v2c) >> 31;
            x[q] = (sg[q] ? -1.0f : 1.0f) * v2c;
          }
#pragma unroll
          for (int q = 0; q < 3; q++) v[q] = phi_of(x[q]);
#pragma unroll
          for (int q = 0; q < 3; q++) at_byte(phi, pos[q]) = copysignf(v[q], u2f(sg[q] << 31));
        }
        const uint64_t bal = __ballot(dec);
        if (lane == 0) {
          const int word = (j * T + wave * 64) >> 5;
          decBits[word] = (uint32_t)bal;
          decBits[word + 1] = (uint32_t)(bal >> 32);
        }
      }
      if (changed) s_changed[t & 1] = 1;
      __syncthreads();
    }

    // ---- converged: CRC-8 of the decisions (linear in the bits, crc8.h:5-6)
    int ok = 0;
    if (status == 1) {
      uint32_t crc = 0;
      for (int j = 0; j < S::CPT; j++) {
        int k = tid + j * T;
        if (k < N && ((decBits[k >> 5] >> (k & 31)) & 1u)) crc ^= L.crcBasis[k];
      }
      for (int o = 32; o > 0; o >>= 1) crc ^= __shfl_xor(crc, o);
      if (lane == 0) s_crc[wave] = crc;
      __syncthreads();
      uint32_t tot = 0;
      for (int i = 0; i < T / 64; i++) tot ^= s_crc[i];
      ok = (tot == pd.crc);
      if (ok) {
        int slot = -1;
        if (tid == 0) {
          slot = atomicAdd(a.slot_count, 1);
          s_item = slot;  // reuse as broadcast
        }
        __syncthreads();
        slot = s_item;
        if (slot < a.slot_capacity) {
          uint8_t *dst = a.slots + (size_t)slot * N;
          for (int j = 0; j < S::CPT; j++) {
            int k = tid + j * T;
            if (k < N) dst[k] = (uint8_t)((decBits[k >> 5] >> (k & 31)) & 1u);
          }
        }
        if (tid == 0) {
          pd.rec[level] = make_int4(iters, 2, slot, 0);
          atomicMin(&st->best, level);
        }
      }
    }
    if (!ok && mode1 && status != 3) {
      // ParallelMode 1 keeps every run code's final decisions: they are the
      // decisions carried into the next code
      if (tid == 0) s_item = atomicAdd(a.slot_count, 1);
      __syncthreads();
      const int slot = s_item;
      if (slot < a.slot_capacity) {
        uint8_t *dst = a.slots + (size_t)slot * N;
        for (int j = 0; j < S::CPT; j++) {
          const int k = tid + j * T;
          if (k < N) dst[k] = (uint8_t)((decBits[k >> 5] >> (k & 31)) & 1u);
        }
      }
      if (tid == 0) pd.rec[level] = make_int4(iters, status == 1 ? 1 : 0, slot, 0);
    } else if (tid == 0 && !ok) {
      pd.rec[level] = make_int4(iters, status == 1 ? 1 : (status == 3 ? 3 : 0), -1, 0);
    }
    if (tid == 0 && mode1) pd.bcase[level] = bcaseSlot;
    __syncthreads();
  }
}

// ------------------------------------------------------------------ ParallelMode 1 resolution
// After a speculative ParallelMode 1 launch, walks each bitplane's codes from
// the queue's first code with the decisions actually carried from code to
// code (LdpcaDecoder.cpp:646-771 with beliefPropagation_OpenMP).  A code that
// recorded the repeat case fails after 4 iterations when the carried
// decisions equal its recorded ones; otherwise its speculative outcome
// stands.  The first accepted code becomes `best`.  When the walk reaches a
// code the launch cancelled or never ran, the bitplane is queued again from
// that code with the carried decisions now known (rerun_count counts them).
// One wave per bitplane.
// A slot index at or beyond slot_capacity (its decisions were not stored),
// for a failed, mismatching or accepted code or a repeat case, adds 1 << 20
// to rerun_count, which the host reports as exhausted slots.
__global__ __launch_bounds__(64) void mode1_resolve_kernel(const PlaneDesc *planes, int nplanes, int numCodes, int N,
                                                           const uint8_t *slots, int slot_capacity, int *rerun_count) {
  const int p = blockIdx.x;
  if (p >= nplanes) return;
  const PlaneDesc d = planes[p];
  PlaneState *st = d.state;
  const int run0 = st->run, lane = threadIdx.x;
  for (int L = run0; L < numCodes - 1; L++) {
    const int4 r = d.rec[L];
    if (r.y == REC_NOT_RUN || r.y == REC_CANCELLED) {
      if (lane == 0) {
        st->run = L;
        st->best = numCodes - 1;
        atomicAdd(rerun_count, 1);
      }
      return;
    }
    const int bs = L != run0 ? d.bcase[L] : 0;
    if (bs - 1 >= slot_capacity || r.z < 0 || r.z >= slot_capacity) {  // every record reached here names a slot
      if (lane == 0) atomicAdd(rerun_count, 1 << 20);
      return;
    }
    bool repeat = false;
    if (bs > 0) {
      const uint8_t *b = slots + (size_t)(bs - 1) * N;
      int diff = 0;
      for (int k = lane; k < N; k += 64) diff |= b[k] != d.carry[k];
      repeat = __ballot(diff) == 0;
    }
    if (!repeat && r.y == REC_ACCEPTED) {
      if (lane == 0) {
        st->best = L;
        st->run = numCodes;
      }
      return;
    }
    const int fin = repeat ? bs - 1 : r.z;
    if (repeat && lane == 0) d.rec[L] = make_int4(4, REC_FAILED, bs - 1, 0);
    const uint8_t *src = slots + (size_t)fin * N;
    for (int k = lane; k < N; k += 64) d.carry[k] = src[k];
  }
  if (lane == 0) {
    st->best = numCodes - 1;  // every BP code failed: full-rate inverse
    st->run = numCodes;
  }
}

// ------------------------------------------------------------------ full rate
// invH_mul (LdpcaDecoder.cpp:624-643) for every bitplane whose BP codes all
// failed, in two launches: invh_slot_kernel gives each such bitplane its
// decision slot (one thread per bitplane), then invh_kernel multiplies.

__global__ void invh_slot_kernel(const PlaneDesc *planes, int nplanes, int numCodes, int *slot_count) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= nplanes) return;
  const PlaneDesc d = planes[p];
  if (d.state->best < numCodes - 1) return;
  d.rec[numCodes - 1] = make_int4(0, REC_ACCEPTED, atomicAdd(slot_count, 1), 1);
}

// Grid (ceil(N / INVH_ROWS), INVH_COLS).  Every wave finds the bitplanes that
// need the inverse with a ballot over the plane list; column y takes the
// k-th of them for k = y, y + INVH_COLS, ...  Workgroup x computes rows
// [x * INVH_ROWS, + INVH_ROWS) of the product, INVH_ROWS / 4 per wave: a row
// of invH (N bits, MSB-first bytes, rows padded to 32-bit words) is read as
// little-endian words by consecutive lanes, ANDed with the byte-swapped
// syndrome words (a common byte permutation leaves the parity of the AND
// unchanged), and the parity is the popcount of the wave's ballot.
constexpr int INVH_ROWS = 64, INVH_COLS = 8, INVH_THREADS = 256;
constexpr int INVH_MAX_WORDS = (6336 + 31) / 32;  // <= 4 words per lane

__global__ __launch_bounds__(INVH_THREADS) void invh_kernel(const PlaneDesc *planes, int nplanes, LadderDev L,
                                                            uint8_t *slots, int slot_capacity) {
  __shared__ uint32_t sw[INVH_MAX_WORDS];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int N = L.N, nw = (N + 31) / 32, C = L.numCodes;
  const int row0 = blockIdx.x * INVH_ROWS + wave * (INVH_ROWS / 4);
  int k = 0;  // bitplanes needing the inverse seen so far (wave-uniform)
  for (int base = 0; base < nplanes; base += 64) {
    const int p = base + lane;
    const bool need = p < nplanes && !(planes[p].state->best < C - 1);
    uint64_t mask = __ballot(need);
    while (mask) {
      const int b = base + __builtin_ctzll(mask);
      mask &= mask - 1;
      if (k++ % INVH_COLS != (int)blockIdx.y) continue;
      const PlaneDesc d = planes[b];
      gbyte_t *acc = as_global(d.acc);
      // full-rate syndrome s_k = a_k ^ a_{k-1} (identity transmission pattern),
      // as MSB-first 32-bit words, bits at k >= N cleared
      __syncthreads();  // the previous bitplane's words are no longer read
      for (int t = tid; t < nw; t += INVH_THREADS) {
        uint32_t a = 0;
        for (int q = 0; q < 4; q++) {
          const int byte = 4 * t + q;
          a = (a << 8) | (byte < N / 8 ? (uint32_t)acc[byte] : 0u);
        }
        const uint32_t prev = t > 0 ? (uint32_t)acc[4 * t - 1] & 1u : 0u;
        uint32_t s = a ^ (a >> 1) ^ (prev << 31);
        const int valid = N - 32 * t;
        if (valid < 32) s &= ~0u << (32 - valid);
        sw[t] = __builtin_bswap32(s);
      }
      __syncthreads();
      uint32_t w[4];
#pragma unroll
      for (int q = 0; q < 4; q++) w[q] = lane + 64 * q < nw ? sw[lane + 64 * q] : 0u;
      const int slot = d.rec[C - 1].z;
      uint8_t *dst = slot >= 0 && slot < slot_capacity ? slots + (size_t)slot * N : nullptr;
      uint32_t mine = 0;  // parity of row row0 + lane (lane < INVH_ROWS / 4)
#pragma unroll 4
      for (int i = 0; i < INVH_ROWS / 4; i++) {
        const int j = row0 + i;
        if (j >= N) break;  // wave-uniform
        const uint32_t *row = (const uint32_t *)(L.invH + (size_t)j * nw * 4);
        uint32_t x = 0;
#pragma unroll
        for (int q = 0; q < 4; q++)
          if (lane + 64 * q < nw) x ^= as_global(row)[lane + 64 * q] & w[q];
        const uint32_t par = __popcll(__ballot(__popc(x) & 1u)) & 1u;
        if (lane == i) mine = par;
      }
      if (dst && lane < INVH_ROWS / 4 && row0 + lane < N) dst[row0 + lane] = (uint8_t)mine;
    }
  }
}

// One workgroup per bitplane: its accepted code's decisions from their slot
// to out + p * N (16-byte words; slots and out are 16-byte aligned).
__global__ void gather_accepted_kernel(const PlaneDesc *planes, int nplanes, int N, const uint8_t *slots,
                                       int slot_capacity, uint8_t *out) {
  const int p = blockIdx.x;
  if (p >= nplanes) return;
  const PlaneDesc d = planes[p];
  const int best = d.state->best;
  if (best < 0) return;
  const int4 r = d.rec[best];
  if (r.y != REC_ACCEPTED || r.z < 0 || r.z >= slot_capacity) return;
  const uint4 *src = (const uint4 *)(slots + (size_t)r.z * N);
  uint4 *dst = (uint4 *)(out + (size_t)p * N);
  for (int i = threadIdx.x; i < N / 16; i += blockDim.x) dst[i] = src[i];
}

// Clears n ints.  Used instead of hipMemsetAsync for the launch counters so a
// captured step graph holds kernel nodes only.
__global__ void zero_ints_kernel(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ------------------------------------------------------------------ host side
hipError_t launch_mode1_resolve(const PlaneDesc *planes, int n, int numCodes, int N, const uint8_t *slots,
                                int slot_capacity, int *rerun_count, hipStream_t s) {
  hipLaunchKernelGGL(mode1_resolve_kernel, dim3(n), dim3(64), 0, s, planes, n, numCodes, N, slots, slot_capacity,
                     rerun_count);
  return hipGetLastError();
}

hipError_t launch_zero_ints(int *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 63) / 64), dim3(64), 0, s, p, n);
  return hipGetLastError();
}

template <int N, int T>
static hipError_t launch_bp(const BpArgs &a, int grid, hipStream_t s) {
  if (a.mode == 1) hipLaunchKernelGGL((bp_persistent_kernel<N, T, true>), dim3(grid), dim3(T), 0, s, a);
  else hipLaunchKernelGGL((bp_persistent_kernel<N, T, false>), dim3(grid), dim3(T), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_bp_any(int N, int threads, const BpArgs &a, int grid, hipStream_t s) {
  if (!a.lad.regular) return launch_bp_irregular(a, grid, s);
  if (N == 6336) {
    if (threads == 704) return launch_bp<6336, 704>(a, grid, s);
    if (threads == 512) return launch_bp<6336, 512>(a, grid, s);
    return launch_bp<6336, 1024>(a, grid, s);
  }
  if (threads == 512) return launch_bp<1584, 512>(a, grid, s);
  return launch_bp<1584, 256>(a, grid, s);
}

hipError_t launch_ce_llr(const PlaneDesc *d_planes, int n, int fixed_start, const Ladder &l1584,
                         const Ladder &l6336, hipStream_t s) {
  hipLaunchKernelGGL(ce_llr_kernel, dim3(n), dim3(256), 0, s, d_planes, n, fixed_start, l1584.d_codeRate,
                     l6336.d_codeRate, l1584.numCodes, l6336.numCodes);
  return hipGetLastError();
}

hipError_t launch_invh(const PlaneDesc *d_planes, int n, const LadderDev &L, uint8_t *slots, int *slot_count,
                       int slot_capacity, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(invh_slot_kernel, dim3((n + 63) / 64), dim3(64), 0, s, d_planes, n, L.numCodes, slot_count);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(invh_kernel, dim3((L.N + INVH_ROWS - 1) / INVH_ROWS, INVH_COLS), dim3(INVH_THREADS), 0, s,
                     d_planes, n, L, slots, slot_capacity);
  return hipGetLastError();
}

hipError_t launch_gather_accepted(const PlaneDesc *d_planes, int n, int N, const uint8_t *slots, int slot_capacity,
                                  uint8_t *out, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(gather_accepted_kernel, dim3(n), dim3(128), 0, s, d_planes, n, N, slots, slot_capacity, out);
  return hipGetLastError();
}

LadderDev ladder_dev(const Ladder &L) {
  LadderDev d;
  d.N = L.N; d.E = L.E; d.numCodes = L.numCodes;
  d.regular = L.regular ? 1 : 0;
  d.colRows = L.d_colRows; d.rowPtr = L.d_rowPtr; d.tx = L.d_tx;
  d.colEdges = L.d_colEdges; d.jc = L.d_jc; d.colSlots = L.d_colSlots; d.colSlotsPerThread = L.colSlotsPerThread;
  d.syndLength = L.d_syndLength; d.invH = L.d_invH; d.crcBasis = L.d_crcBasis;
  return d;
}

}  // namespace dvc
