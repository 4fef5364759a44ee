// ldpca_irreg.hip -- sum-product BP on an irregular LDPCA ladder, whose
// columns have different degrees (irReg1584.lad: 66 codes, N = 1584,
// E = 7596, column degrees 2..21; the reference's alternative ladder at
// DvcDec/src/main.cpp:84-85).
//
// The arithmetic is beliefPropagation's (LdpcaDecoder.cpp:462-580) and
// beliefPropagation_OpenMP's (:164-306), exactly as the regular kernel
// (ldpca.hip) evaluates it: phi(x) = -logf(tanhf((x + 1e-9f)/2)) by
// dvc_bp_phi_fast, check sums in ascending edge order, variable sums
// LLR + c2v of the column's edges in jc order, the same decision, parity and
// CRC rules, and the same handling of NaN likelihood ratios.  What differs
// is the column walk: a column's edges come from the ladder's column pointers
// (jc) and per-edge (row, rank in row) words, and the variable pass runs in
// two sweeps over them -- the first forms each c2v and the column total and
// parks the c2v in the edge's message slot, the second turns the slot into
// the next v2c's phi -- because a degree-21 column's messages do not fit in
// registers.  Columns are dealt to the workgroup's threads by the ladder
// upload so that each thread carries nearly the same number of edges.
//
// Batched form and cancellation are those of bp_persistent_kernel: a
// persistent grid walks (bitplane, code offset) items in offset-major order,
// skips codes above the lowest accepted one and stops running ones once a
// lower code is accepted.  LDS: 42 KB per codeword (3 per CU).
#include <stdio.h>
#include "dvc_internal.h"
#include "dvc_math.h"
#include "ldpca_kernels.h"

#pragma clang fp contract(off)

namespace dvc {

constexpr int IR_NMAX = 1584, IR_EMAX = 8192;

__device__ __forceinline__ uint32_t ir_f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float ir_u2f(uint32_t u) { return __uint_as_float(u); }

// Accumulated-syndrome bit i (MSB-first bytes, ByteToBit LdpcaDecoder.cpp:81-107).
__device__ __forceinline__ uint32_t ir_acc_bit(const uint8_t *acc, int i) {
  return (acc[i >> 3] >> (7 - (i & 7))) & 1u;
}

template <bool MODE1>
__global__ __launch_bounds__(kIrregThreads) void bp_irregular_kernel(BpArgs a) {
  constexpr int T = kIrregThreads;
  // phi's logf table: (invc * 2^-k, k*ln2 + logc) for each (k, i) that phi's
  // arguments reach, read with per-lane indices
  __shared__ __align__(16) double s_phitab[2 * DVC_PHI_LOGTAB];
  // by check row; during the variable pass an edge's c2v.  3 spare slots at
  // the end, so that the check pass may read a row's last slots unconditionally
  __shared__ float phi[IR_EMAX + 3];
  __shared__ float rowTot[IR_NMAX];
  __shared__ uint16_t rowPtr[IR_NMAX + 2];
  __shared__ uint8_t dec[IR_NMAX];
  __shared__ uint32_t synBits[(IR_NMAX + 31) / 32];
  __shared__ uint32_t parBits[2][(IR_NMAX + 31) / 32];
  __shared__ int s_item, s_cancel;
  __shared__ uint32_t s_crc[T / 64];
  __shared__ int s_changed[2];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const LadderDev L = a.lad;
  const int N = L.N, E = L.E, S = L.colSlotsPerThread, synWords = (N + 31) / 32;
  const int16_t *mySlots = L.colSlots + (size_t)tid * S;
  for (int e = tid; e < DVC_PHI_LOGTAB; e += T) dvc_phi_logtab_entry(e, s_phitab);

  for (;;) {
    // ---- grab the next item (thread 0), broadcast through LDS
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
    if (!MODE1 && taint) {  // a NaN LLR: every code fails after max_iter (as in ldpca.hip)
      if (tid == 0) pd.rec[level] = make_int4(a.max_iter, REC_FAILED, -1, 0);
      __syncthreads();
      continue;
    }
    auto phi_of = [&](float x) {
      const float v = dvc_bp_phi_fast(x, s_phitab);
      return (MODE1 && taint && x != x) ? x : v;
    };
    const bool mode1 = MODE1, exact = level == run;
    const int m = L.syndLength[level];
    const uint32_t *gedge = L.colEdges + (size_t)level * E;
    const uint16_t *growPtr = L.rowPtr + (size_t)level * (N + 1);
    const float *gLLR = pd.LLR;

    // ---- row starts and the syndrome (LdpcaDecoder.cpp:684-693)
    for (int i = tid; i <= m; i += T) rowPtr[i] = growPtr[i];
    for (int wd = tid; wd < synWords; wd += T) {
      uint32_t bits = 0;
      for (int q = 0; q < 32; q++) {
        const int k = wd * 32 + q;
        if (k < m) {
          const int cur = L.tx[(size_t)level * N + k];
          uint32_t s = ir_acc_bit(pd.acc, cur);
          if (k > 0) s ^= ir_acc_bit(pd.acc, L.tx[(size_t)level * N + k - 1]);
          bits |= s << q;
        }
      }
      synBits[wd] = bits;
    }
    // ---- variable-to-check messages start at the channel LLR (:478-481)
    for (int i = 0; i < S; i++) {
      const int k = mySlots[i];
      if (k < 0) break;
      const float v = gLLR[k];
      const uint32_t s = ir_f2u(v) >> 31;
      const float mag = (s ? -1.0f : 1.0f) * v;
      const float f = ir_u2f((ir_f2u(phi_of(mag)) & 0x7fffffffu) | (s << 31));
      for (int l = L.jc[k]; l < L.jc[k + 1]; l++) {
        const uint32_t e = gedge[l];
        phi[growPtr[e & 0xffffu] + (e >> 16)] = f;
      }
    }
    __syncthreads();

    int status = 0, iters = 0;
    int sameCount = 0, bcaseSlot = 0;
    for (int t = 0;; t++) {
      // ---- check pass: |phi| sums and sign products; parity of the previous
      // decisions (:495-515, :551-568)
      int fail = 0;
      const uint32_t *parPrev = parBits[(t + 1) & 1];
      for (int r = tid; r < m; r += T) {
        const int e0 = rowPtr[r], e1 = rowPtr[r + 1];
        const uint32_t syn = (synBits[r >> 5] >> (r & 31)) & 1u;
        uint32_t sx = 0u;  // XOR of the row's message words: the sign product in bit 31
        float sum = 0.0f;
        // groups of four loads without conditions, then the last 1..3 slots
        // read unconditionally too (rows are contiguous, phi has spare slots)
        // with those past the row's end taken as +0: sum + 0.0f is sum bit for
        // bit (sum is never -0)
        int j = e0;
        for (; j + 4 <= e1; j += 4) {
          uint32_t f[4];
#pragma unroll
          for (int i = 0; i < 4; i++) f[i] = ir_f2u(phi[j + i]);
#pragma unroll
          for (int i = 0; i < 4; i++) sum += ir_u2f(f[i] & 0x7fffffffu);
          sx ^= f[0] ^ f[1] ^ f[2] ^ f[3];
        }
        if (j < e1) {
          uint32_t f[3];
#pragma unroll
          for (int i = 0; i < 3; i++) f[i] = ir_f2u(phi[j + i]);
          f[1] = j + 1 < e1 ? f[1] : 0u;
          f[2] = j + 2 < e1 ? f[2] : 0u;
#pragma unroll
          for (int i = 0; i < 3; i++) sum += ir_u2f(f[i] & 0x7fffffffu);
          sx ^= f[0] ^ f[1] ^ f[2];
        }
        const uint32_t sgn = syn ^ (sx >> 31);
        rowTot[r] = ir_u2f(ir_f2u(sum) | (sgn << 31));
        if (t > 0 && (syn ^ ((parPrev[r >> 5] >> (r & 31)) & 1u))) fail = 1;
      }
      for (int wd = tid; wd < synWords; wd += T) parBits[t & 1][wd] = 0u;
      if (mode1 && t > 0) sameCount = s_changed[(t - 1) & 1] ? 0 : sameCount + 1;
      if (tid == 0) {
        s_changed[t & 1] = 0;
        s_cancel = (t % a.poll == 0) &&
                   __hip_atomic_load(&st->best, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < level;
      }
      const int anyFail = __syncthreads_or(fail);
      if (mode1 && !exact && t == 4 && sameCount == 3) {
        // iterations 0-3 decided alike: keep the decisions for the resolution
        if (tid == 0) s_item = atomicAdd(a.slot_count, 1);
        __syncthreads();
        const int bs = s_item;
        if (bs < a.slot_capacity) {
          uint8_t *dst = a.slots + (size_t)bs * N;
          for (int k = tid; k < N; k += T) dst[k] = dec[k];
        }
        bcaseSlot = bs + 1;
        __syncthreads();
      }
      if (mode1 && sameCount == 4) { status = 0; iters = t; break; }
      if (t > 0 && !anyFail) { status = 1; iters = t; break; }
      if (s_cancel) { status = 3; iters = t; break; }
      if (t == a.max_iter) { status = 0; iters = t; break; }

      // ---- variable pass (:512-548): c2v of every edge and the column total
      // in jc order, the decision, then each edge's next v2c
      bool changed = false;
      for (int i = 0; i < S; i++) {
        const int k = mySlots[i];
        if (k < 0) break;
        const int l0 = L.jc[k], l1 = L.jc[k + 1];
        float tmp = gLLR[k];
        for (int l = l0; l < l1; l++) {
          const uint32_t e = gedge[l];
          const int row = e & 0xffffu, pos = rowPtr[row] + (e >> 16);
          const uint32_t R = ir_f2u(rowTot[row]), ph = ir_f2u(phi[pos]);
          const float x = ir_u2f(R & 0x7fffffffu) - ir_u2f(ph & 0x7fffffffu);
          const float v = phi_of(x);
          const float c2v = ((R ^ ph) >> 31) ? v * -1.0f : v * 1.0f;
          tmp += c2v;
          phi[pos] = c2v;
        }
        uint32_t d;
        if (!mode1) {
          d = ir_f2u(tmp) >> 31;
        } else {
          d = tmp < 0.0f;  // LLR_overall < 0 (LdpcaDecoder.cpp:257-260)
          if (t == 0) changed |= !exact || pd.carry[k] != (uint8_t)d;
          else changed |= dec[k] != (uint8_t)d;
        }
        dec[k] = (uint8_t)d;
        for (int l = l0; l < l1; l++) {
          const uint32_t e = gedge[l];
          const int row = e & 0xffffu, pos = rowPtr[row] + (e >> 16);
          const float v2c = tmp - phi[pos];
          const uint32_t sg = ir_f2u(v2c) >> 31;
          const float v = phi_of((sg ? -1.0f : 1.0f) * v2c);
          phi[pos] = ir_u2f((ir_f2u(v) & 0x7fffffffu) | (sg << 31));
          if (d) atomicXor(&parBits[t & 1][row >> 5], 1u << (row & 31));
        }
      }
      if (changed) s_changed[t & 1] = 1;
      __syncthreads();
    }

    // ---- converged: CRC-8 of the decisions (linear in the bits, crc8.h:5-6)
    int ok = 0;
    if (status == 1) {
      uint32_t crc = 0;
      for (int k = tid; k < N; k += T)
        if (dec[k]) crc ^= L.crcBasis[k];
      for (int o = 32; o > 0; o >>= 1) crc ^= __shfl_xor(crc, o);
      if (lane == 0) s_crc[wave] = crc;
      __syncthreads();
      uint32_t tot = 0;
      for (int i = 0; i < T / 64; i++) tot ^= s_crc[i];
      ok = (tot == pd.crc);
      if (ok) {
        if (tid == 0) s_item = atomicAdd(a.slot_count, 1);
        __syncthreads();
        const int slot = s_item;
        if (slot < a.slot_capacity) {
          uint8_t *dst = a.slots + (size_t)slot * N;
          for (int k = tid; k < N; k += T) dst[k] = dec[k];
        }
        if (tid == 0) {
          pd.rec[level] = make_int4(iters, REC_ACCEPTED, slot, 0);
          atomicMin(&st->best, level);
        }
      }
    }
    if (!ok && mode1 && status != 3) {
      // ParallelMode 1 keeps every run code's final decisions (the next code's carry)
      if (tid == 0) s_item = atomicAdd(a.slot_count, 1);
      __syncthreads();
      const int slot = s_item;
      if (slot < a.slot_capacity) {
        uint8_t *dst = a.slots + (size_t)slot * N;
        for (int k = tid; k < N; k += T) dst[k] = dec[k];
      }
      if (tid == 0) pd.rec[level] = make_int4(iters, status == 1 ? REC_CRC_MISMATCH : REC_FAILED, slot, 0);
    } else if (tid == 0 && !ok) {
      pd.rec[level] = make_int4(iters, status == 1 ? REC_CRC_MISMATCH : (status == 3 ? REC_CANCELLED : REC_FAILED),
                                -1, 0);
    }
    if (tid == 0 && mode1) pd.bcase[level] = bcaseSlot;
    __syncthreads();
  }
}

hipError_t launch_bp_irregular(const BpArgs &a, int grid, hipStream_t s) {
  if (a.lad.N > IR_NMAX || a.lad.E > IR_EMAX || !a.lad.colSlots) return hipErrorInvalidValue;
  if (a.mode == 1)
    hipLaunchKernelGGL(bp_irregular_kernel<true>, dim3(grid), dim3(kIrregThreads), 0, s, a);
  else
    hipLaunchKernelGGL(bp_irregular_kernel<false>, dim3(grid), dim3(kIrregThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace dvc
