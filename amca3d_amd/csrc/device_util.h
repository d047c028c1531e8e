// Device helpers shared by the CA step kernels: coherent loads, 1-bit cell
// maps, slot geometry, block scans, the decoupled look-back and XCD-aware
// block numbering.
#pragma once
#include <hip/hip_runtime.h>
#include <climits>

#include "plain_step.h"

namespace amca {

// A look-back flag, published after its payload.  The payload is written with
// agent-scope atomic stores (through to the coherent level, sc1), so the
// flag needs only those stores to have completed before it is issued:
// s_waitcnt vmcnt(0) (gfx9 counts stores in vmcnt; 0x0F70 = vmcnt(0), expcnt
// and lgkmcnt not waited).  A release store would also write back the whole
// L2 of the XCD (buffer_wbl2), which the payload does not need.  The signal
// fences keep the compiler from moving the stores across the wait.
__device__ __forceinline__ void st_flag_after_payload(unsigned* p, unsigned v) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// A look-back flag read: its payload is read only after the flag's value is
// tested, with agent-scope atomic loads, so no acquire (which would also
// invalidate the L2) is needed.
__device__ __forceinline__ unsigned ld_flag(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_relaxed(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// loads that bypass the compute unit's L1: words that atomics update during
// the same kernel (claims, liquid map) are read at the L2
__device__ __forceinline__ int ld_coherent(const int* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned ld_coherent(const unsigned* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void slot_delta(int s, int& di, int& dj, int& dk) {
  int idx = s < 13 ? s : s + 1;  // skip the centre of the 3x3x3 block
  dk = idx / 9 - 1;
  dj = (idx / 3) % 3 - 1;
  di = idx % 3 - 1;
}

__device__ __forceinline__ void cell_ijk(const DevParams& P, int c, int& i, int& j, int& k) {
  k = c / P.nxyG;
  int rem = c - k * P.nxyG;
  j = rem / P.nxG;
  i = rem - j * P.nxG;
}

__device__ __forceinline__ unsigned long long ld_coherent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---- 4x4x4 brick map: word 0 liquid, word 1 claimed ----------------------
__device__ __forceinline__ int brick_word(const DevParams& P, int i, int j, int k) {
  return 2 * (((k >> 2) * P.nbY + (j >> 2)) * P.nbX + (i >> 2));
}
__device__ __forceinline__ unsigned long long brick_bit(int i, int j, int k) {
  return 1ull << (((k & 3) << 4) | ((j & 3) << 2) | (i & 3));
}
// Claim words in the same 4x4x4 bricks, 64 int32 (two 128-byte lines) per brick,
// so that a 3x3x3 neighbourhood touches few lines.
__device__ __forceinline__ long long claim_index(const DevParams& P, int i, int j, int k) {
  return 64ll * (((long long)(k >> 2) * P.nbY + (j >> 2)) * P.nbX + (i >> 2)) +
         (((k & 3) << 4) | ((j & 3) << 2) | (i & 3));
}
__device__ __forceinline__ void brick_clear(const DevParams& P, unsigned long long* bm, int word,
                                            int i, int j, int k) {
  atomicAnd(&bm[brick_word(P, i, j, k) + word], ~brick_bit(i, j, k));
}
__device__ __forceinline__ void brick_set(const DevParams& P, unsigned long long* bm, int word,
                                          int i, int j, int k) {
  atomicOr(&bm[brick_word(P, i, j, k) + word], brick_bit(i, j, k));
}

// ---- per-work-group staging of map-word updates ----------------------------
// Bits a list kernel sets in 64-bit map words are gathered per work-group in
// an LDS table keyed by word index and applied with one memory atomic per
// distinct word when the work-group is done (every device-scope atomic runs at
// the memory side, so this saves most of them: neighbouring entries set bits
// in the same words).  Only later kernels read those words.  A word that finds
// no free entry within kStageProbe probes goes straight to memory.
constexpr int kStageBits = 9;
constexpr int kStageSlots = 1 << kStageBits;
constexpr int kStageProbe = 8;
struct WordStage {
  int key[kStageSlots];
  unsigned long long bits[kStageSlots];
};
__device__ __forceinline__ void stage_init(WordStage& s) {
  for (int i = threadIdx.x; i < kStageSlots; i += blockDim.x) {
    s.key[i] = -1;
    s.bits[i] = 0ull;
  }
}
// false: no entry for `word` (the caller applies the bits itself)
__device__ __forceinline__ bool stage_add(WordStage& s, int word, unsigned long long b) {
  unsigned h = ((unsigned)word * 2654435761u) >> (32 - kStageBits);  // Knuth's multiplicative hash
  for (int p = 0; p < kStageProbe; p++) {
    const int k = atomicCAS(&s.key[h], -1, word);
    if (k == -1 || k == word) {
      atomicOr(&s.bits[h], b);
      return true;
    }
    h = (h + 1) & (kStageSlots - 1);
  }
  return false;
}
// after a barrier: one atomic OR per used entry
__device__ __forceinline__ void stage_flush_or(const WordStage& s, unsigned long long* words) {
  for (int i = threadIdx.x; i < kStageSlots; i += blockDim.x)
    if (s.key[i] >= 0) atomicOr(&words[s.key[i]], s.bits[i]);
}

// The 26 neighbour bits of interior cell (i, j, k) from word `kWord` of the
// brick map, in slot order (k, j, i, centre skipped).  The 3x3x3 block spans
// bricks bx0..bx1 (bx1 = bx0 or bx0 + 1) per axis: the 8 corner words are
// loaded once and each of the 9 rows takes its 3 bits from the two x-adjacent
// nibbles.
template <int kWord, bool kCoherent>
__device__ __forceinline__ unsigned brick_bits26(const DevParams& P, const unsigned long long* bm,
                                                 int i, int j, int k) {
  const int bx0 = (i - 1) >> 2, bx1 = (i + 1) >> 2;
  const int by0 = (j - 1) >> 2, by1 = (j + 1) >> 2;
  const int bz0 = (k - 1) >> 2, bz1 = (k + 1) >> 2;
  unsigned long long w[8];
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const int bx = (q & 1) ? bx1 : bx0;
    const int by = (q & 2) ? by1 : by0;
    const int bz = (q & 4) ? bz1 : bz0;
    const unsigned long long* p = bm + 2 * ((bz * P.nbY + by) * P.nbX + bx) + kWord;
    w[q] = kCoherent ? ld_coherent(p) : *p;
  }
  const int xs = (i - 1) & 3;  // first x of the block inside brick bx0
  unsigned m = 0;
  int s = 0;
#pragma unroll
  for (int dk = -1; dk <= 1; dk++) {
    const int z = k + dk;
    const bool hz = (z >> 2) != bz0;
#pragma unroll
    for (int dj = -1; dj <= 1; dj++) {
      const int y = j + dj;
      const bool hy = (y >> 2) != by0;
      const unsigned long long lo = hz ? (hy ? w[6] : w[4]) : (hy ? w[2] : w[0]);
      const unsigned long long hi = hz ? (hy ? w[7] : w[5]) : (hy ? w[3] : w[1]);
      const int sh = ((z & 3) << 4) | ((y & 3) << 2);
      const unsigned row8 = (unsigned)((lo >> sh) & 0xfull) | ((unsigned)((hi >> sh) & 0xfull) << 4);
      const unsigned b = (row8 >> xs) & 7u;
      if (dk == 0 && dj == 0) {
        m |= (b & 1u) << s;
        m |= ((b >> 2) & 1u) << (s + 1);
        s += 2;
      } else {
        m |= b << s;
        s += 3;
      }
    }
  }
  return m;
}

// ---- per-cell slots (brick-major, the claim words' order) -----------------
__device__ __forceinline__ unsigned fast_div(unsigned n, const FastDiv& f) {
  const unsigned t = __umulhi(f.m, n);
  return (t + ((n - t) >> f.s1)) >> f.s2;
}
// Brick coordinates of brick b (b < 2^32).
__device__ __forceinline__ void brick_xyz(const DevParams& P, long long b, int& bx, int& by,
                                          int& bz) {
  const unsigned ub = (unsigned)b;
  const unsigned r = fast_div(ub, P.fdX);
  bx = (int)(ub - r * (unsigned)P.nbX);
  const unsigned z = fast_div(r, P.fdY);
  by = (int)(r - z * (unsigned)P.nbY);
  bz = (int)z;
}

// wave-wide sum
__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Exclusive block scan of packed counter pairs (hi 32 | lo 32).  Both halves
// stay below 2^32, so packed addition never carries between them.
__device__ inline unsigned long long block_scan_excl(unsigned long long x, unsigned long long* total,
                                              unsigned long long* sm) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  unsigned long long v = x;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    unsigned long long y = __shfl_up(v, o, 64);
    if (lane >= o) v += y;
  }
  if (lane == 63) sm[w] = v;
  __syncthreads();
  unsigned long long base = 0, tot = 0;
  for (int q = 0; q < nw; q++) {
    unsigned long long s = sm[q];
    if (q < w) base += s;
    tot += s;
  }
  __syncthreads();
  *total = tot;
  return base + v - x;
}

// Single-pass decoupled look-back, run by the 64 lanes of the block's first
// wave: publishes this block's aggregate, then probes 64 predecessors at a
// time, summing aggregates back to the nearest published inclusive prefix.
// Returns the exclusive prefix (to every lane).  `epoch` makes entries of
// earlier launches invisible.
__device__ inline unsigned long long lookback(unsigned* flag, unsigned long long* agg,
                                       unsigned long long* pref, int bid, unsigned epoch,
                                       unsigned long long aggregate) {
  const int lane = threadIdx.x;
  const unsigned tag = (epoch & 0x3fffffffu) << 2;
  if (bid == 0) {
    if (lane == 0) {
      st_relaxed(&pref[0], aggregate);
      st_flag_after_payload(&flag[0], tag | 2u);
    }
    return 0ull;
  }
  if (lane == 0) {
    st_relaxed(&agg[bid], aggregate);
    st_flag_after_payload(&flag[bid], tag | 1u);
  }
  unsigned long long excl = 0;
  int base = bid - 1;
  while (true) {
    const int p = base - lane;
    bool ready, is_pref;
    if (p >= 0) {
      const unsigned f = ld_flag(&flag[p]);
      ready = (f & ~3u) == tag && (f & 3u) != 0u;
      is_pref = ready && (f & 3u) == 2u;
    } else {
      ready = true;  // before block 0: an empty prefix
      is_pref = true;
    }
    const unsigned long long rm = __ballot(ready);
    const unsigned long long pm = __ballot(is_pref);
    unsigned long long need = ~0ull;
    int first = 64;
    if (pm) {
      first = __builtin_ctzll(pm);
      need = (first == 63) ? ~0ull : ((1ull << (first + 1)) - 1ull);
    }
    if ((rm & need) != need) {
      __builtin_amdgcn_s_sleep(1);
      continue;
    }
    unsigned long long v = 0;
    if (p >= 0 && lane <= first) v = (lane == first) ? ld_relaxed(&pref[p]) : ld_relaxed(&agg[p]);
    excl += wave_sum(v);
    if (pm) break;
    base -= 64;
  }
  if (lane == 0) {
    st_relaxed(&pref[bid], excl + aggregate);
    st_flag_after_payload(&flag[bid], tag | 2u);
  }
  return excl;
}

// Contiguous chunk of a list of n entries owned by block `bid` of `nb`,
// rounded to whole tiles of `tile` entries.
__device__ __forceinline__ void chunk_of(int n, int bid, int nb, int tile, int& lo, int& hi) {
  int per = (n + nb - 1) / nb;
  per = (per + tile - 1) / tile * tile;
  long long l = (long long)bid * per;
  lo = (int)(l < n ? l : n);
  long long h = l + per;
  hi = (int)(h < n ? h : n);
}

// Block index with the eight XCDs each owning a contiguous range: work-groups
// are dispatched to XCDs round-robin, so block b runs on XCD b % 8.
__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x;
  if (nb & 7) return blockIdx.x;
  return (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
}

}  // namespace amca
