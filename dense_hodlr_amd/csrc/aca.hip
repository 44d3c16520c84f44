/*
 * Level-batched partial-pivot ACA fused with kernel-entry assembly.
 *
 * Restates partialPivACA_LowRankApprox (src/lowRank.cpp:117-279) and
 * chooseNextRowCol (:291-298) for every off-diagonal block of a set of tree
 * levels at once.  One ACA iteration of every active block is three launches:
 *
 *   row phase   residual row r = A(i,:) - W(i,:) V^T (kernel row evaluated on
 *               the fly), written to V(:,k); per-workgroup argmax |r| and the
 *               first unchosen column with |r| > minPivot
 *   col phase   pick the pivot column j (first max, else the first unchosen
 *               one), residual column c = A(:,j) - V(j,:) W^T, write
 *               w = c / r_j to W(:,k); argmax |c| and the first unchosen row
 *   finalize    the incremental Frobenius norm (:234-239), epsilon (:241),
 *               the next row and the stopping rule (:149, :226-227, :246);
 *               the last workgroup of a block to arrive does the scalar logic
 *               on partials reduced in a fixed order.
 *
 * Stopping test, one iteration late.  The norm update for term k needs the
 * dot products V^T v_k and W^T w_k.  The row and column passes of iteration
 * k+1 stream every stored column of V and W anyway, v_k and w_k included, so
 * they accumulate those dot products exactly in the same pass; the finalize
 * of iteration k+1 first completes epsilon for term k and, if term k ended the
 * reference's loop, discards iteration k+1's speculative results.  Each
 * iteration therefore reads each panel once.
 *
 * The residual sums run sequentially over l with the product rounded before
 * the sum (contraction off), the order and rounding of the reference's Eigen
 * loop `sum += W(i,l) * V.col(l)`, so residual entries and the stored factors
 * are bit-identical to the CPU restatement's.
 */
#include "aca.h"

#include <algorithm>
#include <vector>

namespace hodlr {

namespace {

constexpr int TPB = 256;   // threads per workgroup
constexpr int LCH = 16;    // l-chunk of the transposed partial reduction
constexpr int DOT_WIN = 128;  // columns whose dot products are combined per workgroup barrier
constexpr int FIN_L = 64;  // dot products per finalize workgroup (x4 slot groups)

__device__ inline void argmax_merge(double& bv, int& bi, double v, int i) {
  // larger value wins; on equal values the lower index wins (first max)
  if (v > bv || (v == bv && i >= 0 && (bi < 0 || i < bi))) {
    bv = v;
    bi = i;
  }
}

__device__ inline bool bit_test(const uint32_t* bits, int i) {
  return (bits[i >> 5] >> (i & 31)) & 1u;
}

__device__ inline void bit_set(uint32_t* bits, int i) { atomicOr(&bits[i >> 5], 1u << (i & 31)); }

/* Workgroup reduction of (max |.|, argmax, first-unchosen index). */
struct PivRed {
  double v;
  int i;
  int first;
};

__device__ PivRed wg_reduce_piv(PivRed p, PivRed* sh) {
  const int t = threadIdx.x;
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(p.v, off);
    const int oi = __shfl_xor(p.i, off);
    const int of = __shfl_xor(p.first, off);
    argmax_merge(p.v, p.i, ov, oi);
    if (of >= 0 && (p.first < 0 || of < p.first)) p.first = of;
  }
  const int lane = t & 63, wave = t >> 6;
  if (lane == 0) sh[wave] = p;
  __syncthreads();
  if (t == 0) {
    PivRed r = sh[0];
    for (int w = 1; w < TPB / 64; w++) {
      argmax_merge(r.v, r.i, sh[w].v, sh[w].i);
      if (sh[w].first >= 0 && (r.first < 0 || sh[w].first < r.first)) r.first = sh[w].first;
    }
    sh[0] = r;
  }
  __syncthreads();
  PivRed r = sh[0];
  __syncthreads();
  return r;
}

/* Partial slot layout: (cap + 4) doubles per workgroup slot:
 * [0, cap) dot partials, cap: max |.|, cap+1: argmax, cap+2: first unchosen. */
__device__ inline double* slot_ptr(const AcaLevels& L, const AcaBlock& b, int slot) {
  const int cap = L.cap[b.level];
  return L.P[b.level] + (size_t)(b.part_off + slot) * (size_t)(cap + 4);
}

__device__ inline void read_piv(const double* s, int cap, double& v, int& i, int& f) {
  v = s[cap];
  i = (int)s[cap + 1];
  f = (int)s[cap + 2];
}

/* Workgroup-parallel reduction of the pivot partials in slots
 * [s0, s0 + nslot): (max |.|, argmax, first unchosen).  p0 holds slot
 * s0 + threadIdx.x, already read ((-1, -1, -1) past nslot). */
__device__ PivRed reduce_slots(const AcaLevels& L, const AcaBlock& b, int s0, int nslot,
                               PivRed* sh, PivRed p0) {
  const int cap = L.cap[b.level];
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
  argmax_merge(p.v, p.i, p0.v, p0.i);
  if (p0.first >= 0) p.first = p0.first;
  for (int w = threadIdx.x + TPB; w < nslot; w += TPB) {
    double v;
    int i, f;
    read_piv(slot_ptr(L, b, s0 + w), cap, v, i, f);
    argmax_merge(p.v, p.i, v, i);
    if (f >= 0 && (p.first < 0 || f < p.first)) p.first = f;
  }
  return wg_reduce_piv(p, sh);
}

__device__ PivRed reduce_slots(const AcaLevels& L, const AcaBlock& b, int s0, int nslot,
                               PivRed* sh) {
  const int cap = L.cap[b.level];
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
  for (int w = threadIdx.x; w < nslot; w += TPB) {
    double v;
    int i, f;
    read_piv(slot_ptr(L, b, s0 + w), cap, v, i, f);
    argmax_merge(p.v, p.i, v, i);
    if (f >= 0 && (p.first < 0 || f < p.first)) p.first = f;
  }
  return wg_reduce_piv(p, sh);
}

/* The pivot column of this iteration from the row-phase partials: >= 0 the
 * column, -1 a row jump (max |r| <= minPivot), -2 no admissible column. */
__device__ int decide_column(const AcaLevels& L, const AcaBlock& b, const uint32_t* colbits,
                             int min_pivot, PivRed* sh) {
  const PivRed r = reduce_slots(L, b, 0, b.row_nwg, sh);
  if (r.v <= (double)min_pivot || r.v < HODLR_PIVOT_FLOOR) return -1;
  if (!bit_test(colbits, r.i)) return r.i;
  return r.first >= 0 ? r.first : -2;
}

/* Deterministic workgroup sum (fixed reduction shape). */
__device__ double wg_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.0;
  for (int w = 0; w < TPB / 64; w++) r += sh[w];
  __syncthreads();
  return r;
}

/* Raw buffer access (CDNA: 32-bit offsets from a 48-bit base, no format
 * conversion; the resource's fourth word is the gfx9 raw-buffer setting). */
constexpr int BUF_RSRC_WORD3 = 0x00020000;
__device__ __forceinline__ double buf_load_f64(__amdgpu_buffer_rsrc_t r, int voff, int soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

__device__ __forceinline__ double read_lane_f64(double v, int lane) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, lane);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

/* Strided slot sums: S(p, g, n) = sum over w = g, g+4, g+8, ... < n of
 * p[w * stride], as eight accumulators: value m (w = g + 4m) of the first
 * 8 F values, F = the groups of eight that fit whole (w + 28 < n), goes to
 * accumulator m mod 8, the rest to accumulator 0, each accumulator in
 * increasing m; then ((a0 + a1) + (a2 + a3)) + ((a4 + a5) + (a6 + a7)).
 * The loads of a round are all issued before its additions (the values sit
 * in the caches of other compute units or in HBM: one memory latency per
 * round instead of one per eight values).  Values past n are +0.0, which
 * leaves an accumulator unchanged: an accumulator starts at +0.0 and a
 * round-to-nearest sum is -0.0 only when both terms are.
 * Loads: buffer resource at the block's first slot, the lane's value offset
 * in a register, the slot's offset (w * stride bytes, uniform: g and m are
 * the same across the wave) in a scalar register. */
__device__ __forceinline__ int strided_full_groups(int n, int g) {
  return n - g - 29 >= 0 ? (n - g - 29) / 32 + 1 : 0;
}

/* zero: +0.0 that the compiler cannot fold (the caller's 0.0 * m for an
 * m >= 0, which for all it knows may be -0.0).  With a value it knows to be
 * +0.0 for the skipped loads it moves the first addition of each
 * accumulator (0.0 + v) next to its load, which then waits for that load
 * before the later ones are issued. */
template <int M>
__device__ __forceinline__ void strided_load(__amdgpu_buffer_rsrc_t rs, int voff, int g, int n,
                                             int stride_b, int m0, double zero, double (&v)[M]) {
#pragma unroll
  for (int q = 0; q < M; q++) {
    const int w = g + 4 * (m0 + q);
    v[q] = w < n ? buf_load_f64(rs, voff, w * stride_b) : zero;
  }
}

/* resource over a block's slots [s0, ...) */
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slot_rsrc(const AcaLevels& L, const AcaBlock& b,
                                                            int s0) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)slot_ptr(L, b, s0), (short)0, 0x7fffffff,
                                          BUF_RSRC_WORD3);
}

/* m0 a multiple of 8 */
template <int M>
__device__ __forceinline__ void strided_add(const double (&v)[M], int m0, int full,
                                            double (&acc)[8]) {
#pragma unroll
  for (int q = 0; q < M; q++) {
    if (m0 + q < 8 * full)
      acc[q & 7] += v[q];
    else
      acc[0] += v[q];
  }
}

__device__ __forceinline__ double strided_tree(const double (&acc)[8]) {
  return ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
}

/* s + w * v with the product rounded before the sum (no fused multiply-add):
 * the arithmetic of the reference's residual loop.  HIP's __dadd_rn /
 * __dmul_rn are plain operators that the default contraction would fuse. */
__device__ __forceinline__ double add_mul_rn(double s, double w, double v) {
#pragma clang fp contract(off)
  const double prod = w * v;
  return s + prod;
}

/* One pass over the k stored columns of panel P (leading dimension ld) for
 * the JP panel rows idx[q] this thread owns (-1: none):
 *   RES:  s[q] = sum_l bc[l*ld] * P(idx[q], l), in order of l with separately
 *         rounded multiply and add (the reference's residual loop);
 *   dots: out[l] = workgroup sum of sum_q P(idx[q], l) * pv[q]  (l < k).
 * Full chunks of LCH columns are unrolled so all their loads issue first.
 *
 * The dot products are summed without a workgroup barrier per chunk: each
 * wave transposes its threads' products through its own columns of red
 * (a wave runs in lockstep, so its LDS writes are visible to its own later
 * reads) into sixteen part sums of sixteen threads each, kept in wsum; every
 * DOT_WIN columns one barrier pair combines the parts of the four waves in a
 * fixed tree.  Parts and tree are those of a 256-thread transposed
 * reduction, so the sums do not depend on the window. */
template <int JP, bool RES, bool NARROW>
__device__ __forceinline__ void panel_pass(const double* __restrict__ P,
                                           const double* __restrict__ bc, int ld, int k,
                                           const int (&idx)[JP], const double (&pv)[JP],
                                           bool dots, double (&s)[JP], double (*red)[TPB],
                                           double (*wsum)[DOT_WIN], double* out) {
  static_assert(TPB == 256 && LCH == 16, "sixteen parts of sixteen threads per column");
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  // Columns in pieces of H, software-pipelined over two register buffers:
  // once a piece's values are used, the piece two ahead is loaded into its
  // registers, so loads stay in flight during the arithmetic and the
  // dot-product reductions (one memory round trip per pass instead of one
  // per chunk).  H keeps the bytes in flight per thread at 8 JP H = 256
  // whatever the rows per thread: 8 columns for four rows, 16 for one or
  // two (a piece is then a whole reduction chunk).  Columns beyond k (the
  // last piece's tail) are neither loaded nor used.
  constexpr int H = (JP >= 4 ? 8 : 16) / (NARROW ? 2 : 1);
  static_assert(LCH % H == 0, "whole pieces per reduction chunk");
  // (RES) the values bc[l ld] of a piece are loaded by one instruction,
  // lane ll taking column c0 + ll, and broadcast from that lane when used
  double va[H][JP], vb[H][JP], ba = 0.0, bb = 0.0;
  // Loads through buffer resources rebased at each piece's first column:
  // the row offsets stay in JP registers and each column's offset is
  // uniform (a scalar operand), so no 64-bit address is formed or held per
  // load.  Offsets reach H ld rows of 8 bytes: ld < 2^24 (checked on the
  // host).
  int voff[JP];
#pragma unroll
  for (int q = 0; q < JP; q++) voff[q] = idx[q] >= 0 ? idx[q] * 8 : 0;
  auto load_piece = [&](double(&v)[H][JP], double& b, int c0) {
    const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
        (void*)(P + (size_t)c0 * ld), (short)0, 0x7fffffff, BUF_RSRC_WORD3);
    if (RES) b = (lane < H && c0 + lane < k) ? bc[(size_t)(c0 + lane) * ld] : 0.0;
#pragma unroll
    for (int ll = 0; ll < H; ll++) {
      const bool ok = c0 + ll < k;
      const int so = ll * ld * 8;
#pragma unroll
      for (int q = 0; q < JP; q++) v[ll][q] = (ok && idx[q] >= 0) ? buf_load_f64(rp, voff[q], so) : 0.0;
    }
  };
  auto use_piece = [&](const double(&v)[H][JP], double b, int c0) {
    const int off = c0 % LCH;
#pragma unroll
    for (int ll = 0; ll < H; ll++) {
      if (c0 + ll < k) {
        const double bl = RES ? read_lane_f64(b, ll) : 0.0;
        double p = 0.0;
#pragma unroll
        for (int q = 0; q < JP; q++) {
          if (RES) s[q] = add_mul_rn(s[q], bl, v[ll][q]);
          p += v[ll][q] * pv[q];
        }
        if (dots) red[off + ll][t] = p;
      }
    }
  };
  // the dot products of chunk [l0, l0 + LCH), once its pieces are used
  auto reduce_chunk = [&](int l0) {
    const int lc = min(LCH, k - l0);
    __builtin_amdgcn_wave_barrier();
    // part 4 wave + pp of column ll: threads 64 wave + 16 pp + u, u < 16
    const int ll = lane >> 2, pp = lane & 3;
    if (ll < lc) {
      double sum = 0.0;
#pragma unroll
      for (int u = 0; u < 16; u++) sum += red[ll][wave * 64 + pp * 16 + u];
      wsum[wave * 4 + pp][l0 % DOT_WIN + ll] = sum;
    }
    __builtin_amdgcn_wave_barrier();
    if ((l0 + LCH) % DOT_WIN == 0 || l0 + LCH >= k) {
      __syncthreads();
      const int wb = l0 / DOT_WIN * DOT_WIN, we = l0 + lc;
      if (wb + t < we) {
        const int x = t;
        double a[8];
#pragma unroll
        for (int q = 0; q < 8; q++) a[q] = wsum[q][x] + wsum[q + 8][x];
        // the xor tree over parts (offsets 8, 4, 2, 1) as seen by part 0
        out[wb + x] = ((a[0] + a[4]) + (a[2] + a[6])) + ((a[1] + a[5]) + (a[3] + a[7]));
      }
      __syncthreads();
    }
  };
  load_piece(va, ba, 0);
  load_piece(vb, bb, H);
  for (int c0 = 0; c0 < k; c0 += 2 * H) {
    use_piece(va, ba, c0);
    load_piece(va, ba, c0 + 2 * H);
    if (dots && ((c0 + H) % LCH == 0 || c0 + H >= k)) reduce_chunk(c0 / LCH * LCH);
    if (c0 + H < k) {
      use_piece(vb, bb, c0 + H);
      load_piece(vb, bb, c0 + 3 * H);
      if (dots && ((c0 + 2 * H) % LCH == 0 || c0 + 2 * H >= k)) reduce_chunk((c0 + H) / LCH * LCH);
    }
  }
}

__global__ void __launch_bounds__(TPB) aca_init_kernel(const AcaBlock* blocks, AcaState* st,
                                                       uint32_t* bits, int nblocks) {
  int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nblocks) return;
  AcaState s;
  s.k = 0;
  s.cur_row = 0;
  s.status = (blocks[b].m > 0 && blocks[b].n > 0 && !blocks[b].skip) ? 0 : 1;
  s.cur_col = -1;
  s.frob2 = 0.0;
  s.eps = 1.0;
  s.iters = 0;
  s.pending = 0;
  s.fin_chunks = 1;
  s.pad_ = 0;
  st[b] = s;
  bits[blocks[b].bits_row0] |= 1u;  // row 0 is chosen first (lowRank.cpp:159)
}

template <class SRC, int JPT, bool NARROW>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(NARROW ? 3 : 2)))
    aca_row_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   const AcaState* __restrict__ st, const uint32_t* __restrict__ bits,
                   const SRC* __restrict__ srcs, int min_pivot) {
  __shared__ double red[LCH][TPB];
  __shared__ double wsum[16][DOT_WIN];
  __shared__ PivRed pr_sh[TPB / 64];
  const int bidx = wg2blk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int local = blockIdx.x - b.row_wg0;
  const int t = threadIdx.x;
  const int ld = L.ld;
  const int k = s.k;
  const int i = s.cur_row;
  const bool pend = s.pending != 0;
  const double* __restrict__ Wp = L.W[b.level] + b.w_row0;
  double* __restrict__ Vp = L.V[b.level] + b.v_row0;
  const SRC src = srcs[b.inst];
  const uint32_t* colbits = bits + b.bits_col0;
  const int jbase = local * TPB * JPT + t;
  double a[JPT], sacc[JPT], pv[JPT];
  int jj[JPT];
#pragma unroll
  for (int q = 0; q < JPT; q++) {
    jj[q] = jbase + q * TPB;
    const bool ok = jj[q] < b.n;
    a[q] = ok ? entry_at(src, b.i0 + i, b.j0 + jj[q]) : 0.0;
    pv[q] = (ok && pend) ? Vp[(size_t)jj[q] + (size_t)(k - 1) * ld] : 0.0;
    if (!ok) jj[q] = -1;
    sacc[q] = 0.0;
  }
  double* slot = slot_ptr(L, b, local);
  // residual row and, for the pending term k-1, the dots V(:,l) . v_{k-1}
  panel_pass<JPT, true, NARROW>(Vp, Wp + i, ld, k, jj, pv, pend, sacc, red, wsum, slot);
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
#pragma unroll
  for (int q = 0; q < JPT; q++) {
    if (jj[q] >= 0) {
      const double r = a[q] - sacc[q];  // lowRank.cpp:167
      Vp[(size_t)jj[q] + (size_t)k * ld] = r;
      const double ar = fabs(r);
      argmax_merge(p.v, p.i, ar, jj[q]);
      if (!bit_test(colbits, jj[q]) && ar > (double)min_pivot && ar >= HODLR_PIVOT_FLOOR &&
          (p.first < 0 || jj[q] < p.first))
        p.first = jj[q];
    }
  }
  const PivRed r = wg_reduce_piv(p, pr_sh);
  if (t == 0) {
    const int cap = L.cap[b.level];
    slot[cap] = r.v;
    slot[cap + 1] = (double)r.i;
    slot[cap + 2] = (double)r.first;
  }
}

template <class SRC, int IPT, bool NARROW>
__global__ void __launch_bounds__(TPB) __attribute__((amdgpu_waves_per_eu(NARROW ? 3 : 2)))
    aca_col_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   AcaState* __restrict__ st, const uint32_t* __restrict__ bits,
                   const SRC* __restrict__ srcs, int min_pivot) {
  __shared__ double red[LCH][TPB];
  __shared__ double wsum[16][DOT_WIN];
  __shared__ PivRed pr_sh[TPB / 64];
  const int bidx = wg2blk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int local = blockIdx.x - b.col_wg0;
  const int t = threadIdx.x;
  const int ld = L.ld;
  const int k = s.k;
  const bool pend = s.pending != 0;
  const int j = decide_column(L, b, bits + b.bits_col0, min_pivot, pr_sh);
  if (local == 0 && t == 0) {
    st[bidx].cur_col = j;
    // the finalize's chunks of dot products l < k - 1, fixed here so that
    // its workgroups agree whether or not its last one has updated k
    st[bidx].fin_chunks = max(1, (k - 1 + FIN_L - 1) / FIN_L);
  }
  double* __restrict__ Wp = L.W[b.level] + b.w_row0;
  const double* __restrict__ Vp = L.V[b.level] + b.v_row0;
  const int ibase = local * TPB * IPT + t;
  double a[IPT], sacc[IPT], pw[IPT];
  int ii[IPT];
  double* slot = slot_ptr(L, b, b.row_nwg + local);
  if (j < 0) {
    // no column this iteration; the pending term still needs W(:,l) . w_{k-1}
    if (!pend) return;
#pragma unroll
    for (int q = 0; q < IPT; q++) {
      ii[q] = ibase + q * TPB;
      const bool ok = ii[q] < b.m;
      pw[q] = ok ? Wp[(size_t)ii[q] + (size_t)(k - 1) * ld] : 0.0;
      if (!ok) ii[q] = -1;
      sacc[q] = 0.0;
    }
    panel_pass<IPT, false, NARROW>(Wp, nullptr, ld, k, ii, pw, true, sacc, red, wsum, slot);
    return;
  }
  const SRC src = srcs[b.inst];
  const uint32_t* rowbits = bits + b.bits_row0;
  const double pivot = 1.0 / Vp[(size_t)j + (size_t)k * ld];  // lowRank.cpp:208
#pragma unroll
  for (int q = 0; q < IPT; q++) {
    ii[q] = ibase + q * TPB;
    const bool ok = ii[q] < b.m;
    a[q] = ok ? entry_at(src, b.i0 + ii[q], b.j0 + j) : 0.0;
    pw[q] = (ok && pend) ? Wp[(size_t)ii[q] + (size_t)(k - 1) * ld] : 0.0;
    if (!ok) ii[q] = -1;
    sacc[q] = 0.0;
  }
  panel_pass<IPT, true, NARROW>(Wp, Vp + j, ld, k, ii, pw, pend, sacc, red, wsum, slot);
  PivRed p;
  p.v = -1.0;
  p.i = -1;
  p.first = -1;
#pragma unroll
  for (int q = 0; q < IPT; q++) {
    if (ii[q] >= 0) {
      const double c = a[q] - sacc[q];  // lowRank.cpp:215
      Wp[(size_t)ii[q] + (size_t)k * ld] = pivot * c;  // :230
      const double ac = fabs(c);
      argmax_merge(p.v, p.i, ac, ii[q]);
      if (!bit_test(rowbits, ii[q]) && ac > (double)min_pivot && ac >= HODLR_PIVOT_FLOOR &&
          (p.first < 0 || ii[q] < p.first))
        p.first = ii[q];
    }
  }
  const PivRed r = wg_reduce_piv(p, pr_sh);
  if (t == 0) {
    const int cap = L.cap[b.level];
    slot[cap] = r.v;
    slot[cap + 1] = (double)r.i;
    slot[cap + 2] = (double)r.first;
  }
}

__global__ void __launch_bounds__(TPB)
    aca_fin_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, const int* __restrict__ wg2blk,
                   const int* __restrict__ wg2chunk, const int* __restrict__ nchunk,
                   AcaState* __restrict__ st, uint32_t* __restrict__ bits,
                   double* __restrict__ fin_part, int* __restrict__ counters, double tol) {
  __shared__ double red[TPB / 64];
  __shared__ double part[4][FIN_L];
  __shared__ PivRed pr_sh[TPB / 64];
  __shared__ int last;
  __shared__ int first_row;
  const int bidx = wg2blk[blockIdx.x];
  const int chunk = wg2chunk[blockIdx.x];
  const AcaBlock b = blocks[bidx];
  const AcaState s = st[bidx];
  if (s.status != 0) return;
  const int t = threadIdx.x;
  const int k = s.k;
  const int j = s.cur_col;
  const bool pend = s.pending != 0;
  // chunks that hold dot products this iteration (l < k - 1, fixed by the
  // column phase); the others leave at once, so early iterations of blocks
  // with large capacity do not pay for their empty chunks
  const int nact = min(nchunk[bidx], s.fin_chunks);
  // A block with few partial slots packs four chunks into one workgroup, a
  // wave per chunk (each thread sums its value's slots itself): up to rank
  // 257 the block then needs one workgroup and no device-scope fence.
  const bool packed = aca_fin_packed(b);
  const int nwg = packed ? (nact + 3) / 4 : nact;
  if (chunk >= nwg) return;

  // cross term of the pending term k-1 (lowRank.cpp:234-237): sum_{l<k-1} e_l d_l.
  // FIN_L values of l per chunk, each summed over the slots in 4 groups
  // (slots w = g mod 4), the groups' sums added in order.
  double cross = 0.0;
  if (packed) {
    // wave w: chunk 4 chunk + w, with the arithmetic of a one-chunk
    // workgroup of the unpacked layout (whose other waves hold zeros)
    const int lane = t & 63, w = t >> 6, lc = 4 * chunk + w;
    double v = 0.0;
    if (pend && lc < nact) {
      const int l = lc * FIN_L + lane;
      const bool on = l < k - 1;
      const int stride_b = (L.cap[b.level] + 4) * (int)sizeof(double);
      const int voff = on ? l * (int)sizeof(double) : 0;  // lanes past k - 1: discarded
      const __amdgpu_buffer_rsrc_t rr = slot_rsrc(L, b, 0), rc = slot_rsrc(L, b, b.row_nwg);
      const double zero = 0.0 * (double)b.m;  // +0.0 (m >= 0)
      // at most 32 slots a phase: eight values per slot group, the 32 loads
      // of a phase issued at once; then ((S0 + S1) + S2) + S3 over the four
      // groups, as the unpacked layout adds its groups' sums
      auto phase_sum = [&](__amdgpu_buffer_rsrc_t rs, int n) {
        double v4[4][8];
#pragma unroll
        for (int g = 0; g < 4; g++) strided_load<8>(rs, voff, g, n, stride_b, 0, zero, v4[g]);
        double r = 0.0;
#pragma unroll
        for (int g = 0; g < 4; g++) {
          double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
          strided_add<8>(v4[g], 0, strided_full_groups(n, g), acc);
          r = g == 0 ? strided_tree(acc) : r + strided_tree(acc);
        }
        // the sum is complete here: otherwise the compiler moves its
        // additions past the next phase's loads, holding both phases' 64
        // values at once (registers for three waves per SIMD instead of seven)
        asm volatile("" : : "v"(r));
        return r;
      };
      const double d = phase_sum(rr, b.row_nwg);
      const double e = phase_sum(rc, b.col_nwg);
      if (on) v = d * e;
    }
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    // (the unpacked workgroup's final sum over its four waves: this wave's
    // sum, then the three zero sums of the others)
    const double z = 0.0;
    v = (((0.0 + v) + z) + z) + z;
    if (lane == 0) part[0][w] = v;
    __syncthreads();
    if (nact > 4) {
      if (t == 0) {
        for (int q = 0; q < 4 && 4 * chunk + q < nact; q++)
          fin_part[(size_t)bidx * MAX_FIN_CHUNKS + 4 * chunk + q] = part[0][q];
        __threadfence();
        const int old = atomicAdd(&counters[bidx], 1);
        last = (old == nwg - 1);
      }
      __syncthreads();
      if (!last) return;
      __threadfence();
    }
  } else {
    if (pend) {
      const int ll = t & (FIN_L - 1), g = t / FIN_L;
      const int l = chunk * FIN_L + ll;
      const bool on = l < k - 1;
      const int stride_b = (L.cap[b.level] + 4) * (int)sizeof(double);
      const int voff = on ? l * (int)sizeof(double) : 0;  // lanes past k - 1: discarded
      const __amdgpu_buffer_rsrc_t rr = slot_rsrc(L, b, 0), rc = slot_rsrc(L, b, b.row_nwg);
      const int gu = __builtin_amdgcn_readfirstlane(g);  // g is one wave's
      const double zero = 0.0 * (double)b.m;  // +0.0 (m >= 0)
      // slot group g of both phases, 16 values of each per round (up to
      // 64 slots a phase in one round)
      double ad[8] = {0, 0, 0, 0, 0, 0, 0, 0}, ae[8] = {0, 0, 0, 0, 0, 0, 0, 0};
      const int fd = strided_full_groups(b.row_nwg, gu), fe = strided_full_groups(b.col_nwg, gu);
      const int nm = (max(b.row_nwg, b.col_nwg) - gu + 3) / 4;
#pragma unroll 1
      for (int m0 = 0; m0 < nm; m0 += 16) {
        double vd[16], ve[16];
        strided_load<16>(rr, voff, gu, b.row_nwg, stride_b, m0, zero, vd);
        strided_load<16>(rc, voff, gu, b.col_nwg, stride_b, m0, zero, ve);
        strided_add<16>(vd, m0, fd, ad);
        strided_add<16>(ve, m0, fe, ae);
      }
      double d = 0.0, e = 0.0;
      if (on) {
        d = strided_tree(ad);
        e = strided_tree(ae);
      }
      part[g][ll] = d;
      __syncthreads();
      if (g == 0) d = part[0][ll] + part[1][ll] + part[2][ll] + part[3][ll];
      __syncthreads();
      part[g][ll] = e;
      __syncthreads();
      if (g == 0) {
        e = part[0][ll] + part[1][ll] + part[2][ll] + part[3][ll];
        cross = (l < k - 1) ? d * e : 0.0;
      }
    }
    cross = wg_sum(cross, red);
    // With one chunk this workgroup is the block's only one: no partials to
    // publish, no device-scope fences (each writes back this XCD's L2).
    if (nact > 1) {
      if (t == 0) {
        fin_part[(size_t)bidx * MAX_FIN_CHUNKS + chunk] = cross;
        __threadfence();
        const int old = atomicAdd(&counters[bidx], 1);
        last = (old == nact - 1);
      }
      __syncthreads();
      if (!last) return;
      // the last workgroup of the block: all its threads reduce, thread 0 decides
      __threadfence();
    }
  }
  const int cap = L.cap[b.level];
  auto cross_at = [&](int c) {
    return packed ? (nact <= 4 ? part[0][c] : fin_part[(size_t)bidx * MAX_FIN_CHUNKS + c])
                  : (nact == 1 ? cross : fin_part[(size_t)bidx * MAX_FIN_CHUNKS + c]);
  };
  // The first term of each sum below and this thread's column-phase pivot
  // slot are loaded together, one memory latency for all; the loops take
  // the terms past TPB (same order: the first term, then the loop's).
  // Absent terms are a +0.0 the compiler cannot fold (see strided_load).
  const double zero = 0.0 * (double)b.m;
  double x0 = zero, y0 = zero, z0 = zero;
  if (pend) {
    if (t < nact) x0 = cross_at(t);
    if (t < b.row_nwg) y0 = slot_ptr(L, b, t)[k - 1];
    if (t < b.col_nwg) z0 = slot_ptr(L, b, b.row_nwg + t)[k - 1];
  }
  PivRed p0;
  p0.v = -1.0;
  p0.i = -1;
  p0.first = -1;
  if (j >= 0 && t < b.col_nwg) read_piv(slot_ptr(L, b, b.row_nwg + t), cap, p0.v, p0.i, p0.first);
  double crossSum = 0.0, nv2 = 0.0, nw2 = 0.0;
  if (pend) {
    crossSum += x0;
    for (int c = t + TPB; c < nact; c += TPB) crossSum += cross_at(c);
    nv2 += y0;
    for (int w = t + TPB; w < b.row_nwg; w += TPB) nv2 += slot_ptr(L, b, w)[k - 1];
    nw2 += z0;
    for (int w = t + TPB; w < b.col_nwg; w += TPB) nw2 += slot_ptr(L, b, b.row_nwg + w)[k - 1];
    crossSum = wg_sum(crossSum, red);
    nv2 = wg_sum(nv2, red);
    nw2 = wg_sum(nw2, red);
  }
  uint32_t* rowbits = bits + b.bits_row0;
  uint32_t* colbits = bits + b.bits_col0;
  PivRed cr;
  cr.v = -1.0;
  cr.i = -1;
  cr.first = -1;
  if (j >= 0) cr = reduce_slots(L, b, b.row_nwg, b.col_nwg, pr_sh, p0);
  if (j == -1) {  // first unchosen row for a jump
    if (t == 0) first_row = 0x7fffffff;
    __syncthreads();
    const int nw = (b.m + 31) >> 5;
    int mine = -1;
    for (int w = t; w < nw && mine < 0; w += TPB) {
      uint32_t x = ~rowbits[w];
      if (w == nw - 1 && (b.m & 31)) x &= (1u << (b.m & 31)) - 1u;
      if (x) mine = w * 32 + __ffs(x) - 1;
    }
    if (mine >= 0) atomicMin(&first_row, mine);
    __syncthreads();
  }
  if (t != 0) return;
  counters[bidx] = 0;
  AcaState ns = s;
  ns.iters = s.iters + 1;
  ns.pending = 0;
  if (pend) {
    // complete epsilon of term k-1 (:238-241) and the loop test (:149)
    ns.frob2 = s.frob2 + 2.0 * crossSum + nw2 * nv2;
    ns.eps = sqrt(nw2) * sqrt(nv2) / sqrt(ns.frob2);
    if (!((ns.eps > tol) || (k < b.min_rank))) {
      ns.status = 1;  // term k-1 was the last; this iteration is discarded
      st[bidx] = ns;
      return;
    }
  }
  if (j == -1) {
    // max |r| <= minPivot: continue from the first unchosen row (:187-195)
    const int nr = first_row;
    if (nr == 0x7fffffff) {
      ns.status = 1;
    } else {
      ns.cur_row = nr;
      bit_set(rowbits, nr);
    }
    st[bidx] = ns;
    return;
  }
  if (j == -2) {  // no admissible column (:200-202)
    ns.status = 1;
    st[bidx] = ns;
    return;
  }
  bit_set(colbits, j);  // :206
  // next row (:218-227)
  const int next = (cr.i >= 0 && !bit_test(rowbits, cr.i)) ? cr.i : cr.first;
  if (next < 0) {  // term k is discarded (:226-227)
    ns.status = 1;
    st[bidx] = ns;
    return;
  }
  // term k is stored (:230-231); its epsilon completes next iteration
  ns.k = k + 1;
  ns.cur_row = next;
  bit_set(rowbits, next);
  const int upper = min(b.m, b.n);
  if (ns.k == b.max_rank || ns.k >= upper) {
    ns.status = 1;  // the loop ends whatever epsilon is (:246-247, :149)
  } else {
    ns.pending = 1;
    if (ns.k >= cap) ns.status = 2;  // storage full: the host grows the panels
  }
  st[bidx] = ns;
}

/* Post-processing of the returned factors (lowRank.cpp:250-274):
 * rank 0 -> one zero column; rank >= min(m,n)-1 -> the dense block with an
 * identity factor; otherwise the speculative column k is cleared.  One
 * workgroup per block. */
template <class SRC>
__global__ void __launch_bounds__(TPB)
    aca_post_kernel(AcaLevels L, const AcaBlock* __restrict__ blocks, AcaState* __restrict__ st,
                    const SRC* __restrict__ srcs, int* __restrict__ ranks) {
  const int bidx = blockIdx.x;
  const AcaBlock b = blocks[bidx];
  if (b.skip == 2) {  // compressed by another method
    if (threadIdx.x == 0) ranks[bidx] = -1;
    return;
  }
  const int k = st[bidx].k;  // (0 for max_rank 0: one zero column below)
  const int ld = L.ld;
  const int cap = L.cap[b.level];
  double* Wp = L.W[b.level] + b.w_row0;
  double* Vp = L.V[b.level] + b.v_row0;
  const int upper = min(b.m, b.n);
  if (k == 0) {
    for (int i = threadIdx.x; i < b.m; i += TPB) Wp[i] = 0.0;
    for (int jj = threadIdx.x; jj < b.n; jj += TPB) Vp[jj] = 0.0;
    if (threadIdx.x == 0) ranks[bidx] = 1;
    return;
  }
  if (k >= upper - 1) {
    const SRC src = srcs[b.inst];
    if (b.n <= b.m) {  // W = block (m x n), V = I_n
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.n; e += TPB) {
        const int ii = (int)(e % b.m), jj = (int)(e / b.m);
        Wp[(size_t)ii + (size_t)jj * ld] = entry_at(src, b.i0 + ii, b.j0 + jj);
      }
      for (size_t e = threadIdx.x; e < (size_t)b.n * b.n; e += TPB) {
        const int r = (int)(e % b.n), c = (int)(e / b.n);
        Vp[(size_t)r + (size_t)c * ld] = (r == c) ? 1.0 : 0.0;
      }
      if (threadIdx.x == 0) ranks[bidx] = b.n;
    } else {  // W = I_m, V = block^T (n x m)
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.m; e += TPB) {
        const int r = (int)(e % b.m), c = (int)(e / b.m);
        Wp[(size_t)r + (size_t)c * ld] = (r == c) ? 1.0 : 0.0;
      }
      for (size_t e = threadIdx.x; e < (size_t)b.m * b.n; e += TPB) {
        const int ii = (int)(e % b.m), jj = (int)(e / b.m);
        Vp[(size_t)jj + (size_t)ii * ld] = entry_at(src, b.i0 + ii, b.j0 + jj);
      }
      if (threadIdx.x == 0) ranks[bidx] = b.m;
    }
    return;
  }
  if (k < cap) {  // column k may hold a discarded iteration's values
    for (int i = threadIdx.x; i < b.m; i += TPB) Wp[(size_t)i + (size_t)k * ld] = 0.0;
    for (int jj = threadIdx.x; jj < b.n; jj += TPB) Vp[(size_t)jj + (size_t)k * ld] = 0.0;
  }
  if (threadIdx.x == 0) ranks[bidx] = k;
}

}  // namespace

void aca_launch_init(const AcaBlock* d_blocks, AcaState* d_state, uint32_t* d_bits, int nblocks,
                     hipStream_t s) {
  aca_init_kernel<<<(nblocks + TPB - 1) / TPB, TPB, 0, s>>>(d_blocks, d_state, d_bits, nblocks);
  HODLR_HIP_CHECK(hipGetLastError());
}

// Launches of more than ACA_NARROW_WGS workgroups (several per CU) take the
// variant with half-width pieces: fewer registers, three workgroups per CU
// instead of two, which hides the latency of the many short workgroups of
// the early iterations; smaller launches keep the wider pieces, whose loads
// in flight per wave a lone workgroup per CU needs.  Same arithmetic.
constexpr int ACA_NARROW_WGS = 768;

template <int JPT>
static void launch_row(const AcaLevels& L, const AcaLaunch& a, const EntrySource* d_srcs,
                       hipStream_t s) {
  if (a.n_row_wg > ACA_NARROW_WGS)
    aca_row_kernel<EntrySource, JPT, true><<<a.n_row_wg, TPB, 0, s>>>(
        L, a.blocks, a.wg2blk_row, a.state, a.bits, d_srcs, a.min_pivot);
  else
    aca_row_kernel<EntrySource, JPT, false><<<a.n_row_wg, TPB, 0, s>>>(
        L, a.blocks, a.wg2blk_row, a.state, a.bits, d_srcs, a.min_pivot);
}

template <int IPT>
static void launch_col(const AcaLevels& L, const AcaLaunch& a, const EntrySource* d_srcs,
                       hipStream_t s) {
  if (a.n_col_wg > ACA_NARROW_WGS)
    aca_col_kernel<EntrySource, IPT, true><<<a.n_col_wg, TPB, 0, s>>>(
        L, a.blocks, a.wg2blk_col, a.state, a.bits, d_srcs, a.min_pivot);
  else
    aca_col_kernel<EntrySource, IPT, false><<<a.n_col_wg, TPB, 0, s>>>(
        L, a.blocks, a.wg2blk_col, a.state, a.bits, d_srcs, a.min_pivot);
}

void aca_launch_iteration(const AcaLevels& L, const AcaLaunch& a, const EntrySource* d_srcs,
                          hipStream_t s, hipEvent_t* ev) {
  if (ev) HODLR_HIP_CHECK(hipEventRecord(ev[0], s));
  switch (a.jpt) {
    case 1: launch_row<1>(L, a, d_srcs, s); break;
    case 2: launch_row<2>(L, a, d_srcs, s); break;
    default: launch_row<4>(L, a, d_srcs, s);
  }
  HODLR_HIP_CHECK(hipGetLastError());
  if (ev) HODLR_HIP_CHECK(hipEventRecord(ev[1], s));
  switch (a.ipt) {
    case 1: launch_col<1>(L, a, d_srcs, s); break;
    case 2: launch_col<2>(L, a, d_srcs, s); break;
    default: launch_col<4>(L, a, d_srcs, s);
  }
  HODLR_HIP_CHECK(hipGetLastError());
  if (ev) HODLR_HIP_CHECK(hipEventRecord(ev[2], s));
  aca_fin_kernel<<<a.n_fin_wg, TPB, 0, s>>>(L, a.blocks, a.wg2blk_fin, a.wg2chunk_fin, a.nchunk,
                                             a.state, a.bits, a.fin_part, a.counters, a.tol);
  HODLR_HIP_CHECK(hipGetLastError());
}

void aca_launch_post(const AcaLevels& L, const AcaBlock* d_blocks, AcaState* d_state,
                     const EntrySource* d_srcs, int* d_ranks, int nblocks, hipStream_t s) {
  aca_post_kernel<EntrySource><<<nblocks, TPB, 0, s>>>(L, d_blocks, d_state, d_srcs, d_ranks);
  HODLR_HIP_CHECK(hipGetLastError());
}

}  // namespace hodlr
