// Second-order energy over a sparse first-order space, on the GPU.
//
// The dense-target PT2 of fcg_pt2.hip forms sigma on every configuration of
// each target Fock block; at many clusters those blocks are astronomically
// large (32^12 configurations at 12 clusters of M=32), and only the
// configurations that some screened contribution reaches matter.  This
// engine follows the reference's own formulation instead
// (compute_pt2_energy, src/tpsci_pt2_energy.jl:10-175, with _pt2_job
// :178-266 and contract_matvec_thread, src/tpsci_matvec_thread.jl:279-818):
// contributions are generated per (ket configuration J, term t) as every
// bra state combination I_A of the term's active clusters,
//
//     x = (I_A, J on the other clusters),   H_t[x, J] = s_t sum_p V[p] prod_a Γ_a[p_a, I_a, J_a],
//
// kept when |H_t[x, J]| > thresh / max_r |c[J, r]| (_collect_significant_thread!
// :830-895), and summed per x before sigma_x^2 / (E0 - Hd_x) is summed over
// the x outside the variational space (project_out!, _sum_pt2 :269-282).
//
// Work and pruning, all bounds being upper bounds of |H_t[x, J]| so that
// nothing that passes the element test is skipped (the reference's
// upper_bound_thread / upper_bound2_thread, :897-1283):
//   items   (ket Fock block F, term t) pairs whose target F + Δ_t has every
//           cluster sector in ClusterOps (the job filter, :49-79) and whose
//           Γ blocks exist (missing blocks contribute nothing, as in
//           contract_matvec_thread); kept when
//           B_t = sum_p |V[p]| prod_a max|Γ_a[p_a, :, :]|  satisfies
//           B_t * max_{J in F} |c_J| > thresh;
//   kets    each Fock block's kets are ordered by max_r |c| descending, so
//           the kets with B_t |c_J| > thresh are a prefix; each of them is
//           kept when its own bound
//           B_t(J) = sum_p |V[p]| prod_a max_I |Γ_a[p_a, I, J_a]|
//           satisfies B_t(J) |c_J| > thresh.  A light screening kernel
//           counts these per item at high occupancy; only items with a kept
//           ket reach the contraction kernel;
//   indices kept kets with equal states on the active clusters share every
//           element (a group); bra states I_a whose per-index bound cannot
//           pass for the group's largest |c| are dropped before the
//           contraction (upper_bound2_thread's survivor lists).
// The kept contraction is the reference's chain (contract the last active
// cluster first), tiled over the last cluster's surviving states so the
// intermediates stay in LDS.  The coarse stages (items, screening) keep a
// wider rounding margin than the per-ket and per-index tests, so a stage
// never drops what a later one would keep.
//
// Accumulation is deterministic and streams in memory: the items with a kept
// ket are ordered by their target Fock block (a stable sort), a count pass
// gives every item's number of kept contributions, and the items are cut
// into chunks at target-block boundaries, each holding at most a fixed
// number of contributions (the cut depends on the counts only).  Per chunk a
// fill pass writes every kept contribution (key = target Fock sectors |
// states of x, value = H_t[x,J] c[J,:]) at positions fixed by prefix sums, a
// stable radix sort on the 128-bit keys keeps equal keys in that order, and
// each x's contributions are summed in it.  Every x lies in one target block,
// so a chunk holds all of its contributions, and e2 does not depend on
// scheduling or on the memory free at the time of the call.
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "fcg_internal.h"

namespace {

constexpr int PT = 256;          // threads per contraction workgroup
constexpr int SW = 64;           // threads per screening workgroup (one wavefront)
constexpr int KCH = PT;          // candidate kets per batch of the contraction kernel
constexpr int ARENA = 2048;      // chain intermediates (doubles)
constexpr int VMAX = 1024;       // staged integrals (doubles)
constexpr int SMAX = 64;         // bra states per cluster the survivor lists hold
constexpr int NPMAX = 64;        // leading Γ dimension the bound weights hold
constexpr double BOUND_SLACK = 1.0 + 1e-10;   // bounds are inflated against rounding
constexpr double SCREEN_SLACK = 1.0 + 1e-8;   // ... and the coarse stages by more

struct PTerm {
  int32_t nbody;
  int32_t cl[4];
  int32_t op[4];
  int32_t dims[4];
  int32_t odd;    // bit q: operator string q has odd fermion parity
  int32_t group;  // TransferConfig group
  int64_t ioff;
};

struct Key128 {
  uint64_t hi, lo;  // hi: target Fock sector codes, lo: states
};

struct KeyDecomposer {
  __host__ __device__ rocprim::tuple<uint64_t &, uint64_t &> operator()(Key128 &k) const {
    return rocprim::tuple<uint64_t &, uint64_t &>(k.hi, k.lo);
  }
};

// Sort key of a contribution: K = tloc * 2^sN + states, tloc the index of
// its target Fock block among the chunk's targets, sN = sbits * N in [1, 64].
__host__ __device__ __forceinline__ Key128 chunk_key(uint64_t tloc, uint64_t states, int sN) {
  return Key128{sN == 64 ? tloc : tloc >> (64 - sN), sN == 64 ? states : states | (tloc << sN)};
}
__host__ __device__ __forceinline__ uint64_t chunk_key_states(const Key128 &k, int sN) {
  return sN == 64 ? k.lo : k.lo & ((uint64_t(1) << sN) - 1);
}
__host__ __device__ __forceinline__ uint64_t chunk_key_target(const Key128 &k, int sN) {
  return sN == 64 ? k.hi : (k.hi << (64 - sN)) | (k.lo >> sN);
}

struct Pt2Args {
  int N, S, R, nbody, fbits, sbits;
  double thresh;
  // ket space
  int64_t n_fock, n_cfg;
  const int64_t *fstart;  // [n_fock+1]
  const int16_t *fock;    // [n_fock][N][2]
  const uint8_t *cst;     // [n_cfg][N] 0-based states
  const int32_t *kord;    // [n_cfg] each Fock block's kets by max_r |c| descending
  const double *csort;    // [n_cfg] their max_r |c|
  const double *thr;      // [n_cfg] thresh / max_r |c| (the reference's thresh_curr), by ket
  const double *cmaxF;    // [n_fock]
  const double *c;        // [R][n_cfg]
  // operator
  const PTerm *terms;
  const int32_t *tc_start;  // [n_tc+1]
  const int16_t *tc;        // [n_tc][N][2]
  int n_tc;
  const double *ints;
  const int32_t *gblock;  // [N][n_ops][S*S][S*S]
  int n_ops;
  const int64_t *gb_off;
  const int32_t *gb_dims;
  const double *gdata;
  const int32_t *nstates;  // [N][S*S]
  const int64_t *bmax_off;  // per Γ block: np values of max_{I,J} |Γ[p,I,J]|
  const double *bmax;
  const int64_t *cmx_off;   // per Γ block: np x m_ket values of max_I |Γ[p,I,J]|
  const double *cmx;
  // items
  int64_t *item_count;      // count pass: items per ket Fock block
  const int64_t *item_ptr;  // fill pass
  int32_t *item_f, *item_t;
  double *item_b;           // fill pass: the item bound B_t (without |c|)
  uint64_t *item_key;       // fill pass: target Fock sector codes
  int64_t fb_base;
  // screening and contraction
  const int32_t *items_f, *items_t;
  const double *items_b;
  int64_t item_base, n_items;
  int32_t *nsurv;             // screening: kept kets per item
  const uint32_t *slot_item;  // contraction slot -> item (items ordered by target)
  const uint32_t *slot_tloc;  // contraction slot -> its target's index in its chunk
  int sN;                     // state bits of a key (sbits * N)
  int64_t slot_base, n_slots;
  int64_t *emit_count;        // count pass: kept contributions per slot
  const int64_t *emit_ptr;    // fill pass: first position of each slot
  int64_t emit_base;          // ... less this (the chunk's first position)
  Key128 *ekey;
  double *eval;               // [E][R]
  unsigned long long *stats;  // [0] groups kept, [1] chain flops, [2] overflow
};

__device__ __forceinline__ int sector_code(int na, int nb, int S) { return na * S + nb; }

__device__ __forceinline__ int32_t gblk(const Pt2Args &a, int c, int op, int cb, int ck) {
  const int SS = a.S * a.S;
  return a.gblock[(((int64_t)c * a.n_ops + op) * SS + cb) * SS + ck];
}

// item bound B_t = sum_p |V[p]| prod_q bmax_q[p_q] for a term whose Γ blocks exist
__device__ double item_bound(const Pt2Args &a, const PTerm &T, const int32_t *blk) {
  if (T.nbody == 1) return a.bmax[a.bmax_off[blk[0]]];
  const double *m0 = a.bmax + a.bmax_off[blk[0]];
  const double *m1 = a.bmax + a.bmax_off[blk[1]];
  const double *m2 = T.nbody > 2 ? a.bmax + a.bmax_off[blk[2]] : nullptr;
  const double *m3 = T.nbody > 3 ? a.bmax + a.bmax_off[blk[3]] : nullptr;
  const int d0 = T.dims[0], d1 = T.dims[1];
  const int d2 = T.nbody > 2 ? T.dims[2] : 1, d3 = T.nbody > 3 ? T.dims[3] : 1;
  const double *V = a.ints + T.ioff;
  double B = 0.0;
  for (int p3 = 0; p3 < d3; p3++)
    for (int p2 = 0; p2 < d2; p2++) {
      const double f23 = (m2 ? m2[p2] : 1.0) * (m3 ? m3[p3] : 1.0);
      if (f23 == 0.0) continue;
      for (int p1 = 0; p1 < d1; p1++) {
        const double f = f23 * m1[p1];
        if (f == 0.0) continue;
        const double *Vp = V + d0 * (p1 + d1 * (p2 + d2 * p3));
        double s = 0.0;
        for (int p0 = 0; p0 < d0; p0++) s += fabs(Vp[p0]) * m0[p0];
        B += s * f;
      }
    }
  return B;
}

// ---- items: (ket Fock block, term) pairs that can contribute ----
template <bool FILL>
__global__ __launch_bounds__(PT) void k_pt2_items(Pt2Args a) {
  __shared__ int16_t sF[2 * FCG_MAXN];
  __shared__ int32_t sScan[PT];
  __shared__ int64_t sBase;
  const int tid = threadIdx.x;
  const int64_t fj = a.fb_base + blockIdx.x;
  const int N = a.N, S = a.S;
  if (tid < 2 * N) sF[tid] = a.fock[fj * 2 * N + tid];
  if (tid == 0) sBase = 0;
  __syncthreads();
  const double cm = a.cmaxF[fj];
  int64_t total = 0;
  for (int g0 = 0; g0 < a.n_tc; g0 += PT) {
    const int g = g0 + tid;
    int cnt = 0;
    int32_t tlist[8];
    double blist[8];
    int nl = 0;
    bool more = false;
    uint64_t tkey = 0;
    if (g < a.n_tc && cm > 0.0) {
      const int16_t *d = a.tc + (int64_t)g * 2 * N;
      bool ok = true;
      for (int c = 0; c < N && ok; c++) {
        const int na = sF[2 * c] + d[2 * c], nb = sF[2 * c + 1] + d[2 * c + 1];
        ok = na >= 0 && nb >= 0 && na < S && nb < S &&
             a.nstates[(int64_t)c * S * S + sector_code(na, nb, S)] > 0;
        if (ok) tkey |= (uint64_t)sector_code(na, nb, S) << (a.fbits * c);
      }
      if (ok) {
        for (int t = a.tc_start[g]; t < a.tc_start[g + 1]; t++) {
          const PTerm T = a.terms[t];
          if (T.nbody > a.nbody) continue;
          int32_t blk[4];
          bool have = true;
          for (int q = 0; q < T.nbody; q++) {
            const int c = T.cl[q];
            const int ck = sector_code(sF[2 * c], sF[2 * c + 1], S);
            const int cb = sector_code(sF[2 * c] + d[2 * c], sF[2 * c + 1] + d[2 * c + 1], S);
            blk[q] = gblk(a, c, T.op[q], cb, ck);
            have = have && blk[q] >= 0;
          }
          if (!have) continue;
          const double B = item_bound(a, T, blk);
          if (!(B * SCREEN_SLACK * cm > a.thresh) || B == 0.0) continue;
          if (FILL) {
            if (nl < 8) {
              tlist[nl] = t;
              blist[nl++] = B;
            } else {
              more = true;
            }
          }
          cnt++;
        }
      }
    }
    if (!FILL) {
      total += cnt;
    } else {
      // ordered compaction: block-wide exclusive scan of the counts
      sScan[tid] = cnt;
      __syncthreads();
      for (int off = 1; off < PT; off <<= 1) {
        const int v = tid >= off ? sScan[tid - off] : 0;
        __syncthreads();
        sScan[tid] += v;
        __syncthreads();
      }
      int64_t at = a.item_ptr[fj] + sBase + sScan[tid] - cnt;
      if (cnt > 0) {
        if (!more) {
          for (int i = 0; i < nl; i++) {
            a.item_f[at + i] = (int32_t)fj;
            a.item_t[at + i] = tlist[i];
            a.item_b[at + i] = blist[i];
            a.item_key[at + i] = tkey;
          }
        } else {
          // more kept terms than the register list holds: recompute in order
          const int16_t *d = a.tc + (int64_t)g * 2 * N;
          for (int t = a.tc_start[g]; t < a.tc_start[g + 1]; t++) {
            const PTerm T = a.terms[t];
            if (T.nbody > a.nbody) continue;
            int32_t blk[4];
            bool have = true;
            for (int q = 0; q < T.nbody; q++) {
              const int c = T.cl[q];
              const int ck = sector_code(sF[2 * c], sF[2 * c + 1], S);
              const int cb = sector_code(sF[2 * c] + d[2 * c], sF[2 * c + 1] + d[2 * c + 1], S);
              blk[q] = gblk(a, c, T.op[q], cb, ck);
              have = have && blk[q] >= 0;
            }
            if (!have) continue;
            const double B = item_bound(a, T, blk);
            if (!(B * SCREEN_SLACK * cm > a.thresh) || B == 0.0) continue;
            a.item_f[at] = (int32_t)fj;
            a.item_t[at] = t;
            a.item_b[at] = B;
            a.item_key[at] = tkey;
            at++;
          }
        }
      }
      __syncthreads();
      if (tid == PT - 1) sBase += sScan[PT - 1];
      __syncthreads();
    }
  }
  if (!FILL) {
    sScan[tid] = (int32_t)total;
    __syncthreads();
    if (tid == 0) {
      int64_t s = 0;
      for (int k = 0; k < PT; k++) s += sScan[k];
      a.item_count[fj] = s;
    }
  }
}

// What the per-ket bound of one item needs: the term's integrals and, per
// active cluster, the column maxima of its Γ block.
struct KetBound {
  int nb, cl[4], np[4], d0, d1, d2, d3;
  const double *V;
  const double *cmx[4];  // cmx of the Γ block; column J_q at np[q] * J_q
};

__device__ __forceinline__ void ket_bound_setup(const Pt2Args &a, const PTerm &T, int fj,
                                                KetBound &k) {
  const int N = a.N, S = a.S;
  const int16_t *F = a.fock + (int64_t)fj * 2 * N;
  const int16_t *D = a.tc + (int64_t)T.group * 2 * N;
  k.nb = T.nbody;
#pragma unroll
  for (int q = 0; q < 4; q++) {  // constant trip count: the arrays stay in registers
    k.cl[q] = 0;
    k.np[q] = 1;
    k.cmx[q] = nullptr;
    if (q < T.nbody) {
      const int c = T.cl[q];
      const int ck = sector_code(F[2 * c], F[2 * c + 1], S);
      const int cb = sector_code(F[2 * c] + D[2 * c], F[2 * c + 1] + D[2 * c + 1], S);
      const int blk = gblk(a, c, T.op[q], cb, ck);
      k.cl[q] = c;
      k.np[q] = a.gb_dims[3 * blk];
      k.cmx[q] = a.cmx + a.cmx_off[blk];
    }
  }
  k.d0 = T.dims[0];
  k.d1 = T.nbody > 1 ? T.dims[1] : 1;
  k.d2 = T.nbody > 2 ? T.dims[2] : 1;
  k.d3 = T.nbody > 3 ? T.dims[3] : 1;
  k.V = a.ints + T.ioff;
}

// B_t(J) = sum_p |V[p]| prod_q max_I |Γ_q[p_q, I, J_q]| (for one body: max_I |H[I, J]|)
__device__ __forceinline__ double ket_bound(const KetBound &k, const uint8_t *st) {
  const double *c0 = k.cmx[0] + (int64_t)k.np[0] * st[k.cl[0]];
  if (k.nb == 1) return c0[0];
  const double *c1 = k.cmx[1] + (int64_t)k.np[1] * st[k.cl[1]];
  const double *c2 = k.nb > 2 ? k.cmx[2] + (int64_t)k.np[2] * st[k.cl[2]] : nullptr;
  const double *c3 = k.nb > 3 ? k.cmx[3] + (int64_t)k.np[3] * st[k.cl[3]] : nullptr;
  double B = 0.0;
  for (int p3 = 0; p3 < k.d3; p3++) {
    const double f3 = c3 ? c3[p3] : 1.0;
    if (f3 == 0.0) continue;
    for (int p2 = 0; p2 < k.d2; p2++) {
      const double f2 = f3 * (c2 ? c2[p2] : 1.0);
      if (f2 == 0.0) continue;
      for (int p1 = 0; p1 < k.d1; p1++) {
        const double f1 = f2 * c1[p1];
        if (f1 == 0.0) continue;
        const double *Vp = k.V + k.d0 * (p1 + k.d1 * (p2 + k.d2 * p3));
        double s = 0.0;
        for (int p0 = 0; p0 < k.d0; p0++) s += fabs(Vp[p0]) * c0[p0];
        B += s * f1;
      }
    }
  }
  return B;
}

// ---- screening: kets per item that pass their own bound (one wavefront per item) ----
__global__ __launch_bounds__(SW) void k_pt2_screen(Pt2Args a) {
  const int64_t it = a.item_base + blockIdx.x;
  if (it >= a.n_items) return;  // uniform
  const int lane = threadIdx.x;
  const int fj = a.items_f[it];
  const PTerm T = a.terms[a.items_t[it]];
  KetBound kb;
  ket_bound_setup(a, T, fj, kb);
  const double Bi = a.items_b[it] * SCREEN_SLACK;
  const int64_t k0 = a.fstart[fj], k1 = a.fstart[fj + 1];
  int count = 0;
  for (int64_t base = k0; base < k1; base += SW) {
    const int64_t k = base + lane;
    bool cand = false, pass = false;
    if (k < k1) {
      const double cj = a.csort[k];
      cand = cj * Bi > a.thresh;
      if (cand) {
        const double B = ket_bound(kb, a.cst + (int64_t)a.kord[k] * a.N);
        pass = B != 0.0 && B * SCREEN_SLACK * cj > a.thresh;
      }
    }
    count += __popcll(__ballot(pass));
    if (!__all(cand)) break;  // kets are ordered by |c|: no later candidate
  }
  if (lane == 0) a.nsurv[it] = count;
}

// block-wide exclusive scan of one value per thread (wavefront shuffles, then
// the wavefront totals through LDS); returns the total.  Two barriers.
__device__ __forceinline__ int64_t block_scan(int64_t v, int64_t *sWave, int64_t *out_excl) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int64_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sWave[w] = x;
  __syncthreads();
  int64_t pre = 0, tot = 0;
  for (int k = 0; k < PT / 64; k++) {
    const int64_t t = sWave[k];
    if (k < w) pre += t;
    tot += t;
  }
  *out_excl = pre + x - v;
  __syncthreads();  // sWave is free again
  return tot;
}

// ---- contraction and emission: every kept contribution of an item ----
// One workgroup per slot (an item with a kept ket).  COUNT:
// emit_count[slot] = number of kept contributions; FILL: writes them from
// emit_ptr[slot] - emit_base on, in the order (candidate batch, group in
// active-state order, tile, output element, ket in |c| order), the same in
// both passes.
//
// The chain contracts the last active cluster first.  Its intermediates are
// tiled over the surviving states of the last cluster (T = 1) or, when a
// 4-body term's stage after the second cluster would not fit, over the
// pairs of surviving states of the last two clusters (T = 2), so that every
// stored stage fits half the arena and several workgroups share a CU.
template <bool FILL, typename TC>
__global__ __launch_bounds__(PT) void k_pt2_emit(Pt2Args a) {
  __shared__ double sV[VMAX];
  __shared__ TC sA[ARENA];           // chain intermediates in the contraction's precision
  __shared__ double sW[4][NPMAX];     // per-cluster bound weights w_q[p]
  __shared__ double sCM[4][NPMAX];    // column maxima of the group's ket states
  __shared__ int16_t sSurv[4][SMAX];  // surviving bra states per active cluster
  __shared__ uint8_t sFlag[4][SMAX];
  __shared__ int sNS[4];
  __shared__ int32_t sBKet[KCH];      // kept kets of the batch, in candidate order
  __shared__ double sBCab[KCH];       // their max_r |c|
  __shared__ double sBThr[KCH];       // their element thresholds
  __shared__ uint64_t sBLo[KCH];      // their state keys with the active clusters cleared
  __shared__ uint32_t sKey[KCH];      // active-state keys, sorted with sPos
  __shared__ int16_t sPos[KCH];
  __shared__ int sGrp[KCH + 1];
  __shared__ int sNG, sDone;
  __shared__ int64_t sWave[PT / 64];
  __shared__ uint64_t sTloc;
  __shared__ int sBlk[4], sNp[4], sMb[4], sMk[4], sCl[4], sDims[4];
  __shared__ double sSign;
  __shared__ int64_t sEmit;  // kept contributions so far (count) / next position (fill)
  __shared__ int sSkip;

  const int tid = threadIdx.x;
  const int64_t slot = a.slot_base + blockIdx.x;
  if (slot >= a.n_slots) return;  // uniform
  const int64_t it = a.slot_item[slot];
  const int fj = a.items_f[it];
  const PTerm T = a.terms[a.items_t[it]];
  const int N = a.N, S = a.S, nb = T.nbody;
  const int16_t *F = a.fock + (int64_t)fj * 2 * N;
  const int16_t *D = a.tc + (int64_t)T.group * 2 * N;
  if (tid == 0) {
    sTloc = FILL ? a.slot_tloc[slot] : 0;
    int par = 0, nv = 1, skip = 0;
#pragma unroll
    for (int q = 0; q < 4; q++) {  // constant trip count: the term stays in registers
      sCl[q] = q < nb ? T.cl[q] : 0;
      sDims[q] = q < nb ? T.dims[q] : 1;
      if (q < nb) {
        const int c = T.cl[q];
        const int ck = sector_code(F[2 * c], F[2 * c + 1], S);
        const int cb = sector_code(F[2 * c] + D[2 * c], F[2 * c + 1] + D[2 * c + 1], S);
        const int blk = gblk(a, c, T.op[q], cb, ck);
        sBlk[q] = blk;
        sNp[q] = a.gb_dims[3 * blk];
        sMb[q] = a.gb_dims[3 * blk + 1];
        sMk[q] = a.gb_dims[3 * blk + 2];
        if ((T.odd >> q) & 1) {
          int ne = 0;
          for (int c2 = 0; c2 < c; c2++) ne += F[2 * c2] + F[2 * c2 + 1];
          par ^= ne & 1;
        }
        if (nb > 1) nv *= T.dims[q];
        // shapes the staged chain cannot hold are reported, not computed
        skip |= sMb[q] > SMAX || sNp[q] > NPMAX || sMk[q] > 255;
      }
    }
    sSign = par ? -1.0 : 1.0;
    sEmit = FILL ? a.emit_ptr[slot] - a.emit_base : 0;
    skip |= nv > VMAX;
    sSkip = skip;
    if (skip && !FILL) atomicAdd(&a.stats[2], 1ull);
  }
  __syncthreads();
  if (sSkip) {
    if (!FILL && tid == 0) a.emit_count[slot] = 0;
    return;
  }
  const double sign = sSign;
  int nv = 1;
  if (nb > 1)
    for (int q = 0; q < nb; q++) nv *= sDims[q];
  for (int e = tid; e < nv; e += PT) sV[e] = nb > 1 ? a.ints[T.ioff + e] : 1.0;
  KetBound kb;
  ket_bound_setup(a, T, fj, kb);
  kb.V = sV;  // the staged integrals (read after the batch loop's first barrier)
  const double Bi = a.items_b[it] * SCREEN_SLACK;
  const int64_t k0 = a.fstart[fj], k1 = a.fstart[fj + 1];
  unsigned long long groups_kept = 0, flops = 0;
  for (int64_t cb0 = k0; cb0 < k1; cb0 += KCH) {
    __syncthreads();
    // candidates of this batch and their own bounds
    const int64_t k = cb0 + tid;
    bool cand = false;
    int pass = 0;
    int32_t J = 0;
    double cj = 0.0;
    if (k < k1) {
      cj = a.csort[k];
      cand = cj * Bi > a.thresh;
      if (cand) {
        J = a.kord[k];
        const double B = ket_bound(kb, a.cst + (int64_t)J * N);
        pass = B != 0.0 && B * BOUND_SLACK * cj > a.thresh;
      }
    }
    if (tid == PT - 1) sDone = !cand;  // kets are ordered by |c|
    int64_t pos;
    const int nbat = (int)block_scan(pass, sWave, &pos);
    if (pass) {
      sBKet[pos] = J;
      sBCab[pos] = cj;
      sBThr[pos] = a.thr[J];
      const uint8_t *st = a.cst + (int64_t)J * N;
      uint32_t key = 0;  // states below 255 (checked on the host): never the sentinel
      uint64_t lo = 0;
      for (int cc = 0; cc < N; cc++) lo |= (uint64_t)st[cc] << (a.sbits * cc);
#pragma unroll
      for (int q = 0; q < 4; q++)
        if (q < nb) {
          key |= (uint32_t)st[sCl[q]] << (8 * q);
          lo &= ~(((uint64_t(1) << a.sbits) - 1) << (a.sbits * sCl[q]));
        }
      sBLo[pos] = lo;
      sKey[pos] = key;
      sPos[pos] = (int16_t)pos;
    }
    int P2 = 1;
    if (nbat > 1) {
      P2 = 2;
      while (P2 < nbat) P2 <<= 1;
      if (tid >= nbat && tid < P2) {
        sKey[tid] = 0xFFFFFFFFu;
        sPos[tid] = (int16_t)tid;
      }
    }
    __syncthreads();
    if (nbat > 1) {
      // bitonic sort of (key, position) over the next power of two
      for (int kk = 2; kk <= P2; kk <<= 1)
        for (int j = kk >> 1; j > 0; j >>= 1) {
          const int i = tid;
          if (i < P2) {
            const int l = i ^ j;
            if (l > i) {
              const bool up = (i & kk) == 0;
              const uint32_t ki = sKey[i], kl = sKey[l];
              const int vi = sPos[i], vl = sPos[l];
              const bool gt = ki > kl || (ki == kl && vi > vl);
              if (gt == up) {
                sKey[i] = kl;
                sKey[l] = ki;
                sPos[i] = (int16_t)vl;
                sPos[l] = (int16_t)vi;
              }
            }
          }
          __syncthreads();
        }
    }
    if (tid == 0) {
      int ng = 0;
      for (int j = 0; j < nbat; j++)
        if (j == 0 || sKey[j] != sKey[j - 1]) sGrp[ng++] = j;
      sGrp[ng] = nbat;
      sNG = ng;
    }
    __syncthreads();
    const int ng = sNG;
    for (int gi = 0; gi < ng; gi++) {
      const int j0 = sGrp[gi], j1 = sGrp[gi + 1];
      // the group's active ket states; its largest |c| is its first member's
      // (the batch is in |c| order and the sort keeps that order in a group)
      const uint8_t *st0 = a.cst + (int64_t)sBKet[sPos[j0]] * N;
      const double cm = sBCab[sPos[j0]];
      // column maxima max_I |Γ_q[p, I, J_q]| for the group's ket states
      for (int e = tid; e < 4 * NPMAX; e += PT) {
        const int q = e / NPMAX, p = e % NPMAX;
        if (q < nb && p < sNp[q])
          sCM[q][p] = a.cmx[a.cmx_off[sBlk[q]] + (int64_t)sNp[q] * st0[sCl[q]] + p];
      }
      __syncthreads();
      // per-cluster weights w_q[p_q] = sum over the other indices of
      // |V| prod_{r != q} colmax_r
      if (nb > 1) {
        for (int e = tid; e < 4 * NPMAX; e += PT) {
          const int q = e / NPMAX, pq = e % NPMAX;
          if (q >= nb || pq >= sNp[q]) continue;
          // the other three indices (unused ones of extent 1) as an odometer
          int dA = 1, dB = 1, dC = 1, sA_ = 0, sB_ = 0, sC_ = 0, rA = 0, rB = 0, rC = 0;
          int sq = 1, acc = 1, n_other = 0;
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const int dr = r < nb ? sDims[r] : 1;
            if (r == q) {
              sq = acc;
            } else {
              if (n_other == 0) { dA = dr; sA_ = acc; rA = r; }
              else if (n_other == 1) { dB = dr; sB_ = acc; rB = r; }
              else { dC = dr; sC_ = acc; rC = r; }
              n_other++;
            }
            acc *= dr;
          }
          const double *cA = sCM[rA], *cB = sCM[rB], *cC = sCM[rC];
          const bool hB = nb > 2, hC = nb > 3;
          double s = 0.0;
          for (int iC = 0; iC < dC; iC++) {
            const double fC = hC ? cC[iC] : 1.0;
            for (int iB = 0; iB < dB; iB++) {
              const double fB = fC * (hB ? cB[iB] : 1.0);
              const int oB = pq * sq + iB * sB_ + iC * sC_;
              double t = 0.0;
              for (int iA = 0; iA < dA; iA++) t += fabs(sV[oB + iA * sA_]) * cA[iA];
              s += t * fB;
            }
          }
          sW[q][pq] = s;
        }
      } else if (tid == 0) {
        sW[0][0] = 1.0;
      }
      __syncthreads();
      // survivor lists: bra states I_q whose bound can pass
      {
        const int q = tid / SMAX, I = tid % SMAX;
        if (q < nb && I < sMb[q]) {
          const int Jq = st0[sCl[q]];
          const int np = sNp[q];
          const double *Gi = a.gdata + a.gb_off[sBlk[q]] + (int64_t)np * (I + (int64_t)sMb[q] * Jq);
          double s = 0.0;
          for (int p = 0; p < np; p++) s += sW[q][p] * fabs(Gi[p]);
          sFlag[q][I] = s != 0.0 && s * BOUND_SLACK * cm > a.thresh;
        }
      }
      __syncthreads();
      if (tid < nb) {
        const int q = tid;
        int n = 0;
        for (int I = 0; I < sMb[q]; I++)
          if (sFlag[q][I]) sSurv[q][n++] = (int16_t)I;
        sNS[q] = n;
      }
      __syncthreads();
      bool empty = false;
      for (int q = 0; q < nb; q++) empty |= sNS[q] == 0;
      if (empty) continue;  // uniform; nothing below the last barrier is rewritten first
      groups_kept++;
      // tiling (see above).  Stored stage sizes per tile state: after the
      // last cluster pre_full = prod_{q<last} d_q; with T = 2 after cluster
      // last-1 pre_full / d_{last-1}; after an untiled cluster q (the
      // highest untiled down to 1) prod_{r<q} d_r * prod_{r=q}^{hi} |S_r|;
      // cluster 0 is contracted per output during the emission
      const int last = nb - 1;
      int pre_full = 1;
      for (int q = 0; q < last; q++) pre_full *= nb > 1 ? sDims[q] : 1;
      int TT = 1, need = 1;
      if (nb > 1) {
        for (TT = 1; TT <= 2; TT++) {
          if (TT == 2 && nb < 3) break;
          int pre = pre_full;
          need = pre;
          if (TT == 2) {
            pre /= sDims[last - 1];
            need = max(need, pre);
          }
          int ps = 1;
          for (int q = last - TT; q >= 1; q--) {
            pre /= sDims[q];
            ps *= sNS[q];
            need = max(need, pre * ps);
          }
          if (need <= ARENA / 2) break;
        }
      }
      if (need > ARENA / 2 || TT > 2) {  // uniform: flags the call, computes nothing wrong
        if (tid == 0 && !FILL) atomicAdd(&a.stats[2], 1ull);
        __syncthreads();
        continue;
      }
      const int nsA = nb > 1 && TT == 2 ? sNS[last - 1] : 1;  // combination m = iA + nsA * iL
      const int ntile = nb == 1 ? 1 : sNS[last] * nsA;
      const int tk = nb == 1 ? 1 : min(ntile, (ARENA / 2) / need);
      const int J0 = st0[sCl[0]];
      const double *G0 = a.gdata + a.gb_off[sBlk[0]];
      for (int t0 = 0; t0 < ntile; t0 += tk) {
        const int tn = nb == 1 ? 1 : min(tk, ntile - t0);
        TC *cur = sA, *nxt = sA + ARENA / 2;
        int cols = 1;
        if (nb > 1) {
          int pre = pre_full;
          {  // the last cluster: cur[r + pre * i] = sum_p V[r + pre * p] Γ[p, S[m_L], J]
            const int q = last;
            const int np = sNp[q], mb = sMb[q];
            const int Jq = st0[sCl[q]];
            const double *G = a.gdata + a.gb_off[sBlk[q]];
            for (int e = tid; e < pre * tn; e += PT) {
              const int r = e % pre, i = e / pre;
              const int iL = (t0 + i) / nsA;
              const double *Gi = G + (int64_t)np * (sSurv[q][iL] + (int64_t)mb * Jq);
              TC v = 0;
              for (int p = 0; p < np; p++) v += (TC)sV[r + pre * p] * (TC)Gi[p];
              cur[e] = v;
            }
            flops += 2ull * pre * tn * np;
          }
          cols = tn;
          __syncthreads();
          if (TT == 2) {  // cluster last-1 at the tile's own state: no expansion
            const int q = last - 1;
            const int np = sNp[q], mb = sMb[q];
            const int Jq = st0[sCl[q]];
            const double *G = a.gdata + a.gb_off[sBlk[q]];
            const int pre2 = pre / np;
            for (int e = tid; e < pre2 * tn; e += PT) {
              const int r2 = e % pre2, i = e / pre2;
              const int iA = (t0 + i) % nsA;
              const double *Gi = G + (int64_t)np * (sSurv[q][iA] + (int64_t)mb * Jq);
              TC v = 0;
              for (int p = 0; p < np; p++) v += cur[r2 + pre2 * p + pre * i] * (TC)Gi[p];
              nxt[e] = v;
            }
            flops += 2ull * pre2 * tn * np;
            __syncthreads();
            TC *tmp = cur;
            cur = nxt;
            nxt = tmp;
            pre = pre2;
          }
          for (int q = last - TT; q >= 1; q--) {
            // nxt[r2 + pre2 * (i + ns * col)] = sum_p cur[r2 + pre2 * p + pre * col] Γ_q[p, S_q[i], J]
            const int np = sNp[q], mb = sMb[q];
            const int Jq = st0[sCl[q]];
            const double *G = a.gdata + a.gb_off[sBlk[q]];
            const int pre2 = pre / np, ns = sNS[q];
            for (int e = tid; e < pre2 * ns * cols; e += PT) {
              const int r2 = e % pre2, rem = e / pre2, i = rem % ns, col = rem / ns;
              const double *Gi = G + (int64_t)np * (sSurv[q][i] + (int64_t)mb * Jq);
              TC v = 0;
              for (int p = 0; p < np; p++) v += cur[r2 + pre2 * p + pre * col] * (TC)Gi[p];
              nxt[e] = v;
            }
            flops += 2ull * pre2 * ns * cols * np;
            __syncthreads();
            TC *tmp = cur;
            cur = nxt;
            nxt = tmp;
            pre = pre2;
            cols *= ns;
          }
        }
        const int ns0 = sNS[0];
        const int nout = ns0 * cols;
        flops += 2ull * nout * (nb == 1 ? 0 : sNp[0]);
        // output e = i0 + ns0 * col, col = i_1 + |S_1| (i_2 + ... (tile index));
        // kept contributions counted (and, filling, written in this order)
        const int64_t base = sEmit;  // read before the scan's barriers
        int64_t excl0 = 0;           // this thread's first position (pass 1)
        for (int pass2 = 0; pass2 < (FILL ? 2 : 1); pass2++) {
          int64_t wpos = base + excl0, cnt = 0;
          for (int e = tid; e < nout; e += PT) {
            const int i0 = e % ns0, col = e / ns0;
            double x;
            if (nb == 1) {
              x = (double)(TC)G0[sSurv[0][i0] + (int64_t)sMb[0] * J0];  // 2-D block (np = 1)
            } else {
              const int np = sNp[0];
              const double *Gi = G0 + (int64_t)np * (sSurv[0][i0] + (int64_t)sMb[0] * J0);
              TC v = 0;
              for (int p = 0; p < np; p++) v += cur[p + np * col] * (TC)Gi[p];
              x = (double)v;
            }
            const double ax = fabs(x);
            if (pass2 == 0) {
              for (int j = j0; j < j1; j++) cnt += ax > sBThr[sPos[j]];
              continue;
            }
            // key: target Fock sectors | bra states on the active clusters,
            // the ket's states elsewhere
            uint64_t act = (uint64_t)sSurv[0][i0] << (a.sbits * sCl[0]);
            if (nb > 1) {
              int rem = col;
              for (int q = 1; q <= last - TT; q++) {
                act |= (uint64_t)sSurv[q][rem % sNS[q]] << (a.sbits * sCl[q]);
                rem /= sNS[q];
              }
              const int m = t0 + rem;
              act |= (uint64_t)sSurv[last][m / nsA] << (a.sbits * sCl[last]);
              if (TT == 2) act |= (uint64_t)sSurv[last - 1][m % nsA] << (a.sbits * sCl[last - 1]);
            }
            const double v = sign * x;
            for (int j = j0; j < j1; j++) {
              const int pj = sPos[j];
              if (!(ax > sBThr[pj])) continue;  // _collect_significant_thread!'s test
              const int64_t Jm = sBKet[pj];
              a.ekey[wpos] = chunk_key(sTloc, sBLo[pj] | act, a.sN);
              for (int r = 0; r < a.R; r++)
                a.eval[wpos * a.R + r] = v * a.c[(int64_t)r * a.n_cfg + Jm];
              wpos++;
            }
          }
          if (pass2 == 0) {
            const int64_t tot = block_scan(cnt, sWave, &excl0);
            // block_scan's barriers: every thread has read sEmit
            if (tid == 0) sEmit = base + tot;
          }
        }
        __syncthreads();
      }
    }
    if (sDone) break;  // uniform: written before the batch's scan barriers
  }
  __syncthreads();
  if (!FILL && tid == 0) a.emit_count[slot] = sEmit;
  if (!FILL && tid == 0 && groups_kept) atomicAdd(&a.stats[0], groups_kept);
  // every thread counted the same (uniform) chain flops
  if (!FILL && tid == 0 && flops) atomicAdd(&a.stats[1], flops);
}

__global__ void k_gather_keys(int64_t n, const uint32_t *idx, const uint64_t *key,
                              uint64_t *out) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    out[i] = key[idx[i]];
}

// per Γ block: bmax[p] = max_{I,J} |Γ[p,I,J]|, cmx[p + np J] = max_I |Γ[p,I,J]|
__global__ void k_gamma_max(int64_t n_blocks, const int64_t *gb_off, const int32_t *gb_dims,
                            const double *gdata, const int64_t *bmax_off, double *bmax,
                            const int64_t *cmx_off, double *cmx) {
  const int64_t b = blockIdx.x;
  if (b >= n_blocks) return;
  const int np = gb_dims[3 * b], mb = gb_dims[3 * b + 1], mk = gb_dims[3 * b + 2];
  const double *G = gdata + gb_off[b];
  for (int e = threadIdx.x; e < np * mk; e += blockDim.x) {
    const int p = e % np, J = e / np;
    double m = 0.0;
    for (int I = 0; I < mb; I++) m = fmax(m, fabs(G[p + (int64_t)np * (I + (int64_t)mb * J)]));
    cmx[cmx_off[b] + e] = m;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < np; p += blockDim.x) {
    double m = 0.0;
    for (int J = 0; J < mk; J++) m = fmax(m, cmx[cmx_off[b] + p + (int64_t)np * J]);
    bmax[bmax_off[b] + p] = m;
  }
}

// segment heads of the sorted keys (x changes)
__global__ void k_heads(int64_t n, const Key128 *k, uint32_t *head) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    head[i] = (i == 0 || k[i].hi != k[i - 1].hi || k[i].lo != k[i - 1].lo) ? 1u : 0u;
}

__global__ void k_seg_start(int64_t n, const uint32_t *head, const uint64_t *segid,
                            int64_t *seg_start) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    if (head[i]) seg_start[segid[i]] = i;
}

__device__ __forceinline__ void neumaier(double &sum, double &comp, double v) {
  const double t = sum + v;
  comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
  sum = t;
}

struct SumArgs {
  int N, S, R, n_ops, h0_op, fbits, sbits, sN;
  int64_t n_seg;
  const int64_t *seg_start;  // [n_seg + 1]
  const uint64_t *tkeys;     // the chunk's target Fock keys by index
  const double *sval;        // one root: values in sorted order (perm unused)
  const uint64_t *perm;      // several roots: sorted position -> emission index
  const Key128 *skey;        // sorted keys
  const double *eval;        // [E][R] in emission order
  const Key128 *vkey;        // variational configurations, sorted
  int64_t n_v;
  const int64_t *tvar;       // per chunk target: [start, end) of its Fock block in vkey
  const int32_t *gblock;
  const int64_t *gb_off;
  const int32_t *gb_dims;
  const double *gdata;
  const double *h0tab;       // [N][S*S][2^sbits]: H0_c[x, x] by (cluster, sector, state)
  const double *e0;
  double *part;              // [blocks][R][2]
  unsigned long long *n_out; // configurations outside the variational space
};

// the H0 diagonal of every (cluster, sector, state) in one table, so that a
// configuration's Hd is N table reads instead of N chains of Γ-block lookups
__global__ void k_h0_table(SumArgs a, double *tab) {
  const int SS = a.S * a.S, M = 1 << a.sbits;
  const int64_t n = (int64_t)a.N * SS * M;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i / ((int64_t)SS * M)), code = (int)((i / M) % SS), xs = (int)(i % M);
    const int blk = a.gblock[(((int64_t)c * a.n_ops + a.h0_op) * SS + code) * SS + code];
    double v = 0.0;
    if (blk >= 0) {
      const int mb = a.gb_dims[3 * blk + 1];
      if (xs < mb) v = a.gdata[a.gb_off[blk] + xs + (int64_t)mb * xs];
    }
    tab[i] = v;
  }
}

// per x: sigma = sum of its contributions in sorted order; outside the
// variational space, sigma^2 / (E0 - Hd) with Hd = sum_c H0_c[x_c, x_c]
__global__ __launch_bounds__(256) void k_pt2_seg_sum(SumArgs a) {
  __shared__ double ss[256], sc[256];
  __shared__ unsigned long long sOut;
  if (threadIdx.x == 0) sOut = 0;
  __syncthreads();
  const int R = a.R;
  double sum[FCG_MAX_R], comp[FCG_MAX_R];
  for (int r = 0; r < FCG_MAX_R; r++) sum[r] = comp[r] = 0.0;
  unsigned long long nout = 0;
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n_seg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = a.seg_start[s], e = a.seg_start[s + 1];
    const Key128 kc = a.skey[b];
    const uint64_t tl = chunk_key_target(kc, a.sN);
    const Key128 k{a.tkeys[tl], chunk_key_states(kc, a.sN)};
    // variational configurations are projected out (searched among the
    // target Fock block's own, an empty range when it holds none)
    const int64_t v1 = a.tvar[2 * tl + 1];
    int64_t lo = a.tvar[2 * tl], hi = v1;
    while (lo < hi) {
      const int64_t mid = (lo + hi) / 2;
      if (a.vkey[mid].lo < k.lo) lo = mid + 1;
      else hi = mid;
    }
    if (lo < v1 && a.vkey[lo].lo == k.lo) continue;
    nout++;
    double hd = 0.0;
    const int SS = a.S * a.S;
    for (int c = 0; c < a.N; c++) {
      const int code = (int)((k.hi >> (a.fbits * c)) & ((1ull << a.fbits) - 1));
      const int xs = (int)((k.lo >> (a.sbits * c)) & ((1ull << a.sbits) - 1));
      hd += a.h0tab[(((int64_t)c * SS + code) << a.sbits) + xs];
    }
    for (int r = 0; r < R; r++) {
      double v = 0.0;
      if (a.sval)
        for (int64_t i = b; i < e; i++) v += a.sval[i];
      else
        for (int64_t i = b; i < e; i++) v += a.eval[a.perm[i] * R + r];
      neumaier(sum[r], comp[r], v * v / (a.e0[r] - hd));
    }
  }
  if (nout) atomicAdd(&sOut, nout);
  for (int r = 0; r < R; r++) {
    ss[threadIdx.x] = sum[r];
    sc[threadIdx.x] = comp[r];
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0, c = 0.0;
      for (int t = 0; t < (int)blockDim.x; t++) {
        neumaier(s, c, ss[t]);
        c += sc[t];
      }
      a.part[((int64_t)blockIdx.x * R + r) * 2] = s;
      a.part[((int64_t)blockIdx.x * R + r) * 2 + 1] = c;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && sOut) atomicAdd(a.n_out, sOut);
}

// open sigma: per x (segment), its full key and sigma_r = the sum of its
// contributions in sorted order (the order k_pt2_seg_sum sums them in)
__global__ __launch_bounds__(256) void k_open_out(SumArgs a, uint64_t *khi, uint64_t *klo,
                                                  double *sig) {
  for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < a.n_seg;
       s += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = a.seg_start[s], e = a.seg_start[s + 1];
    const Key128 kc = a.skey[b];
    khi[s] = a.tkeys[chunk_key_target(kc, a.sN)];
    klo[s] = chunk_key_states(kc, a.sN);
    for (int r = 0; r < a.R; r++) {
      double v = 0.0;
      if (a.sval)
        for (int64_t i = b; i < e; i++) v += a.sval[i];
      else
        for (int64_t i = b; i < e; i++) v += a.eval[a.perm[i] * a.R + r];
      sig[s * a.R + r] = v;
    }
  }
}

__global__ void k_iota(int64_t n, uint64_t *p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = (uint64_t)i;
}

template <typename T>
struct Buf {
  fcg_ctx *ctx = nullptr;
  T *p = nullptr;
  explicit Buf(fcg_ctx *c) : ctx(c) {}
  ~Buf() {
    if (p) fcg_dfree(ctx, p);
  }
  bool alloc(int64_t n) {
    if (p) fcg_dfree(ctx, p);
    p = static_cast<T *>(fcg_dmalloc(ctx, sizeof(T) * std::max<int64_t>(n, 1)));
    return p != nullptr;
  }
  int upload(const T *src, int64_t n) {
    if (!alloc(n)) return FCG_ENOMEM;
    if (n > 0)
      FCG_HIP(hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream), ctx);
    return FCG_OK;
  }
};

// FCG_VERBOSE=1 in the environment: phase timings on stderr
bool verbose() {
  static const bool v = [] {
    const char *e = std::getenv("FCG_VERBOSE");
    return e && e[0] == '1';
  }();
  return v;
}

void vlog(const char *what, std::chrono::steady_clock::time_point t0, double x = -1.0) {
  if (!verbose()) return;
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (x >= 0.0) std::fprintf(stderr, "[fcg pt2] %-28s %10.3f s  %.6g\n", what, s, x);
  else std::fprintf(stderr, "[fcg pt2] %-28s %10.3f s\n", what, s);
  std::fflush(stderr);
}

int bits_for(int64_t n) {  // bits to hold values 0 .. n-1
  int b = 0;
  while ((int64_t(1) << b) < n) b++;
  return std::max(b, 1);
}

constexpr int64_t FB_CHUNK = 65535;
constexpr int64_t ITEM_CHUNK = 1 << 24;      // workgroups per screening launch
constexpr int64_t SLOT_CHUNK = 1 << 22;      // workgroups per contraction launch
constexpr int64_t CHUNK_E = int64_t(1) << 27;  // contributions per accumulation chunk
constexpr int SUM_BLOCKS = 1024;

// contributions per accumulation chunk: FCG_PT2_CHUNK in the environment
// lowers it for devices with little free memory
int64_t chunk_limit() {
  const char *e = std::getenv("FCG_PT2_CHUNK");
  if (e) {
    const long long v = std::atoll(e);
    if (v > 0) return v;
  }
  return CHUNK_E;
}

double seconds_since(std::chrono::steady_clock::time_point t) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

// Configurations reached by the sparse open matvec, in the engine's order:
// packed keys (Fock sector codes, states) and sigma (R per configuration)
struct fcg_open_result {
  int N = 0, R = 0, S = 0, fbits = 0, sbits = 0;
  std::vector<uint64_t> khi, klo;
  std::vector<double> sig;
};

static int pt2_sparse_run(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t var, int R,
                          const double *c, const double *e0, int nbody, int h0_op,
                          double thresh, int precision, double *e2,
                          fcg_pt2_sparse_stats *st, fcg_open_result *open);

extern "C" int fcg_tpsci_pt2_sparse(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t var,
                                    int R, const double *c, const double *e0, int nbody,
                                    int h0_op, double thresh, double *e2,
                                    fcg_pt2_sparse_stats *st) {
  return fcg_tpsci_pt2_sparse_ex(ctx, h, o, var, R, c, e0, nbody, h0_op, thresh, FCG_FP64, e2, st);
}

extern "C" int fcg_tpsci_pt2_sparse_ex(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t var,
                                       int R, const double *c, const double *e0, int nbody,
                                       int h0_op, double thresh, int precision, double *e2,
                                       fcg_pt2_sparse_stats *st) {
  return pt2_sparse_run(ctx, h, o, var, R, c, e0, nbody, h0_op, thresh, precision, e2, st,
                        nullptr);
}

extern "C" int fcg_open_sigma_sparse(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t ket,
                                     int nbody, int R, const double *c, double thresh,
                                     fcg_open_result_t *out) {
  if (!out || R < 1 || R > FCG_MAX_R) return FCG_EINVAL;
  *out = nullptr;
  auto res = std::make_unique<fcg_open_result>();
  std::vector<double> e0(R, 0.0), e2(R, 0.0);
  const int rc = pt2_sparse_run(ctx, h, o, ket, R, c, e0.data(), nbody, 0, thresh, FCG_FP64,
                                e2.data(), nullptr, res.get());
  if (rc) return rc;
  *out = res.release();
  return FCG_OK;
}

extern "C" int fcg_open_result_size(fcg_open_result_t res, int64_t *n) {
  if (!res || !n) return FCG_EINVAL;
  *n = (int64_t)res->klo.size();
  return FCG_OK;
}

extern "C" int fcg_open_result_copy(fcg_open_result_t res, int16_t *fock, int16_t *cfg,
                                    double *sigma) {
  if (!res || !fock || !cfg || !sigma) return FCG_EINVAL;
  const int N = res->N, R = res->R, S = res->S;
  const int64_t n = (int64_t)res->klo.size();
  const uint64_t fm = (uint64_t(1) << res->fbits) - 1, sm = (uint64_t(1) << res->sbits) - 1;
  for (int64_t x = 0; x < n; x++) {
    for (int cc = 0; cc < N; cc++) {
      const int code = (int)((res->khi[x] >> (res->fbits * cc)) & fm);
      fock[(x * N + cc) * 2] = (int16_t)(code / S);
      fock[(x * N + cc) * 2 + 1] = (int16_t)(code % S);
      cfg[x * N + cc] = (int16_t)(((res->klo[x] >> (res->sbits * cc)) & sm) + 1);
    }
    for (int r = 0; r < R; r++) sigma[(int64_t)r * n + x] = res->sig[x * R + r];
  }
  return FCG_OK;
}

extern "C" void fcg_open_result_destroy(fcg_open_result_t res) { delete res; }

static int pt2_sparse_run(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t var, int R,
                          const double *c, const double *e0, int nbody, int h0_op,
                          double thresh, int precision, double *e2,
                          fcg_pt2_sparse_stats *st, fcg_open_result *open) {
  if (!ctx || !h || !o || !var || !c || !e0 || !e2 || R < 1 || R > FCG_MAX_R || nbody < 1 ||
      nbody > 4 || h0_op < 0 || h0_op >= o->n_ops || !(thresh >= 0.0) ||
      (precision != FCG_FP64 && precision != FCG_FP32))
    return FCG_EINVAL;
  const bool fp32 = precision == FCG_FP32;
  if (!var->tpsci) {
    ctx->err = "fcg_tpsci_pt2_sparse needs a TPSCI space";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  const auto t0 = std::chrono::steady_clock::now();
  const int N = var->N, S = o->S;
  hipStream_t s = ctx->stream;
  // configurations of the space, grouped by Fock block: read straight from
  // the uploaded configurations while the space has not been cut into boxes
  // (the boxes serve plans, not this engine), else enumerated from the boxes
  // (which keep the caller's order)
  const int64_t n_cfg = var->nvar, n_fock = var->n_fock;
  if (n_cfg >= (int64_t(1) << 31)) {
    ctx->err = "fcg_tpsci_pt2_sparse: more than 2^31 configurations";
    return FCG_EOVERFLOW;
  }
  std::vector<uint8_t> cst((size_t)std::max<int64_t>(n_cfg, 1) * N);
  std::vector<int64_t> fstart(n_fock + 1, 0);
  const bool unboxed = var->n_box < 0;
  if (unboxed) {
    // lo holds the configurations (1-based, n_cfg x N), var_off the Fock
    // blocks' first configurations
    if ((int64_t)var->var_off.size() != n_fock + 1 || var->var_off[n_fock] != n_cfg ||
        (int64_t)var->lo.size() != n_cfg * N) {
      ctx->err = "fcg_tpsci_pt2_sparse: inconsistent TPSCI space";
      return FCG_EINVAL;
    }
    for (int64_t f = 0; f <= n_fock; f++) {
      fstart[f] = var->var_off[f];
      if ((f == 0 && fstart[0] != 0) || (f > 0 && fstart[f] < fstart[f - 1])) {
        ctx->err = "fcg_tpsci_pt2_sparse: Fock block starts must ascend from 0";
        return FCG_EINVAL;
      }
    }
    int over = 0;
#pragma omp parallel for schedule(static) reduction(| : over)
    for (int64_t e = 0; e < n_cfg * N; e++) {
      const int st0 = var->lo[e] - 1;
      over |= st0 < 0 || st0 > 255;
      cst[e] = (uint8_t)st0;
    }
    if (over) {
      ctx->err = "fcg_tpsci_pt2_sparse: state index outside 1..256";
      return FCG_EOVERFLOW;
    }
  }
  for (int64_t b = 0; !unboxed && b < var->n_box; b++) {
    const int32_t f = var->box_fock[b];
    if (b > 0 && f < var->box_fock[b - 1]) {
      ctx->err = "fcg_tpsci_pt2_sparse: Fock blocks must be contiguous";
      return FCG_EINVAL;
    }
    fstart[f + 1] += var->var_off[b + 1] - var->var_off[b];
  }
  int over = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(| : over)
  for (int64_t b = 0; b < (unboxed ? 0 : var->n_box); b++) {
    const int64_t v0 = var->var_off[b], d = var->var_off[b + 1] - v0;
    for (int64_t e = 0; e < d; e++) {
      int64_t r = e;
      for (int cc = 0; cc < N; cc++) {
        const int rk = var->rank[b * N + cc];
        const int st0 = (int)(r % rk) + var->lo[b * N + cc] - 1;
        r /= rk;
        over |= st0 > 255;
        cst[(v0 + e) * N + cc] = (uint8_t)st0;
      }
    }
  }
  if (over) {
    ctx->err = "fcg_tpsci_pt2_sparse: state index above 256";
    return FCG_EOVERFLOW;
  }
  if (!unboxed)
    for (int64_t f = 0; f < n_fock; f++) fstart[f + 1] += fstart[f];
  // every state inside its sector's basis; every term's Γ blocks shaped as
  // its integrals (the kernels index both without further checks)
  int bad_sector = 0, bad_state = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : bad_sector, bad_state)
  for (int64_t f = 0; f < n_fock; f++) {
    for (int cc = 0; cc < N; cc++) {
      const int na = var->fock[(f * N + cc) * 2], nb = var->fock[(f * N + cc) * 2 + 1];
      if (na < 0 || nb < 0 || na >= S || nb >= S) {
        bad_sector = 1;
        continue;
      }
      const int m = o->nstates[(int64_t)cc * S * S + na * S + nb];
      for (int64_t j = fstart[f]; j < fstart[f + 1]; j++)
        bad_state |= cst[j * N + cc] >= m || cst[j * N + cc] >= 255;
    }
  }
  if (bad_sector) {
    ctx->err = "fcg_tpsci_pt2_sparse: Fock sector outside the clusters";
    return FCG_EINVAL;
  }
  if (bad_state) {
    ctx->err = "fcg_tpsci_pt2_sparse: state index outside its sector's basis";
    return FCG_EINVAL;
  }
  {
    const int64_t SS = (int64_t)S * S;
    std::vector<uint8_t> seen((size_t)N * o->n_ops, 0);
    for (int t = 0; t < h->n_terms; t++)
      for (int q = 0; q < h->nbody[t]; q++) {
        const int cl = h->clusters[4 * t + q], op = h->ops[4 * t + q];
        if (cl < 0 || cl >= N || op < 0 || op >= o->n_ops) {
          ctx->err = "fcg_tpsci_pt2_sparse: term cluster or operator out of range";
          return FCG_EINVAL;
        }
        if (seen[(size_t)cl * o->n_ops + op]) continue;
        seen[(size_t)cl * o->n_ops + op] = 1;
        const int want = h->nbody[t] == 1 ? 1 : h->dims[4 * t + q];
        for (int64_t e = 0; e < SS * SS; e++) {
          const int blk = o->gblock[((int64_t)cl * o->n_ops + op) * SS * SS + e];
          if (blk >= 0 && o->gb_dims[3 * blk] != want) {
            ctx->err = "fcg_tpsci_pt2_sparse: a Γ block's leading dimension differs from its "
                       "term's integrals";
            return FCG_EINVAL;
          }
        }
      }
  }
  // key widths: sector codes and states of every cluster must fit 64 bits each
  int max_state = 1;
  for (int v : o->nstates) max_state = std::max(max_state, v);
  const int fbits = bits_for((int64_t)S * S), sbits = bits_for(max_state);
  if (fbits * N > 64 || sbits * N > 64) {
    ctx->err = "fcg_tpsci_pt2_sparse: configuration keys need more than 128 bits";
    return FCG_EOVERFLOW;
  }  if (open) {
    open->N = N;
    open->R = R;
    open->S = S;
    open->fbits = fbits;
    open->sbits = sbits;
  }

  // screening thresholds per ket (the reference's thresh / max_r |c|), and
  // each Fock block's kets ordered by max_r |c| descending (ties by position)
  std::vector<double> cabs(std::max<int64_t>(n_cfg, 1), 0.0), thr(std::max<int64_t>(n_cfg, 1));
  std::vector<double> cmaxF(std::max<int64_t>(n_fock, 1), 0.0);
  std::vector<int32_t> kord(std::max<int64_t>(n_cfg, 1));
  std::vector<double> csort(std::max<int64_t>(n_cfg, 1));
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < n_cfg; j++) {
    double m = 0.0;
    for (int r = 0; r < R; r++) m = std::max(m, std::fabs(c[(int64_t)r * n_cfg + j]));
    cabs[j] = m;
    thr[j] = m > 0.0 ? thresh / m : INFINITY;
  }
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t f = 0; f < n_fock; f++) {
    const int64_t j0 = fstart[f], j1 = fstart[f + 1];
    std::iota(kord.begin() + j0, kord.begin() + j1, (int32_t)j0);
    std::sort(kord.begin() + j0, kord.begin() + j1, [&](int32_t x, int32_t y) {
      return cabs[x] > cabs[y] || (cabs[x] == cabs[y] && x < y);
    });
    for (int64_t j = j0; j < j1; j++) csort[j] = cabs[kord[j]];
    cmaxF[f] = j1 > j0 ? csort[j0] : 0.0;
  }
  // variational keys, sorted: the Fock blocks by their sector codes, each
  // block's states sorted in parallel (blocks have distinct sector codes)
  std::vector<uint64_t> fkey(std::max<int64_t>(n_fock, 1));
  for (int64_t f = 0; f < n_fock; f++) {
    uint64_t hi = 0;
    for (int cc = 0; cc < N; cc++)
      hi |= (uint64_t)(var->fock[(f * N + cc) * 2] * S + var->fock[(f * N + cc) * 2 + 1])
            << (fbits * cc);
    fkey[f] = hi;
  }
  std::vector<int64_t> forder(std::max<int64_t>(n_fock, 1)), vstart(n_fock + 1, 0);
  std::iota(forder.begin(), forder.begin() + n_fock, (int64_t)0);
  std::sort(forder.begin(), forder.begin() + n_fock,
            [&](int64_t x, int64_t y) { return fkey[x] < fkey[y]; });
  for (int64_t k = 0; k < n_fock; k++)
    vstart[k + 1] = vstart[k] + (fstart[forder[k] + 1] - fstart[forder[k]]);
  std::vector<Key128> vk(std::max<int64_t>(n_cfg, 1));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t k = 0; k < n_fock; k++) {
    const int64_t f = forder[k];
    Key128 *out = vk.data() + vstart[k];
    for (int64_t j = fstart[f]; j < fstart[f + 1]; j++) {
      uint64_t lo = 0;
      for (int cc = 0; cc < N; cc++) lo |= (uint64_t)cst[j * N + cc] << (sbits * cc);
      out[j - fstart[f]] = Key128{fkey[f], lo};
    }
    std::sort(out, out + (fstart[f + 1] - fstart[f]),
              [](const Key128 &x, const Key128 &y) { return x.lo < y.lo; });
  }
  // term table
  std::vector<PTerm> terms(h->n_terms);
  for (int g = 0; g < h->n_tc; g++)
    for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) {
      PTerm &T = terms[t];
      T = PTerm{};
      T.nbody = h->nbody[t];
      T.group = g;
      T.ioff = h->ioff[t];
      for (int q = 0; q < 4; q++) {
        T.cl[q] = h->clusters[4 * t + q];
        T.op[q] = h->ops[4 * t + q];
        T.dims[q] = h->dims[4 * t + q];
        if (q < T.nbody && h->parity[4 * t + q] == 1) T.odd |= 1 << q;
      }
    }
  // per-Γ-block maxima
  const int64_t n_blocks = (int64_t)o->gb_off.size();
  std::vector<int64_t> bmax_off(n_blocks + 1, 0), cmx_off(n_blocks + 1, 0);
  for (int64_t b = 0; b < n_blocks; b++) {
    bmax_off[b + 1] = bmax_off[b] + o->gb_dims[3 * b];
    cmx_off[b + 1] = cmx_off[b] + (int64_t)o->gb_dims[3 * b] * o->gb_dims[3 * b + 2];
  }
  vlog("host preparation", t0);
  Buf<int64_t> d_fstart(ctx), d_bmax_off(ctx), d_cmx_off(ctx), d_item_count(ctx), d_item_ptr(ctx);
  Buf<int16_t> d_fock(ctx), d_tc(ctx);
  Buf<uint8_t> d_cst(ctx);
  Buf<int32_t> d_kord(ctx);
  Buf<double> d_csort(ctx), d_thr(ctx), d_cmaxF(ctx), d_c(ctx), d_bmax(ctx), d_cmx(ctx), d_e0(ctx);
  Buf<PTerm> d_terms(ctx);
  Buf<int32_t> d_tcs(ctx), d_gblock(ctx);
  Buf<Key128> d_vkey(ctx);
  Buf<unsigned long long> d_stats(ctx);
  int rc = d_fstart.upload(fstart.data(), n_fock + 1);
  if (!rc) rc = d_fock.upload(var->fock.data(), (int64_t)var->fock.size());
  if (!rc) rc = d_cst.upload(cst.data(), n_cfg * N);
  if (!rc) rc = d_kord.upload(kord.data(), n_cfg);
  if (!rc) rc = d_csort.upload(csort.data(), n_cfg);
  if (!rc) rc = d_thr.upload(thr.data(), n_cfg);
  if (!rc) rc = d_cmaxF.upload(cmaxF.data(), n_fock);
  if (!rc) rc = d_c.upload(c, (int64_t)R * n_cfg);
  if (!rc) rc = d_e0.upload(e0, R);
  if (!rc) rc = d_terms.upload(terms.data(), (int64_t)terms.size());
  if (!rc) rc = d_tcs.upload(h->tc_start.data(), (int64_t)h->tc_start.size());
  if (!rc) rc = d_tc.upload(h->tc.data(), (int64_t)h->tc.size());
  if (!rc) rc = d_gblock.upload(o->gblock.data(), (int64_t)o->gblock.size());
  if (!rc) rc = d_bmax_off.upload(bmax_off.data(), n_blocks + 1);
  if (!rc) rc = d_cmx_off.upload(cmx_off.data(), n_blocks + 1);
  if (!rc) rc = d_vkey.upload(vk.data(), n_cfg);
  std::vector<int32_t> nstates(o->nstates.begin(), o->nstates.end());
  Buf<int32_t> d_nstates(ctx);
  if (!rc) rc = d_nstates.upload(nstates.data(), (int64_t)nstates.size());
  if (rc) return rc;
  if (!d_bmax.alloc(bmax_off[n_blocks]) || !d_cmx.alloc(cmx_off[n_blocks]) ||
      !d_item_count.alloc(n_fock) || !d_stats.alloc(4)) {
    ctx->err = "PT2 workspace allocation failed";
    return FCG_ENOMEM;
  }
  FCG_HIP(hipMemsetAsync(d_stats.p, 0, 4 * sizeof(unsigned long long), s), ctx);
  if (n_blocks > 0)
    hipLaunchKernelGGL(k_gamma_max, dim3((unsigned)n_blocks), dim3(256), 0, s, n_blocks,
                       o->d_gb_off, o->d_gb_dims, o->d_gdata, d_bmax_off.p, d_bmax.p,
                       d_cmx_off.p, d_cmx.p);
  Pt2Args a{};
  a.N = N;
  a.S = S;
  a.R = R;
  a.nbody = nbody;
  a.fbits = fbits;
  a.sbits = sbits;
  a.thresh = thresh;
  a.n_fock = n_fock;
  a.n_cfg = n_cfg;
  a.fstart = d_fstart.p;
  a.fock = d_fock.p;
  a.cst = d_cst.p;
  a.kord = d_kord.p;
  a.csort = d_csort.p;
  a.thr = d_thr.p;
  a.cmaxF = d_cmaxF.p;
  a.c = d_c.p;
  a.terms = d_terms.p;
  a.tc_start = d_tcs.p;
  a.tc = d_tc.p;
  a.n_tc = h->n_tc;
  a.ints = h->d_ints;
  a.gblock = d_gblock.p;
  a.n_ops = o->n_ops;
  a.gb_off = o->d_gb_off;
  a.gb_dims = o->d_gb_dims;
  a.gdata = o->d_gdata;
  a.nstates = d_nstates.p;
  a.bmax_off = d_bmax_off.p;
  a.bmax = d_bmax.p;
  a.cmx_off = d_cmx_off.p;
  a.cmx = d_cmx.p;
  a.item_count = d_item_count.p;
  a.stats = d_stats.p;
  // items: count, prefix, fill
  for (int64_t f0 = 0; f0 < n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    hipLaunchKernelGGL(k_pt2_items<false>, dim3((unsigned)std::min(FB_CHUNK, n_fock - f0)),
                       dim3(PT), 0, s, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<int64_t> icnt(std::max<int64_t>(n_fock, 1));
  FCG_HIP(hipMemcpyAsync(icnt.data(), d_item_count.p, sizeof(int64_t) * n_fock,
                         hipMemcpyDeviceToHost, s), ctx);
  FCG_HIP(hipStreamSynchronize(s), ctx);
  std::vector<int64_t> iptr(n_fock + 1, 0);
  for (int64_t f = 0; f < n_fock; f++) iptr[f + 1] = iptr[f] + icnt[f];
  const int64_t n_items = iptr[n_fock];
  if (n_items >= (int64_t(1) << 32)) {
    ctx->err = "fcg_tpsci_pt2_sparse: more than 2^32 (ket Fock block, term) items";
    return FCG_EOVERFLOW;
  }
  vlog("items counted", t0, (double)n_items);
  Buf<int32_t> d_item_f(ctx), d_item_t(ctx), d_nsurv(ctx);
  Buf<double> d_item_b(ctx);
  Buf<uint64_t> d_item_key(ctx);
  if ((rc = d_item_ptr.upload(iptr.data(), n_fock + 1))) return rc;
  if (!d_item_f.alloc(n_items) || !d_item_t.alloc(n_items) || !d_item_b.alloc(n_items) ||
      !d_item_key.alloc(n_items) || !d_nsurv.alloc(n_items)) {
    ctx->err = "PT2 item list allocation failed";
    return FCG_ENOMEM;
  }
  a.item_ptr = d_item_ptr.p;
  a.item_f = d_item_f.p;
  a.item_t = d_item_t.p;
  a.item_b = d_item_b.p;
  a.item_key = d_item_key.p;
  for (int64_t f0 = 0; f0 < n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    hipLaunchKernelGGL(k_pt2_items<true>, dim3((unsigned)std::min(FB_CHUNK, n_fock - f0)),
                       dim3(PT), 0, s, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  // screening: kept kets per item
  a.items_f = d_item_f.p;
  a.items_t = d_item_t.p;
  a.items_b = d_item_b.p;
  a.n_items = n_items;
  a.nsurv = d_nsurv.p;
  for (int64_t i0 = 0; i0 < n_items; i0 += ITEM_CHUNK) {
    a.item_base = i0;
    hipLaunchKernelGGL(k_pt2_screen, dim3((unsigned)std::min(ITEM_CHUNK, n_items - i0)),
                       dim3(SW), 0, s, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<int32_t> nsurv(std::max<int64_t>(n_items, 1));
  FCG_HIP(hipMemcpyAsync(nsurv.data(), d_nsurv.p, sizeof(int32_t) * n_items,
                         hipMemcpyDeviceToHost, s), ctx);
  FCG_HIP(hipStreamSynchronize(s), ctx);
  std::vector<uint32_t> kept;
  int64_t kept_kets = 0;
  for (int64_t i = 0; i < n_items; i++)
    if (nsurv[i] > 0) {
      kept.push_back((uint32_t)i);
      kept_kets += nsurv[i];
    }
  const int64_t n_slots = (int64_t)kept.size();
  std::vector<int32_t>().swap(nsurv);
  vlog("items screened, kept", t0, (double)n_slots);
  vlog("... (item, ket) pairs kept", t0, (double)kept_kets);
  // slots: the kept items ordered by target Fock block (stable)
  Buf<uint32_t> d_kept(ctx), d_slot(ctx);
  Buf<uint64_t> d_kkey(ctx), d_skkey(ctx);
  if ((rc = d_kept.upload(kept.data(), n_slots))) return rc;
  if (!d_slot.alloc(n_slots) || !d_kkey.alloc(n_slots) || !d_skkey.alloc(n_slots)) {
    ctx->err = "PT2 slot allocation failed";
    return FCG_ENOMEM;
  }
  std::vector<uint64_t> slot_key(std::max<int64_t>(n_slots, 1));
  Buf<char> d_tmp(ctx);
  if (n_slots > 0) {
    hipLaunchKernelGGL(k_gather_keys, dim3(1024), dim3(256), 0, s, n_slots, d_kept.p,
                       d_item_key.p, d_kkey.p);
    size_t tb = 0;
    FCG_HIP(rocprim::radix_sort_pairs(nullptr, tb, d_kkey.p, d_skkey.p, d_kept.p, d_slot.p,
                                      (size_t)n_slots, 0u, 64u, s),
            ctx);
    if (!d_tmp.alloc((int64_t)tb)) {
      ctx->err = "PT2 slot sort workspace allocation failed";
      return FCG_ENOMEM;
    }
    FCG_HIP(rocprim::radix_sort_pairs(d_tmp.p, tb, d_kkey.p, d_skkey.p, d_kept.p, d_slot.p,
                                      (size_t)n_slots, 0u, 64u, s),
            ctx);
    FCG_HIP(hipMemcpyAsync(slot_key.data(), d_skkey.p, sizeof(uint64_t) * n_slots,
                           hipMemcpyDeviceToHost, s), ctx);
  }
  // contraction, count pass
  Buf<int64_t> d_ecount(ctx), d_eptr(ctx);
  if (!d_ecount.alloc(n_slots)) {
    ctx->err = "PT2 emission counts allocation failed";
    return FCG_ENOMEM;
  }
  a.slot_item = d_slot.p;
  a.n_slots = n_slots;
  a.emit_count = d_ecount.p;
  const auto te = std::chrono::steady_clock::now();
  for (int64_t i0 = 0; i0 < n_slots; i0 += SLOT_CHUNK) {
    a.slot_base = i0;
    if (fp32)
      hipLaunchKernelGGL((k_pt2_emit<false, float>),
                         dim3((unsigned)std::min(SLOT_CHUNK, n_slots - i0)), dim3(PT), 0, s, a);
    else
      hipLaunchKernelGGL((k_pt2_emit<false, double>),
                         dim3((unsigned)std::min(SLOT_CHUNK, n_slots - i0)), dim3(PT), 0, s, a);
    if (verbose()) {
      FCG_HIP(hipStreamSynchronize(s), ctx);
      vlog("contraction count chunk", t0, (double)std::min(i0 + SLOT_CHUNK, n_slots));
    }
  }
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<int64_t> ecnt(std::max<int64_t>(n_slots, 1));
  FCG_HIP(hipMemcpyAsync(ecnt.data(), d_ecount.p, sizeof(int64_t) * n_slots,
                         hipMemcpyDeviceToHost, s), ctx);
  unsigned long long stc[4];
  FCG_HIP(hipMemcpyAsync(stc, d_stats.p, sizeof(stc), hipMemcpyDeviceToHost, s), ctx);
  FCG_HIP(hipStreamSynchronize(s), ctx);
  double emit_s = seconds_since(te);
  if (stc[2]) {
    ctx->err = "fcg_tpsci_pt2_sparse: a term's shape exceeds the staged contraction";
    return FCG_EOVERFLOW;
  }
  std::vector<int64_t> eptr(n_slots + 1, 0);
  for (int64_t i = 0; i < n_slots; i++) eptr[i + 1] = eptr[i] + ecnt[i];
  const int64_t E = eptr[n_slots];
  vlog("contributions counted", t0, (double)E);
  // chunks of whole target blocks, at most CHUNK_E contributions unless one
  // block alone holds more
  std::vector<int64_t> cut{0};
  {
    const int64_t limit = chunk_limit();
    int64_t i = 0;
    while (i < n_slots) {
      int64_t j = i;
      while (j < n_slots) {
        int64_t k = j + 1;  // end of the target block starting at j
        while (k < n_slots && slot_key[k] == slot_key[j]) k++;
        if (j > i && eptr[k] - eptr[i] > limit) break;
        j = k;
      }
      cut.push_back(j);
      i = j;
    }
  }
  const int n_chunks = (int)cut.size() - 1;
  int64_t maxE = 0;
  for (int ch = 0; ch < n_chunks; ch++) maxE = std::max(maxE, eptr[cut[ch + 1]] - eptr[cut[ch]]);
  if ((rc = d_eptr.upload(eptr.data(), n_slots + 1))) return rc;
  // each chunk's targets, and every slot's index among its chunk's targets
  // (the sort key holds that index, not the Fock sectors)
  std::vector<int64_t> tstart{0};
  std::vector<uint64_t> tkeys;
  std::vector<uint32_t> tloc(std::max<int64_t>(n_slots, 1));
  for (int ch = 0; ch < n_chunks; ch++) {
    uint32_t n = 0;
    for (int64_t i = cut[ch]; i < cut[ch + 1]; i++) {
      if (i == cut[ch] || slot_key[i] != slot_key[i - 1]) {
        tkeys.push_back(slot_key[i]);
        n++;
      }
      tloc[i] = n - 1;
    }
    tstart.push_back((int64_t)tkeys.size());
  }
  // each target's Fock block among the sorted variational keys (empty when
  // the space has no configuration there)
  std::vector<int64_t> tvar(2 * std::max<size_t>(tkeys.size(), 1), 0);
  {
    std::vector<uint64_t> sk(std::max<int64_t>(n_fock, 1));
    for (int64_t k = 0; k < n_fock; k++) sk[k] = fkey[forder[k]];
    for (size_t t = 0; t < tkeys.size(); t++) {
      const auto it = std::lower_bound(sk.begin(), sk.begin() + n_fock, tkeys[t]);
      const int64_t k = it - sk.begin();
      if (k < n_fock && sk[k] == tkeys[t]) {
        tvar[2 * t] = vstart[k];
        tvar[2 * t + 1] = vstart[k + 1];
      }
    }
  }
  Buf<uint32_t> d_tloc(ctx);
  Buf<uint64_t> d_tkeys(ctx);
  Buf<int64_t> d_tvar(ctx);
  if ((rc = d_tloc.upload(tloc.data(), n_slots))) return rc;
  if ((rc = d_tkeys.upload(tkeys.data(), (int64_t)tkeys.size()))) return rc;
  if ((rc = d_tvar.upload(tvar.data(), (int64_t)tvar.size()))) return rc;
  const int sN = sbits * N;
  a.slot_tloc = d_tloc.p;
  a.sN = sN;
  // one root: values sorted with their keys; several: an index permutation
  const bool direct = R == 1;
  Buf<Key128> d_ekey(ctx), d_skey(ctx);
  Buf<double> d_eval(ctx), d_sval(ctx);
  Buf<uint64_t> d_idx(ctx), d_perm(ctx), d_segid(ctx);
  Buf<uint32_t> d_head(ctx);
  Buf<int64_t> d_seg(ctx);
  if (!d_ekey.alloc(maxE) || !d_eval.alloc(maxE * R) || !d_skey.alloc(maxE) ||
      !d_idx.alloc(direct ? 0 : maxE) || !d_perm.alloc(direct ? 0 : maxE) ||
      !d_sval.alloc(direct ? maxE : 0) || !d_head.alloc(maxE) || !d_segid.alloc(maxE) ||
      !d_seg.alloc(maxE + 1)) {
    ctx->err = "PT2 emission buffers (" + std::to_string(maxE) + " contributions) do not fit";
    return FCG_ENOMEM;
  }
  size_t sort_bytes = 0, scan_bytes = 0;
  if (direct)
    FCG_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, d_ekey.p, d_skey.p, d_eval.p,
                                      d_sval.p, (size_t)std::max<int64_t>(maxE, 1),
                                      KeyDecomposer{}, 0u, 128u, s),
            ctx);
  else
    FCG_HIP(rocprim::radix_sort_pairs(nullptr, sort_bytes, d_ekey.p, d_skey.p, d_idx.p,
                                      d_perm.p, (size_t)std::max<int64_t>(maxE, 1),
                                      KeyDecomposer{}, 0u, 128u, s),
            ctx);
  FCG_HIP(rocprim::exclusive_scan(nullptr, scan_bytes, d_head.p, d_segid.p, (uint64_t)0,
                                  (size_t)std::max<int64_t>(maxE, 1), rocprim::plus<uint64_t>(),
                                  s),
          ctx);
  if (!d_tmp.alloc((int64_t)std::max(sort_bytes, scan_bytes))) {
    ctx->err = "PT2 sort workspace allocation failed";
    return FCG_ENOMEM;
  }
  Buf<double> d_part(ctx);
  Buf<unsigned long long> d_nout(ctx);
  if (!d_part.alloc((int64_t)SUM_BLOCKS * R * 2) || !d_nout.alloc(1)) {
    ctx->err = "PT2 sum buffers allocation failed";
    return FCG_ENOMEM;
  }
  FCG_HIP(hipMemsetAsync(d_nout.p, 0, sizeof(unsigned long long), s), ctx);
  a.emit_ptr = d_eptr.p;
  a.ekey = d_ekey.p;
  a.eval = d_eval.p;
  SumArgs sa{};
  sa.N = N;
  sa.S = S;
  sa.R = R;
  sa.n_ops = o->n_ops;
  sa.h0_op = h0_op;
  sa.fbits = fbits;
  sa.sbits = sbits;
  sa.sN = sN;
  sa.seg_start = d_seg.p;
  sa.perm = direct ? nullptr : d_perm.p;
  sa.sval = direct ? d_sval.p : nullptr;
  sa.skey = d_skey.p;
  sa.eval = d_eval.p;
  sa.vkey = d_vkey.p;
  sa.n_v = n_cfg;
  sa.gblock = d_gblock.p;
  sa.gb_off = o->d_gb_off;
  sa.gb_dims = o->d_gb_dims;
  sa.gdata = o->d_gdata;
  Buf<double> d_h0tab(ctx);
  if (!d_h0tab.alloc(((int64_t)N * S * S) << sbits)) {
    ctx->err = "PT2 H0 diagonal table allocation failed";
    return FCG_ENOMEM;
  }
  hipLaunchKernelGGL(k_h0_table, dim3(256), dim3(256), 0, s, sa, d_h0tab.p);
  sa.h0tab = d_h0tab.p;
  sa.e0 = d_e0.p;
  sa.part = d_part.p;
  sa.n_out = d_nout.p;
  std::vector<double> sum(R, 0.0), comp(R, 0.0);
  std::vector<double> part((size_t)SUM_BLOCKS * R * 2);
  int64_t n_seg_total = 0;
  double sort_s = 0.0;
  for (int ch = 0; ch < n_chunks; ch++) {
    const int64_t s0 = cut[ch], s1 = cut[ch + 1];
    const int64_t Ec = eptr[s1] - eptr[s0];
    if (Ec == 0) continue;
    const auto tf = std::chrono::steady_clock::now();
    a.emit_base = eptr[s0];
    a.n_slots = s1;
    for (int64_t i0 = s0; i0 < s1; i0 += SLOT_CHUNK) {
      a.slot_base = i0;
      if (fp32)
        hipLaunchKernelGGL((k_pt2_emit<true, float>), dim3((unsigned)std::min(SLOT_CHUNK, s1 - i0)),
                           dim3(PT), 0, s, a);
      else
        hipLaunchKernelGGL((k_pt2_emit<true, double>),
                           dim3((unsigned)std::min(SLOT_CHUNK, s1 - i0)), dim3(PT), 0, s, a);
    }
    FCG_HIP(hipGetLastError(), ctx);
    if (verbose()) {
      FCG_HIP(hipStreamSynchronize(s), ctx);
      vlog("contributions written", t0, (double)Ec);
    }
    FCG_HIP(hipStreamSynchronize(s), ctx);
    emit_s += seconds_since(tf);
    const auto ts = std::chrono::steady_clock::now();
    // stable sort of the chunk's contributions by their state bits alone: the
    // emission wrote them grouped by target Fock block (slots in target
    // order), so a stable sort keeps every (target, states) group contiguous
    // and the target bits above sN need no passes of their own
    const unsigned end_bit = (unsigned)sN;
    sa.tkeys = d_tkeys.p + tstart[ch];
    sa.tvar = d_tvar.p + 2 * tstart[ch];
    size_t tb = sort_bytes;
    if (direct) {
      FCG_HIP(rocprim::radix_sort_pairs(d_tmp.p, tb, d_ekey.p, d_skey.p, d_eval.p, d_sval.p,
                                        (size_t)Ec, KeyDecomposer{}, 0u, end_bit, s),
              ctx);
    } else {
      hipLaunchKernelGGL(k_iota, dim3(1024), dim3(256), 0, s, Ec, d_idx.p);
      FCG_HIP(rocprim::radix_sort_pairs(d_tmp.p, tb, d_ekey.p, d_skey.p, d_idx.p, d_perm.p,
                                        (size_t)Ec, KeyDecomposer{}, 0u, end_bit, s),
              ctx);
    }
    // segments of equal keys: heads, ids by scan, starts
    hipLaunchKernelGGL(k_heads, dim3(2048), dim3(256), 0, s, Ec, d_skey.p, d_head.p);
    size_t sb = scan_bytes;
    FCG_HIP(rocprim::exclusive_scan(d_tmp.p, sb, d_head.p, d_segid.p, (uint64_t)0, (size_t)Ec,
                                    rocprim::plus<uint64_t>(), s),
            ctx);
    uint64_t last_id = 0;
    uint32_t last_head = 0;
    FCG_HIP(hipMemcpyAsync(&last_id, d_segid.p + Ec - 1, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           s), ctx);
    FCG_HIP(hipMemcpyAsync(&last_head, d_head.p + Ec - 1, sizeof(uint32_t),
                           hipMemcpyDeviceToHost, s), ctx);
    FCG_HIP(hipStreamSynchronize(s), ctx);
    const int64_t n_seg = (int64_t)last_id + last_head;
    hipLaunchKernelGGL(k_seg_start, dim3(2048), dim3(256), 0, s, Ec, d_head.p, d_segid.p,
                       d_seg.p);
    FCG_HIP(hipMemcpyAsync(d_seg.p + n_seg, &Ec, sizeof(int64_t), hipMemcpyHostToDevice, s), ctx);
    sa.n_seg = n_seg;
    if (open) {
      // open sigma: every configuration of the chunk with its key and sigma
      Buf<uint64_t> d_khi(ctx), d_klo(ctx);
      Buf<double> d_sig(ctx);
      if (!d_khi.alloc(std::max<int64_t>(n_seg, 1)) || !d_klo.alloc(std::max<int64_t>(n_seg, 1)) ||
          !d_sig.alloc(std::max<int64_t>(n_seg * R, 1))) {
        ctx->err = "open sigma output allocation failed";
        return FCG_ENOMEM;
      }
      hipLaunchKernelGGL(k_open_out, dim3(SUM_BLOCKS), dim3(256), 0, s, sa, d_khi.p, d_klo.p,
                         d_sig.p);
      FCG_HIP(hipGetLastError(), ctx);
      const size_t at = open->klo.size();
      open->khi.resize(at + n_seg);
      open->klo.resize(at + n_seg);
      open->sig.resize((at + n_seg) * R);
      FCG_HIP(hipMemcpyAsync(open->khi.data() + at, d_khi.p, sizeof(uint64_t) * n_seg,
                             hipMemcpyDeviceToHost, s), ctx);
      FCG_HIP(hipMemcpyAsync(open->klo.data() + at, d_klo.p, sizeof(uint64_t) * n_seg,
                             hipMemcpyDeviceToHost, s), ctx);
      FCG_HIP(hipMemcpyAsync(open->sig.data() + at * R, d_sig.p, sizeof(double) * n_seg * R,
                             hipMemcpyDeviceToHost, s), ctx);
      FCG_HIP(hipStreamSynchronize(s), ctx);
      sort_s += seconds_since(ts);
      n_seg_total += n_seg;
      continue;
    }
    hipLaunchKernelGGL(k_pt2_seg_sum, dim3(SUM_BLOCKS), dim3(256), 0, s, sa);
    FCG_HIP(hipGetLastError(), ctx);
    FCG_HIP(hipMemcpyAsync(part.data(), d_part.p, sizeof(double) * part.size(),
                           hipMemcpyDeviceToHost, s), ctx);
    FCG_HIP(hipStreamSynchronize(s), ctx);
    sort_s += seconds_since(ts);
    n_seg_total += n_seg;
    for (int r = 0; r < R; r++)
      for (int blk = 0; blk < SUM_BLOCKS; blk++) {
        const double v = part[((size_t)blk * R + r) * 2];
        const double t = sum[r] + v;
        comp[r] += std::fabs(sum[r]) >= std::fabs(v) ? (sum[r] - t) + v : (v - t) + sum[r];
        sum[r] = t;
        comp[r] += part[((size_t)blk * R + r) * 2 + 1];
      }
    if (verbose()) vlog("chunk summed, configurations", t0, (double)n_seg);
  }
  unsigned long long nout = 0;
  FCG_HIP(hipMemcpyAsync(&nout, d_nout.p, sizeof(nout), hipMemcpyDeviceToHost, s), ctx);
  FCG_HIP(hipStreamSynchronize(s), ctx);
  for (int r = 0; r < R; r++) e2[r] = sum[r] + comp[r];
  if (st) {
    st->n_items = n_items;
    st->n_groups = (int64_t)stc[0];
    st->n_contributions = E;
    st->n_configs = n_seg_total;
    st->n_first_order = (int64_t)nout;
    st->chain_flops = (double)stc[1];
    st->seconds = seconds_since(t0);
    st->n_items_kept = n_slots;
    st->n_kets_kept = kept_kets;
    st->n_chunks = n_chunks;
    st->contraction_seconds = emit_s;
    st->sort_seconds = sort_s;
  }
  return FCG_OK;
}
