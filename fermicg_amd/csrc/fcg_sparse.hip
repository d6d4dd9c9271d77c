// Device-side join and Hamiltonian build for spaces whose boxes are all single
// configurations (sparse TPSCI spaces: the rank-1 configuration sets of large
// cluster counts).  The host join of fcg_plan_host.cpp enumerates (box, box,
// term) tasks explicitly, which at ~1e10 connected triples does not fit in
// host memory; here the host only lists connected Fock pairs and the GPU does
// the rest.
//
// Per bra Fock block (one workgroup), per connected ket Fock block
// F_ket = F_bra - TransferConfig (src/tpsci_outer.jl:479-481), every
// (bra configuration I, ket configuration J) pair is tested against every
// term of the TransferConfig group with the reference's check_term rule
// (src/type_ClusteredTerm.jl:92-176: states equal on the clusters the term
// does not touch).  Two passes:
//   count  the structurally nonzero (I, J) per row I;
//   fill   the element H[I, J] = sum over passing terms of the reference's
//          contract_matrix_element (src/tpsci_inner.jl:10-216, _contract
//          :270-328) with the sign of compute_terms_state_sign
//          (src/type_ClusteredTerm.jl:181-197), written at its row position.
// Rows are the panels of the sigma apply with db = 1 (columns in ket Fock
// pair order, then ket configuration order), so k_apply, the diagonal and
// Davidson run unchanged.  Positions come from per-row bitmaps and popcounts,
// so the layout and every value are deterministic.
#include <algorithm>
#include <chrono>
#include <numeric>

#include "fcg_internal.h"

namespace {

struct SparseTerm {
  int32_t nbody;
  int32_t cl[4];
  int32_t op[4];
  int32_t dims[4];
  int32_t odd;   // bit a: operator string a has odd fermion parity
  int32_t pad;
  int64_t ioff;
  uint64_t inact[2];  // 0xFF per cluster byte the term does not touch
};

struct SparseArgs {
  int N, S, n_ops, nbody;
  const int64_t *fstart;  // [n_fock+1] configurations (rows) of each Fock block
  const int16_t *fock;    // [n_fock][N][2]
  const int16_t *cfg;     // [n_cfg][N] 1-based states
  const int64_t *fp_ptr;  // [n_fock+1] connected ket Fock blocks of each bra block
  const int32_t *fp_fk, *fp_g;
  const int32_t *tc_start;
  const SparseTerm *terms;
  const double *ints;
  const int32_t *gblock;  // [N][n_ops][S*S][S*S]
  const int64_t *gb_off;
  const int32_t *gb_dims;
  const double *gdata;
  int64_t fb_base;
  int32_t *row_nnz;                  // count pass
  unsigned long long *stats;         // [0] triples, [1] flops, [2] missing blocks
  const int64_t *row_start;          // fill pass
  // fill pass: each TransferConfig group's terms in stage-sized chunks
  // (chunk k of group g ends at term chunk_end[grp_chunk[g] + k]) and each
  // term's integral offset in the stage (-1 dropped by nbody, -2 too large
  // for the stage: read from global memory)
  const int32_t *grp_chunk, *chunk_end, *voff;
  double *K;
  int32_t *colidx;
  int64_t *diag_koff;
  int32_t *diag_col;
};

constexpr int ST = 256;  // threads per workgroup
constexpr int MB = FCG_SPARSE_MAX_BLOCK;

__device__ __forceinline__ void pack_states(const int16_t *c, int N, uint64_t (&p)[2]) {
  uint64_t w0 = 0, w1 = 0;
  for (int k = 0; k < N; k++) {
    const uint64_t v = (uint64_t)(uint16_t)c[k] & 0xFFu;
    if (k < 8) w0 |= v << (8 * k);
    else w1 |= v << (8 * (k - 8));
  }
  p[0] = w0;
  p[1] = w1;
}

__device__ __forceinline__ int state_of(const uint64_t *p, int c) {
  return (int)((p[c >> 3] >> (8 * (c & 7))) & 0xFFu);
}

// sum_p V[p0,p1,..] g0[p0] g1[p1] ...; the first Γ column (D0 values) is held
// in registers, V is read in order (from LDS when staged there)
template <int D0>
__device__ __forceinline__ double contract_v(const double *V, const double *const *g,
                                             const int32_t *d, int nbody) {
  double g0[D0];
#pragma unroll
  for (int k = 0; k < D0; k++) g0[k] = g[0][k];
  double acc = 0.0;
  if (nbody == 2) {
    for (int p1 = 0; p1 < d[1]; p1++) {
      const double *Vp = V + D0 * p1;
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D0; k++) s += Vp[k] * g0[k];
      acc += s * g[1][p1];
    }
  } else if (nbody == 3) {
    for (int p2 = 0; p2 < d[2]; p2++) {
      double s2 = 0.0;
      for (int p1 = 0; p1 < d[1]; p1++) {
        const double *Vp = V + D0 * (p1 + d[1] * p2);
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D0; k++) s += Vp[k] * g0[k];
        s2 += s * g[1][p1];
      }
      acc += s2 * g[2][p2];
    }
  } else {
    for (int p3 = 0; p3 < d[3]; p3++) {
      double s3 = 0.0;
      for (int p2 = 0; p2 < d[2]; p2++) {
        double s2 = 0.0;
        for (int p1 = 0; p1 < d[1]; p1++) {
          const double *Vp = V + D0 * (p1 + d[1] * (p2 + d[2] * p3));
          double s = 0.0;
#pragma unroll
          for (int k = 0; k < D0; k++) s += Vp[k] * g0[k];
          s2 += s * g[1][p1];
        }
        s3 += s2 * g[2][p2];
      }
      acc += s3 * g[3][p3];
    }
  }
  return acc;
}

constexpr int SV = 1536;   // staged integrals of a TransferConfig group (doubles)
constexpr int NZMAX = 2048;  // nonzeros of one Fock pair listed for the fill (more: dense walk)
constexpr int SVT = 64;    // staged terms per group

// per (Fock pair, term): where each active cluster's Γ block starts, its
// leading dimension and bra state count, and the sign; blk < 0 marks a
// missing block
struct TermPair {
  int64_t base[4];
  int32_t np[4], mb[4];
  int32_t ok;
  float sgn;
};

__device__ void term_pair(const SparseArgs &a, const SparseTerm &T, const int16_t *fb,
                          const int16_t *fk, TermPair *out) {
  const int SS = a.S * a.S;
  out->ok = 1;
  for (int q = 0; q < T.nbody; q++) {
    const int c = T.cl[q];
    const int cb = fb[2 * c] * a.S + fb[2 * c + 1], ck = fk[2 * c] * a.S + fk[2 * c + 1];
    const int blk = a.gblock[(((int64_t)c * a.n_ops + T.op[q]) * SS + cb) * SS + ck];
    out->base[q] = blk;  // block ids first: independent loads
  }
  for (int q = 0; q < T.nbody; q++)
    if (out->base[q] < 0) {
      out->ok = 0;
      return;
    }
  for (int q = 0; q < T.nbody; q++) {
    const int blk = (int)out->base[q];
    out->base[q] = a.gb_off[blk];
    out->np[q] = a.gb_dims[3 * blk];
    out->mb[q] = a.gb_dims[3 * blk + 1];
  }
  int par = 0;
  for (int q = 0; q < T.nbody; q++) {
    if (!((T.odd >> q) & 1)) continue;
    int ne = 0;
    for (int c = 0; c < T.cl[q]; c++) ne += fk[2 * c] + fk[2 * c + 1];
    par ^= ne & 1;
  }
  out->sgn = par ? -1.f : 1.f;
}

// sum_p V[p0,p1,p2,p3] g0[p0] g1[p1] g2[p2] g3[p3] for a fixed shape: every
// Γ column is loaded into registers once, V (from LDS) is read in order
template <int D0, int D1, int D2, int D3>
__device__ __forceinline__ double contract_fixed(const double *V, const double *const *g) {
  double a0[D0], a1[D1], a2[D2], a3[D3];
#pragma unroll
  for (int k = 0; k < D0; k++) a0[k] = g[0][k];
#pragma unroll
  for (int k = 0; k < D1; k++) a1[k] = g[1][k];
#pragma unroll
  for (int k = 0; k < D2; k++) a2[k] = D2 > 1 ? g[2][k] : 1.0;
#pragma unroll
  for (int k = 0; k < D3; k++) a3[k] = D3 > 1 ? g[3][k] : 1.0;
  double acc = 0.0;
#pragma unroll
  for (int p3 = 0; p3 < D3; p3++) {
    double s3 = 0.0;
#pragma unroll
    for (int p2 = 0; p2 < D2; p2++) {
      double s2 = 0.0;
#pragma unroll
      for (int p1 = 0; p1 < D1; p1++) {
        const double *Vp = V + D0 * (p1 + D1 * (p2 + D2 * p3));
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < D0; k++) s += Vp[k] * a0[k];
        s2 += s * a1[p1];
      }
      s3 += s2 * a2[p2];
    }
    acc += s3 * a3[p3];
  }
  return acc;
}

// any shape: column-major V over nbody dims
__device__ __forceinline__ double contract_any(const double *V, const double *const *g,
                                               const int32_t *d, int nbody) {
  const int d0 = d[0], d1 = d[1], d2 = nbody > 2 ? d[2] : 1, d3 = nbody > 3 ? d[3] : 1;
  double acc = 0.0;
  for (int p3 = 0; p3 < d3; p3++)
    for (int p2 = 0; p2 < d2; p2++)
      for (int p1 = 0; p1 < d1; p1++) {
        const double *Vp = V + d0 * (p1 + d1 * (p2 + d2 * p3));
        double s = 0.0;
        for (int p0 = 0; p0 < d0; p0++) s += Vp[p0] * g[0][p0];
        double f = g[1][p1];
        if (nbody > 2) f *= g[2][p2];
        if (nbody > 3) f *= g[3][p3];
        acc += s * f;
      }
  return acc;
}

// the shapes of 4-orbital clusters; 0 = no fixed shape
__device__ __forceinline__ int shape_id(const SparseTerm &T) {
  const int *d = T.dims;
  if (T.nbody == 4 && d[0] == 4 && d[1] == 4 && d[2] == 4 && d[3] == 4) return 1;
  if (T.nbody == 3) {
    if (d[0] == 16 && d[1] == 4 && d[2] == 4) return 2;
    if (d[0] == 4 && d[1] == 16 && d[2] == 4) return 3;
    if (d[0] == 4 && d[1] == 4 && d[2] == 16) return 4;
  }
  if (T.nbody == 2) {
    if (d[0] == 16 && d[1] == 16) return 5;
    if (d[0] == 4 && d[1] == 4) return 6;
  }
  return 0;
}

// sum_p V[p0,p1] g0[p0] g1[p1] with the second Γ column (D1 values) in
// registers (two-cluster terms whose first string has three operators)
template <int D1>
__device__ __forceinline__ double contract_v1(const double *V, const double *const *g, int d0) {
  double g1[D1];
#pragma unroll
  for (int k = 0; k < D1; k++) g1[k] = g[1][k];
  double acc = 0.0;
  for (int p0 = 0; p0 < d0; p0++) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D1; k++) s += V[p0 + d0 * k] * g1[k];
    acc += s * g[0][p0];
  }
  return acc;
}

// Launch shapes of the join and fill: workgroup threads, the largest Fock
// block (configurations), staged integrals (doubles) and terms per group
// chunk, and the longest listed nonzero list per Fock pair.  Spaces whose
// Fock blocks hold at most 64 configurations take one wavefront per bra
// block with small stages: a Fock pair has ~90 nonzeros, which leave most of
// a 256-thread workgroup idle, and the smaller LDS footprint (10 KB against
// 38 KB) fits four times as many workgroups per CU.
template <int ST_, int MB_, int SV_, int SVT_, int NZ_, int WPE_>
struct RowsCfg {
  static constexpr int ST = ST_, MB = MB_, SV = SV_, SVT = SVT_, NZ = NZ_, WPE = WPE_;
};
// WPE: the waves per SIMD asked of the register allocator (the small
// variant then fits 128 VGPRs without scratch: four waves per SIMD, which its
// 10 KB of LDS also allows)
using BigRows = RowsCfg<ST, MB, SV, SVT, NZMAX, 4>;
using SmallRows = RowsCfg<64, 64, 512, 16, 1024, 5>;

template <bool FILL, class C>
__global__ __launch_bounds__(C::ST, C::WPE) void k_sparse_rows(SparseArgs a) {
  constexpr int ST = C::ST, MB = C::MB, SV = C::SV, SVT = C::SVT, NZMAX = C::NZ;
  __shared__ uint64_t sPI[MB][2], sPJ[MB][2];
  __shared__ uint64_t sBits[MB][MB / 64];
  __shared__ int32_t sCur[MB];
  __shared__ int16_t sFb[2 * FCG_MAXN], sFk[2 * FCG_MAXN];
  __shared__ unsigned long long sTrip, sFlop, sMiss;
  __shared__ double sV[FILL ? SV : 1];
  __shared__ int32_t sVoff[FILL ? SVT : 1];
  __shared__ TermPair sTP[FILL ? SVT : 1];
  __shared__ uint16_t sNZ[FILL ? NZMAX : 1];  // the pair's nonzeros, (I << 8) | J
  __shared__ int32_t sWave[ST / 64];
  const int tid = threadIdx.x;
  const int N = a.N;
  const int64_t fb = a.fb_base + blockIdx.x;
  const int64_t b0 = a.fstart[fb];
  const int nb = (int)(a.fstart[fb + 1] - b0);
  for (int I = tid; I < nb; I += ST) {
    pack_states(a.cfg + (b0 + I) * N, N, sPI[I]);
    sCur[I] = 0;
  }
  if (tid < 2 * N) sFb[tid] = a.fock[fb * 2 * N + tid];
  if (tid == 0) sTrip = sFlop = sMiss = 0;
  unsigned long long trip = 0, flop = 0, miss = 0;
  __syncthreads();
  for (int64_t q = a.fp_ptr[fb]; q < a.fp_ptr[fb + 1]; q++) {
    const int64_t fk = a.fp_fk[q];
    const int g = a.fp_g[q];
    const int64_t k0 = a.fstart[fk];
    const int nk = (int)(a.fstart[fk + 1] - k0);
    for (int J = tid; J < nk; J += ST) pack_states(a.cfg + (k0 + J) * N, N, sPJ[J]);
    if (tid < 2 * N) sFk[tid] = a.fock[fk * 2 * N + tid];
    for (int e = tid; e < nb * (MB / 64); e += ST) sBits[e / (MB / 64)][e % (MB / 64)] = 0;
    __syncthreads();
    const int t0 = a.tc_start[g], t1 = a.tc_start[g + 1];
    // structural nonzeros: some term passes check_term
    for (int e = tid; e < nb * nk; e += ST) {
      const int I = e / nk, J = e % nk;
      bool nz = false;
      for (int t = t0; t < t1; t++) {
        const SparseTerm &T = a.terms[t];
        if (T.nbody > a.nbody) continue;
        if (((sPI[I][0] ^ sPJ[J][0]) & T.inact[0]) | ((sPI[I][1] ^ sPJ[J][1]) & T.inact[1]))
          continue;
        nz = true;
        if (!FILL) {
          trip++;
          uint64_t v = 1;
          if (T.nbody > 1)
            for (int k = 0; k < T.nbody; k++) v *= (uint64_t)T.dims[k];
          flop += 2 * v;
        }
      }
      if (nz) atomicOr(reinterpret_cast<unsigned long long *>(&sBits[I][J >> 6]),
                       1ull << (J & 63));
    }
    __syncthreads();
    if (FILL) {
      // list the pair's nonzeros (row-major, J ascending), so that the
      // element loops below give every thread a nonzero instead of walking
      // all nb x nk candidates; a pair with more than NZMAX walks them
      int n_list = 0;
      {
        int cnt = 0;
        if (tid < nb)
          for (int k = 0; k < MB / 64; k++) cnt += __popcll(sBits[tid][k]);
        const int lane = tid & 63, w = tid >> 6;
        int x = cnt;
        for (int off = 1; off < 64; off <<= 1) {
          const int y = __shfl_up(x, off, 64);
          if (lane >= off) x += y;
        }
        if (lane == 63) sWave[w] = x;
        __syncthreads();
        int pre = 0;
        for (int k = 0; k < ST / 64; k++) {
          if (k < w) pre += sWave[k];
          n_list += sWave[k];
        }
        if (n_list <= NZMAX && tid < nb) {
          int at = pre + x - cnt;
          for (int k = 0; k < MB / 64; k++) {
            uint64_t wb = sBits[tid][k];
            while (wb) {
              const int j = 64 * k + __builtin_ctzll(wb);
              wb &= wb - 1;
              sNZ[at++] = (uint16_t)((tid << 8) | j);
            }
          }
        }
        __syncthreads();
      }
      const bool listed = n_list <= NZMAX;
      const int n_walk = listed ? n_list : nb * nk;
      // the group's terms in chunks that fit the stage (integrals and
      // per-pair Γ block data in LDS); element values accumulate in K over
      // the chunks in term order
      for (int c0 = t0, first = 1, kc = a.grp_chunk[g]; c0 < t1; first = 0, kc++) {
        const int c1 = a.chunk_end[kc];
        for (int t = c0 + tid; t < c1; t += ST) {
          sVoff[t - c0] = a.voff[t];
          term_pair(a, a.terms[t], sFb, sFk, &sTP[t - c0]);
        }
        __syncthreads();
        for (int t = c0; t < c1; t++) {
          const int off = sVoff[t - c0];
          const SparseTerm &T = a.terms[t];
          if (off < 0 || T.nbody == 1) continue;  // skipped, or integrals left in global memory
          int sz = 1;
          for (int k = 0; k < T.nbody; k++) sz *= T.dims[k];
          for (int e = tid; e < sz; e += ST) sV[off + e] = a.ints[T.ioff + e];
        }
        __syncthreads();
#pragma unroll 1
        for (int e = tid; e < n_walk; e += ST) {
          const int code = listed ? (int)sNZ[e] : (((e / nk) << 8) | (e % nk));
          const int I = code >> 8, J = code & 255;
          const uint64_t w = sBits[I][J >> 6];
          if (!((w >> (J & 63)) & 1ull)) continue;
          int below = __popcll(w & ((1ull << (J & 63)) - 1ull));
          for (int k = 0; k < (J >> 6); k++) below += __popcll(sBits[I][k]);
          const int64_t pos = a.row_start[b0 + I] + sCur[I] + below;
          double v = 0.0;
          bool ms = false;
          for (int t = c0; t < c1; t++) {
            const SparseTerm &T = a.terms[t];
            const int off = sVoff[t - c0];
            if (off == -1) continue;
            if (((sPI[I][0] ^ sPJ[J][0]) & T.inact[0]) | ((sPI[I][1] ^ sPJ[J][1]) & T.inact[1]))
              continue;
            const TermPair &tp = sTP[t - c0];
            if (!tp.ok) {
              ms = true;
              continue;
            }
            const double *g[4];
            for (int q = 0; q < T.nbody; q++) {
              const int c = T.cl[q];
              g[q] = a.gdata + tp.base[q] +
                     (int64_t)tp.np[q] * ((state_of(sPI[I], c) - 1) +
                                          (int64_t)tp.mb[q] * (state_of(sPJ[J], c) - 1));
            }
            const int d0 = T.dims[0];
            double x;
            switch (T.nbody == 1 ? 7 : off < 0 ? 0 : shape_id(T)) {
              case 1: x = contract_fixed<4, 4, 4, 4>(sV + off, g); break;
              case 2: x = contract_fixed<16, 4, 4, 1>(sV + off, g); break;
              case 3: x = contract_fixed<4, 16, 4, 1>(sV + off, g); break;
              case 4: x = contract_fixed<4, 4, 16, 1>(sV + off, g); break;
              case 5: x = contract_fixed<16, 16, 1, 1>(sV + off, g); break;
              case 6: x = contract_fixed<4, 4, 1, 1>(sV + off, g); break;
              case 7: x = g[0][0]; break;
              default:
                if (off < 0) x = contract_any(a.ints + T.ioff, g, T.dims, T.nbody);
                else if (d0 == 4) x = contract_v<4>(sV + off, g, T.dims, T.nbody);
                else if (d0 == 16) x = contract_v<16>(sV + off, g, T.dims, T.nbody);
                else if (T.nbody == 2 && T.dims[1] == 4) x = contract_v1<4>(sV + off, g, d0);
                else x = contract_any(sV + off, g, T.dims, T.nbody);
            }
            v += (double)tp.sgn * x;
          }
          if (ms) miss++;
          if (first) {
            a.K[pos] = v;
            a.colidx[pos] = (int32_t)(k0 + J);
            if (fk == fb && J == I) {
              a.diag_koff[b0 + I] = a.row_start[b0 + I];
              a.diag_col[b0 + I] = sCur[I] + below;
            }
          } else {
            a.K[pos] += v;
          }
        }
        __syncthreads();
        c0 = c1;
      }
    }
    for (int I = tid; I < nb; I += ST) {
      int n = 0;
      for (int k = 0; k < MB / 64; k++) n += __popcll(sBits[I][k]);
      sCur[I] += n;
    }
    __syncthreads();
  }
  if (!FILL)
    for (int I = tid; I < nb; I += ST) a.row_nnz[b0 + I] = sCur[I];
  if (trip) atomicAdd(&sTrip, trip);
  if (flop) atomicAdd(&sFlop, flop);
  if (miss) atomicAdd(&sMiss, miss);
  __syncthreads();
  if (tid == 0) {
    if (sTrip) atomicAdd(&a.stats[0], sTrip);
    if (sFlop) atomicAdd(&a.stats[1], sFlop);
    if (sMiss) atomicAdd(&a.stats[2], sMiss);
  }
}

// ---- connected Fock pairs on the device ----
// Fock configurations packed 4 bits per electron count (32 counts in two
// words); ket table: open addressing, key and Fock index per slot.
struct FockSlot {
  uint64_t w0, w1;
  int64_t f;
};

__device__ __forceinline__ uint64_t dmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  return x ^ (x >> 33);
}

struct PairArgs {
  int N, S;
  int64_t n_fock, fb_base;
  const int16_t *fock;     // bra Fock configurations [n_fock][N][2]
  const int16_t *tc;       // TransferConfigs of the live groups [n_live][N][2]
  const int32_t *live;     // group index of each live TransferConfig
  int n_live;
  const FockSlot *table;
  int64_t cap;
  int32_t *count;          // count pass: pairs per bra Fock block
  const int64_t *ptr;      // fill pass
  int32_t *fk, *g;
};

template <bool FILL>
__global__ __launch_bounds__(ST) void k_fock_pairs(PairArgs a) {
  __shared__ int16_t sF[2 * FCG_MAXN];
  __shared__ int32_t sScan[ST];
  __shared__ int32_t sBase;
  const int tid = threadIdx.x;
  const int64_t fb = a.fb_base + blockIdx.x;
  const int N2 = 2 * a.N;
  if (tid < N2) sF[tid] = a.fock[fb * N2 + tid];
  if (tid == 0) sBase = 0;
  __syncthreads();
  int total = 0;
  for (int g0 = 0; g0 < a.n_live; g0 += ST) {
    const int j = g0 + tid;
    int32_t hit = -1;
    if (j < a.n_live) {
      const int16_t *tc = a.tc + (int64_t)j * N2;
      uint64_t w0 = 0, w1 = 0;
      bool ok = true;
      for (int i = 0; i < N2; i++) {
        const int v = sF[i] - tc[i];
        ok = ok && v >= 0 && v < a.S;
        const uint64_t nib = (uint64_t)(v & 0xF);
        if (i < 16) w0 |= nib << (4 * i);
        else w1 |= nib << (4 * (i - 16));
      }
      if (ok) {
        uint64_t h = dmix64(w0 ^ dmix64(w1)) & (uint64_t)(a.cap - 1);
        while (a.table[h].f >= 0) {
          if (a.table[h].w0 == w0 && a.table[h].w1 == w1) {
            hit = (int32_t)a.table[h].f;
            break;
          }
          h = (h + 1) & (uint64_t)(a.cap - 1);
        }
      }
    }
    if (!FILL) {
      total += hit >= 0;
    } else {
      // ordered compaction: block-wide exclusive scan of the hit flags
      sScan[tid] = hit >= 0;
      __syncthreads();
      for (int off = 1; off < ST; off <<= 1) {
        const int v = tid >= off ? sScan[tid - off] : 0;
        __syncthreads();
        sScan[tid] += v;
        __syncthreads();
      }
      if (hit >= 0) {
        const int64_t at = a.ptr[fb] + sBase + sScan[tid] - 1;
        a.fk[at] = hit;
        a.g[at] = a.live[j];
      }
      __syncthreads();
      if (tid == ST - 1) sBase += sScan[ST - 1];
      __syncthreads();
    }
  }
  if (!FILL) {
    sScan[tid] = total;
    __syncthreads();
    if (tid == 0) {
      int s = 0;
      for (int k = 0; k < ST; k++) s += sScan[k];
      a.count[fb] = s;
    }
  }
}

// device scratch of the plan build, stream-ordered on the building context
template <typename T>
struct DBuf {
  fcg_ctx *ctx = nullptr;
  T *p = nullptr;
  bool alloc(fcg_ctx *c, int64_t n) {
    ctx = c;
    p = static_cast<T *>(fcg_dmalloc(c, sizeof(T) * std::max<int64_t>(n, 1)));
    return p != nullptr;
  }
  ~DBuf() {
    if (p) fcg_dfree(ctx, p);
  }
};

template <typename T>
int upload(fcg_ctx *ctx, DBuf<T> &b, const T *src, int64_t n) {
  if (!b.alloc(ctx, n)) {
    ctx->err = "device allocation failed in the sparse plan";
    return FCG_ENOMEM;
  }
  if (n > 0)
    FCG_HIP(hipMemcpyAsync(b.p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream), ctx);
  return FCG_OK;
}

constexpr int64_t FB_CHUNK = 65536;  // bra Fock blocks per launch

uint64_t hmix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  return x ^ (x >> 33);
}

// Connected Fock pairs (bra block, ket block F_bra - TransferConfig, group),
// in group order per bra block, as CSR on the device.  Same list as the
// host fcg_fock_pairs; N <= 16 clusters, electron counts below 16.
int device_fock_pairs(fcg_ctx *ctx, const fcg_ham *h, const fcg_space *sp, int nbody, int S,
                      DBuf<int64_t> &d_ptr, DBuf<int32_t> &d_fk, DBuf<int32_t> &d_g,
                      int64_t *n_out) {
  const int N = sp->N;
  int64_t cap = 16;
  while (cap < 2 * sp->n_fock) cap <<= 1;
  std::vector<FockSlot> table(cap, FockSlot{0, 0, -1});
  for (int64_t f = 0; f < sp->n_fock; f++) {
    uint64_t w[2] = {0, 0};
    for (int i = 0; i < 2 * N; i++)
      w[i >> 4] |= (uint64_t)(sp->fock[f * 2 * N + i] & 0xF) << (4 * (i & 15));
    uint64_t hh = hmix64(w[0] ^ hmix64(w[1])) & (cap - 1);
    while (table[hh].f >= 0) hh = (hh + 1) & (cap - 1);
    table[hh] = FockSlot{w[0], w[1], f};
  }
  std::vector<int32_t> live;
  std::vector<int16_t> tcl;
  for (int g = 0; g < h->n_tc; g++) {
    bool ok = false;
    for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) ok |= h->nbody[t] <= nbody;
    if (!ok) continue;
    live.push_back(g);
    tcl.insert(tcl.end(), &h->tc[(int64_t)g * 2 * N], &h->tc[(int64_t)(g + 1) * 2 * N]);
  }
  DBuf<FockSlot> d_table;
  DBuf<int16_t> d_fock, d_tc;
  DBuf<int32_t> d_live, d_count;
  int rc = upload(ctx, d_table, table.data(), cap);
  if (!rc) rc = upload(ctx, d_fock, sp->fock.data(), (int64_t)sp->fock.size());
  if (!rc) rc = upload(ctx, d_tc, tcl.data(), (int64_t)tcl.size());
  if (!rc) rc = upload(ctx, d_live, live.data(), (int64_t)live.size());
  if (rc) return rc;
  if (!d_count.alloc(ctx, sp->n_fock)) {
    ctx->err = "device allocation failed in the sparse plan";
    return FCG_ENOMEM;
  }
  PairArgs a{};
  a.N = N;
  a.S = S;
  a.n_fock = sp->n_fock;
  a.fock = d_fock.p;
  a.tc = d_tc.p;
  a.live = d_live.p;
  a.n_live = (int)live.size();
  a.table = d_table.p;
  a.cap = cap;
  a.count = d_count.p;
  for (int64_t f0 = 0; f0 < sp->n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    hipLaunchKernelGGL(k_fock_pairs<false>, dim3((unsigned)std::min<int64_t>(FB_CHUNK, sp->n_fock - f0)),
                       dim3(ST), 0, ctx->stream, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<int32_t> cnt(sp->n_fock);
  FCG_HIP(hipMemcpyAsync(cnt.data(), d_count.p, sizeof(int32_t) * sp->n_fock,
                         hipMemcpyDeviceToHost, ctx->stream), ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  std::vector<int64_t> ptr(sp->n_fock + 1, 0);
  for (int64_t f = 0; f < sp->n_fock; f++) ptr[f + 1] = ptr[f] + cnt[f];
  const int64_t n = ptr.back();
  rc = upload(ctx, d_ptr, ptr.data(), (int64_t)ptr.size());
  if (rc) return rc;
  if (!d_fk.alloc(ctx, n) || !d_g.alloc(ctx, n)) {
    ctx->err = "device allocation failed in the sparse plan";
    return FCG_ENOMEM;
  }
  a.ptr = d_ptr.p;
  a.fk = d_fk.p;
  a.g = d_g.p;
  for (int64_t f0 = 0; f0 < sp->n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    hipLaunchKernelGGL(k_fock_pairs<true>, dim3((unsigned)std::min<int64_t>(FB_CHUNK, sp->n_fock - f0)),
                       dim3(ST), 0, ctx->stream, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  *n_out = n;
  return FCG_OK;
}

}  // namespace

int fcg_device_fock_pairs(fcg_ctx *ctx, const fcg_ham *h, const fcg_space *sp, int nbody, int S,
                          int64_t **ptr, int32_t **fk, int32_t **g, int64_t *n) {
  *ptr = nullptr;
  *fk = *g = nullptr;
  *n = 0;
  if (sp->N > 16 || S > 16) {
    ctx->err = "device Fock pairs: at most 16 clusters and 15 electrons per cluster and spin";
    return FCG_EOVERFLOW;
  }
  DBuf<int64_t> d_ptr;
  DBuf<int32_t> d_fk, d_g;
  if (int rc = device_fock_pairs(ctx, h, sp, nbody, S, d_ptr, d_fk, d_g, n)) return rc;
  *ptr = d_ptr.p;
  *fk = d_fk.p;
  *g = d_g.p;
  d_ptr.p = nullptr;  // ownership passes to the caller
  d_fk.p = nullptr;
  d_g.p = nullptr;
  return FCG_OK;
}

bool fcg_sparse_plan_fits(const fcg_space *sp) {
  const int N = sp->N;
  int64_t run = 0;
  for (int64_t b = 0; b < sp->n_box; b++) {
    if (b > 0 && sp->box_fock[b] < sp->box_fock[b - 1]) return false;
    run = (b > 0 && sp->box_fock[b] == sp->box_fock[b - 1]) ? run + 1 : 1;
    if (run > FCG_SPARSE_MAX_BLOCK) return false;
    for (int c = 0; c < N; c++)
      if (sp->lo[b * N + c] > 255) return false;
  }
  return true;
}

int fcg_create_sparse_plan(fcg_ctx *ctx, fcg_ham *h, fcg_ops *o, fcg_space *sp, int nbody,
                           int precisions, fcg_plan **out, int64_t max_bytes, int64_t *need) {
  const auto t0 = std::chrono::steady_clock::now();
  const int N = sp->N, S = o->S;
  const int64_t n_row = sp->n_box;
  // configurations of each Fock block are consecutive rows
  std::vector<int64_t> fstart(sp->n_fock + 1, 0);
  for (int64_t b = 0; b < n_row; b++) {
    const int32_t f = sp->box_fock[b];
    if (b > 0 && f < sp->box_fock[b - 1]) {
      ctx->err = "sparse plan: Fock blocks must be contiguous";
      return FCG_EINVAL;
    }
    fstart[f + 1]++;
    for (int c = 0; c < N; c++)
      if (sp->lo[b * N + c] > 255) {
        ctx->err = "sparse plan: state index above 255";
        return FCG_EOVERFLOW;
      }
  }
  int64_t max_block = 0;
  for (int64_t f = 0; f < sp->n_fock; f++) {
    if (fstart[f + 1] > FCG_SPARSE_MAX_BLOCK) {
      ctx->err = "sparse plan: more than 256 configurations in a Fock block";
      return FCG_EOVERFLOW;
    }
    max_block = std::max(max_block, fstart[f + 1]);
    fstart[f + 1] += fstart[f];
  }
  const bool small = max_block <= SmallRows::MB;
  const int cfg_sv = small ? SmallRows::SV : BigRows::SV;
  const int cfg_svt = small ? SmallRows::SVT : BigRows::SVT;
  std::vector<int64_t> fp_ptr;
  std::vector<int32_t> fp_fk, fp_g;
  DBuf<int64_t> d_fp_ptr;
  DBuf<int32_t> d_fp_fk, d_fp_g;
  const bool dev_pairs = S <= 16 && N <= 16;
  int64_t n_fp = 0;
  if (!dev_pairs) {
    fcg_fock_pairs(h, sp, sp, nbody, S, &fp_ptr, &fp_fk, &fp_g);
    n_fp = fp_ptr.back();
  } else if (int rc0 = device_fock_pairs(ctx, h, sp, nbody, S, d_fp_ptr, d_fp_fk, d_fp_g, &n_fp)) {
    return rc0;
  }
  // term table for the device join
  std::vector<SparseTerm> terms(h->n_terms);
  for (int t = 0; t < h->n_terms; t++) {
    SparseTerm &T = terms[t];
    T = SparseTerm{};
    T.nbody = h->nbody[t];
    T.ioff = h->ioff[t];
    T.inact[0] = T.inact[1] = 0;
    for (int c = 0; c < N; c++) {
      bool act = false;
      for (int q = 0; q < T.nbody; q++) act |= h->clusters[4 * t + q] == c;
      if (!act) T.inact[c >> 3] |= 0xFFull << (8 * (c & 7));
    }
    for (int q = 0; q < 4; q++) {
      T.cl[q] = h->clusters[4 * t + q];
      T.op[q] = h->ops[4 * t + q];
      T.dims[q] = h->dims[4 * t + q];
      if (q < T.nbody && h->parity[4 * t + q] == 1) T.odd |= 1 << q;
    }
  }
  DBuf<int64_t> d_fstart;
  DBuf<int32_t> d_tcs, d_gblock, d_nnz;
  DBuf<int16_t> d_fock;
  DBuf<SparseTerm> d_terms;
  DBuf<int32_t> d_grp_chunk, d_chunk_end, d_voff;
  DBuf<unsigned long long> d_stats;
  int rc = upload(ctx, d_fstart, fstart.data(), (int64_t)fstart.size());
  if (!rc && !dev_pairs) rc = upload(ctx, d_fp_ptr, fp_ptr.data(), (int64_t)fp_ptr.size());
  if (!rc && !dev_pairs) rc = upload(ctx, d_fp_fk, fp_fk.data(), n_fp);
  if (!rc && !dev_pairs) rc = upload(ctx, d_fp_g, fp_g.data(), n_fp);
  if (!rc) rc = upload(ctx, d_tcs, h->tc_start.data(), (int64_t)h->tc_start.size());
  if (!rc) rc = upload(ctx, d_gblock, o->gblock.data(), (int64_t)o->gblock.size());
  if (!rc) rc = upload(ctx, d_fock, sp->fock.data(), (int64_t)sp->fock.size());
  if (!rc) rc = upload(ctx, d_terms, terms.data(), (int64_t)terms.size());
  // each group's terms in chunks that fit the fill pass's stage: integrals
  // of at most SV doubles and at most SVT terms per chunk, in term order
  std::vector<int32_t> grp_chunk(h->n_tc + 1, 0), chunk_end, voff(terms.size(), 0);
  for (int g = 0; g < h->n_tc; g++) {
    grp_chunk[g] = (int32_t)chunk_end.size();
    const int t1 = h->tc_start[g + 1];
    for (int c0 = h->tc_start[g]; c0 < t1;) {
      int off = 0, t = c0;
      for (; t < t1 && t - c0 < cfg_svt; t++) {
        const SparseTerm &T = terms[t];
        int sz = 0;
        if (T.nbody > 1) {
          sz = 1;
          for (int k = 0; k < T.nbody; k++) sz *= T.dims[k];
        }
        if (T.nbody > nbody) sz = 0;
        if (sz > cfg_sv) {  // larger than the stage: integrals read from global memory
          voff[t] = -2;
          continue;
        }
        if (off + sz > cfg_sv) break;
        voff[t] = T.nbody > nbody ? -1 : off;
        off += sz;
      }
      chunk_end.push_back(t);
      c0 = t;
    }
  }
  grp_chunk[h->n_tc] = (int32_t)chunk_end.size();
  if (!rc) rc = upload(ctx, d_grp_chunk, grp_chunk.data(), (int64_t)grp_chunk.size());
  if (!rc) rc = upload(ctx, d_chunk_end, chunk_end.data(), (int64_t)chunk_end.size());
  if (!rc) rc = upload(ctx, d_voff, voff.data(), (int64_t)voff.size());
  if (rc) return rc;
  if (!d_nnz.alloc(ctx, n_row) || !d_stats.alloc(ctx, 3)) {
    ctx->err = "device allocation failed in the sparse plan";
    return FCG_ENOMEM;
  }
  FCG_HIP(hipMemsetAsync(d_stats.p, 0, 3 * sizeof(unsigned long long), ctx->stream), ctx);
  SparseArgs a{};
  a.N = N;
  a.S = S;
  a.n_ops = o->n_ops;
  a.nbody = nbody;
  a.fstart = d_fstart.p;
  a.fock = d_fock.p;
  a.cfg = sp->d_lo;
  a.fp_ptr = d_fp_ptr.p;
  a.fp_fk = d_fp_fk.p;
  a.fp_g = d_fp_g.p;
  a.tc_start = d_tcs.p;
  a.terms = d_terms.p;
  a.grp_chunk = d_grp_chunk.p;
  a.chunk_end = d_chunk_end.p;
  a.voff = d_voff.p;
  a.ints = h->d_ints;
  a.gblock = d_gblock.p;
  a.gb_off = o->d_gb_off;
  a.gb_dims = o->d_gb_dims;
  a.gdata = o->d_gdata;
  a.row_nnz = d_nnz.p;
  a.stats = d_stats.p;
  for (int64_t f0 = 0; f0 < sp->n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    const unsigned nblk = (unsigned)std::min<int64_t>(FB_CHUNK, sp->n_fock - f0);
    if (small)
      hipLaunchKernelGGL((k_sparse_rows<false, SmallRows>), dim3(nblk), dim3(SmallRows::ST), 0,
                         ctx->stream, a);
    else
      hipLaunchKernelGGL((k_sparse_rows<false, BigRows>), dim3(nblk), dim3(BigRows::ST), 0,
                         ctx->stream, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<int32_t> nnz(n_row);
  unsigned long long st[3];
  FCG_HIP(hipMemcpyAsync(nnz.data(), d_nnz.p, sizeof(int32_t) * n_row, hipMemcpyDeviceToHost,
                         ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(st, d_stats.p, sizeof(st), hipMemcpyDeviceToHost, ctx->stream), ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  // row panels (db = 1): widths padded to whole quads
  std::vector<int64_t> col_ptr(n_row + 1, 0);
  for (int64_t b = 0; b < n_row; b++)
    col_ptr[b + 1] = col_ptr[b] + (nnz[b] + FCG_QUAD - 1) / FCG_QUAD * FCG_QUAD;
  const int64_t n_cols = col_ptr[n_row];
  if (n_cols > INT64_MAX / 32) return FCG_EOVERFLOW;
  int64_t reserved = 0;
  {
    const int64_t bytes = fcg_plan_bytes(n_cols, n_cols, n_row, precisions | (1 << FCG_FP64));
    if (int rc2 = fcg_plan_reserve(ctx, bytes, max_bytes, need)) return rc2;
    reserved = bytes;
  }
  fcg_plan *pl = new fcg_plan();
  pl->reserved_bytes = reserved;
  pl->ctx = ctx;
  pl->sp = pl->ksp = sp;
  pl->n_pair = n_fp;
  pl->n_task = (int64_t)st[0];
  pl->n_kel = n_cols;
  pl->n_cols = n_cols;
  pl->build_flops = (double)st[1];
  pl->precisions = precisions | (1 << FCG_FP64);
  std::vector<int32_t> order;
  std::vector<int64_t> var_off(n_row + 1);
  std::iota(var_off.begin(), var_off.end(), 0);
  fcg_apply_order(var_off, col_ptr, &order, &pl->n_multi, &pl->n_single_slots);
  auto fail = [&](int code, const char *what) {
    if (code == FCG_EDEVICE || code == FCG_ENOMEM) ctx->err = std::string("sparse plan: ") + what;
    fcg_plan_destroy(pl);
    return code;
  };
  auto alloc = [&](auto **p, int64_t n) {
    using T = std::remove_reference_t<decltype(**p)>;
    *p = static_cast<T *>(fcg_dmalloc(ctx, sizeof(T) * std::max<int64_t>(n, 1)));
    return *p != nullptr;
  };
  if (!alloc(&pl->d_row_koff, n_row) || !alloc(&pl->d_col_ptr, n_row + 1) ||
      !alloc(&pl->d_order, (int64_t)order.size()) || !alloc(&pl->d_colidx, n_cols) ||
      !alloc(&pl->d_K64, n_cols) || !alloc(&pl->d_diag_koff, n_row) ||
      !alloc(&pl->d_diag_col, n_row))
    return fail(FCG_ENOMEM, "device allocation");
  if ((pl->precisions & (1 << FCG_FP32)) && !alloc(&pl->d_K32, n_cols))
    return fail(FCG_ENOMEM, "device allocation");
  if ((pl->precisions & (1 << FCG_BF16X2)) && !alloc(&pl->d_Kbf, n_cols))
    return fail(FCG_ENOMEM, "device allocation");
  if ((pl->precisions & (1 << FCG_BF16)) && !alloc(&pl->d_Kb16, n_cols))
    return fail(FCG_ENOMEM, "device allocation");
  hipStream_t s = ctx->stream;
  if (hipMemcpyAsync(pl->d_row_koff, col_ptr.data(), sizeof(int64_t) * n_row,
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(pl->d_col_ptr, col_ptr.data(), sizeof(int64_t) * (n_row + 1),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemcpyAsync(pl->d_order, order.data(), sizeof(int32_t) * order.size(),
                     hipMemcpyHostToDevice, s) != hipSuccess ||
      hipMemsetAsync(pl->d_colidx, 0, sizeof(int32_t) * std::max<int64_t>(n_cols, 1), s) !=
          hipSuccess ||
      hipMemsetAsync(pl->d_K64, 0, sizeof(double) * std::max<int64_t>(n_cols, 1), s) !=
          hipSuccess ||
      hipMemsetAsync(pl->d_diag_koff, 0xFF, sizeof(int64_t) * std::max<int64_t>(n_row, 1), s) !=
          hipSuccess ||
      hipMemsetAsync(pl->d_diag_col, 0, sizeof(int32_t) * std::max<int64_t>(n_row, 1), s) !=
          hipSuccess)
    return fail(FCG_EDEVICE, "upload");
  a.row_start = pl->d_col_ptr;
  a.K = pl->d_K64;
  a.colidx = pl->d_colidx;
  a.diag_koff = pl->d_diag_koff;
  a.diag_col = pl->d_diag_col;
  if (hipMemsetAsync(d_stats.p, 0, 3 * sizeof(unsigned long long), s) != hipSuccess)
    return fail(FCG_EDEVICE, "upload");
  for (int64_t f0 = 0; f0 < sp->n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    const unsigned nblk = (unsigned)std::min<int64_t>(FB_CHUNK, sp->n_fock - f0);
    if (small)
      hipLaunchKernelGGL((k_sparse_rows<true, SmallRows>), dim3(nblk), dim3(SmallRows::ST), 0, s,
                         a);
    else
      hipLaunchKernelGGL((k_sparse_rows<true, BigRows>), dim3(nblk), dim3(BigRows::ST), 0, s, a);
  }
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = fcg_launch_convert(pl, s);
  if (e == hipSuccess) e = hipMemcpyAsync(st, d_stats.p, sizeof(st), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    ctx->err = std::string("sparse plan build: ") + hipGetErrorString(e);
    fcg_plan_destroy(pl);
    return FCG_EDEVICE;
  }
  if (st[2]) {
    ctx->err = "sparse plan: a term needs a Γ block that ClusterOps lacks";
    fcg_plan_destroy(pl);
    return FCG_EMISSING_BLOCK;
  }
  pl->build_us =
      std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  *out = pl;
  return FCG_OK;
}
