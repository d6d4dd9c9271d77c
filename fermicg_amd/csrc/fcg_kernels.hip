// Device kernels of libfcg_hip.so (gfx950, wave64).
//
// k_build_blocks  one workgroup per (bra box, ket box) pair: builds the dense
//                 Hamiltonian block sum_t s_t op_t (x) S_inactive in FP64.
//                 op_t is the reference's build_dense_H_term chain
//                 (src/tucker_build_dense_H_term.jl:19-172: compress each
//                 active cluster's Γ with the Tucker factors, then contract
//                 the integrals one cluster at a time), S are the inactive
//                 clusters' factor overlaps U_bra' U_ket
//                 (src/tucker_contract_dense_H_with_state.jl:82-158), or, with
//                 identity factors, a Kronecker delta of the absolute state
//                 indices (the inactive-state test of check_term).  With
//                 identity factors (TPSCI) every element equals the
//                 reference's contract_matrix_element (src/tpsci_inner.jl:
//                 10-216).
// k_apply         sigma = H c: one wavefront per bra box streams that box's
//                 row of blocks, stored side by side as one panel (db x sum
//                 dk, in column quads, panel_index), with a column -> variable
//                 gather of c.  HBM-bound; owner-computes with a fixed-order
//                 reduction, so results are bit-reproducible.
#include "fcg_internal.h"

namespace {

struct BuildArgs {
  const int32_t *pair_B, *pair_K;
  const int64_t *pair_t0, *pair_koff;
  const int32_t *pair_col;
  const TaskDev *tasks;
  const TermDev *terms;
  const double *ints;
  const int64_t *gb_off;
  const int32_t *gb_dims;
  const double *gdata;
  const int16_t *loB, *hiB, *rankB, *loK, *hiK, *rankK;  // bra / ket boxes
  const int64_t *foffB, *foffK;
  const double *facB, *facK;
  int N;
  const int64_t *pair_list;  // pair index of workgroup pair_base + blockIdx.x
  int64_t pair_base;
  double *K;
  // screening (nullptr: none): a term's element enters only when its
  // magnitude exceeds thr[ket variable] (_collect_significant_thread!)
  const double *thr;
  const int64_t *kvar;
  // fused sigma (fsig != nullptr): fsig[bra var, r] += block x fc[ket var, r]
  // instead of storing the block (column-major, nb / nk variables)
  const double *fc;
  double *fsig;
  int fR;
  int64_t nb_var, nk_var;
  const int64_t *bvar;
  // bra variables reached by a nonzero element that passed the screen
  uint8_t *hit;
  // fused sigma: this pair's db x fR products go to fpart + fyoff[pair]
  double *fpart;
  const int64_t *fyoff;
};

__device__ __forceinline__ double fac(const double *F, int64_t off, int nrg, int x, int i) {
  return off < 0 ? (x == i ? 1.0 : 0.0) : F[off + x + (int64_t)nrg * i];
}

constexpr int BT = 256;                       // threads of the build kernel
constexpr int KPT = FCG_MAX_BOX * FCG_MAX_BOX / BT;  // block elements per thread
constexpr int BT_TINY = 64, KPT_TINY = FCG_TINY_ELEMS / BT_TINY;  // the tiny class
static_assert(FCG_MAX_BOX <= 256, "state digits are bytes");
static_assert(FCG_MAX_BOX * FCG_MAX_R <= FCG_LDS_X_SMALL,
              "fused sigma stages c and y in LDS (fused plans use no tiny class)");

// LX / LG / LT / LS: LDS capacities (doubles): the chain's arena holds 2 LX
// (a stage's input at one end, its output at the other), the staged Γ
// slices LG, the factor-compression rows LT and the overlaps LS; the host
// puts each
// pair in the smallest class that holds it.  MINW: waves per SIMD that the
// class's LDS allows (sets the register budget).  NT threads hold EPT block
// elements each (NT * EPT >= the class's largest db * dk).
template <int LX, int LG, int LT, int LS, int MINW, int NT, int EPT>
__global__ __launch_bounds__(NT, MINW) void k_build_blocks(BuildArgs a) {
  __shared__ double sA[2 * LX];
  __shared__ double sG[LG];
  __shared__ double sT[LT];
  __shared__ double sS[LS];
  __shared__ int sLoB[FCG_MAXN], sLoK[FCG_MAXN], sNB[FCG_MAXN], sNK[FCG_MAXN];
  __shared__ int sRB[FCG_MAXN], sRK[FCG_MAXN], sKind[FCG_MAXN], sSoff[FCG_MAXN];
  __shared__ int64_t sFB[FCG_MAXN], sFK[FCG_MAXN];
  __shared__ int sAct[FCG_MAXN];
  __shared__ int sStride[FCG_MAXN];
  // per pair: each bra row's / ket column's state digit per cluster, and the
  // screening threshold of each ket column; per task: each row's and column's
  // share of the chain-output index and the key of its inactive digits
  __shared__ uint8_t sDigB[FCG_MAX_BOX * FCG_MAXN], sDigK[FCG_MAX_BOX * FCG_MAXN];
  __shared__ int2 sRowT[FCG_MAX_BOX], sColT[FCG_MAX_BOX];
  __shared__ double sThr[FCG_MAX_BOX];
  __shared__ int sFast;
  // per task, per chain stage s: the cluster, Γ index count, Γ block and
  // integral stride of the term position contracted at stage s (in LDS, so
  // that loops over stages index them without a private stack array)
  __shared__ int sCh[4][4];

  const int tid = threadIdx.x;
  const int64_t p = a.pair_list[a.pair_base + blockIdx.x];
  const int64_t B = a.pair_B[p], K = a.pair_K[p];
  const int N = a.N;
  if (tid < N) {
    const int c = tid;
    sLoB[c] = a.loB[B * N + c] - 1;
    sLoK[c] = a.loK[K * N + c] - 1;
    sNB[c] = a.hiB[B * N + c] - a.loB[B * N + c] + 1;
    sNK[c] = a.hiK[K * N + c] - a.loK[K * N + c] + 1;
    sRB[c] = a.rankB[B * N + c];
    sRK[c] = a.rankK[K * N + c];
    sFB[c] = a.foffB[B * N + c];
    sFK[c] = a.foffK[K * N + c];
  }
  __syncthreads();
  if (tid == 0) {
    int off = 0;
    for (int c = 0; c < N; c++) {
      const bool same = sLoB[c] == sLoK[c] && sNB[c] == sNK[c];
      if (sFB[c] < 0 && sFK[c] < 0) sKind[c] = 0;  // identity: equal states
      else if (same) { sKind[c] = 1; sSoff[c] = off; off += sRB[c] * sRK[c]; }
      else sKind[c] = 2;
    }
  }
  __syncthreads();
  int db = 1, dk = 1;
  for (int c = 0; c < N; c++) { db *= sRB[c]; dk *= sRK[c]; }
  // overlaps of inactive-capable clusters: S = U_bra' U_ket
  for (int c = 0; c < N; c++) {
    if (sKind[c] != 1) continue;
    const int rb = sRB[c], rk = sRK[c], nrg = sNB[c];
    for (int e = tid; e < rb * rk; e += NT) {
      const int i = e % rb, j = e / rb;
      double s = 0.0;
      for (int x = 0; x < nrg; x++)
        s += fac(a.facB, sFB[c], nrg, x, i) * fac(a.facK, sFK[c], nrg, x, j);
      sS[sSoff[c] + e] = s;
    }
  }
  for (int x0 = tid; x0 < db + dk; x0 += NT) {  // mixed-radix digits, cluster 0 fastest
    const bool row = x0 < db;
    int x = row ? x0 : x0 - db;
    uint8_t *dg = row ? sDigB + x * N : sDigK + x * N;
    for (int c = 0; c < N; c++) {
      const int rc = row ? sRB[c] : sRK[c];
      dg[c] = (uint8_t)(x % rc);
      x /= rc;
    }
  }
  if (a.thr)
    for (int q = tid; q < dk; q += NT) sThr[q] = a.thr[a.kvar[K] + q];
  // this thread's block elements e = tid + NT k (column-major, db rows) are
  // visited as (row r, column q), advanced by (dr, dq) per step
  const int r0 = tid % db, q0 = tid / db, dr = NT % db, dq = NT / db;
  double acc[EPT];
#pragma unroll
  for (int k = 0; k < EPT; k++) acc[k] = 0.0;
  uint32_t hm = 0;  // bit k: element k received a nonzero screened contribution
  static_assert(EPT <= 32, "hit mask holds one bit per element of a thread");
  __syncthreads();

  for (int64_t t = a.pair_t0[p]; t < a.pair_t0[p + 1]; t++) {
    const int32_t meta = a.tasks[t].meta;
    const TermDev *tm = a.terms + a.tasks[t].term;
    const int nb = tm->nbody;
    double *cur = sA;
    if (tid < nb) {
      const int tp = task_order(meta, tid);
      int str = 1;
      for (int j = 0; j < tp; j++) str *= tm->dims[j];
      sCh[tid][0] = tm->cl[tp];
      sCh[tid][1] = nb == 1 ? 1 : tm->dims[tp];
      sCh[tid][2] = a.tasks[t].blk[tp];
      sCh[tid][3] = str;
    }
    if (tid < N) sAct[tid] = -1;
    __syncthreads();
    int rest = 1;
    if (nb == 1) {
      if (tid == 0) cur[0] = 1.0;
    } else {
      // integrals with their axes in chain order (stage s contracts term
      // position task_order(meta, s)); original layout is column-major in
      // term order
      for (int s = 0; s < nb; s++) rest *= sCh[s][1];
      const int64_t ioff = tm->ioff;
      for (int e = tid; e < rest; e += NT) {
        int r = e, src = 0;
        for (int s = 0; s < nb; s++) {
          const int d = sCh[s][1];
          src += (r % d) * sCh[s][3];
          r /= d;
        }
        cur[e] = a.ints[ioff + src];
      }
    }
    if (tid == 0) {
      int st = 1;
      for (int s = 0; s < nb; s++) {
        const int c = sCh[s][0];
        sAct[c] = s;
        sStride[c] = st;
        st *= sRB[c] * sRK[c];
      }
      // the separable form needs every inactive cluster to be a Kronecker
      // delta of absolute state indices (identity factors on both sides)
      int fast = 1;
      for (int c = 0; c < N; c++)
        if (sAct[c] < 0 && sKind[c] != 0) fast = 0;
      sFast = fast;
    }
    for (int x0 = tid; x0 < db + dk; x0 += NT) {
      // element (r, q) of the task's term reads cur[rowoff[r] + coloff[q]]
      // and is nonzero when rowkey[r] == colkey[q]: the inactive clusters'
      // bra states shifted to ket numbering, in the ket digits' mixed radix
      // (-1 when one falls outside the ket box)
      const bool row = x0 < db;
      const int x = row ? x0 : x0 - db;
      const uint8_t *dg = row ? sDigB + x * N : sDigK + x * N;
      int off = 0, st = 1;
      for (int s = 0; s < nb; s++) {
        const int c = sCh[s][0];
        off += (row ? dg[c] : sRB[c] * dg[c]) * st;
        st *= sRB[c] * sRK[c];
      }
      int key = 0, w = 1;
      bool out = false;
      for (int c = 0; c < N; c++) {
        bool act = false;
        for (int s = 0; s < nb; s++) act |= sCh[s][0] == c;
        if (act) continue;
        int v = dg[c];
        if (row) {
          v += sLoB[c] - sLoK[c];
          out |= v < 0 || v >= sRK[c];
        }
        key += v * w;
        w *= sRK[c];
      }
      (row ? sRowT[x] : sColT[x]) = make_int2(off, out ? -1 : key);
    }
    for (int s = 0; s < nb; s++) {
      const int c = sCh[s][0];
      const int blk = sCh[s][2];
      const int np = sCh[s][1];
      const int mb = a.gb_dims[3 * blk + 1];
      const double *G = a.gdata + a.gb_off[blk];
      const int rb = sRB[c], rk = sRK[c], lb = sLoB[c], lk = sLoK[c];
      const int nI = sNB[c], nJ = sNK[c];
      const int64_t fB = sFB[c], fK = sFK[c];
      if (fB < 0 && fK < 0) {
        for (int e = tid; e < np * rb * rk; e += NT) {
          const int pp = e % np, ij = e / np, i = ij % rb, j = ij / rb;
          sG[e] = G[pp + (int64_t)np * ((lb + i) + (int64_t)mb * (lk + j))];
        }
        __syncthreads();
      } else {
        int Pc = LT / (nI * rk);
        if (Pc > np) Pc = np;
        for (int p0 = 0; p0 < np; p0 += Pc) {
          const int pc = min(Pc, np - p0);
          for (int e = tid; e < pc * nI * rk; e += NT) {
            const int pp = e % pc, r2 = e / pc, I = r2 % nI, j = r2 / nI;
            double v = 0.0;
            const double *Gp = G + (p0 + pp) + (int64_t)np * (lb + I);
            if (fK < 0) {
              v = Gp[(int64_t)np * mb * (lk + j)];
            } else {
              for (int J = 0; J < nJ; J++)
                v += Gp[(int64_t)np * mb * (lk + J)] * a.facK[fK + J + (int64_t)nJ * j];
            }
            sT[e] = v;
          }
          __syncthreads();
          for (int e = tid; e < pc * rb * rk; e += NT) {
            const int pp = e % pc, r2 = e / pc, i = r2 % rb, j = r2 / rb;
            double v = 0.0;
            if (fB < 0) {
              v = sT[pp + pc * (i + nI * j)];
            } else {
              for (int I = 0; I < nI; I++)
                v += a.facB[fB + I + (int64_t)nI * i] * sT[pp + pc * (I + nI * j)];
            }
            sG[(p0 + pp) + np * (i + rb * j)] = v;
          }
          __syncthreads();
        }
      }
      // contract the leading index: nxt[q, P] = sum_p cur[p, q] g[p, P],
      // nxt at the arena's other end
      const int rest2 = rest / np, PP = rb * rk, osz = rest2 * PP;
      double *nxt = cur == sA ? sA + 2 * LX - osz : sA;
      for (int e = tid; e < osz; e += NT) {
        const int q = e % rest2, P = e / rest2;
        double v = 0.0;
        for (int pp = 0; pp < np; pp++) v += cur[pp + np * q] * sG[pp + np * P];
        nxt[e] = v;
      }
      __syncthreads();
      cur = nxt;
      rest = osz;
    }
    // Kronecker accumulate into this thread's block elements
    const double sign = (meta & 1) ? -1.0 : 1.0;
    if (sFast) {
      int r = r0, q = q0;
      for (int k = 0; k < EPT; k++) {
        if (q >= dk) break;
        const int2 R = sRowT[r], C = sColT[q];
        if (R.y == C.y) {
          const double x = cur[R.x + C.x];
          if (!a.thr || fabs(x) > sThr[q]) {
            acc[k] += sign * x;
            if (x != 0.0) hm |= 1u << k;
          }
        }
        r += dr;
        q += dq;
        if (r >= db) { r -= db; q++; }
      }
    } else {
      int r = r0, q = q0;
      for (int k = 0; k < EPT; k++) {
        if (q >= dk) break;
        const uint8_t *dB = sDigB + r * N, *dK = sDigK + q * N;
        double val = sign;
        int opidx = 0;
        for (int c = 0; c < N; c++) {
          const int ic = dB[c], jc = dK[c];
          if (sAct[c] >= 0) {
            opidx += (ic + sRB[c] * jc) * sStride[c];
          } else if (sKind[c] == 0) {
            if (sLoB[c] + ic != sLoK[c] + jc) { val = 0.0; break; }
          } else if (sKind[c] == 1) {
            val *= sS[sSoff[c] + ic + sRB[c] * jc];
          } else {
            val = 0.0;
            break;
          }
        }
        if (val != 0.0) {
          const double x = cur[opidx];
          if (!a.thr || fabs(x) > sThr[q]) {
            acc[k] += val * x;
            if (x != 0.0) hm |= 1u << k;
          }
        }
        r += dr;
        q += dq;
        if (r >= db) { r -= db; q++; }
      }
    }
    __syncthreads();
  }
  if (a.hit && hm) {
    const int64_t bv0 = a.bvar[B];
#pragma unroll
    for (int k = 0; k < EPT; k++)
      if ((hm >> k) & 1u) a.hit[bv0 + (tid + NT * k) % db] = 1;
  }
  if (a.fsig) {
    // fused sigma: y[row, r] = sum_q block[row, q] c[q, r] in column order.
    // The block's columns pass through the free Γ stage (sG) in chunks; each
    // (row, root) entry of y belongs to one thread, which adds the chunk's
    // products in order, so y is the same however the threads are
    // scheduled.  The pair's y goes to its own slot (no atomics); the bra
    // box's slots are summed in pair order by k_fused_sum.
    double *sY = sA, *sC = sA + LX;
    const int fR = a.fR, nY = db * fR;
    const int64_t kv0 = a.kvar[K];
    for (int e = tid; e < nY; e += NT) sY[e] = 0.0;
    for (int e = tid; e < dk * fR; e += NT)
      sC[e] = a.fc[(int64_t)(e / dk) * a.nk_var + kv0 + e % dk];
    const int CH = LG / db;  // columns per chunk (LG >= FCG_MAX_BOX for the fused classes)
    for (int c0 = 0; c0 < dk; c0 += CH) {
      const int c1 = min(dk, c0 + CH);
      __syncthreads();
      {
        int r = r0, q = q0;
        for (int k = 0; k < EPT; k++) {
          if (q >= dk) break;
          if (q >= c0 && q < c1) sG[r + db * (q - c0)] = acc[k];
          r += dr;
          q += dq;
          if (r >= db) { r -= db; q++; }
        }
      }
      __syncthreads();
      for (int e = tid; e < nY; e += NT) {
        const int r = e / fR;
        const double *cq = sC + (e % fR) * dk;
        double y = sY[e];
        for (int q = c0; q < c1; q++) y += sG[r + db * (q - c0)] * cq[q];
        sY[e] = y;
      }
    }
    double *slot = a.fpart + a.fyoff[p];
    for (int e = tid; e < nY; e += NT) slot[e] = sY[e];
    return;
  }
  double *Kout = a.K + a.pair_koff[p];
  const int64_t col0 = a.pair_col[p];
#pragma unroll
  for (int k = 0; k < EPT; k++) {
    const int e = tid + NT * k;
    if (e < db * dk) Kout[panel_index(db, e % db, col0 + e / db)] = acc[k];
  }
}

// ---- sigma apply ----
// Element conversion of the stored Hamiltonian blocks.
__device__ __forceinline__ double kval(double v) { return v; }
__device__ __forceinline__ float kval(float v) { return v; }
__device__ __forceinline__ float kval(uint32_t u) {  // hi bf16 | lo bf16
  return __uint_as_float(u & 0xffff0000u) + __uint_as_float(u << 16);
}
__device__ __forceinline__ float kval(uint16_t u) {  // bf16
  return __uint_as_float((uint32_t)u << 16);
}

// Apply variants, picked per plan, precision and root count by timing each
// once on the plan's first sigma (fcg_sigma_device).  Every variant reads the
// same elements in the same order per lane and reduces in the same order, so
// the result does not depend on the choice.
//   multi-row boxes (k_apply): wavefronts per workgroup, quads in flight;
//     32-bit blocks load branch-free from a clamped quad, 64-bit ones
//     predicated (measured faster each way);
//   one-row boxes (k_apply_rows): quads in flight.
struct MultiVariant { int wpb, unroll; bool nt; };
constexpr MultiVariant kMulti[] = {{4, 16, false}, {1, 16, false}, {4, 8, false},
                                   {4, 16, true}, {1, 16, true}};
struct RowsVariant { int unroll; bool nt; };
constexpr RowsVariant kRows[] = {{8, false}, {4, false}, {8, true}, {4, true}};
constexpr int N_MULTI_VARIANTS = 5, N_ROWS_VARIANTS = 4;
constexpr int TUNED_MAX_R = 4;  // larger root counts use the default variants
template <typename TK>
constexpr int default_multi() { return sizeof(TK) == 8 ? 1 : 0; }
template <typename TK>
constexpr bool apply_clamp() { return sizeof(TK) != 8; }

template <typename T> struct Vec2;
template <> struct Vec2<float> { using type = float2; };
template <> struct Vec2<double> { using type = double2; };

// an element of c rounded to bf16 (FCG_FP32_BF16C): read as the FP32 value
// whose upper 16 bits it holds
struct bf16c {
  uint16_t u;
  __device__ operator float() const { return __uint_as_float((uint32_t)u << 16); }
};

// R consecutive values from global memory (16- or 8-byte vector loads when
// R allows; the interleaved vector keeps a variable's R roots contiguous)
template <typename TC, int R>
__device__ __forceinline__ void load_row(const TC *__restrict__ p, TC (&v)[R]) {
  if constexpr (sizeof(TC) == 4 && R % 4 == 0) {
#pragma unroll
    for (int h = 0; h < R / 4; h++) {
      const float4 t = reinterpret_cast<const float4 *>(p)[h];
      v[4 * h] = t.x; v[4 * h + 1] = t.y; v[4 * h + 2] = t.z; v[4 * h + 3] = t.w;
    }
  } else if constexpr (sizeof(TC) == 2 && R % 4 == 0) {
#pragma unroll
    for (int h = 0; h < R / 4; h++) {
      const uint2 t = reinterpret_cast<const uint2 *>(p)[h];
      v[4 * h].u = (uint16_t)t.x;
      v[4 * h + 1].u = (uint16_t)(t.x >> 16);
      v[4 * h + 2].u = (uint16_t)t.y;
      v[4 * h + 3].u = (uint16_t)(t.y >> 16);
    }
  } else if constexpr (sizeof(TC) == 2 && R % 2 == 0) {
#pragma unroll
    for (int h = 0; h < R / 2; h++) {
      const uint32_t t = reinterpret_cast<const uint32_t *>(p)[h];
      v[2 * h].u = (uint16_t)t;
      v[2 * h + 1].u = (uint16_t)(t >> 16);
    }
  } else if constexpr (sizeof(TC) == 2) {
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = p[r];
  } else if constexpr (R % 2 == 0) {
    using V2 = typename Vec2<TC>::type;
#pragma unroll
    for (int h = 0; h < R / 2; h++) {
      const V2 t = reinterpret_cast<const V2 *>(p)[h];
      v[2 * h] = t.x;
      v[2 * h + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = p[r];
  }
}

// R consecutive staged values (paired vector LDS reads when R is even)
template <typename TC, int R>
__device__ __forceinline__ void stage_read(const TC *p, TC (&v)[R]) {
  if constexpr (R % 2 == 0) {
    using V2 = typename Vec2<TC>::type;
    const V2 *q = reinterpret_cast<const V2 *>(p);
#pragma unroll
    for (int h = 0; h < R / 2; h++) {
      const V2 t = q[h];
      v[2 * h] = t.x;
      v[2 * h + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = p[r];
  }
}

__device__ __forceinline__ void wave_sync() {
  // order this wavefront's LDS accesses (all lanes of a wave run in lockstep)
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads of data read once per sigma (blocks, column indices)
typedef float nf4 __attribute__((ext_vector_type(4)));
typedef unsigned nu4 __attribute__((ext_vector_type(4)));
typedef unsigned nu2 __attribute__((ext_vector_type(2)));
typedef double nd2 __attribute__((ext_vector_type(2)));
typedef int ni4 __attribute__((ext_vector_type(4)));
// NT: non-temporal (the stream does not displace the gathered vector from
// the caches); a variant choice, measured faster for some precisions and
// plans and slower for others
template <bool NT, typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
  if constexpr (NT) return __builtin_nontemporal_load(p);
  else return *p;
}

// One column quad of one panel row as loaded from HBM (16 or 32 bytes).
template <typename TK> struct RawQuad;
template <> struct RawQuad<float> {
  float4 v;
  template <bool NT> __device__ void load(const float *p) {
    const nf4 t = ld_stream<NT>(reinterpret_cast<const nf4 *>(p));
    v = make_float4(t.x, t.y, t.z, t.w);
  }
  __device__ void zero() { v = make_float4(0.f, 0.f, 0.f, 0.f); }
  __device__ void get(float (&k)[4]) const { k[0] = v.x; k[1] = v.y; k[2] = v.z; k[3] = v.w; }
};
template <> struct RawQuad<uint32_t> {
  uint4 v;
  template <bool NT> __device__ void load(const uint32_t *p) {
    const nu4 t = ld_stream<NT>(reinterpret_cast<const nu4 *>(p));
    v = make_uint4(t.x, t.y, t.z, t.w);
  }
  __device__ void zero() { v = make_uint4(0u, 0u, 0u, 0u); }
  __device__ void get(float (&k)[4]) const {
    k[0] = kval(v.x); k[1] = kval(v.y); k[2] = kval(v.z); k[3] = kval(v.w);
  }
};
template <> struct RawQuad<uint16_t> {  // four bf16 in 8 bytes
  uint2 v;
  template <bool NT> __device__ void load(const uint16_t *p) {
    const nu2 t = ld_stream<NT>(reinterpret_cast<const nu2 *>(p));
    v = make_uint2(t.x, t.y);
  }
  __device__ void zero() { v = make_uint2(0u, 0u); }
  __device__ void get(float (&k)[4]) const {
    k[0] = __uint_as_float(v.x << 16); k[1] = __uint_as_float(v.x & 0xffff0000u);
    k[2] = __uint_as_float(v.y << 16); k[3] = __uint_as_float(v.y & 0xffff0000u);
  }
};
template <> struct RawQuad<double> {
  double2 a, b;
  template <bool NT> __device__ void load(const double *p) {
    const nd2 x = ld_stream<NT>(reinterpret_cast<const nd2 *>(p));
    const nd2 y = ld_stream<NT>(reinterpret_cast<const nd2 *>(p) + 1);
    a = make_double2(x.x, x.y);
    b = make_double2(y.x, y.y);
  }
  __device__ void zero() { a = b = make_double2(0.0, 0.0); }
  __device__ void get(double (&k)[4]) const { k[0] = a.x; k[1] = a.y; k[2] = b.x; k[3] = b.y; }
};

// Column quads staged per chunk: the per-wave stage holds 4*CG*R entries
// within 8 KB, and at least the 64*R partial sums of the final reduction.
template <typename TC, int R>
constexpr int stage_quads() {
  constexpr int q = 8192 / (FCG_QUAD * R * (int)sizeof(TC));
  return q > 64 ? 64 : q;
}

// sigma[var[b]+i, r] = sum_j panel_b[i, j] c[colidx_b[j], r] for every bra
// box b.  One wavefront per bra box, boxes taken in `order` (largest panel
// first, so the long rows start early and similar rows share a workgroup);
// lane l = i + db*s (s < sl = 64/db) owns
// panel row i and every sl-th column quad, loading each quad with one
// 16-byte (FP32, BF16X2) or two 16-byte (FP64) loads; consecutive lanes read
// consecutive quads, so a wave streams contiguous panel memory.  The column
// indices and vector entries of the next CG quads are gathered cooperatively
// (one quad per lane) one chunk ahead and staged in a per-wave LDS buffer,
// read back with vector LDS loads broadcast across the lanes of a quad.
// Slices are summed in a fixed order through LDS, so results are
// reproducible bit for bit.
template <typename TK, typename TA, typename TC, int R, int WPB, int UNR, bool CLAMP, bool NT>
__global__ __launch_bounds__(64 * WPB) void k_apply(int64_t n_box,
                                                          const int32_t *__restrict__ order,
                                                          const int64_t *__restrict__ var,
                                               const int64_t *__restrict__ row_koff,
                                               const int64_t *__restrict__ col_ptr,
                                               const int32_t *__restrict__ colidx,
                                               const TK *__restrict__ K, const TC *__restrict__ c,
                                               int64_t nvar_s, double *__restrict__ sigma) {
  static_assert(sizeof(TA) == sizeof(TC), "stage buffer doubles as reduction buffer");
  constexpr int APPLY_UNROLL = UNR;
  constexpr int APPLY_WPB = WPB;
  constexpr int CG = stage_quads<TC, R>();
  static_assert(CG * FCG_QUAD >= 64, "reduction needs 64*R staged entries");
  __shared__ __align__(16) TC stage[APPLY_WPB][CG * FCG_QUAD * R];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * APPLY_WPB + wave;
  const bool live = w < n_box && order[w] >= 0;
  const int64_t b = live ? order[w] : 0;
  const int db = live ? (int)(var[b + 1] - var[b]) : 1;
  const int sl = 64 / db;
  const int i = lane % db, s = lane / db;
  const bool act = s < sl;
  TA acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = (TA)0;
  TC *st = stage[wave];
  if (live) {  // uniform per wave
    const int64_t c0 = col_ptr[b];
    const int ngrp = (int)((col_ptr[b + 1] - c0) / FCG_QUAD);
    const TK *Kr = K + row_koff[b] + FCG_QUAD * i;
    const int32_t *ci = colidx + c0;
    const int nk = (CG + sl - 1) / sl;  // quad steps per chunk (uniform)
    TC nv[FCG_QUAD][R];
    auto gather = [&](int g) {
      const bool ok = lane < CG && g < ngrp;
      const int4 idx = ok ? *reinterpret_cast<const int4 *>(ci + (int64_t)FCG_QUAD * g)
                          : make_int4(0, 0, 0, 0);
      const int col[FCG_QUAD] = {idx.x, idx.y, idx.z, idx.w};
#pragma unroll
      for (int q = 0; q < FCG_QUAD; q++) {
        if (ok) {
          load_row<TC, R>(c + (int64_t)col[q] * R, nv[q]);
        } else {
#pragma unroll
          for (int r = 0; r < R; r++) nv[q][r] = (TC)0;
        }
      }
    };
    gather(lane);
    for (int base = 0; base < ngrp; base += CG) {
      wave_sync();
      if (lane < CG) {
#pragma unroll
        for (int q = 0; q < FCG_QUAD; q++)
#pragma unroll
          for (int r = 0; r < R; r++) st[(FCG_QUAD * lane + q) * R + r] = nv[q][r];
      }
      wave_sync();
      gather(base + CG + lane);  // next chunk, in flight during this one
      const int lim = act ? min(CG, ngrp - base) : 0;
      const TK *Kc = Kr + (int64_t)FCG_QUAD * db * base;
      for (int k0 = 0; k0 < nk; k0 += APPLY_UNROLL) {
        RawQuad<TK> kv[APPLY_UNROLL];
        if constexpr (CLAMP) {
          // unconditional loads from a clamped (valid) quad, then a select:
          // no branch per load
          const int jmax = max(lim - 1, 0);
#pragma unroll
          for (int u = 0; u < APPLY_UNROLL; u++)
            kv[u].template load<NT>(Kc + (uint32_t)(FCG_QUAD * db * min(s + sl * (k0 + u), jmax)));
#pragma unroll
          for (int u = 0; u < APPLY_UNROLL; u++)
            if (s + sl * (k0 + u) >= lim) kv[u].zero();
        } else {
#pragma unroll
          for (int u = 0; u < APPLY_UNROLL; u++) {
            const int j = s + sl * (k0 + u);
            if (j < lim) kv[u].template load<NT>(Kc + (uint32_t)(FCG_QUAD * db * j));
            else kv[u].zero();
          }
        }
#pragma unroll
        for (int u = 0; u < APPLY_UNROLL; u++) {
          const int j = min(s + sl * (k0 + u), CG - 1);
          TA kq[FCG_QUAD];
          kv[u].get(kq);
#pragma unroll
          for (int q = 0; q < FCG_QUAD; q++) {
            TC cvj[R];
            stage_read<TC, R>(st + (FCG_QUAD * j + q) * R, cvj);
#pragma unroll
            for (int r = 0; r < R; r++) acc[r] += kq[q] * (TA)cvj[r];
          }
        }
      }
    }
  }
  // fixed-order slice reduction in this wave's stage buffer
  wave_sync();
  TA *red = reinterpret_cast<TA *>(st);
#pragma unroll
  for (int r = 0; r < R; r++) red[lane * R + r] = act ? acc[r] : (TA)0;
  wave_sync();
  if (live && lane < db) {
#pragma unroll
    for (int r = 0; r < R; r++) {
      TA v = (TA)0;
      for (int s2 = 0; s2 < sl; s2++) v += red[(lane + db * s2) * R + r];
      sigma[(int64_t)r * nvar_s + var[b] + lane] = (double)v;
    }
  }
}

// One-row boxes (db = 1, the rows of sparse TPSCI spaces): no row shares a
// column with another, so there is nothing to stage; the wave streams the
// row's columns in slabs of 64*4U, lane l taking columns l, l+64, ... of a
// slab (4U in flight), so each block and index load is coalesced and
// neighbouring lanes gather neighbouring columns: the runs of consecutive
// ket configurations in a row share cache lines within one gather
// instruction (1.45x over quad-per-lane gathers on the 2e6-configuration
// rank-1 space).  Rows come from `order` (XCD-grouped slots, -1 = empty);
// the 64 lane sums are combined by a fixed butterfly.
template <typename TK, typename TA, typename TC, int R, int U, bool NT>
__global__ __launch_bounds__(64 * FCG_CSR_WPB) void k_apply_rows(
    int64_t n_slots, const int32_t *__restrict__ order, const int64_t *__restrict__ var,
    const int64_t *__restrict__ row_koff, const int64_t *__restrict__ col_ptr,
    const int32_t *__restrict__ colidx, const TK *__restrict__ K, const TC *__restrict__ c,
    int64_t nvar_s, double *__restrict__ sigma) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int64_t w = (int64_t)blockIdx.x * FCG_CSR_WPB + wave;
  const int32_t b = w < n_slots ? order[w] : -1;
  if (b < 0) return;  // uniform per wave
  const int64_t c0 = col_ptr[b];
  const int ngrp = (int)((col_ptr[b + 1] - c0) / FCG_QUAD);
  const TK *Kr = K + row_koff[b];
  const int32_t *ci = colidx + c0;
  TA acc[R];
#pragma unroll
  for (int r = 0; r < R; r++) acc[r] = (TA)0;
  const int ncol = FCG_QUAD * ngrp;
  const TK *Kl = Kr + lane;
  const int32_t *cl = ci + lane;
  constexpr int NE = 4 * U;  // columns in flight per lane
  for (int s0 = 0; s0 < ncol; s0 += 64 * NE) {
    TK kv[NE];
    int idx[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
      const int j = s0 + 64 * e;
      if (j + lane < ncol) {
        kv[e] = ld_stream<NT>(Kl + j);
        idx[e] = ld_stream<NT>(cl + j);
      } else {
        kv[e] = (TK)0;
        idx[e] = 0;
      }
    }
#pragma unroll
    for (int e = 0; e < NE; e++) {
      TC cv[R];
      load_row<TC, R>(c + (int64_t)idx[e] * R, cv);
      const TA kq = kval(kv[e]);
#pragma unroll
      for (int r = 0; r < R; r++) acc[r] += kq * (TA)cv[r];
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    TA v = acc[r];
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    acc[r] = v;
  }
  if (lane == 0) {
#pragma unroll
    for (int r = 0; r < R; r++) sigma[(int64_t)r * nvar_s + var[b]] = (double)acc[r];
  }
}

__device__ __forceinline__ uint32_t bf16_rne_bits(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return u & 0xffff0000u;
}

__global__ void k_convert(int64_t n, const double *__restrict__ K64, float *__restrict__ K32,
                          uint32_t *__restrict__ Kbf, uint16_t *__restrict__ Kb16) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float f = (float)K64[e];
    if (K32) K32[e] = f;
    if (Kbf) {
      const uint32_t hi = bf16_rne_bits(f);
      const float r = f - __uint_as_float(hi);
      const uint32_t lo = bf16_rne_bits(r);
      Kbf[e] = hi | (lo >> 16);
    }
    if (Kb16) Kb16[e] = (uint16_t)(bf16_rne_bits(f) >> 16);
  }
}

// y[i*R + r] = x[r*n + i]: the trial vectors in the apply's interleaved
// layout (a variable's R roots contiguous), converted to TC
template <typename TC>
__global__ void k_interleave(int64_t n, int R, const double *__restrict__ x, TC *__restrict__ y) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = (TC)x[(e % R) * n + e / R];
}

// the same, each value rounded to bf16 (round to nearest even)
__global__ void k_interleave_bf16(int64_t n, int R, const double *__restrict__ x,
                                  bf16c *__restrict__ y) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e].u = (uint16_t)(bf16_rne_bits((float)x[(e % R) * n + e / R]) >> 16);
}

__global__ void k_diag(int64_t n_box, const int64_t *__restrict__ var,
                       const int64_t *__restrict__ dkoff, const int32_t *__restrict__ dcol,
                       const double *__restrict__ K, double *__restrict__ hd) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_box) return;
  const int64_t v0 = var[b];
  const int db = (int)(var[b + 1] - v0);
  for (int i = 0; i < db; i++)
    hd[v0 + i] = dkoff[b] >= 0 ? K[dkoff[b] + panel_index(db, i, (int64_t)dcol[b] + i)] : 0.0;
}

}  // namespace

hipError_t fcg_launch_build(const fcg_plan *pl, const fcg_ham *h, const fcg_ops *o,
                            hipStream_t s, const FusedSigma *fused) {
  const fcg_space *sp = pl->sp;
  BuildArgs a;
  a.pair_B = pl->d_pair_B;
  a.pair_K = pl->d_pair_K;
  a.pair_t0 = pl->d_pair_t0;
  a.pair_koff = pl->d_pair_koff;
  a.pair_col = pl->d_pair_col;
  a.tasks = pl->d_tasks;
  a.terms = h->d_terms;
  a.ints = h->d_ints;
  a.gb_off = o->d_gb_off;
  a.gb_dims = o->d_gb_dims;
  a.gdata = o->d_gdata;
  const fcg_space *kp = pl->ksp;
  a.loB = sp->d_lo;
  a.hiB = sp->d_hi;
  a.rankB = sp->d_rank;
  a.foffB = sp->d_foff;
  a.facB = sp->d_factors;
  a.loK = kp->d_lo;
  a.hiK = kp->d_hi;
  a.rankK = kp->d_rank;
  a.foffK = kp->d_foff;
  a.facK = kp->d_factors;
  a.N = sp->N;
  a.K = pl->d_K64;
  a.thr = pl->d_thr;
  a.kvar = kp->d_var;
  a.pair_list = pl->d_pair_list;
  a.fc = fused ? fused->c : nullptr;
  a.fsig = fused ? fused->sigma : nullptr;
  a.fR = fused ? fused->R : 0;
  a.hit = fused ? fused->hit : nullptr;
  a.fpart = fused ? fused->part : nullptr;
  a.fyoff = fused ? fused->yoff : nullptr;
  a.nb_var = sp->nvar;
  a.nk_var = kp->nvar;
  a.bvar = sp->d_var;
  const int64_t chunk = 1 << 20;
  // the list holds the tiny class's pairs [0, n_tiny), then the small LDS
  // class's (n_small), the middle class's (n_mid) and the full class's
  const int64_t small0 = pl->n_tiny, mid0 = small0 + pl->n_small, full0 = mid0 + pl->n_mid;
  for (int64_t base = 0; base < pl->n_pair;) {
    a.pair_base = base;
    const int64_t lim = base < small0 ? small0 : base < mid0 ? mid0 : base < full0 ? full0 : pl->n_pair;
    const int64_t end = std::min(base + chunk, lim);
    const int64_t n = end - base;
    if (base < small0)
      hipLaunchKernelGGL((k_build_blocks<FCG_LDS_X_TINY, FCG_LDS_X_TINY, FCG_LDS_T_TINY, FCG_LDS_S_TINY, 3, BT_TINY,
                                         KPT_TINY>),
                         dim3((unsigned)n), dim3(BT_TINY), 0, s, a);
    else if (base < mid0)
      hipLaunchKernelGGL((k_build_blocks<FCG_LDS_X_SMALL, FCG_LDS_X_SMALL, FCG_LDS_T_SMALL, FCG_LDS_S_SMALL, 3, BT,
                                         KPT>),
                         dim3((unsigned)n), dim3(BT), 0, s, a);
    else if (base < full0)
      hipLaunchKernelGGL((k_build_blocks<FCG_LDS_X_MID, FCG_LDS_G_MID, FCG_LDS_T_MID, FCG_LDS_S_MID, 2, BT, KPT>),
                         dim3((unsigned)n), dim3(BT), 0, s, a);
    else
      hipLaunchKernelGGL((k_build_blocks<FCG_LDS_X, FCG_LDS_X, FCG_LDS_T, FCG_LDS_S, 1, BT, KPT>),
                         dim3((unsigned)n), dim3(BT), 0, s, a);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    base = end;
  }
  return hipSuccess;
}

// sigma[bra var, r] += sum over the bra box's pairs (pair order, i.e. ket box
// order) of their fused products; one workgroup per bra box
__global__ __launch_bounds__(256) void k_fused_sum(const int64_t *__restrict__ row_ptr,
                                                   const int64_t *__restrict__ yoff,
                                                   const int64_t *__restrict__ bvar,
                                                   const double *__restrict__ part, int fR,
                                                   int64_t nb_var, double *__restrict__ sigma) {
  const int64_t B = blockIdx.x;
  const int64_t v0 = bvar[B];
  const int nY = (int)(bvar[B + 1] - v0) * fR;
  const int64_t q0 = row_ptr[B], q1 = row_ptr[B + 1];
  for (int e = threadIdx.x; e < nY; e += blockDim.x) {
    double y = 0.0;
    for (int64_t q = q0; q < q1; q++) y += part[yoff[q] + e];
    sigma[(int64_t)(e % fR) * nb_var + v0 + e / fR] += y;
  }
}

hipError_t fcg_launch_fused_sum(const fcg_plan *pl, const FusedSigma *f, const int64_t *row_ptr,
                                hipStream_t s) {
  const fcg_space *sp = pl->sp;
  if (sp->n_box == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fused_sum, dim3((unsigned)sp->n_box), dim3(128), 0, s, row_ptr, f->yoff,
                     sp->d_var, f->part, f->R, sp->nvar, f->sigma);
  return hipGetLastError();
}

hipError_t fcg_launch_convert(const fcg_plan *pl, hipStream_t s) {
  if (!pl->d_K32 && !pl->d_Kbf && !pl->d_Kb16) return hipSuccess;
  hipLaunchKernelGGL(k_convert, dim3(4096), dim3(256), 0, s, pl->n_kel, pl->d_K64, pl->d_K32,
                     pl->d_Kbf, pl->d_Kb16);
  return hipGetLastError();
}

template <typename TK, typename TA, typename TC, int R, int V>
void launch_multi_v(const fcg_plan *pl, const TK *K, const TC *c, double *sigma, hipStream_t s) {
  constexpr int W = kMulti[V].wpb, U = kMulti[V].unroll;
  constexpr bool NT = kMulti[V].nt;
  hipLaunchKernelGGL((k_apply<TK, TA, TC, R, W, U, apply_clamp<TK>(), NT>),
                     dim3((unsigned)((pl->n_multi + W - 1) / W)), dim3(64 * W), 0, s, pl->n_multi,
                     pl->d_order, pl->sp->d_var, pl->d_row_koff, pl->d_col_ptr, pl->d_colidx, K, c,
                     pl->sp->nvar, sigma);
}

template <typename TK, typename TA, typename TC, int R, int V>
void launch_rows_v(const fcg_plan *pl, const TK *K, const TC *c, double *sigma, hipStream_t s) {
  hipLaunchKernelGGL((k_apply_rows<TK, TA, TC, R, kRows[V].unroll, kRows[V].nt>),
                     dim3((unsigned)(pl->n_single_slots / FCG_CSR_WPB)), dim3(64 * FCG_CSR_WPB), 0,
                     s, pl->n_single_slots, pl->d_order + pl->n_multi, pl->sp->d_var,
                     pl->d_row_koff, pl->d_col_ptr, pl->d_colidx, K, c, pl->sp->nvar, sigma);
}

// parts: bit 0 multi-row boxes, bit 1 one-row boxes
template <typename TK, typename TA, typename TC, int R>
void launch_r(const fcg_plan *pl, int vm, int vr, int parts, const TK *K, const TC *c,
              double *sigma, hipStream_t s) {
  if constexpr (sizeof(TC) == 2) {
    // bf16 c is gathered by the one-row apply only (multi-row boxes stage an
    // FP32 c, launched separately)
  } else if ((parts & 1) && pl->n_multi > 0) {
    if constexpr (R <= TUNED_MAX_R) {
      if (vm == 0) launch_multi_v<TK, TA, TC, R, 0>(pl, K, c, sigma, s);
      else if (vm == 1) launch_multi_v<TK, TA, TC, R, 1>(pl, K, c, sigma, s);
      else if (vm == 2) launch_multi_v<TK, TA, TC, R, 2>(pl, K, c, sigma, s);
      else if (vm == 3) launch_multi_v<TK, TA, TC, R, 3>(pl, K, c, sigma, s);
      else launch_multi_v<TK, TA, TC, R, 4>(pl, K, c, sigma, s);
    } else {
      launch_multi_v<TK, TA, TC, R, default_multi<TK>()>(pl, K, c, sigma, s);
    }
  }
  if ((parts & 2) && pl->n_single_slots > 0) {
    if constexpr (R <= TUNED_MAX_R) {
      if (vr == 0) launch_rows_v<TK, TA, TC, R, 0>(pl, K, c, sigma, s);
      else if (vr == 1) launch_rows_v<TK, TA, TC, R, 1>(pl, K, c, sigma, s);
      else if (vr == 2) launch_rows_v<TK, TA, TC, R, 2>(pl, K, c, sigma, s);
      else launch_rows_v<TK, TA, TC, R, 3>(pl, K, c, sigma, s);
    } else {
      launch_rows_v<TK, TA, TC, R, 0>(pl, K, c, sigma, s);
    }
  }
}

template <typename TK, typename TA, typename TC>
hipError_t launch_by_r(const fcg_plan *pl, int R, int vm, int vr, int parts, const TK *K,
                       const TC *c, double *sigma, hipStream_t s) {
  switch (R) {
#define CASE(RR) \
  case RR: launch_r<TK, TA, TC, RR>(pl, vm, vr, parts, K, c, sigma, s); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
    CASE(9) CASE(10) CASE(11) CASE(12) CASE(13) CASE(14) CASE(15) CASE(16)
#undef CASE
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

int fcg_apply_default_variant(int precision) {
  return precision == FCG_FP64 ? default_multi<double>() : default_multi<float>();
}
int fcg_apply_n_variants(int *n_multi, int *n_rows, int *max_r) {
  *n_multi = N_MULTI_VARIANTS;
  *n_rows = N_ROWS_VARIANTS;
  *max_r = TUNED_MAX_R;
  return 0;
}

hipError_t fcg_launch_apply(const fcg_plan *pl, int R, const double *c, double *sigma,
                            int precision, hipStream_t s, hipEvent_t before_apply, int vm, int vr,
                            void *ci_buf, int parts) {
  const fcg_space *sp = pl->sp;
  if (sp->n_box == 0) {
    if (before_apply) (void)hipEventRecord(before_apply, s);
    return hipSuccess;
  }
  const int64_t nk = pl->ksp->nvar;
  if (precision == FCG_FP64) {
    double *ci = static_cast<double *>(ci_buf);
    hipLaunchKernelGGL(k_interleave<double>, dim3(1024), dim3(256), 0, s, nk, R, c, ci);
    if (before_apply) (void)hipEventRecord(before_apply, s);
    return launch_by_r<double, double, double>(pl, R, vm, vr, parts, pl->d_K64, ci, sigma, s);
  }
  if (precision == FCG_FP32_BF16C) {
    // FP32 blocks; the one-row apply gathers c rounded to bf16, multi-row
    // boxes stage an FP32 c.  ci_buf holds R * nk doubles: the FP32 copy
    // (when multi-row boxes run) at its start, the bf16 copy after it at an
    // 8-byte boundary (it fits for R * nk >= 2; a one-value c with both kinds
    // of boxes takes the FP32 copy for both)
    const int64_t n = nk * R;
    const bool multi = (parts & 1) && pl->n_multi > 0;
    const bool rows = (parts & 2) && pl->n_single_slots > 0;
    char *base = static_cast<char *>(ci_buf);
    float *cf = reinterpret_cast<float *>(base);
    const int64_t off = multi ? (4 * n + 7) / 8 * 8 : 0;
    bf16c *cb = rows && off + 2 * n <= 8 * n ? reinterpret_cast<bf16c *>(base + off) : nullptr;
    if (multi || (rows && !cb))
      hipLaunchKernelGGL(k_interleave<float>, dim3(1024), dim3(256), 0, s, nk, R, c, cf);
    if (cb) hipLaunchKernelGGL(k_interleave_bf16, dim3(1024), dim3(256), 0, s, nk, R, c, cb);
    if (before_apply) (void)hipEventRecord(before_apply, s);
    hipError_t e = hipSuccess;
    if (multi) e = launch_by_r<float, float, float>(pl, R, vm, vr, 1, pl->d_K32, cf, sigma, s);
    if (e == hipSuccess && rows)
      e = cb ? launch_by_r<float, float, bf16c>(pl, R, vm, vr, 2, pl->d_K32, cb, sigma, s)
             : launch_by_r<float, float, float>(pl, R, vm, vr, 2, pl->d_K32, cf, sigma, s);
    return e;
  }
  float *ci = static_cast<float *>(ci_buf);
  hipLaunchKernelGGL(k_interleave<float>, dim3(1024), dim3(256), 0, s, nk, R, c, ci);
  if (before_apply) (void)hipEventRecord(before_apply, s);
  if (precision == FCG_FP32)
    return launch_by_r<float, float, float>(pl, R, vm, vr, parts, pl->d_K32, ci, sigma, s);
  if (precision == FCG_BF16)
    return launch_by_r<uint16_t, float, float>(pl, R, vm, vr, parts, pl->d_Kb16, ci, sigma, s);
  return launch_by_r<uint32_t, float, float>(pl, R, vm, vr, parts, pl->d_Kbf, ci, sigma, s);
}

hipError_t fcg_launch_diag(const fcg_plan *pl, double *hd, hipStream_t s) {
  const fcg_space *sp = pl->sp;
  hipLaunchKernelGGL(k_diag, dim3((unsigned)((sp->n_box + 255) / 256)), dim3(256), 0, s,
                     sp->n_box, sp->d_var, pl->d_diag_koff, pl->d_diag_col, pl->d_K64, hd);
  return hipGetLastError();
}
