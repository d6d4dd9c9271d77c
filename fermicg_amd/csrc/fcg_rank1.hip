// Matrix-free sigma for TPSCI spaces of rank <= 1.
//
// A configuration of rank <= 1 has every cluster in the lowest state of its
// sector except at most one (the spaces the large benchmark systems use:
// 9.0e6 configurations at 12 clusters of M=32, whose stored Hamiltonian
// would hold 2.4e11-4.3e11 elements).  For such spaces the reference's
// tps_ci_matvec (src/tpsci_outer.jl:462-527, contract_matrix_element
// src/tpsci_inner.jl:10-216) factorises.  Per bra Fock block F_b, ket Fock
// block F_k = F_b - Δ and term t on active clusters A with integrals V and
// sign s (compute_terms_state_sign, src/type_ClusteredTerm.jl:181-197),
// writing g_a = Γ_a[:, 0, 0] (state 0 = the sector's lowest state):
//
//   e      = sum_p V[p] prod_a g_a[p_a]                          (ref <- ref)
//   W_a    = V contracted with g_b on every b != a               (length n_a)
//   W_ab   = V contracted with g_c on every c not in {a, b}      (n_a x n_b)
//   u_b    = sum_{j>=1} Γ_b[:, 0, j] c_(b,j)                      (length n_b)
//
//   y_ref   += s (e c_ref + sum_b W_b . u_b)
//   y_(a,i) += s (sum_j (W_a . Γ_a[:, i, j]) c_(a,j) [j = 0: c_ref]
//                 + sum_{b != a} Γ_a[:, i, 0]' W_ab u_b)          (a in A, i >= 1)
//   y_(x,i) += s e c_(x,i)                                        (x not in A)
//
// -- every other (bra, ket) pair breaks check_term (src/type_ClusteredTerm.jl:
// 92-176) on an inactive cluster.  Vectors live in a layout that holds every
// rank <= 1 configuration of the space's Fock blocks (ref, then cluster by
// cluster states 1..M-1); the caller's configurations are scattered into it
// and gathered back, so a subset of the rank <= 1 configurations works too.
//
// One workgroup per bra Fock block owns that block's rows of sigma in LDS;
// each output value is updated by one thread in a fixed term order, so the
// result is bit-reproducible.  T = double (FP64) or float (the mixed form:
// FP32 arithmetic on an FP32 copy of the Γ tensors and integrals).
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fcg_internal.h"

struct R1Term {
  int32_t nbody;
  int32_t cl[4];
  int32_t op[4];
  int32_t dims[4];
  int32_t odd;
  int32_t pad;
  int64_t ioff;
};

struct fcg_rank1 {
  int N = 0, S = 0, n_ops = 0, nbody = 4;
  int64_t n_fock = 0, n_lay = 0, n_var = 0, n_pairs = 0;
  int lmax = 0;            // largest block layout
  int64_t *d_foff = nullptr;   // [n_fock+1]
  int32_t *d_meta = nullptr;   // [n_fock][2N]: M_c then slot offset of cluster c
  int16_t *d_fock = nullptr;   // [n_fock][N][2]
  int64_t *d_fp_ptr = nullptr;
  int32_t *d_fp_fk = nullptr, *d_fp_g = nullptr;
  int32_t *d_tcs = nullptr;
  R1Term *d_terms = nullptr;
  int32_t *d_gblock = nullptr;
  int64_t *d_perm = nullptr;   // [n_var] caller's configuration -> layout slot
  float *d_g32 = nullptr;      // FP32 copy of the Γ data
  float *d_i32 = nullptr;      // FP32 copy of the integrals
  unsigned int *d_err = nullptr;
  uint8_t *d_nzb = nullptr;    // [n_fock] per-call flags of nonzero ket blocks
  bool checked = false;
};

int fcg_rank1_check(fcg_ctx *ctx, fcg_plan *pl);

namespace {

constexpr int RT = 256;      // threads per workgroup
constexpr int RC = 4;        // roots per launch
constexpr int LMAX = 512;    // layout slots of one Fock block
constexpr int VMX = 1024;    // staged integrals
constexpr int NPX = 64;      // largest leading dimension of a Γ block

template <typename T>
struct R1Args {
  int N, S, n_ops, nbody, R;
  int64_t fb_base, n_lay;
  const int64_t *foff;
  const int32_t *meta;
  const int16_t *fock;
  const int64_t *fp_ptr;
  const int32_t *fp_fk, *fp_g;
  const int32_t *tc_start;
  const R1Term *terms;
  const T *ints;
  const int32_t *gblock;
  const int64_t *gb_off;
  const int32_t *gb_dims;
  const T *gdata;
  const T *x;  // [R][n_lay] (R <= RC in this launch)
  T *y;        // [R][n_lay]
  double *diag;  // DIAG: [n_lay]
  const uint8_t *nzb;  // [n_fock] ket blocks with a nonzero coefficient (others skipped)
  unsigned int *err;
};

// decode column-major index v over nb dims into p[0..nb)
__device__ __forceinline__ void decode(int v, int nb, const int32_t *dims, int *p) {
  for (int q = 0; q < nb; q++) {
    p[q] = v % dims[q];
    v /= dims[q];
  }
}

template <typename T, bool DIAG>
__global__ __launch_bounds__(RT) void k_rank1(R1Args<T> a) {
  __shared__ T sY[LMAX * RC];
  __shared__ T sX[LMAX * RC];
  __shared__ T sV[VMX];
  __shared__ T sG0[4][NPX];     // Γ_q[:, 0, 0]
  __shared__ T sWa[4][NPX];     // W_a
  __shared__ T sWab[6][256];    // W_ab of the 6 cluster pairs, n_a * n_b <= 256
  __shared__ T sU[4][NPX * RC];  // u_b[p][r]
  __shared__ T sE;
  __shared__ T sRed[RT];
  __shared__ int sMb[FCG_MAXN], sOb[FCG_MAXN], sMk[FCG_MAXN], sOk[FCG_MAXN];
  __shared__ int sBlk[4], sNp[4], sMbq[4], sMkq[4];
  __shared__ int sSlotA[4];   // layout offset of each active cluster's states in F_b
  __shared__ int sLb, sLk;
  __shared__ int64_t sFk0;

  const int tid = threadIdx.x;
  const int64_t fb = a.fb_base + blockIdx.x;
  const int N = a.N, S = a.S, R = a.R;
  const int16_t *Fb = a.fock + fb * 2 * N;
  if (tid < N) {
    sMb[tid] = a.meta[fb * 2 * N + tid];
    sOb[tid] = a.meta[fb * 2 * N + N + tid];
  }
  if (tid == 0) sLb = (int)(a.foff[fb + 1] - a.foff[fb]);
  __syncthreads();
  const int Lb = sLb;
  for (int e = tid; e < Lb * RC; e += RT) sY[e] = (T)0;
  for (int64_t q = a.fp_ptr[fb]; q < a.fp_ptr[fb + 1]; q++) {
    const int64_t fk = a.fp_fk[q];
    if (DIAG && fk != fb) continue;  // only the Δ = 0 pair reaches the diagonal
    if (!DIAG && !a.nzb[fk]) continue;  // a ket block of zeros contributes nothing
    const int g = a.fp_g[q];
    const int16_t *Fk = a.fock + fk * 2 * N;
    __syncthreads();
    if (tid < N) {
      sMk[tid] = a.meta[fk * 2 * N + tid];
      sOk[tid] = a.meta[fk * 2 * N + N + tid];
    }
    if (tid == 0) {
      sLk = (int)(a.foff[fk + 1] - a.foff[fk]);
      sFk0 = a.foff[fk];
    }
    __syncthreads();
    const int Lk = sLk;
    if (!DIAG)
      for (int e = tid; e < Lk * RC; e += RT) {
        const int i = e % Lk, r = e / Lk;
        sX[e] = r < R ? a.x[(int64_t)r * a.n_lay + sFk0 + i] : (T)0;
      }
    for (int t = a.tc_start[g]; t < a.tc_start[g + 1]; t++) {
      const R1Term Tm = a.terms[t];
      const int nb = Tm.nbody;
      if (nb > a.nbody) continue;
      __syncthreads();
      if (tid == 0) {
        bool ok = true;
        for (int qq = 0; qq < nb; qq++) {
          const int c = Tm.cl[qq];
          const int cb = Fb[2 * c] * S + Fb[2 * c + 1], ck = Fk[2 * c] * S + Fk[2 * c + 1];
          const int SS = S * S;
          const int blk = a.gblock[(((int64_t)c * a.n_ops + Tm.op[qq]) * SS + cb) * SS + ck];
          sBlk[qq] = blk;
          if (blk < 0) {
            ok = false;
            continue;
          }
          sNp[qq] = a.gb_dims[3 * blk];
          sMbq[qq] = a.gb_dims[3 * blk + 1];
          sMkq[qq] = a.gb_dims[3 * blk + 2];
          sSlotA[qq] = sOb[c];
          ok = ok && sNp[qq] <= NPX && sMbq[qq] == sMb[c] && sMkq[qq] == sMk[c];
        }
        if (!ok) {
          atomicOr(a.err, 1u);
          sBlk[0] = -1;
        }
      }
      __syncthreads();
      if (sBlk[0] < 0) continue;  // uniform; reported through err
      // sign from the ket Fock configuration
      T sg = (T)1;
      {
        int par = 0;
        for (int qq = 0; qq < nb; qq++) {
          if (!((Tm.odd >> qq) & 1)) continue;
          int ne = 0;
          for (int c2 = 0; c2 < Tm.cl[qq]; c2++) ne += Fk[2 * c2] + Fk[2 * c2 + 1];
          par ^= ne & 1;
        }
        if (par) sg = (T)-1;
      }
      int nv = 1;
      if (nb > 1)
        for (int qq = 0; qq < nb; qq++) nv *= Tm.dims[qq];
      for (int e = tid; e < nv; e += RT) sV[e] = nb > 1 ? a.ints[Tm.ioff + e] : (T)1;
      for (int e = tid; e < 4 * NPX; e += RT) {
        const int qq = e / NPX, p = e % NPX;
        if (qq < nb && p < sNp[qq]) sG0[qq][p] = a.gdata[a.gb_off[sBlk[qq]] + p];
      }
      __syncthreads();
      // e: fixed-order strided partial sums, then a fixed tree
      {
        T part = (T)0;
        for (int v = tid; v < nv; v += RT) {
          int p[4];
          decode(v, nb, Tm.dims, p);
          T f = sV[v];
          for (int qq = 0; qq < nb; qq++) f *= sG0[qq][nb > 1 ? p[qq] : 0];
          part += f;
        }
        sRed[tid] = part;
        __syncthreads();
        for (int s2 = RT / 2; s2 > 0; s2 >>= 1) {
          if (tid < s2) sRed[tid] += sRed[tid + s2];
          __syncthreads();
        }
        if (tid == 0) sE = sRed[0];
      }
      // W_a (one thread per entry) and W_ab (one thread per entry), each a
      // fixed-order sum over the indices of the other clusters
      int strd[4];
      {
        int st = 1;
        for (int qq = 0; qq < nb; qq++) {
          strd[qq] = st;
          st *= nb > 1 ? Tm.dims[qq] : 1;
        }
      }
      for (int e = tid; e < 4 * NPX; e += RT) {
        const int qq = e / NPX, pq = e % NPX;
        if (qq >= nb || pq >= sNp[qq]) continue;
        T s = (T)0;
        if (nb == 1) {
          s = (T)1;
        } else {
          const int nrest = nv / Tm.dims[qq];
          for (int m = 0; m < nrest; m++) {
            int rem = m, v = pq * strd[qq];
            T f = (T)1;
            for (int r2 = 0; r2 < nb; r2++) {
              if (r2 == qq) continue;
              const int pr2 = rem % Tm.dims[r2];
              rem /= Tm.dims[r2];
              v += pr2 * strd[r2];
              f *= sG0[r2][pr2];
            }
            s += sV[v] * f;
          }
        }
        sWa[qq][pq] = s;
      }
      if (!DIAG && nb > 1) {
        int pr = 0;
        for (int qa = 0; qa < nb; qa++)
          for (int qb = qa + 1; qb < nb; qb++, pr++) {
            const int na = sNp[qa], nbb = sNp[qb];
            const int nrest = nv / (Tm.dims[qa] * Tm.dims[qb]);
            for (int e = tid; e < na * nbb; e += RT) {
              const int pa = e % na, pb = e / na;
              T s = (T)0;
              for (int m = 0; m < nrest; m++) {
                int rem = m, v = pa * strd[qa] + pb * strd[qb];
                T f = (T)1;
                for (int r2 = 0; r2 < nb; r2++) {
                  if (r2 == qa || r2 == qb) continue;
                  const int pr2 = rem % Tm.dims[r2];
                  rem /= Tm.dims[r2];
                  v += pr2 * strd[r2];
                  f *= sG0[r2][pr2];
                }
                s += sV[v] * f;
              }
              sWab[pr][e] = s;
            }
          }
      }
      // u_b[p][r] = sum_{j>=1} Γ_b[p, 0, j] x_(b,j)
      if (!DIAG)
        for (int e = tid; e < 4 * NPX * RC; e += RT) {
          const int qq = e / (NPX * RC), rem = e % (NPX * RC), p = rem % NPX, r = rem / NPX;
          if (qq >= nb || p >= sNp[qq]) continue;
          const int c = Tm.cl[qq], np = sNp[qq], mb = sMbq[qq], mk = sMkq[qq];
          const T *G = a.gdata + a.gb_off[sBlk[qq]];
          T s = (T)0;
          for (int j = 1; j < mk; j++) s += G[p + (int64_t)np * (int64_t)mb * j] * sX[(sOk[c] + j - 1) + Lk * r];
          sU[qq][p + NPX * r] = s;
        }
      __syncthreads();
      const T e0 = sE * sg;
      // updates: ref (threads 0..RC-1), active clusters' states, the rest
      if (tid < RC) {
        const int r = tid;
        if (DIAG) {
          if (r == 0) sY[0] += e0;
        } else {
          T v = sE * sX[0 + Lk * r];
          for (int qq = 0; qq < nb; qq++)
            for (int p = 0; p < sNp[qq]; p++) v += sWa[qq][p] * sU[qq][p + NPX * r];
          sY[0 + Lb * r] += sg * v;
        }
      }
      // active cluster qq, bra state i >= 1: thread per (qq, i)
      for (int e = tid; e < 4 * LMAX; e += RT) {
        const int qq = e / LMAX, i = e % LMAX;
        if (qq >= nb || i < 1 || i >= sMbq[qq]) continue;
        const int c = Tm.cl[qq], np = sNp[qq], mb = sMbq[qq], mk = sMkq[qq];
        const T *G = a.gdata + a.gb_off[sBlk[qq]];
        const int slot = sOb[c] + i - 1;
        if (DIAG) {
          // I = J = (c, i): W_a . Γ_a[:, i, i]
          T d = (T)0;
          const T *Gij = G + (int64_t)np * (i + (int64_t)mb * i);
          for (int p = 0; p < np; p++) d += sWa[qq][p] * Gij[p];
          sY[slot] += sg * d;
          continue;
        }
        T acc[RC];
        for (int r = 0; r < RC; r++) acc[r] = (T)0;
        for (int j = 0; j < mk; j++) {
          const T *Gij = G + (int64_t)np * (i + (int64_t)mb * j);
          T w = (T)0;
          for (int p = 0; p < np; p++) w += sWa[qq][p] * Gij[p];
          const int xs = j == 0 ? 0 : sOk[c] + j - 1;
          for (int r = 0; r < RC; r++) acc[r] += w * sX[xs + Lk * r];
        }
        // cross terms: Γ_a[:, i, 0]' W_ab u_b over the other active clusters
        if (nb > 1) {
          const T *Gi0 = G + (int64_t)np * i;
          int pr = 0;
          for (int qa = 0; qa < nb; qa++)
            for (int qb = qa + 1; qb < nb; qb++, pr++) {
              if (qa != qq && qb != qq) continue;
              const int other = qa == qq ? qb : qa;
              const int na = sNp[qa];
              const int no = sNp[other];
              for (int r = 0; r < RC; r++) {
                T z = (T)0;
                for (int pa = 0; pa < np; pa++) {
                  T t2 = (T)0;
                  for (int po = 0; po < no; po++) {
                    // W_ab stored as [p_qa + n_qa * p_qb]
                    const T w = qa == qq ? sWab[pr][pa + na * po] : sWab[pr][po + na * pa];
                    t2 += w * sU[other][po + NPX * r];
                  }
                  z += Gi0[pa] * t2;
                }
                acc[r] += z;
              }
            }
        }
        for (int r = 0; r < RC; r++) sY[slot + Lb * r] += sg * acc[r];
      }
      // clusters outside the term: (x, i) <- (x, i) with e
      for (int e = tid; e < Lb - 1; e += RT) {
        const int slot = e + 1;
        // which cluster holds this slot
        int x = 0;
        while (x + 1 < N && sOb[x + 1] <= slot) x++;
        bool act = false;
        for (int qq = 0; qq < nb; qq++) act |= Tm.cl[qq] == x;
        if (act) continue;
        if (DIAG) {
          sY[slot] += e0;
        } else {
          const int i = slot - sOb[x] + 1;  // state of cluster x (>= 1)
          const int xs = sOk[x] + i - 1;    // the same state in the ket block
          for (int r = 0; r < RC; r++) sY[slot + Lb * r] += e0 * sX[xs + Lk * r];
        }
      }
    }
  }
  __syncthreads();
  const int64_t o0 = a.foff[fb];
  if (DIAG) {
    for (int i = tid; i < Lb; i += RT) a.diag[o0 + i] = (double)sY[i];
  } else {
    for (int e = tid; e < Lb * R; e += RT) {
      const int i = e % Lb, r = e / Lb;
      a.y[(int64_t)r * a.n_lay + o0 + i] = sY[i + Lb * r];
    }
  }
}

// x_lay[r][perm[i]] = c[r][i] (layout zeroed first); sigma[r][i] = y_lay[r][perm[i]]
template <typename T>
__global__ void k_r1_scatter(int64_t n, int R, int64_t n_lay, const int64_t *perm, const double *c,
                             T *x) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, r = e / n;
    x[r * n_lay + perm[i]] = (T)c[e];
  }
}

template <typename T>
__global__ void k_r1_gather(int64_t n, int R, int64_t n_lay, const int64_t *perm, const T *y,
                            double *s) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * R;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e % n, r = e / n;
    s[e] = (double)y[r * n_lay + perm[i]];
  }
}

__global__ void k_r1_gather_diag(int64_t n, const int64_t *perm, const double *d, double *hd) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    hd[i] = d[perm[i]];
}

// nzb[f] = 1 when some root's coefficient in Fock block f's layout is nonzero
template <typename T>
__global__ void k_r1_nonzero(int64_t n_fock, int R, int64_t n_lay, const int64_t *foff,
                             const T *x, uint8_t *nzb) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n_fock) return;
  uint8_t nz = 0;
  for (int r = 0; r < R && !nz; r++)
    for (int64_t i = foff[f]; i < foff[f + 1]; i++)
      if (x[(int64_t)r * n_lay + i] != (T)0) {
        nz = 1;
        break;
      }
  nzb[f] = nz;
}

__global__ void k_to_float(int64_t n, const double *a, float *b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

__global__ void k_zero(int64_t n, double *p) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

constexpr int64_t FB_CHUNK = 65535;

template <typename T>
int launch_rank1(fcg_ctx *ctx, fcg_plan *pl, R1Args<T> a, bool diag) {
  const fcg_rank1 *r1 = pl->r1;
  for (int64_t f0 = 0; f0 < r1->n_fock; f0 += FB_CHUNK) {
    a.fb_base = f0;
    const unsigned nb = (unsigned)std::min(FB_CHUNK, r1->n_fock - f0);
    if (diag)
      hipLaunchKernelGGL((k_rank1<T, true>), dim3(nb), dim3(RT), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((k_rank1<T, false>), dim3(nb), dim3(RT), 0, ctx->stream, a);
  }
  FCG_HIP(hipGetLastError(), ctx);
  return FCG_OK;
}

template <typename T>
R1Args<T> base_args(const fcg_plan *pl, const fcg_ham *h, const fcg_ops *o) {
  const fcg_rank1 *r1 = pl->r1;
  R1Args<T> a{};
  a.N = r1->N;
  a.S = r1->S;
  a.n_ops = r1->n_ops;
  a.nbody = r1->nbody;
  a.n_lay = r1->n_lay;
  a.foff = r1->d_foff;
  a.meta = r1->d_meta;
  a.fock = r1->d_fock;
  a.fp_ptr = r1->d_fp_ptr;
  a.fp_fk = r1->d_fp_fk;
  a.fp_g = r1->d_fp_g;
  a.tc_start = r1->d_tcs;
  a.terms = r1->d_terms;
  a.gblock = r1->d_gblock;
  a.gb_off = o->d_gb_off;
  a.gb_dims = o->d_gb_dims;
  a.err = r1->d_err;
  a.nzb = r1->d_nzb;
  return a;
}

}  // namespace

void fcg_rank1_free(fcg_ctx *ctx, fcg_plan *pl) {
  fcg_rank1 *r = pl->r1;
  if (!r) return;
  fcg_dfree(ctx, r->d_foff);
  fcg_dfree(ctx, r->d_meta);
  fcg_dfree(ctx, r->d_fock);
  fcg_dfree(ctx, r->d_fp_ptr);
  fcg_dfree(ctx, r->d_fp_fk);
  fcg_dfree(ctx, r->d_fp_g);
  fcg_dfree(ctx, r->d_tcs);
  fcg_dfree(ctx, r->d_terms);
  fcg_dfree(ctx, r->d_gblock);
  fcg_dfree(ctx, r->d_perm);
  fcg_dfree(ctx, r->d_g32);
  fcg_dfree(ctx, r->d_i32);
  fcg_dfree(ctx, r->d_err);
  fcg_dfree(ctx, r->d_nzb);
  delete r;
  pl->r1 = nullptr;
}

namespace {
int ensure_r1_buffers(fcg_ctx *ctx, int64_t n) {
  if (n <= ctx->r1_elems) return FCG_OK;
  if (ctx->d_r1x) fcg_dfree(ctx, ctx->d_r1x);
  if (ctx->d_r1y) fcg_dfree(ctx, ctx->d_r1y);
  ctx->d_r1x = static_cast<double *>(fcg_dmalloc(ctx, sizeof(double) * std::max<int64_t>(n, 1)));
  ctx->d_r1y = static_cast<double *>(fcg_dmalloc(ctx, sizeof(double) * std::max<int64_t>(n, 1)));
  ctx->r1_elems = 0;
  if (!ctx->d_r1x || !ctx->d_r1y) return FCG_ENOMEM;
  ctx->r1_elems = n;
  return FCG_OK;
}
}  // namespace

int fcg_rank1_sigma(fcg_ctx *ctx, fcg_plan *pl, int R, const double *c, double *sigma,
                    int precision) {
  fcg_rank1 *r1 = pl->r1;
  const fcg_ham *h = pl->r1_ham;
  const fcg_ops *o = pl->r1_ops;
  const int64_t L = r1->n_lay, n = r1->n_var;
  if (int rc = ensure_r1_buffers(ctx, L * R)) {
    ctx->err = "matrix-free sigma buffers allocation failed";
    return rc;
  }
  hipStream_t s = ctx->stream;
  const bool f32 = precision != FCG_FP64;
  // layout vectors (the slots no configuration maps to stay zero)
  hipLaunchKernelGGL(k_zero, dim3(1024), dim3(256), 0, s, L * R, ctx->d_r1x);
  if (f32)
    hipLaunchKernelGGL(k_r1_scatter<float>, dim3(1024), dim3(256), 0, s, n, R, L, r1->d_perm, c,
                       reinterpret_cast<float *>(ctx->d_r1x));
  else
    hipLaunchKernelGGL(k_r1_scatter<double>, dim3(1024), dim3(256), 0, s, n, R, L, r1->d_perm, c,
                       ctx->d_r1x);
  const unsigned nfb = (unsigned)((r1->n_fock + 255) / 256);
  if (f32)
    hipLaunchKernelGGL(k_r1_nonzero<float>, dim3(nfb), dim3(256), 0, s, r1->n_fock, R, L,
                       r1->d_foff, reinterpret_cast<const float *>(ctx->d_r1x), r1->d_nzb);
  else
    hipLaunchKernelGGL(k_r1_nonzero<double>, dim3(nfb), dim3(256), 0, s, r1->n_fock, R, L,
                       r1->d_foff, ctx->d_r1x, r1->d_nzb);
  FCG_HIP(hipEventRecord(ctx->ev0, s), ctx);
  for (int r0 = 0; r0 < R; r0 += RC) {
    const int rn = std::min(RC, R - r0);
    if (f32) {
      R1Args<float> a = base_args<float>(pl, h, o);
      a.R = rn;
      a.ints = r1->d_i32;
      a.gdata = r1->d_g32;
      a.x = reinterpret_cast<const float *>(ctx->d_r1x) + (int64_t)r0 * L;
      a.y = reinterpret_cast<float *>(ctx->d_r1y) + (int64_t)r0 * L;
      if (int rc = launch_rank1<float>(ctx, pl, a, false)) return rc;
    } else {
      R1Args<double> a = base_args<double>(pl, h, o);
      a.R = rn;
      a.ints = h->d_ints;
      a.gdata = o->d_gdata;
      a.x = ctx->d_r1x + (int64_t)r0 * L;
      a.y = ctx->d_r1y + (int64_t)r0 * L;
      if (int rc = launch_rank1<double>(ctx, pl, a, false)) return rc;
    }
  }
  FCG_HIP(hipEventRecord(ctx->ev1, s), ctx);
  if (!r1->checked) {  // the first sigma visits every term: missing Γ blocks show here
    if (int rc = fcg_rank1_check(ctx, pl)) return rc;
    r1->checked = true;
  }
  if (f32)
    hipLaunchKernelGGL(k_r1_gather<float>, dim3(1024), dim3(256), 0, s, n, R, L, r1->d_perm,
                       reinterpret_cast<const float *>(ctx->d_r1y), sigma);
  else
    hipLaunchKernelGGL(k_r1_gather<double>, dim3(1024), dim3(256), 0, s, n, R, L, r1->d_perm,
                       ctx->d_r1y, sigma);
  FCG_HIP(hipGetLastError(), ctx);
  return FCG_OK;
}

int fcg_rank1_diag(fcg_ctx *ctx, fcg_plan *pl, double *hd) {
  fcg_rank1 *r1 = pl->r1;
  const int64_t L = r1->n_lay;
  if (int rc = ensure_r1_buffers(ctx, L)) {
    ctx->err = "matrix-free diagonal buffers allocation failed";
    return rc;
  }
  R1Args<double> a = base_args<double>(pl, pl->r1_ham, pl->r1_ops);
  a.R = 1;
  a.ints = pl->r1_ham->d_ints;
  a.gdata = pl->r1_ops->d_gdata;
  a.diag = ctx->d_r1y;
  if (int rc = launch_rank1<double>(ctx, pl, a, true)) return rc;
  if (int rc = fcg_rank1_check(ctx, pl)) return rc;
  hipLaunchKernelGGL(k_r1_gather_diag, dim3(1024), dim3(256), 0, ctx->stream, r1->n_var,
                     r1->d_perm, ctx->d_r1y, hd);
  FCG_HIP(hipGetLastError(), ctx);
  return FCG_OK;
}

extern "C" int fcg_plan_create_matrix_free(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o,
                                           fcg_space_t sp, int nbody, int precisions,
                                           fcg_plan_t *out) {
  if (!ctx || !h || !o || !sp || !out || nbody < 1 || nbody > 4) return FCG_EINVAL;
  if (!sp->tpsci) {
    ctx->err = "matrix-free plans need a TPSCI space";
    return FCG_EINVAL;
  }
  if (precisions & ~((1 << FCG_FP64) | (1 << FCG_FP32))) {
    ctx->err = "matrix-free plans compute in FP64 and FP32 only";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  if (int rc = fcg_form_boxes(ctx, o, sp)) return rc;
  const int N = sp->N, S = o->S;
  const int64_t n_fock = sp->n_fock, n_var = sp->nvar;
  // layout: per Fock block the ground configuration, then cluster by
  // cluster its states 1..M-1
  std::vector<int64_t> foff(n_fock + 1, 0);
  std::vector<int32_t> meta((size_t)n_fock * 2 * N);
  int lmax = 0;
  for (int64_t f = 0; f < n_fock; f++) {
    int off = 1;
    for (int c = 0; c < N; c++) {
      const int na = sp->fock[(f * N + c) * 2], nb = sp->fock[(f * N + c) * 2 + 1];
      const int m = (na >= 0 && nb >= 0 && na < S && nb < S)
                        ? o->nstates[(int64_t)c * S * S + na * S + nb] : 0;
      if (m < 1) {
        ctx->err = "matrix-free plan: a Fock block's sector has no states";
        return FCG_EINVAL;
      }
      meta[f * 2 * N + c] = m;
      meta[f * 2 * N + N + c] = off;
      off += m - 1;
    }
    foff[f + 1] = foff[f] + off;
    lmax = std::max(lmax, off);
  }
  if (lmax > LMAX) {
    ctx->err = "matrix-free plan: a Fock block holds more than " + std::to_string(LMAX) +
               " rank <= 1 configurations";
    return FCG_EOVERFLOW;
  }
  // the caller's configurations in the layout
  std::vector<int64_t> perm(std::max<int64_t>(n_var, 1));
  std::vector<uint8_t> used(std::max<int64_t>(foff[n_fock], 1), 0);
  for (int64_t b = 0; b < sp->n_box; b++) {
    const int64_t f = sp->box_fock[b];
    const int64_t v0 = sp->var_off[b], d = sp->var_off[b + 1] - v0;
    for (int64_t e = 0; e < d; e++) {
      int64_t r = e;
      int exc = -1, st = 0;
      for (int c = 0; c < N; c++) {
        const int rk = sp->rank[b * N + c];
        const int s0 = (int)(r % rk) + sp->lo[b * N + c] - 1;
        r /= rk;
        if (s0 != 0) {
          if (exc >= 0) {
            ctx->err = "matrix-free plan: a configuration has more than one excited cluster";
            return FCG_EINVAL;
          }
          exc = c;
          st = s0;
        }
      }
      const int64_t slot = foff[f] + (exc < 0 ? 0 : meta[f * 2 * N + N + exc] + st - 1);
      if (exc >= 0 && st >= meta[f * 2 * N + exc]) {
        ctx->err = "matrix-free plan: state outside its sector";
        return FCG_EINVAL;
      }
      if (used[slot]) {
        ctx->err = "matrix-free plan: a configuration appears twice";
        return FCG_EINVAL;
      }
      used[slot] = 1;
      perm[v0 + e] = slot;
    }
  }
  fcg_rank1 *r1 = new fcg_rank1();
  r1->N = N;
  r1->S = S;
  r1->n_ops = o->n_ops;
  r1->nbody = nbody;
  r1->n_fock = n_fock;
  r1->n_lay = foff[n_fock];
  r1->n_var = n_var;
  r1->lmax = lmax;
  std::vector<R1Term> terms(h->n_terms);
  for (int t = 0; t < h->n_terms; t++) {
    R1Term &T = terms[t];
    T = R1Term{};
    T.nbody = h->nbody[t];
    T.ioff = h->ioff[t];
    int nv = 1;
    for (int q = 0; q < 4; q++) {
      T.cl[q] = h->clusters[4 * t + q];
      T.op[q] = h->ops[4 * t + q];
      T.dims[q] = h->dims[4 * t + q];
      if (q < T.nbody && h->parity[4 * t + q] == 1) T.odd |= 1 << q;
      if (q < T.nbody && T.nbody > 1) nv *= T.dims[q];
    }
    if (nv > VMX) {
      delete r1;
      ctx->err = "matrix-free plan: a term's integrals exceed the staged size";
      return FCG_EOVERFLOW;
    }
    if (T.nbody > 1)
      for (int qa = 0; qa < T.nbody; qa++)
        for (int qb = qa + 1; qb < T.nbody; qb++)
          if (T.dims[qa] * T.dims[qb] > 256 || T.dims[qa] > NPX) {
            delete r1;
            ctx->err = "matrix-free plan: a term's index pair exceeds the staged size";
            return FCG_EOVERFLOW;
          }
  }
  {  // every term's Γ blocks shaped as its integrals (the kernel relies on it)
    const int64_t SS = (int64_t)S * S;
    std::vector<uint8_t> seen((size_t)N * o->n_ops, 0);
    for (int t = 0; t < h->n_terms; t++)
      for (int q = 0; q < h->nbody[t]; q++) {
        const int cl = h->clusters[4 * t + q], op = h->ops[4 * t + q];
        if (cl < 0 || cl >= N || op < 0 || op >= o->n_ops) {
          delete r1;
          ctx->err = "matrix-free plan: term cluster or operator out of range";
          return FCG_EINVAL;
        }
        if (seen[(size_t)cl * o->n_ops + op]) continue;
        seen[(size_t)cl * o->n_ops + op] = 1;
        const int want = h->nbody[t] == 1 ? 1 : h->dims[4 * t + q];
        for (int64_t e = 0; e < SS * SS; e++) {
          const int blk = o->gblock[((int64_t)cl * o->n_ops + op) * SS * SS + e];
          if (blk >= 0 && o->gb_dims[3 * blk] != want) {
            delete r1;
            ctx->err = "matrix-free plan: a Γ block's leading dimension differs from its "
                       "term's integrals";
            return FCG_EINVAL;
          }
        }
      }
  }
  fcg_plan *pl = new fcg_plan();
  pl->ctx = ctx;
  pl->sp = pl->ksp = sp;
  pl->r1 = r1;
  pl->r1_ham = h;
  pl->r1_ops = o;
  pl->precisions = precisions | (1 << FCG_FP64);
  pl->holds_blocks = true;  // can apply sigma (matrix-free)
  auto up = [&](auto **dst, const auto *src, int64_t cnt) -> int {
    using T = std::remove_reference_t<decltype(**dst)>;
    *dst = static_cast<T *>(fcg_dmalloc(ctx, sizeof(T) * std::max<int64_t>(cnt, 1)));
    if (!*dst) return FCG_ENOMEM;
    if (cnt > 0 &&
        hipMemcpyAsync(*dst, src, sizeof(T) * cnt, hipMemcpyHostToDevice, ctx->stream) !=
            hipSuccess)
      return FCG_EDEVICE;
    return FCG_OK;
  };
  int rc = up(&r1->d_foff, foff.data(), n_fock + 1);
  if (!rc) rc = up(&r1->d_meta, meta.data(), (int64_t)meta.size());
  if (!rc) rc = up(&r1->d_fock, sp->fock.data(), (int64_t)sp->fock.size());
  if (!rc) rc = up(&r1->d_tcs, h->tc_start.data(), (int64_t)h->tc_start.size());
  if (!rc) rc = up(&r1->d_terms, terms.data(), (int64_t)terms.size());
  if (!rc) rc = up(&r1->d_gblock, o->gblock.data(), (int64_t)o->gblock.size());
  if (!rc) rc = up(&r1->d_perm, perm.data(), n_var);
  if (!rc) {
    r1->d_err = static_cast<unsigned int *>(fcg_dmalloc(ctx, sizeof(unsigned int)));
    r1->d_nzb = static_cast<uint8_t *>(fcg_dmalloc(ctx, std::max<int64_t>(n_fock, 1)));
    rc = r1->d_err && r1->d_nzb ? FCG_OK : FCG_ENOMEM;
  }
  if (!rc && hipMemsetAsync(r1->d_err, 0, sizeof(unsigned int), ctx->stream) != hipSuccess)
    rc = FCG_EDEVICE;
  if (!rc) rc = fcg_device_fock_pairs(ctx, h, sp, nbody, S, &r1->d_fp_ptr, &r1->d_fp_fk,
                                      &r1->d_fp_g, &r1->n_pairs);
  if (!rc && (pl->precisions & (1 << FCG_FP32))) {
    const int64_t ng = o->n_gdata;
    int64_t ni = 0;
    for (int t = 0; t < h->n_terms; t++) {
      int64_t sz = 1;
      if (h->nbody[t] > 1)
        for (int q = 0; q < h->nbody[t]; q++) sz *= h->dims[4 * t + q];
      ni = std::max(ni, h->ioff[t] + sz);
    }
    r1->d_g32 = static_cast<float *>(fcg_dmalloc(ctx, sizeof(float) * std::max<int64_t>(ng, 1)));
    r1->d_i32 = static_cast<float *>(fcg_dmalloc(ctx, sizeof(float) * std::max<int64_t>(ni, 1)));
    if (!r1->d_g32 || !r1->d_i32) {
      rc = FCG_ENOMEM;
    } else {
      hipLaunchKernelGGL(k_to_float, dim3(1024), dim3(256), 0, ctx->stream, ng, o->d_gdata,
                         r1->d_g32);
      hipLaunchKernelGGL(k_to_float, dim3(1024), dim3(256), 0, ctx->stream, ni, h->d_ints,
                         r1->d_i32);
    }
  }
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FCG_EDEVICE;
  if (rc) {
    if (ctx->err.empty()) ctx->err = "matrix-free plan: device allocation or upload failed";
    fcg_plan_destroy(pl);
    return rc;
  }
  pl->n_pair = r1->n_pairs;
  fcg_plan_register(pl);
  *out = pl;
  return FCG_OK;
}

// errors the kernels flagged (a missing or mis-shaped Γ block)
int fcg_rank1_check(fcg_ctx *ctx, fcg_plan *pl) {
  unsigned int e = 0;
  FCG_HIP(hipMemcpyAsync(&e, pl->r1->d_err, sizeof(e), hipMemcpyDeviceToHost, ctx->stream), ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  if (e) {
    ctx->err = "matrix-free sigma: a term needs a Γ block that ClusterOps lacks or mis-shapes";
    return FCG_EMISSING_BLOCK;
  }
  return FCG_OK;
}
