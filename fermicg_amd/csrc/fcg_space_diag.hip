// Diagonal of H over the configurations of a TPSCI space without a plan.
//
// The reference's compute_diagonal (src/tpsci_outer.jl:694-715) sums, for
// every configuration, contract_matrix_element (src/type_ClusteredTerm.jl)
// of each term in the zero TransferConfig group with bra = ket = that
// configuration.  Only that group reaches the diagonal: a term of another
// group changes some cluster's Fock sector.  Here one thread takes one
// configuration; the threads of a wavefront share a Fock block (the space
// lists its configurations block by block), so they walk the same term list
// and read the same integrals, while each gathers its own Γ columns
// Γ[:, s, s].  Nothing is stored beyond the inputs and the result, so the
// diagonal of a first-order space (millions of configurations, for which a
// stored plan would hold every off-diagonal element too) costs one pass.
#include <algorithm>
#include <string>
#include <vector>

#include "fcg_internal.h"

namespace {

constexpr int SD_THREADS = 256;

// Fock block of configuration i: largest f with fstart[f] <= i
__device__ inline int64_t fock_of(const int64_t *__restrict__ fstart, int64_t n_fock, int64_t i) {
  int64_t lo = 0, hi = n_fock - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (fstart[mid] <= i) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

__global__ void __launch_bounds__(SD_THREADS)
k_space_diag(int64_t n, int N, int64_t n_fock, const int64_t *__restrict__ fstart,
             const int16_t *__restrict__ cfg, const int32_t *__restrict__ task_ptr,
             const TaskDev *__restrict__ tasks, const TermDev *__restrict__ terms,
             const double *__restrict__ ints, const double *__restrict__ gdata,
             const int64_t *__restrict__ gb_off, const int32_t *__restrict__ gb_dims,
             double *__restrict__ hd) {
  const int64_t stride = (int64_t)gridDim.x * SD_THREADS;
  for (int64_t i = (int64_t)blockIdx.x * SD_THREADS + threadIdx.x; i < n; i += stride) {
    const int64_t f = fock_of(fstart, n_fock, i);
    const int16_t *st = cfg + i * N;
    double acc = 0.0;
    for (int32_t k = task_ptr[f]; k < task_ptr[f + 1]; k++) {
      const TaskDev tk = tasks[k];
      const TermDev T = terms[tk.term];
      // Γ[:, s, s] of each active cluster (column-major (np, M, M) blocks)
      const double *g[4] = {nullptr, nullptr, nullptr, nullptr};
#pragma unroll
      for (int a = 0; a < 4; a++) {
        if (a < T.nbody) {
          const int blk = tk.blk[a];
          const int64_t np = gb_dims[3 * blk], mb = gb_dims[3 * blk + 1];
          const int64_t s = st[T.cl[a]] - 1;
          g[a] = gdata + gb_off[blk] + np * (s + mb * s);
        }
      }
      double me = 0.0;
      if (T.nbody == 1) {
        acc += g[0][0];
        continue;
      }
      const double *V = ints + T.ioff;
      const int d0 = T.dims[0], d1 = T.dims[1];
      if (T.nbody == 2) {
        for (int j = 0; j < d1; j++) {
          const double tmp = g[1][j];
          for (int p = 0; p < d0; p++) me += g[0][p] * V[p + d0 * j] * tmp;
        }
      } else if (T.nbody == 3) {
        const int d2 = T.dims[2];
        for (int q = 0; q < d2; q++)
          for (int j = 0; j < d1; j++) {
            const double tmp = g[1][j] * g[2][q];
            const double *Vj = V + (int64_t)d0 * (j + (int64_t)d1 * q);
            for (int p = 0; p < d0; p++) me += g[0][p] * Vj[p] * tmp;
          }
      } else {
        const int d2 = T.dims[2], d3 = T.dims[3];
        for (int l = 0; l < d3; l++)
          for (int q = 0; q < d2; q++) {
            const double tmp2 = g[2][q] * g[3][l];
            for (int j = 0; j < d1; j++) {
              const double tmp = g[1][j] * tmp2;
              const double *Vj = V + (int64_t)d0 * (j + (int64_t)d1 * (q + (int64_t)d2 * l));
              for (int p = 0; p < d0; p++) me += g[0][p] * Vj[p] * tmp;
            }
          }
      }
      acc += (tk.meta & 1) ? -me : me;
    }
    hd[i] = acc;
  }
}

// sign of a term on a Fock configuration (compute_terms_state_sign,
// src/type_ClusteredTerm.jl:181-197): odd-parity operators pick up the
// parity of the electrons on the clusters before theirs
int diag_term_sign(const fcg_ham *h, int t, const int16_t *fk) {
  int s = 0;
  for (int a = 0; a < h->nbody[t]; a++) {
    if (h->parity[4 * t + a] != 1) continue;
    int ne = 0;
    for (int c = 0; c < h->clusters[4 * t + a]; c++) ne += fk[2 * c] + fk[2 * c + 1];
    s ^= ne & 1;
  }
  return s ? -1 : 1;
}

struct DevBuf {
  fcg_ctx *ctx;
  void *p = nullptr;
  explicit DevBuf(fcg_ctx *c) : ctx(c) {}
  ~DevBuf() {
    if (p) fcg_dfree(ctx, p);
  }
  template <class T>
  int upload(const T *src, int64_t n) {
    if (n <= 0) return FCG_OK;
    p = fcg_dmalloc(ctx, sizeof(T) * n);
    if (!p) return FCG_ENOMEM;
    if (src && hipMemcpyAsync(p, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream) !=
                   hipSuccess) {
      ctx->err = "space diagonal: upload failed";
      return FCG_EDEVICE;
    }
    return FCG_OK;
  }
  template <class T>
  T *get() const {
    return static_cast<T *>(p);
  }
};

}  // namespace

extern "C" int fcg_space_diagonal(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t sp,
                                  double *hd) {
  if (!ctx || !h || !o || !sp || !hd) return FCG_EINVAL;
  if (!sp->tpsci) {
    ctx->err = "fcg_space_diagonal needs a TPSCI space (configurations, no Tucker factors)";
    return FCG_EINVAL;
  }
  if (h->N != sp->N || o->N != sp->N) {
    ctx->err = "cluster counts of the Hamiltonian, ClusterOps and space differ";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  const int N = sp->N, S = o->S;
  const int64_t n_fock = sp->n_fock;
  // configurations and Fock block starts, from the upload (boxes not formed)
  // or expanded from the boxes (first cluster fastest within a box)
  std::vector<int16_t> cfg_boxed;
  std::vector<int64_t> fstart_boxed;
  const int16_t *cfg;
  const int64_t *fstart;
  int64_t n;
  if (sp->n_box < 0) {
    cfg = sp->lo.data();
    fstart = sp->var_off.data();
    n = sp->var_off[n_fock];
  } else {
    n = sp->nvar;
    cfg_boxed.resize((size_t)n * N);
    fstart_boxed.assign(n_fock + 1, n);
    for (int64_t b = sp->n_box - 1; b >= 0; b--) fstart_boxed[sp->box_fock[b]] = sp->var_off[b];
    for (int64_t f = n_fock - 1; f >= 0; f--)
      fstart_boxed[f] = std::min(fstart_boxed[f], fstart_boxed[f + 1]);
    for (int64_t b = 0; b < sp->n_box; b++) {
      const int16_t *lo = &sp->lo[b * N], *rk = &sp->rank[b * N];
      for (int64_t v = sp->var_off[b]; v < sp->var_off[b + 1]; v++) {
        int64_t r = v - sp->var_off[b];
        for (int c = 0; c < N; c++) {
          cfg_boxed[v * N + c] = (int16_t)(lo[c] + r % rk[c]);
          r /= rk[c];
        }
      }
    }
    cfg = cfg_boxed.data();
    fstart = fstart_boxed.data();
  }
  if (n == 0) return FCG_OK;
  // the zero TransferConfig group
  int g0 = -1;
  for (int g = 0; g < h->n_tc && g0 < 0; g++) {
    bool zero = true;
    for (int i = 0; i < 2 * N; i++) zero = zero && h->tc[(int64_t)g * 2 * N + i] == 0;
    if (zero) g0 = g;
  }
  // its terms on each Fock block: Γ blocks at (F_c, F_c) and the sign
  std::vector<int32_t> task_ptr(n_fock + 1, 0);
  std::vector<TaskDev> tasks;
  for (int64_t f = 0; f < n_fock; f++) {
    task_ptr[f] = (int32_t)tasks.size();
    if (g0 < 0 || fstart[f + 1] == fstart[f]) continue;
    const int16_t *F = &sp->fock[f * 2 * N];
    for (int c = 0; c < N; c++)
      if (F[2 * c] < 0 || F[2 * c + 1] < 0 || F[2 * c] >= S || F[2 * c + 1] >= S) {
        ctx->err = "Fock sector outside the cluster";
        return FCG_EINVAL;
      }
    for (int t = h->tc_start[g0]; t < h->tc_start[g0 + 1]; t++) {
      TaskDev d{};
      d.term = t;
      d.meta = diag_term_sign(h, t, F) < 0 ? 1 : 0;
      for (int a = 0; a < 4; a++) d.blk[a] = -1;
      for (int a = 0; a < h->nbody[t]; a++) {
        const int c = h->clusters[4 * t + a];
        const int code = F[2 * c] * S + F[2 * c + 1];
        const int blk = o->block(c, h->ops[4 * t + a], code, code);
        const int M = o->nstates[(int64_t)c * S * S + code];
        if (blk < 0) {
          ctx->err = "missing Γ block for term " + std::to_string(t) + " cluster " +
                     std::to_string(c);
          return FCG_EMISSING_BLOCK;
        }
        const int32_t *gd = &o->gb_dims[3 * (int64_t)blk];
        const int np = h->nbody[t] == 1 ? 1 : h->dims[4 * t + a];
        if (gd[0] != np || gd[1] != M || gd[2] != M) {
          ctx->err = "Γ block shape does not match term " + std::to_string(t);
          return FCG_EINVAL;
        }
        d.blk[a] = blk;
      }
      tasks.push_back(d);
    }
  }
  task_ptr[n_fock] = (int32_t)tasks.size();
  // state indices within each block's sector
  for (int64_t f = 0; f < n_fock; f++) {
    const int16_t *F = &sp->fock[f * 2 * N];
    for (int64_t i = fstart[f]; i < fstart[f + 1]; i++)
      for (int c = 0; c < N; c++) {
        const int M = o->nstates[(int64_t)c * S * S + F[2 * c] * S + F[2 * c + 1]];
        if (cfg[i * N + c] < 1 || cfg[i * N + c] > M) {
          ctx->err = "state index out of range";
          return FCG_EINVAL;
        }
      }
  }
  DevBuf d_cfg(ctx), d_fs(ctx), d_tp(ctx), d_tk(ctx), d_hd(ctx);
  {
    int rc = d_cfg.upload(cfg, n * N);
    if (!rc) rc = d_fs.upload(fstart, n_fock + 1);
    if (!rc) rc = d_tp.upload(task_ptr.data(), n_fock + 1);
    if (!rc) rc = d_tk.upload(tasks.data(), (int64_t)tasks.size());
    if (!rc) rc = d_hd.upload<double>(nullptr, n);
    if (rc) return rc;
  }
  const int64_t want = (n + SD_THREADS - 1) / SD_THREADS;
  const unsigned grid = (unsigned)std::min<int64_t>(want, 1 << 20);
  hipLaunchKernelGGL(k_space_diag, dim3(grid), dim3(SD_THREADS), 0, ctx->stream, n, N, n_fock,
                     d_fs.get<int64_t>(), d_cfg.get<int16_t>(), d_tp.get<int32_t>(),
                     tasks.empty() ? nullptr : d_tk.get<TaskDev>(), h->d_terms, h->d_ints,
                     o->d_gdata, o->d_gb_off, o->d_gb_dims, d_hd.get<double>());
  FCG_HIP(hipGetLastError(), ctx);
  FCG_HIP(hipMemcpyAsync(hd, d_hd.get<double>(), sizeof(double) * n, hipMemcpyDeviceToHost,
                         ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}
