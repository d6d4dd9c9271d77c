// Second-order (Epstein-Nesbet-type, H0 = sum of Hcmf) energy of a TPSCI
// state over its first-order space, on the GPU.
//
// Follows the reference's compute_pt2_energy (src/tpsci_pt2_energy.jl:10-175):
//   * target Fock blocks are F_ket + TransferConfig over the state's Fock
//     blocks, kept when every cluster sector exists (:49-79);
//   * sigma on the target blocks (the open matvec, _pt2_job :178-266) over
//     every configuration of each target block, with the reference's
//     screening: a term's contribution from ket configuration j enters only
//     when its magnitude exceeds thresh_foi / max_r |c[j, r]|
//     (_collect_significant_thread!, src/tpsci_matvec_thread.jl:830-895; the
//     reference's prescreen bounds only skip contributions that fail this
//     test);
//   * configurations of the variational space are projected out
//     (project_out!, src/tpsci_outer.jl:869-889);
//   * denominators from the Hcmf diagonal (compute_diagonal! with "Hcmf",
//     src/tpsci_outer.jl:735-757) and E0 = <0|H0|0> supplied by the caller;
//   * e2[r] = sum_x sigma[x,r]^2 / (E0[r] - Hd[x]) (_sum_pt2 :269-282),
//     accumulated with Neumaier-compensated sums in a fixed order.
#include <chrono>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "fcg_internal.h"

namespace {

constexpr int PT2_BLOCKS = 1024;

// Hd[x] = sum_c H0_c[i_c, i_c] over the configurations of each box
__global__ void k_h0_diag(int64_t n_box, int N, const int64_t *__restrict__ var,
                            const int16_t *__restrict__ rank, const int64_t *__restrict__ hoff,
                            const int32_t *__restrict__ hstride, const double *__restrict__ gdata,
                            double *__restrict__ hd) {
  const int64_t b = (int64_t)blockIdx.x;
  if (b >= n_box) return;
  const int64_t v0 = var[b];
  const int d = (int)(var[b + 1] - v0);
  for (int e = threadIdx.x; e < d; e += blockDim.x) {
    int r = e;
    double s = 0.0;
    for (int c = 0; c < N; c++) {
      const int rk = rank[b * N + c];
      const int ic = r % rk;
      r /= rk;
      s += gdata[hoff[b * N + c] + (int64_t)ic * hstride[b * N + c]];
    }
    hd[v0 + e] = s;
  }
}

__device__ __forceinline__ void neumaier(double &sum, double &comp, double v) {
  const double t = sum + v;
  comp += fabs(sum) >= fabs(v) ? (sum - t) + v : (v - t) + sum;
  sum = t;
}

// per-block compensated partial sums of sigma^2 / (E0 - Hd) over unmasked x
__global__ __launch_bounds__(256) void k_pt2_sum(int64_t n, int R, const double *__restrict__ sigma,
                                                 const double *__restrict__ hd,
                                                 const uint8_t *__restrict__ inspace,
                                                 const double *__restrict__ e0,
                                                 double *__restrict__ part) {
  __shared__ double ss[256], sc[256];
  for (int r = 0; r < R; r++) {
    double sum = 0.0, comp = 0.0;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n;
         x += (int64_t)gridDim.x * blockDim.x) {
      if (inspace[x]) continue;
      const double v = sigma[(int64_t)r * n + x];
      neumaier(sum, comp, v * v / (e0[r] - hd[x]));
    }
    ss[threadIdx.x] = sum;
    sc[threadIdx.x] = comp;
    __syncthreads();
    if (threadIdx.x == 0) {  // fixed-order block reduction
      double s = 0.0, c = 0.0;
      for (int t = 0; t < (int)blockDim.x; t++) {
        neumaier(s, c, ss[t]);
        c += sc[t];
      }
      part[((int64_t)blockIdx.x * R + r) * 2] = s;
      part[((int64_t)blockIdx.x * R + r) * 2 + 1] = c;
    }
    __syncthreads();
  }
}

// scratch of one PT2 call, stream-ordered on its context
template <typename T>
struct DevBuf {
  fcg_ctx *ctx = nullptr;
  T *p = nullptr;
  explicit DevBuf(fcg_ctx *c) : ctx(c) {}
  ~DevBuf() {
    if (p) fcg_dfree(ctx, p);
  }
  bool alloc(int64_t n) {
    p = static_cast<T *>(fcg_dmalloc(ctx, sizeof(T) * std::max<int64_t>(n, 1)));
    return p != nullptr;
  }
};

}  // namespace

// thresholds of the reference's _collect_significant_thread!
// (src/tpsci_matvec_thread.jl:830-895): thresh / max_r |c[j, r]| per ket
// variable (infinite when the coefficients vanish: nothing is collected)
int fcg_screen_thresholds(fcg_ctx *ctx, int64_t n, int R, const double *c, double thresh,
                          double **d_thr) {
  std::vector<double> thr(std::max<int64_t>(n, 1));
  for (int64_t j = 0; j < n; j++) {
    double m = 0.0;
    for (int r = 0; r < R; r++) m = std::max(m, std::fabs(c[(int64_t)r * n + j]));
    thr[j] = m > 0.0 ? thresh / m : INFINITY;
  }
  *d_thr = static_cast<double *>(fcg_dmalloc(ctx, sizeof(double) * thr.size()));
  if (!*d_thr) return FCG_ENOMEM;
  FCG_HIP(hipMemcpyAsync(*d_thr, thr.data(), sizeof(double) * thr.size(), hipMemcpyHostToDevice,
                         ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}

extern "C" int fcg_tpsci_pt2(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t var, int R,
                             const double *c, const double *e0, int nbody, int h0_op,
                             double thresh, double *e2, fcg_pt2_stats *st) {
  if (!ctx || !h || !o || !var || !c || !e0 || !e2 || R < 1 || R > FCG_MAX_R || nbody < 1 ||
      nbody > 4 || h0_op < 0 || h0_op >= o->n_ops || !(thresh >= 0.0))
    return FCG_EINVAL;
  if (!var->tpsci) {
    ctx->err = "fcg_tpsci_pt2 needs a TPSCI space";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  const auto t0 = std::chrono::steady_clock::now();
  if (int rc = fcg_form_boxes(ctx, o, var)) return rc;
  fcg_space *tsp = new fcg_space();
  tsp->ctx = ctx;
  std::unique_ptr<fcg_space, void (*)(fcg_space *)> tguard(tsp, fcg_space_destroy);
  Pt2Target tg;
  if (int rc = fcg_pt2_target(h, o, var, h0_op, tsp, &tg, &ctx->err)) return rc;
  if (int rc = fcg_space_to_device(tsp)) return rc;
  const int N = var->N;
  const int64_t nt = tsp->nvar, n_tf = tsp->n_fock;
  const std::vector<uint8_t> &inspace = tg.inspace;
  const std::vector<int64_t> &hoff = tg.hoff;
  const std::vector<int32_t> &hstride = tg.hstride;
  DevBuf<double> dthr(ctx);
  if (thresh > 0.0)
    if (int rc = fcg_screen_thresholds(ctx, var->nvar, R, c, thresh, &dthr.p)) return rc;
  DevBuf<double> dc(ctx), ds(ctx), dhd(ctx), de0(ctx), dpart(ctx);
  DevBuf<uint8_t> din(ctx);
  DevBuf<int64_t> dhoff(ctx);
  DevBuf<int32_t> dhstr(ctx);
  if (!dc.alloc((int64_t)R * var->nvar) || !ds.alloc((int64_t)R * nt) || !dhd.alloc(nt) ||
      !de0.alloc(R) || !dpart.alloc((int64_t)PT2_BLOCKS * R * 2) || !din.alloc(nt) ||
      !dhoff.alloc(tsp->n_box * N) || !dhstr.alloc(tsp->n_box * N)) {
    ctx->err = "PT2 workspace allocation failed";
    return FCG_ENOMEM;
  }
  FCG_HIP(hipMemcpyAsync(dc.p, c, sizeof(double) * R * var->nvar, hipMemcpyHostToDevice, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(de0.p, e0, sizeof(double) * R, hipMemcpyHostToDevice, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(din.p, inspace.data(), nt, hipMemcpyHostToDevice, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(dhoff.p, hoff.data(), sizeof(int64_t) * hoff.size(), hipMemcpyHostToDevice, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(dhstr.p, hstride.data(), sizeof(int32_t) * hstride.size(), hipMemcpyHostToDevice, ctx->stream), ctx);
  // sigma on the first-order space, accumulated during the block build (the
  // blocks are used once, so none is stored)
  FCG_HIP(hipMemsetAsync(ds.p, 0, sizeof(double) * R * nt, ctx->stream), ctx);
  const FusedSigma fused{R, dc.p, ds.p};
  fcg_plan *pl = nullptr;
  if (int rc = fcg_create_plan(ctx, h, o, tsp, var, nbody, 1 << FCG_FP64, &pl, INT64_MAX,
                               nullptr, dthr.p, &fused))
    return rc;
  std::unique_ptr<fcg_plan, void (*)(fcg_plan *)> pguard(pl, fcg_plan_destroy);
  if (tsp->n_box > 0)
    hipLaunchKernelGGL(k_h0_diag, dim3((unsigned)tsp->n_box), dim3(64), 0, ctx->stream,
                       tsp->n_box, N, tsp->d_var, tsp->d_rank, dhoff.p, dhstr.p, o->d_gdata, dhd.p);
  hipLaunchKernelGGL(k_pt2_sum, dim3(PT2_BLOCKS), dim3(256), 0, ctx->stream, nt, R, ds.p, dhd.p,
                     din.p, de0.p, dpart.p);
  FCG_HIP(hipGetLastError(), ctx);
  std::vector<double> part((size_t)PT2_BLOCKS * R * 2);
  FCG_HIP(hipMemcpyAsync(part.data(), dpart.p, sizeof(double) * part.size(), hipMemcpyDeviceToHost, ctx->stream), ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  for (int r = 0; r < R; r++) {
    double s = 0.0, comp = 0.0;
    for (int blk = 0; blk < PT2_BLOCKS; blk++) {
      const double v = part[((size_t)blk * R + r) * 2];
      const double t = s + v;
      comp += std::fabs(s) >= std::fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
      comp += part[((size_t)blk * R + r) * 2 + 1];
    }
    e2[r] = s + comp;
  }
  if (st) {
    int64_t nin = 0;
    for (uint8_t m : inspace) nin += m;
    st->n_target_fock = n_tf;
    st->n_target_var = nt;
    st->n_first_order = nt - nin;
    st->block_pairs = pl->n_pair;
    st->block_elems = pl->n_kel;
    st->build_flops = pl->build_flops;
    st->build_seconds = pl->build_us * 1e-6;
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  return FCG_OK;
}

// Open matvec with the reference's screening (open_matvec_thread,
// src/tpsci_matvec_thread.jl:11-129, its element test
// _collect_significant_thread! :830-895): sigma over the configurations of
// `bra` from the state (ket, c), built and applied in one pass (no block is
// stored), and hit[x] = 1 for every bra configuration that some term's
// element reached with |element| > thresh / max_r |c[j, r]| -- the keys the
// reference inserts, whether or not their summed sigma cancels to zero.
extern "C" int fcg_open_sigma(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t bra,
                              fcg_space_t ket, int nbody, int R, const double *c, double thresh,
                              double *sigma, uint8_t *hit) {
  if (!ctx || !h || !o || !bra || !ket || !c || !sigma || R < 1 || R > FCG_MAX_R || nbody < 1 ||
      nbody > 4 || !(thresh >= 0.0))
    return FCG_EINVAL;
  if (!bra->tpsci || !ket->tpsci) {
    ctx->err = "fcg_open_sigma needs TPSCI spaces";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  if (int rc = fcg_form_boxes(ctx, o, bra)) return rc;
  if (int rc = fcg_form_boxes(ctx, o, ket)) return rc;
  const int64_t nb = bra->nvar, nk = ket->nvar;
  DevBuf<double> dthr(ctx), dc(ctx), ds(ctx);
  DevBuf<uint8_t> dhit(ctx);
  // thresholds always: a ket variable with vanishing coefficients gets an
  // infinite threshold and reaches nothing, as in the reference
  if (int rc = fcg_screen_thresholds(ctx, nk, R, c, thresh, &dthr.p)) return rc;
  if (!dc.alloc((int64_t)R * nk) || !ds.alloc((int64_t)R * nb) || !dhit.alloc(nb)) {
    ctx->err = "open sigma workspace allocation failed";
    return FCG_ENOMEM;
  }
  FCG_HIP(hipMemcpyAsync(dc.p, c, sizeof(double) * R * nk, hipMemcpyHostToDevice, ctx->stream),
          ctx);
  FCG_HIP(hipMemsetAsync(ds.p, 0, sizeof(double) * R * std::max<int64_t>(nb, 1), ctx->stream),
          ctx);
  FCG_HIP(hipMemsetAsync(dhit.p, 0, std::max<int64_t>(nb, 1), ctx->stream), ctx);
  FusedSigma fused{R, dc.p, ds.p};
  fused.hit = dhit.p;
  fcg_plan *pl = nullptr;
  if (nb > 0 && nk > 0) {
    if (int rc = fcg_create_plan(ctx, h, o, bra, ket, nbody, 1 << FCG_FP64, &pl, INT64_MAX,
                                 nullptr, dthr.p, &fused))
      return rc;
  }
  std::unique_ptr<fcg_plan, void (*)(fcg_plan *)> pguard(pl, fcg_plan_destroy);
  if (nb > 0) {
    FCG_HIP(hipMemcpyAsync(sigma, ds.p, sizeof(double) * R * nb, hipMemcpyDeviceToHost,
                           ctx->stream),
            ctx);
    if (hit)
      FCG_HIP(hipMemcpyAsync(hit, dhit.p, nb, hipMemcpyDeviceToHost, ctx->stream), ctx);
  }
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}
