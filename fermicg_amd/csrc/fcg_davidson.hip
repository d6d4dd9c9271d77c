// Block Davidson over a sigma plan, vectors resident in HBM.
//
// Replaces the external BlockDavidson.eigs used by tps_ci_davidson
// (src/tpsci_outer.jl:364-453) and ci_solve (src/tucker_outer.jl:68-215),
// with the reference's options (conv_thresh on the residual norm,
// max_ss_vecs per root, lindep_thresh, diagonal preconditioner).
//
// Restarts are thick: the subspace collapses onto its lowest min(2R, 16) Ritz
// vectors.  Mixed precision: expansion vectors get their sigma in FP32 or
// split-BF16; every refresh_k iterations, and before convergence is
// accepted, the subspace collapses onto those Ritz vectors and their sigma
// is recomputed in FP64 (one launch), so the residual that decides
// convergence is always an FP64 one.
// Orthogonalisation (classical Gram-Schmidt, twice) and the Rayleigh-Ritz
// projection are FP64; the small projected eigenproblem is solved on the host
// by cyclic Jacobi.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "fcg_internal.h"

void fcg_blas_destroy(fcg_ctx *ctx) {
  if (ctx && ctx->blas) {
    rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
  }
}

namespace {

__global__ void k_residual(int64_t n, int R, const double *X, const double *AX,
                           const double *theta, double *res) {
  const int64_t N = n * R;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n);
    res[e] = AX[e] - theta[r] * X[e];
  }
}

__global__ void k_precond(int64_t n, int R, const double *hd, const double *theta, double *t) {
  const int64_t N = n * R;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < N;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int r = (int)(e / n);
    double d = theta[r] - hd[e - (int64_t)r * n];
    if (fabs(d) < 1e-8) d = d < 0 ? -1e-8 : 1e-8;
    t[e] = t[e] / d;
  }
}

void jacobi_eigh(int m, std::vector<double> &A, std::vector<double> &w, std::vector<double> &V) {
  V.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; i++) V[i + (size_t)m * i] = 1.0;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int j = 0; j < m; j++)
      for (int i = 0; i < j; i++) off += A[i + (size_t)m * j] * A[i + (size_t)m * j];
    if (off < 1e-30) break;
    for (int p = 0; p < m; p++)
      for (int q = p + 1; q < m; q++) {
        const double apq = A[p + (size_t)m * q];
        if (fabs(apq) < 1e-300) continue;
        const double app = A[p + (size_t)m * p], aqq = A[q + (size_t)m * q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
        const double c = 1.0 / sqrt(1.0 + t * t), s = t * c;
        for (int k = 0; k < m; k++) {
          const double akp = A[k + (size_t)m * p], akq = A[k + (size_t)m * q];
          A[k + (size_t)m * p] = c * akp - s * akq;
          A[k + (size_t)m * q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; k++) {
          const double apk = A[p + (size_t)m * k], aqk = A[q + (size_t)m * k];
          A[p + (size_t)m * k] = c * apk - s * aqk;
          A[q + (size_t)m * k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; k++) {
          const double vkp = V[k + (size_t)m * p], vkq = V[k + (size_t)m * q];
          V[k + (size_t)m * p] = c * vkp - s * vkq;
          V[k + (size_t)m * q] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(m);
  for (int i = 0; i < m; i++) idx[i] = i;
  std::sort(idx.begin(), idx.end(),
            [&](int a, int b) { return A[a + (size_t)m * a] < A[b + (size_t)m * b]; });
  std::vector<double> V2((size_t)m * m);
  w.resize(m);
  for (int j = 0; j < m; j++) {
    w[j] = A[idx[j] + (size_t)m * idx[j]];
    for (int k = 0; k < m; k++) V2[k + (size_t)m * j] = V[k + (size_t)m * idx[j]];
  }
  V.swap(V2);
}

struct Dav {
  fcg_ctx *ctx;
  rocblas_handle h;
  int64_t n;
  double *V = nullptr, *W = nullptr, *X = nullptr, *AX = nullptr, *T = nullptr;
  double *hd = nullptr, *small = nullptr, *theta = nullptr;
  ~Dav() {
    for (double *p : {V, W, X, AX, T, hd, small, theta})
      if (p) (void)hipFree(p);
  }
};

#define BL(x)                                                  \
  do {                                                         \
    if ((x) != rocblas_status_success) {                       \
      ctx->err = std::string("rocBLAS: ") + #x;                \
      return FCG_EDEVICE;                                      \
    }                                                          \
  } while (0)

// Orthogonalise the k columns of T against V[:, :m] (CGS twice) and among
// themselves; normalise.  Columns below lindep are dropped (compacted).
int orthonormalize(Dav &d, int m, int k, double lindep, int *kept) {
  fcg_ctx *ctx = d.ctx;
  const double one = 1.0, zero = 0.0, mone = -1.0;
  const int64_t n = d.n;
  for (int pass = 0; pass < 2 && m > 0; pass++) {
    BL(rocblas_dgemm(d.h, rocblas_operation_transpose, rocblas_operation_none, m, k, n, &one,
                     d.V, n, d.T, n, &zero, d.small, m));
    BL(rocblas_dgemm(d.h, rocblas_operation_none, rocblas_operation_none, n, k, m, &mone, d.V,
                     n, d.small, m, &one, d.T, n));
  }
  int out = 0;
  for (int j = 0; j < k; j++) {
    double *tj = d.T + (int64_t)j * n;
    for (int pass = 0; pass < 2 && out > 0; pass++) {
      BL(rocblas_dgemv(d.h, rocblas_operation_transpose, n, out, &one, d.T, n, tj, 1, &zero,
                       d.small, 1));
      BL(rocblas_dgemv(d.h, rocblas_operation_none, n, out, &mone, d.T, n, d.small, 1, &one, tj,
                       1));
    }
    // re-orthogonalise against V once more for the new column
    if (m > 0) {
      BL(rocblas_dgemv(d.h, rocblas_operation_transpose, n, m, &one, d.V, n, tj, 1, &zero,
                       d.small, 1));
      BL(rocblas_dgemv(d.h, rocblas_operation_none, n, m, &mone, d.V, n, d.small, 1, &one, tj,
                       1));
    }
    double nrm = 0.0;
    BL(rocblas_dnrm2(d.h, n, tj, 1, &nrm));
    if (!(nrm > lindep)) continue;
    const double inv = 1.0 / nrm;
    BL(rocblas_dscal(d.h, n, &inv, tj, 1));
    if (out != j)
      FCG_HIP(hipMemcpyAsync(d.T + (int64_t)out * n, tj, sizeof(double) * n,
                             hipMemcpyDeviceToDevice, ctx->stream),
              ctx);
    out++;
  }
  *kept = out;
  return FCG_OK;
}

}  // namespace

extern "C" int fcg_davidson(fcg_ctx_t ctx, fcg_plan_t pl, int R, const double *v0,
                            const fcg_davidson_opts *o, double *evals, double *evecs,
                            fcg_davidson_stats *st) {
  if (!ctx || !pl || !v0 || !o || !evals || !evecs || R < 1 || R > FCG_MAX_R) return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (!ctx->blas) {
    rocblas_handle hb;
    if (rocblas_create_handle(&hb) != rocblas_status_success) {
      ctx->err = "rocblas_create_handle failed";
      return FCG_EDEVICE;
    }
    ctx->blas = hb;
  }
  Dav d;
  d.ctx = ctx;
  d.h = (rocblas_handle)ctx->blas;
  BL(rocblas_set_stream(d.h, ctx->stream));
  BL(rocblas_set_pointer_mode(d.h, rocblas_pointer_mode_host));
  const int64_t n = pl->sp->nvar;
  d.n = n;
  const int low = o->precision;
  const bool mixed = low != FCG_FP64;
  if (mixed && !(pl->precisions & (1 << low))) {
    ctx->err = "plan lacks the requested sigma precision";
    return FCG_EINVAL;
  }
  // thick restart keeps the lowest KR Ritz vectors (one sigma launch holds
  // at most FCG_MAX_R columns)
  const int KR = std::min(2 * R, FCG_MAX_R);
  const int mmax = std::max(KR + R, o->max_ss_vecs * R);
  const int refresh = std::max(1, o->refresh_k);
  auto alloc = [&](double **p, int64_t cnt) {
    return hipMalloc((void **)p, sizeof(double) * std::max<int64_t>(cnt, 1)) == hipSuccess;
  };
  if (!alloc(&d.V, n * mmax) || !alloc(&d.W, n * mmax) || !alloc(&d.X, n * KR) ||
      !alloc(&d.AX, n * KR) || !alloc(&d.T, n * KR) || !alloc(&d.hd, n) ||
      !alloc(&d.small, (int64_t)mmax * mmax) || !alloc(&d.theta, R)) {
    ctx->err = "Davidson workspace allocation failed";
    return FCG_ENOMEM;
  }
  FCG_HIP(fcg_launch_diag(pl, d.hd, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(d.T, v0, sizeof(double) * n * R, hipMemcpyHostToDevice, ctx->stream),
          ctx);
  int kept = 0;
  if (int rc = orthonormalize(d, 0, R, o->lindep_thresh, &kept)) return rc;
  if (kept < R) {
    ctx->err = "start vectors are linearly dependent";
    return FCG_EINVAL;
  }
  FCG_HIP(hipMemcpyAsync(d.V, d.T, sizeof(double) * n * R, hipMemcpyDeviceToDevice, ctx->stream),
          ctx);
  fcg_davidson_stats stats{};
  auto sigma = [&](const double *in, double *out, int k, int prec) -> int {
    int rc = fcg_sigma_device(ctx, pl, k, in, out, prec);
    if (prec == FCG_FP64) stats.sigma_fp64 += k;
    else stats.sigma_low += k;
    return rc;
  };
  if (int rc = sigma(d.V, d.W, R, FCG_FP64)) return rc;
  int m = R;
  int ne = R;  // leading subspace columns whose sigma is FP64
  int since_refresh = 0;
  std::vector<double> G, w, Y, theta(R), res(R);
  const double one = 1.0, zero = 0.0;
  int converged = 0;
  for (int it = 1; it <= o->max_iter; it++) {
    stats.iterations = it;
    BL(rocblas_dgemm(d.h, rocblas_operation_transpose, rocblas_operation_none, m, m, n, &one, d.V,
                     n, d.W, n, &zero, d.small, m));
    G.resize((size_t)m * m);
    FCG_HIP(hipMemcpyAsync(G.data(), d.small, sizeof(double) * m * m, hipMemcpyDeviceToHost,
                           ctx->stream),
            ctx);
    FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
    // G[i,j] = v_i' w_j.  Columns below ne carry FP64 sigma: a coupling of
    // an exact column i with a reduced-precision column j is taken as
    // v_j' w_i (= v_i' H v_j exactly), the rest is symmetrised by averaging
    for (int j = 0; j < m; j++)
      for (int i = 0; i < j; i++) {
        const double a = (i < ne && j >= ne)
                             ? G[j + (size_t)m * i]
                             : 0.5 * (G[i + (size_t)m * j] + G[j + (size_t)m * i]);
        G[i + (size_t)m * j] = G[j + (size_t)m * i] = a;
      }
    jacobi_eigh(m, G, w, Y);
    for (int r = 0; r < R; r++) theta[r] = w[r];
    const int nk = std::min(KR, m);
    FCG_HIP(hipMemcpyAsync(d.small, Y.data(), sizeof(double) * m * nk, hipMemcpyHostToDevice,
                           ctx->stream),
            ctx);
    FCG_HIP(hipMemcpyAsync(d.theta, theta.data(), sizeof(double) * R, hipMemcpyHostToDevice,
                           ctx->stream),
            ctx);
    // lowest nk Ritz vectors X = V Y and their images AX = W Y
    BL(rocblas_dgemm(d.h, rocblas_operation_none, rocblas_operation_none, n, nk, m, &one, d.V, n,
                     d.small, m, &zero, d.X, n));
    BL(rocblas_dgemm(d.h, rocblas_operation_none, rocblas_operation_none, n, nk, m, &one, d.W, n,
                     d.small, m, &zero, d.AX, n));
    hipLaunchKernelGGL(k_residual, dim3(1024), dim3(256), 0, ctx->stream, n, R, d.X, d.AX,
                       d.theta, d.T);
    double rmax = 0.0;
    for (int r = 0; r < R; r++) {
      BL(rocblas_dnrm2(d.h, n, d.T + (int64_t)r * n, 1, &res[r]));
      rmax = std::max(rmax, res[r]);
    }
    since_refresh++;
    if (rmax < o->conv_thresh && ne == m) {
      converged = 1;
      break;
    }
    const bool want_refresh =
        mixed && ne < m && (rmax < o->conv_thresh || since_refresh >= refresh);
    if (want_refresh || m + R > mmax) {
      const bool exact = ne == m;
      m = nk;
      if (!exact) {
        // collapse onto the re-orthonormalised Ritz vectors; their sigma is
        // recomputed in FP64 so the residual that decides convergence is exact
        FCG_HIP(hipMemcpyAsync(d.T, d.X, sizeof(double) * n * nk, hipMemcpyDeviceToDevice,
                               ctx->stream),
                ctx);
        if (int rc = orthonormalize(d, 0, nk, 1e-300, &kept)) return rc;
        m = kept;
        FCG_HIP(hipMemcpyAsync(d.V, d.T, sizeof(double) * n * m, hipMemcpyDeviceToDevice,
                               ctx->stream),
                ctx);
        if (int rc = sigma(d.V, d.W, m, FCG_FP64)) return rc;
        since_refresh = 0;
      } else {
        // exact subspace: thick restart V <- V Y, W <- W Y
        FCG_HIP(hipMemcpyAsync(d.V, d.X, sizeof(double) * n * nk, hipMemcpyDeviceToDevice,
                               ctx->stream),
                ctx);
        FCG_HIP(hipMemcpyAsync(d.W, d.AX, sizeof(double) * n * nk, hipMemcpyDeviceToDevice,
                               ctx->stream),
                ctx);
      }
      ne = m;
      continue;
    }
    // correction vectors from the residuals of the R lowest roots
    if (o->precond)
      hipLaunchKernelGGL(k_precond, dim3(1024), dim3(256), 0, ctx->stream, n, R, d.hd, d.theta,
                         d.T);
    if (int rc = orthonormalize(d, m, R, o->lindep_thresh, &kept)) return rc;
    if (kept == 0) {
      ctx->err = "Davidson: all correction vectors linearly dependent";
      break;
    }
    FCG_HIP(hipMemcpyAsync(d.V + (int64_t)m * n, d.T, sizeof(double) * n * kept,
                           hipMemcpyDeviceToDevice, ctx->stream),
            ctx);
    const int prec = mixed ? low : FCG_FP64;
    if (int rc = sigma(d.V + (int64_t)m * n, d.W + (int64_t)m * n, kept, prec)) return rc;
    m += kept;
    if (!mixed) ne = m;
  }
  FCG_HIP(hipMemcpyAsync(evecs, d.X, sizeof(double) * n * R, hipMemcpyDeviceToHost, ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  for (int r = 0; r < R; r++) {
    evals[r] = theta[r];
    if (r < 16) stats.residual[r] = res[r];
  }
  stats.converged = converged;
  if (st) *st = stats;
  return converged ? FCG_OK : FCG_ENOTCONV;
}
