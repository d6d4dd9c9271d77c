// Block Davidson over a sigma plan, vectors resident in HBM.
//
// Replaces the external BlockDavidson.eigs used by tps_ci_davidson
// (src/tpsci_outer.jl:364-453) and ci_solve (src/tucker_outer.jl:68-215),
// with the reference's options (conv_thresh on the residual norm,
// max_ss_vecs per root, lindep_thresh, diagonal preconditioner).
//
// Restarts are thick: the subspace collapses onto its lowest min(2R, 16) Ritz
// vectors.  Mixed precision: expansion vectors get their sigma in FP32,
// split-BF16 or BF16; at every restart, before convergence is accepted and
// (optionally) every refresh_k iterations, the subspace collapses onto those
// Ritz vectors and their sigma is recomputed in FP64 (one launch), so the
// residual that decides convergence is always an FP64 one.
// Orthogonalisation (block Gram-Schmidt against the subspace, twice, then
// symmetric (Loewdin) orthonormalisation of the new block) and the
// Rayleigh-Ritz projection are FP64 on the GPU with this file's own
// tall-skinny kernels: a split-K Gram product reduced in a fixed order (so
// results are reproducible) and a row-parallel product with a small matrix.
// The small projected eigenproblems (Rayleigh-Ritz and the Loewdin
// orthonormalisation) are solved on the GPU by parallel cyclic Jacobi in one
// workgroup (subspaces up to EMAX vectors; larger ones on the host), so an
// iteration synchronises with the host twice: for the residual norms that
// decide convergence and restarts, and for the count of correction vectors
// kept after the linear-dependence test.
#include <algorithm>
#include <cmath>
#include <vector>

#include "fcg_internal.h"

namespace {

constexpr int GT = 16;            // Gram output tile (GT x GT per workgroup)
constexpr int GROWS = 64;         // rows staged in LDS per step
constexpr int64_t GCHUNK = 4096;  // rows per split-K chunk
constexpr int EMAX = 96;          // largest subspace whose eigenproblem runs on the GPU
constexpr int ET = 1024;          // threads of the eigensolver workgroup

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

// part[chunk][i + mv*j] = sum over the chunk's rows of V[r,i] W[r,j]
__global__ __launch_bounds__(256) void k_gram_partial(int64_t n, const double *V, int mv,
                                                      const double *W, int mw, double *part) {
  __shared__ double sV[GROWS][GT + 1];
  __shared__ double sW[GROWS][GT + 1];
  const int tid = threadIdx.x, ti = tid % GT, tj = tid / GT;
  const int i0 = blockIdx.y * GT, j0 = blockIdx.z * GT;
  const int64_t r_begin = (int64_t)blockIdx.x * GCHUNK;
  const int64_t r_end = std::min<int64_t>(n, r_begin + GCHUNK);
  double acc = 0.0;
  for (int64_t r0 = r_begin; r0 < r_end; r0 += GROWS) {
    for (int e = tid; e < GROWS * GT; e += 256) {
      const int rr = e % GROWS, cc = e / GROWS;
      const int64_t r = r0 + rr;
      sV[rr][cc] = (r < r_end && i0 + cc < mv) ? V[r + n * (i0 + cc)] : 0.0;
      sW[rr][cc] = (r < r_end && j0 + cc < mw) ? W[r + n * (j0 + cc)] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int rr = 0; rr < GROWS; rr++) acc += sV[rr][ti] * sW[rr][tj];
    __syncthreads();
  }
  const int i = i0 + ti, j = j0 + tj;
  if (i < mv && j < mw) part[(int64_t)blockIdx.x * mv * mw + i + (int64_t)mv * j] = acc;
}

// part[chunk][i + mv*j] as k_gram_partial, for mv, mw <= GW: one workgroup
// per row chunk covers every output, so V and W are read once
constexpr int GW = 64, GR = 32;
__global__ __launch_bounds__(256) void k_gram_partial_wide(int64_t n, const double *V, int mv,
                                                           const double *W, int mw,
                                                           double *part) {
  __shared__ double sV[GR][GW + 1];
  __shared__ double sW[GR][GW + 1];
  const int tid = threadIdx.x;
  const int cnt = mv * mw;
  const int64_t r_begin = (int64_t)blockIdx.x * GCHUNK;
  const int64_t r_end = std::min<int64_t>(n, r_begin + GCHUNK);
  double acc[GW * GW / 256];
#pragma unroll
  for (int k = 0; k < GW * GW / 256; k++) acc[k] = 0.0;
  for (int64_t r0 = r_begin; r0 < r_end; r0 += GR) {
    for (int e = tid; e < GR * GW; e += 256) {
      const int rr = e % GR, cc = e / GR;
      const int64_t r = r0 + rr;
      sV[rr][cc] = (r < r_end && cc < mv) ? V[r + n * cc] : 0.0;
      sW[rr][cc] = (r < r_end && cc < mw) ? W[r + n * cc] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < GW * GW / 256; k++) {
      const int o = tid + 256 * k;
      if (o < cnt) {
        const int i = o % mv, j = o / mv;
        double a = acc[k];
        for (int rr = 0; rr < GR; rr++) a += sV[rr][i] * sW[rr][j];
        acc[k] = a;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < GW * GW / 256; k++) {
    const int o = tid + 256 * k;
    if (o < cnt) part[(int64_t)blockIdx.x * cnt + o] = acc[k];
  }
}

// out[i + ld*j] = sum over chunks of part[chunk][i + ma*j], one wavefront per
// output, in a fixed order (strided partial sums, then a shuffle tree)
__global__ __launch_bounds__(64) void k_gram_reduce(int nchunk, int ma, int cnt,
                                                    const double *part, double *out, int ld) {
  const int e = blockIdx.x;
  if (e >= cnt) return;
  double s = 0.0;
  for (int c = threadIdx.x; c < nchunk; c += 64) s += part[(int64_t)c * cnt + e];
  for (int k = 32; k >= 1; k >>= 1) s += __shfl_xor(s, k, 64);
  if (threadIdx.x == 0) out[(e % ma) + (int64_t)ld * (e / ma)] = s;
}

// Y[r, c] = beta Y[r, c] + alpha sum_i A[r, i] B[i, c]  (B is k x nc, k <= 256, nc <= 16)
__global__ __launch_bounds__(256) void k_gemm_small(int64_t n, int k, int nc, const double *A,
                                                    const double *B, double alpha, double beta,
                                                    double *Y) {
  __shared__ double sB[FCG_MAX_R * 16 * 16];
  for (int e = threadIdx.x; e < k * nc; e += blockDim.x) sB[e] = B[e];
  __syncthreads();
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    double acc[16];
#pragma unroll
    for (int c = 0; c < 16; c++) acc[c] = 0.0;
    for (int i = 0; i < k; i++) {
      const double a = A[r + n * i];
#pragma unroll
      for (int c = 0; c < 16; c++)
        if (c < nc) acc[c] += a * sB[i + k * c];
    }
    for (int c = 0; c < nc; c++) {
      const double prev = beta == 0.0 ? 0.0 : beta * Y[r + n * c];
      Y[r + n * c] = prev + alpha * acc[c];
    }
  }
}

// Symmetric eigendecomposition of the m x m matrix G (column-major with
// leading dimension ldg, m <= EMAX) in one workgroup: parallel cyclic Jacobi, each sweep m'-1 rounds of
// m'/2 disjoint rotations (round-robin pairing, m' = m rounded up to even),
// the rotation of a pair as in the serial method, until the off-diagonal
// norm is below 1e-15 of the Frobenius norm (at most 100 sweeps).  Then:
//   mode 0 (Rayleigh-Ritz): G is first symmetrised -- an element coupling an
//     FP64 column i < ne with a reduced-precision column j >= ne is taken as
//     G[j, i], the rest averaged; out = the eigenvectors of the nk lowest
//     eigenvalues (m x nk), theta = the R lowest eigenvalues;
//   mode 1 (orthonormalisation): out = U[:, c] / sqrt(w_c) over the
//     eigenvalues w_c > floor2 and > 0, largest first (m x kept), *kept_out.
__global__ __launch_bounds__(ET) void k_small_eigh(int m, int ne, const double *G, int ldg,
                                                  int mode, int nk, int R, double floor2,
                                                  double *out, double *theta, int *kept_out) {
  __shared__ double A[EMAX][EMAX + 1];  // A[row][col]
  __shared__ double V[EMAX][EMAX + 1];
  __shared__ double sC[EMAX / 2], sS[EMAX / 2];
  __shared__ int sP[EMAX / 2], sQ[EMAX / 2];
  __shared__ double sRed[ET / 64], sOff, sNorm;
  __shared__ int sOrd[EMAX];
  const int tid = threadIdx.x;
  for (int e = tid; e < m * m; e += ET) {
    const int i = e % m, j = e / m;
    double a = G[i + (int64_t)ldg * j];
    if (mode == 0 && i != j) {
      const int lo = min(i, j), hi = max(i, j);
      a = (lo < ne && hi >= ne) ? G[hi + (int64_t)ldg * lo]
                                : 0.5 * (G[i + (int64_t)ldg * j] + G[j + (int64_t)ldg * i]);
    }
    A[i][j] = a;
    V[i][j] = i == j ? 1.0 : 0.0;
  }
  __syncthreads();
  const int mp = m + (m & 1);
  auto offnorm = [&](bool full) {  // sum of squares off the diagonal (and all of A)
    double o = 0.0, f = 0.0;
    for (int e = tid; e < m * m; e += ET) {
      const int i = e % m, j = e / m;
      const double a = A[i][j];
      if (i != j) o += a * a;
      f += a * a;
    }
    for (int k = 32; k >= 1; k >>= 1) {
      o += __shfl_xor(o, k, 64);
      f += __shfl_xor(f, k, 64);
    }
    if ((tid & 63) == 0) sRed[tid >> 6] = o;
    __syncthreads();
    if (tid == 0) {
      double t = 0.0;
      for (int k = 0; k < ET / 64; k++) t += sRed[k];
      sOff = t;
    }
    __syncthreads();
    if (full) {
      if ((tid & 63) == 0) sRed[tid >> 6] = f;
      __syncthreads();
      if (tid == 0) {
        double t = 0.0;
        for (int k = 0; k < ET / 64; k++) t += sRed[k];
        sNorm = t;
      }
      __syncthreads();
    }
  };
  offnorm(true);
  const double stop = fmax(1e-30, 1e-30 * sNorm);  // (1e-15)^2 of the squared norm
  for (int sweep = 0; sweep < 100 && sOff > stop && m > 1; sweep++) {
    for (int rnd = 0; rnd < mp - 1; rnd++) {
      if (tid < mp / 2) {
        const int i = tid;
        auto player = [&](int pos) { return pos == 0 ? 0 : 1 + (pos - 1 + rnd) % (mp - 1); };
        int p = player(i), q = player(mp - 1 - i);
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, sn = 0.0;
        if (q < m) {
          const double apq = A[p][q];
          if (fabs(apq) >= 1e-300) {
            const double app = A[p][p], aqq = A[q][q];
            const double tau = (aqq - app) / (2.0 * apq);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            sn = t * c;
          }
        }
        sC[i] = c;
        sS[i] = sn;
        sP[i] = p;
        sQ[i] = q < m ? q : -1;
      }
      __syncthreads();
      // columns p, q of A and V
      for (int e = tid; e < (mp / 2) * m; e += ET) {
        const int i = e / m, k = e % m, q = sQ[i];
        if (q < 0 || sS[i] == 0.0) continue;
        const int p = sP[i];
        const double c = sC[i], sn = sS[i];
        const double akp = A[k][p], akq = A[k][q];
        A[k][p] = c * akp - sn * akq;
        A[k][q] = sn * akp + c * akq;
        const double vkp = V[k][p], vkq = V[k][q];
        V[k][p] = c * vkp - sn * vkq;
        V[k][q] = sn * vkp + c * vkq;
      }
      __syncthreads();
      // rows p, q of A
      for (int e = tid; e < (mp / 2) * m; e += ET) {
        const int i = e / m, k = e % m, q = sQ[i];
        if (q < 0 || sS[i] == 0.0) continue;
        const int p = sP[i];
        const double c = sC[i], sn = sS[i];
        const double apk = A[p][k], aqk = A[q][k];
        A[p][k] = c * apk - sn * aqk;
        A[q][k] = sn * apk + c * aqk;
      }
      __syncthreads();
    }
    offnorm(false);
  }
  // eigenvalues ascending (ties by index): rank of each diagonal element
  for (int i = tid; i < m; i += ET) {
    const double wi = A[i][i];
    int rank = 0;
    for (int j = 0; j < m; j++) {
      const double wj = A[j][j];
      rank += wj < wi || (wj == wi && j < i);
    }
    sOrd[rank] = i;
  }
  __syncthreads();
  if (mode == 0) {
    for (int e = tid; e < m * nk; e += ET) {
      const int k = e % m, c = e / m;
      out[k + (int64_t)m * c] = V[k][sOrd[c]];
    }
    if (tid < R) theta[tid] = A[sOrd[tid]][sOrd[tid]];
  } else {
    __shared__ int sKeep[EMAX];
    __shared__ int sNk;
    if (tid == 0) {
      int nkp = 0;
      for (int j = m - 1; j >= 0; j--) {  // largest first
        const double wj = A[sOrd[j]][sOrd[j]];
        if (wj > floor2 && wj > 0.0) sKeep[nkp++] = sOrd[j];
      }
      sNk = nkp;
      *kept_out = nkp;
    }
    __syncthreads();
    for (int e = tid; e < m * sNk; e += ET) {
      const int k = e % m, c = e / m;
      const int j = sKeep[c];
      out[k + (int64_t)m * c] = V[k][j] / sqrt(A[j][j]);
    }
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
  fcg_ctx *ctx = nullptr;
  int64_t n = 0;
  int nchunk = 0;
  double *V = nullptr, *W = nullptr, *X = nullptr, *AX = nullptr, *T = nullptr, *T2 = nullptr;
  double *hd = nullptr, *small = nullptr, *theta = nullptr, *part = nullptr;
  double *G = nullptr;      // V'W of the subspace, leading dimension mmax
  int *kept_dev = nullptr;  // [2]: columns kept by the two orthonormalisation rounds
  int deferred = -1;        // columns the last unchecked second round must have kept
  ~Dav() {  // stream-ordered: waits for no other context's work
    for (double *p : {V, W, X, AX, T, T2, hd, small, theta, part, G})
      if (p) fcg_dfree(ctx, p);
    if (kept_dev) fcg_dfree(ctx, kept_dev);
  }
};

// the second orthonormalisation round keeps every column of an orthonormal
// block; its count is read at the next synchronisation instead of at once
int check_deferred(Dav &d, const int *kept_host) {
  if (d.deferred >= 0 && kept_host[1] != d.deferred) {
    d.ctx->err = "Davidson: a correction block lost rank in its second orthonormalisation";
    return FCG_ENOTCONV;
  }
  d.deferred = -1;
  return FCG_OK;
}

// dev_out (ma x mb, column-major with leading dimension ld, default ma) =
// A' B over all n rows; optionally copied to host (then ld must be ma)
int gram(Dav &d, const double *A, int ma, const double *B, int mb, double *dev_out,
         std::vector<double> *host_out, int ld = 0) {
  fcg_ctx *ctx = d.ctx;
  if (ld <= 0) ld = ma;
  if (ma <= GW && mb <= GW) {
    hipLaunchKernelGGL(k_gram_partial_wide, dim3((unsigned)d.nchunk), dim3(256), 0, ctx->stream,
                       d.n, A, ma, B, mb, d.part);
  } else {
    const dim3 grid((unsigned)d.nchunk, (unsigned)((ma + GT - 1) / GT),
                    (unsigned)((mb + GT - 1) / GT));
    hipLaunchKernelGGL(k_gram_partial, grid, dim3(256), 0, ctx->stream, d.n, A, ma, B, mb, d.part);
  }
  const int cnt = ma * mb;
  hipLaunchKernelGGL(k_gram_reduce, dim3((unsigned)cnt), dim3(64), 0, ctx->stream, d.nchunk, ma,
                     cnt, d.part, dev_out, ld);
  FCG_HIP(hipGetLastError(), ctx);
  if (host_out) {
    host_out->resize(cnt);
    FCG_HIP(hipMemcpyAsync(host_out->data(), dev_out, sizeof(double) * cnt, hipMemcpyDeviceToHost,
                           ctx->stream),
            ctx);
    FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  }
  return FCG_OK;
}

// Y = beta Y + alpha A B  (A: n x k resident, B: k x nc on the device)
int gemm(Dav &d, const double *A, int k, const double *B, int nc, double alpha, double beta,
         double *Y) {
  fcg_ctx *ctx = d.ctx;
  const unsigned blocks = (unsigned)std::min<int64_t>(2048, (d.n + 255) / 256);
  hipLaunchKernelGGL(k_gemm_small, dim3(blocks), dim3(256), 0, ctx->stream, d.n, k, nc, A, B,
                     alpha, beta, Y);
  FCG_HIP(hipGetLastError(), ctx);
  return FCG_OK;
}

// Orthonormalise the k columns of T against V[:, :m] and among themselves:
// two rounds of (block Gram-Schmidt against V, then symmetric orthonormal-
// isation S^-1/2 of the block).  The first round drops directions whose norm
// after projection is below lindep; kept columns are compacted to the front.
int orthonormalize(Dav &d, int m, int k, double lindep, int *kept) {
  fcg_ctx *ctx = d.ctx;
  std::vector<double> S, w, U, M;
  int cur = k;
  for (int round = 0; round < 2; round++) {
    if (m > 0) {
      if (int rc = gram(d, d.V, m, d.T, cur, d.small, nullptr)) return rc;
      if (int rc = gemm(d, d.V, m, d.small, cur, -1.0, 1.0, d.T)) return rc;
    }
    const double floor2 = round == 0 ? lindep * lindep : 0.0;
    if (cur <= EMAX) {  // on the GPU: M = U[:, kept] w^-1/2 in place of S
      if (int rc = gram(d, d.T, cur, d.T, cur, d.small, nullptr)) return rc;
      hipLaunchKernelGGL(k_small_eigh, dim3(1), dim3(ET), 0, ctx->stream, cur, 0, d.small, cur,
                         1, 0, 0, floor2, d.small, nullptr, d.kept_dev + round);
      FCG_HIP(hipGetLastError(), ctx);
      int nk = cur;
      if (round == 0) {
        FCG_HIP(hipMemcpyAsync(&nk, d.kept_dev, sizeof(int), hipMemcpyDeviceToHost, ctx->stream),
                ctx);
        FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
      } else {
        d.deferred = cur;
      }
      if (nk == 0) {
        *kept = 0;
        return FCG_OK;
      }
      if (int rc = gemm(d, d.T, cur, d.small, nk, 1.0, 0.0, d.T2)) return rc;
      FCG_HIP(hipMemcpyAsync(d.T, d.T2, sizeof(double) * d.n * nk, hipMemcpyDeviceToDevice,
                             ctx->stream),
              ctx);
      cur = nk;
      continue;
    }
    if (int rc = gram(d, d.T, cur, d.T, cur, d.small, &S)) return rc;
    jacobi_eigh(cur, S, w, U);
    std::vector<int> keep;
    for (int j = cur - 1; j >= 0; j--)  // largest first
      if (w[j] > floor2 && w[j] > 0.0) keep.push_back(j);
    const int nk = (int)keep.size();
    if (nk == 0) {
      *kept = 0;
      return FCG_OK;
    }
    M.assign((size_t)cur * nk, 0.0);
    for (int c = 0; c < nk; c++) {
      const double sc = 1.0 / std::sqrt(w[keep[c]]);
      for (int i = 0; i < cur; i++) M[i + (size_t)cur * c] = U[i + (size_t)cur * keep[c]] * sc;
    }
    FCG_HIP(hipMemcpyAsync(d.small, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice,
                           ctx->stream),
            ctx);
    if (int rc = gemm(d, d.T, cur, d.small, nk, 1.0, 0.0, d.T2)) return rc;
    FCG_HIP(hipMemcpyAsync(d.T, d.T2, sizeof(double) * d.n * nk, hipMemcpyDeviceToDevice,
                           ctx->stream),
            ctx);
    cur = nk;
  }
  *kept = cur;
  return FCG_OK;
}

}  // namespace

extern "C" int fcg_davidson(fcg_ctx_t ctx, fcg_plan_t pl, int R, const double *v0,
                            const fcg_davidson_opts *o, double *evals, double *evecs,
                            fcg_davidson_stats *st) {
  if (!ctx || !pl || !v0 || !o || !evals || !evecs || R < 1 || R > FCG_MAX_R) return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (pl->ksp != pl->sp || !pl->holds_blocks) {
    ctx->err = "Davidson needs a plan within one space that holds its blocks";
    return FCG_EINVAL;
  }
  fcg_plan_note_user(pl, ctx);
  Dav d;
  d.ctx = ctx;
  const int64_t n = pl->sp->nvar;
  d.n = n;
  d.nchunk = (int)((n + GCHUNK - 1) / GCHUNK);
  const int low = o->precision;
  if (low < FCG_FP64 || low >= FCG_N_PREC) {
    ctx->err = "unknown sigma precision";
    return FCG_EINVAL;
  }
  const bool mixed = low != FCG_FP64;
  if (mixed && !(pl->precisions & (1 << low))) {
    ctx->err = "plan lacks the requested sigma precision";
    return FCG_EINVAL;
  }
  // thick restart keeps the lowest KR Ritz vectors (one sigma launch holds
  // at most FCG_MAX_R columns); the subspace is capped at 256 vectors
  const int KR = std::min(2 * R, FCG_MAX_R);
  const int mmax = std::max(KR + R, std::min(256, o->max_ss_vecs * R));
  // refresh period; 0 = only at restarts and when the residual is below the
  // threshold
  const int refresh = o->refresh_k > 0 ? o->refresh_k : INT32_MAX;
  auto alloc = [&](double **p, int64_t cnt) {
    *p = static_cast<double *>(fcg_dmalloc(ctx, sizeof(double) * std::max<int64_t>(cnt, 1)));
    return *p != nullptr;
  };
  if (!alloc(&d.V, n * mmax) || !alloc(&d.W, n * mmax) || !alloc(&d.X, n * KR) ||
      !alloc(&d.AX, n * KR) || !alloc(&d.T, n * KR) || !alloc(&d.T2, n * KR) ||
      !alloc(&d.hd, n) || !alloc(&d.small, (int64_t)mmax * mmax) || !alloc(&d.theta, R) ||
      !alloc(&d.G, (int64_t)mmax * mmax) ||
      !alloc(&d.part, (int64_t)d.nchunk * mmax * mmax) ||
      !(d.kept_dev = static_cast<int *>(fcg_dmalloc(ctx, 2 * sizeof(int))))) {
    ctx->err = "Davidson workspace allocation failed";
    return FCG_ENOMEM;
  }
  if (o->precond_diag)
    FCG_HIP(hipMemcpyAsync(d.hd, o->precond_diag, sizeof(double) * n, hipMemcpyHostToDevice,
                           ctx->stream),
            ctx);
  else if (pl->r1) {
    if (int rc = fcg_rank1_diag(ctx, pl, d.hd)) return rc;
  } else {
    FCG_HIP(fcg_launch_diag(pl, d.hd, ctx->stream), ctx);
  }
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
  std::vector<double> G, w, Y, theta(R), res(R), rr;
  int kept_host[2] = {0, 0};
  int gm = 0;  // leading columns of the subspace whose V'W entries d.G holds
  int converged = 0;
  for (int it = 1; it <= o->max_iter; it++) {
    stats.iterations = it;
    const int nk = std::min(KR, m);
    if (m <= EMAX) {
      // G = V'W, updated by the blocks of the columns added since the last
      // step, and its eigenvectors on the GPU (symmetrised there: see
      // k_small_eigh)
      if (gm < m) {
        if (int rc = gram(d, d.V, m, d.W + (int64_t)gm * n, m - gm, d.G + (int64_t)gm * mmax,
                          nullptr, mmax))
          return rc;
        if (gm > 0)
          if (int rc = gram(d, d.V + (int64_t)gm * n, m - gm, d.W, gm, d.G + gm, nullptr, mmax))
            return rc;
        gm = m;
      }
      hipLaunchKernelGGL(k_small_eigh, dim3(1), dim3(ET), 0, ctx->stream, m, ne, d.G, mmax, 0,
                         nk, R, 0.0, d.small, d.theta, nullptr);
      FCG_HIP(hipGetLastError(), ctx);
    } else {
      gm = 0;
      if (int rc = gram(d, d.V, m, d.W, m, d.small, &G)) return rc;
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
      FCG_HIP(hipMemcpyAsync(d.small, Y.data(), sizeof(double) * m * nk, hipMemcpyHostToDevice,
                             ctx->stream),
              ctx);
      FCG_HIP(hipMemcpyAsync(d.theta, theta.data(), sizeof(double) * R, hipMemcpyHostToDevice,
                             ctx->stream),
              ctx);
    }
    // lowest nk Ritz vectors X = V Y and their images AX = W Y
    if (int rc = gemm(d, d.V, m, d.small, nk, 1.0, 0.0, d.X)) return rc;
    if (int rc = gemm(d, d.W, m, d.small, nk, 1.0, 0.0, d.AX)) return rc;
    hipLaunchKernelGGL(k_residual, dim3(1024), dim3(256), 0, ctx->stream, n, R, d.X, d.AX,
                       d.theta, d.T);
    // the one synchronisation of the step: residual norms, Ritz values and
    // the deferred orthonormalisation count come back together
    FCG_HIP(hipMemcpyAsync(theta.data(), d.theta, sizeof(double) * R, hipMemcpyDeviceToHost,
                           ctx->stream),
            ctx);
    FCG_HIP(hipMemcpyAsync(kept_host, d.kept_dev, sizeof(kept_host), hipMemcpyDeviceToHost,
                           ctx->stream),
            ctx);
    if (int rc = gram(d, d.T, R, d.T, R, d.small, &rr)) return rc;
    if (int rc = check_deferred(d, kept_host)) return rc;
    double rmax = 0.0;
    for (int r = 0; r < R; r++) {
      res[r] = std::sqrt(std::max(0.0, rr[r + (size_t)R * r]));
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
      gm = 0;  // V and W are replaced
      if (!exact) {
        // collapse onto the re-orthonormalised Ritz vectors; their sigma is
        // recomputed in FP64 so the residual that decides convergence is exact
        FCG_HIP(hipMemcpyAsync(d.T, d.X, sizeof(double) * n * nk, hipMemcpyDeviceToDevice,
                               ctx->stream),
                ctx);
        if (int rc = orthonormalize(d, 0, nk, 0.0, &kept)) return rc;
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
  FCG_HIP(hipMemcpyAsync(kept_host, d.kept_dev, sizeof(kept_host), hipMemcpyDeviceToHost,
                         ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  if (int rc = check_deferred(d, kept_host)) return rc;
  for (int r = 0; r < R; r++) {
    evals[r] = theta[r];
    if (r < 16) stats.residual[r] = res[r];
  }
  stats.converged = converged;
  if (st) *st = stats;
  return converged ? FCG_OK : FCG_ENOTCONV;
}
