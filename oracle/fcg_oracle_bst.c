/* CPU oracle: FP64 C restatement of the reference's BST sigma build.
 * TEST INFRASTRUCTURE ONLY -- see fcg_oracle.h.
 *
 * Follows (reference file:line):
 *   build_sigma!                  src/tucker_inner.jl:165-245 (one job per bra block)
 *   form_sigma_block!             src/tucker_inner.jl:253-304
 *   check_term (TuckerConfig)     src/type_ClusteredTerm.jl:92-176
 *   compute_terms_state_sign      src/type_ClusteredTerm.jl:181-197
 *   build_dense_H_term 1B-4B      src/tucker_build_dense_H_term.jl:5-174
 *     g[p,k,l] = sum_IJ Ul[I,k] Γ[p,I,J] Ur[J,l], then the integrals are
 *     contracted one active cluster at a time into op(ket..., bra...)
 *   contract_dense_H_with_state   src/tucker_contract_dense_H_with_state.jl:4-329
 *     overlaps S = Ub' Uk on inactive clusters (transform_basis,
 *     src/hosvd.jl:634-670), then bra(active, rest) += s op2' ket(active, rest)
 */
#define _POSIX_C_SOURCE 199309L
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "fcg_oracle.h"

#define MAXN 32

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static uint64_t hash_t(const int16_t *f, int len) {
  uint64_t h = 1469598103934665603ULL;
  for (int i = 0; i < len; i++) { h ^= (uint16_t)f[i]; h *= 1099511628211ULL; }
  return h;
}

typedef struct { int64_t cap; int64_t *slot; const int16_t *fock; int N; } fmap_t;

static void fmap_build(fmap_t *m, const int16_t *fock, int64_t n, int N) {
  m->cap = 16;
  while (m->cap < 2 * n) m->cap <<= 1;
  m->slot = (int64_t *)malloc(sizeof(int64_t) * m->cap);
  for (int64_t i = 0; i < m->cap; i++) m->slot[i] = -1;
  m->fock = fock; m->N = N;
  for (int64_t i = 0; i < n; i++) {
    uint64_t h = hash_t(fock + i * 2 * N, 2 * N) & (m->cap - 1);
    while (m->slot[h] >= 0) h = (h + 1) & (m->cap - 1);
    m->slot[h] = i;
  }
}

static int64_t fmap_find(const fmap_t *m, const int16_t *key) {
  uint64_t h = hash_t(key, 2 * m->N) & (m->cap - 1);
  while (m->slot[h] >= 0) {
    if (!memcmp(m->fock + m->slot[h] * 2 * m->N, key, sizeof(int16_t) * 2 * m->N)) return m->slot[h];
    h = (h + 1) & (m->cap - 1);
  }
  return -1;
}

static double ufac(const orc_bst *b, int64_t blk, int c, int N, int x, int i) {
  int64_t off = b->foff[blk * N + c];
  int nrg = b->hi[blk * N + c] - b->lo[blk * N + c] + 1;
  return off < 0 ? (x == i ? 1.0 : 0.0) : b->factors[off + x + (int64_t)nrg * i];
}

/* g[p + np*(k + rb*l)] = sum_{I,J} Ul[I,k] Γ[p, lo_b+I, lo_k+J] Ur[J,l] */
static void compress_gamma(const orc_bst *b, int N, int c, int64_t B, int64_t K, const double *G,
                           int np, int mb, double *g, double *tmp) {
  const int lb = b->lo[B * N + c] - 1, lk = b->lo[K * N + c] - 1;
  const int nI = b->hi[B * N + c] - lb, nJ = b->hi[K * N + c] - lk;
  const int rb = b->rank[B * N + c], rk = b->rank[K * N + c];
  /* tmp[p,k,J] = sum_I Ul[I,k] Γ[p,I,J] */
  for (int J = 0; J < nJ; J++)
    for (int k = 0; k < rb; k++)
      for (int p = 0; p < np; p++) {
        double s = 0.0;
        for (int I = 0; I < nI; I++)
          s += ufac(b, B, c, N, I, k) * G[p + (int64_t)np * ((lb + I) + (int64_t)mb * (lk + J))];
        tmp[p + np * (k + rb * J)] = s;
      }
  /* g[p,k,l] = sum_J Ur[J,l] tmp[p,k,J] */
  for (int l = 0; l < rk; l++)
    for (int k = 0; k < rb; k++)
      for (int p = 0; p < np; p++) {
        double s = 0.0;
        for (int J = 0; J < nJ; J++) s += ufac(b, K, c, N, J, l) * tmp[p + np * (k + rb * J)];
        g[p + np * (k + rb * l)] = s;
      }
}

typedef struct { double *a, *b, *g, *tmp, *op, *ket2, *ket3; } scratch_t;

/* op^T in (bra active multi-index) x (ket active multi-index) layout:
 * opT[I_A + dbA * J_A], I_A and J_A column-major over active clusters. */
static int build_op(const orc_ops *o, const orc_ham *h, const orc_bst *b, int N, int t,
                    const int16_t *fb, const int16_t *fk, int64_t B, int64_t K, scratch_t *s,
                    double *opT, int *dbA_out, int *dkA_out) {
  const int nb = h->nbody[t], S = o->S, SS = S * S;
  const int32_t *cl = h->clusters + 4 * t;
  double *cur = s->a, *nxt = s->b;
  int rest = 1;
  if (nb == 1) cur[0] = 1.0;
  else {
    for (int a = 0; a < nb; a++) rest *= h->dims[4 * t + a];
    memcpy(cur, h->ints + h->ioff[t], sizeof(double) * rest);
  }
  int dbA = 1, dkA = 1;
  int rbs[4], rks[4];
  for (int a = 0; a < nb; a++) {
    const int c = cl[a];
    const int cb = fb[2 * c] * S + fb[2 * c + 1], ck = fk[2 * c] * S + fk[2 * c + 1];
    const int64_t blk = o->gblock[(((int64_t)c * o->n_ops + h->ops[4 * t + a]) * SS + cb) * SS + ck];
    if (blk < 0) return -3;
    const int np = o->gb_dims[3 * blk], mb = o->gb_dims[3 * blk + 1];
    const int rb = b->rank[B * N + c], rk = b->rank[K * N + c];
    compress_gamma(b, N, c, B, K, o->gdata + o->gb_off[blk], np, mb, s->g, s->tmp);
    /* contract the leading integral index with g: nxt[q, (k,l)] */
    const int rest2 = rest / np, PP = rb * rk;
    for (int P = 0; P < PP; P++)
      for (int q = 0; q < rest2; q++) {
        double v = 0.0;
        for (int p = 0; p < np; p++) v += cur[p + np * q] * s->g[p + np * P];
        nxt[q + rest2 * P] = v;
      }
    double *x = cur; cur = nxt; nxt = x;
    rest = rest2 * PP;
    rbs[a] = rb; rks[a] = rk;
    dbA *= rb; dkA *= rk;
  }
  /* cur[(k1,l1),(k2,l2),...] -> opT[(k1..kn),(l1..ln)] */
  for (int e = 0; e < rest; e++) {
    int r = e, ib = 0, ik = 0, sb = 1, sk = 1;
    for (int a = 0; a < nb; a++) {
      const int P = r % (rbs[a] * rks[a]);
      r /= rbs[a] * rks[a];
      ib += (P % rbs[a]) * sb; sb *= rbs[a];
      ik += (P / rbs[a]) * sk; sk *= rks[a];
    }
    opT[ib + (int64_t)dbA * ik] = cur[e];
  }
  *dbA_out = dbA; *dkA_out = dkA;
  return 0;
}

static double term_sign(const orc_ham *h, int t, const int16_t *fk) {
  int s = 0;
  for (int a = 0; a < h->nbody[t]; a++) {
    if (h->parity[4 * t + a] != 1) continue;
    int ne = 0;
    for (int c = 0; c < h->clusters[4 * t + a]; c++) ne += fk[2 * c] + fk[2 * c + 1];
    s ^= ne & 1;
  }
  return s ? -1.0 : 1.0;
}

/* contract_dense_H_with_state for one root-block: bra_core += s opT * ket'
 * where ket' = ket core transformed by inactive overlaps. */
static void contract(const orc_bst *b, int N, const int32_t *cl, int nb, int64_t B, int64_t K,
                     double sign, const double *opT, int dbA, int dkA, const double *ket,
                     double *bra, scratch_t *s) {
  int rb[MAXN], rk[MAXN], act[MAXN];
  for (int c = 0; c < N; c++) { rb[c] = b->rank[B * N + c]; rk[c] = b->rank[K * N + c]; act[c] = 0; }
  for (int a = 0; a < nb; a++) act[cl[a]] = 1;
  /* transform_basis: inactive mode c: (rk) -> (rb) with S = Ub' Uk */
  int dims[MAXN];
  int64_t tot = 1;
  for (int c = 0; c < N; c++) { dims[c] = rk[c]; tot *= rk[c]; }
  double *cur = s->ket2, *nxt = s->ket3;
  memcpy(cur, ket, sizeof(double) * tot);
  for (int c = 0; c < N; c++) {
    if (act[c]) continue;
    const int nrg = b->hi[B * N + c] - b->lo[B * N + c] + 1;
    int64_t inner = 1, outer = 1;
    for (int d = 0; d < c; d++) inner *= dims[d];
    for (int d = c + 1; d < N; d++) outer *= dims[d];
    for (int64_t o2 = 0; o2 < outer; o2++)
      for (int i = 0; i < rb[c]; i++)
        for (int64_t in = 0; in < inner; in++) {
          double v = 0.0;
          for (int j = 0; j < rk[c]; j++) {
            double S = 0.0;
            for (int x = 0; x < nrg; x++) S += ufac(b, B, c, N, x, i) * ufac(b, K, c, N, x, j);
            v += S * cur[in + inner * (j + (int64_t)dims[c] * o2)];
          }
          nxt[in + inner * (i + (int64_t)rb[c] * o2)] = v;
        }
    dims[c] = rb[c];
    double *x = cur; cur = nxt; nxt = x;
  }
  /* bra[I] += sign * sum_{J_A} opT[I_A, J_A] ket'[J_A, I_inactive] */
  int64_t dbt = 1;
  for (int c = 0; c < N; c++) dbt *= rb[c];
  for (int64_t e = 0; e < dbt; e++) {
    int64_t r = e, ia = 0, sa = 1, kin = 0, sk = 1;
    int64_t strides[MAXN];
    int64_t st = 1;
    for (int c = 0; c < N; c++) { strides[c] = st; st *= dims[c]; }
    for (int c = 0; c < N; c++) {
      const int i = (int)(r % rb[c]);
      r /= rb[c];
      if (act[c]) { ia += i * sa; sa *= rb[c]; }
      else kin += i * strides[c];
    }
    (void)sk;
    double v = 0.0;
    for (int jA = 0; jA < dkA; jA++) {
      int64_t rr = jA, kidx = kin;
      for (int c = 0; c < N; c++) {
        if (!act[c]) continue;
        kidx += (rr % rk[c]) * strides[c];
        rr /= rk[c];
      }
      v += opT[ia + (int64_t)dbA * jA] * cur[kidx];
    }
    bra[e] += sign * v;
  }
}

int orc_bst_sigma_rows(const orc_ops *o, const orc_ham *h, const orc_bst *b, int R,
                       const double *c, int64_t nvar, const int64_t *rows, int64_t n_rows,
                       double *out, int nthreads, double *t_build, double *t_contract) {
  const int N = o->n_clusters;
  if (N > MAXN) return -1;
  fmap_t fm;
  fmap_build(&fm, b->fock, b->n_fock, N);
  /* blocks of each Fock configuration */
  int64_t *fcount = (int64_t *)calloc(b->n_fock + 1, sizeof(int64_t));
  for (int64_t k = 0; k < b->n_box; k++) fcount[b->box_fock[k] + 1]++;
  for (int64_t f = 0; f < b->n_fock; f++) fcount[f + 1] += fcount[f];
  int64_t *flist = (int64_t *)malloc(sizeof(int64_t) * (b->n_box > 0 ? b->n_box : 1));
  int64_t *fill = (int64_t *)calloc(b->n_fock, sizeof(int64_t));
  for (int64_t k = 0; k < b->n_box; k++) {
    const int f = b->box_fock[k];
    flist[fcount[f] + fill[f]++] = k;
  }
  int64_t *row_off = (int64_t *)malloc(sizeof(int64_t) * (n_rows + 1));
  row_off[0] = 0;
  for (int64_t q = 0; q < n_rows; q++)
    row_off[q + 1] = row_off[q] + (b->var_off[rows[q] + 1] - b->var_off[rows[q]]);
  const int64_t nout = row_off[n_rows];
  memset(out, 0, sizeof(double) * nout * R);
  double tb = 0.0, tc = 0.0;
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel reduction(+ : tb, tc)
  {
    scratch_t s;
    const size_t big = 1 << 17;
    s.a = malloc(8 * big); s.b = malloc(8 * big); s.g = malloc(8 * big); s.tmp = malloc(8 * big);
    s.op = malloc(8 * big); s.ket2 = malloc(8 * big); s.ket3 = malloc(8 * big);
#pragma omp for schedule(dynamic, 1)
    for (int64_t q = 0; q < n_rows; q++) {
      const int64_t B = rows[q];
      const int16_t *fb = b->fock + (int64_t)b->box_fock[B] * 2 * N;
      int16_t fk[2 * MAXN];
      for (int g = 0; g < h->n_tc; g++) {
        const int16_t *tcg = h->tc + (int64_t)g * 2 * N;
        for (int i = 0; i < 2 * N; i++) fk[i] = fb[i] - tcg[i];
        const int64_t fki = fmap_find(&fm, fk);
        if (fki < 0) continue;
        for (int64_t kk = fcount[fki]; kk < fcount[fki + 1]; kk++) {
          const int64_t K = flist[kk];
          for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) {
            /* check_term on TuckerConfigs: inactive ranges equal */
            const int32_t *cl = h->clusters + 4 * t;
            int ok = 1;
            for (int cc = 0; cc < N && ok; cc++) {
              int a_ = 0;
              for (int a = 0; a < h->nbody[t]; a++) a_ |= cl[a] == cc;
              if (a_) continue;
              ok = b->lo[B * N + cc] == b->lo[K * N + cc] && b->hi[B * N + cc] == b->hi[K * N + cc];
            }
            if (!ok) continue;
            const double t0 = now_s();
            int dbA, dkA;
            int rc = build_op(o, h, b, N, t, fb, fk, B, K, &s, s.op, &dbA, &dkA);
            const double t1 = now_s();
            if (rc) { err = rc; continue; }
            const double sign = term_sign(h, t, fk);
            for (int r = 0; r < R; r++)
              contract(b, N, cl, h->nbody[t], B, K, sign, s.op, dbA, dkA,
                       c + (int64_t)r * nvar + b->var_off[K], out + (int64_t)r * nout + row_off[q], &s);
            tb += t1 - t0;
            tc += now_s() - t1;
          }
        }
      }
    }
    free(s.a); free(s.b); free(s.g); free(s.tmp); free(s.op); free(s.ket2); free(s.ket3);
  }
  if (t_build) *t_build = tb;
  if (t_contract) *t_contract = tc;
  free(fcount); free(flist); free(fill); free(row_off); free(fm.slot);
  return err;
}
