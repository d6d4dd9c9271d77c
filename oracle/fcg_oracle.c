/* CPU oracle: FP64 C restatement of the reference's TPSCI sigma build.
 * TEST INFRASTRUCTURE ONLY -- see fcg_oracle.h.
 *
 * Follows (reference file:line):
 *   check_term                 src/type_ClusteredTerm.jl:92-176
 *   compute_terms_state_sign   src/type_ClusteredTerm.jl:181-197
 *   contract_matrix_element    src/tpsci_inner.jl:10-216, _contract :270-328
 *   tps_ci_matvec              src/tpsci_outer.jl:462-527
 *   build_full_H               src/tpsci_outer.jl:10-50
 *   compute_diagonal           src/tpsci_outer.jl:694-715
 *   FockConfig - TransferConfig  src/Indexing.jl:21-23
 */
#include "fcg_oracle.h"

#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXN 64

/* ---- Fock-config hash (open addressing over the int16 tuple) ---- */
typedef struct {
  int64_t cap;
  int64_t *slot;
  const int16_t *fock;
  int N;
} fock_map;

static uint64_t hash_tuple(const int16_t *f, int len) {
  uint64_t h = 1469598103934665603ULL;
  for (int i = 0; i < len; i++) {
    h ^= (uint16_t)f[i];
    h *= 1099511628211ULL;
  }
  return h;
}

static void fock_map_build(fock_map *m, const int16_t *fock, int64_t n, int N) {
  m->cap = 16;
  while (m->cap < 2 * n) m->cap <<= 1;
  m->slot = (int64_t *)malloc(sizeof(int64_t) * m->cap);
  for (int64_t i = 0; i < m->cap; i++) m->slot[i] = -1;
  m->fock = fock;
  m->N = N;
  for (int64_t i = 0; i < n; i++) {
    uint64_t h = hash_tuple(fock + i * 2 * N, 2 * N) & (m->cap - 1);
    while (m->slot[h] >= 0) h = (h + 1) & (m->cap - 1);
    m->slot[h] = i;
  }
}

static int64_t fock_map_find(const fock_map *m, const int16_t *key) {
  uint64_t h = hash_tuple(key, 2 * m->N) & (m->cap - 1);
  while (m->slot[h] >= 0) {
    if (memcmp(m->fock + m->slot[h] * 2 * m->N, key, sizeof(int16_t) * 2 * m->N) == 0)
      return m->slot[h];
    h = (h + 1) & (m->cap - 1);
  }
  return -1;
}

/* ---- term logic ---- */

/* check_term: inactive clusters have equal Fock sector and equal state;
 * active clusters satisfy fock_bra = fock_ket + delta. */
static int check_term(const orc_ham *h, int t, int N, const int16_t *fb, const int16_t *bra,
                      const int16_t *fk, const int16_t *ket) {
  const int32_t *cl = h->clusters + 4 * t;
  int k = h->nbody[t];
  for (int ci = 0; ci < N; ci++) {
    int active = 0;
    for (int a = 0; a < k; a++) active |= (cl[a] == ci);
    if (active) continue;
    if (fb[2 * ci] != fk[2 * ci] || fb[2 * ci + 1] != fk[2 * ci + 1]) return 0;
    if (bra[ci] != ket[ci]) return 0;
  }
  for (int a = 0; a < k; a++) {
    int ci = cl[a];
    const int16_t *d = h->delta + 8 * t + 2 * a;
    if (fb[2 * ci] != fk[2 * ci] + d[0] || fb[2 * ci + 1] != fk[2 * ci + 1] + d[1]) return 0;
  }
  return 1;
}

/* compute_terms_state_sign: (-1)^(electrons in clusters before each odd op),
 * electrons counted in the KET Fock config. */
static double state_sign(const orc_ham *h, int t, const int16_t *fk) {
  int s = 0;
  const int32_t *cl = h->clusters + 4 * t;
  for (int a = 0; a < h->nbody[t]; a++) {
    if (h->parity[4 * t + a] != 1) continue;
    int ne = 0;
    for (int ci = 0; ci < cl[a]; ci++) ne += fk[2 * ci] + fk[2 * ci + 1];
    s ^= (ne & 1);
  }
  return s ? -1.0 : 1.0;
}

static const double *gamma_block(const orc_ops *o, int c, int op, const int16_t *fb,
                                  const int16_t *fk, const int32_t **dims) {
  int S = o->S, SS = S * S;
  int cb = fb[2 * c] * S + fb[2 * c + 1];
  int ck = fk[2 * c] * S + fk[2 * c + 1];
  if (fb[2 * c] < 0 || fb[2 * c + 1] < 0 || fk[2 * c] < 0 || fk[2 * c + 1] < 0 ||
      fb[2 * c] >= S || fb[2 * c + 1] >= S || fk[2 * c] >= S || fk[2 * c + 1] >= S)
    return NULL;
  int64_t id = o->gblock[(((int64_t)c * o->n_ops + op) * SS + cb) * SS + ck];
  if (id < 0) return NULL;
  *dims = o->gb_dims + 3 * id;
  return o->gdata + o->gb_off[id];
}

/* contract_matrix_element; returns 0 on success, -3 on a missing block. */
static int matrix_element(const orc_ops *o, const orc_ham *h, int t, const int16_t *fb,
                          const int16_t *bra, const int16_t *fk, const int16_t *ket,
                          double *out) {
  int k = h->nbody[t];
  const int32_t *cl = h->clusters + 4 * t;
  const int32_t *op = h->ops + 4 * t;
  const double *g[4];
  for (int a = 0; a < k; a++) {
    const int32_t *d;
    const double *blk = gamma_block(o, cl[a], op[a], fb, fk, &d);
    if (!blk) return -3;
    int npd = d[0], mb = d[1];
    int I = bra[cl[a]] - 1, J = ket[cl[a]] - 1;
    g[a] = blk + (int64_t)npd * (I + (int64_t)mb * J);
  }
  if (k == 1) {
    *out = g[0][0];
    return 0;
  }
  const double *V = h->ints + h->ioff[t];
  const int32_t *dm = h->dims + 4 * t;
  double me = 0.0;
  if (k == 2) {
    for (int j = 0; j < dm[1]; j++) {
      double tmp = g[1][j];
      for (int i = 0; i < dm[0]; i++) me += g[0][i] * V[i + dm[0] * j] * tmp;
    }
  } else if (k == 3) {
    for (int kk = 0; kk < dm[2]; kk++)
      for (int j = 0; j < dm[1]; j++) {
        double tmp = g[1][j] * g[2][kk];
        const double *Vj = V + dm[0] * (j + (int64_t)dm[1] * kk);
        for (int i = 0; i < dm[0]; i++) me += g[0][i] * Vj[i] * tmp;
      }
  } else {
    for (int l = 0; l < dm[3]; l++)
      for (int kk = 0; kk < dm[2]; kk++) {
        double tmp2 = g[2][kk] * g[3][l];
        for (int j = 0; j < dm[1]; j++) {
          double tmp = g[1][j] * tmp2;
          const double *Vj = V + dm[0] * (j + (int64_t)dm[1] * (kk + (int64_t)dm[2] * l));
          for (int i = 0; i < dm[0]; i++) me += g[0][i] * Vj[i] * tmp;
        }
      }
  }
  *out = state_sign(h, t, fk) * me;
  return 0;
}

static int n_clusters_of(const orc_ops *o) { return o->n_clusters; }

/* sigma row for one bra config (the body of do_job, tpsci_outer.jl:475-510) */
static int sigma_row(const orc_ops *o, const orc_ham *h, const orc_space *sp,
                     const fock_map *fm, int64_t fbi, int64_t bi, int R, const double *c,
                     double *out /* R */) {
  int N = n_clusters_of(o);
  const int16_t *fb = sp->fock + fbi * 2 * N;
  const int16_t *bra = sp->cfg + bi * N;
  int16_t fk[2 * MAXN];
  for (int r = 0; r < R; r++) out[r] = 0.0;
  for (int g = 0; g < h->n_tc; g++) {
    const int16_t *tc = h->tc + (int64_t)g * 2 * N;
    for (int i = 0; i < 2 * N; i++) fk[i] = fb[i] - tc[i];
    int64_t fki = fock_map_find(fm, fk);
    if (fki < 0) continue;
    for (int64_t kj = sp->fock_start[fki]; kj < sp->fock_start[fki + 1]; kj++) {
      const int16_t *ket = sp->cfg + kj * N;
      for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) {
        if (!check_term(h, t, N, fb, bra, fk, ket)) continue;
        double me;
        int rc = matrix_element(o, h, t, fb, bra, fk, ket, &me);
        if (rc) return rc;
        for (int r = 0; r < R; r++) out[r] += me * c[(int64_t)r * sp->n_cfg + kj];
      }
    }
  }
  return 0;
}

static int64_t *cfg_to_fock(const orc_space *sp) {
  int64_t *f = (int64_t *)malloc(sizeof(int64_t) * (sp->n_cfg > 0 ? sp->n_cfg : 1));
  for (int64_t fi = 0; fi < sp->n_fock; fi++)
    for (int64_t i = sp->fock_start[fi]; i < sp->fock_start[fi + 1]; i++) f[i] = fi;
  return f;
}

int orc_tpsci_sigma_rows(const orc_ops *o, const orc_ham *h, const orc_space *sp, int R,
                         const double *c, const int64_t *rows, int64_t n_rows,
                         double *sigma_rows, int nthreads) {
  if (o->n_clusters > MAXN) return -1;
  fock_map fm;
  fock_map_build(&fm, sp->fock, sp->n_fock, o->n_clusters);
  int64_t *c2f = cfg_to_fock(sp);
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t q = 0; q < n_rows; q++) {
    double out[64];
    int64_t bi = rows[q];
    int rc = sigma_row(o, h, sp, &fm, c2f[bi], bi, R, c, out);
    if (rc) err = rc;
    for (int r = 0; r < R; r++) sigma_rows[(int64_t)r * n_rows + q] = out[r];
  }
  free(c2f);
  free(fm.slot);
  return err;
}

int orc_tpsci_sigma(const orc_ops *o, const orc_ham *h, const orc_space *sp, int R,
                    const double *c, double *sigma, int nthreads) {
  if (R > 64) return -1;
  int64_t *rows = (int64_t *)malloc(sizeof(int64_t) * (sp->n_cfg > 0 ? sp->n_cfg : 1));
  for (int64_t i = 0; i < sp->n_cfg; i++) rows[i] = i;
  int rc = orc_tpsci_sigma_rows(o, h, sp, R, c, rows, sp->n_cfg, sigma, nthreads);
  free(rows);
  return rc;
}

int orc_build_full_H(const orc_ops *o, const orc_ham *h, const orc_space *sp, double *H,
                     int nthreads) {
  int N = o->n_clusters;
  if (N > MAXN) return -1;
  fock_map fm;
  fock_map_build(&fm, sp->fock, sp->n_fock, N);
  int64_t n = sp->n_cfg;
  int64_t *c2f = cfg_to_fock(sp);
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t bi = 0; bi < n; bi++) {
    const int16_t *fb = sp->fock + c2f[bi] * 2 * N;
    const int16_t *bra = sp->cfg + bi * N;
    for (int64_t kj = 0; kj <= bi; kj++) {
      const int16_t *fk = sp->fock + c2f[kj] * 2 * N;
      const int16_t *ket = sp->cfg + kj * N;
      int16_t tc[2 * MAXN];
      for (int i = 0; i < 2 * N; i++) tc[i] = fb[i] - fk[i];
      double acc = 0.0;
      for (int g = 0; g < h->n_tc; g++) {
        if (memcmp(h->tc + (int64_t)g * 2 * N, tc, sizeof(int16_t) * 2 * N) != 0) continue;
        for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) {
          if (!check_term(h, t, N, fb, bra, fk, ket)) continue;
          double me;
          int rc = matrix_element(o, h, t, fb, bra, fk, ket, &me);
          if (rc) err = rc;
          else acc += me;
        }
      }
      H[bi * n + kj] = acc;
      H[kj * n + bi] = acc;
    }
  }
  free(c2f);
  free(fm.slot);
  return err;
}

int orc_tpsci_diagonal(const orc_ops *o, const orc_ham *h, const orc_space *sp, int zero_tc,
                       double *hd) {
  int N = o->n_clusters;
  for (int64_t fi = 0; fi < sp->n_fock; fi++) {
    const int16_t *f = sp->fock + fi * 2 * N;
    for (int64_t i = sp->fock_start[fi]; i < sp->fock_start[fi + 1]; i++) {
      const int16_t *cfg = sp->cfg + i * N;
      double acc = 0.0;
      for (int t = h->tc_start[zero_tc]; t < h->tc_start[zero_tc + 1]; t++) {
        double me;
        int rc = matrix_element(o, h, t, f, cfg, f, cfg, &me);
        if (rc) return rc;
        acc += me;
      }
      hd[i] = acc;
    }
  }
  return 0;
}
