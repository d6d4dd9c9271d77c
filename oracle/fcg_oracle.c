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
 *   compute_pt2_energy         src/tpsci_pt2_energy.jl:10-175, _pt2_job :178-266,
 *                              _sum_pt2 :269-282 (dense target blocks, and
 *                              over the configurations reached)
 *   contract_matvec_thread     src/tpsci_matvec_thread.jl:279-820 (prescreen=false)
 *   _collect_significant_thread!  src/tpsci_matvec_thread.jl:830-895
 *   project_out!               src/tpsci_outer.jl:869-889
 *   compute_diagonal! (op)     src/tpsci_outer.jl:735-757
 */
#include "fcg_oracle.h"

#include <math.h>
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

/* ---- second-order energy over the first-order space ---- */

/* insert-or-find on an open-addressing table of Fock tuples stored in `keys` */
static int64_t fock_intern(int64_t *slot, int64_t cap, int16_t *keys, int64_t *n_keys,
                           const int16_t *f, int N) {
  uint64_t hsh = hash_tuple(f, 2 * N) & (cap - 1);
  while (slot[hsh] >= 0) {
    if (memcmp(keys + slot[hsh] * 2 * N, f, sizeof(int16_t) * 2 * N) == 0) return slot[hsh];
    hsh = (hsh + 1) & (cap - 1);
  }
  memcpy(keys + *n_keys * 2 * N, f, sizeof(int16_t) * 2 * N);
  slot[hsh] = *n_keys;
  return (*n_keys)++;
}

/* contract_matvec_thread for one term and one ket configuration
 * (src/tpsci_matvec_thread.jl:279-820 with prescreen=false): the
 * integrals contracted with the ket column of each active cluster's Γ,
 * giving T[I1..Ik] over the active clusters' bra states; elements above
 * thresh / max|c_ket| (_collect_significant_thread!, :830-895) are added,
 * times the state sign and c_ket, to the dense sigma of the target Fock
 * block.  Returns -3 when a 1-body Γ block is missing (the reference reads
 * it unconditionally); multi-body terms with a missing block are skipped
 * (their `haskey(...) || return`). */
/* configuration -> sigma[R] map of one target Fock block, keyed by the
 * configuration's linear index in the block (first cluster fastest);
 * open addressing, insertion order kept in `order` */
typedef struct {
  int64_t cap, n, R, adds;  /* adds: elements accumulated (kept contributions) */
  int64_t *key;   /* [cap], -1 = empty */
  int64_t *idx;   /* [cap] entry of the slot */
  int64_t *order; /* [n] keys by insertion */
  double *val;    /* [n][R] */
} lin_map;

static void lin_map_init(lin_map *m, int R) {
  m->cap = 1024;
  m->n = 0;
  m->adds = 0;
  m->R = R;
  m->key = (int64_t *)malloc(sizeof(int64_t) * m->cap);
  m->idx = (int64_t *)malloc(sizeof(int64_t) * m->cap);
  for (int64_t i = 0; i < m->cap; i++) m->key[i] = -1;
  m->order = (int64_t *)malloc(sizeof(int64_t) * (m->cap / 2));
  m->val = (double *)malloc(sizeof(double) * (m->cap / 2) * R);
}

static void lin_map_free(lin_map *m) {
  free(m->key);
  free(m->idx);
  free(m->order);
  free(m->val);
}

static int64_t lin_map_slot(const lin_map *m, int64_t k) {
  uint64_t hsh = ((uint64_t)k * 0x9E3779B97F4A7C15ull) >> 17;
  int64_t i = (int64_t)(hsh & (uint64_t)(m->cap - 1));
  while (m->key[i] >= 0 && m->key[i] != k) i = (i + 1) & (m->cap - 1);
  return i;
}

/* the sigma entry of configuration k, inserted as zeros when absent */
static double *lin_map_get(lin_map *m, int64_t k) {
  int64_t i = lin_map_slot(m, k);
  if (m->key[i] == k) return m->val + m->idx[i] * m->R;
  if (2 * (m->n + 1) > m->cap) {
    const int64_t old = m->cap;
    int64_t *ok = m->key, *oi = m->idx;
    m->cap *= 2;
    m->key = (int64_t *)malloc(sizeof(int64_t) * m->cap);
    m->idx = (int64_t *)malloc(sizeof(int64_t) * m->cap);
    for (int64_t j = 0; j < m->cap; j++) m->key[j] = -1;
    for (int64_t j = 0; j < old; j++)
      if (ok[j] >= 0) {
        const int64_t s = lin_map_slot(m, ok[j]);
        m->key[s] = ok[j];
        m->idx[s] = oi[j];
      }
    free(ok);
    free(oi);
    m->order = (int64_t *)realloc(m->order, sizeof(int64_t) * (m->cap / 2));
    m->val = (double *)realloc(m->val, sizeof(double) * (m->cap / 2) * m->R);
    i = lin_map_slot(m, k);
  }
  m->key[i] = k;
  m->idx[i] = m->n;
  m->order[m->n] = k;
  double *v = m->val + m->n * m->R;
  for (int64_t r = 0; r < m->R; r++) v[r] = 0.0;
  m->n++;
  return v;
}

static int pt2_term(const orc_ops *o, const orc_ham *h, int t, const int16_t *fx,
                    const int16_t *fk, const int16_t *ket, const double *cket /* R */, int R,
                    double thresh, const int64_t *stride, int64_t nx, double *sig,
                    unsigned char *hit /* nullable: marks inserted configurations */,
                    lin_map *hm /* non-null: accumulate here instead of sig */,
                    double **w1, double **w2, int64_t *wcap) {
  const int N = o->n_clusters;
  const int k = h->nbody[t];
  const int32_t *cl = h->clusters + 4 * t;
  const int32_t *op = h->ops + 4 * t;
  const double *col[4];
  int np[4], mb[4];
  for (int a = 0; a < k; a++) {
    const int32_t *d;
    const double *blk = gamma_block(o, cl[a], op[a], fx, fk, &d);
    if (!blk) return k == 1 ? -3 : 0;
    np[a] = d[0];
    mb[a] = d[1];
    col[a] = blk + (int64_t)d[0] * d[1] * (ket[cl[a]] - 1);
  }
  /* workspace: the largest intermediate of the contraction below */
  int64_t need = k == 1 ? mb[0] : 1;
  if (k > 1) {
    for (int m = k; m >= 0; m--) {
      int64_t sz = 1;
      for (int a = 0; a < k; a++) sz *= a < m ? np[a] : mb[a];
      need = sz > need ? sz : need;
    }
  }
  if (need > *wcap) {
    free(*w1);
    free(*w2);
    *w1 = (double *)malloc(sizeof(double) * need);
    *w2 = (double *)malloc(sizeof(double) * need);
    *wcap = need;
  }
  /* Y[p1..pm, I(m+1)..Ik], contracted from the last active cluster inwards */
  double *Y = *w1, *Z = *w2;
  int64_t ny;
  if (k == 1) {
    ny = mb[0];
    for (int64_t i = 0; i < ny; i++) Y[i] = col[0][i];
  } else {
    const double *V = h->ints + h->ioff[t];
    ny = 1;
    for (int a = 0; a < k; a++) ny *= np[a];
    for (int64_t i = 0; i < ny; i++) Y[i] = V[i];
    int64_t B = 1;
    for (int m = k - 1; m >= 0; m--) {
      int64_t A = 1;
      for (int a = 0; a < m; a++) A *= np[a];
      const int P = np[m], M = mb[m];
      for (int64_t b = 0; b < B; b++)
        for (int i = 0; i < M; i++)
          for (int64_t a = 0; a < A; a++) {
            double acc = 0.0;
            for (int p = 0; p < P; p++) acc += Y[a + A * (p + (int64_t)P * b)] * col[m][p + (int64_t)P * i];
            Z[a + A * (i + (int64_t)M * b)] = acc;
          }
      B *= M;
      double *tmp = Y;
      Y = Z;
      Z = tmp;
    }
    ny = B;
  }
  const double sgn = state_sign(h, t, fk);
  double cmax = 0.0;
  for (int r = 0; r < R; r++) cmax = fabs(cket[r]) > cmax ? fabs(cket[r]) : cmax;
  const double tcur = thresh / cmax;
  /* base linear index: the ket configuration with the active clusters zeroed */
  int64_t base = 0;
  for (int c = 0; c < N; c++) {
    int act = 0;
    for (int a = 0; a < k; a++) act |= cl[a] == c;
    if (!act) base += (int64_t)(ket[c] - 1) * stride[c];
  }
  for (int64_t e = 0; e < ny; e++) {
    const double v = Y[e];
    if (!(fabs(v) > tcur)) continue;
    int64_t lin = base, r0 = e;
    for (int a = 0; a < k; a++) {
      lin += (r0 % mb[a]) * stride[cl[a]];
      r0 /= mb[a];
    }
    if (hm) {
      hm->adds++;
      double *acc = lin_map_get(hm, lin);
      for (int r = 0; r < R; r++) acc[r] += sgn * v * cket[r];
    } else {
      for (int r = 0; r < R; r++) sig[lin + nx * r] += sgn * v * cket[r];
    }
    if (hit) hit[lin] = 1;
  }
  return 0;
}

int orc_tpsci_pt2(const orc_ops *o, const orc_ham *h, const orc_space *sp, int R,
                  const double *c, const double *e0, int nbody, int h0_op, double thresh,
                  double *e2, int64_t *n_jobs, int nthreads) {
  const int N = o->n_clusters, S = o->S;
  if (N > MAXN || R < 1 || R > 64) return -1;
  /* jobs: target Fock blocks F_ket + TransferConfig (tpsci_pt2_energy.jl:49-79) */
  const int64_t n_cand = sp->n_fock * (int64_t)h->n_tc;
  int64_t cap = 16;
  while (cap < 2 * n_cand) cap <<= 1;
  int64_t *slot = (int64_t *)malloc(sizeof(int64_t) * cap);
  for (int64_t i = 0; i < cap; i++) slot[i] = -1;
  int16_t *jfock = (int16_t *)malloc(sizeof(int16_t) * 2 * N * (n_cand > 0 ? n_cand : 1));
  int64_t *cand_job = (int64_t *)malloc(sizeof(int64_t) * (n_cand > 0 ? n_cand : 1));
  int64_t nj = 0;
  for (int64_t f = 0; f < sp->n_fock; f++) {
    const int16_t *fk = sp->fock + f * 2 * N;
    for (int g = 0; g < h->n_tc; g++) {
      int16_t fx[2 * MAXN];
      int ok = 1;
      for (int i = 0; i < 2 * N && ok; i++) {
        fx[i] = fk[i] + h->tc[(int64_t)g * 2 * N + i];
        ok = fx[i] >= 0 && fx[i] < S;
      }
      for (int ci = 0; ci < N && ok; ci++) {
        const int32_t *d;
        ok = gamma_block(o, ci, 0, fx, fx, &d) != NULL;
      }
      cand_job[f * h->n_tc + g] = ok ? fock_intern(slot, cap, jfock, &nj, fx, N) : -1;
    }
  }
  fock_map vm;
  fock_map_build(&vm, sp->fock, sp->n_fock, N);
  double *je2 = (double *)calloc((size_t)(nj > 0 ? nj : 1) * R, sizeof(double));
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t j = 0; j < nj; j++) {
    const int16_t *fx = jfock + j * 2 * N;
    int64_t stride[MAXN], nx = 1;
    int M[MAXN];
    for (int ci = 0; ci < N; ci++) {
      const int32_t *d = NULL;
      if (!gamma_block(o, ci, 0, fx, fx, &d)) continue; /* excluded when the job was formed */
      M[ci] = d[1];
      stride[ci] = nx;
      nx *= M[ci];
    }
    double *sig = (double *)calloc((size_t)nx * R, sizeof(double));
    double *w1 = NULL, *w2 = NULL;
    int64_t wcap = 0;
    double cket[64];
    int rc = 0;
    /* _pt2_job: every (TransferConfig group, ket Fock block) feeding F_x */
    for (int64_t f = 0; f < sp->n_fock && !rc; f++) {
      for (int g = 0; g < h->n_tc && !rc; g++) {
        if (cand_job[f * h->n_tc + g] != j) continue;
        const int16_t *fk = sp->fock + f * 2 * N;
        for (int t = h->tc_start[g]; t < h->tc_start[g + 1] && !rc; t++) {
          if (h->nbody[t] > nbody) continue;
          for (int64_t kj = sp->fock_start[f]; kj < sp->fock_start[f + 1] && !rc; kj++) {
            for (int r = 0; r < R; r++) cket[r] = c[(int64_t)r * sp->n_cfg + kj];
            rc = pt2_term(o, h, t, fx, fk, sp->cfg + kj * N, cket, R, thresh, stride, nx, sig,
                          NULL, NULL, &w1, &w2, &wcap);
          }
        }
      }
    }
    /* project_out! (src/tpsci_outer.jl:869-889): drop variational configurations */
    unsigned char *out = (unsigned char *)calloc((size_t)nx, 1);
    const int64_t vf = fock_map_find(&vm, fx);
    if (vf >= 0)
      for (int64_t i = sp->fock_start[vf]; i < sp->fock_start[vf + 1]; i++) {
        int64_t lin = 0;
        for (int ci = 0; ci < N; ci++) lin += (int64_t)(sp->cfg[i * N + ci] - 1) * stride[ci];
        out[lin] = 1;
      }
    /* compute_diagonal! with the H0 operator (src/tpsci_outer.jl:735-757) */
    const double *hb[MAXN];
    for (int ci = 0; ci < N && !rc; ci++) {
      const int32_t *d;
      hb[ci] = gamma_block(o, ci, h0_op, fx, fx, &d);
      if (!hb[ci]) rc = -3;
    }
    /* _sum_pt2 (src/tpsci_pt2_energy.jl:269-282) */
    if (!rc)
      for (int64_t x = 0; x < nx; x++) {
        if (out[x]) continue;
        double hd = 0.0;
        int64_t rem = x;
        for (int ci = 0; ci < N; ci++) {
          const int64_t i = rem % M[ci];
          rem /= M[ci];
          hd += hb[ci][i * (M[ci] + 1)];
        }
        for (int r = 0; r < R; r++) {
          const double s = sig[x + nx * r];
          je2[j * R + r] += s * s / (e0[r] - hd);
        }
      }
    if (rc) err = rc;
    free(out);
    free(w1);
    free(w2);
    free(sig);
  }
  for (int r = 0; r < R; r++) {
    double s = 0.0, comp = 0.0;
    for (int64_t j = 0; j < nj; j++) {
      const double v = je2[j * R + r], t = s + v;
      comp += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    e2[r] = s + comp;
  }
  if (n_jobs) *n_jobs = nj;
  free(je2);
  free(vm.slot);
  free(cand_job);
  free(jfock);
  free(slot);
  return err;
}

/* compute_pt2_energy over a sparse first-order space: the jobs, terms and
 * element test of orc_tpsci_pt2, with each target block's sigma held in a
 * map over the configurations reached (lin_map) instead of a dense array,
 * and ket configurations whose coefficients are all zero skipped (their
 * threshold thresh / 0 admits no element).  So it serves target Fock blocks
 * too large to form (many clusters) for states with few nonzero
 * coefficients.  Per job the configurations are summed in the order they
 * were first reached.  n_x (nullable): configurations reached outside the
 * variational space; n_contrib (nullable): elements kept.  Returns 0, -3 on a missing 1-body or H0 block, -4 when
 * a target block has more than 2^62 configurations. */
int orc_tpsci_pt2_sparse(const orc_ops *o, const orc_ham *h, const orc_space *sp, int R,
                         const double *c, const double *e0, int nbody, int h0_op, double thresh,
                         double *e2, int64_t *n_jobs, int64_t *n_x, int64_t *n_contrib,
                         int nthreads) {
  const int N = o->n_clusters, S = o->S;
  if (N > MAXN || R < 1 || R > 64) return -1;
  /* ket Fock blocks holding a nonzero coefficient */
  int64_t *act = (int64_t *)malloc(sizeof(int64_t) * (sp->n_fock > 0 ? sp->n_fock : 1));
  int64_t na = 0;
  for (int64_t f = 0; f < sp->n_fock; f++) {
    int nz = 0;
    for (int64_t kj = sp->fock_start[f]; kj < sp->fock_start[f + 1] && !nz; kj++)
      for (int r = 0; r < R; r++) nz |= c[(int64_t)r * sp->n_cfg + kj] != 0.0;
    if (nz) act[na++] = f;
  }
  /* jobs: target Fock blocks F_ket + TransferConfig (tpsci_pt2_energy.jl:49-79) */
  const int64_t n_cand = na * (int64_t)h->n_tc;
  int64_t cap = 16;
  while (cap < 2 * n_cand) cap <<= 1;
  int64_t *slot = (int64_t *)malloc(sizeof(int64_t) * cap);
  for (int64_t i = 0; i < cap; i++) slot[i] = -1;
  int16_t *jfock = (int16_t *)malloc(sizeof(int16_t) * 2 * N * (n_cand > 0 ? n_cand : 1));
  int64_t *cand_job = (int64_t *)malloc(sizeof(int64_t) * (n_cand > 0 ? n_cand : 1));
  int64_t nj = 0;
  for (int64_t ai = 0; ai < na; ai++) {
    const int16_t *fk = sp->fock + act[ai] * 2 * N;
    for (int g = 0; g < h->n_tc; g++) {
      int16_t fx[2 * MAXN];
      int ok = 1;
      for (int i = 0; i < 2 * N && ok; i++) {
        fx[i] = fk[i] + h->tc[(int64_t)g * 2 * N + i];
        ok = fx[i] >= 0 && fx[i] < S;
      }
      for (int ci = 0; ci < N && ok; ci++) {
        const int32_t *d;
        ok = gamma_block(o, ci, 0, fx, fx, &d) != NULL;
      }
      cand_job[ai * h->n_tc + g] = ok ? fock_intern(slot, cap, jfock, &nj, fx, N) : -1;
    }
  }
  /* the (ket block, TransferConfig) pairs of each job, in (block, group) order */
  int64_t *jstart = (int64_t *)calloc((size_t)nj + 1, sizeof(int64_t));
  for (int64_t i = 0; i < n_cand; i++)
    if (cand_job[i] >= 0) jstart[cand_job[i] + 1]++;
  for (int64_t j = 0; j < nj; j++) jstart[j + 1] += jstart[j];
  int64_t *fill = (int64_t *)malloc(sizeof(int64_t) * (nj > 0 ? nj : 1));
  for (int64_t j = 0; j < nj; j++) fill[j] = jstart[j];
  int64_t *pairs = (int64_t *)malloc(sizeof(int64_t) * (jstart[nj] > 0 ? jstart[nj] : 1));
  for (int64_t i = 0; i < n_cand; i++)
    if (cand_job[i] >= 0) pairs[fill[cand_job[i]]++] = i;
  fock_map vm;
  fock_map_build(&vm, sp->fock, sp->n_fock, N);
  double *je2 = (double *)calloc((size_t)(nj > 0 ? nj : 1) * R, sizeof(double));
  int err = 0;
  int64_t nxs = 0, ncs = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : nxs, ncs)
  for (int64_t j = 0; j < nj; j++) {
    const int16_t *fx = jfock + j * 2 * N;
    int64_t stride[MAXN], nx = 1;
    int M[MAXN];
    int rc = 0;
    for (int ci = 0; ci < N; ci++) {
      const int32_t *d = NULL;
      if (!gamma_block(o, ci, 0, fx, fx, &d)) continue; /* excluded when the job was formed */
      M[ci] = d[1];
      stride[ci] = nx;
      if (nx > ((int64_t)1 << 62) / M[ci]) rc = -4;
      nx *= M[ci];
    }
    lin_map hm;
    lin_map_init(&hm, R);
    double *w1 = NULL, *w2 = NULL;
    int64_t wcap = 0;
    double cket[64];
    /* _pt2_job: every (TransferConfig group, ket Fock block) feeding F_x */
    for (int64_t pi = jstart[j]; pi < jstart[j + 1] && !rc; pi++) {
      const int64_t f = act[pairs[pi] / h->n_tc];
      const int g = (int)(pairs[pi] % h->n_tc);
      const int16_t *fk = sp->fock + f * 2 * N;
      for (int t = h->tc_start[g]; t < h->tc_start[g + 1] && !rc; t++) {
        if (h->nbody[t] > nbody) continue;
        for (int64_t kj = sp->fock_start[f]; kj < sp->fock_start[f + 1] && !rc; kj++) {
          double cmax = 0.0;
          for (int r = 0; r < R; r++) {
            cket[r] = c[(int64_t)r * sp->n_cfg + kj];
            cmax = fabs(cket[r]) > cmax ? fabs(cket[r]) : cmax;
          }
          if (cmax == 0.0) continue;
          rc = pt2_term(o, h, t, fx, fk, sp->cfg + kj * N, cket, R, thresh, stride, nx, NULL,
                        NULL, &hm, &w1, &w2, &wcap);
        }
      }
    }
    /* project_out! (src/tpsci_outer.jl:869-889): drop variational configurations */
    unsigned char *drop = (unsigned char *)calloc((size_t)(hm.n > 0 ? hm.n : 1), 1);
    const int64_t vf = fock_map_find(&vm, fx);
    if (vf >= 0 && !rc)
      for (int64_t i = sp->fock_start[vf]; i < sp->fock_start[vf + 1]; i++) {
        int64_t lin = 0;
        for (int ci = 0; ci < N; ci++) lin += (int64_t)(sp->cfg[i * N + ci] - 1) * stride[ci];
        const int64_t sl = lin_map_slot(&hm, lin);
        if (hm.key[sl] == lin) drop[hm.idx[sl]] = 1;
      }
    /* compute_diagonal! with the H0 operator (src/tpsci_outer.jl:735-757) */
    const double *hb[MAXN];
    for (int ci = 0; ci < N && !rc; ci++) {
      const int32_t *d;
      hb[ci] = gamma_block(o, ci, h0_op, fx, fx, &d);
      if (!hb[ci]) rc = -3;
    }
    /* _sum_pt2 (src/tpsci_pt2_energy.jl:269-282) */
    if (!rc)
      for (int64_t e = 0; e < hm.n; e++) {
        if (drop[e]) continue;
        nxs++;
        double hd = 0.0;
        int64_t rem = hm.order[e];
        for (int ci = 0; ci < N; ci++) {
          const int64_t i = rem % M[ci];
          rem /= M[ci];
          hd += hb[ci][i * (M[ci] + 1)];
        }
        for (int r = 0; r < R; r++) {
          const double sv = hm.val[e * R + r];
          je2[j * R + r] += sv * sv / (e0[r] - hd);
        }
      }
    if (rc) err = rc;
    ncs += hm.adds;
    free(drop);
    lin_map_free(&hm);
    free(w1);
    free(w2);
  }
  for (int r = 0; r < R; r++) {
    double s = 0.0, comp = 0.0;
    for (int64_t j = 0; j < nj; j++) {
      const double v = je2[j * R + r], t = s + v;
      comp += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
      s = t;
    }
    e2[r] = s + comp;
  }
  if (n_jobs) *n_jobs = nj;
  if (n_x) *n_x = nxs;
  if (n_contrib) *n_contrib = ncs;
  free(je2);
  free(vm.slot);
  free(pairs);
  free(fill);
  free(jstart);
  free(cand_job);
  free(jfock);
  free(slot);
  free(act);
  return err;
}

/* Open matvec (open_matvec_thread, src/tpsci_matvec_thread.jl:11-129, with
 * prescreen=false) on caller-listed target Fock blocks: every (TransferConfig
 * group, ket Fock block) with F_ket + TransferConfig = F_x feeds block x
 * (the job list of :25-52); each term (nbody or fewer clusters) on each ket
 * configuration adds its elements above thresh / max_r |c_ket| (pt2_term).
 * Block j covers configurations off[j] .. off[j+1]-1 (first cluster fastest;
 * the sizes must match the "H" blocks, else -2).  sig is [R][off[n]]
 * (root-major), hit[x] = 1 for the configurations the reference inserts
 * into its output (some element passed the threshold). */
int orc_tpsci_open_sigma(const orc_ops *o, const orc_ham *h, const orc_space *sp, int R,
                         const double *c, int nbody, double thresh, int64_t n_target,
                         const int16_t *target_fock, const int64_t *off, double *sig,
                         unsigned char *hit, int nthreads) {
  const int N = o->n_clusters, S = o->S;
  if (N > MAXN || R < 1 || R > 64) return -1;
  const int64_t total = off[n_target];
  for (int64_t x = 0; x < total * R; x++) sig[x] = 0.0;
  for (int64_t x = 0; x < total; x++) hit[x] = 0;
  fock_map tm;
  fock_map_build(&tm, target_fock, n_target, N);
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t j = 0; j < n_target; j++) {
    const int16_t *fx = target_fock + j * 2 * N;
    int64_t stride[MAXN], nx = 1;
    int rc = 0;
    for (int ci = 0; ci < N; ci++) {
      const int32_t *d = NULL;
      if (!gamma_block(o, ci, 0, fx, fx, &d)) { rc = -2; break; }
      stride[ci] = nx;
      nx *= d[1];
    }
    if (!rc && nx != off[j + 1] - off[j]) rc = -2;
    double *js = rc ? NULL : (double *)calloc((size_t)nx * R, sizeof(double));
    double *w1 = NULL, *w2 = NULL;
    int64_t wcap = 0;
    double cket[64];
    for (int64_t f = 0; f < sp->n_fock && !rc; f++) {
      const int16_t *fk = sp->fock + f * 2 * N;
      for (int g = 0; g < h->n_tc && !rc; g++) {
        int16_t fy[2 * MAXN];
        int ok = 1;
        for (int i = 0; i < 2 * N && ok; i++) {
          fy[i] = fk[i] + h->tc[(int64_t)g * 2 * N + i];
          ok = fy[i] >= 0 && fy[i] < S;
        }
        if (!ok || fock_map_find(&tm, fy) != j) continue;
        for (int t = h->tc_start[g]; t < h->tc_start[g + 1] && !rc; t++) {
          if (h->nbody[t] > nbody) continue;
          for (int64_t kj = sp->fock_start[f]; kj < sp->fock_start[f + 1] && !rc; kj++) {
            for (int r = 0; r < R; r++) cket[r] = c[(int64_t)r * sp->n_cfg + kj];
            rc = pt2_term(o, h, t, fx, fk, sp->cfg + kj * N, cket, R, thresh, stride, nx, js,
                          hit + off[j], NULL, &w1, &w2, &wcap);
          }
        }
      }
    }
    if (!rc)
      for (int r = 0; r < R; r++)
        for (int64_t x = 0; x < nx; x++) sig[(int64_t)r * total + off[j] + x] = js[x + nx * r];
    if (rc) {
#pragma omp critical
      err = rc;
    }
    free(js);
    free(w1);
    free(w2);
  }
  free(tm.slot);
  return err;
}
