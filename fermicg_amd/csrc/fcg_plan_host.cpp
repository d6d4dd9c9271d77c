// Host side of plan construction: which (bra box, ket box, term) triples
// are connected, and the layout of the Hamiltonian blocks in HBM.
//
// The connection rule is the reference's check_term
// (src/type_ClusteredTerm.jl:92-176) lifted from single configurations to
// boxes: clusters a term does not touch must carry the same Fock sector and
// the same state range on both sides, and the active clusters' Fock sectors
// must differ by the term's delta, i.e. fock_ket = fock_bra - TransferConfig
// (src/Indexing.jl:21-23, the lookup in tps_ci_matvec,
// src/tpsci_outer.jl:479-481).  The fermionic sign is the reference's
// compute_terms_state_sign (src/type_ClusteredTerm.jl:181-197), taken from
// the ket Fock configuration.
#include <omp.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <numeric>
#include <string>
#include <unordered_map>

#include "fcg_internal.h"

namespace {

std::string fock_key(const int16_t *f, int N) {
  std::string k(2 * N, '\0');
  for (int i = 0; i < 2 * N; i++) k[i] = (char)(f[i] + 1);
  return k;
}

int term_sign(const fcg_ham *h, int t, const int16_t *fk) {
  int s = 0;
  for (int a = 0; a < h->nbody[t]; a++) {
    if (h->parity[4 * t + a] != 1) continue;
    int ne = 0;
    for (int c = 0; c < h->clusters[4 * t + a]; c++) ne += fk[2 * c] + fk[2 * c + 1];
    s ^= ne & 1;
  }
  return s ? -1 : 1;
}

struct LocalTask {
  int32_t B, K;
  TaskDev t;
};

}  // namespace

namespace {

void push_box(fcg_space *sp, int32_t f, const int16_t *lo, const int16_t *hi) {
  const int N = sp->N;
  sp->box_fock.push_back(f);
  int64_t d = 1;
  for (int c = 0; c < N; c++) {
    sp->lo.push_back(lo[c]);
    sp->hi.push_back(hi[c]);
    sp->rank.push_back((int16_t)(hi[c] - lo[c] + 1));
    sp->foff.push_back(-1);
    d *= hi[c] - lo[c] + 1;
  }
  sp->var_off.push_back(sp->var_off.back() + d);
}

// Every configuration of Fock block f (product of M[c] states, first cluster
// fastest) as boxes of at most FCG_MAX_BOX configurations, each a contiguous
// run of that order.
void push_full_block(fcg_space *sp, int32_t f, const std::vector<int> &M) {
  const int N = sp->N;
  int16_t lo[FCG_MAXN], hi[FCG_MAXN];
  int k = 0;
  int64_t pre = 1;
  while (k < N && pre * M[k] <= FCG_MAX_BOX) pre *= M[k++];
  if (k == N) {
    for (int c = 0; c < N; c++) { lo[c] = 1; hi[c] = (int16_t)M[c]; }
    push_box(sp, f, lo, hi);
    return;
  }
  const int chunk = (int)(FCG_MAX_BOX / pre);
  std::vector<int> slow(N, 1);
  while (true) {
    for (int a = 1; a <= M[k]; a += chunk) {
      for (int c = 0; c < N; c++) {
        if (c < k) { lo[c] = 1; hi[c] = (int16_t)M[c]; }
        else if (c == k) { lo[c] = (int16_t)a; hi[c] = (int16_t)std::min(M[k], a + chunk - 1); }
        else { lo[c] = hi[c] = (int16_t)slow[c]; }
      }
      push_box(sp, f, lo, hi);
    }
    int c = k + 1;
    for (; c < N; c++) {
      if (++slow[c] <= M[c]) break;
      slow[c] = 1;
    }
    if (c >= N) break;
  }
}

int sector_states(const fcg_ops *o, const int16_t *fk, int N, std::vector<int> *M,
                  std::string *err) {
  const int S = o->S;
  for (int c = 0; c < N; c++) {
    if (fk[2 * c] < 0 || fk[2 * c + 1] < 0 || fk[2 * c] >= S || fk[2 * c + 1] >= S) {
      *err = "Fock sector outside the cluster";
      return FCG_EINVAL;
    }
    (*M)[c] = o->nstates[(int64_t)c * S * S + fk[2 * c] * S + fk[2 * c + 1]];
    if ((*M)[c] <= 0) {
      *err = "Fock block uses a sector missing from ClusterOps";
      return FCG_EMISSING_BLOCK;
    }
  }
  return FCG_OK;
}

}  // namespace

int fcg_tpsci_boxes(const fcg_ops *o, int N, int64_t n_fock, const int16_t *fock,
                    const int64_t *fock_start, int64_t n_cfg, const int16_t *cfg,
                    fcg_space *sp, std::string *err) {
  sp->N = N;
  sp->tpsci = true;
  sp->n_fock = n_fock;
  sp->fock.assign(fock, fock + n_fock * 2 * N);
  if (fock_start[0] != 0 || fock_start[n_fock] != n_cfg) {
    *err = "fock_start must cover [0, n_cfg)";
    return FCG_EINVAL;
  }
  sp->var_off.assign(1, 0);
  std::vector<int> M(N);
  for (int64_t f = 0; f < n_fock; f++) {
    const int16_t *fk = fock + f * 2 * N;
    if (int rc = sector_states(o, fk, N, &M, err)) return rc;
    int64_t prod = 1;
    for (int c = 0; c < N; c++) prod *= M[c];
    const int64_t s0 = fock_start[f], s1 = fock_start[f + 1];
    for (int64_t i = s0; i < s1; i++)
      for (int c = 0; c < N; c++)
        if (cfg[i * N + c] < 1 || cfg[i * N + c] > M[c]) {
          *err = "state index out of range";
          return FCG_EINVAL;
        }
    bool full = (s1 - s0 == prod);
    if (full) {  // configurations must enumerate the product, first cluster fastest
      std::vector<int> idx(N, 1);
      for (int64_t i = s0; i < s1 && full; i++) {
        for (int c = 0; c < N; c++)
          if (cfg[i * N + c] != idx[c]) { full = false; break; }
        for (int c = 0; c < N; c++) {
          if (++idx[c] <= M[c]) break;
          idx[c] = 1;
        }
      }
    }
    if (full) {
      push_full_block(sp, (int32_t)f, M);
    } else {
      int16_t lo[FCG_MAXN];
      for (int64_t i = s0; i < s1; i++) {
        for (int c = 0; c < N; c++) lo[c] = cfg[i * N + c];
        push_box(sp, (int32_t)f, lo, lo);
      }
    }
    if (sp->var_off.back() != s1) {
      *err = "internal: box partition does not match the configuration order";
      return FCG_EINVAL;
    }
  }
  sp->n_box = (int64_t)sp->box_fock.size();
  sp->nvar = sp->var_off.back();
  return FCG_OK;
}

int fcg_full_space(const fcg_ops *o, int N, int64_t n_fock, const int16_t *fock, fcg_space *sp,
                   std::string *err) {
  sp->N = N;
  sp->tpsci = true;
  sp->n_fock = n_fock;
  sp->fock.assign(fock, fock + n_fock * 2 * N);
  sp->var_off.assign(1, 0);
  std::vector<int> M(N);
  for (int64_t f = 0; f < n_fock; f++) {
    if (int rc = sector_states(o, fock + f * 2 * N, N, &M, err)) return rc;
    push_full_block(sp, (int32_t)f, M);
  }
  sp->n_box = (int64_t)sp->box_fock.size();
  sp->nvar = sp->var_off.back();
  return FCG_OK;
}

// Validate one space's boxes against the ClusterOps state counts.
static int check_space(const fcg_ops *o, const fcg_space *sp, std::string *err) {
  const int N = sp->N, S = o->S;
  for (int64_t b = 0; b < sp->n_box; b++) {
    if (sp->var_off[b + 1] - sp->var_off[b] > FCG_MAX_BOX) {
      *err = "block dimension exceeds " + std::to_string(FCG_MAX_BOX);
      return FCG_EOVERFLOW;
    }
    const int16_t *F = &sp->fock[(int64_t)sp->box_fock[b] * 2 * N];
    for (int c = 0; c < N; c++) {
      if (F[2 * c] < 0 || F[2 * c + 1] < 0 || F[2 * c] >= S || F[2 * c + 1] >= S) {
        *err = "Fock sector outside the cluster";
        return FCG_EINVAL;
      }
      const int M = o->nstates[(int64_t)c * S * S + F[2 * c] * S + F[2 * c + 1]];
      if (sp->hi[b * N + c] > M || sp->lo[b * N + c] < 1) {
        *err = "box " + std::to_string(b) + " range outside the sector's " + std::to_string(M) +
               " states on cluster " + std::to_string(c);
        return FCG_EINVAL;
      }
    }
  }
  if (sp->nvar > INT32_MAX) {
    *err = "more than 2^31 variables";
    return FCG_EOVERFLOW;
  }
  return FCG_OK;
}

int fcg_build_host_plan(const fcg_ham *h, const fcg_ops *o, const fcg_space *bra,
                        const fcg_space *ket, int nbody, bool layout, int class_cap,
                        HostPlan *out, std::string *err) {
  const int N = bra->N, S = o->S;
  if (h->N != N || o->N != N || ket->N != N) {
    *err = "cluster counts of ham/ops/spaces differ";
    return FCG_EINVAL;
  }
  if (int rc = check_space(o, bra, err)) return rc;
  if (ket != bra)
    if (int rc = check_space(o, ket, err)) return rc;
  // ket Fock configurations and their boxes
  std::unordered_map<std::string, int64_t> fmap;
  fmap.reserve(ket->n_fock * 2);
  for (int64_t f = 0; f < ket->n_fock; f++) fmap.emplace(fock_key(&ket->fock[f * 2 * N], N), f);
  std::vector<std::vector<int32_t>> kboxes(ket->n_fock), bboxes(bra->n_fock);
  for (int64_t k = 0; k < ket->n_box; k++) kboxes[ket->box_fock[k]].push_back((int32_t)k);
  for (int64_t b = 0; b < bra->n_box; b++) bboxes[bra->box_fock[b]].push_back((int32_t)b);
  auto dim_b = [&](int64_t b) { return bra->var_off[b + 1] - bra->var_off[b]; };
  auto dim_k = [&](int64_t k) { return ket->var_off[k + 1] - ket->var_off[k]; };
  // terms of each TransferConfig grouped by active-cluster mask
  struct Group { uint32_t mask; std::vector<int> terms; };
  std::vector<std::vector<Group>> groups(h->n_tc);
  for (int g = 0; g < h->n_tc; g++) {
    for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) {
      if (h->nbody[t] > nbody) continue;
      uint32_t m = 0;
      for (int a = 0; a < h->nbody[t]; a++) m |= 1u << h->clusters[4 * t + a];
      auto it = std::find_if(groups[g].begin(), groups[g].end(),
                             [&](const Group &x) { return x.mask == m; });
      if (it == groups[g].end()) groups[g].push_back({m, {t}});
      else it->terms.push_back(t);
    }
  }
  std::vector<RawVec<LocalTask>> per_fock(bra->n_fock);
  int errcode = FCG_OK;
  std::string errmsg;
#pragma omp parallel for schedule(dynamic, 4)
  for (int64_t fb = 0; fb < bra->n_fock; fb++) {
    if (errcode || bboxes[fb].empty()) continue;
    const int16_t *F = &bra->fock[fb * 2 * N];
    RawVec<LocalTask> loc;
    int16_t fk[2 * FCG_MAXN];
    for (int g = 0; g < h->n_tc && !errcode; g++) {
      const int16_t *tc = &h->tc[(int64_t)g * 2 * N];
      bool ok = true;
      for (int i = 0; i < 2 * N; i++) {
        fk[i] = F[i] - tc[i];
        if (fk[i] < 0 || fk[i] >= S) ok = false;
      }
      if (!ok) continue;
      auto it = fmap.find(fock_key(fk, N));
      if (it == fmap.end()) continue;
      const int64_t fki = it->second;
      for (const Group &G : groups[g]) {
        // Γ blocks and sign per term (depend only on the Fock pair)
        std::vector<TaskDev> tt;
        for (int t : G.terms) {
          TaskDev d{};
          d.term = t;
          d.meta = term_sign(h, t, fk) < 0 ? 1 : 0;
          for (int a = 0; a < 4; a++) d.blk[a] = -1;
          for (int a = 0; a < h->nbody[t]; a++) {
            const int c = h->clusters[4 * t + a];
            const int blk = o->block(c, h->ops[4 * t + a], F[2 * c] * S + F[2 * c + 1],
                                     fk[2 * c] * S + fk[2 * c + 1]);
            if (blk < 0) {
#pragma omp critical
              {
                errcode = FCG_EMISSING_BLOCK;
                errmsg = "missing Γ block for term " + std::to_string(t) + " cluster " +
                         std::to_string(c);
              }
            } else {
              const int32_t *gd = &o->gb_dims[3 * (int64_t)blk];
              const int np_t = h->nbody[t] == 1 ? 1 : h->dims[4 * t + a];
              const int Mb = o->nstates[(int64_t)c * S * S + F[2 * c] * S + F[2 * c + 1]];
              const int Mk = o->nstates[(int64_t)c * S * S + fk[2 * c] * S + fk[2 * c + 1]];
              if (gd[0] != np_t || gd[1] != Mb || gd[2] != Mk) {
#pragma omp critical
                {
                  errcode = FCG_EINVAL;
                  errmsg = "Γ block shape disagrees with term " + std::to_string(t) +
                           " / sector sizes on cluster " + std::to_string(c);
                }
              }
            }
            d.blk[a] = blk;
          }
          tt.push_back(d);
        }
        for (int32_t B : bboxes[fb])
          for (int32_t K : kboxes[fki]) {
            // inactive clusters: equal TuckerConfig range (reference
            // check_term); with identity factors on both sides (TPSCI boxes)
            // overlapping ranges suffice -- the states must then be equal,
            // which the block builder enforces element by element
            bool eq = true;
            for (int c = 0; c < N && eq; c++) {
              if (G.mask >> c & 1u) continue;
              const int64_t ib = (int64_t)B * N + c, ik = (int64_t)K * N + c;
              if (bra->foff[ib] < 0 && ket->foff[ik] < 0)
                eq = std::max(bra->lo[ib], ket->lo[ik]) <= std::min(bra->hi[ib], ket->hi[ik]);
              else
                eq = bra->lo[ib] == ket->lo[ik] && bra->hi[ib] == ket->hi[ik];
            }
            if (!eq) continue;
            for (const TaskDev &d : tt) loc.push_back({B, K, d});
          }
      }
    }
    std::stable_sort(loc.begin(), loc.end(), [](const LocalTask &a, const LocalTask &b) {
      return a.B != b.B ? a.B < b.B : a.K < b.K;
    });
    per_fock[fb].swap(loc);
  }
  if (errcode) {
    *err = errmsg;
    return errcode;
  }
  // assemble pairs ordered by bra box then ket box, and their tasks
  // contiguous in that order: each bra Fock block's runs of equal (B, K)
  // are counted, offset by prefix sums and written in parallel
  struct PairRec { int32_t B, K; int64_t t0, t1; };
  const int64_t nfb = bra->n_fock;
  std::vector<int64_t> pair_off(nfb + 1, 0), task_off(nfb + 1, 0);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t fb = 0; fb < nfb; fb++) {
    const auto &v = per_fock[fb];
    int64_t n = 0;
    for (size_t i = 0; i < v.size(); i++)
      if (i == 0 || v[i].B != v[i - 1].B || v[i].K != v[i - 1].K) n++;
    pair_off[fb + 1] = n;
    task_off[fb + 1] = (int64_t)v.size();
  }
  for (int64_t fb = 0; fb < nfb; fb++) {
    pair_off[fb + 1] += pair_off[fb];
    task_off[fb + 1] += task_off[fb];
  }
  RawVec<PairRec> pairs(pair_off[nfb]);
  out->tasks.resize(task_off[nfb]);
#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t fb = 0; fb < nfb; fb++) {
    auto &v = per_fock[fb];
    int64_t pp = pair_off[fb], tt = task_off[fb];
    for (size_t i = 0; i < v.size(); i++) {
      if (i == 0 || v[i].B != v[i - 1].B || v[i].K != v[i - 1].K) {
        pairs[pp++] = {v[i].B, v[i].K, tt, tt};
      }
      out->tasks[tt++] = v[i].t;
      pairs[pp - 1].t1 = tt;
    }
    RawVec<LocalTask>().swap(v);
  }
  auto pair_less = [](const PairRec &a, const PairRec &b) {
    return a.B != b.B ? a.B < b.B : a.K < b.K;
  };
  if (!std::is_sorted(pairs.begin(), pairs.end(), pair_less)) {
    // bra boxes not grouped by Fock block: sort the pairs, then regroup
    // their tasks in the new order
    std::stable_sort(pairs.begin(), pairs.end(), pair_less);
    RawVec<TaskDev> moved(out->tasks.size());
    int64_t at = 0;
    for (PairRec &r : pairs) {
      const int64_t n = r.t1 - r.t0;
      std::copy(out->tasks.begin() + r.t0, out->tasks.begin() + r.t1, moved.begin() + at);
      r.t0 = at;
      r.t1 = at + n;
      at += n;
    }
    out->tasks.swap(moved);
  }
  // pair tables (tasks are contiguous in pair order, so a pair's first task
  // is its t0); row b of the bra holds pairs [row_ptr[b], row_ptr[b + 1])
  const int64_t np = (int64_t)pairs.size(), nbb = bra->n_box;
  out->pair_B.resize(np);
  out->pair_K.resize(np);
  out->pair_t0.resize(np + 1);
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; p++) {
    out->pair_B[p] = pairs[p].B;
    out->pair_K[p] = pairs[p].K;
    out->pair_t0[p] = pairs[p].t0;
  }
  out->pair_t0[np] = np ? pairs[np - 1].t1 : 0;
  out->row_ptr.assign(nbb + 1, 0);
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b <= nbb; b++)
    out->row_ptr[b] =
        std::lower_bound(out->pair_B.begin(), out->pair_B.end(), (int32_t)b) - out->pair_B.begin();
  // parallel pass: each task's chain order, LDS limits, flops
  double flops = 0.0;
  int64_t elems = 0;
  std::atomic<int> lim_err{FCG_OK};
  const char *lim_msg = nullptr;
  RawVec<uint8_t> lds_class(np);  // 3 tiny, 2 small, 1 middle, 0 full
#pragma omp parallel reduction(+ : flops, elems)
  {
    // chain orders by task shape (nbody, Γ index counts, rank products): a
    // direct-mapped cache per thread, since the same shapes recur across
    // pairs and the search walks up to 24 permutations
    struct ChainOrder {
      uint64_t k0 = ~0ull, k1 = ~0ull;
      int64_t arena = 0;
      double fl = 0.0;
      int best[4] = {0, 1, 2, 3};
    };
    std::vector<ChainOrder> chain_cache(4096);
    int64_t m_arena = 0;
    int m_best[4] = {0, 1, 2, 3};
    double m_fl = 0.0;
#pragma omp for schedule(dynamic, 1024)
    for (int64_t p = 0; p < np; p++) {
      if (lim_err) continue;
      const PairRec &r = pairs[p];
      auto fail = [&](const char *m) {
#pragma omp critical
        {
          if (!lim_err) {
            lim_err = FCG_EOVERFLOW;
            lim_msg = m;
          }
        }
      };
      const int64_t db = dim_b(r.B), dk = dim_k(r.K);
      const int64_t B = r.B, K = r.K;
      elems += db * dk;
      int64_t ssum = 0;
      for (int c = 0; c < N; c++)
        if (bra->lo[B * N + c] == ket->lo[K * N + c] && bra->hi[B * N + c] == ket->hi[K * N + c] &&
            !(bra->foff[B * N + c] < 0 && ket->foff[K * N + c] < 0))
          ssum += (int64_t)bra->rank[B * N + c] * ket->rank[K * N + c];
      if (ssum > FCG_LDS_S) {
        fail("inactive-cluster overlaps exceed LDS budget");
        continue;
      }
      // largest LDS use of the pair's tasks: staged Γ slices (X), the chain
      // arena (A: a stage's input and output side by side, and the staged
      // integrals), factor-compression rows (T), overlaps (S)
      int64_t need_x = 0, need_a = 0, need_t = 0;
      for (int64_t t = r.t0; t < r.t1; t++) {
        TaskDev &d = out->tasks[t];
        const int tm = d.term, nb = h->nbody[tm];
        int64_t np_[4] = {0, 0, 0, 0}, rr_[4] = {0, 0, 0, 0};
        int64_t rest0 = 1;
        const char *lds_err = nullptr;
        for (int a = 0; a < nb; a++) {
          const int c = h->clusters[4 * tm + a];
          np_[a] = (nb == 1) ? 1 : h->dims[4 * tm + a];
          rr_[a] = (int64_t)bra->rank[B * N + c] * ket->rank[K * N + c];
          rest0 *= np_[a];
          const int64_t nI = bra->hi[B * N + c] - bra->lo[B * N + c] + 1;
          const int64_t nJ = ket->hi[K * N + c] - ket->lo[K * N + c] + 1;
          const int64_t rk = ket->rank[K * N + c];
          if (np_[a] * rr_[a] > FCG_LDS_X || nI * rk > FCG_LDS_T) lds_err = "Γ slice exceeds LDS budget";
          need_x = std::max(need_x, np_[a] * rr_[a]);
          need_t = std::max(need_t, nI * rk);
          const bool ident = bra->foff[B * N + c] < 0 && ket->foff[K * N + c] < 0;
          if (!ident) flops += 2.0 * np_[a] * (double)(nI * nJ * rk + nI * bra->rank[B * N + c] * rk);
        }
        if (lds_err) {
          fail(lds_err);
          break;
        }
        if (rest0 > 2 * FCG_LDS_X) {
          fail("term integrals exceed LDS budget");
          break;
        }
        // chain order: the permutation whose stages need the smallest arena
        // (input and output of a stage side by side; ties: fewest flops)
        uint64_t k0 = (uint64_t)nb, k1 = 0;
        bool keyed = true;
        for (int a = 0; a < 4; a++) {
          keyed = keyed && np_[a] < 256 && rr_[a] < 65536;
          k0 |= (uint64_t)(np_[a] & 255) << (8 + 8 * a);
          k1 |= (uint64_t)(rr_[a] & 65535) << (16 * a);
        }
        ChainOrder &co =
            chain_cache[((k0 * 0x9E3779B97F4A7C15ull) ^ (k1 * 0xC2B2AE3D27D4EB4Full)) >> 52];
        if (keyed && co.k0 == k0 && co.k1 == k1) {
          m_arena = co.arena;
          m_fl = co.fl;
          std::copy(co.best, co.best + 4, m_best);
        } else {
          int perm[4] = {0, 1, 2, 3};
          int64_t best_arena = INT64_MAX;
          double best_fl = 0.0;
          do {
            int64_t rest = rest0, arena = rest0;
            double fl = 0.0;
            for (int s2 = 0; s2 < nb; s2++) {
              const int a = perm[s2];
              const int64_t next = rest / np_[a] * rr_[a];
              fl += 2.0 * np_[a] * (double)next;
              arena = std::max(arena, rest + next);
              rest = next;
            }
            if (arena < best_arena || (arena == best_arena && fl < best_fl)) {
              best_arena = arena;
              best_fl = fl;
              std::copy(perm, perm + 4, m_best);
            }
          } while (std::next_permutation(perm, perm + nb));
          m_arena = best_arena;
          m_fl = best_fl;
          if (keyed) {
            co.k0 = k0;
            co.k1 = k1;
            co.arena = m_arena;
            co.fl = m_fl;
            std::copy(m_best, m_best + 4, co.best);
          }
        }
        if (m_arena > 2 * FCG_LDS_X) {
          fail("chain intermediate exceeds LDS budget");
          break;
        }
        need_a = std::max(need_a, m_arena);
        for (int s2 = 0; s2 < 4; s2++) d.meta |= (s2 < nb ? m_best[s2] : 0) << (8 + 4 * s2);
        flops += m_fl + 2.0 * db * dk;
      }
      // the tiny class's LDS cannot stage a fused sigma's c and y
      // need_x: staged Γ slice (LG), need_a: chain arena (2 LX)
      auto fits = [&](int lx, int lg, int lt, int ls) {
        return need_x <= lg && need_a <= 2 * lx && need_t <= lt && ssum <= ls;
      };
      const int cls =
          layout && db * dk <= FCG_TINY_ELEMS &&
                  fits(FCG_LDS_X_TINY, FCG_LDS_X_TINY, FCG_LDS_T_TINY, FCG_LDS_S_TINY)
              ? 3
          : fits(FCG_LDS_X_SMALL, FCG_LDS_X_SMALL, FCG_LDS_T_SMALL, FCG_LDS_S_SMALL)
              ? 2
              : fits(FCG_LDS_X_MID, FCG_LDS_G_MID, FCG_LDS_T_MID, FCG_LDS_S_MID);
      lds_class[p] = (uint8_t)std::min(cls, class_cap);
    }
  }
  if (lim_err.load()) {
    *err = lim_msg;
    return lim_err.load();
  }
  // LDS classes, each in pair order: 3 tiny, 2 small, 1 middle, 0 full (a
  // stable partition over per-thread chunks)
  {
    const int nth = std::max(1, omp_get_max_threads());
    const int64_t chunk = (np + nth - 1) / nth;
    std::vector<int64_t> cnt(4 * (size_t)nth, 0);
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nth; t++)
      for (int64_t p = t * chunk; p < std::min(np, (t + 1) * chunk); p++) cnt[lds_class[p] * nth + t]++;
    int64_t at = 0;
    for (int cls = 3; cls >= 0; cls--) {
      const int64_t start = at;
      for (int t = 0; t < nth; t++) {
        const int64_t n = cnt[cls * nth + t];
        cnt[cls * nth + t] = at;
        at += n;
      }
      if (cls == 3) out->n_tiny = at - start;
      if (cls == 2) out->n_small = at - start;
      if (cls == 1) out->n_mid = at - start;
    }
    out->pair_list.resize(np);
#pragma omp parallel for schedule(static, 1)
    for (int t = 0; t < nth; t++) {
      int64_t pos[4] = {cnt[t], cnt[nth + t], cnt[2 * nth + t], cnt[3 * nth + t]};
      for (int64_t p = t * chunk; p < std::min(np, (t + 1) * chunk); p++)
        out->pair_list[pos[lds_class[p]]++] = p;
    }
  }
  out->flops = flops;
  if (!layout) {
    out->n_kel = elems;
    return FCG_OK;
  }
  // row panels: the blocks of one bra row side by side form a db x ncol
  // panel (ncol padded to a multiple of FCG_QUAD with zero columns), stored
  // in column quads (panel_index); colidx maps its columns to ket variables
  out->pair_col.resize(np);
  out->pair_koff.resize(np);
  out->diag_col.assign(nbb, 0);
  out->diag_koff.assign(nbb, -1);
  std::vector<int64_t> row_cols(nbb, 0);
  std::vector<uint8_t> has_diag(nbb, 0);
  int wide = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(| : wide)
  for (int64_t b = 0; b < nbb; b++) {
    int64_t w = 0;
    for (int64_t p = out->row_ptr[b]; p < out->row_ptr[b + 1]; p++) {
      const int64_t dk = dim_k(out->pair_K[p]);
      if (w + dk > INT32_MAX / FCG_MAX_BOX) {
        wide = 1;
        break;
      }
      out->pair_col[p] = (int32_t)w;
      if (ket == bra && out->pair_K[p] == b) {
        out->diag_col[b] = (int32_t)w;
        has_diag[b] = 1;
      }
      w += dk;
    }
    row_cols[b] = w;
  }
  if (wide) {
    *err = "row panel too wide";
    return FCG_EOVERFLOW;
  }
  out->row_koff.assign(nbb, 0);
  out->col_ptr.assign(nbb + 1, 0);
  int64_t ncol = 0, koff = 0;
  for (int64_t b = 0; b < nbb; b++) {
    const int64_t w = (row_cols[b] + FCG_QUAD - 1) / FCG_QUAD * FCG_QUAD;
    out->col_ptr[b] = ncol;
    out->row_koff[b] = koff;
    if (has_diag[b]) out->diag_koff[b] = koff;
    ncol += w;
    koff += dim_b(b) * w;
  }
  out->col_ptr[nbb] = ncol;
#pragma omp parallel for schedule(static)
  for (int64_t p = 0; p < np; p++) out->pair_koff[p] = out->row_koff[out->pair_B[p]];
  out->colidx.resize(ncol);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t b = 0; b < nbb; b++) {
    int64_t at = out->col_ptr[b];
    for (int64_t q = out->row_ptr[b]; q < out->row_ptr[b + 1]; q++) {
      const int32_t K = out->pair_K[q];
      for (int64_t j = 0; j < dim_k(K); j++) out->colidx[at++] = (int32_t)(ket->var_off[K] + j);
    }
    for (; at < out->col_ptr[b + 1]; at++) out->colidx[at] = 0;  // padding columns
  }
  fcg_apply_order(bra->var_off, out->col_ptr, &out->order, &out->n_multi, &out->n_single_slots);
  out->n_kel = koff;
  return FCG_OK;
}

// Target Fock blocks follow compute_pt2_energy's job list
// (src/tpsci_pt2_energy.jl:49-79): F_ket + TransferConfig with every
// cluster sector in range and present in ClusterOps.
int fcg_pt2_target(const fcg_ham *h, const fcg_ops *o, const fcg_space *var, int h0_op,
                   fcg_space *tsp, Pt2Target *out, std::string *err) {
  const int N = var->N, S = o->S;
  auto key_of = [N](const int16_t *f) {
    std::string k(2 * N, '\0');
    for (int i = 0; i < 2 * N; i++) k[i] = (char)(f[i] + 1);
    return k;
  };
  std::unordered_map<std::string, int64_t> tmap;
  std::vector<int16_t> tfock, fx(2 * N);
  for (int64_t f = 0; f < var->n_fock; f++) {
    const int16_t *F = &var->fock[f * 2 * N];
    for (int g = 0; g < h->n_tc; g++) {
      const int16_t *tc = &h->tc[(int64_t)g * 2 * N];
      bool ok = true;
      for (int c = 0; c < N && ok; c++) {
        const int a = F[2 * c] + tc[2 * c], b = F[2 * c + 1] + tc[2 * c + 1];
        ok = a >= 0 && b >= 0 && a < S && b < S && o->nstates[(int64_t)c * S * S + a * S + b] > 0;
        fx[2 * c] = (int16_t)a;
        fx[2 * c + 1] = (int16_t)b;
      }
      if (!ok) continue;
      if (tmap.emplace(key_of(fx.data()), (int64_t)tfock.size() / (2 * N)).second)
        tfock.insert(tfock.end(), fx.begin(), fx.end());
    }
  }
  const int64_t n_tf = (int64_t)tfock.size() / (2 * N);
  if (int rc = fcg_full_space(o, N, n_tf, tfock.data(), tsp, err)) return rc;
  // first variable of each target Fock block
  std::vector<int64_t> fvar0(n_tf, -1);
  for (int64_t b = 0; b < tsp->n_box; b++)
    if (fvar0[tsp->box_fock[b]] < 0) fvar0[tsp->box_fock[b]] = tsp->var_off[b];
  // variational configurations (project_out!, src/tpsci_outer.jl:869-889);
  // full blocks are enumerated first cluster fastest
  out->inspace.assign(tsp->nvar, 0);
  std::vector<int> M(N);
  for (int64_t b = 0; b < var->n_box; b++) {
    const int16_t *F = &var->fock[(int64_t)var->box_fock[b] * 2 * N];
    auto it = tmap.find(key_of(F));
    if (it == tmap.end()) continue;  // no TransferConfig maps this block onto itself
    for (int c = 0; c < N; c++) M[c] = o->nstates[(int64_t)c * S * S + F[2 * c] * S + F[2 * c + 1]];
    const int64_t d = var->var_off[b + 1] - var->var_off[b];
    for (int64_t e = 0; e < d; e++) {
      int64_t r = e, lin = 0, stride = 1;
      for (int c = 0; c < N; c++) {
        const int rk = var->rank[b * N + c];
        const int ic = (int)(r % rk) + var->lo[b * N + c] - 1;
        r /= rk;
        lin += ic * stride;
        stride *= M[c];
      }
      const int64_t x = fvar0[it->second] + lin;
      if (x < 0 || x >= tsp->nvar) {
        *err = "internal: variational configuration outside its target block";
        return FCG_EINVAL;
      }
      out->inspace[x] = 1;
    }
  }
  // diagonal of H0 = sum_c op h0_op (compute_diagonal!, src/tpsci_outer.jl:735-757)
  out->hoff.assign(tsp->n_box * N, 0);
  out->hstride.assign(tsp->n_box * N, 0);
  for (int64_t b = 0; b < tsp->n_box; b++) {
    const int16_t *F = &tsp->fock[(int64_t)tsp->box_fock[b] * 2 * N];
    for (int c = 0; c < N; c++) {
      const int code = F[2 * c] * S + F[2 * c + 1];
      const int blk = o->block(c, h0_op, code, code);
      if (blk < 0) {
        *err = "H0 block missing on cluster " + std::to_string(c);
        return FCG_EMISSING_BLOCK;
      }
      const int m = o->gb_dims[3 * blk + 1];
      if (o->gb_dims[3 * blk] != 1 || o->gb_dims[3 * blk + 2] != m || tsp->hi[b * N + c] > m) {
        *err = "H0 block of cluster " + std::to_string(c) + " is not square over the sector";
        return FCG_EINVAL;
      }
      out->hoff[b * N + c] = o->gb_off[blk] + (int64_t)(tsp->lo[b * N + c] - 1) * (m + 1);
      out->hstride[b * N + c] = m + 1;
    }
  }
  return FCG_OK;
}

// Connected Fock pairs for a space of single-configuration rows: for each bra
// Fock block, the ket Fock blocks F_bra - TransferConfig present in `ket`,
// with the TransferConfig group, in group order (the lookup of
// tps_ci_matvec, src/tpsci_outer.jl:479-481).  Groups without a term of at
// most nbody clusters are skipped.  Fock configurations are packed into two
// 64-bit words (4 bits per electron count) and looked up in an
// open-addressing table.
namespace {
struct FockKey {
  uint64_t w[2];
  bool operator==(const FockKey &o) const { return w[0] == o.w[0] && w[1] == o.w[1]; }
};
inline uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  return x ^ (x >> 33);
}
}  // namespace

int fcg_fock_pairs(const fcg_ham *h, const fcg_space *bra, const fcg_space *ket, int nbody,
                   int S, std::vector<int64_t> *ptr, std::vector<int32_t> *fk_out,
                   std::vector<int32_t> *g_out) {
  const int N = bra->N;
  if (S > 16 || N > 16) {  // 4-bit counts, 32 counts in two words
    // general case: string keys
    std::unordered_map<std::string, int32_t> fmap;
    for (int64_t f = 0; f < ket->n_fock; f++) fmap.emplace(fock_key(&ket->fock[f * 2 * N], N), (int32_t)f);
    ptr->assign(bra->n_fock + 1, 0);
    fk_out->clear();
    g_out->clear();
    int16_t fk[2 * FCG_MAXN];
    for (int64_t fb = 0; fb < bra->n_fock; fb++) {
      const int16_t *F = &bra->fock[fb * 2 * N];
      for (int g = 0; g < h->n_tc; g++) {
        bool ok = false;
        for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++) ok |= h->nbody[t] <= nbody;
        for (int i = 0; i < 2 * N && ok; i++) {
          fk[i] = F[i] - h->tc[(int64_t)g * 2 * N + i];
          ok = fk[i] >= 0 && fk[i] < S;
        }
        if (!ok) continue;
        auto it = fmap.find(fock_key(fk, N));
        if (it == fmap.end()) continue;
        fk_out->push_back(it->second);
        g_out->push_back(g);
      }
      (*ptr)[fb + 1] = (int64_t)fk_out->size();
    }
    return FCG_OK;
  }
  auto pack = [N](const int16_t *f) {
    FockKey k{{0, 0}};
    for (int i = 0; i < 2 * N; i++) k.w[i >> 4] |= (uint64_t)(f[i] & 0xF) << (4 * (i & 15));
    return k;
  };
  // ket table: key and Fock index side by side (one cache line per probe)
  struct Slot {
    FockKey k;
    int64_t f;
  };
  int64_t cap = 16;
  while (cap < 2 * ket->n_fock) cap <<= 1;
  std::vector<Slot> table(cap, Slot{{{0, 0}}, -1});
  for (int64_t f = 0; f < ket->n_fock; f++) {
    const FockKey k = pack(&ket->fock[f * 2 * N]);
    uint64_t hsh = mix64(k.w[0] ^ mix64(k.w[1])) & (cap - 1);
    while (table[hsh].f >= 0) hsh = (hsh + 1) & (cap - 1);
    table[hsh] = Slot{k, f};
  }
  std::vector<int> live;
  for (int g = 0; g < h->n_tc; g++)
    for (int t = h->tc_start[g]; t < h->tc_start[g + 1]; t++)
      if (h->nbody[t] <= nbody) {
        live.push_back(g);
        break;
      }
  std::vector<std::vector<int32_t>> per(bra->n_fock);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t fb = 0; fb < bra->n_fock; fb++) {
    const int16_t *F = &bra->fock[fb * 2 * N];
    int16_t fk[2 * FCG_MAXN];
    std::vector<int32_t> &v = per[fb];
    for (int g : live) {
      const int16_t *tc = &h->tc[(int64_t)g * 2 * N];
      bool ok = true;
      for (int i = 0; i < 2 * N && ok; i++) {
        fk[i] = F[i] - tc[i];
        ok = fk[i] >= 0 && fk[i] < S;
      }
      if (!ok) continue;
      const FockKey k = pack(fk);
      uint64_t hsh = mix64(k.w[0] ^ mix64(k.w[1])) & (cap - 1);
      while (table[hsh].f >= 0) {
        if (table[hsh].k == k) {
          v.push_back((int32_t)table[hsh].f);
          v.push_back(g);
          break;
        }
        hsh = (hsh + 1) & (cap - 1);
      }
    }
  }
  ptr->assign(bra->n_fock + 1, 0);
  for (int64_t fb = 0; fb < bra->n_fock; fb++) (*ptr)[fb + 1] = (*ptr)[fb] + (int64_t)per[fb].size() / 2;
  fk_out->resize((*ptr)[bra->n_fock]);
  g_out->resize((*ptr)[bra->n_fock]);
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t fb = 0; fb < bra->n_fock; fb++) {
    int64_t at = (*ptr)[fb];
    for (size_t i = 0; i < per[fb].size(); i += 2, at++) {
      (*fk_out)[at] = per[fb][i];
      (*g_out)[at] = per[fb][i + 1];
    }
  }
  return FCG_OK;
}

// Boxes of several rows: largest panel first (ties by box index), so the long
// rows start early.  One-row boxes: in space order, which keeps rows of one
// Fock block (sharing their ket Fock blocks) together; workgroup g of
// FCG_CSR_WPB rows runs on XCD g mod 8, so slot g is given the row group
// (g mod 8) * G + g / 8 and each XCD's L2 sees one contiguous run of rows.
void fcg_apply_order(const std::vector<int64_t> &var_off, const std::vector<int64_t> &col_ptr,
                     std::vector<int32_t> *order, int64_t *n_multi, int64_t *n_single_slots) {
  const int64_t n_box = (int64_t)var_off.size() - 1;
  std::vector<int32_t> multi, single;
  for (int64_t b = 0; b < n_box; b++)
    (var_off[b + 1] - var_off[b] > 1 ? multi : single).push_back((int32_t)b);
  std::stable_sort(multi.begin(), multi.end(), [&](int32_t a, int32_t b) {
    return (var_off[a + 1] - var_off[a]) * (col_ptr[a + 1] - col_ptr[a]) >
           (var_off[b + 1] - var_off[b]) * (col_ptr[b + 1] - col_ptr[b]);
  });
  const int64_t ns = (int64_t)single.size();
  const int64_t nw = (ns + FCG_CSR_WPB - 1) / FCG_CSR_WPB;
  const int64_t g8 = (nw + FCG_XCDS - 1) / FCG_XCDS;
  const int64_t slots = ns ? g8 * FCG_XCDS * FCG_CSR_WPB : 0;
  *order = multi;
  order->resize(multi.size() + slots, -1);
  for (int64_t w = 0; w < slots; w++) {
    const int64_t g = w / FCG_CSR_WPB, k = w % FCG_CSR_WPB;
    const int64_t idx = ((g % FCG_XCDS) * g8 + g / FCG_XCDS) * FCG_CSR_WPB + k;
    (*order)[multi.size() + w] = idx < ns ? single[idx] : -1;
  }
  *n_multi = (int64_t)multi.size();
  *n_single_slots = slots;
}
