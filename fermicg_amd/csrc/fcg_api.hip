// C ABI of libfcg_hip.so: handles, uploads, plan creation, sigma, diagonal.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>

#include "fcg_internal.h"

void *fcg_dmalloc(fcg_ctx *ctx, size_t bytes) {
  void *p = nullptr;
  if (hipMallocAsync(&p, std::max<size_t>(bytes, 1), ctx->stream) != hipSuccess) {
    ctx->err = "hipMallocAsync failed (" + std::to_string(bytes) + " bytes)";
    return nullptr;
  }
  return p;
}

void fcg_dfree(fcg_ctx *ctx, void *p) {
  if (p) (void)hipFreeAsync(p, ctx->stream);
}

void fcg_plan_note_user(fcg_plan *pl, fcg_ctx *ctx) {
  if (ctx == pl->ctx) return;
  std::lock_guard<std::mutex> g(pl->users_mu);
  if (std::find(pl->users.begin(), pl->users.end(), ctx) == pl->users.end())
    pl->users.push_back(ctx);
}

namespace {
std::mutex g_plans_mu;
std::vector<fcg_plan *> g_plans;  // live plans
}  // namespace

void fcg_plan_register(fcg_plan *pl) {
  {
    // room for the user contexts, so that noting one allocates nothing
    std::lock_guard<std::mutex> g2(pl->users_mu);
    pl->users.reserve(16);
  }
  std::lock_guard<std::mutex> g(g_plans_mu);
  g_plans.push_back(pl);
}

void fcg_plan_unregister(fcg_plan *pl) {
  std::lock_guard<std::mutex> g(g_plans_mu);
  auto it = std::find(g_plans.begin(), g_plans.end(), pl);
  if (it != g_plans.end()) g_plans.erase(it);
}

// a context being destroyed stops being a user of every live plan (its
// stream is synchronised by fcg_destroy itself)
static void forget_user(fcg_ctx *ctx) {
  std::lock_guard<std::mutex> g(g_plans_mu);
  for (fcg_plan *pl : g_plans) {
    std::lock_guard<std::mutex> g2(pl->users_mu);
    auto it = std::find(pl->users.begin(), pl->users.end(), ctx);
    if (it != pl->users.end()) pl->users.erase(it);
  }
}

namespace {

template <typename T>
int dev_alloc(fcg_ctx *ctx, T **dst, int64_t n) {
  *dst = nullptr;
  if (n <= 0) return FCG_OK;
  *dst = static_cast<T *>(fcg_dmalloc(ctx, sizeof(T) * n));
  return *dst ? FCG_OK : FCG_ENOMEM;
}

template <typename T>
int dev_upload(fcg_ctx *ctx, T **dst, const T *src, int64_t n) {
  if (int rc = dev_alloc(ctx, dst, n)) return rc;
  if (n > 0)
    FCG_HIP(hipMemcpyAsync(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice, ctx->stream), ctx);
  return FCG_OK;
}

template <typename T>
void dev_free(fcg_ctx *ctx, T *&p) {
  if (p) fcg_dfree(ctx, p);
  p = nullptr;
}

#define TRY(x)              \
  do {                      \
    int rc_ = (x);          \
    if (rc_) return rc_;    \
  } while (0)

int ensure_scratch(fcg_ctx *ctx, int64_t n) {
  if (n <= ctx->scratch_elems) return FCG_OK;
  dev_free(ctx, ctx->d_c);
  dev_free(ctx, ctx->d_s);
  ctx->scratch_elems = 0;
  TRY(dev_alloc(ctx, &ctx->d_c, n));
  TRY(dev_alloc(ctx, &ctx->d_s, n));
  ctx->scratch_elems = n;
  return FCG_OK;
}

}  // namespace

int fcg_space_to_device(fcg_space *sp) {
  fcg_ctx *ctx = sp->ctx;
  const int64_t nb = sp->n_box * sp->N;
  TRY(dev_upload(ctx, &sp->d_lo, sp->lo.data(), nb));
  TRY(dev_upload(ctx, &sp->d_hi, sp->hi.data(), nb));
  TRY(dev_upload(ctx, &sp->d_rank, sp->rank.data(), nb));
  TRY(dev_upload(ctx, &sp->d_foff, sp->foff.data(), nb));
  TRY(dev_upload(ctx, &sp->d_var, sp->var_off.data(), sp->n_box + 1));
  if (!sp->factors.empty())
    TRY(dev_upload(ctx, &sp->d_factors, sp->factors.data(), (int64_t)sp->factors.size()));
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}

// Scratch of fcg_sigma_device_noalloc: the interleaved copy of c ([n_var][R]
// FP64 or FP32) in a file-scope device array, which each GPU holds until the
// process ends.  Host code never uses the array's name (a host-side
// stand-in, not a device address): fcg_create looks up its address on each
// device once, with that device current, and the entry point passes that
// address to the kernels.  Every call on one GPU shares the array.
constexpr int64_t FCG_NOALLOC_SCRATCH_DOUBLES = int64_t(1) << 23;  // 64 MiB
static __device__ __align__(256) double scratch[FCG_NOALLOC_SCRATCH_DOUBLES];
constexpr int FCG_MAX_DEVICES = 64;

namespace {
std::mutex g_scratch_mu;
void *g_scratch_dev[FCG_MAX_DEVICES] = {};  // address of `scratch` on each device

// the device must be current
int scratch_address(fcg_ctx *ctx) {
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch_dev[ctx->device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      ctx->err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return FCG_EDEVICE;
    }
    g_scratch_dev[ctx->device] = ptr;
  }
  ctx->noalloc_scratch = g_scratch_dev[ctx->device];
  return FCG_OK;
}
}  // namespace

extern "C" {

int fcg_create(int device, fcg_ctx_t *out) {
  if (!out) return FCG_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n ||
      device >= FCG_MAX_DEVICES)
    return FCG_EDEVICE;
  fcg_ctx *ctx = new fcg_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&ctx->stream) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      scratch_address(ctx) != FCG_OK) {
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return FCG_EDEVICE;
  }
  *out = ctx;
  return FCG_OK;
}

void fcg_destroy(fcg_ctx_t ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  dev_free(ctx, ctx->d_c);
  dev_free(ctx, ctx->d_s);
  if (ctx->d_ci) fcg_dfree(ctx, ctx->d_ci);
  if (ctx->d_r1x) fcg_dfree(ctx, ctx->d_r1x);
  if (ctx->d_r1y) fcg_dfree(ctx, ctx->d_r1y);
  (void)hipStreamSynchronize(ctx->stream);
  forget_user(ctx);
  for (auto &e : ctx->slots)
    if (e) (void)hipEventDestroy(e);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *fcg_last_error(fcg_ctx_t ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int fcg_ops_upload(fcg_ctx_t ctx, int n_clusters, int S, int n_ops, const int32_t *gblock,
                   int64_t n_blocks, const int64_t *gb_off, const int32_t *gb_dims,
                   const double *gdata, int64_t n_gdata, fcg_ops_t *out) {
  if (!ctx || !out || n_clusters <= 0 || n_clusters > FCG_MAXN || S <= 1 || n_ops <= 0)
    return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  fcg_ops *o = new fcg_ops();
  o->ctx = ctx;
  o->N = n_clusters;
  o->S = S;
  o->n_ops = n_ops;
  const int64_t SS = (int64_t)S * S;
  o->gblock.assign(gblock, gblock + (int64_t)n_clusters * n_ops * SS * SS);
  o->gb_off.assign(gb_off, gb_off + n_blocks);
  o->gb_dims.assign(gb_dims, gb_dims + 3 * n_blocks);
  for (int64_t b = 0; b < n_blocks; b++) {
    const int64_t sz = (int64_t)gb_dims[3 * b] * gb_dims[3 * b + 1] * gb_dims[3 * b + 2];
    if (gb_off[b] < 0 || gb_off[b] + sz > n_gdata) {
      ctx->err = "Γ block " + std::to_string(b) + " outside gdata";
      delete o;
      return FCG_EINVAL;
    }
  }
  o->nstates.assign(n_clusters * SS, 0);
  for (int c = 0; c < n_clusters; c++)
    for (int code = 0; code < SS; code++) {
      const int blk = o->block(c, 0 /* "H" */, code, code);
      if (blk >= 0) o->nstates[c * SS + code] = gb_dims[3 * blk + 1];
    }
  o->n_gdata = n_gdata;
  int rc = dev_upload(ctx, &o->d_gdata, gdata, n_gdata);
  if (!rc) rc = dev_upload(ctx, &o->d_gb_off, gb_off, n_blocks);
  if (!rc) rc = dev_upload(ctx, &o->d_gb_dims, gb_dims, 3 * n_blocks);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FCG_EDEVICE;
  if (rc) {
    fcg_ops_destroy(o);
    return rc;
  }
  *out = o;
  return FCG_OK;
}

void fcg_ops_destroy(fcg_ops_t o) {
  if (!o) return;
  dev_free(o->ctx, o->d_gdata);
  dev_free(o->ctx, o->d_gb_off);
  dev_free(o->ctx, o->d_gb_dims);
  delete o;
}

int fcg_ham_upload(fcg_ctx_t ctx, int n_clusters, int n_terms, const int32_t *nbody,
                   const int32_t *clusters, const int32_t *ops, const int32_t *parity,
                   const int32_t *dims, const int16_t *delta, const int64_t *ioff,
                   const double *ints, int64_t n_ints, int n_tc, const int16_t *tc,
                   const int32_t *tc_start, fcg_ham_t *out) {
  if (!ctx || !out || n_clusters <= 0 || n_clusters > FCG_MAXN || n_terms < 0 || n_tc < 0)
    return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  fcg_ham *h = new fcg_ham();
  h->ctx = ctx;
  h->N = n_clusters;
  h->n_terms = n_terms;
  h->n_tc = n_tc;
  h->nbody.assign(nbody, nbody + n_terms);
  h->clusters.assign(clusters, clusters + 4 * n_terms);
  h->ops.assign(ops, ops + 4 * n_terms);
  h->parity.assign(parity, parity + 4 * n_terms);
  h->dims.assign(dims, dims + 4 * n_terms);
  h->delta.assign(delta, delta + 8 * n_terms);
  h->ioff.assign(ioff, ioff + n_terms);
  h->tc.assign(tc, tc + (int64_t)n_tc * 2 * n_clusters);
  h->tc_start.assign(tc_start, tc_start + n_tc + 1);
  std::vector<TermDev> td(n_terms);
  for (int t = 0; t < n_terms; t++) {
    const int nb = nbody[t];
    if (nb < 1 || nb > 4 || tc_start[n_tc] != n_terms) {
      ctx->err = "bad term table";
      delete h;
      return FCG_EINVAL;
    }
    int64_t sz = 1;
    td[t].nbody = nb;
    for (int a = 0; a < 4; a++) {
      td[t].cl[a] = clusters[4 * t + a];
      td[t].dims[a] = dims[4 * t + a];
      if (a < nb) {
        if (clusters[4 * t + a] < 0 || clusters[4 * t + a] >= n_clusters ||
            (a > 0 && clusters[4 * t + a] <= clusters[4 * t + a - 1])) {
          ctx->err = "term clusters must be ascending and in range";
          delete h;
          return FCG_EINVAL;
        }
        if (nb > 1) sz *= dims[4 * t + a];
      }
    }
    td[t].pad = 0;
    td[t].ioff = ioff[t];
    if (nb > 1 && (ioff[t] < 0 || ioff[t] + sz > n_ints)) {
      ctx->err = "term integrals outside ints";
      delete h;
      return FCG_EINVAL;
    }
  }
  int rc = dev_upload(ctx, &h->d_terms, td.data(), n_terms);
  if (!rc) rc = dev_upload(ctx, &h->d_ints, ints, n_ints);
  if (!rc && hipStreamSynchronize(ctx->stream) != hipSuccess) rc = FCG_EDEVICE;
  if (rc) {
    fcg_ham_destroy(h);
    return rc;
  }
  *out = h;
  return FCG_OK;
}

void fcg_ham_destroy(fcg_ham_t h) {
  if (!h) return;
  dev_free(h->ctx, h->d_terms);
  dev_free(h->ctx, h->d_ints);
  delete h;
}

int fcg_tpsci_space_upload(fcg_ctx_t ctx, int n_clusters, int64_t n_fock, const int16_t *fock,
                           const int64_t *fock_start, int64_t n_cfg, const int16_t *cfg,
                           fcg_space_t *out) {
  // The space needs ClusterOps state counts to recognise full blocks; they
  // are taken at plan time, so here the caller's arrays are only kept.
  if (!ctx || !out || n_clusters <= 0 || n_clusters > FCG_MAXN || n_fock < 0 || n_cfg < 0)
    return FCG_EINVAL;
  fcg_space *sp = new fcg_space();
  sp->ctx = ctx;
  sp->N = n_clusters;
  sp->tpsci = true;
  sp->n_fock = n_fock;
  sp->fock.assign(fock, fock + n_fock * 2 * n_clusters);
  // stash configurations in lo/hi until boxes are formed against an ops table
  sp->var_off.assign(fock_start, fock_start + n_fock + 1);
  sp->lo.assign(cfg, cfg + n_cfg * n_clusters);
  sp->nvar = n_cfg;
  sp->n_box = -1;  // boxes not formed yet
  *out = sp;
  return FCG_OK;
}

int fcg_bst_space_upload(fcg_ctx_t ctx, int n_clusters, int64_t n_fock, const int16_t *fock,
                         int64_t n_box, const int32_t *box_fock, const int16_t *box_lo,
                         const int16_t *box_hi, const int16_t *box_rank, const int64_t *box_foff,
                         const double *factors, int64_t n_factors, fcg_space_t *out) {
  if (!ctx || !out || n_clusters <= 0 || n_clusters > FCG_MAXN || n_fock < 0 || n_box < 0)
    return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  const int N = n_clusters;
  fcg_space *sp = new fcg_space();
  sp->ctx = ctx;
  sp->N = N;
  sp->n_fock = n_fock;
  sp->n_box = n_box;
  sp->fock.assign(fock, fock + n_fock * 2 * N);
  sp->box_fock.assign(box_fock, box_fock + n_box);
  sp->lo.assign(box_lo, box_lo + n_box * N);
  sp->hi.assign(box_hi, box_hi + n_box * N);
  sp->rank.assign(box_rank, box_rank + n_box * N);
  sp->foff.assign(box_foff, box_foff + n_box * N);
  sp->factors.assign(factors, factors + n_factors);
  sp->var_off.assign(1, 0);
  for (int64_t b = 0; b < n_box; b++) {
    if (box_fock[b] < 0 || box_fock[b] >= n_fock) {
      ctx->err = "box Fock index out of range";
      delete sp;
      return FCG_EINVAL;
    }
    int64_t d = 1;
    for (int c = 0; c < N; c++) {
      const int64_t i = b * N + c;
      const int nrg = box_hi[i] - box_lo[i] + 1;
      if (box_lo[i] < 1 || nrg < 1 || box_rank[i] < 1 ||
          (box_foff[i] < 0 && box_rank[i] != nrg) ||
          (box_foff[i] >= 0 && (box_rank[i] > nrg ||
                                box_foff[i] + (int64_t)nrg * box_rank[i] > n_factors))) {
        ctx->err = "bad box " + std::to_string(b) + " cluster " + std::to_string(c);
        delete sp;
        return FCG_EINVAL;
      }
      d *= box_rank[i];
    }
    sp->var_off.push_back(sp->var_off.back() + d);
  }
  sp->nvar = sp->var_off.back();
  int rc = fcg_space_to_device(sp);
  if (rc) {
    fcg_space_destroy(sp);
    return rc;
  }
  *out = sp;
  return FCG_OK;
}

int64_t fcg_space_nvar(fcg_space_t sp) { return sp ? sp->nvar : -1; }

void fcg_space_destroy(fcg_space_t sp) {
  if (!sp) return;
  dev_free(sp->ctx, sp->d_lo);
  dev_free(sp->ctx, sp->d_hi);
  dev_free(sp->ctx, sp->d_rank);
  dev_free(sp->ctx, sp->d_foff);
  dev_free(sp->ctx, sp->d_var);
  dev_free(sp->ctx, sp->d_dim);
  dev_free(sp->ctx, sp->d_factors);
  delete sp;
}

}  // extern "C"

// TPSCI spaces form their boxes at first use, against the ClusterOps table
int fcg_form_boxes(fcg_ctx *ctx, fcg_ops *o, fcg_space *sp) {
  if (!sp->tpsci || sp->n_box >= 0) return FCG_OK;
  std::vector<int64_t> fstart = sp->var_off;
  std::vector<int16_t> cfg = sp->lo;
  std::vector<int16_t> fock = sp->fock;
  sp->lo.clear();
  sp->var_off.clear();
  int rc = fcg_tpsci_boxes(o, sp->N, sp->n_fock, fock.data(), fstart.data(),
                           (int64_t)cfg.size() / sp->N, cfg.data(), sp, &ctx->err);
  if (rc) {
    sp->lo = cfg;
    sp->var_off = fstart;
    sp->box_fock.clear();
    sp->hi.clear();
    sp->rank.clear();
    sp->foff.clear();
    sp->n_box = -1;
    return rc;
  }
  return fcg_space_to_device(sp);
}

namespace {
std::atomic<int64_t> g_plan_reserved{0};
std::atomic<int64_t> g_plan_budget{INT64_MAX};
}  // namespace

int fcg_plan_reserve(fcg_ctx *ctx, int64_t bytes, int64_t max_bytes, int64_t *need) {
  if (need) *need = bytes;
  if (bytes > max_bytes) {
    ctx->err = "plan needs " + std::to_string(bytes) + " device bytes, limit " +
               std::to_string(max_bytes);
    return FCG_ENOMEM;
  }
  int64_t cur = g_plan_reserved.load();
  do {
    if (cur + bytes > g_plan_budget.load()) {
      ctx->err = "plan needs " + std::to_string(bytes) + " device bytes; " +
                 std::to_string(cur) + " of the plan budget " +
                 std::to_string(g_plan_budget.load()) + " are held";
      return FCG_ENOMEM;
    }
  } while (!g_plan_reserved.compare_exchange_weak(cur, cur + bytes));
  return FCG_OK;
}

void fcg_plan_release(int64_t bytes) { g_plan_reserved.fetch_sub(bytes); }

int64_t fcg_plan_bytes(int64_t n_kel, int64_t n_cols, int64_t n_box, int precisions) {
  int64_t per = 8;
  if (precisions & (1 << FCG_FP32)) per += 4;
  if (precisions & (1 << FCG_BF16X2)) per += 4;
  if (precisions & (1 << FCG_BF16)) per += 2;
  return n_kel * per + n_cols * 4 + n_box * 48;
}

int fcg_create_plan(fcg_ctx *ctx, fcg_ham *h, fcg_ops *o, fcg_space *sp, fcg_space *ksp,
                    int nbody, int precisions, fcg_plan **out, int64_t max_bytes, int64_t *need,
                    const double *thr, const FusedSigma *fused) {
  if (!ctx || !h || !o || !sp || !ksp || !out || nbody < 1 || nbody > 4) return FCG_EINVAL;
  // the bf16-c apply reads the FP32 blocks
  if (precisions & (1 << FCG_FP32_BF16C)) precisions |= 1 << FCG_FP32;
  (void)hipSetDevice(ctx->device);
  auto t0 = std::chrono::steady_clock::now();
  TRY(fcg_form_boxes(ctx, o, sp));
  if (ksp != sp) TRY(fcg_form_boxes(ctx, o, ksp));
  if (ksp == sp && sp->tpsci && sp->n_box > 0 && sp->nvar == sp->n_box && !thr && !fused &&
      fcg_sparse_plan_fits(sp)) {
    // every box one configuration: join and build on the device (spaces
    // beyond the device join's limits take the host join below)
    const int rc = fcg_create_sparse_plan(ctx, h, o, sp, nbody, precisions, out, max_bytes, need);
    if (rc == FCG_OK) fcg_plan_register(*out);
    return rc;
  }
  HostPlan hp;
  const bool store = fused == nullptr;
  int rc = fcg_build_host_plan(h, o, sp, ksp, nbody, store, ctx->build_class_cap, &hp, &ctx->err);
  if (rc) return rc;
  // fused sigma: each pair's products get a slot of db x R doubles
  std::vector<int64_t> fyoff;
  if (fused) {
    fyoff.resize(hp.pair_B.size() + 1);
    fyoff[0] = 0;
    for (size_t q = 0; q < hp.pair_B.size(); q++) {
      const int32_t B = hp.pair_B[q];
      fyoff[q + 1] = fyoff[q] + (sp->var_off[B + 1] - sp->var_off[B]) * fused->R;
    }
  }
  int64_t reserved = 0;
  {
    // the pair and task tables (and a fused sigma's pair slots) live on the
    // device during the build
    const int64_t bytes = (fused ? 8 * (fyoff.back() + (int64_t)fyoff.size() + sp->n_box + 1)
                                 : fcg_plan_bytes(hp.n_kel, (int64_t)hp.colidx.size(), sp->n_box,
                                                  precisions | (1 << FCG_FP64))) +
                          (int64_t)hp.pair_B.size() * 44 + (int64_t)hp.tasks.size() * sizeof(TaskDev);
    if (int rc2 = fcg_plan_reserve(ctx, bytes, max_bytes, need)) return rc2;
    reserved = bytes;
  }
  fcg_plan *pl = new fcg_plan();
  pl->reserved_bytes = reserved;
  pl->d_thr = thr;
  pl->ctx = ctx;
  pl->sp = sp;
  pl->ksp = ksp;
  pl->n_pair = (int64_t)hp.pair_B.size();
  pl->n_task = (int64_t)hp.tasks.size();
  pl->n_kel = hp.n_kel;
  pl->n_cols = (int64_t)hp.colidx.size();
  pl->build_flops = hp.flops;
  pl->precisions = precisions | (1 << FCG_FP64);
#define PTRY(x)                 \
  do {                          \
    int rc_ = (x);              \
    if (rc_) {                  \
      fcg_plan_destroy(pl);     \
      return rc_;               \
    }                           \
  } while (0)
  PTRY(dev_upload(ctx, &pl->d_pair_B, hp.pair_B.data(), pl->n_pair));
  PTRY(dev_upload(ctx, &pl->d_pair_K, hp.pair_K.data(), pl->n_pair));
  PTRY(dev_upload(ctx, &pl->d_pair_t0, hp.pair_t0.data(), pl->n_pair + 1));
  PTRY(dev_upload(ctx, &pl->d_tasks, hp.tasks.data(), pl->n_task));
  PTRY(dev_upload(ctx, &pl->d_pair_list, hp.pair_list.data(), pl->n_pair));
  pl->n_tiny = hp.n_tiny;
  pl->n_small = hp.n_small;
  pl->n_mid = hp.n_mid;
  pl->holds_blocks = store;
  PTRY(dev_upload(ctx, &pl->d_row_ptr, hp.row_ptr.data(), sp->n_box + 1));
  FusedSigma fz{};
  struct FusedSlots {  // freed on every return (after the build's synchronisation)
    fcg_ctx *ctx;
    double *part = nullptr;
    int64_t *yoff = nullptr;
    ~FusedSlots() {
      dev_free(ctx, part);
      dev_free(ctx, yoff);
    }
  } slots{ctx};
  if (fused) {
    fz = *fused;
    PTRY(dev_alloc(ctx, &slots.part, std::max<int64_t>(fyoff.back(), 1)));
    PTRY(dev_upload(ctx, &slots.yoff, fyoff.data(), (int64_t)fyoff.size()));
    fz.part = slots.part;
    fz.yoff = slots.yoff;
  }
  if (store) {  // the apply's layout: only for a plan that keeps its blocks
    PTRY(dev_upload(ctx, &pl->d_pair_koff, hp.pair_koff.data(), pl->n_pair));
    PTRY(dev_upload(ctx, &pl->d_pair_col, hp.pair_col.data(), pl->n_pair));
    PTRY(dev_upload(ctx, &pl->d_diag_koff, hp.diag_koff.data(), sp->n_box));
    PTRY(dev_upload(ctx, &pl->d_diag_col, hp.diag_col.data(), sp->n_box));
    PTRY(dev_upload(ctx, &pl->d_row_koff, hp.row_koff.data(), sp->n_box));
    PTRY(dev_upload(ctx, &pl->d_col_ptr, hp.col_ptr.data(), sp->n_box + 1));
    PTRY(dev_upload(ctx, &pl->d_order, hp.order.data(), (int64_t)hp.order.size()));
    pl->n_multi = hp.n_multi;
    pl->n_single_slots = hp.n_single_slots;
    PTRY(dev_upload(ctx, &pl->d_colidx, hp.colidx.data(), pl->n_cols));
  }
  if (store) {
    PTRY(dev_alloc(ctx, &pl->d_K64, pl->n_kel));
    if (pl->precisions & (1 << FCG_FP32)) PTRY(dev_alloc(ctx, &pl->d_K32, pl->n_kel));
    if (pl->precisions & (1 << FCG_BF16X2)) PTRY(dev_alloc(ctx, &pl->d_Kbf, pl->n_kel));
    if (pl->precisions & (1 << FCG_BF16)) PTRY(dev_alloc(ctx, &pl->d_Kb16, pl->n_kel));
  }
  if (store && pl->n_kel > 0) {
    // zero columns pad each row panel to whole quads
    const hipError_t e = hipMemsetAsync(pl->d_K64, 0, sizeof(double) * pl->n_kel, ctx->stream);
    if (e != hipSuccess) {
      ctx->err = std::string("block clear: ") + hipGetErrorString(e);
      fcg_plan_destroy(pl);
      return FCG_EDEVICE;
    }
  }
  if (pl->n_pair > 0) {
    hipError_t e = fcg_launch_build(pl, h, o, ctx->stream, fused ? &fz : nullptr);
    if (e == hipSuccess && fused) e = fcg_launch_fused_sum(pl, &fz, pl->d_row_ptr, ctx->stream);
    if (e == hipSuccess && store) e = fcg_launch_convert(pl, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      ctx->err = std::string("block build: ") + hipGetErrorString(e);
      fcg_plan_destroy(pl);
      return FCG_EDEVICE;
    }
  }
  // device copies of the pair and task bookkeeping are only needed for the
  // build: free them and release their share of the reservation, so that
  // the plan holds exactly fcg_plan_device_bytes
  dev_free(ctx, pl->d_tasks);
  dev_free(ctx, pl->d_pair_list);
  dev_free(ctx, pl->d_pair_t0);
  dev_free(ctx, pl->d_pair_col);
  dev_free(ctx, pl->d_pair_B);
  dev_free(ctx, pl->d_pair_K);
  dev_free(ctx, pl->d_pair_koff);
  dev_free(ctx, pl->d_row_ptr);
  {
    const int64_t keep =
        store ? fcg_plan_bytes(pl->n_kel, pl->n_cols, sp->n_box, pl->precisions) : 0;
    if (pl->reserved_bytes > keep) {
      fcg_plan_release(pl->reserved_bytes - keep);
      pl->reserved_bytes = keep;
    }
  }
  pl->d_thr = nullptr;
  pl->build_us =
      std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
  fcg_plan_register(pl);
  *out = pl;
  return FCG_OK;
#undef PTRY
}

extern "C" {

int fcg_plan_create(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t sp, int nbody,
                    int precisions, fcg_plan_t *out) {
  return fcg_create_plan(ctx, h, o, sp, sp, nbody, precisions, out);
}

int fcg_plan_create_pair(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t bra,
                         fcg_space_t ket, int nbody, int precisions, fcg_plan_t *out) {
  return fcg_create_plan(ctx, h, o, bra, ket, nbody, precisions, out);
}

int fcg_plan_create_pair_screened(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t bra,
                                  fcg_space_t ket, int nbody, int precisions, int R,
                                  const double *c, double thresh, fcg_plan_t *out) {
  if (!ctx || !ket || !c || R < 1 || !(thresh >= 0.0)) return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  TRY(fcg_form_boxes(ctx, o, ket));
  double *thr = nullptr;
  if (thresh > 0.0) TRY(fcg_screen_thresholds(ctx, ket->nvar, R, c, thresh, &thr));
  const int rc = fcg_create_plan(ctx, h, o, bra, ket, nbody, precisions, out, INT64_MAX, nullptr,
                                 thr);
  if (thr) fcg_dfree(ctx, thr);
  return rc;
}

int fcg_plan_create_limited(fcg_ctx_t ctx, fcg_ham_t h, fcg_ops_t o, fcg_space_t sp, int nbody,
                            int precisions, int64_t max_bytes, int64_t *need_bytes,
                            fcg_plan_t *out) {
  if (need_bytes) *need_bytes = 0;
  return fcg_create_plan(ctx, h, o, sp, sp, nbody, precisions, out, max_bytes, need_bytes);
}

int fcg_set_plan_budget(int64_t bytes) {
  if (bytes <= 0) return FCG_EINVAL;
  g_plan_budget.store(bytes);
  return FCG_OK;
}

int fcg_plan_budget(int64_t *budget, int64_t *held) {
  if (budget) *budget = g_plan_budget.load();
  if (held) *held = g_plan_reserved.load();
  return FCG_OK;
}

int fcg_plan_device_bytes(fcg_plan_t pl, int64_t *bytes) {
  if (!pl || !bytes) return FCG_EINVAL;
  *bytes = pl->holds_blocks ? fcg_plan_bytes(pl->n_kel, pl->n_cols, pl->sp->n_box, pl->precisions)
                            : 0;
  return FCG_OK;
}

void fcg_plan_destroy(fcg_plan_t pl) {
  if (!pl) return;
  fcg_plan_unregister(pl);
  // other contexts that read the plan finish first; the plan's own stream
  // orders the frees after its work
  {
    std::lock_guard<std::mutex> g(pl->users_mu);
    for (fcg_ctx *u : pl->users) (void)hipStreamSynchronize(u->stream);
  }
  fcg_ctx *ctx = pl->ctx;
  fcg_rank1_free(ctx, pl);
  dev_free(ctx, pl->d_row_ptr);
  dev_free(ctx, pl->d_pair_B);
  dev_free(ctx, pl->d_pair_K);
  dev_free(ctx, pl->d_pair_t0);
  dev_free(ctx, pl->d_pair_koff);
  dev_free(ctx, pl->d_pair_col);
  dev_free(ctx, pl->d_tasks);
  dev_free(ctx, pl->d_pair_list);
  dev_free(ctx, pl->d_diag_koff);
  dev_free(ctx, pl->d_diag_col);
  dev_free(ctx, pl->d_row_koff);
  dev_free(ctx, pl->d_col_ptr);
  dev_free(ctx, pl->d_order);
  dev_free(ctx, pl->d_colidx);
  dev_free(ctx, pl->d_K64);
  dev_free(ctx, pl->d_K32);
  dev_free(ctx, pl->d_Kbf);
  dev_free(ctx, pl->d_Kb16);
  fcg_plan_release(pl->reserved_bytes);
  delete pl;
}

int fcg_set_build_class_cap(fcg_ctx_t ctx, int cap) {
  if (!ctx || cap < 0 || cap > 3) return FCG_EINVAL;
  ctx->build_class_cap = cap;
  return FCG_OK;
}

int fcg_plan_stats(fcg_plan_t pl, double *o6) {
  if (!pl || !o6) return FCG_EINVAL;
  o6[0] = (double)pl->n_pair;
  o6[1] = (double)pl->n_task;
  o6[2] = (double)pl->n_kel;
  o6[3] = pl->build_flops;
  o6[4] = pl->build_us;
  o6[5] = (double)pl->n_cols;
  return FCG_OK;
}

int fcg_plan_apply_variant(fcg_plan_t pl, int precision, int R, int *multi, int *rows) {
  if (!pl || precision < FCG_FP64 || precision >= FCG_N_PREC || R < 1 || R > FCG_MAX_R)
    return FCG_EINVAL;
  if (multi) *multi = pl->tuned[precision][R] ? pl->tuned_vm[precision][R] : -1;
  if (rows) *rows = pl->tuned[precision][R] ? pl->tuned_vr[precision][R] : -1;
  return FCG_OK;
}

int fcg_apply_variant_counts(int *n_multi, int *n_rows, int *max_r) {
  if (!n_multi || !n_rows || !max_r) return FCG_EINVAL;
  fcg_apply_n_variants(n_multi, n_rows, max_r);
  return FCG_OK;
}

int fcg_plan_set_apply_variant(fcg_plan_t pl, int precision, int R, int multi, int rows) {
  if (!pl || precision < FCG_FP64 || precision >= FCG_N_PREC || R < 1 || R > FCG_MAX_R)
    return FCG_EINVAL;
  int nm = 1, nr = 1, max_r = 0;
  fcg_apply_n_variants(&nm, &nr, &max_r);
  if (multi < 0 || multi >= nm || rows < 0 || rows >= nr) return FCG_EINVAL;
  // above the tuned root count each kernel has one compiled form
  if (R > max_r && (multi != fcg_apply_default_variant(precision) || rows != 0))
    return FCG_EINVAL;
  std::lock_guard<std::mutex> g(pl->tune_mu);
  pl->tuned_vm[precision][R] = (int8_t)multi;
  pl->tuned_vr[precision][R] = (int8_t)rows;
  pl->tuned[precision][R] = true;
  return FCG_OK;
}

}  // extern "C"

// Autotuning of the apply: on the first sigma of a plan for a precision and
// root count (at most the tuned maximum), each variant of each apply kernel
// runs twice on the caller's operands and the faster second run wins.  The
// variants compute identical results, so the outputs of the timing runs are
// simply overwritten.
static int tune_apply(fcg_ctx *ctx, fcg_plan *pl, int R, const double *c, double *s,
                      int precision, void *ci) {
  int nm = 1, nr = 1, max_r = 0;
  fcg_apply_n_variants(&nm, &nr, &max_r);
  int best_m = fcg_apply_default_variant(precision), best_r = 0;
  if (R <= max_r) {
    auto time_it = [&](int vm, int vr, int parts, float *ms) -> int {
      for (int rep = 0; rep < 2; rep++) {
        FCG_HIP(fcg_launch_apply(pl, R, c, s, precision, ctx->stream, ctx->ev0, vm, vr, ci,
                                 parts),
                ctx);
        FCG_HIP(hipEventRecord(ctx->ev1, ctx->stream), ctx);
        FCG_HIP(hipEventSynchronize(ctx->ev1), ctx);
        FCG_HIP(hipEventElapsedTime(ms, ctx->ev0, ctx->ev1), ctx);
      }
      return FCG_OK;
    };
    float best = 1e30f, t = 0.f;
    if (pl->n_multi > 0)
      for (int vm = 0; vm < nm; vm++) {
        TRY(time_it(vm, 0, 1, &t));
        if (t < best) { best = t; best_m = vm; }
      }
    best = 1e30f;
    if (pl->n_single_slots > 0)
      for (int vr = 0; vr < nr; vr++) {
        TRY(time_it(best_m, vr, 2, &t));
        if (t < best) { best = t; best_r = vr; }
      }
  }
  pl->tuned_vm[precision][R] = (int8_t)best_m;
  pl->tuned_vr[precision][R] = (int8_t)best_r;
  pl->tuned[precision][R] = true;
  return FCG_OK;
}

extern "C" {

int fcg_sigma_device(fcg_ctx_t ctx, fcg_plan_t pl, int R, const double *c_dev, double *s_dev,
                     int precision) {
  if (!ctx || !pl || R < 1 || R > FCG_MAX_R || !c_dev || !s_dev) return FCG_EINVAL;
  if (precision < FCG_FP64 || precision >= FCG_N_PREC || !(pl->precisions & (1 << precision))) {
    ctx->err = "precision not built into this plan";
    return FCG_EINVAL;
  }
  if (!pl->holds_blocks) {
    ctx->err = "plan holds no blocks";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  fcg_plan_note_user(pl, ctx);
  if (pl->r1) {  // matrix-free plan of a rank <= 1 space
    ctx->last_launches = 1;
    return fcg_rank1_sigma(ctx, pl, R, c_dev, s_dev, precision);
  }
  const int64_t n = (int64_t)R * pl->ksp->nvar;
  if (ctx->ci_bytes < n * 8) {  // interleaved copy of c, sized for the FP64 form
    // stream-ordered on this context: the old copy's readers ran on it
    if (ctx->d_ci) fcg_dfree(ctx, ctx->d_ci);
    ctx->d_ci = nullptr;
    ctx->ci_bytes = 0;
    double *p = nullptr;
    TRY(dev_alloc(ctx, &p, n));
    ctx->d_ci = p;
    ctx->ci_bytes = n * 8;
  }
  {
    std::lock_guard<std::mutex> g(pl->tune_mu);
    if (!pl->tuned[precision][R])
      TRY(tune_apply(ctx, pl, R, c_dev, s_dev, precision, ctx->d_ci));
  }
  FCG_HIP(fcg_launch_apply(pl, R, c_dev, s_dev, precision, ctx->stream, ctx->ev0,
                           pl->tuned_vm[precision][R], pl->tuned_vr[precision][R], ctx->d_ci),
          ctx);
  FCG_HIP(hipEventRecord(ctx->ev1, ctx->stream), ctx);
  ctx->last_launches = 2;
  return FCG_OK;
}

int fcg_sigma_device_noalloc(fcg_ctx_t ctx, fcg_plan_t pl, int R, const double *c_dev,
                             double *s_dev, int precision) {
  if (!ctx || !pl || R < 1 || R > FCG_MAX_R || !c_dev || !s_dev) return FCG_EINVAL;
  if (precision < FCG_FP64 || precision >= FCG_N_PREC || !(pl->precisions & (1 << precision))) {
    ctx->err = "precision not built into this plan";
    return FCG_EINVAL;
  }
  if (!pl->holds_blocks) {
    ctx->err = "plan holds no blocks";
    return FCG_EINVAL;
  }
  if (pl->r1) {
    ctx->err = "fcg_sigma_device_noalloc: matrix-free plans use layout buffers; use "
               "fcg_sigma_device";
    return FCG_EINVAL;
  }
  const int64_t n = (int64_t)R * pl->ksp->nvar;
  if (n > FCG_NOALLOC_SCRATCH_DOUBLES) {
    ctx->err = "fcg_sigma_device_noalloc: R x n_var = " + std::to_string(n) +
               " exceeds the fixed scratch of " + std::to_string(FCG_NOALLOC_SCRATCH_DOUBLES) +
               " values (64 MiB); use fcg_sigma_device";
    return FCG_EOVERFLOW;
  }
  if (!ctx->noalloc_scratch) {
    ctx->err = "fcg_sigma_device_noalloc: scratch address unknown";
    return FCG_EDEVICE;
  }
  (void)hipSetDevice(ctx->device);
  fcg_plan_note_user(pl, ctx);  // capacity reserved at plan creation
  {
    std::lock_guard<std::mutex> g(pl->tune_mu);
    if (!pl->tuned[precision][R])
      TRY(tune_apply(ctx, pl, R, c_dev, s_dev, precision, ctx->noalloc_scratch));
  }
  FCG_HIP(fcg_launch_apply(pl, R, c_dev, s_dev, precision, ctx->stream, ctx->ev0,
                           pl->tuned_vm[precision][R], pl->tuned_vr[precision][R],
                           ctx->noalloc_scratch),
          ctx);
  FCG_HIP(hipEventRecord(ctx->ev1, ctx->stream), ctx);
  ctx->last_launches = 2;
  return FCG_OK;
}

int fcg_sigma(fcg_ctx_t ctx, fcg_plan_t pl, int R, const double *c, double *sigma,
              int precision) {
  if (!ctx || !pl || R < 1 || R > FCG_MAX_R || !c || !sigma) return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  const int64_t nc = (int64_t)R * pl->ksp->nvar, ns = (int64_t)R * pl->sp->nvar;
  TRY(ensure_scratch(ctx, std::max(nc, ns)));
  FCG_HIP(hipMemcpyAsync(ctx->d_c, c, sizeof(double) * nc, hipMemcpyHostToDevice, ctx->stream),
          ctx);
  TRY(fcg_sigma_device(ctx, pl, R, ctx->d_c, ctx->d_s, precision));
  FCG_HIP(hipMemcpyAsync(sigma, ctx->d_s, sizeof(double) * ns, hipMemcpyDeviceToHost,
                         ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}

int fcg_diagonal(fcg_ctx_t ctx, fcg_plan_t pl, double *hd) {
  if (!ctx || !pl || !hd) return FCG_EINVAL;
  if (pl->ksp != pl->sp || !pl->holds_blocks) {
    ctx->err = "diagonal needs a plan within one space that holds its blocks";
    return FCG_EINVAL;
  }
  (void)hipSetDevice(ctx->device);
  fcg_plan_note_user(pl, ctx);
  const int64_t n = pl->sp->nvar;
  TRY(ensure_scratch(ctx, n));
  if (pl->r1)
    TRY(fcg_rank1_diag(ctx, pl, ctx->d_s));
  else
    FCG_HIP(fcg_launch_diag(pl, ctx->d_s, ctx->stream), ctx);
  FCG_HIP(hipMemcpyAsync(hd, ctx->d_s, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream),
          ctx);
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}

int fcg_last_sigma_time(fcg_ctx_t ctx, double *ms, int *launches) {
  if (!ctx) return FCG_EINVAL;
  FCG_HIP(hipEventSynchronize(ctx->ev1), ctx);
  float t = 0.f;
  FCG_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1), ctx);
  if (ms) *ms = t;
  if (launches) *launches = ctx->last_launches;
  return FCG_OK;
}

int fcg_event_record(fcg_ctx_t ctx, int slot) {
  if (!ctx || slot < 0 || slot >= 8) return FCG_EINVAL;
  (void)hipSetDevice(ctx->device);
  if (!ctx->slots[slot]) FCG_HIP(hipEventCreate(&ctx->slots[slot]), ctx);
  FCG_HIP(hipEventRecord(ctx->slots[slot], ctx->stream), ctx);
  return FCG_OK;
}

int fcg_event_elapsed(fcg_ctx_t ctx, int a, int b, double *ms) {
  if (!ctx || !ms || a < 0 || a >= 8 || b < 0 || b >= 8 || !ctx->slots[a] || !ctx->slots[b])
    return FCG_EINVAL;
  FCG_HIP(hipEventSynchronize(ctx->slots[b]), ctx);
  float t = 0.f;
  FCG_HIP(hipEventElapsedTime(&t, ctx->slots[a], ctx->slots[b]), ctx);
  *ms = t;
  return FCG_OK;
}

int fcg_synchronize(fcg_ctx_t ctx) {
  if (!ctx) return FCG_EINVAL;
  FCG_HIP(hipStreamSynchronize(ctx->stream), ctx);
  return FCG_OK;
}

}  // extern "C"
