// Internal types of libfcg_hip.so (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <sys/mman.h>

#include <cstdint>
#include <cstdlib>
#include <memory>
#include <new>
#include <mutex>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "../../include/fcg.h"

// Limits of the block engine.  A "box" is a tensor-product block of the
// basis: one Fock configuration plus, per cluster, a range of local states
// and a Tucker factor (identity for TPSCI).  Hamiltonian blocks between two
// boxes are built whole in one workgroup, so their sizes are bounded by LDS.
constexpr int FCG_MAXN = 16;          // clusters
constexpr int FCG_MAX_BOX = 64;       // rows/cols of one Hamiltonian block
constexpr int FCG_LDS_X = 4096;       // chain intermediates (doubles)
constexpr int FCG_LDS_T = 2048;       // factor-compression scratch (doubles)
constexpr int FCG_LDS_S = 1024;       // inactive-cluster overlaps (doubles)
// smaller LDS footprints for the block pairs that fit them (several
// workgroups per CU instead of one)
// tiny class: blocks of at most FCG_TINY_ELEMS elements, built by 64 threads
constexpr int FCG_LDS_X_TINY = 256, FCG_LDS_T_TINY = 128, FCG_LDS_S_TINY = 64;
constexpr int FCG_TINY_ELEMS = 512;
constexpr int FCG_LDS_X_SMALL = 1024, FCG_LDS_T_SMALL = 512, FCG_LDS_S_SMALL = 256;
// the middle class stages Γ slices as large as the full class's (its arena
// and factor rows are smaller), so that two workgroups still share a CU
constexpr int FCG_LDS_X_MID = 2048, FCG_LDS_G_MID = 4096, FCG_LDS_T_MID = 768, FCG_LDS_S_MID = 512;
constexpr int FCG_MAX_R = 16;         // roots per sigma launch
constexpr int FCG_N_PREC = 5;  // FCG_FP64 .. FCG_FP32_BF16C
constexpr int FCG_QUAD = 4;           // panel columns per stored column quad
constexpr int FCG_CSR_WPB = 4;        // one-row boxes: wavefronts (rows) per workgroup
constexpr int FCG_XCDS = 8;           // MI355X accelerator complex dies (workgroups round-robin)

// Row panel storage.  The Hamiltonian blocks of one bra box (db rows) sit
// side by side as a db x ncol panel, ncol a multiple of FCG_QUAD; the panel
// is stored as consecutive column quads, each quad row-major (db rows of
// FCG_QUAD consecutive columns), so one 16-byte FP32 load reads four columns
// of one row.
__host__ __device__ inline int64_t panel_index(int db, int row, int64_t col) {
  return (int64_t)FCG_QUAD * db * (col / FCG_QUAD) + (int64_t)FCG_QUAD * row + col % FCG_QUAD;
}

#define FCG_HIP(call, ctx)                                                        \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess) {                                                       \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
      return FCG_EDEVICE;                                                         \
    }                                                                             \
  } while (0)

struct fcg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t slots[8] = {};
  std::string err;
  double last_ms = 0.0;
  int last_launches = 0;
  // scratch for host-pointer sigma calls
  double *d_c = nullptr, *d_s = nullptr;
  int64_t scratch_elems = 0;
  // c interleaved ([n_var][R], FP64 or FP32) for the apply: owned by the
  // context, so plans shared between contexts never share it
  void *d_ci = nullptr;
  int64_t ci_bytes = 0;
  // device address of the fixed scratch array of fcg_sigma_device_noalloc
  // on this context's device
  void *noalloc_scratch = nullptr;
  // layout buffers of matrix-free (rank <= 1) sigma calls
  double *d_r1x = nullptr, *d_r1y = nullptr;
  int64_t r1_elems = 0;
  int build_class_cap = 3;  // largest LDS class plans on this context may use
};

struct TermDev {
  int32_t nbody;
  int32_t cl[4];
  int32_t dims[4];
  int32_t pad;
  int64_t ioff;
};

// meta: bit 0 = negative sign; bits 8+4s.. = term position contracted at
// chain stage s (the host picks the order that keeps intermediates small)
struct TaskDev {
  int32_t term;
  int32_t meta;
  int32_t blk[4];
};
__host__ __device__ inline int task_order(int32_t meta, int s) { return (meta >> (8 + 4 * s)) & 0xF; }

struct fcg_ops {
  fcg_ctx *ctx = nullptr;
  int N = 0, S = 0, n_ops = 0;
  std::vector<int32_t> gblock;   // [N][n_ops][S*S][S*S]
  std::vector<int64_t> gb_off;
  std::vector<int32_t> gb_dims;  // [n_blocks][3]
  std::vector<int32_t> nstates;  // [N][S*S] from the "H" blocks
  double *d_gdata = nullptr;
  int64_t n_gdata = 0;
  int64_t *d_gb_off = nullptr;
  int32_t *d_gb_dims = nullptr;
  int32_t block(int c, int op, int code_b, int code_k) const {
    int SS = S * S;
    return gblock[(((int64_t)c * n_ops + op) * SS + code_b) * SS + code_k];
  }
};

struct fcg_ham {
  fcg_ctx *ctx = nullptr;
  int N = 0, n_terms = 0, n_tc = 0;
  std::vector<int32_t> nbody, clusters, ops, parity, dims, tc_start;
  std::vector<int16_t> delta, tc;
  std::vector<int64_t> ioff;
  TermDev *d_terms = nullptr;
  double *d_ints = nullptr;
};

struct fcg_space {
  fcg_ctx *ctx = nullptr;
  int N = 0;
  bool tpsci = false;
  int64_t n_fock = 0, n_box = 0, nvar = 0;
  std::vector<int16_t> fock;                  // [n_fock][N][2]
  std::vector<int32_t> box_fock;              // [n_box]
  std::vector<int16_t> lo, hi, rank;          // [n_box][N] (1-based ranges)
  std::vector<int64_t> foff;                  // [n_box][N] (-1 = identity)
  std::vector<int64_t> var_off;               // [n_box+1]
  std::vector<double> factors;
  int16_t *d_lo = nullptr, *d_hi = nullptr, *d_rank = nullptr;
  int64_t *d_foff = nullptr, *d_var = nullptr;
  int32_t *d_dim = nullptr;
  double *d_factors = nullptr;
};

struct fcg_plan {
  fcg_ctx *ctx = nullptr;
  // contexts other than ctx that launched work reading this plan: their
  // streams are synchronised before the plan's memory is freed
  std::mutex users_mu;
  std::vector<fcg_ctx *> users;
  fcg_space *sp = nullptr;   // bra space (rows of H, sigma)
  fcg_space *ksp = nullptr;  // ket space (columns of H, c); == sp for H within one space
  int64_t n_pair = 0, n_task = 0, n_kel = 0;
  double build_flops = 0.0, build_us = 0.0;
  int precisions = 1;
  // device
  int64_t *d_row_ptr = nullptr;   // [n_box+1] pairs of each bra box
  int32_t *d_pair_B = nullptr, *d_pair_K = nullptr;
  int64_t *d_pair_t0 = nullptr;   // [n_pair+1]
  int64_t *d_pair_koff = nullptr; // [n_pair] panel of the pair's bra box
  int32_t *d_pair_col = nullptr;  // [n_pair] first panel column of the pair
  TaskDev *d_tasks = nullptr;
  int64_t *d_pair_list = nullptr;  // build only: pair indices by LDS class, smallest first
  int64_t n_tiny = 0, n_small = 0, n_mid = 0;
  bool holds_blocks = true;        // false: built with a fused sigma, nothing stored
  int64_t *d_diag_koff = nullptr; // [n_box] panel holding the (B,B) block or -1
  int32_t *d_diag_col = nullptr;  // [n_box] its first panel column
  int64_t *d_row_koff = nullptr;  // [n_box] first block of each bra row
  int64_t *d_col_ptr = nullptr;   // [n_box+1] columns of each bra row
  // apply schedule: [0, n_multi) boxes of several rows, largest panel first;
  // then n_single_slots slots of one-row boxes (-1 = empty slot) laid out so
  // that each XCD's workgroups take a contiguous run of rows
  int32_t *d_order = nullptr;
  int64_t n_multi = 0, n_single_slots = 0;
  int32_t *d_colidx = nullptr;    // [n_cols] column -> variable index
  int64_t n_cols = 0;
  double *d_K64 = nullptr;
  float *d_K32 = nullptr;
  uint32_t *d_Kbf = nullptr;      // hi bf16 << 16 | lo bf16
  uint16_t *d_Kb16 = nullptr;     // bf16 (the hi part alone)
  int64_t reserved_bytes = 0;     // held against the process-wide plan budget
  // matrix-free plan of a rank <= 1 space (fcg_rank1.hip); nullptr for plans
  // that store their Hamiltonian blocks
  struct fcg_rank1 *r1 = nullptr;
  const fcg_ham *r1_ham = nullptr;
  const fcg_ops *r1_ops = nullptr;
  const double *d_thr = nullptr;  // build only: per-ket-variable screening threshold
  // apply variants chosen by timing on the first sigma of each precision and R
  // (tune_mu: one context times them while others wait)
  std::mutex tune_mu;
  bool tuned[FCG_N_PREC][FCG_MAX_R + 1] = {};
  int8_t tuned_vm[FCG_N_PREC][FCG_MAX_R + 1] = {}, tuned_vr[FCG_N_PREC][FCG_MAX_R + 1] = {};
};

// host plan construction (fcg_plan_host.cpp)
// Allocator whose resize() leaves new trivial elements unwritten (default-
// instead of value-initialised): the host plan's large tables are written in
// full by parallel passes, so a serial zero fill first would only cost time.
// Tables of 2 MiB and more are 2 MiB-aligned and marked for transparent huge
// pages (hosts that enable them on request), so first touch faults in 2 MiB
// pages instead of 4 KiB ones.
template <class T>
struct DefaultInitAlloc : std::allocator<T> {
  static constexpr size_t kHuge = (size_t)2 << 20;
  template <class U>
  struct rebind {
    using other = DefaultInitAlloc<U>;
  };
  DefaultInitAlloc() = default;
  template <class U>
  DefaultInitAlloc(const DefaultInitAlloc<U> &) noexcept {}
  T *allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < kHuge) return std::allocator<T>::allocate(n);
    const size_t sz = (bytes + kHuge - 1) / kHuge * kHuge;
    void *p = std::aligned_alloc(kHuge, sz);
    if (!p) throw std::bad_alloc();
    (void)madvise(p, sz, MADV_HUGEPAGE);  // a hint: refused, the table keeps small pages
    return static_cast<T *>(p);
  }
  void deallocate(T *p, size_t n) {
    if (n * sizeof(T) < kHuge) std::allocator<T>::deallocate(p, n);
    else std::free(p);
  }
  template <class U>
  void construct(U *p) noexcept(std::is_nothrow_default_constructible<U>::value) {
    ::new (static_cast<void *>(p)) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a) {
    ::new (static_cast<void *>(p)) U(std::forward<A>(a)...);
  }
};
template <class T>
using RawVec = std::vector<T, DefaultInitAlloc<T>>;

struct HostPlan {
  RawVec<int32_t> pair_B, pair_K, pair_col, colidx;
  RawVec<int64_t> pair_t0, pair_koff;
  std::vector<int64_t> row_ptr, diag_koff, row_koff, col_ptr;
  std::vector<int32_t> diag_col, order;
  int64_t n_multi = 0, n_single_slots = 0;
  RawVec<TaskDev> tasks;
  // pair indices for the block build: pairs within the tiny class's limits
  // first (n_tiny of them; plans that store their blocks only), then those
  // within the small LDS limits (n_small), then the middle limits (n_mid),
  // then the rest, each in pair order
  RawVec<int64_t> pair_list;
  int64_t n_tiny = 0, n_small = 0, n_mid = 0;
  // stored elements: panel layout included (layout), else sum of db * dk
  int64_t n_kel = 0;
  double flops = 0.0;
};
// apply schedule of a plan (see fcg_plan::d_order)
void fcg_apply_order(const std::vector<int64_t> &var_off, const std::vector<int64_t> &col_ptr,
                     std::vector<int32_t> *order, int64_t *n_multi, int64_t *n_single_slots);
// layout = false: pairs, tasks and LDS classes only (a plan whose blocks are
// used once during the build); pair_col, pair_koff, the row panels, colidx,
// the apply order and the diagonal tables are left empty
// class_cap: largest LDS class index a pair may take (3 tiny .. 0 full)
int fcg_build_host_plan(const fcg_ham *h, const fcg_ops *o, const fcg_space *bra,
                        const fcg_space *ket, int nbody, bool layout, int class_cap,
                        HostPlan *out, std::string *err);
int fcg_tpsci_boxes(const fcg_ops *o, int N, int64_t n_fock, const int16_t *fock,
                    const int64_t *fock_start, int64_t n_cfg, const int16_t *cfg,
                    fcg_space *sp, std::string *err);
// every configuration of each listed Fock block, as boxes (first-order spaces)
int fcg_full_space(const fcg_ops *o, int N, int64_t n_fock, const int16_t *fock, fcg_space *sp,
                   std::string *err);
// First-order space of a TPSCI state for PT2: every configuration of the
// Fock blocks F + TransferConfig (F a Fock block of `var`, every cluster
// sector present in ClusterOps), in first-appearance order, as boxes in
// `tsp` (host side only); inspace[x] marks the configurations of `var`;
// hoff/hstride [n_box][N] address the diagonal of op h0_op for each box's
// first state and its stride.
struct Pt2Target {
  std::vector<uint8_t> inspace;
  std::vector<int64_t> hoff;
  std::vector<int32_t> hstride;
};
int fcg_pt2_target(const fcg_ham *h, const fcg_ops *o, const fcg_space *var, int h0_op,
                   fcg_space *tsp, Pt2Target *out, std::string *err);

// Fock pairs of single-configuration spaces (CSR over bra Fock blocks)
int fcg_fock_pairs(const fcg_ham *h, const fcg_space *bra, const fcg_space *ket, int nbody,
                   int S, std::vector<int64_t> *ptr, std::vector<int32_t> *fk,
                   std::vector<int32_t> *g);
// connected Fock pairs (bra block, ket block bra - TransferConfig, group) of
// a space with itself, as CSR over bra blocks on the device; the caller owns
// the three arrays (fcg_dfree)
int fcg_device_fock_pairs(fcg_ctx *ctx, const fcg_ham *h, const fcg_space *sp, int nbody, int S,
                          int64_t **ptr, int32_t **fk, int32_t **g, int64_t *n);
// plan of a space whose boxes are all single configurations, joined and
// built on the device (fcg_sparse.hip)
int fcg_create_sparse_plan(fcg_ctx *ctx, fcg_ham *h, fcg_ops *o, fcg_space *sp, int nbody,
                           int precisions, fcg_plan **out, int64_t max_bytes, int64_t *need);
// device bytes a plan holds while it is alive (blocks in each precision,
// column indices, row tables), and the extra bytes its build needs
int64_t fcg_plan_bytes(int64_t n_kel, int64_t n_cols, int64_t n_box, int precisions);
// reserve a plan's bytes against max_bytes and the process-wide plan budget
// (FCG_ENOMEM, nothing held, when either would be exceeded); release on destroy
int fcg_plan_reserve(fcg_ctx *ctx, int64_t bytes, int64_t max_bytes, int64_t *need);

// Device memory of the library is stream-ordered on the owning context's
// stream (hipMallocAsync / hipFreeAsync): freeing it waits for no other
// context's work, where hipFree would synchronise the whole device.  Memory
// read on other contexts' streams is freed only after those streams are
// synchronised (plans record such users; uploads and plan creation end
// with a stream synchronisation).  fcg_dmalloc returns nullptr on failure
// with ctx->err set.
void *fcg_dmalloc(fcg_ctx *ctx, size_t bytes);
void fcg_dfree(fcg_ctx *ctx, void *p);
// note that ctx launches work reading pl (no-op for the owning context)
void fcg_plan_note_user(fcg_plan *pl, fcg_ctx *ctx);
void fcg_plan_release(int64_t bytes);
// live plans, so that a destroyed context can be removed from their users
void fcg_plan_register(fcg_plan *pl);
void fcg_plan_unregister(fcg_plan *pl);
constexpr int FCG_SPARSE_MAX_BLOCK = 256;  // configurations per Fock block (device join)
// the device join's limits: at most FCG_SPARSE_MAX_BLOCK configurations per
// Fock block (consecutive), state indices below 256
bool fcg_sparse_plan_fits(const fcg_space *sp);

// per-ket-variable screening thresholds on the device (fcg_pt2.hip)
int fcg_screen_thresholds(fcg_ctx *ctx, int64_t n, int R, const double *c, double thresh,
                          double **d_thr);

// shared helpers (fcg_api.hip)
int fcg_space_to_device(fcg_space *sp);
int fcg_form_boxes(fcg_ctx *ctx, fcg_ops *o, fcg_space *sp);
// max_bytes: refuse (FCG_ENOMEM, nothing allocated, *need set) when the
// plan's device footprint would exceed it
// thr (device, [ket n_var] or nullptr): screened build, see BuildArgs::thr
// fused (or nullptr): sigma += H c during the build, no block stored; the
// plan then holds no blocks (for a sigma needed once, as in PT2)
struct FusedSigma {
  int R;
  const double *c;  // device, ket n_var x R, column-major
  double *sigma;    // device, bra n_var x R, column-major, accumulated into
  // device, bra n_var bytes, or nullptr: set to 1 for every bra variable
  // that some term's nonzero element reached after screening (the keys the
  // reference's open matvec inserts)
  uint8_t *hit = nullptr;
  // filled by fcg_create_plan: each pair's products (db x R, row-major by
  // root) at part + yoff[pair], summed per bra box in pair order afterwards
  // (fixed order: the result does not depend on how workgroups are scheduled)
  double *part = nullptr;
  const int64_t *yoff = nullptr;
};
int fcg_create_plan(fcg_ctx *ctx, fcg_ham *h, fcg_ops *o, fcg_space *sp, fcg_space *ksp,
                    int nbody, int precisions, fcg_plan **out, int64_t max_bytes = INT64_MAX,
                    int64_t *need = nullptr, const double *thr = nullptr,
                    const FusedSigma *fused = nullptr);

// matrix-free sigma and diagonal of rank <= 1 plans (fcg_rank1.hip);
// c, sigma: device, n_var x R column-major
int fcg_rank1_sigma(fcg_ctx *ctx, fcg_plan *pl, int R, const double *c, double *sigma,
                    int precision);
int fcg_rank1_diag(fcg_ctx *ctx, fcg_plan *pl, double *hd);
void fcg_rank1_free(fcg_ctx *ctx, fcg_plan *pl);

// kernels (fcg_kernels.hip)
hipError_t fcg_launch_build(const fcg_plan *pl, const fcg_ham *h, const fcg_ops *o,
                            hipStream_t s, const FusedSigma *fused = nullptr);
hipError_t fcg_launch_convert(const fcg_plan *pl, hipStream_t s);
// fused sigma: sigma[bra var, r] += the bra box's pair products in pair order
hipError_t fcg_launch_fused_sum(const fcg_plan *pl, const FusedSigma *f, const int64_t *row_ptr,
                                hipStream_t s);
// vm / vr: apply variants for multi-row and one-row boxes; parts: bit 0
// multi-row boxes, bit 1 one-row boxes
// ci: the interleave buffer (at least R * ket n_var * 8 bytes)
hipError_t fcg_launch_apply(const fcg_plan *pl, int R, const double *c, double *sigma,
                            int precision, hipStream_t s, hipEvent_t before_apply, int vm, int vr,
                            void *ci, int parts = 3);
int fcg_apply_default_variant(int precision);
int fcg_apply_n_variants(int *n_multi, int *n_rows, int *max_r);
hipError_t fcg_launch_diag(const fcg_plan *pl, double *hd, hipStream_t s);
