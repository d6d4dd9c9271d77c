/* fcg.h -- C ABI of the MI355X sigma-build library (libfcg_hip.so).
 *
 * Drop-in boundary for FermiCG's sigma-build hot path.  Julia host code
 * would ccall these entry points from the same methods the reference
 * exposes; the Python package fermicg_amd binds them with ctypes.
 *
 * Conventions
 *   - return codes: 0 ok, negative FCG_E* on error; fcg_last_error(ctx)
 *     returns a message.  Nothing throws or aborts across the ABI.
 *   - host pointers are read during the call only; outputs are
 *     caller-allocated host memory unless the name ends in _device.
 *   - handles are device-resident and owned by the library; free them with
 *     fcg_destroy_*; destroy dependants (plans) before their inputs, and all
 *     handles before their context.
 *   - one fcg_ctx_t per host thread; calls on one ctx are serialised on its
 *     HIP stream.  Device memory is stream-ordered on the context that made
 *     the handle, so destroying a handle waits for no other context's work;
 *     a plan may serve several contexts at once (each context keeps its own
 *     interleaved copy of c), fcg_plan_destroy first synchronises each of
 *     them, and a context destroyed earlier stops being one of them.
 *   - layouts are column-major: Γ blocks (np, m_bra, m_ket), integrals as
 *     the ClusteredTerm ints array, vectors (n_var x R) with element (i,r)
 *     at r*n_var + i.  State indices and ranges are 1-based (Julia).
 *   - sector code of a cluster Fock sector (na, nb) is na*S + nb.
 */
#ifndef FCG_H
#define FCG_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  FCG_OK = 0,
  FCG_EINVAL = -1,         /* bad argument or inconsistent shapes */
  FCG_ENOMEM = -2,         /* device allocation failed */
  FCG_EMISSING_BLOCK = -3, /* a term needs a Γ block that was not uploaded */
  FCG_EOVERFLOW = -4,      /* a size exceeds a library limit */
  FCG_EDEVICE = -5,        /* HIP runtime error */
  FCG_ENOTCONV = -6,       /* Davidson did not converge within max_iter */
  FCG_PENDING = 1          /* fcg_serve_poll: the solve has not finished (not an error) */
};

/* Stored Hamiltonian-block precisions for the sigma apply.  FCG_FP64 and
 * FCG_FP32 accumulate in their own precision; FCG_BF16X2 stores each
 * element as hi + lo bf16 (the FP32 value to ~16 bits) and FCG_BF16 as the
 * hi bf16 alone, both accumulated in FP32.  FCG_FP32_BF16C applies the FP32
 * blocks to the vector c rounded to bf16 (round to nearest even), which
 * halves the bytes each gathered column of c occupies; FP32 accumulation
 * (a plan built with it keeps the FP32 copy). */
enum { FCG_FP64 = 0, FCG_FP32 = 1, FCG_BF16X2 = 2, FCG_BF16 = 3, FCG_FP32_BF16C = 4 };

typedef struct fcg_ctx *fcg_ctx_t;
typedef struct fcg_ops *fcg_ops_t;
typedef struct fcg_ham *fcg_ham_t;
typedef struct fcg_space *fcg_space_t;
typedef struct fcg_plan *fcg_plan_t;

/* Context: device ordinal, its stream and workspace. */
int fcg_create(int device, fcg_ctx_t *out);
void fcg_destroy(fcg_ctx_t ctx);
const char *fcg_last_error(fcg_ctx_t ctx);

/* ClusterOps of all clusters (replaces Vector{ClusterOps} as built by
 * compute_cluster_ops, reference src/build_local_quantities.jl:40-191, and
 * read by contract_matrix_element, src/tpsci_inner.jl:10-216).
 * gblock[c][op][code_bra][code_ket] = block id or -1; block b has dims
 * gb_dims[3b..3b+2] = (np, m_bra, m_ket) and data at gdata[gb_off[b]]. */
int fcg_ops_upload(fcg_ctx_t ctx, int n_clusters, int S, int n_ops, const int32_t *gblock,
                   int64_t n_blocks, const int64_t *gb_off, const int32_t *gb_dims,
                   const double *gdata, int64_t n_gdata, fcg_ops_t *out);
void fcg_ops_destroy(fcg_ops_t ops);

/* ClusteredOperator (replaces ClusteredOperator.trans with ClusteredTerm1B-4B,
 * reference src/type_ClusteredOperator.jl:5-8, src/type_ClusteredTerm.jl:31-67).
 * Terms ordered by TransferConfig group: group g holds terms
 * [tc_start[g], tc_start[g+1]); tc[g] is its (Δnα,Δnβ) per cluster.
 * Per term t: nbody[t]; clusters/ops/parity/dims [4t..4t+3] (0-based
 * clusters, -1 padding); delta[8t..8t+7]; integrals at ints[ioff[t]]. */
int fcg_ham_upload(fcg_ctx_t ctx, int n_clusters, int n_terms, const int32_t *nbody,
                   const int32_t *clusters, const int32_t *ops, const int32_t *parity,
                   const int32_t *dims, const int16_t *delta, const int64_t *ioff,
                   const double *ints, int64_t n_ints, int n_tc, const int16_t *tc,
                   const int32_t *tc_start, fcg_ham_t *out);
void fcg_ham_destroy(fcg_ham_t ham);

/* TPSCIstate basis (replaces TPSCIstate.data keys, reference
 * src/type_TPSCIstate.jl:11-14; flattening order of get_vector :139-149).
 * fock: [n_fock][N][2]; configurations of Fock block f are
 * cfg[fock_start[f] .. fock_start[f+1]) with 1-based int16 state indices
 * cfg[i*N + c].  The library keeps the caller's order. */
int fcg_tpsci_space_upload(fcg_ctx_t ctx, int n_clusters, int64_t n_fock, const int16_t *fock,
                           const int64_t *fock_start, int64_t n_cfg, const int16_t *cfg,
                           fcg_space_t *out);

/* BSTstate basis (replaces BSTstate.data keys and Tucker factors, reference
 * src/type_BSTstate.jl:13-18, src/hosvd.jl:15-19; flattening order of
 * get_vector src/type_BSTstate.jl:499-515).  Block b: Fock block
 * box_fock[b]; per cluster c the TuckerConfig range
 * box_lo[b*N+c]..box_hi[b*N+c] (1-based, inclusive), Tucker rank
 * box_rank[b*N+c] and factor (range x rank, column-major) at
 * factors[box_foff[b*N+c]] or identity when box_foff is -1 (then rank must
 * equal the range length).  Core b occupies variables
 * [var_off, var_off + prod(rank)) in blocks' order, column-major. */
int fcg_bst_space_upload(fcg_ctx_t ctx, int n_clusters, int64_t n_fock, const int16_t *fock,
                         int64_t n_box, const int32_t *box_fock, const int16_t *box_lo,
                         const int16_t *box_hi, const int16_t *box_rank, const int64_t *box_foff,
                         const double *factors, int64_t n_factors, fcg_space_t *out);
int64_t fcg_space_nvar(fcg_space_t sp);
void fcg_space_destroy(fcg_space_t sp);

/* Sigma plan: connects the space with itself through the Hamiltonian and
 * builds every Hamiltonian block in the space's basis once (replaces
 * cache_hamiltonian, src/tucker_inner.jl:90-159, and the term loop of
 * tps_ci_matvec, src/tpsci_outer.jl:462-527).  nbody (1..4) drops terms on
 * more clusters (build_sigma! nbody keyword).  precisions: bitmask of
 * (1<<FCG_FP64)|(1<<FCG_FP32)|(1<<FCG_BF16X2)|(1<<FCG_BF16) copies kept
 * resident (FP64 always; (1<<FCG_FP32_BF16C) keeps the FP32 copy). */
int fcg_plan_create(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp, int nbody,
                    int precisions, fcg_plan_t *out);
/* Hamiltonian between two spaces (rows: bra, columns: ket): sigma_bra =
 * H c_ket, i.e. the open matvec of the reference (open_matvec_thread,
 * src/tpsci_matvec_thread.jl:11-129) restricted to the configurations of
 * `bra`, without screening.  fcg_diagonal and fcg_davidson need a plan from
 * fcg_plan_create. */
int fcg_plan_create_pair(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t bra,
                         fcg_space_t ket, int nbody, int precisions, fcg_plan_t *out);
/* fcg_plan_create with admission by device footprint: when the plan would
 * hold more than max_bytes of device memory it returns FCG_ENOMEM having
 * allocated nothing; *need_bytes receives the footprint either way.  No
 * counterpart in the reference; for the serving loop, which admits solves
 * as HBM frees up. */
int fcg_plan_create_limited(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp,
                            int nbody, int precisions, int64_t max_bytes, int64_t *need_bytes,
                            fcg_plan_t *out);
/* Matrix-free plan of a TPSCI space whose configurations are all of rank
 * <= 1 (every cluster in the lowest state of its sector except at most
 * one): sigma = H c is formed per call from the Γ tensors and integrals,
 * factorised over the term's active clusters, with nothing stored (the
 * reference's tps_ci_matvec, src/tpsci_outer.jl:462-527, restricted to such
 * spaces; for spaces whose stored Hamiltonian does not fit in HBM).
 * precisions: (1<<FCG_FP64) | optionally (1<<FCG_FP32) (FP32 arithmetic on
 * FP32 copies of Γ and the integrals).  fcg_sigma*, fcg_diagonal and
 * fcg_davidson accept the plan.  FCG_EINVAL if a configuration has two
 * excited clusters; FCG_EOVERFLOW beyond 512 rank <= 1 configurations in a
 * Fock block.  No counterpart in the reference beyond tps_ci_matvec. */
int fcg_plan_create_matrix_free(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp,
                                int nbody, int precisions, fcg_plan_t *out);
/* Device bytes held by a plan (blocks, column indices, row tables). */
int fcg_plan_device_bytes(fcg_plan_t plan, int64_t *bytes);
/* Process-wide budget for the device bytes of all live plans (default
 * unlimited): every plan creation reserves its footprint against it before
 * allocating and returns FCG_ENOMEM, having allocated nothing, when it would
 * not fit; destroying a plan releases its bytes.  fcg_plan_budget reports the
 * budget and the bytes held.  No counterpart in the reference; for the
 * serving loop, which admits concurrent solves by HBM footprint. */
int fcg_set_plan_budget(int64_t bytes);
int fcg_plan_budget(int64_t *budget, int64_t *held);
/* The same with the reference's screening of the open matvec
 * (contract_matvec_thread / _collect_significant_thread!,
 * src/tpsci_matvec_thread.jl:279-895): a term's element from ket variable j
 * enters only when its magnitude exceeds thresh / max_r |c[j, r]|, c being
 * the ket state (host, n_ket x R) the open sigma is for.  thresh = 0 keeps
 * every element. */
int fcg_plan_create_pair_screened(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t bra,
                                  fcg_space_t ket, int nbody, int precisions, int R,
                                  const double *c, double thresh, fcg_plan_t *out);
void fcg_plan_destroy(fcg_plan_t plan);
/* Plan statistics: [0] block pairs, [1] (term, block pair) tasks,
 * [2] Hamiltonian-block elements, [3] build flops, [4] build time (us),
 * [5] panel columns (column-index entries read by one sigma). */
int fcg_plan_stats(fcg_plan_t plan, double *out6);

/* sigma = H c (replaces tps_ci_matvec, src/tpsci_outer.jl:462-527, and
 * build_sigma!, src/tucker_inner.jl:165-245).  c, sigma: host (n_var x R). */
int fcg_sigma(fcg_ctx_t ctx, fcg_plan_t plan, int R, const double *c, double *sigma,
              int precision);
/* Same with device-resident FP64 operands (n_var x R column-major). */
int fcg_sigma_device(fcg_ctx_t ctx, fcg_plan_t plan, int R, const double *c_dev,
                     double *sigma_dev, int precision);

/* fcg_sigma_device that allocates nothing.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops).  The interleaved copy of c that
 * fcg_sigma_device keeps in a per-context allocation lives here in a fixed
 * device array of 2^23 values (64 MiB) that each GPU holds until the process
 * ends; R x (ket n_var) above 2^23 returns FCG_EOVERFLOW.  Every call on one
 * GPU shares that array (it is written in full before it is read), so calls
 * that overlap on one GPU -- from any contexts -- race; serialise them.  The
 * first call of a plan for a (precision, R) also times the apply variants
 * (no allocation).  Results equal fcg_sigma_device's bit for bit. */
int fcg_sigma_device_noalloc(fcg_ctx_t ctx, fcg_plan_t plan, int R, const double *c_dev,
                             double *sigma_dev, int precision);

/* Diagonal of H in the space's basis (replaces compute_diagonal,
 * src/tpsci_outer.jl:694-715). */
int fcg_diagonal(fcg_ctx_t ctx, fcg_plan_t plan, double *hd);
/* The same without a plan, for any TPSCI space (replaces compute_diagonal,
 * src/tpsci_outer.jl:694-715, term by term: every term of the zero
 * TransferConfig group contracted at bra = ket = each configuration).  For
 * spaces whose off-diagonal Hamiltonian is not needed, such as the
 * first-order space of compute_qdpt_energy (src/tpsci_pt2_energy.jl:344).
 * hd: host (n_var), in the space's configuration order.  FCG_EINVAL for a
 * space with Tucker factors. */
int fcg_space_diagonal(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp, double *hd);

/* Davidson for the lowest R eigenpairs (replaces tps_ci_davidson,
 * src/tpsci_outer.jl:364-453, and ci_solve, src/tucker_outer.jl:68-215,
 * whose eigensolver is the external BlockDavidson package).  precision
 * FCG_FP64 runs every sigma in FP64; the other precisions run the
 * expansion sigmas in that precision and recompute the Ritz vectors'
 * sigma and residual in FP64 at every restart, when the reduced-precision
 * residual falls below conv_thresh, and every refresh_k iterations if
 * refresh_k > 0; convergence is decided on the FP64 residual.  v0: host (n_var x R) start vectors (orthonormalised
 * inside); evals (R), evecs (n_var x R) host outputs. */
typedef struct {
  double conv_thresh;   /* residual norm threshold (reference conv_thresh) */
  int max_iter;         /* reference max_iter */
  int max_ss_vecs;      /* subspace vectors per root before restart */
  double lindep_thresh; /* drop expansion vectors below this norm */
  int precision;        /* FCG_FP64 .. FCG_FP32_BF16C */
  int refresh_k;        /* extra FP64 refresh period (0: restarts only) */
  int precond;          /* 1: diagonal (Davidson) preconditioner */
  /* host (n_var) diagonal for the preconditioner, or NULL for the diagonal
   * of H; ci_solve passes the H0 (Hcmf) diagonal of the pseudo-canonical
   * Tucker basis (form_1body_operator_diagonal!, src/tucker_pt.jl:284-355) */
  const double *precond_diag;
} fcg_davidson_opts;

typedef struct {
  int iterations;
  int sigma_fp64;       /* number of FP64 sigma vector products */
  int sigma_low;        /* number of reduced-precision sigma vector products */
  int converged;
  double residual[16];
} fcg_davidson_stats;

int fcg_davidson(fcg_ctx_t ctx, fcg_plan_t plan, int R, const double *v0,
                 const fcg_davidson_opts *opts, double *evals, double *evecs,
                 fcg_davidson_stats *stats);

/* Serving loop: independent Davidson solves submitted as they arrive and run
 * with continuous batching on one GPU.  `slots` solver threads, each with its
 * own context (HIP stream), take queued requests oldest first, so solves of
 * different sizes overlap on the device; each builds its sigma plan under
 * the process-wide plan budget (set to hbm_budget while the server runs, see
 * fcg_set_plan_budget) and, when the plan does not fit beside the resident
 * ones, waits until a running solve frees its plan; a plan larger than the
 * whole budget fails its request with FCG_ENOMEM.  Every solve runs
 * fcg_davidson with `davidson` (its precond_diag ignored: the diagonal of
 * H) and frees its plan.  Solves are deterministic, so a request's results
 * are the same alone and in the mix.  No counterpart in the reference, whose
 * Julia host would call tps_ci_davidson (src/tpsci_outer.jl:364) per
 * request; for the served mix of independent solves. */
typedef struct fcg_server *fcg_server_t;
typedef struct {
  int device;
  int slots;                   /* concurrent solves (1..64) */
  int64_t hbm_budget;          /* device bytes all live plans may hold */
  int nbody;                   /* terms on at most this many clusters (1..4) */
  fcg_davidson_opts davidson;  /* options of every solve */
} fcg_serve_opts;
typedef struct {
  double submitted, started, planned, finished;  /* seconds since fcg_serve_start */
  int64_t plan_bytes;                            /* the request's plan footprint */
} fcg_serve_times;
int fcg_serve_start(const fcg_serve_opts *opts, fcg_server_t *out);
/* Queue a solve of the lowest R eigenpairs of the Hamiltonian `ham` over
 * `sp` from the host start vectors v0 (n_var x R, copied).  The handles
 * must be uploaded and stay alive until the ticket is polled done; several
 * requests may share them.  *ticket identifies the request. */
int fcg_serve_submit(fcg_server_t srv, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t sp, int R,
                     const double *v0, int64_t *ticket);
/* FCG_PENDING while the request is queued or running (wait = 0), else the
 * solve's return code with evals (R), evecs (n_var x R), stats and times
 * filled (each nullable) and the ticket released.  wait = 1 blocks until
 * the request finishes. */
int fcg_serve_poll(fcg_server_t srv, int64_t ticket, int wait, double *evals, double *evecs,
                   fcg_davidson_stats *stats, fcg_serve_times *times);
/* Finish every queued solve, join the slots, restore the plan budget and
 * free the server (results not polled are dropped). */
int fcg_serve_stop(fcg_server_t srv);

/* Second-order energy of a TPSCI state over its first-order space (replaces
 * compute_pt2_energy, reference src/tpsci_pt2_energy.jl:10-175, with its
 * per-Fock-block job _pt2_job :178-266 and sum _sum_pt2 :269-282).
 * var: the state's TPSCI space; c: its coefficients (n_var x R, host,
 * orthonormal columns as the reference's orthonormalize! leaves them);
 * e0[r] = <r|H0|r> with H0 the sum of the clusters' local operator h0_op
 * (the reference's H0 string; 1 = "Hcmf", its default).
 * e2[r] = sum_x sigma[x,r]^2 / (e0[r] - H0_diag[x]) over every
 * configuration x of the Fock blocks F_ket + TransferConfig outside the
 * variational space, with sigma = H c screened as the reference's open
 * matvec does at thresh_foi = thresh (see fcg_plan_create_pair_screened;
 * 0 keeps every contribution). */
typedef struct {
  int64_t n_target_fock;  /* Fock blocks of the first-order space (jobs) */
  int64_t n_target_var;   /* configurations in those blocks */
  int64_t n_first_order;  /* ... outside the variational space */
  int64_t block_pairs;    /* Hamiltonian blocks built */
  int64_t block_elems;    /* their elements */
  double seconds;         /* wall time of the call */
  double build_flops;     /* planner flop count of the block build */
  double build_seconds;   /* wall time of the join and block build */
} fcg_pt2_stats;

int fcg_tpsci_pt2(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t var, int R,
                  const double *c, const double *e0, int nbody, int h0_op, double thresh,
                  double *e2, fcg_pt2_stats *stats);

/* compute_pt2_energy over a sparse first-order space (the reference's own
 * formulation, src/tpsci_pt2_energy.jl:10-282 with contract_matvec_thread
 * and its bounds, src/tpsci_matvec_thread.jl:279-1283): contributions are
 * generated per (ket configuration, term) as every bra state combination of
 * the term's active clusters, pruned by upper bounds that never drop an
 * element passing the test |element| > thresh / max_r |c[j, r]|, kept by that
 * test, summed per first-order configuration x in a fixed order, and e2[r] =
 * sum over x outside the variational space of sigma[x,r]^2 /
 * (e0[r] - H0_diag[x]).  Arguments as fcg_tpsci_pt2; it never forms a target
 * Fock block whole, so it serves many clusters, and it accumulates in
 * chunks of whole target Fock blocks, so device memory bounds the largest
 * target block's contributions rather than all of them (2^27 per chunk, or
 * FCG_PT2_CHUNK from the environment).  e2 is the same on every call with
 * the same inputs and chunk size (fixed summation order).  FCG_EOVERFLOW when the
 * configuration keys need more than 2 x 64 bits or a term's shape exceeds
 * the staged contraction (then use fcg_tpsci_pt2). */
typedef struct {
  int64_t n_items;          /* (ket Fock block, term) pairs passing the first bound */
  int64_t n_groups;         /* ket groups (equal active states) contracted */
  int64_t n_contributions;  /* kept (x, ket, term) contributions */
  int64_t n_configs;        /* distinct x reached */
  int64_t n_first_order;    /* ... outside the variational space */
  double chain_flops;       /* flops of the kept contractions */
  double seconds;           /* wall time of the call */
  int64_t n_items_kept;     /* items with a ket passing its own bound */
  int64_t n_kets_kept;      /* (item, ket) pairs passing the ket bound */
  int64_t n_chunks;         /* accumulation chunks (whole target Fock blocks) */
  double contraction_seconds; /* count and fill passes of the contraction */
  double sort_seconds;      /* sorting and summing the contributions */
} fcg_pt2_sparse_stats;

int fcg_tpsci_pt2_sparse(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t var, int R,
                         const double *c, const double *e0, int nbody, int h0_op, double thresh,
                         double *e2, fcg_pt2_sparse_stats *stats);
/* The same with the contraction's arithmetic chosen: FCG_FP64, or FCG_FP32
 * (Γ and integrals rounded to FP32 as read, the chain and each element in
 * FP32; the threshold test, the values x c, the per-configuration sums and
 * the compensated e2 sum in FP64, the bounds from the FP64 data): the
 * mixed-precision PT2, whose |e2(FP32) - e2(FP64)| a caller reports
 * against this FP64 engine.  No counterpart in the reference (FP64 only). */
int fcg_tpsci_pt2_sparse_ex(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t var, int R,
                            const double *c, const double *e0, int nbody, int h0_op,
                            double thresh, int precision, double *e2,
                            fcg_pt2_sparse_stats *stats);

/* Open matvec of a TPSCI state onto the configurations of `bra` (replaces
 * open_matvec_thread, src/tpsci_matvec_thread.jl:11-129, for the target
 * configurations the caller lists; prescreen does not change the result).
 * c: the state (n_ket x R, host); a term's element from ket configuration j
 * enters when |element| > thresh / max_r |c[j, r]|
 * (_collect_significant_thread!, :830-895).  sigma (n_bra x R) and hit
 * (n_bra bytes, may be NULL) are host outputs: hit[x] = 1 when some element
 * reached x (the keys the reference inserts, kept even when their sum
 * cancels).  Nothing is stored on the device after the call. */
int fcg_open_sigma(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t bra, fcg_space_t ket,
                   int nbody, int R, const double *c, double thresh, double *sigma,
                   uint8_t *hit);

/* Open matvec without a target list (replaces open_matvec_thread,
 * src/tpsci_matvec_thread.jl:11-129, generated as the reference generates it:
 * per ket configuration and term, _open_matvec_thread_job and
 * contract_matvec_thread, :233-818, with its bounds, :897-1283): the
 * configurations reached, in first-order spaces too large to list whole.
 * Same element test as fcg_open_sigma; a configuration is returned when some
 * element reached it (the reference's inserted keys).  The engine of
 * fcg_tpsci_pt2_sparse, its sums kept per configuration instead of reduced to
 * e2.  *out: a result handle (read it with fcg_open_result_size / _copy,
 * free it with fcg_open_result_destroy). */
typedef struct fcg_open_result *fcg_open_result_t;
int fcg_open_sigma_sparse(fcg_ctx_t ctx, fcg_ham_t ham, fcg_ops_t ops, fcg_space_t ket,
                          int nbody, int R, const double *c, double thresh,
                          fcg_open_result_t *out);
/* Number of configurations in the result. */
int fcg_open_result_size(fcg_open_result_t res, int64_t *n);
/* fock (n x N x 2: alpha, beta electrons per cluster), cfg (n x N, 1-based
 * cluster states) and sigma (column-major n x R) of the n configurations,
 * in the engine's deterministic order (chunks of target Fock blocks, within
 * a chunk by states); the caller groups them by Fock configuration. */
int fcg_open_result_copy(fcg_open_result_t res, int16_t *fock, int16_t *cfg, double *sigma);
void fcg_open_result_destroy(fcg_open_result_t res);

/* Apply variants the plan chose for a precision and root count (-1 before
 * its first sigma there).  The first sigma of each (precision, R) times each
 * variant of the apply kernels (workgroup shape, loads in flight) on the
 * caller's operands and keeps the fastest; all variants give identical
 * results.  No counterpart in the reference (its tuning is BLAS's). */
int fcg_plan_apply_variant(fcg_plan_t plan, int precision, int R, int *multi, int *rows);

/* Pin the apply variants for a precision and root count instead of timing
 * them (multi in [0, n_multi), rows in [0, n_rows) of
 * fcg_apply_variant_counts for R <= max_r; above max_r only the precision's
 * default multi-row variant (1 for FP64, 0 otherwise) and rows 0 exist).
 * FCG_EINVAL for any other pair.  No counterpart in the reference. */
int fcg_plan_set_apply_variant(fcg_plan_t plan, int precision, int R, int multi, int rows);

/* Number of multi-row and one-row apply variants, and the largest root
 * count that has all of them.  No counterpart in the reference. */
int fcg_apply_variant_counts(int *n_multi, int *n_rows, int *max_r);

/* LDS classes of the block build: each (bra box, ket box) pair of a plan is
 * built by the smallest class that holds it (3 tiny, 2 small, 1 middle,
 * 0 full).  fcg_set_build_class_cap(ctx, cap), cap in [0, 3], makes plans
 * created on ctx afterwards use at most class `cap` (0 builds every pair in
 * the full class); 3 is the default.  Every class builds the same blocks
 * bit for bit, so this is for testing and measurement.  No counterpart in
 * the reference. */
int fcg_set_build_class_cap(fcg_ctx_t ctx, int cap);

/* Device time of the last fcg_sigma* call's apply kernel (milliseconds,
 * HIP events on the context's stream) and its launch count. */
int fcg_last_sigma_time(fcg_ctx_t ctx, double *ms, int *launches);

/* Stream-ordered timing on the context's stream: record event `slot`
 * (0..7), then read the time between two recorded slots (waits for the
 * later one).  No counterpart in the reference (its timers are host-side
 * @time / TimerOutputs sections, e.g. src/tucker_outer.jl:135,185). */
int fcg_event_record(fcg_ctx_t ctx, int slot);
int fcg_event_elapsed(fcg_ctx_t ctx, int slot_a, int slot_b, double *ms);
int fcg_synchronize(fcg_ctx_t ctx);

#ifdef __cplusplus
}
#endif
#endif
