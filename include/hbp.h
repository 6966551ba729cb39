/*
 * hbp.h - C-ABI of the MI355X-native belief-propagation PEPS backend.
 *
 * Two layers, both with plain pointers and sizes only:
 *
 *  1. Tensor layer: the Tensor Computing Interface (TCI) operations that
 *     tnbp's algorithms issue, on device-resident column-major tensors.  Each
 *     entry point names the min-tci function it replaces
 *     (external/min-tci/include/tci/gqten/...); a C++ TCI header binds these
 *     so that include/tnbp and the apps compile unchanged.
 *
 *  2. PEPS layer: fused entry points behind tnbp's algorithm functions
 *     (include/tnbp/ptns/function/{init,apply,bp,truncation,measure}.h) that
 *     keep the whole state (site tensors V, messages E) resident in HBM and
 *     run each step as a few hand-written CDNA4 kernels.
 *
 * Conventions (same as the reference):
 *  - storage is column-major, first index fastest (min-tci constructor.h:109-110);
 *  - complex numbers are interleaved (re, im) pairs of double or float;
 *  - messages E[m] (m = position in the sorted edge list) run first -> second,
 *    E[m + n_e] second -> first (tnbp bp.h:170-174); index 0 = ket, 1 = bra;
 *  - a site tensor has its bonds in ascending global edge order, then the
 *    physical index (graph.h:40-55, init.h:57-60).
 *
 * Every call returns an hbp_status; on failure hbp_last_error() holds a
 * message.  Calls on one context are issued in order on one HIP stream;
 * calls that return host values synchronise that stream.  A context must not
 * be used from two host threads at once.
 */
#ifndef HBP_H
#define HBP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBP_MAX_ORDER 16

typedef int hbp_status;
enum {
  HBP_OK = 0,
  HBP_ERR_INVALID_ARGUMENT = 1, /* std::invalid_argument in min-tci */
  HBP_ERR_LOGIC = 2,            /* std::logic_error */
  HBP_ERR_BAD_ALLOC = 3,        /* std::bad_alloc */
  HBP_ERR_OUT_OF_RANGE = 4,     /* std::out_of_range */
  HBP_ERR_HIP = 5,              /* HIP runtime failure */
  HBP_ERR_UNSUPPORTED = 6       /* shape or option outside what the kernels take */
};

enum { HBP_C128 = 0, HBP_F64 = 1, HBP_C64 = 2, HBP_F32 = 3 };

/* Arithmetic of the PEPS layer.  FP64: everything in double.  MIXED: site
 * tensors, contractions and gate applies in float on MFMA with float
 * accumulation; messages, normalisation, residual, truncation and <O> in double. */
enum { HBP_PREC_FP64 = 0, HBP_PREC_MIXED = 1 };

typedef struct hbp_ctx hbp_ctx;
typedef struct hbp_tensor hbp_tensor;
typedef struct hbp_peps hbp_peps;

const char *hbp_last_error(void);
const char *hbp_version(void);

/* ---- context: tci::create_context (min-tci gqten/traits.h:25-27) ---- */
hbp_status hbp_ctx_create(int device, hbp_ctx **out);
hbp_status hbp_ctx_destroy(hbp_ctx *ctx);
hbp_status hbp_ctx_synchronize(hbp_ctx *ctx);
/* The HIP stream all work of this context runs on (a hipStream_t). */
hbp_status hbp_ctx_stream(hbp_ctx *ctx, void **stream);
/* Bytes of live allocations of the context (and their peak), each counted at
 * its size class (a power of two from 256 B below 32 MiB, 2 MiB steps above).
 * Freed blocks kept by the context for reuse do not count. */
hbp_status hbp_ctx_memory(hbp_ctx *ctx, size_t *in_use, size_t *peak);
/* Return the context's unused device memory (kept blocks and the pool's
 * reserve) to the device (after a job ends) and reset the peak to the
 * bytes still in use.  No reference counterpart. */
hbp_status hbp_ctx_trim(hbp_ctx *ctx);
/* Kernel launch variants that compute bitwise the same result (rows per
 * wave, loads in flight, occupancy) are timed on the first launch of each
 * shape class and the fastest is kept for the process (HBP_AUTOTUNE=0 in the
 * environment turns this off).  Writes the choices so far, one line per shape
 * class, NUL-terminated, into buf (len bytes); needed receives the size the
 * full report takes.  No reference counterpart. */
hbp_status hbp_autotune_report(char *buf, size_t len, size_t *needed);
/* HIP graphs of hbp_peps_bp_run_batched on this context so far: graphs
 * recorded, replays, and recordings dropped because the iteration needed
 * something a graph cannot hold (the loop then ran without graphs).  No
 * reference counterpart. */
hbp_status hbp_ctx_graph_stats(hbp_ctx *ctx, int64_t *recorded, int64_t *replays,
                               int64_t *dropped);
/* Free and total device memory of a device (hipMemGetInfo). */
hbp_status hbp_device_memory(int device, size_t *free_bytes, size_t *total_bytes);
/* Kernel timing (no reference counterpart; the reference times nothing but
 * one SVD, apps/tci_test/main.cc:39-48).  While enabled, each launch of the
 * hot kernels carries a pair of HIP events in its own dispatch on the context
 * stream (no extra queue packets).  profile_read synchronises, returns launch
 * count, summed kernel duration, and the algorithmic flops and bytes of all
 * timed launches of one kind, and forgets them; disabling drops records not
 * read yet (their events keep the launches' memory alive).  Kinds (the main kernel of
 * each; split-K reductions are not included): HBP_PROF_MODE_PRODUCT,
 * HBP_PROF_MESSAGE (message partials), HBP_PROF_PRODUCT_MESSAGE (mode product
 * merged into the message kernel). */
enum { HBP_PROF_MODE_PRODUCT = 0, HBP_PROF_MESSAGE = 1, HBP_PROF_PRODUCT_MESSAGE = 2 };
hbp_status hbp_ctx_profile(hbp_ctx *ctx, int enable);
hbp_status hbp_ctx_profile_read(hbp_ctx *ctx, int kind, int64_t *count, double *total_ms,
                                double *flops, double *bytes);

/* ==================================================== tensor layer ==== */
/* Zero tensor of the given shape (tci::zeros / gqten tensor(shape)). */
hbp_status hbp_tensor_create(hbp_ctx *ctx, int dtype, int order, const int32_t *shape,
                             hbp_tensor **out);
hbp_status hbp_tensor_destroy(hbp_ctx *ctx, hbp_tensor *t);
/* tci::order / tci::shape / tci::size (min-tci getter.h). shape holds HBP_MAX_ORDER. */
hbp_status hbp_tensor_info(const hbp_tensor *t, int *dtype, int *order, int32_t *shape,
                           size_t *size);
/* Host <-> device copies of the column-major element array
 * (tci::assign_from_range constructor.h:114-163 / tci::to_range). */
hbp_status hbp_tensor_upload(hbp_ctx *ctx, hbp_tensor *t, const void *host);
hbp_status hbp_tensor_download(hbp_ctx *ctx, const hbp_tensor *t, void *host);
/* tci::copy (constructor.h:192-204): deep copy, dst takes src's shape. */
hbp_status hbp_tensor_copy(hbp_ctx *ctx, const hbp_tensor *src, hbp_tensor *dst);
/* tci::reshape (manipulator.h:37-55): metadata only, same size required. */
hbp_status hbp_tensor_reshape(hbp_ctx *ctx, hbp_tensor *t, int order, const int32_t *shape);
/* tci::transpose (manipulator.h:61-79): axis i of out = axis new_order[i] of in. */
hbp_status hbp_transpose(hbp_ctx *ctx, const hbp_tensor *in, const int32_t *new_order,
                         hbp_tensor *out);
/* tci::cplx_conj (manipulator.h:85-101), in place. */
hbp_status hbp_conj(hbp_ctx *ctx, hbp_tensor *t);
/* tci::fill (constructor.h:82-104). */
hbp_status hbp_fill(hbp_ctx *ctx, hbp_tensor *t, double re, double im);
/* tci::contract (min-tci linalg.h:311-323): labels shared by a and b are
 * summed, survivors are laid out in the order of lc (label rules of
 * GqtenCntLabelHelper linalg.h:205-305).  c may alias a or b. */
hbp_status hbp_contract(hbp_ctx *ctx, const hbp_tensor *a, const int32_t *la,
                        const hbp_tensor *b, const int32_t *lb, hbp_tensor *c,
                        const int32_t *lc, int nc);
/* tci::norm (linalg.h:67-72): Frobenius norm. */
hbp_status hbp_norm(hbp_ctx *ctx, const hbp_tensor *t, double *out);
/* tci::normalize (linalg.h:77-84): t /= ||t||_F, returns the norm. */
hbp_status hbp_normalize(hbp_ctx *ctx, hbp_tensor *t, double *norm_out);
/* tci::scale (linalg.h:103-110): t *= (re + i im). */
hbp_status hbp_scale(hbp_ctx *ctx, hbp_tensor *t, double re, double im);
/* tci::linear_combine (linalg.h:468-493): out = sum_i coefs[i] * ins[i];
 * coefs holds n complex numbers as (re, im) pairs. */
hbp_status hbp_linear_combine(hbp_ctx *ctx, int n, const hbp_tensor *const *ins,
                              const double *coefs, hbp_tensor *out);
/* tci::trace (linalg.h:142-149): trace over bond pairs (pairs[2k], pairs[2k+1]). */
hbp_status hbp_trace(hbp_ctx *ctx, const hbp_tensor *in, int npairs, const int32_t *pairs,
                     hbp_tensor *out);
/* tci::diag (linalg.h:49-62): vector of n -> n x n diagonal matrix; a real
 * vector gives a complex matrix when out is complex (tci::to_cplx + diag). */
hbp_status hbp_diag(hbp_ctx *ctx, const hbp_tensor *vec, hbp_tensor *out);
/* tci::get_elem (getter.h:50-80); an order-0 tensor returns its scalar. */
hbp_status hbp_get_elem(hbp_ctx *ctx, const hbp_tensor *t, const int32_t *coors, double *re,
                        double *im);
/* tci::svd (linalg.h:500-512): thin SVD of the matricisation with nrows row
 * bonds; u (rows..., k), s real (k), vdag (k, cols...), k = min(m, n),
 * singular values descending.  Batched one-sided Jacobi in FP64 on the GPU. */
hbp_status hbp_svd(hbp_ctx *ctx, const hbp_tensor *a, int nrows, hbp_tensor *u, hbp_tensor *s,
                   hbp_tensor *vdag);
/* tci::trunc_svd(..., chi_max, s_min) (linalg.h:552-569): keeps the singular
 * values > s_min, at most chi_max and at least one; trunc_err = discarded
 * s^2 / total s^2. */
hbp_status hbp_trunc_svd(hbp_ctx *ctx, const hbp_tensor *a, int nrows, int chi_max,
                         double s_min, hbp_tensor *u, hbp_tensor *s, hbp_tensor *vdag,
                         double *trunc_err);

/* Allocation-free SVD for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference: it is hbp_svd's batched
 * one-sided Jacobi kernel on caller-owned device buffers, with its work
 * arrays in a fixed device scratch array of the library (16 MiB per device,
 * held until the process ends) instead of context allocations.
 * a: m x n complex128 column-major (device); u: m x n, s: n doubles, vdag:
 * n x n, all device memory of the caller; singular values descending, the
 * same bits as hbp_svd on the same matrix.  Needs m >= n, n <= 512 and
 * (m n + n^2) * 16 bytes <= 16 MiB - 256; larger sizes are rejected with
 * HBP_ERR_UNSUPPORTED.  Asynchronous on the context's stream (results are
 * ready after hbp_ctx_synchronize or later work on that stream).  All calls
 * on one device share its scratch, which keeps the last call's contents:
 * calls that overlap on one device (two contexts' streams, two threads)
 * race. */
hbp_status hbp_svd_noalloc(hbp_ctx *ctx, const void *a, int m, int n, void *u, double *s,
                           void *vdag);
/* Device address of a tensor's column-major element array, valid until the
 * tensor is resized or destroyed: lets callers hand the library's tensors
 * to their own kernels or to hbp_svd_noalloc.  No reference counterpart. */
hbp_status hbp_tensor_data(const hbp_tensor *t, void **ptr);
/* A value that changes with every write to the tensor through this C-ABI
 * (process-wide unique), so that a caller can tell whether a tensor changed
 * since it last looked (the fused TCI overloads cache PEPS state by it).  No
 * reference counterpart. */
hbp_status hbp_tensor_version(const hbp_tensor *t, uint64_t *version);

/* ====================================================== PEPS layer ==== */
/* tnbp::InitTensorProductState (init.h:17-95) on one rank: |0...0> with all
 * bonds of dimension 1 and 1x1 messages equal to 1.  edges holds n_edges
 * (first, second) pairs; sites are the sorted distinct endpoints.
 * precision is HBP_PREC_FP64 or HBP_PREC_MIXED. */
hbp_status hbp_peps_create(hbp_ctx *ctx, int n_edges, const int32_t *edges,
                           const int32_t *phys_dims, int precision, hbp_peps **out);
hbp_status hbp_peps_destroy(hbp_ctx *ctx, hbp_peps *p);
/* Execution strategy (no reference counterpart; both give the reference's
 * results).  FUSED: each message is its mode products plus one message
 * contraction, normalised once; the BP loop reuses the condition's messages
 * as the next sweep.  PLAIN: the reference's op sequence one TCI call at a
 * time - a copy of V per message, a norm and a scale pass after every bond,
 * a conjugated copy of V, and a separate condition sweep per BP iteration
 * (bp.h:128-175, :290-350, main.cc:121-159). */
enum { HBP_MODE_FUSED = 0, HBP_MODE_PLAIN = 1 };
hbp_status hbp_peps_set_mode(hbp_peps *p, int mode);
/* Mixed precision only: the BP loop also stops once the residual is below
 * this floor (default 1e-6), since FP32 messages cannot reach FP64 tolerances. */
hbp_status hbp_peps_set_residual_floor(hbp_peps *p, double floor);
/* FUSED mode: BP update schedule.  JACOBI (default) is the reference's: every
 * message of a sweep from the previous sweep's messages (bp.h:65-261, E <- F
 * after all layers, main.cc:138-141).  CHECKERBOARD, on a lattice whose graph
 * is bipartite (square lattices, chains, trees; others fall back to JACOBI),
 * first updates the messages leaving the sites of one colour, then those
 * leaving the other colour from the new ones.  A message leaving a colour-c
 * site depends only on messages leaving the other colour, so Jacobi sweeps
 * advance two independent sequences one step each, and a checkerboard sweep
 * (the same work) advances one of them two steps: about half the sweeps to
 * the same fixed point.  Iteration counts then differ from the reference's;
 * converged results agree within the BP tolerance.  No reference counterpart. */
/* FUSED mode, merged sweep: whether the sites whose target pair sits behind
 * other bonds are kept rearranged (pair first) for the whole BP loop, about
 * one more absorbed lattice of memory.  -1 (default): when it fits with
 * max(16 GiB, 15 %) of the device to spare; 0: never (each sweep rearranges
 * into a work buffer); 1: always (for callers that have budgeted the memory,
 * such as the serving layer's admission).  No reference counterpart. */
hbp_status hbp_peps_set_pair_cache(hbp_peps *p, int policy);
enum { HBP_SCHEDULE_JACOBI = 0, HBP_SCHEDULE_CHECKERBOARD = 1 };
hbp_status hbp_peps_set_schedule(hbp_peps *p, int schedule);
/* FUSED mode: Anderson mixing of the BP loop's message vector (no reference
 * counterpart; default depth 0, off).  Each iteration still computes
 * G = BP(E) and the residual of E as bp.h:266-355 does and stops on the same
 * test, keeping E (main.cc:138-158); once the residual is below `start` the
 * next E is the normalised combination sum_i a_i G_i of the last depth + 1
 * sweep outputs with the real coefficients (sum 1) that minimise the norm of
 * the same combination of their residual vectors G_i - E_i.  A fixed point of
 * the plain iteration is one of the mixed one, so converged results agree
 * within the BP tolerance; the number of sweeps falls.  Applies to whole BP
 * loops run by one call (hbp_peps_bp_run, hbp_peps_bp_iterate); the window
 * does not outlive the call.  depth <= 8. */
hbp_status hbp_peps_set_accel(hbp_peps *p, int depth, double start);
/* FUSED mode only (default on): fix each message's free scalar phase, trace
 * real and positive, and compare messages modulo that phase in the residual.
 * bp.h:321-350 compares phases too; rounding seeds them and each sweep
 * multiplies them along non-backtracking walks (growth ~ z - 1 per sweep), so
 * on large lattices that residual stops falling near 1e-10 and then rises to
 * ~4/pi.  Expectation values do not depend on the phases.  No reference
 * counterpart; PLAIN mode keeps the reference's residual. */
hbp_status hbp_peps_set_phase_gauge(hbp_peps *p, int on);
/* FUSED mode: how AbsorbTPO stores the absorbed tensors.  LAZY keeps each
 * chi-sized V_s with its operator O_s and materialises the absorbed tensor
 * per site when a sweep needs it (the absorbed 32x32 lattice at D = 64 is
 * 550 GB in FP64); EAGER stores the absorbed lattice (apply.h:52-66); AUTO
 * (default) is EAGER while the absorbed lattice takes at most half of the
 * free device memory.  PLAIN mode always absorbs eagerly.  No reference
 * counterpart. */
enum { HBP_ABSORB_AUTO = 0, HBP_ABSORB_LAZY = 1, HBP_ABSORB_EAGER = 2 };
hbp_status hbp_peps_set_absorb_mode(hbp_peps *p, int mode);
/* FUSED mode: how a BP sweep is launched.  PER_SITE issues each site's mode
 * products and messages as separate launches (efficient once bonds reach
 * ~64); MERGED is PER_SITE with the last mode product of each outgoing
 * message computed inside the message kernel (bonds in (32, 64] on both of
 * the two bonds involved; other messages as PER_SITE), which saves writing
 * and re-reading that product, and with every resident site whose bonds are
 * all <= 16 (degree <= 4) in one launch of the small-site kernel per sweep
 * (one workgroup per outgoing message, FP64 sums).  AUTO (default) is MERGED (which
 * falls back to PER_SITE launches wherever bonds are outside (32, 64]).  GROUPED (one
 * launch per sweep level for all sites) is reserved and rejected with
 * HBP_ERR_UNSUPPORTED: its first GPU run left the device unusable, the cause
 * was not found and its code was taken out; hbp_peps_bp_run_batched covers
 * the launch-bound small bonds instead.  No reference counterpart. */
enum { HBP_SWEEP_AUTO = 0, HBP_SWEEP_PER_SITE = 1, HBP_SWEEP_GROUPED = 2, HBP_SWEEP_MERGED = 3 };
hbp_status hbp_peps_set_sweep_kernel(hbp_peps *p, int kind);
hbp_status hbp_peps_num_sites(const hbp_peps *p, int *n_sites, int *n_edges);
/* Shape of site tensor s (order = degree + 1) and of message m (2 x D). */
hbp_status hbp_peps_site_shape(const hbp_peps *p, int site, int *order, int32_t *shape);
/* Download / upload a site tensor or message as complex128 column-major
 * (site index = position in the sorted site list; message index 0..2n_e-1). */
hbp_status hbp_peps_get_site(hbp_ctx *ctx, hbp_peps *p, int site, double *host);
hbp_status hbp_peps_set_site(hbp_ctx *ctx, hbp_peps *p, int site, int order,
                             const int32_t *shape, const double *host);
/* Replace every site by zeros of bond dims bond_dims[edge] (and physical dims
 * phys_dims[site] when not null) and every message by a zero matrix, dropping
 * pending absorptions: the first step of restoring a checkpoint, after which
 * set_site / set_message fill in the saved state (the restart of
 * apps/estimator_restart reads V and E back in the same way, iotps.h:37-105). */
hbp_status hbp_peps_set_bond_dims(hbp_ctx *ctx, hbp_peps *p, const int32_t *bond_dims,
                                  const int32_t *phys_dims);
hbp_status hbp_peps_get_message(hbp_ctx *ctx, hbp_peps *p, int m, int32_t *dim, double *host);
hbp_status hbp_peps_set_message(hbp_ctx *ctx, hbp_peps *p, int m, int32_t dim,
                                const double *host);

/* Site tensors and messages held in a caller's tensors (the C++ TCI binding's
 * fused overloads of tnbp's algorithm functions use these to run the PEPS
 * layer on the drivers' own V and E vectors).  No reference counterpart.
 * bind_sites: site s reads the caller's tensor sites[s] (bonds in ascending
 * edge order, then the physical index; element type of the PEPS precision)
 * without a copy; the PEPS never writes or frees a bound tensor, and a bound
 * tensor must outlive its binding (until the site is replaced or bound again).
 * Bond dims are taken from the tensors; messages are re-laid out (zeros)
 * when they change.
 * load_messages: copies the 2 n_e complex128 D x D message tensors into the
 * PEPS (msgs[m] = E[m] in the reference's numbering, bp.h:170-174).
 * store_messages: copies messages which[i] (or i when which is NULL) into
 * out[i], resizing them.
 * store_sites: writes every site tensor (absorbed form where pending) into
 * out[s] as complex128, resizing it; in FP64 the site is then bound to out[s]. */
hbp_status hbp_peps_bind_sites(hbp_ctx *ctx, hbp_peps *p, const hbp_tensor *const *sites);
hbp_status hbp_peps_load_messages(hbp_ctx *ctx, hbp_peps *p, const hbp_tensor *const *msgs);
hbp_status hbp_peps_store_messages(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *which,
                                   hbp_tensor *const *out);
hbp_status hbp_peps_store_sites(hbp_ctx *ctx, hbp_peps *p, hbp_tensor *const *out);

/* tnbp::AbsorbTPO (apply.h:14-177): per site O_s(t_1..t_z, d_in, d_out) given
 * as complex128 column-major host arrays (ops[s] with shape tpo_shapes[s*8..]),
 * V'[(v o)...,p'] = sum_p O[o..,p,p'] V[v..,p]; messages become E (x) I_t. */
hbp_status hbp_peps_absorb_tpo(hbp_ctx *ctx, hbp_peps *p, const double *const *ops,
                               const int32_t *tpo_shapes);
/* hbp_peps_absorb_tpo for n PEPS of one context at once (no reference
 * counterpart): ops[i] and tpo_shapes[i] are PEPS i's arguments.  Every eager
 * PEPS's operators go up in one copy, all their sites up to 16 M elements are
 * absorbed in one launch per precision, every message expansion in one
 * launch; each PEPS's result is bitwise the one hbp_peps_absorb_tpo gives it. */
hbp_status hbp_peps_absorb_tpo_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps,
                                       const double *const *const *ops,
                                       const int32_t *const *tpo_shapes);

/* tnbp::BeliefPropagation over all layers followed by E <- F
 * (bp.h:65-261, apps/kicked_ising_lattice/main.cc:128-141). */
hbp_status hbp_peps_bp_sweep(hbp_ctx *ctx, hbp_peps *p);
/* tnbp::BeliefPropagationCondition (bp.h:266-355): mean Frobenius distance
 * of recomputed and stored normalised messages. */
hbp_status hbp_peps_bp_condition(hbp_ctx *ctx, hbp_peps *p, double *residual);
/* The driver's BP loop (main.cc:121-159): repeat sweep + condition until
 * residual < tol or max_iter.  The condition's recomputed messages are the
 * next sweep's messages, so each iteration after the first costs one sweep.
 * residuals (may be NULL) receives max_iter values; iters the count run. */
hbp_status hbp_peps_bp_run(hbp_ctx *ctx, hbp_peps *p, int max_iter, double tol, int *iters,
                           double *residuals);
/* hbp_peps_bp_run for n PEPS of one context at once (no reference
 * counterpart; each PEPS gets the iterations, residuals and state that
 * hbp_peps_bp_run gives it alone).  The PEPS in fused mode with the Jacobi
 * schedule and no mixing iterate in lockstep, one BP iteration of all of
 * them per replay of a HIP graph: those with site degree <= 4 and every
 * bond <= 16 (HBP_BATCH_SMALL) in one persistent launch per precision over
 * every site of every such PEPS (one workgroup per outgoing message and
 * stride, launch grid picked by the autotuner), a finalisation and a residual launch; the
 * others (HBP_BATCH_GRAPH) with their own fused sweeps (per-site launches,
 * variants picked by the autotuner), bitwise as hbp_peps_bp_run.  The first
 * iteration runs unrecorded; the graph of an iteration is recorded once per
 * message-set parity and replayed until a PEPS leaves (it converged or
 * reached max_iter).  The other PEPS (plain mode, checkerboard schedule,
 * mixing) run hbp_peps_bp_run one after another.  residuals (may be NULL):
 * n rows of max_iter values; iters, converged, batched (each may be NULL):
 * n values, batched[i] = HBP_BATCH_SMALL, HBP_BATCH_GRAPH or 0 (on its
 * own). */
enum { HBP_BATCH_ALONE = 0, HBP_BATCH_SMALL = 1, HBP_BATCH_GRAPH = 2 };
hbp_status hbp_peps_bp_run_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps, int max_iter,
                                   double tol, int *iters, double *residuals, int *converged,
                                   int *batched);
/* The same loop in pieces, for callers that interleave work or report
 * progress: up to n iterations; `first` starts a loop (main.cc:121), `last`
 * marks the piece that ends at max_bp_iterations.  converged = 1 once the
 * residual fell below tol.  Consecutive pieces give the results of one call
 * of hbp_peps_bp_run. */
hbp_status hbp_peps_bp_iterate(hbp_ctx *ctx, hbp_peps *p, int n, double tol, int first,
                               int last, int *iters, double *residuals, int *converged);
/* tnbp::Truncation (truncation.h:17-222 with SquareRootAndInverse root.h:65-148):
 * bond_dims / trunc_errs (may be NULL) receive one value per edge. */
hbp_status hbp_peps_truncate(hbp_ctx *ctx, hbp_peps *p, int chi_max, double sv_min,
                             int32_t *bond_dims, double *trunc_errs);
/* hbp_peps_truncate for n PEPS of one context at once (no reference
 * counterpart): every message's SVD, the roots, T = Ra^T Rb, its SVD and the
 * projectors of all n PEPS in one launch each (the SVDs in two when some PEPS
 * has bonds above 64 and others do not), the projectors of every site
 * with bonds <= 16 in one launch per precision, the other sites site by site,
 * all new messages in one launch.  PEPS i keeps chi_max[i] values above
 * sv_min[i]; its bond dims and truncation errors are written at
 * bond_dims / trunc_errs + (n_edges of PEPS 0 .. i-1) (each may be NULL).
 * Each PEPS's result is bitwise the one hbp_peps_truncate gives it alone. */
hbp_status hbp_peps_truncate_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps,
                                     const int32_t *chi_max, const double *sv_min,
                                     int32_t *bond_dims, double *trunc_errs);
/* tnbp::Measure for single sites (measure.h:14-94): ops are n complex 2x2
 * (d x d) operators O(in, out) column-major; out receives n complex values. */
hbp_status hbp_peps_measure_sites(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *sites,
                                  const double *ops, double *out);
/* tnbp::Measure for nearest-neighbour pairs (measure.h:99-300): ops are n
 * complex (d,d,d,d) operators indexed (a_in, b_in, a_out, b_out). */
hbp_status hbp_peps_measure_edges(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *pairs,
                                  const double *ops, double *out);
/* Product operators O_1 ... O_k on simple paths of neighbouring sites (k up to
 * 63), under the BP environment of the path: an extension of tnbp::Measure
 * (measure.h:14-300 stops at two sites; paulitotpo.h:39-153 drops weight >= 3).
 * Path t is sites[offsets[t] .. offsets[t+1]) (site labels, consecutive sites
 * adjacent, none repeated); ops holds one complex (d,d) operator (in, out) per
 * path site in the same order.  out: n complex values <O_1 ... O_k>.
 * k = 1 equals hbp_peps_measure_sites; k = 2 equals hbp_peps_measure_edges
 * with the product operator. */
hbp_status hbp_peps_measure_paths(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *offsets,
                                  const int32_t *sites, const double *ops, double *out);

#ifdef __cplusplus
}
#endif
#endif /* HBP_H */
