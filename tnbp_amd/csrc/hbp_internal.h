// Internal declarations shared by the hbp HIP sources (not part of the C-ABI).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "hbp.h"

namespace hbp {

struct Error : std::runtime_error {
  hbp_status code;
  Error(hbp_status c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HBP_HIP(call)                                                                        \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      throw ::hbp::Error(HBP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));  \
  } while (0)

#define HBP_REQUIRE(cond, code, msg)                                                         \
  do {                                                                                       \
    if (!(cond)) throw ::hbp::Error((code), (msg));                                          \
  } while (0)

inline size_t dtype_size(int dtype) {
  switch (dtype) {
    case HBP_C128: return 16;
    case HBP_F64: return 8;
    case HBP_C64: return 8;
    case HBP_F32: return 4;
  }
  throw Error(HBP_ERR_INVALID_ARGUMENT, "unknown dtype");
}
inline bool dtype_complex(int dtype) { return dtype == HBP_C128 || dtype == HBP_C64; }

// ------------------------------------------------------------- context
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipMemPool_t pool = nullptr;
  size_t in_use = 0, peak = 0;
  // HIP graphs of batched BP loops: recorded, replays, recordings dropped
  int64_t graphs_recorded = 0, graph_replays = 0, graphs_dropped = 0;
  // grow-only scratch for split-K partials and reductions
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  // grow-only arena for partials whose reduction is deferred (batched)
  void *arena = nullptr;
  size_t arena_bytes = 0;
  // Pinned staging of uploads: kStageSlots slots of kStageSlotBytes used in
  // turn.  An upload that fits a slot is copied there and its DMA ordered on
  // the stream, with the slot's event recorded after it; the host waits only
  // when the slot's previous copy (kStageSlots uploads back) is still in
  // flight, instead of for the whole stream.  Larger uploads wait for the
  // stream as before.
  static constexpr int kStageSlots = 64;
  static constexpr size_t kStageSlotBytes = 256 << 10;
  char *stage = nullptr;
  hipEvent_t stage_ev[kStageSlots] = {};
  bool stage_busy[kStageSlots] = {};
  int stage_next = 0;

  void *alloc(size_t bytes);
  void release(void *p, size_t bytes);
  // device memory this context can still allocate: the device's free memory
  // plus what its pool and its block cache hold in reserve
  size_t free_bytes(size_t *total = nullptr);
  // Released blocks are kept and handed out again for the same size class:
  // large blocks (>= 32 MiB, 2 MiB steps) because every step of a PEPS run
  // frees and re-allocates ~D^4 d-sized site tensors and getting that memory
  // back from the driver each time stalls the host for seconds; small blocks
  // (powers of two from 256 B, at most 256 MiB of them kept) because a run at
  // small bond dimensions allocates and frees work arrays thousands of times
  // per step, and each pool round trip costs more host time than the kernels
  // it serves.  Reuse is safe because all of a context's work is ordered on
  // its one stream.  flush_blocks() gives the cache back (trim, failed
  // allocations, destroy).
  std::multimap<size_t, void *> free_blocks;
  std::unordered_map<void *, size_t> block_sizes;  // live cached-class block -> size
  size_t cached = 0, cached_small = 0;
  void flush_blocks();
  void *get_scratch(size_t bytes);
  // the arena with at least `bytes` (contents not kept when it grows)
  void *get_arena(size_t bytes);
  void sync();

  // kernel timing with HIP events on this context's stream, per kernel class
  struct ProfRec {
    hipEvent_t a, b;
    int kind;
    double flops, bytes;
    bool timed;  // the bracket's main kernel carried a and b in its dispatch
  };
  bool prof = false;
  std::vector<ProfRec> prof_recs;
  // totals of records already folded (their events destroyed), per kind:
  // launches, ms, flops, bytes
  double prof_acc[3][4] = {{0}};
  hipEvent_t prof_event();
  void prof_clear();  // destroys the events of every open record
  // folds the older half of the open records into prof_acc once more than
  // kProfMaxLive are open, so that a long profiled run keeps a bounded
  // number of events (each holds a completion signal and pins its dispatch)
  static constexpr size_t kProfMaxLive = 4096;
  void prof_fold();
  // returns the index of the open record, or -1 when profiling is off
  int prof_begin(int kind, double flops, double bytes);
  void prof_end(int idx);
  // host -> device copy of a small host buffer that may go out of scope on
  // return: enqueued on the context stream and waited for
  void upload(void *dst, const void *src, size_t bytes);
};

struct Tensor {
  int dtype = HBP_C128;
  int order = 0;
  int32_t shape[HBP_MAX_ORDER] = {0};
  size_t size = 1;
  void *data = nullptr;
  size_t bytes = 0;
  // changes on every write through the C-ABI (process-wide unique values), so
  // that a caller can tell whether a tensor changed since it last looked
  uint64_t version = 0;
};
// a fresh version for a tensor that is being written
void touch(Tensor *t);

size_t shape_size(int order, const int32_t *shape);
void tensor_resize(Ctx *ctx, Tensor *t, int dtype, int order, const int32_t *shape);

// Kernel timing slot of the calling thread: Ctx::prof_begin arms it with a
// (start, stop) event pair and the next main-kernel launch of a profiled
// bracket (mode product, message partials, merged product-message) carries
// the pair in its own dispatch (hipExtLaunchKernelGGL), so that timing puts
// no extra packets into the queue.  Launches made while the autotuner times
// its variants do not take the events.
struct LaunchTiming {
  hipEvent_t start = nullptr, stop = nullptr;
  bool armed = false, used = false;
};
LaunchTiming &launch_timing();
// true while tuned_launch times variants on this thread
bool &launch_tuning();
// true while this thread records a context stream into a HIP graph.  What a
// graph cannot hold raises HBP_ERR_UNSUPPORTED instead of being recorded:
// a device allocation that the context's block cache cannot serve, a host
// wait (Ctx::sync, Ctx::upload), an autotuner trial; timing brackets are
// skipped and released blocks stay in the cache (no stream-ordered free).
bool &graph_capturing();
template <typename K, typename... Args>
inline void timed_launch(K kern, dim3 grid, dim3 block, hipStream_t st, Args... args) {
  LaunchTiming &t = launch_timing();
  if (t.armed && !t.used && !launch_tuning()) {
    hipExtLaunchKernelGGL(kern, grid, block, 0, st, t.start, t.stop, 0, args...);
    t.used = true;
  } else {
    hipLaunchKernelGGL(kern, grid, block, 0, st, args...);
  }
}

// ------------------------------------------------------------ kernels
// Out[q, j] = scale * sum_i T[q, i] E(i, j), with the row q = l + L r of a
// column-major (L, Din, R) tensor, E(i, j) = E[i + Din j] (e_trans = 0) or
// E[j + Dout i] (e_trans = 1), complex<double> E; T / Out of scalar S.
// norm_partials (may be null) receives one sum |Out|^2 per block.
template <typename S>
void launch_mode_product(hipStream_t st, const void *T, const double2 *E, void *Out, int64_t L,
                         int Din, int Dout, int64_t R, int e_trans, double scale,
                         double *norm_partials, int *n_partials);

// Partial sums of M[i, j] = sum_q T[q, i] conj(V[q, j]) (conjV = 1) or
// T[q, i] V[q, j] (conjV = 0) over row chunks; T is a (L, D1, R) view and V a
// (L, D2, R) view of scalar S, with r strides ldT_R, ldV_R (normally L*D).
// Returns the number of chunks written (each D1*D2 complex<double>).
// wL (L entries) and wR (R entries), when given, weight row (l, r) of T by
// wL[l] wR[r]: a mode product with a diagonal matrix folded into the load.
template <typename S>
int message_partials(hipStream_t st, const void *T, const void *V, double2 *partials,
                     int64_t L, int D1, int D2, int64_t R, int64_t ldT_R, int64_t ldV_R,
                     int64_t max_chunks, int conjV, const double2 *wL = nullptr,
                     const double2 *wR = nullptr, int Da = 0, int64_t Sb = 0);
// (Da, Sb): column c of a row at L (c % Da) + Sb (c / Da); Da = 0 is the plain view.
// w[i] = prod_b diag(E_b)[c_b] over the column-major split of i into dims[0..nb);
// n may exceed prod(dims) by a trailing unit-weight factor.
// Row weights of a diagonal environment: w[i] = prod_b E[diag_off[b] + c_b (dims[b] + 1)]
// over the column-major digits c_b of i (dims[0], ..., dims[nb-1], then a unit
// factor for indices that carry no message).  Up to WEIGHT_BATCH vectors
// are built in one launch.
struct WeightArgs {
  int nb;
  int64_t dims[HBP_MAX_ORDER];
  int64_t diag_off[HBP_MAX_ORDER];  // offset of the message's (0,0) element
};
constexpr int WEIGHT_BATCH = 8;
struct WeightBatch {
  WeightArgs a[WEIGHT_BATCH];
  int64_t n[WEIGHT_BATCH];
  double2 *w[WEIGHT_BATCH];
  int count;
};
void launch_build_weights_batch(hipStream_t st, const double2 *E, const WeightBatch &b);
void launch_build_weights(hipStream_t st, const double2 *E, double2 *w, int64_t n, int nb,
                          const int64_t *dims, const int64_t *diag_off);
int64_t message_num_chunks(int64_t Q);
// Partials of M[i, i'] = sum_{o, a'} (sum_a P[o; i, a] E(a, a')) conj V[o; i', a']
// for adjacent bonds k (i, dim Dk) and j (a, dim Dj), both <= 64, of two
// column-major tensors of the same shape viewed as (L, i, a, R) when k_first
// and (L, a, i, R) otherwise, E(a, a') = E[a + Dj a'].  Returns the number of
// partials written (each Dk*Dk complex of scalar S, at most max_blocks).
template <typename S>
int product_message_partials(hipStream_t st, const void *P, const void *V, const double2 *E,
                             void *partials, int64_t L, int64_t R, int Dk, int Dj,
                             int k_first, int max_blocks);
// Partials of both messages of a bond pair (x, y) at the front of a tensor
// whose environment is diagonal: slices X_o (Dx x Dy, contiguous, o < nO),
// M_x = sum_o w_o X_o diag(ey) X_o^H, M_y = sum_o w_o X_o^T diag(ex) conj(X_o),
// ex / ey read as E[c (D + 1)] (diagonals of the messages), wo may be null
// (all ones).  Returns the number of partials (each Dx*Dx resp. Dy*Dy complex
// of scalar S, at most max_blocks).
template <typename S>
int pair_gram_partials(hipStream_t st, const void *X, int64_t nO, int Dx, int Dy,
                       const double2 *Ex, const double2 *Ey, const double2 *wo, void *part_x,
                       void *part_y, int max_blocks);
// Y[c + Cn (l + L r)] = X[l + L (c + Cn r)] for complex elements of elem_bytes
// (8 or 16): moves the middle index block of an (L, Cn, R) view to the front.
void launch_swap_blocks(hipStream_t st, const void *X, void *Y, int64_t L, int64_t Cn,
                        int64_t R, int elem_bytes);
// M = sum over chunks (fixed order), complex<double>; complex<float>
// partials are summed in double.
// rho[p, p'] = sum_l w[l] T[l + L p] conj V[l + L p'] over the rows l < L for
// a physical index of dimension 2 (w may be null): FP64 partials of 4 entries
// per block into `partials`; returns the block count (<= max_blocks)
template <typename S>
int rho_partials(hipStream_t st, const void *T, const void *V, const double2 *w, int64_t L,
                 double2 *partials, int max_blocks);
// up to RED_BATCH fixed-order reductions of partials (FP64, or FP32 stored as
// float2) into FP64 messages, in one launch
constexpr int RED_BATCH = 8;
struct RedBatchJob {
  const void *partials;
  double2 *M;
  int nch, n, f32;
};
struct RedBatch {
  RedBatchJob j[RED_BATCH];
  int count;
};
void launch_reduce_batch(hipStream_t st, const RedBatch &b);
void launch_reduce_partials(hipStream_t st, const double2 *partials, int nchunks, int n,
                            double2 *M);
void launch_reduce_partials(hipStream_t st, const float2 *partials, int nchunks, int n,
                            double2 *M);

// Batched small-bond sweep (hbp_batched.hip): one work item per site of a
// PEPS in the batch; all of the site's outgoing raw messages (unnormalised,
// complex double, D x D column-major) from its tensor V (bonds then the
// physical index) and the incoming messages E + in_off[k] (D x D each).
constexpr int SMALL_D_MAX = 16;
struct SmallSiteJob {
  const void *V;
  const double2 *E;
  double2 *raw;
  int64_t in_off[4];
  int64_t out_off[4];
  int dims[5];  // bond dims, then the physical dim at index z
  int z;
};
template <typename S>
void launch_sweep_small(hipStream_t st, const SmallSiteJob *jobs_dev, int njobs, void *scratch,
                        int64_t slab_elems, int grid, const double2 *E_base = nullptr,
                        double2 *raw_base = nullptr);
// Truncation of one site with bonds <= SMALL_D_MAX (truncation.h:156-217):
// out = normalise(V x_0 P_0 x_1 ... x_{z-1} P_{z-1}) with the projector of
// every bond (dims[k] x keep[k], complex double: P[i + dims[k] j], or
// P[j + keep[k] i] where trans[k]), one workgroup per site.
struct SmallProjJob {
  const void *V;
  void *out;
  const double2 *P[4];
  int dims[5];  // bond dims before the projectors, then the physical dim at index z
  int keep[4];
  int trans[4];
  int z;
};
template <typename S>
void launch_project_small(hipStream_t st, const SmallProjJob *jobs_dev, int njobs, void *scratch,
                          int64_t slab_elems, int grid);
// Products of operators along paths (measure.h:14-94 site by site, the path's
// transfer matrix carried from site to site), bonds <= SMALL_D_MAX, one
// workgroup per path.  Site i of a path: T = V x_{b != kin, kout} E_in(b)
// [x_kin X] [x_phys op]; M = <T, conj V> over all but the bond t = kout, else
// kin, else the physical index; X <- M / |M|_F with lg += log |M|_F while
// kout >= 0; at the last site val = sum X .* M (or the trace of M on a
// one-site path).
struct SmallChainSite {
  const void *V;
  const double2 *op;  // (d, d) or null
  int64_t in_off[4];  // incoming messages of the bonds (offsets into E)
  int dims[5];        // bond dims, then the physical dim at index z
  int z, kin, kout;
};
struct SmallChainJob {
  int first, n;  // its sites: sites[first .. first + n)
  int slot;      // val[slot], lg[slot]
};
template <typename S>
void launch_chain_small(hipStream_t st, const SmallChainSite *sites, const SmallChainJob *jobs,
                        int njobs, const double2 *E, void *scratch, int64_t slab_elems,
                        double2 *val, double *lg, int grid);
// Grids the batched sweep is launched with: GRID_MULT[v] workgroups per CU
// (variant v), at most one per work item; scratch for the largest.
constexpr int SMALL_GRID_VARIANTS = 3;
int small_sweep_grid(int variant, int njobs);
// k_sweep_small's work items per site: one per outgoing message (degree <= 4)
constexpr int SMALL_SWEEP_ITEMS = 4;
// The variant for (precision, njobs class, slab class): timed on first use by
// running every variant on the real work (each computes the same messages;
// the chosen one runs last), remembered for the process.  Must not run while
// the stream is being captured.
template <typename S>
int launch_sweep_small_tuned(hipStream_t st, const SmallSiteJob *jobs_dev, int njobs,
                             void *scratch, int64_t slab_elems, const double2 *E_base = nullptr,
                             double2 *raw_base = nullptr);

// the autotuner's choices so far, one line per shape class
std::string autotune_report();
int mode_product_col_blocks(int Dout);
constexpr int MSG_ROWS_PER_STAGE = 32;
constexpr int MSG_OUT_ROWS = 64;
constexpr int RED_ENTRIES = 16;

// Elementwise / reduction helpers (complex<double> unless noted).
void launch_norm2_partials(hipStream_t st, const void *x, size_t n, int dtype, double *partials,
                           int *nblocks);
void launch_sum_partials(hipStream_t st, const double *partials, int n, double *out);
void launch_scale(hipStream_t st, void *x, size_t n, int dtype, double re, double im,
                  const double *dev_scale_inv_sqrt);
void launch_conj(hipStream_t st, void *x, size_t n, int dtype);
void launch_fill(hipStream_t st, void *x, size_t n, int dtype, double re, double im);
void launch_convert(hipStream_t st, const void *x, int dtype_in, void *y, int dtype_out,
                    size_t n);
void launch_permute(hipStream_t st, const void *in, void *out, int elem_bytes, int order,
                    const int32_t *in_shape, const int32_t *perm);
void launch_axpy_c128(hipStream_t st, const double2 *x, double2 *y, size_t n, double re,
                      double im);
void launch_trace(hipStream_t st, const double2 *in, double2 *out, int order,
                  const int32_t *shape, int npairs, const int32_t *pairs);

// Batched one-sided Jacobi SVD of complex<double> matrices.
//   A[b]: m x n column-major at A + offA[b]; results U[b] (m x k), s[b] (k),
//   Vd[b] (k x n), k = min(m, n), descending.  work needs
//   batch * (m*n + n*n) complex<double>.
struct SvdJob {
  int64_t a_off, u_off, s_off, vd_off, work_off;
  int m, n;
};
// Matrices with m n + n^2 <= SVD_LDS_ELEMS (64 x 64 and smaller) run with
// their working columns in LDS (max_m: the largest m of the batch, 0 when
// unknown); larger ones in `work` (job.work_off, m n + n^2 elements each).
constexpr int64_t SVD_LDS_ELEMS = 8192;
void launch_batched_svd(hipStream_t st, const double2 *A, double2 *U, double *s, double2 *Vd,
                        double2 *work, const SvdJob *jobs_dev, int batch, int max_n, int max_m);
// one SvdJob written to device memory by a kernel (no host copy)
void launch_store_svd_job(hipStream_t st, const SvdJob &job, SvdJob *dst);

// Small batched complex GEMM: C[b] = op(A[b]) op(B[b]) for matrices up to a few
// hundred on a side (truncation projectors and the generic TCI contract).
struct GemmJob {
  int64_t a_off, b_off, c_off;
  int m, n, k;
  int lda, ldb, ldc;
  int opa, opb;  // 0 = N, 1 = T, 2 = C (conjugate transpose), 3 = conj (no transpose)
  double2 alpha;
};
void launch_batched_gemm(hipStream_t st, const double2 *A, const double2 *B, double2 *C,
                         const GemmJob *jobs_dev, int batch, int max_m, int max_n);

}  // namespace hbp

// opaque C handles
struct hbp_ctx : hbp::Ctx {};
struct hbp_tensor : hbp::Tensor {};
