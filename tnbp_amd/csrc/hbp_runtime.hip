// Context, device memory, and the tensor (TCI) layer of the C-ABI.
//
// tci::contract is dispatched by the shape of the contraction onto the two
// hand-written MFMA kernels:
//   * one operand is a matrix sharing one bond with the other -> mode product;
//   * both operands share every bond but one (or two) at the same positions
//     -> message contraction (sum over all shared bonds, split-K, FP64 partials);
//   * otherwise the operands are permuted into a GEMM whose narrow side (<= 128)
//     becomes the mode product's output bond.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <set>
#include <string>

#include "hbp_internal.h"

namespace {
thread_local std::string g_last_error;
}

// Device scratch of hbp_svd_noalloc: its job record (first 256 bytes) and the
// one-sided Jacobi work arrays (a copy of A, m x n, and the rotation
// accumulator, n x n, complex128).  Every device the library runs on holds
// one copy until the process ends.  In host code `scratch` only names the
// symbol: its device address on each device comes from hipGetSymbolAddress
// (bind_noalloc_scratch), and only that address is passed to kernels.
static constexpr size_t kNoAllocScratchBytes = (size_t)16 << 20;
static constexpr size_t kNoAllocJobBytes = 256;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {nullptr};

// Looks up this device's copy of `scratch` once; the device must be current.
void bind_noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || ptr == nullptr)
    throw ::hbp::Error(HBP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") +
                                        hipGetErrorString(e));
  g_scratch_ptr[device] = ptr;
}
}  // namespace

#define HBP_API_BEGIN try {
#define HBP_API_END                                                    \
  }                                                                    \
  catch (const ::hbp::Error &e) {                                      \
    g_last_error = e.what();                                           \
    return e.code;                                                     \
  }                                                                    \
  catch (const std::bad_alloc &e) {                                    \
    g_last_error = "bad_alloc";                                        \
    return HBP_ERR_BAD_ALLOC;                                          \
  }                                                                    \
  catch (const std::exception &e) {                                    \
    g_last_error = e.what();                                           \
    return HBP_ERR_LOGIC;                                              \
  }                                                                    \
  return HBP_OK;

namespace hbp {

void set_last_error(const std::string &s) { g_last_error = s; }

void touch(Tensor *t) {
  static std::atomic<uint64_t> counter{0};
  t->version = ++counter;
}

size_t shape_size(int order, const int32_t *shape) {
  size_t n = 1;
  for (int d = 0; d < order; ++d) n *= (size_t)shape[d];
  return n;
}

namespace {
constexpr size_t kBigBlock = (size_t)32 << 20;
constexpr size_t kBlockStep = (size_t)2 << 20;
constexpr size_t kSmallCacheCap = (size_t)256 << 20;
// size class of a request: 2 MiB steps from 32 MiB, powers of two below
size_t block_class(size_t bytes) {
  if (bytes >= kBigBlock) return (bytes + kBlockStep - 1) / kBlockStep * kBlockStep;
  size_t c = 256;
  while (c < bytes) c <<= 1;
  return c;
}
}  // namespace

void *Ctx::alloc(size_t bytes) {
  if (bytes == 0) bytes = 16;
  bytes = block_class(bytes);
  auto it = free_blocks.find(bytes);
  if (it != free_blocks.end()) {
    void *p = it->second;
    free_blocks.erase(it);
    cached -= bytes;
    if (bytes < kBigBlock) cached_small -= bytes;
    block_sizes[p] = bytes;
    in_use += bytes;
    peak = std::max(peak, in_use);
    return p;
  }
  HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED,
              "device allocation of " + std::to_string(bytes) + " bytes while a graph is recorded");
  void *p = nullptr;
  hipError_t e = hipMallocFromPoolAsync(&p, bytes, pool, stream);
  if (e != hipSuccess && cached > 0) {
    (void)hipGetLastError();
    flush_blocks();
    e = hipMallocFromPoolAsync(&p, bytes, pool, stream);
  }
  if (e != hipSuccess) {
    (void)hipGetLastError();
    throw Error(HBP_ERR_BAD_ALLOC, std::string("device allocation of ") + std::to_string(bytes) +
                                       " bytes failed: " + hipGetErrorString(e));
  }
  block_sizes[p] = bytes;
  in_use += bytes;
  peak = std::max(peak, in_use);
  return p;
}

void Ctx::release(void *p, size_t bytes) {
  if (!p) return;
  auto it = block_sizes.find(p);
  // the size recorded at allocation decides the class, not the caller's
  const size_t b = it != block_sizes.end() ? it->second : (bytes == 0 ? (size_t)16 : bytes);
  if (it != block_sizes.end()) block_sizes.erase(it);
  in_use -= std::min(in_use, b);
  if (b >= kBigBlock || cached_small + b <= kSmallCacheCap || graph_capturing()) {
    free_blocks.emplace(b, p);
    cached += b;
    if (b < kBigBlock) cached_small += b;
    return;
  }
  HBP_HIP(hipFreeAsync(p, stream));
}

void Ctx::flush_blocks() {
  for (auto &kv : free_blocks) (void)hipFreeAsync(kv.second, stream);
  free_blocks.clear();
  cached = 0;
  cached_small = 0;
}

void *Ctx::get_scratch(size_t bytes) {
  if (bytes > scratch_bytes) {
    HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED, "scratch growth while a graph is recorded");
    if (scratch) release(scratch, scratch_bytes);
    size_t nb = std::max(bytes, scratch_bytes * 2);
    nb = (nb + 4095) & ~(size_t)4095;
    scratch = alloc(nb);
    scratch_bytes = nb;
  }
  return scratch;
}

void *Ctx::get_arena(size_t bytes) {
  if (bytes > arena_bytes) {
    HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED, "arena growth while a graph is recorded");
    if (arena) release(arena, arena_bytes);
    arena_bytes = (std::max(bytes, arena_bytes * 2) + 4095) & ~(size_t)4095;
    arena = alloc(arena_bytes);
  }
  return arena;
}

void Ctx::sync() {
  HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED, "host wait while a graph is recorded");
  HBP_HIP(hipStreamSynchronize(stream));
}

size_t Ctx::free_bytes(size_t *total) {
  size_t fr = 0, tot = 0;
  HBP_HIP(hipMemGetInfo(&fr, &tot));
  uint64_t reserved = 0, used = 0;
  HBP_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved));
  HBP_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used));
  if (total) *total = tot;
  return fr + (size_t)(reserved > used ? reserved - used : 0) + cached;
}

// Events are not recycled: an event that a kernel dispatch carried keeps that
// dispatch (and the device memory its arguments point to) alive until the
// event is destroyed, so read-out records destroy theirs.
hipEvent_t Ctx::prof_event() {
  hipEvent_t e;
  HBP_HIP(hipEventCreate(&e));
  return e;
}

void Ctx::prof_clear() {
  for (auto &r : prof_recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  prof_recs.clear();
  for (auto &k : prof_acc)
    for (double &v : k) v = 0.0;
}

void Ctx::prof_fold() {
  const size_t n = prof_recs.size() / 2;
  for (size_t i = 0; i < n; ++i) {
    ProfRec &r = prof_recs[i];
    if (r.timed && r.kind >= 0 && r.kind < 3) {
      HBP_HIP(hipEventSynchronize(r.b));
      float t = 0.0f;
      HBP_HIP(hipEventElapsedTime(&t, r.a, r.b));
      prof_acc[r.kind][0] += 1.0;
      prof_acc[r.kind][1] += t;
      prof_acc[r.kind][2] += r.flops;
      prof_acc[r.kind][3] += r.bytes;
    }
    HBP_HIP(hipEventDestroy(r.a));
    HBP_HIP(hipEventDestroy(r.b));
  }
  prof_recs.erase(prof_recs.begin(), prof_recs.begin() + (std::ptrdiff_t)n);
}

LaunchTiming &launch_timing() {
  static thread_local LaunchTiming t;
  return t;
}
bool &launch_tuning() {
  static thread_local bool on = false;
  return on;
}
bool &graph_capturing() {
  static thread_local bool on = false;
  return on;
}

int Ctx::prof_begin(int kind, double flops, double bytes) {
  if (!prof || graph_capturing()) return -1;
  if (prof_recs.size() >= kProfMaxLive) prof_fold();
  ProfRec r{prof_event(), prof_event(), kind, flops, bytes, false};
  LaunchTiming &t = launch_timing();
  t.start = r.a;
  t.stop = r.b;
  t.armed = true;
  t.used = false;
  prof_recs.push_back(r);
  return (int)prof_recs.size() - 1;
}

void Ctx::prof_end(int idx) {
  if (idx < 0) return;
  LaunchTiming &t = launch_timing();
  // a bracket whose kernel did not take the events (none of the three main
  // kernels ran in it) is not counted
  prof_recs[idx].timed = t.armed && t.used;
  t = LaunchTiming{};
}

void Ctx::upload(void *dst, const void *src, size_t bytes) {
  HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED, "upload while a graph is recorded");
  if (bytes == 0) return;
  if (bytes <= kStageSlotBytes) {
    if (!stage) {
      HBP_HIP(hipHostMalloc((void **)&stage, kStageSlots * kStageSlotBytes, hipHostMallocDefault));
      for (int k = 0; k < kStageSlots; ++k)
        HBP_HIP(hipEventCreateWithFlags(&stage_ev[k], hipEventDisableTiming));
    }
    const int k = stage_next;
    stage_next = (k + 1) % kStageSlots;
    if (stage_busy[k]) HBP_HIP(hipEventSynchronize(stage_ev[k]));
    char *h = stage + (size_t)k * kStageSlotBytes;
    std::memcpy(h, src, bytes);
    HBP_HIP(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, stream));
    HBP_HIP(hipEventRecord(stage_ev[k], stream));
    stage_busy[k] = true;
    return;
  }
  HBP_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream));
  HBP_HIP(hipStreamSynchronize(stream));
}

void tensor_resize(Ctx *ctx, Tensor *t, int dtype, int order, const int32_t *shape) {
  HBP_REQUIRE(order >= 0 && order <= HBP_MAX_ORDER, HBP_ERR_INVALID_ARGUMENT,
              "tensor order out of range");
  for (int d = 0; d < order; ++d)
    HBP_REQUIRE(shape[d] >= 1, HBP_ERR_INVALID_ARGUMENT, "bond dimension must be >= 1");
  const size_t n = shape_size(order, shape);
  const size_t bytes = n * dtype_size(dtype);
  if (t->data == nullptr || t->bytes < bytes || t->bytes > 4 * bytes + 4096) {
    if (t->data) ctx->release(t->data, t->bytes);
    t->data = ctx->alloc(bytes);
    t->bytes = bytes;
  }
  t->dtype = dtype;
  t->order = order;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) t->shape[d] = d < order ? shape[d] : 0;
  t->size = n;
  touch(t);
}

// ------------------------------------------------------------ reductions
double device_norm2(Ctx *ctx, const void *x, size_t n, int dtype, double *dev_out) {
  double *parts = (double *)ctx->get_scratch(1024 * sizeof(double) + 64);
  int nb = 0;
  launch_norm2_partials(ctx->stream, x, n, dtype, parts, &nb);
  launch_sum_partials(ctx->stream, parts, nb, dev_out);
  return 0.0;
}

static double norm_sync(Ctx *ctx, const Tensor *t) {
  double *dv = (double *)ctx->alloc(sizeof(double));
  device_norm2(ctx, t->data, t->size, t->dtype, dv);
  double h = 0.0;
  HBP_HIP(hipMemcpyAsync(&h, dv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->release(dv, sizeof(double));
  ctx->sync();
  return std::sqrt(h);
}

// ------------------------------------------------------------- contract
struct ContractPlan {
  std::vector<int> shared;  // labels summed
  std::vector<int> ax_a, ax_b;
  std::vector<int> free_a, free_b;  // axes of a / b that survive, in their order
  std::vector<int> perm;            // output axis i = concat(free_a, free_b)[perm[i]]
};

static ContractPlan plan_labels(const Tensor *a, const int32_t *la, const Tensor *b,
                                const int32_t *lb, const int32_t *lc, int nc) {
  ContractPlan p;
  std::set<int> sa, sb, sc;
  for (int i = 0; i < a->order; ++i)
    HBP_REQUIRE(sa.insert(la[i]).second, HBP_ERR_INVALID_ARGUMENT, "bond_labs_a has duplicate label.");
  for (int i = 0; i < b->order; ++i)
    HBP_REQUIRE(sb.insert(lb[i]).second, HBP_ERR_INVALID_ARGUMENT, "bond_labs_b has duplicate label.");
  for (int i = 0; i < nc; ++i)
    HBP_REQUIRE(sc.insert(lc[i]).second, HBP_ERR_INVALID_ARGUMENT, "bond_labs_c has duplicate label.");
  std::set<int> surv;
  for (int x : sa)
    if (sb.count(x)) p.shared.push_back(x);
    else surv.insert(x);
  for (int x : sb)
    if (!sa.count(x)) surv.insert(x);
  HBP_REQUIRE(surv == sc, HBP_ERR_INVALID_ARGUMENT,
              "bond_labs_c does not match the set of survivor labels.");
  for (int x : p.shared) {
    const int ia = (int)(std::find(la, la + a->order, x) - la);
    const int ib = (int)(std::find(lb, lb + b->order, x) - lb);
    HBP_REQUIRE(a->shape[ia] == b->shape[ib], HBP_ERR_INVALID_ARGUMENT,
                "contracted bond dimensions differ");
    p.ax_a.push_back(ia);
    p.ax_b.push_back(ib);
  }
  std::vector<int> free_labels;
  for (int i = 0; i < a->order; ++i)
    if (!sb.count(la[i])) {
      p.free_a.push_back(i);
      free_labels.push_back(la[i]);
    }
  for (int i = 0; i < b->order; ++i)
    if (!sa.count(lb[i])) {
      p.free_b.push_back(i);
      free_labels.push_back(lb[i]);
    }
  for (int i = 0; i < nc; ++i)
    p.perm.push_back((int)(std::find(free_labels.begin(), free_labels.end(), lc[i]) -
                           free_labels.begin()));
  return p;
}

static bool is_identity(const std::vector<int> &v) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] != (int)i) return false;
  return true;
}

// Permute t (dtype C128) into a fresh buffer with axis order `axes`.
static void *permuted(Ctx *ctx, const Tensor *t, const std::vector<int> &axes, size_t *bytes) {
  *bytes = t->size * 16;
  void *buf = ctx->alloc(*bytes);
  launch_permute(ctx->stream, t->data, buf, 16, t->order, t->shape, axes.data());
  return buf;
}

// C(M x N) = A(M x K) B(K x N), all column-major complex<double>.
static void gemm_c128(Ctx *ctx, const double2 *A, const double2 *B, double2 *C, int64_t M,
                      int64_t K, int64_t N) {
  if (N <= 128) {
    launch_mode_product<double>(ctx->stream, A, B, C, M, (int)K, (int)N, 1, 0, 1.0, nullptr,
                                nullptr);
  } else if (M <= 128) {
    // C^T = B^T A^T: B^T is the (1, K, N) view of B, A^T enters transposed
    launch_mode_product<double>(ctx->stream, B, A, C, 1, (int)K, (int)M, N, 1, 1.0, nullptr,
                                nullptr);
  } else {
    HBP_REQUIRE(K <= INT32_MAX && M <= INT32_MAX && N <= INT32_MAX, HBP_ERR_UNSUPPORTED,
                "gemm too large");
    GemmJob j{};
    j.m = (int)M; j.n = (int)N; j.k = (int)K;
    j.lda = (int)M; j.ldb = (int)K; j.ldc = (int)M;
    j.alpha = make_double2(1.0, 0.0);
    GemmJob *dj = (GemmJob *)ctx->alloc(sizeof(GemmJob));
    ctx->upload(dj, &j, sizeof(j));
    launch_batched_gemm(ctx->stream, A, B, C, dj, 1, (int)M, (int)N);
    ctx->release(dj, sizeof(GemmJob));
  }
}

// sum over all shared bonds except positions kept; T and V have identical
// shapes and label positions; output M[i, j] over one kept position k.
static void message_c128(Ctx *ctx, const void *T, const void *V, const int32_t *shape,
                         int order, int k, int conjV, double2 *M) {
  int64_t L = 1, R = 1;
  for (int d = 0; d < k; ++d) L *= shape[d];
  for (int d = k + 1; d < order; ++d) R *= shape[d];
  const int D = shape[k];
  HBP_REQUIRE(D <= 128, HBP_ERR_UNSUPPORTED, "message: bond > 128");
  const int64_t nch = message_num_chunks(L * R);
  double2 *parts = (double2 *)ctx->get_scratch((size_t)nch * D * D * 16 + 64);
  const int n = message_partials<double>(ctx->stream, T, V, parts, L, D, D, R, L * D, L * D,
                                         nch, conjV);
  launch_reduce_partials(ctx->stream, parts, n, D * D, M);
}

// Element types other than complex128 (the apps' single-precision build,
// real tensors) in contract, svd, linear_combine and trace: the operands are
// converted to complex128 (exactly), the complex128 kernels run, and each
// result is rounded to the operands' type: complex when an operand is
// complex, single precision when every operand is.
static bool single_dtype(int d) { return d == HBP_C64 || d == HBP_F32; }
static int common_dtype(int a, int b) {
  const bool cplx = dtype_complex(a) || dtype_complex(b);
  const bool single = single_dtype(a) && single_dtype(b);
  return cplx ? (single ? HBP_C64 : HBP_C128) : (single ? HBP_F32 : HBP_F64);
}
// x itself when complex128, else its complex128 copy in tmp
static const Tensor *as_c128(Ctx *ctx, const Tensor *x, Tensor *tmp) {
  if (x->dtype == HBP_C128) return x;
  tensor_resize(ctx, tmp, HBP_C128, x->order, x->shape);
  launch_convert(ctx->stream, x->data, x->dtype, tmp->data, HBP_C128, x->size);
  return tmp;
}
// round a complex128 (or float64) result to dtype in place
static void round_to(Ctx *ctx, Tensor *t, int dtype) {
  if (t->dtype == dtype) return;
  const size_t bytes = std::max<size_t>(t->size, 1) * dtype_size(dtype);
  void *d = ctx->alloc(bytes);
  launch_convert(ctx->stream, t->data, t->dtype, d, dtype, t->size);
  ctx->release(t->data, t->bytes);
  t->data = d;
  t->bytes = bytes;
  t->dtype = dtype;
}
static void drop_tmp(Ctx *ctx, Tensor *tmp) {
  if (tmp->data) ctx->release(tmp->data, tmp->bytes);
  tmp->data = nullptr;
}

static void contract_c128(Ctx *ctx, const Tensor *a, const int32_t *la, const Tensor *b,
                          const int32_t *lb, Tensor *c, const int32_t *lc, int nc);
static void contract_impl(Ctx *ctx, const Tensor *a, const int32_t *la, const Tensor *b,
                          const int32_t *lb, Tensor *c, const int32_t *lc, int nc) {
  if (a->dtype == HBP_C128 && b->dtype == HBP_C128) {
    contract_c128(ctx, a, la, b, lb, c, lc, nc);
    return;
  }
  const int dt = common_dtype(a->dtype, b->dtype);
  Tensor ta, tb;
  const Tensor *pa = as_c128(ctx, a, &ta), *pb = as_c128(ctx, b, &tb);
  contract_c128(ctx, pa, la, pb, lb, c, lc, nc);
  drop_tmp(ctx, &ta);
  drop_tmp(ctx, &tb);
  round_to(ctx, c, dt);
}

static void contract_c128(Ctx *ctx, const Tensor *a, const int32_t *la, const Tensor *b,
                          const int32_t *lb, Tensor *c, const int32_t *lc, int nc) {
  HBP_REQUIRE(nc <= HBP_MAX_ORDER, HBP_ERR_INVALID_ARGUMENT, "contract: output order too large");
  ContractPlan p = plan_labels(a, la, b, lb, lc, nc);
  int32_t out_shape[HBP_MAX_ORDER];
  std::vector<int32_t> fshape;
  for (int i : p.free_a) fshape.push_back(a->shape[i]);
  for (int i : p.free_b) fshape.push_back(b->shape[i]);
  for (int i = 0; i < nc; ++i) out_shape[i] = fshape[p.perm[i]];
  const size_t out_n = shape_size(nc, out_shape);
  double2 *out = (double2 *)ctx->alloc(out_n * 16);

  bool done = false;
  // (1) mode product: b is a matrix sharing exactly one bond with a, and the
  //     output keeps a's bond order with b's free bond in place of the shared one
  for (int swap = 0; swap < 2 && !done; ++swap) {
    const Tensor *X = swap ? b : a;
    const Tensor *Mx = swap ? a : b;
    const std::vector<int> &axX = swap ? p.ax_b : p.ax_a;
    const std::vector<int> &axM = swap ? p.ax_a : p.ax_b;
    const std::vector<int> &freeX = swap ? p.free_b : p.free_a;
    const std::vector<int> &freeM = swap ? p.free_a : p.free_b;
    if (Mx->order != 2 || p.shared.size() != 1 || freeM.size() != 1) continue;
    const int k = axX[0];
    // output axis order must be X's axes with axis k replaced by M's free bond
    std::vector<int> want;  // index into concat(free_a, free_b) for each X axis
    const int nfa = (int)p.free_a.size();
    for (int d = 0; d < X->order; ++d) {
      if (d == k) want.push_back(swap ? 0 : nfa);
      else {
        const int pos = (int)(std::find(freeX.begin(), freeX.end(), d) - freeX.begin());
        want.push_back(swap ? nfa > 0 ? 1 + pos : pos : pos);
      }
    }
    if (swap) {  // concat order is (free_a = M's free bond, free_b = X's axes)
      want.clear();
      for (int d = 0; d < X->order; ++d) {
        if (d == k) want.push_back(0);
        else {
          const int pos = (int)(std::find(freeX.begin(), freeX.end(), d) - freeX.begin());
          want.push_back(1 + pos);
        }
      }
    }
    if (want != p.perm) continue;
    int64_t L = 1, R = 1;
    for (int d = 0; d < k; ++d) L *= X->shape[d];
    for (int d = k + 1; d < X->order; ++d) R *= X->shape[d];
    const int Din = X->shape[k];
    const int Dout = Mx->shape[freeM[0]];
    if (Dout > 128) continue;
    // E(i, j) with i the shared bond of M: e_trans when the shared bond is M's axis 1
    const int e_trans = axM[0] == 1 ? 1 : 0;
    launch_mode_product<double>(ctx->stream, X->data, (const double2 *)Mx->data, out, L, Din,
                                Dout, R, e_trans, 1.0, nullptr, nullptr);
    done = true;
  }
  // (2) message: same order, all axes shared at equal positions except one
  if (!done && a->order == b->order && nc == 2 && (int)p.shared.size() == a->order - 1) {
    bool same = true;
    for (size_t s = 0; s < p.shared.size(); ++s) same &= p.ax_a[s] == p.ax_b[s];
    if (same && p.free_a.size() == 1 && p.free_b.size() == 1 && p.free_a[0] == p.free_b[0]) {
      const int k = p.free_a[0];
      bool shapes = true;
      for (int d = 0; d < a->order; ++d) shapes &= a->shape[d] == b->shape[d];
      if (shapes && a->shape[k] <= 128) {
        // M[i, j] = sum a[.. i ..] b[.. j ..]  (no conjugation: TCI operands are explicit)
        double2 *M = out;
        double2 *tmp = nullptr;
        if (!is_identity(p.perm)) tmp = (double2 *)ctx->alloc(out_n * 16), M = tmp;
        message_c128(ctx, a->data, b->data, a->shape, a->order, k, 0, M);
        if (tmp) {
          const int32_t sh[2] = {a->shape[k], b->shape[k]};
          const int32_t pr[2] = {1, 0};
          launch_permute(ctx->stream, tmp, out, 16, 2, sh, pr);
          ctx->release(tmp, out_n * 16);
        }
        done = true;
      }
    }
  }
  // (3) full contraction to a scalar: a dot product over identical layouts
  if (!done && nc == 0 && a->order == b->order) {
    bool same = true;
    for (size_t s = 0; s < p.shared.size(); ++s) same &= p.ax_a[s] == p.ax_b[s];
    if (same) {
      const int32_t sh[1] = {(int32_t)std::min<size_t>(a->size, INT32_MAX)};
      if (a->size <= INT32_MAX) {
        // (a->size, 1) view with the unit bond as the "message" bond
        const int32_t sh2[2] = {sh[0], 1};
        message_c128(ctx, a->data, b->data, sh2, 2, 1, 0, out);
        done = true;
      }
    }
  }
  // (4) general: permute into A (M x K), B (K x N), GEMM, permute the result
  if (!done) {
    std::vector<int> pa = p.free_a, pb = p.ax_b;
    pa.insert(pa.end(), p.ax_a.begin(), p.ax_a.end());
    pb.insert(pb.end(), p.free_b.begin(), p.free_b.end());
    int64_t M = 1, K = 1, N = 1;
    for (int i : p.free_a) M *= a->shape[i];
    for (int i : p.ax_a) K *= a->shape[i];
    for (int i : p.free_b) N *= b->shape[i];
    size_t ba = 0, bb = 0;
    const void *A = a->data, *B = b->data;
    void *Ab = nullptr, *Bb = nullptr;
    if (!is_identity(pa)) A = Ab = permuted(ctx, a, pa, &ba);
    if (!is_identity(pb)) B = Bb = permuted(ctx, b, pb, &bb);
    double2 *Cm = out;
    double2 *tmp = nullptr;
    if (!is_identity(p.perm)) tmp = (double2 *)ctx->alloc(out_n * 16), Cm = tmp;
    gemm_c128(ctx, (const double2 *)A, (const double2 *)B, Cm, M, K, N);
    if (tmp) {
      launch_permute(ctx->stream, tmp, out, 16, (int)fshape.size(), fshape.data(),
                     p.perm.data());
      ctx->release(tmp, out_n * 16);
    }
    if (Ab) ctx->release(Ab, ba);
    if (Bb) ctx->release(Bb, bb);
  }
  // install the result (c may alias a or b: they are read by now in stream order)
  if (c->data) ctx->release(c->data, c->bytes);
  c->data = out;
  c->bytes = out_n * 16;
  c->dtype = HBP_C128;
  c->order = nc;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) c->shape[d] = d < nc ? out_shape[d] : 0;
  c->size = out_n;
}

// --------------------------------------------------------------- SVD
void svd_impl(Ctx *ctx, const Tensor *a, int nrows, Tensor *u, Tensor *s, Tensor *vd,
              int keep_max, double s_min, double *trunc_err) {
  HBP_REQUIRE(a->dtype == HBP_C128, HBP_ERR_UNSUPPORTED, "svd: complex128 only");
  HBP_REQUIRE(nrows >= 0 && nrows <= a->order, HBP_ERR_INVALID_ARGUMENT, "svd: bad row count");
  int64_t m = 1, n = 1;
  for (int d = 0; d < nrows; ++d) m *= a->shape[d];
  for (int d = nrows; d < a->order; ++d) n *= a->shape[d];
  if (m < n) {
    // one-sided Jacobi orthogonalises columns: run it on A^H (n x m) and swap
    // the factors, A = (A^H)^H = Vd'^H S U'^H
    Tensor ah;
    const int32_t sh2[2] = {(int32_t)m, (int32_t)n}, pr[2] = {1, 0};
    const int32_t shh[2] = {(int32_t)n, (int32_t)m};
    tensor_resize(ctx, &ah, HBP_C128, 2, shh);
    launch_permute(ctx->stream, a->data, ah.data, 16, 2, sh2, pr);
    launch_conj(ctx->stream, ah.data, ah.size, HBP_C128);
    Tensor u2, vd2;
    svd_impl(ctx, &ah, 1, &u2, s, &vd2, keep_max, s_min, trunc_err);
    const int keep = s->shape[0];
    // u = vd2^H (m x keep), vd = u2^H (keep x n)
    int32_t ush[HBP_MAX_ORDER], vsh[HBP_MAX_ORDER];
    for (int d = 0; d < nrows; ++d) ush[d] = a->shape[d];
    ush[nrows] = keep;
    tensor_resize(ctx, u, HBP_C128, nrows + 1, ush);
    const int32_t s_vd2[2] = {keep, (int32_t)m}, s_u2[2] = {(int32_t)n, keep};
    launch_permute(ctx->stream, vd2.data, u->data, 16, 2, s_vd2, pr);
    launch_conj(ctx->stream, u->data, u->size, HBP_C128);
    vsh[0] = keep;
    for (int d = nrows; d < a->order; ++d) vsh[d - nrows + 1] = a->shape[d];
    tensor_resize(ctx, vd, HBP_C128, a->order - nrows + 1, vsh);
    launch_permute(ctx->stream, u2.data, vd->data, 16, 2, s_u2, pr);
    launch_conj(ctx->stream, vd->data, vd->size, HBP_C128);
    ctx->release(ah.data, ah.bytes);
    ctx->release(u2.data, u2.bytes);
    ctx->release(vd2.data, vd2.bytes);
    return;
  }
  HBP_REQUIRE(n <= 512 && m <= (1 << 20), HBP_ERR_UNSUPPORTED, "svd: matrix too large");
  const int k = (int)std::min(m, n);
  double2 *U = (double2 *)ctx->alloc((size_t)m * k * 16);
  double *S = (double *)ctx->alloc((size_t)k * 8);
  double2 *Vd = (double2 *)ctx->alloc((size_t)k * n * 16);
  double2 *W = (double2 *)ctx->alloc((size_t)(m * n + n * n) * 16);
  SvdJob job{0, 0, 0, 0, 0, (int)m, (int)n};
  SvdJob *dj = (SvdJob *)ctx->alloc(sizeof(SvdJob));
  ctx->upload(dj, &job, sizeof(job));
  launch_batched_svd(ctx->stream, (const double2 *)a->data, U, S, Vd, W, dj, 1, (int)n, (int)m);
  std::vector<double> hs(k);
  HBP_HIP(hipMemcpyAsync(hs.data(), S, k * 8, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(W, (size_t)(m * n + n * n) * 16);
  ctx->release(dj, sizeof(SvdJob));
  int keep = k;
  if (keep_max >= 0) {
    int cnt = 0;
    for (int i = 0; i < k; ++i)
      if (hs[i] > s_min) ++cnt;
    keep = std::max(1, std::min(cnt, keep_max));
    double tot = 0.0, dis = 0.0;
    for (int i = 0; i < k; ++i) {
      tot += hs[i] * hs[i];
      if (i >= keep) dis += hs[i] * hs[i];
    }
    if (trunc_err) *trunc_err = tot > 0.0 ? dis / tot : 0.0;
  }
  // u (rows..., keep): the first keep columns of U
  int32_t ush[HBP_MAX_ORDER], vsh[HBP_MAX_ORDER];
  for (int d = 0; d < nrows; ++d) ush[d] = a->shape[d];
  ush[nrows] = keep;
  tensor_resize(ctx, u, HBP_C128, nrows + 1, ush);
  HBP_HIP(hipMemcpyAsync(u->data, U, (size_t)m * keep * 16, hipMemcpyDeviceToDevice, ctx->stream));
  const int32_t ssh[1] = {keep};
  tensor_resize(ctx, s, HBP_F64, 1, ssh);
  HBP_HIP(hipMemcpyAsync(s->data, S, (size_t)keep * 8, hipMemcpyDeviceToDevice, ctx->stream));
  vsh[0] = keep;
  for (int d = nrows; d < a->order; ++d) vsh[d - nrows + 1] = a->shape[d];
  tensor_resize(ctx, vd, HBP_C128, a->order - nrows + 1, vsh);
  if (keep == k) {
    HBP_HIP(hipMemcpyAsync(vd->data, Vd, (size_t)k * n * 16, hipMemcpyDeviceToDevice, ctx->stream));
  } else {
    HBP_HIP(hipMemcpy2DAsync(vd->data, (size_t)keep * 16, Vd, (size_t)k * 16, (size_t)keep * 16,
                             (size_t)n, hipMemcpyDeviceToDevice, ctx->stream));
  }
  ctx->release(U, (size_t)m * k * 16);
  ctx->release(S, (size_t)k * 8);
  ctx->release(Vd, (size_t)k * n * 16);
}

}  // namespace hbp

using namespace hbp;

// ======================================================================= C-ABI
extern "C" {

const char *hbp_last_error(void) { return g_last_error.c_str(); }
const char *hbp_version(void) { return "hbp 0.1 gfx950"; }

hbp_status hbp_ctx_create(int device, hbp_ctx **out) {
  HBP_API_BEGIN
  HBP_REQUIRE(out, HBP_ERR_INVALID_ARGUMENT, "null out");
  int n = 0;
  HBP_HIP(hipGetDeviceCount(&n));
  HBP_REQUIRE(device >= 0 && device < n, HBP_ERR_INVALID_ARGUMENT, "no such device");
  HBP_HIP(hipSetDevice(device));
  bind_noalloc_scratch(device);
  hbp_ctx *c = new hbp_ctx();
  c->device = device;
  HBP_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  hipMemPoolProps props{};
  props.allocType = hipMemAllocationTypePinned;
  props.location.type = hipMemLocationTypeDevice;
  props.location.id = device;
  HBP_HIP(hipMemPoolCreate(&c->pool, &props));
  uint64_t thr = UINT64_MAX;
  HBP_HIP(hipMemPoolSetAttribute(c->pool, hipMemPoolAttrReleaseThreshold, &thr));
  *out = c;
  HBP_API_END
}

hbp_status hbp_ctx_destroy(hbp_ctx *ctx) {
  HBP_API_BEGIN
  if (!ctx) return HBP_OK;
  HBP_HIP(hipSetDevice(ctx->device));
  if (ctx->scratch) ctx->release(ctx->scratch, ctx->scratch_bytes);
  if (ctx->arena) ctx->release(ctx->arena, ctx->arena_bytes);
  ctx->flush_blocks();
  HBP_HIP(hipStreamSynchronize(ctx->stream));
  ctx->prof_clear();
  if (ctx->stage) {
    for (int k = 0; k < Ctx::kStageSlots; ++k) (void)hipEventDestroy(ctx->stage_ev[k]);
    (void)hipHostFree(ctx->stage);
    ctx->stage = nullptr;
  }
  // give the reserve back before the pool goes, so that a context created
  // next in this process sees the memory as free
  HBP_HIP(hipMemPoolTrimTo(ctx->pool, 0));
  HBP_HIP(hipMemPoolDestroy(ctx->pool));
  HBP_HIP(hipStreamDestroy(ctx->stream));
  delete ctx;
  HBP_API_END
}

hbp_status hbp_ctx_synchronize(hbp_ctx *ctx) {
  HBP_API_BEGIN
  ctx->sync();
  HBP_API_END
}

hbp_status hbp_ctx_stream(hbp_ctx *ctx, void **stream) {
  HBP_API_BEGIN
  *stream = (void *)ctx->stream;
  HBP_API_END
}

hbp_status hbp_ctx_memory(hbp_ctx *ctx, size_t *in_use, size_t *peak) {
  HBP_API_BEGIN
  if (in_use) *in_use = ctx->in_use;
  if (peak) *peak = ctx->peak;
  HBP_API_END
}

hbp_status hbp_autotune_report(char *buf, size_t len, size_t *needed) {
  HBP_API_BEGIN
  const std::string r = autotune_report();
  if (needed) *needed = r.size() + 1;
  if (buf && len > 0) {
    const size_t n = std::min(len - 1, r.size());
    std::memcpy(buf, r.data(), n);
    buf[n] = 0;
  }
  HBP_API_END
}

hbp_status hbp_ctx_trim(hbp_ctx *ctx) {
  HBP_API_BEGIN
  if (ctx->scratch) {
    ctx->release(ctx->scratch, ctx->scratch_bytes);
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  if (ctx->arena) {
    ctx->release(ctx->arena, ctx->arena_bytes);
    ctx->arena = nullptr;
    ctx->arena_bytes = 0;
  }
  ctx->flush_blocks();
  ctx->sync();
  HBP_HIP(hipMemPoolTrimTo(ctx->pool, 0));
  ctx->peak = ctx->in_use;
  HBP_API_END
}

hbp_status hbp_ctx_graph_stats(hbp_ctx *ctx, int64_t *recorded, int64_t *replays,
                               int64_t *dropped) {
  HBP_API_BEGIN
  HBP_REQUIRE(ctx, HBP_ERR_INVALID_ARGUMENT, "null context");
  if (recorded) *recorded = ctx->graphs_recorded;
  if (replays) *replays = ctx->graph_replays;
  if (dropped) *dropped = ctx->graphs_dropped;
  HBP_API_END
}

hbp_status hbp_device_memory(int device, size_t *free_bytes, size_t *total_bytes) {
  HBP_API_BEGIN
  int cur = 0;
  HBP_HIP(hipGetDevice(&cur));
  HBP_HIP(hipSetDevice(device));
  size_t f = 0, t = 0;
  const hipError_t e = hipMemGetInfo(&f, &t);
  HBP_HIP(hipSetDevice(cur));
  HBP_HIP(e);
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  HBP_API_END
}

hbp_status hbp_ctx_profile(hbp_ctx *ctx, int enable) {
  HBP_API_BEGIN
  ctx->prof = enable != 0;
  // records not read by now are dropped (their events pin device memory)
  if (!ctx->prof) {
    ctx->sync();
    ctx->prof_clear();
  }
  HBP_API_END
}

hbp_status hbp_ctx_profile_read(hbp_ctx *ctx, int kind, int64_t *count, double *total_ms,
                                double *flops, double *bytes) {
  HBP_API_BEGIN
  ctx->sync();
  int64_t n = 0;
  double ms = 0.0, fl = 0.0, by = 0.0;
  std::vector<Ctx::ProfRec> keep;
  for (auto &r : ctx->prof_recs) {
    if (r.kind != kind) {
      keep.push_back(r);
      continue;
    }
    if (r.timed) {
      float t = 0.0f;
      HBP_HIP(hipEventElapsedTime(&t, r.a, r.b));
      ++n;
      ms += t;
      fl += r.flops;
      by += r.bytes;
    }
    HBP_HIP(hipEventDestroy(r.a));
    HBP_HIP(hipEventDestroy(r.b));
  }
  ctx->prof_recs.swap(keep);
  if (kind >= 0 && kind < 3) {
    double *acc = ctx->prof_acc[kind];
    n += (int64_t)acc[0];
    ms += acc[1];
    fl += acc[2];
    by += acc[3];
    for (int i = 0; i < 4; ++i) acc[i] = 0.0;
  }
  if (count) *count = n;
  if (total_ms) *total_ms = ms;
  if (flops) *flops = fl;
  if (bytes) *bytes = by;
  HBP_API_END
}

hbp_status hbp_tensor_create(hbp_ctx *ctx, int dtype, int order, const int32_t *shape,
                             hbp_tensor **out) {
  HBP_API_BEGIN
  (void)dtype_size(dtype);
  hbp_tensor *t = new hbp_tensor();
  try {
    tensor_resize(ctx, t, dtype, order, shape);
    launch_fill(ctx->stream, t->data, t->size, dtype, 0.0, 0.0);
  } catch (...) {
    delete t;
    throw;
  }
  *out = t;
  HBP_API_END
}

hbp_status hbp_tensor_destroy(hbp_ctx *ctx, hbp_tensor *t) {
  HBP_API_BEGIN
  if (!t) return HBP_OK;
  if (t->data) ctx->release(t->data, t->bytes);
  delete t;
  HBP_API_END
}

hbp_status hbp_tensor_info(const hbp_tensor *t, int *dtype, int *order, int32_t *shape,
                           size_t *size) {
  HBP_API_BEGIN
  if (dtype) *dtype = t->dtype;
  if (order) *order = t->order;
  if (shape)
    for (int d = 0; d < HBP_MAX_ORDER; ++d) shape[d] = t->shape[d];
  if (size) *size = t->size;
  HBP_API_END
}

hbp_status hbp_tensor_upload(hbp_ctx *ctx, hbp_tensor *t, const void *host) {
  HBP_API_BEGIN
  HBP_HIP(hipMemcpyAsync(t->data, host, t->size * dtype_size(t->dtype), hipMemcpyHostToDevice,
                         ctx->stream));
  ctx->sync();
  touch(t);
  HBP_API_END
}

hbp_status hbp_tensor_download(hbp_ctx *ctx, const hbp_tensor *t, void *host) {
  HBP_API_BEGIN
  HBP_HIP(hipMemcpyAsync(host, t->data, t->size * dtype_size(t->dtype), hipMemcpyDeviceToHost,
                         ctx->stream));
  ctx->sync();
  HBP_API_END
}

hbp_status hbp_tensor_copy(hbp_ctx *ctx, const hbp_tensor *src, hbp_tensor *dst) {
  HBP_API_BEGIN
  if (src == dst) return HBP_OK;
  tensor_resize(ctx, dst, src->dtype, src->order, src->shape);
  HBP_HIP(hipMemcpyAsync(dst->data, src->data, src->size * dtype_size(src->dtype),
                         hipMemcpyDeviceToDevice, ctx->stream));
  HBP_API_END
}

hbp_status hbp_tensor_reshape(hbp_ctx *ctx, hbp_tensor *t, int order, const int32_t *shape) {
  HBP_API_BEGIN
  (void)ctx;
  HBP_REQUIRE(order >= 0 && order <= HBP_MAX_ORDER, HBP_ERR_INVALID_ARGUMENT, "bad order");
  HBP_REQUIRE(shape_size(order, shape) == t->size, HBP_ERR_INVALID_ARGUMENT,
              "reshape must keep the number of elements");
  t->order = order;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) t->shape[d] = d < order ? shape[d] : 0;
  touch(t);
  HBP_API_END
}

hbp_status hbp_transpose(hbp_ctx *ctx, const hbp_tensor *in, const int32_t *new_order,
                         hbp_tensor *out) {
  HBP_API_BEGIN
  std::vector<int> seen(in->order, 0);
  for (int d = 0; d < in->order; ++d) {
    HBP_REQUIRE(new_order[d] >= 0 && new_order[d] < in->order && !seen[new_order[d]],
                HBP_ERR_INVALID_ARGUMENT, "transpose: not a permutation");
    seen[new_order[d]] = 1;
  }
  int32_t sh[HBP_MAX_ORDER];
  for (int d = 0; d < in->order; ++d) sh[d] = in->shape[new_order[d]];
  const size_t bytes = in->size * dtype_size(in->dtype);
  void *buf = ctx->alloc(bytes);
  launch_permute(ctx->stream, in->data, buf, (int)dtype_size(in->dtype), in->order, in->shape,
                 new_order);
  if (out->data) ctx->release(out->data, out->bytes);
  out->data = buf;
  out->bytes = bytes;
  out->dtype = in->dtype;
  out->size = in->size;
  out->order = in->order;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) out->shape[d] = d < in->order ? sh[d] : 0;
  touch(out);
  HBP_API_END
}

hbp_status hbp_conj(hbp_ctx *ctx, hbp_tensor *t) {
  HBP_API_BEGIN
  launch_conj(ctx->stream, t->data, t->size, t->dtype);
  touch(t);
  HBP_API_END
}

hbp_status hbp_fill(hbp_ctx *ctx, hbp_tensor *t, double re, double im) {
  HBP_API_BEGIN
  launch_fill(ctx->stream, t->data, t->size, t->dtype, re, im);
  touch(t);
  HBP_API_END
}

hbp_status hbp_contract(hbp_ctx *ctx, const hbp_tensor *a, const int32_t *la,
                        const hbp_tensor *b, const int32_t *lb, hbp_tensor *c,
                        const int32_t *lc, int nc) {
  HBP_API_BEGIN
  contract_impl(ctx, a, la, b, lb, c, lc, nc);
  touch(c);
  HBP_API_END
}

hbp_status hbp_norm(hbp_ctx *ctx, const hbp_tensor *t, double *out) {
  HBP_API_BEGIN
  *out = norm_sync(ctx, t);
  HBP_API_END
}

hbp_status hbp_normalize(hbp_ctx *ctx, hbp_tensor *t, double *norm_out) {
  HBP_API_BEGIN
  double *dv = (double *)ctx->alloc(sizeof(double));
  device_norm2(ctx, t->data, t->size, t->dtype, dv);
  launch_scale(ctx->stream, t->data, t->size, t->dtype, 1.0, 0.0, dv);
  if (norm_out) {
    double h = 0.0;
    HBP_HIP(hipMemcpyAsync(&h, dv, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    *norm_out = std::sqrt(h);
  }
  ctx->release(dv, sizeof(double));
  touch(t);
  HBP_API_END
}

hbp_status hbp_scale(hbp_ctx *ctx, hbp_tensor *t, double re, double im) {
  HBP_API_BEGIN
  launch_scale(ctx->stream, t->data, t->size, t->dtype, re, im, nullptr);
  touch(t);
  HBP_API_END
}

hbp_status hbp_linear_combine(hbp_ctx *ctx, int n, const hbp_tensor *const *ins,
                              const double *coefs, hbp_tensor *out) {
  HBP_API_BEGIN
  HBP_REQUIRE(n >= 1, HBP_ERR_INVALID_ARGUMENT, "linear_combine: no inputs");
  int dt = ins[0]->dtype;
  for (int i = 0; i < n; ++i) {
    HBP_REQUIRE(ins[i]->size == ins[0]->size, HBP_ERR_INVALID_ARGUMENT,
                "linear_combine: shapes differ");
    dt = common_dtype(dt, ins[i]->dtype);
  }
  // the coefficients are complex: the result is complex in any case
  dt = single_dtype(dt) ? HBP_C64 : HBP_C128;
  const size_t bytes = ins[0]->size * 16;
  double2 *acc = (double2 *)ctx->alloc(bytes);
  launch_fill(ctx->stream, acc, ins[0]->size, HBP_C128, 0.0, 0.0);
  for (int i = 0; i < n; ++i) {
    Tensor ti;
    const Tensor *x = as_c128(ctx, ins[i], &ti);
    launch_axpy_c128(ctx->stream, (const double2 *)x->data, acc, ins[0]->size,
                     coefs[2 * i], coefs[2 * i + 1]);
    drop_tmp(ctx, &ti);
  }
  const int order = ins[0]->order;
  int32_t sh[HBP_MAX_ORDER];
  for (int d = 0; d < order; ++d) sh[d] = ins[0]->shape[d];
  if (out->data) ctx->release(out->data, out->bytes);
  out->data = acc;
  out->bytes = bytes;
  out->dtype = HBP_C128;
  out->order = order;
  out->size = ins[0]->size;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) out->shape[d] = d < order ? sh[d] : 0;
  round_to(ctx, out, dt);
  touch(out);
  HBP_API_END
}

hbp_status hbp_trace(hbp_ctx *ctx, const hbp_tensor *in, int npairs, const int32_t *pairs,
                     hbp_tensor *out) {
  HBP_API_BEGIN
  const int dt = in->dtype;
  Tensor tin;
  const Tensor *x = as_c128(ctx, in, &tin);
  std::vector<int> used(in->order, 0);
  for (int p = 0; p < npairs; ++p) {
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    HBP_REQUIRE(i >= 0 && j >= 0 && i < in->order && j < in->order && i != j && !used[i] &&
                    !used[j] && in->shape[i] == in->shape[j],
                HBP_ERR_INVALID_ARGUMENT, "trace: bad bond pair");
    used[i] = used[j] = 1;
  }
  int32_t sh[HBP_MAX_ORDER];
  int order = 0;
  for (int d = 0; d < in->order; ++d)
    if (!used[d]) sh[order++] = in->shape[d];
  const size_t n = shape_size(order, sh);
  double2 *buf = (double2 *)ctx->alloc(n * 16);
  launch_trace(ctx->stream, (const double2 *)x->data, buf, in->order, in->shape, npairs, pairs);
  drop_tmp(ctx, &tin);
  if (out->data) ctx->release(out->data, out->bytes);
  out->data = buf;
  out->bytes = n * 16;
  out->dtype = HBP_C128;
  out->order = order;
  out->size = n;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) out->shape[d] = d < order ? sh[d] : 0;
  round_to(ctx, out, dt);
  touch(out);
  HBP_API_END
}

hbp_status hbp_diag(hbp_ctx *ctx, const hbp_tensor *vec, hbp_tensor *out) {
  HBP_API_BEGIN
  const int64_t n = (int64_t)vec->size;
  HBP_REQUIRE(n <= (1 << 15), HBP_ERR_UNSUPPORTED, "diag: vector too long");
  // a real vector into a complex tensor gives the complex matrix (tci::diag
  // then to_cplx in one call)
  const int dtype = (!dtype_complex(vec->dtype) && dtype_complex(out->dtype))
                        ? (vec->dtype == HBP_F32 ? HBP_C64 : HBP_C128)
                        : vec->dtype;
  const int32_t sh[2] = {(int32_t)n, (int32_t)n};
  const size_t es = dtype_size(dtype);
  void *buf = ctx->alloc((size_t)n * n * es);
  launch_fill(ctx->stream, buf, (size_t)n * n, dtype, 0.0, 0.0);
  // scatter the vector onto the diagonal: element i goes to i + n i
  void *conv = ctx->alloc((size_t)n * es);
  launch_convert(ctx->stream, vec->data, vec->dtype, conv, dtype, (size_t)n);
  HBP_HIP(hipMemcpy2DAsync(buf, (n + 1) * es, conv, es, es, (size_t)n, hipMemcpyDeviceToDevice,
                           ctx->stream));
  ctx->release(conv, (size_t)n * es);
  if (out->data && out != vec) ctx->release(out->data, out->bytes);
  if (out == vec && out->data) ctx->release(out->data, out->bytes);
  out->data = buf;
  out->bytes = (size_t)n * n * es;
  out->dtype = dtype;
  out->order = 2;
  out->size = (size_t)n * n;
  for (int d = 0; d < HBP_MAX_ORDER; ++d) out->shape[d] = d < 2 ? sh[d] : 0;
  touch(out);
  HBP_API_END
}

hbp_status hbp_get_elem(hbp_ctx *ctx, const hbp_tensor *t, const int32_t *coors, double *re,
                        double *im) {
  HBP_API_BEGIN
  size_t off = 0, stride = 1;
  for (int d = 0; d < t->order; ++d) {
    HBP_REQUIRE(coors[d] >= 0 && coors[d] < t->shape[d], HBP_ERR_OUT_OF_RANGE,
                "get_elem: coordinate out of range");
    off += (size_t)coors[d] * stride;
    stride *= (size_t)t->shape[d];
  }
  double2 v = make_double2(0.0, 0.0);
  const size_t es = dtype_size(t->dtype);
  unsigned char host[16] = {0};
  HBP_HIP(hipMemcpyAsync(host, (const unsigned char *)t->data + off * es, es,
                         hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  switch (t->dtype) {
    case HBP_C128: std::memcpy(&v, host, 16); break;
    case HBP_F64: std::memcpy(&v.x, host, 8); break;
    case HBP_C64: {
      float f[2];
      std::memcpy(f, host, 8);
      v = make_double2(f[0], f[1]);
      break;
    }
    default: {
      float f;
      std::memcpy(&f, host, 4);
      v.x = f;
    }
  }
  *re = v.x;
  *im = v.y;
  HBP_API_END
}

// svd_impl on a complex copy of a complex64 tensor; u, vdag rounded to
// complex64 and s to float32 (tensor_traits of complex<float>: real_ten_t
// float).  Real tensors are refused (the reference decomposes complex ones).
static void svd_any(Ctx *ctx, const Tensor *a, int nrows, Tensor *u, Tensor *s, Tensor *vd,
                    int keep_max, double s_min, double *trunc_err) {
  HBP_REQUIRE(dtype_complex(a->dtype), HBP_ERR_UNSUPPORTED, "svd: complex tensors only");
  Tensor ta;
  svd_impl(ctx, as_c128(ctx, a, &ta), nrows, u, s, vd, keep_max, s_min, trunc_err);
  drop_tmp(ctx, &ta);
  if (a->dtype == HBP_C64) {
    round_to(ctx, u, HBP_C64);
    round_to(ctx, vd, HBP_C64);
    round_to(ctx, s, HBP_F32);
  }
}

hbp_status hbp_svd(hbp_ctx *ctx, const hbp_tensor *a, int nrows, hbp_tensor *u, hbp_tensor *s,
                   hbp_tensor *vdag) {
  HBP_API_BEGIN
  svd_any(ctx, a, nrows, u, s, vdag, -1, 0.0, nullptr);
  HBP_API_END
}

hbp_status hbp_svd_noalloc(hbp_ctx *ctx, const void *a, int m, int n, void *u, double *s,
                           void *vdag) {
  HBP_API_BEGIN
  HBP_REQUIRE(ctx && a && u && s && vdag, HBP_ERR_INVALID_ARGUMENT, "svd_noalloc: null argument");
  HBP_REQUIRE(m >= 1 && n >= 1, HBP_ERR_INVALID_ARGUMENT, "svd_noalloc: empty matrix");
  HBP_REQUIRE(m >= n, HBP_ERR_UNSUPPORTED,
              "svd_noalloc: needs rows >= columns (pass the conjugate transpose and swap the "
              "factors, or use hbp_svd)");
  HBP_REQUIRE(n <= 512, HBP_ERR_UNSUPPORTED, "svd_noalloc: more than 512 columns");
  const size_t work = ((size_t)m * n + (size_t)n * n) * sizeof(double2);
  HBP_REQUIRE(work <= kNoAllocScratchBytes - kNoAllocJobBytes, HBP_ERR_UNSUPPORTED,
              "svd_noalloc: (m n + n^2) complex128 work exceeds its 16 MiB device scratch "
              "(m = " + std::to_string(m) + ", n = " + std::to_string(n) + ")");
  HBP_REQUIRE(ctx->device >= 0 && ctx->device < kMaxDevices && g_scratch_ptr[ctx->device],
              HBP_ERR_LOGIC, "svd_noalloc: device scratch not bound");
  HBP_HIP(hipSetDevice(ctx->device));
  char *base = (char *)g_scratch_ptr[ctx->device];
  SvdJob *dj = (SvdJob *)base;
  double2 *W = (double2 *)(base + kNoAllocJobBytes);
  // the kernel initialises every work element it reads (A's copy, V = I)
  launch_store_svd_job(ctx->stream, SvdJob{0, 0, 0, 0, 0, m, n}, dj);
  launch_batched_svd(ctx->stream, (const double2 *)a, (double2 *)u, s, (double2 *)vdag, W, dj, 1,
                     n, m);
  HBP_API_END
}

hbp_status hbp_tensor_data(const hbp_tensor *t, void **ptr) {
  HBP_API_BEGIN
  HBP_REQUIRE(t && ptr, HBP_ERR_INVALID_ARGUMENT, "tensor_data: null argument");
  *ptr = t->data;
  HBP_API_END
}

hbp_status hbp_tensor_version(const hbp_tensor *t, uint64_t *version) {
  HBP_API_BEGIN
  HBP_REQUIRE(t && version, HBP_ERR_INVALID_ARGUMENT, "tensor_version: null argument");
  *version = t->version;
  HBP_API_END
}

hbp_status hbp_trunc_svd(hbp_ctx *ctx, const hbp_tensor *a, int nrows, int chi_max,
                         double s_min, hbp_tensor *u, hbp_tensor *s, hbp_tensor *vdag,
                         double *trunc_err) {
  HBP_API_BEGIN
  HBP_REQUIRE(chi_max >= 1, HBP_ERR_INVALID_ARGUMENT, "trunc_svd: chi_max < 1");
  svd_any(ctx, a, nrows, u, s, vdag, chi_max, s_min, trunc_err);
  HBP_API_END
}

}  // extern "C"
