// PEPS layer: the state (site tensors V, messages E) lives in HBM and each
// tnbp algorithm step runs as a handful of kernel launches.
//
//   AbsorbTPO            apply.h:14-177        k_absorb_site, k_expand_messages
//   BeliefPropagation    bp.h:65-261           mode products + message kernel
//   BP condition         bp.h:266-355          reused from the next sweep (same messages)
//   Truncation           truncation.h:17-222   batched Jacobi SVD, k_roots, k_rr, k_projectors
//   Measure (site/edge)  measure.h:14-300      mode products + message kernel, k_site_ratio,
//                                              k_edge_ratio
//
// Normalisation: the reference normalises T after every absorbed bond
// (bp.h:153).  Those are positive scalings of a tensor whose final message is
// normalised again (bp.h:166), so they change rounding only; the fused path
// normalises each message once, in FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>
#include <type_traits>
#include <vector>

#include "hbp_internal.h"

namespace hbp {
void set_last_error(const std::string &s);
void svd_impl(Ctx *ctx, const Tensor *a, int nrows, Tensor *u, Tensor *s, Tensor *vd,
              int keep_max, double s_min, double *trunc_err);
double device_norm2(Ctx *ctx, const void *x, size_t n, int dtype, double *dev_out);

namespace {

__device__ inline double wsum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}
__device__ inline double bsum(double v, double *sh) {
  v = wsum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += sh[w];
  __syncthreads();
  return t;
}

struct MsgJob {
  int64_t raw_off;   // raw message (complex<double>) in the raw buffer
  int64_t new_off;   // destination in the new message set
  int64_t old_off;   // stored message to compare against (residual)
  int D;
};

// unit complex c with c * tr = |tr| (1 when the trace vanishes)
__device__ inline double2 phase_of_trace(double re, double im) {
  const double a = sqrt(re * re + im * im);
  return a > 0.0 ? make_double2(re / a, -im / a) : make_double2(1.0, 0.0);
}

// normalise every raw message (FP64) and, when old != null, the Frobenius
// distance between it and the normalised stored message (bp.h:321-350).
// phase_gauge != 0 also fixes each message's free scalar phase (trace real
// and positive) on both sides of the distance.
// Normalise one raw message (FP64) into out (and out2 when given) and, when
// old != null, return the Frobenius distance between it and the normalised
// stored message (bp.h:321-350); phase_gauge != 0 also fixes each message's
// free scalar phase (trace real and positive) on both sides of the distance.
// Called by every thread of a block; the result is valid in thread 0.
__device__ double finalize_one(const double2 *__restrict__ raw, double2 *__restrict__ out,
                               const double2 *__restrict__ old, int D, int phase_gauge,
                               double2 *__restrict__ out2, double *sh) {
  const int n = D * D;
  double s = 0.0, so = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const double2 v = raw[e];
    s += v.x * v.x + v.y * v.y;
    if (old) {
      const double2 w = old[e];
      so += w.x * w.x + w.y * w.y;
    }
  }
  s = bsum(s, sh);
  if (old) so = bsum(so, sh);
  double2 c = make_double2(1.0, 0.0), co = make_double2(1.0, 0.0);
  if (phase_gauge) {
    double tr = 0.0, ti = 0.0, tor = 0.0, toi = 0.0;
    for (int i = threadIdx.x; i < D; i += blockDim.x) {
      const double2 v = raw[(int64_t)i * (D + 1)];
      tr += v.x;
      ti += v.y;
      if (old) {
        const double2 w = old[(int64_t)i * (D + 1)];
        tor += w.x;
        toi += w.y;
      }
    }
    tr = bsum(tr, sh);
    ti = bsum(ti, sh);
    c = phase_of_trace(tr, ti);
    if (old) {
      tor = bsum(tor, sh);
      toi = bsum(toi, sh);
      co = phase_of_trace(tor, toi);
    }
  }
  const double inv = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
  const double invo = so > 0.0 ? 1.0 / sqrt(so) : 1.0;
  c = make_double2(c.x * inv, c.y * inv);
  co = make_double2(co.x * invo, co.y * invo);
  double d = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) {
    const double2 v = raw[e];
    const double2 m = make_double2(v.x * c.x - v.y * c.y, v.x * c.y + v.y * c.x);
    out[e] = m;
    if (out2) out2[e] = m;
    if (old) {
      const double2 w = old[e];
      const double dx = m.x - (w.x * co.x - w.y * co.y), dy = m.y - (w.x * co.y + w.y * co.x);
      d += dx * dx + dy * dy;
    }
  }
  if (old) d = bsum(d, sh);
  return sqrt(d);
}

// normalise every raw message of one PEPS (offsets into the raw, out and old
// buffers per job) and the distance of each to its stored message
__global__ __launch_bounds__(256) void k_finalize_messages(const double2 *__restrict__ raw,
                                                            double2 *__restrict__ out,
                                                            const double2 *__restrict__ old,
                                                            const MsgJob *__restrict__ jobs,
                                                            double *__restrict__ dist,
                                                            int phase_gauge,
                                                            double2 *__restrict__ out2) {
  __shared__ double sh[8];
  const MsgJob j = jobs[blockIdx.x];
  const double d = finalize_one(raw + j.raw_off, out + j.new_off, old ? old + j.old_off : nullptr,
                                j.D, phase_gauge, out2 ? out2 + j.new_off : nullptr, sh);
  if (old && threadIdx.x == 0) dist[blockIdx.x] = d;
}

// The same for the messages of several PEPS (a job per message, with its
// own buffers), each PEPS's residual its distances' mean (bp.h:342-350).
struct MsgJobB {
  const double2 *raw;
  double2 *out;
  const double2 *old;
  int D, phase_gauge;
};
__global__ __launch_bounds__(256) void k_finalize_batched(const MsgJobB *__restrict__ jobs,
                                                           double *__restrict__ dist) {
  __shared__ double sh[8];
  const MsgJobB j = jobs[blockIdx.x];
  const double d = finalize_one(j.raw, j.out, j.old, j.D, j.phase_gauge, nullptr, sh);
  if (j.old && threadIdx.x == 0) dist[blockIdx.x] = d;
}
// res[p] = vol[p] * sum of dist[first[p] .. first[p + 1]) in a fixed order
__global__ void k_residual_batched(const double *__restrict__ dist, const int *__restrict__ first,
                                   const double *__restrict__ vol, double *__restrict__ res) {
  __shared__ double sh[8];
  const int p = blockIdx.x;
  double s = 0.0;
  for (int i = first[p] + threadIdx.x; i < first[p + 1]; i += blockDim.x) s += dist[i] * vol[p];
  s = bsum(s, sh);
  if (threadIdx.x == 0) res[p] = s;
}

__global__ void k_scatter_dist(const double *__restrict__ d, const int *__restrict__ pos, int n,
                               double *__restrict__ dist) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) dist[pos[i]] = d[i];
}

__global__ void k_residual(const double *__restrict__ dist, int n, double vol,
                           double *__restrict__ out) {
  __shared__ double sh[8];
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += dist[i] * vol;
  s = bsum(s, sh);
  if (threadIdx.x == 0) out[0] = s;
}

// ------------------------------------------------ Anderson mixing (BP loop)
// The message set is one vector of n complex numbers (all 2 n_e messages,
// contiguous).  Up to AA_MAX + 1 vectors of the mixing window per call.
constexpr int AA_MAX = 8;
constexpr int AA_BLOCKS = 512;
struct AAVecs {
  const double2 *v[AA_MAX + 1];
  double c[AA_MAX + 1];
  int n;
};

// f = G - E, and per block the partial sums of Re<f, h_i> for the window's
// residual vectors h_0 .. h_{nh-1} followed by <f, f>:
// part[(nh + 1) * block + i]; the blocks' partials are added by k_aa_sum.
__global__ __launch_bounds__(256) void k_aa_diff_dots(const double2 *__restrict__ G,
                                                       const double2 *__restrict__ E,
                                                       double2 *__restrict__ f, int64_t n,
                                                       AAVecs h, double *__restrict__ part) {
  __shared__ double sh[8];
  double s[AA_MAX + 2];
  for (int i = 0; i <= AA_MAX + 1; ++i) s[i] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double2 g = G[e], x = E[e];
    const double2 d = make_double2(g.x - x.x, g.y - x.y);
    f[e] = d;
#pragma unroll
    for (int i = 0; i < AA_MAX + 1; ++i)
      if (i < h.n) {
        const double2 w = h.v[i][e];
        s[i] += d.x * w.x + d.y * w.y;
      }
    s[AA_MAX + 1] += d.x * d.x + d.y * d.y;
  }
#pragma unroll
  for (int i = 0; i < AA_MAX + 2; ++i) {
    if (i > h.n) break;  // uniform over the block
    const double t = bsum(i < h.n ? s[i] : s[AA_MAX + 1], sh);
    if (threadIdx.x == 0) part[(int64_t)(h.n + 1) * blockIdx.x + i] = t;
  }
}

// out[i] = sum over the blocks of part[(m) * b + i], in block order
__global__ void k_aa_sum(const double *__restrict__ part, int m, int blocks,
                         double *__restrict__ out) {
  const int i = threadIdx.x;
  if (i >= m) return;
  double s = 0.0;
  for (int b = 0; b < blocks; ++b) s += part[(int64_t)m * b + i];
  out[i] = s;
}

// out = sum_i c_i v_i (real coefficients: Hermitian messages stay Hermitian)
__global__ __launch_bounds__(256) void k_aa_combine(AAVecs g, double2 *__restrict__ out,
                                                     int64_t n) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    double re = 0.0, im = 0.0;
    for (int i = 0; i < g.n; ++i) {
      const double2 v = g.v[i][e];
      re += g.c[i] * v.x;
      im += g.c[i] * v.y;
    }
    out[e] = make_double2(re, im);
  }
}

// ------------------------------------------------------------- absorb
struct AbsorbArgs {
  int z;
  int64_t vstride[HBP_MAX_ORDER];  // strides of V (bonds then phys)
  int vdim[HBP_MAX_ORDER];         // D_k of V
  int64_t ostride[HBP_MAX_ORDER];  // strides of O (bonds, d_in, d_out)
  int odim[HBP_MAX_ORDER];         // t_k
  int64_t nstride[HBP_MAX_ORDER];  // strides of V' (fused bonds, d_out)
  int ndim[HBP_MAX_ORDER];         // D_k t_k, then d_out
  int din, dout;
  // n < 2^32: the index is split with 32-bit division by invariant integers
  // (multiply-high and shift, magic m and shift s per divisor); 64-bit
  // division had made this HBM-bound kernel instruction-bound (~4x slower)
  int fast;
  uint32_t nm[HBP_MAX_ORDER], ns[HBP_MAX_ORDER];  // for ndim
  uint32_t vm[HBP_MAX_ORDER], vs[HBP_MAX_ORDER];  // for vdim
};

// Division of any 32-bit n by an invariant d >= 1 (Granlund and Montgomery,
// "Division by invariant integers using multiplication", 1994):
// s = ceil(log2 d), m = floor(2^32 (2^s - d) / d) + 1, n / d = (mulhi(n, m) + n) >> s
// with the sum taken in 64 bits.
static void fast_div_params(uint32_t d, uint32_t *m, uint32_t *s) {
  uint32_t sh = 0;
  while ((uint64_t(1) << sh) < d) ++sh;
  *m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << sh) - d)) / d + 1);
  *s = sh;
}
__device__ __forceinline__ uint32_t fast_div(uint32_t n, uint32_t m, uint32_t s) {
  return (uint32_t)(((uint64_t)__umulhi(n, m) + n) >> s);
}

// V'[(v o)..., p'] = sum_p O[o..., p, p'] V[v..., p]  (apply.h:52-66), element e
template <typename C>
__device__ __forceinline__ C absorb_elem(const C *__restrict__ V, const double2 *__restrict__ O,
                                         size_t e, const AbsorbArgs &a) {
  {
    int64_t vb = 0, ob = 0;
    int pp;
    if (a.fast) {
      uint32_t rem = (uint32_t)e;
      for (int k = 0; k < a.z; ++k) {
        const uint32_t q = fast_div(rem, a.nm[k], a.ns[k]);
        const uint32_t f = rem - q * (uint32_t)a.ndim[k];
        rem = q;
        const uint32_t o = fast_div(f, a.vm[k], a.vs[k]);
        const uint32_t v = f - o * (uint32_t)a.vdim[k];
        vb += (int64_t)v * a.vstride[k];
        ob += (int64_t)o * a.ostride[k];
      }
      pp = (int)rem;
    } else {
      size_t rem = e;
      for (int k = 0; k < a.z; ++k) {
        const int f = (int)(rem % a.ndim[k]);
        rem /= a.ndim[k];
        const int v = f % a.vdim[k], o = f / a.vdim[k];
        vb += v * a.vstride[k];
        ob += o * a.ostride[k];
      }
      pp = (int)rem;
    }
    ob += pp * a.ostride[a.z + 1];
    double re = 0.0, im = 0.0;
    for (int p = 0; p < a.din; ++p) {
      const C v = V[vb + p * a.vstride[a.z]];
      const double2 o = O[ob + p * a.ostride[a.z]];
      re += o.x * (double)v.x - o.y * (double)v.y;
      im += o.x * (double)v.y + o.y * (double)v.x;
    }
    C out;
    out.x = re;
    out.y = im;
    return out;
  }
}

template <typename C>
__global__ void k_absorb_site(const C *__restrict__ V, const double2 *__restrict__ O,
                              C *__restrict__ Vn, size_t n, AbsorbArgs a) {
  for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (size_t)gridDim.x * blockDim.x)
    Vn[e] = absorb_elem(V, O, e, a);
}

// Every site of an absorption in one launch: job j owns blocks
// [first[j], first[j + 1]) of the grid, ABS_EPB elements each (the last one
// partial); a block finds its job by binary search over first (first[0] = 0,
// first[njobs] = the grid size, each job at least one block).
constexpr int ABS_EPB = 1024;
struct AbsorbJob {
  const void *V;
  const double2 *O;
  void *Vn;
  size_t n;
  AbsorbArgs a;
};
template <typename C>
__global__ __launch_bounds__(256) void k_absorb_batch(const AbsorbJob *__restrict__ jobs,
                                                      const int *__restrict__ first, int njobs) {
  const int b = blockIdx.x;
  int lo = 0, hi = njobs;  // first[lo] <= b < first[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid;
  }
  const AbsorbJob &j = jobs[lo];
  const size_t e0 = (size_t)(b - first[lo]) * ABS_EPB;
  const C *V = (const C *)j.V;
  C *Vn = (C *)j.Vn;
  for (int i = threadIdx.x; i < ABS_EPB; i += blockDim.x) {
    const size_t e = e0 + i;
    if (e < j.n) Vn[e] = absorb_elem(V, j.O, e, j.a);
  }
}

// E'[(e0 i0), (e1 i1)] = E[e0, e1] delta(i0, i1)  (apply.h:110-128)
// E_e <- E_e (x) I_t for every message in one launch: block b expands
// message jobs[b] (apply.h:110-128), element (e0 + D i0, e1 + D i1) of the
// (D t) x (D t) result = E[e0, e1] if i0 == i1 else 0.
struct ExpandJob {
  const double2 *src;  // the old message (D x D)
  double2 *dst;        // the new one (D t x D t)
  int D, t;
};
__global__ __launch_bounds__(256) void k_expand_messages(const ExpandJob *__restrict__ jobs) {
  const ExpandJob j = jobs[blockIdx.x];
  const int D = j.D, Dn = j.D * j.t;
  for (int e = threadIdx.x; e < Dn * Dn; e += blockDim.x) {
    const int r = e % Dn, c = e / Dn;
    const int e0 = r % D, i0 = r / D, e1 = c % D, i1 = c / D;
    j.dst[e] = (i0 == i1) ? j.src[e0 + D * e1] : make_double2(0.0, 0.0);
  }
}

// Copies of n-element blocks to two places each (after truncation both
// directed messages of an edge are diag(Lambda)): block b does jobs[b].
struct CopyJob {
  int64_t src;         // offset in src
  double2 *dst0, *dst1;
  int n;
};
__global__ __launch_bounds__(256) void k_copy_twice(const double2 *__restrict__ src,
                                                    const CopyJob *__restrict__ jobs) {
  const CopyJob j = jobs[blockIdx.x];
  for (int e = threadIdx.x; e < j.n; e += blockDim.x) {
    const double2 v = src[j.src + e];
    j.dst0[e] = v;
    j.dst1[e] = v;
  }
}

// ---------------------------------------------------------- truncation
struct RootJob {
  int64_t u_off, s_off, vd_off;  // SVD of the message
  int64_t r_off, si_off;         // outputs R, S (D x D)
  int D;
  double sv_min;                 // of the PEPS the message belongs to
};

// SquareRootAndInverse (root.h:65-148): R = U diag(d) Vd, S = Vd^H diag(f) U^H
// with d = sqrt(e) if e > sv_min else 0, f = d / (d d) if d > sqrt(sv_min) else 0.
__global__ __launch_bounds__(256) void k_roots(const double2 *__restrict__ U,
                                                const double *__restrict__ sv,
                                                const double2 *__restrict__ Vd,
                                                double2 *__restrict__ out,
                                                const RootJob *__restrict__ jobs) {
  const RootJob j = jobs[blockIdx.x];
  const int D = j.D;
  const double sv_min = j.sv_min;
  __shared__ double dd[512], ff[512];
  for (int c = threadIdx.x; c < D; c += blockDim.x) {
    const double e = sv[j.s_off + c];
    const double d = fabs(e) > sv_min ? sqrt(e) : 0.0;
    dd[c] = d;
    ff[c] = fabs(d) > sqrt(sv_min) ? d / (d * d) : 0.0;
  }
  __syncthreads();
  const double2 *u = U + j.u_off;
  const double2 *v = Vd + j.vd_off;
  for (int e = threadIdx.x; e < D * D; e += blockDim.x) {
    const int r = e % D, c = e / D;
    double rr = 0.0, ri = 0.0, sr = 0.0, si = 0.0;
    for (int k = 0; k < D; ++k) {
      // R[r,c] = sum_k U[r,k] d_k Vd[k,c]
      const double2 a = u[r + (int64_t)D * k], b = v[k + (int64_t)D * c];
      rr += dd[k] * (a.x * b.x - a.y * b.y);
      ri += dd[k] * (a.x * b.y + a.y * b.x);
      // S[r,c] = sum_k conj(Vd[k,r]) f_k conj(U[c,k])
      const double2 x = v[k + (int64_t)D * r], y = u[c + (int64_t)D * k];
      sr += ff[k] * (x.x * y.x - x.y * y.y);
      si += -ff[k] * (x.x * y.y + x.y * y.x);
    }
    out[j.r_off + e] = make_double2(rr, ri);
    out[j.si_off + e] = make_double2(sr, si);
  }
}

struct EdgeJob {
  int64_t ra_off, sa_off, rb_off, sb_off;  // in the roots buffer
  int64_t t_off;                           // T = Ra^T Rb (D x D)
  int64_t u_off, s_off, vd_off;            // SVD of T
  int64_t pa_off, pb_off, e_off;           // projectors and new messages
  int D;
  int chi;                                 // chi_max and s_min of the PEPS
  double s_min;
};

// T = Ra^T Rb, normalised (truncation.h:110-121)
__global__ __launch_bounds__(256) void k_rr(const double2 *__restrict__ roots,
                                             double2 *__restrict__ T,
                                             const EdgeJob *__restrict__ jobs) {
  __shared__ double sh[8];
  const EdgeJob j = jobs[blockIdx.x];
  const int D = j.D;
  const double2 *ra = roots + j.ra_off, *rb = roots + j.rb_off;
  double s = 0.0;
  for (int e = threadIdx.x; e < D * D; e += blockDim.x) {
    const int r = e % D, c = e / D;
    double xr = 0.0, xi = 0.0;
    for (int k = 0; k < D; ++k) {
      const double2 a = ra[k + (int64_t)D * r], b = rb[k + (int64_t)D * c];
      xr += a.x * b.x - a.y * b.y;
      xi += a.x * b.y + a.y * b.x;
    }
    T[j.t_off + e] = make_double2(xr, xi);
    s += xr * xr + xi * xi;
  }
  s = bsum(s, sh);
  const double inv = s > 0.0 ? 1.0 / sqrt(s) : 1.0;
  __syncthreads();
  for (int e = threadIdx.x; e < D * D; e += blockDim.x) {
    const double2 v = T[j.t_off + e];
    T[j.t_off + e] = make_double2(v.x * inv, v.y * inv);
  }
}

// trunc_svd(T, chi, s_min) selection, E = diag(L) for both directions,
// Pa = Sa^T X sqrt(L) (D x keep), Pb = sqrt(L) Y Sb (keep x D)  (truncation.h:123-217)
__global__ __launch_bounds__(256) void k_projectors(const double2 *__restrict__ roots,
                                                     const double2 *__restrict__ U,
                                                     const double *__restrict__ sv,
                                                     const double2 *__restrict__ Vd,
                                                     double2 *__restrict__ proj,
                                                     const EdgeJob *__restrict__ jobs,
                                                     int *__restrict__ keep_out,
                                                     double *__restrict__ err_out) {
  const EdgeJob j = jobs[blockIdx.x];
  const int D = j.D, chi = j.chi;
  const double s_min = j.s_min;
  __shared__ int keep_sh;
  __shared__ double sq[512];
  if (threadIdx.x == 0) {
    int cnt = 0;
    double tot = 0.0, dis = 0.0;
    for (int c = 0; c < D; ++c)
      if (sv[j.s_off + c] > s_min) ++cnt;
    int keep = cnt < chi ? cnt : chi;
    if (keep < 1) keep = 1;
    for (int c = 0; c < D; ++c) {
      const double s = sv[j.s_off + c];
      tot += s * s;
      if (c >= keep) dis += s * s;
    }
    keep_sh = keep;
    keep_out[blockIdx.x] = keep;
    err_out[blockIdx.x] = tot > 0.0 ? dis / tot : 0.0;
  }
  __syncthreads();
  const int keep = keep_sh;
  for (int c = threadIdx.x; c < keep; c += blockDim.x) sq[c] = sqrt(sv[j.s_off + c]);
  __syncthreads();
  const double2 *sa = roots + j.sa_off, *sb = roots + j.sb_off;
  const double2 *X = U + j.u_off, *Y = Vd + j.vd_off;  // X: D x D (first keep cols), Y: D x D
  for (int e = threadIdx.x; e < D * keep; e += blockDim.x) {
    const int i = e % D, c = e / D;
    double pr = 0.0, pi = 0.0;
    for (int x = 0; x < D; ++x) {
      const double2 a = sa[x + (int64_t)D * i], b = X[x + (int64_t)D * c];
      pr += a.x * b.x - a.y * b.y;
      pi += a.x * b.y + a.y * b.x;
    }
    proj[j.pa_off + e] = make_double2(pr * sq[c], pi * sq[c]);
  }
  for (int e = threadIdx.x; e < keep * D; e += blockDim.x) {
    const int c = e % keep, i = e / keep;
    double pr = 0.0, pi = 0.0;
    for (int x = 0; x < D; ++x) {
      const double2 a = Y[c + (int64_t)D * x], b = sb[x + (int64_t)D * i];
      pr += a.x * b.x - a.y * b.y;
      pi += a.x * b.y + a.y * b.x;
    }
    proj[j.pb_off + e] = make_double2(pr * sq[c], pi * sq[c]);
  }
  for (int e = threadIdx.x; e < keep * keep; e += blockDim.x) {
    const int r = e % keep, c = e / keep;
    proj[j.e_off + e] = make_double2(r == c ? sv[j.s_off + r] : 0.0, 0.0);
  }
}

// ----------------------------------------------------------- measure
// <O> = sum rho[p,p'] O[p,p'] / sum rho[p,p]  (measure.h:72-89)
__global__ void k_site_ratio(const double2 *__restrict__ rho, const double2 *__restrict__ ops,
                             const int *__restrict__ rho_idx, int d, int n,
                             double2 *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double2 *r = rho + (int64_t)rho_idx[t] * d * d;
  const double2 *o = ops + (int64_t)t * d * d;
  double nr = 0.0, ni = 0.0, dr = 0.0, di = 0.0;
  for (int p = 0; p < d; ++p)
    for (int q = 0; q < d; ++q) {
      const double2 a = r[p + d * q], b = o[p + d * q];
      nr += a.x * b.x - a.y * b.y;
      ni += a.x * b.y + a.y * b.x;
      if (p == q) {
        dr += a.x;
        di += a.y;
      }
    }
  const double den = dr * dr + di * di;
  out[t] = make_double2((nr * dr + ni * di) / den, (ni * dr - nr * di) / den);
}

// R = sum AA[p,p',i,j] O[p,x,p',y] BB[x,y,i,j], S = sum (tr AA)(tr BB)  (measure.h:255-294)
// AA is stored as d*d blocks M_{p p'}(i, j), block index p + d p'.
__global__ __launch_bounds__(256) void k_edge_ratio(const double2 *__restrict__ AA,
                                                     const double2 *__restrict__ BB,
                                                     const double2 *__restrict__ ops,
                                                     const int64_t *__restrict__ env_off,
                                                     const int *__restrict__ Dm, int d,
                                                     double2 *__restrict__ out) {
  __shared__ double sh[8];
  const int t = blockIdx.x;
  const int D = Dm[t];
  const int64_t DD = (int64_t)D * D;
  const double2 *A = AA + env_off[t], *B = BB + env_off[t];
  const double2 *O = ops + (int64_t)t * d * d * d * d;
  double rr = 0.0, ri = 0.0, sr = 0.0, si = 0.0;
  const int64_t Dd = (int64_t)D * d;
  // element (p, p', i, j) of an environment sits at (i + D p) + Dd (j + D p')
  auto at = [&](int pp, int qq, int i, int j) { return (i + (int64_t)D * pp) + Dd * (j + (int64_t)D * qq); };
  for (int64_t e = threadIdx.x; e < DD; e += blockDim.x) {
    const int i = (int)(e % D), j = (int)(e / D);
    double tar = 0.0, tai = 0.0, tbr = 0.0, tbi = 0.0;
    for (int p = 0; p < d; ++p) {
      const double2 a = A[at(p, p, i, j)], b = B[at(p, p, i, j)];
      tar += a.x; tai += a.y; tbr += b.x; tbi += b.y;
    }
    sr += tar * tbr - tai * tbi;
    si += tar * tbi + tai * tbr;
    for (int p = 0; p < d; ++p)
      for (int pp = 0; pp < d; ++pp) {
        const double2 a = A[at(p, pp, i, j)];
        for (int x = 0; x < d; ++x)
          for (int y = 0; y < d; ++y) {
            const double2 o = O[p + d * (x + d * (pp + d * y))];
            const double2 b = B[at(x, y, i, j)];
            const double ar = a.x * o.x - a.y * o.y, ai = a.x * o.y + a.y * o.x;
            rr += ar * b.x - ai * b.y;
            ri += ar * b.y + ai * b.x;
          }
      }
  }
  rr = bsum(rr, sh);
  ri = bsum(ri, sh);
  sr = bsum(sr, sh);
  si = bsum(si, sh);
  if (threadIdx.x == 0) {
    const double den = sr * sr + si * si;
    out[t] = make_double2((rr * sr + ri * si) / den, (ri * sr - rr * si) / den);
  }
}

// Path transfer matrix X (n complex): X /= ||X||_F, lg = (init ? 0 : lg) + log ||X||_F.
__global__ __launch_bounds__(256) void k_chain_norm(double2 *__restrict__ X, int n,
                                                     double *__restrict__ lg, int init) {
  __shared__ double sh[8];
  double s = 0.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) s += X[e].x * X[e].x + X[e].y * X[e].y;
  s = bsum(s, sh);
  const double nrm = sqrt(s);
  const double inv = nrm > 0.0 ? 1.0 / nrm : 1.0;
  for (int e = threadIdx.x; e < n; e += blockDim.x) X[e] = make_double2(X[e].x * inv, X[e].y * inv);
  if (threadIdx.x == 0) *lg = (init ? 0.0 : *lg) + (nrm > 0.0 ? log(nrm) : 0.0);
}

// Closing a path: val = sum_ij X[i, j] M[i, j] (X = null: trace of the d x d M)
__global__ __launch_bounds__(256) void k_chain_close(const double2 *__restrict__ X,
                                                      const double2 *__restrict__ M, int n, int d,
                                                      double2 *__restrict__ val,
                                                      double *__restrict__ lg, int init) {
  __shared__ double sh[8];
  double re = 0.0, im = 0.0;
  if (X) {
    for (int e = threadIdx.x; e < n; e += blockDim.x) {
      const double2 a = X[e], b = M[e];
      re += a.x * b.x - a.y * b.y;
      im += a.x * b.y + a.y * b.x;
    }
  } else {
    for (int q = threadIdx.x; q < d; q += blockDim.x) {
      re += M[q + d * q].x;
      im += M[q + d * q].y;
    }
  }
  re = bsum(re, sh);
  im = bsum(im, sh);
  if (threadIdx.x == 0) {
    *val = make_double2(re, im);
    if (init) *lg = 0.0;
  }
}

// out[t] = (val[t] / val[den[t]]) exp(lg[t] - lg[den[t]])
__global__ void k_chain_ratio(const double2 *__restrict__ val, const double *__restrict__ lg,
                              const int *__restrict__ den, int n, double2 *__restrict__ out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double2 a = val[t], b = val[den[t]];
  const double sc = exp(lg[t] - lg[den[t]]);
  const double dd = b.x * b.x + b.y * b.y;
  out[t] = make_double2(sc * (a.x * b.x + a.y * b.y) / dd, sc * (a.y * b.x - a.x * b.y) / dd);
}

inline int gridn(size_t n) {
  size_t g = (n + 255) / 256;
  return (int)std::max<size_t>(1, std::min<size_t>(g, 8192));
}

}  // namespace

// ================================================================ state
struct Peps {
  int precision = HBP_PREC_FP64;
  int n_sites = 0, n_edges = 0;
  std::vector<int> ea, eb;                 // edge endpoints (site addresses)
  std::vector<int> site_label;             // address -> label
  std::vector<std::vector<int>> bonds;     // address -> incident edge ids ascending
  std::vector<int> phys;                   // address -> d
  std::vector<int> bond_dim;               // edge -> D
  std::vector<void *> V;                   // address -> device tensor (S precision)
  std::vector<size_t> V_bytes;
  // site s's memory belongs to a caller's tensor (hbp_peps_bind_sites): never
  // released or written by the PEPS; replacing the site only drops the flag
  std::vector<char> V_borrowed;
  double2 *E = nullptr;                    // all 2 n_e messages
  size_t E_bytes = 0;
  std::vector<int64_t> E_off;              // message -> offset (complex<double>)
  // message m is exactly diagonal (after init, AbsorbTPO's E (x) I and
  // Truncation's diag(L)); a BP sweep produces dense messages
  std::vector<char> E_diag;
  // Lazy absorption: AbsorbTPO keeps V_s and its operator O_s; the absorbed
  // tensor is materialised into abs_buf when a kernel needs it, so only the
  // chi-sized state stays resident (the absorbed lattice is ~D^4 d per site).
  int absorb_mode = HBP_ABSORB_AUTO;
  int sweep_kernel = HBP_SWEEP_AUTO;
  std::vector<char> pending;
  std::vector<void *> pend_O;
  std::vector<size_t> pend_O_bytes;
  std::vector<AbsorbArgs> pend_args;
  void *abs_buf = nullptr;
  size_t abs_bytes = 0;
  int abs_site = -1;
  // work buffers, each as large as the largest site tensor
  void *buf[3] = {nullptr, nullptr, nullptr};
  size_t buf_bytes = 0;
  // Site tensors with the target pair of their first half moved to the front
  // (merged sweep): built on first use and reused by every sweep while the
  // sites stay unchanged (the sites are constant during a BP loop); dropped by
  // anything that changes a site.  Only for resident sites, and only while
  // device memory stays ample.
  std::vector<void *> Vpair;
  std::vector<size_t> Vpair_bytes;
  int pair_cache = -1;  // -1 undecided (decided once per set of sites), 0 off, 1 on
  // Small-site pass of the fused sweep: device job tables of the sites with
  // bonds <= SMALL_D_MAX (per colour filter: all, colour 0, colour 1) and the
  // host copies they were made from; a table is uploaded again only when its
  // jobs change (site tensors or message layout), so the sweeps of a BP loop
  // reuse it (and a recorded sweep needs no upload).
  SmallSiteJob *stab[3] = {nullptr, nullptr, nullptr};
  size_t stab_bytes[3] = {0, 0, 0};
  std::vector<SmallSiteJob> stab_host[3];
  // The lazily absorbed sites that pass takes, materialised together into one
  // buffer (site s at byte offset sabs_off[s]) on the first sweep after the
  // sites change and reused by the later sweeps; dropped with the pair-front
  // copies by anything that changes a site.
  void *sabs = nullptr;
  size_t sabs_bytes = 0;
  std::vector<int64_t> sabs_off;
  // caller's choice (hbp_peps_set_pair_cache): -1 by free device memory, 0 never, 1 always
  int pair_cache_policy = -1;
  int mode = HBP_MODE_FUSED;
  // mixed precision: BP stops once the residual is below this (FP32 messages
  // cannot reach the FP64 tolerances), or below the caller's tolerance
  double mixed_floor = 1.0e-6;
  // Fused mode fixes the free scalar phase of every message (trace real and
  // positive).  The residual of bp.h:321-350 compares normalised messages
  // including that phase, which rounding seeds and every sweep multiplies along
  // non-backtracking walks (growth ~ z - 1 per sweep on a square lattice), so
  // without the fix the residual stalls near 4/pi once it is below ~1e-10 on
  // large lattices.  Expectation values and projectors do not depend on it.
  bool phase_gauge = true;
  // BP update schedule of the fused mode: JACOBI computes every message from
  // the previous sweep's (bp.h:65-261 with E <- F after all layers);
  // CHECKERBOARD (bipartite lattices) first updates the messages leaving the
  // sites of colour 0, then those leaving colour 1 from the new ones
  int schedule = HBP_SCHEDULE_JACOBI;
  // Anderson mixing of the fused BP loop (hbp_peps_set_accel): window depth
  // (0: off, the plain fixed-point iteration E <- BP(E)) and the residual
  // below which the loop starts mixing
  int aa_depth = 0;
  double aa_start = 1.0e-2;
  std::vector<char> colour;  // site -> 0 / 1; empty until computed, or not bipartite
  int bipartite = -1;        // -1 unknown, 0 no, 1 yes

  int elem_bytes() const { return precision == HBP_PREC_FP64 ? 16 : 8; }
  int site_dtype() const { return precision == HBP_PREC_FP64 ? HBP_C128 : HBP_C64; }
  int z(int s) const { return (int)bonds[s].size(); }
  int in_msg(int s, int k) const {  // message entering site s on its bond k
    const int m = bonds[s][k];
    return ea[m] == s ? m + n_edges : m;
  }
  int out_msg(int s, int k) const {  // message leaving site s along bond k
    const int m = bonds[s][k];
    return ea[m] == s ? m : m + n_edges;
  }
  void shape(int s, int32_t *sh) const {
    for (int k = 0; k < z(s); ++k) sh[k] = bond_dim[bonds[s][k]];
    sh[z(s)] = phys[s];
  }
  size_t site_size(int s) const {
    size_t n = phys[s];
    for (int m : bonds[s]) n *= (size_t)bond_dim[m];
    return n;
  }
  int msg_dim(int m) const { return bond_dim[m % n_edges]; }
};

static void layout_messages(Ctx *ctx, Peps *p, double2 **newE, size_t *bytes,
                            std::vector<int64_t> *off) {
  off->resize(2 * p->n_edges);
  int64_t o = 0;
  for (int m = 0; m < 2 * p->n_edges; ++m) {
    (*off)[m] = o;
    const int64_t D = p->msg_dim(m);
    o += D * D;
  }
  *bytes = (size_t)std::max<int64_t>(o, 1) * 16;
  *newE = (double2 *)ctx->alloc(*bytes);
}

static void ensure_work(Ctx *ctx, Peps *p, size_t bytes) {
  if (bytes <= p->buf_bytes) return;
  // blocks that outlive a recording must not come from the cache then (a
  // recorded graph may use them as its temporaries)
  HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED, "work buffers grow while a graph is recorded");
  for (int i = 0; i < 3; ++i)
    if (p->buf[i]) ctx->release(p->buf[i], p->buf_bytes);
  for (int i = 0; i < 3; ++i) p->buf[i] = ctx->alloc(bytes);
  p->buf_bytes = bytes;
}

// V'_s = O_s applied to V_s (apply.h:52-66) into dst
static void materialize(Ctx *ctx, Peps *p, int s, void *dst) {
  const size_t n = p->site_size(s);
  if (p->precision == HBP_PREC_FP64)
    hipLaunchKernelGGL(k_absorb_site<double2>, dim3(gridn(n)), dim3(256), 0, ctx->stream,
                       (const double2 *)p->V[s], (const double2 *)p->pend_O[s], (double2 *)dst, n,
                       p->pend_args[s]);
  else
    hipLaunchKernelGGL(k_absorb_site<float2>, dim3(gridn(n)), dim3(256), 0, ctx->stream,
                       (const float2 *)p->V[s], (const double2 *)p->pend_O[s], (float2 *)dst, n,
                       p->pend_args[s]);
  HBP_HIP(hipGetLastError());
}

// The site tensor as the algorithms see it: V_s itself, or its absorbed form
// materialised into abs_buf (reused while the same site is asked for again).
static const void *site_data(Ctx *ctx, Peps *p, int s) {
  if (!p->pending[s]) return p->V[s];
  if (p->abs_site == s) return p->abs_buf;
  const size_t bytes = p->site_size(s) * p->elem_bytes();
  if (bytes > p->abs_bytes) {
    HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED,
                "absorption buffer grows while a graph is recorded");
    if (p->abs_buf) ctx->release(p->abs_buf, p->abs_bytes);
    p->abs_buf = ctx->alloc(bytes);
    p->abs_bytes = bytes;
  }
  materialize(ctx, p, s, p->abs_buf);
  p->abs_site = s;
  return p->abs_buf;
}

// The absorptions of `jobs` (sites of p, any sizes) in one launch of
// k_absorb_batch; the same arithmetic per element as materialize().
static void absorb_batch(Ctx *ctx, Peps *p, const std::vector<AbsorbJob> &jobs) {
  const int nj = (int)jobs.size();
  if (nj == 0) return;
  std::vector<int> first(nj + 1, 0);
  for (int j = 0; j < nj; ++j) {
    const size_t nb = (jobs[j].n + ABS_EPB - 1) / ABS_EPB;
    HBP_REQUIRE(nb >= 1 && (size_t)first[j] + nb < (size_t)INT32_MAX, HBP_ERR_UNSUPPORTED,
                "absorb: batched grid too large");
    first[j + 1] = first[j] + (int)nb;
  }
  AbsorbJob *dj = (AbsorbJob *)ctx->alloc(nj * sizeof(AbsorbJob));
  int *dfirst = (int *)ctx->alloc((nj + 1) * sizeof(int));
  ctx->upload(dj, jobs.data(), nj * sizeof(AbsorbJob));
  ctx->upload(dfirst, first.data(), (nj + 1) * sizeof(int));
  if (p->precision == HBP_PREC_FP64)
    hipLaunchKernelGGL(k_absorb_batch<double2>, dim3((unsigned)first[nj]), dim3(256), 0,
                       ctx->stream, dj, dfirst, nj);
  else
    hipLaunchKernelGGL(k_absorb_batch<float2>, dim3((unsigned)first[nj]), dim3(256), 0,
                       ctx->stream, dj, dfirst, nj);
  HBP_HIP(hipGetLastError());
  ctx->release(dfirst, (nj + 1) * sizeof(int));
  ctx->release(dj, nj * sizeof(AbsorbJob));
}

// Give up site s's current memory (a borrowed site is only forgotten).
static void release_site(Ctx *ctx, Peps *p, int s) {
  if (!p->V_borrowed.empty() && p->V_borrowed[s]) {
    p->V_borrowed[s] = 0;
    return;
  }
  ctx->release(p->V[s], p->V_bytes[s]);
}

// Store the absorbed form of site s as its tensor and drop the operator.
static void make_resident(Ctx *ctx, Peps *p, int s) {
  if (!p->pending[s]) return;
  const size_t bytes = p->site_size(s) * p->elem_bytes();
  void *nv = ctx->alloc(bytes);
  materialize(ctx, p, s, nv);
  release_site(ctx, p, s);
  p->V[s] = nv;
  p->V_bytes[s] = bytes;
  ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
  p->pend_O[s] = nullptr;
  p->pending[s] = 0;
  if (p->abs_site == s) p->abs_site = -1;
}

// Drops the copies derived from the sites: pair-front site tensors and the
// small-site pass's materialised lazily absorbed sites.
static void drop_pair_cache(Ctx *ctx, Peps *p) {
  for (size_t s = 0; s < p->Vpair.size(); ++s)
    if (p->Vpair[s]) {
      ctx->release(p->Vpair[s], p->Vpair_bytes[s]);
      p->Vpair[s] = nullptr;
      p->Vpair_bytes[s] = 0;
    }
  p->pair_cache = -1;
  if (p->sabs) {
    ctx->release(p->sabs, p->sabs_bytes);
    p->sabs = nullptr;
    p->sabs_bytes = 0;
  }
  p->sabs_off.clear();
}

// Replace site s's tensor (already allocated at the current shape) and drop
// any pending operator.
static void set_resident(Ctx *ctx, Peps *p, int s, void *nv, size_t bytes) {
  drop_pair_cache(ctx, p);
  release_site(ctx, p, s);
  p->V[s] = nv;
  p->V_bytes[s] = bytes;
  if (p->pending[s]) {
    ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
    p->pend_O[s] = nullptr;
    p->pending[s] = 0;
  }
  if (p->abs_site == s) p->abs_site = -1;
}

// Row weights of a mode product with diagonal messages: w[i] = prod_j
// diag(E_in(j))[c_j] over the bonds j in [j0, j1) of site s, i split
// column-major over their dims, times a trailing unit factor of `unit`.
static double2 *diag_weights(Ctx *ctx, Peps *p, int s, int j0, int j1, int64_t unit,
                             int64_t *n_out) {
  int64_t dims[HBP_MAX_ORDER], offs[HBP_MAX_ORDER];
  int nb = 0;
  int64_t n = unit;
  for (int j = j0; j < j1; ++j) {
    dims[nb] = p->bond_dim[p->bonds[s][j]];
    offs[nb] = p->E_off[p->in_msg(s, j)];
    n *= dims[nb];
    ++nb;
  }
  *n_out = n;
  if (nb == 0) return nullptr;
  double2 *w = (double2 *)ctx->alloc((size_t)n * 16);
  launch_build_weights(ctx->stream, p->E, w, n, nb, dims, offs);
  return w;
}

// Deferred reductions: message kernels write their split-K partials into the
// context's arena and the fixed-order reductions run together, up to
// RED_BATCH per launch (at small bond dimensions a sweep is bound by the
// number of launches, not by their work).  The arena is capped, so large
// messages still reduce one or two at a time; flush() before anything reads
// the messages.  The sums are bitwise those of one reduction per message.
struct Reducer {
  static constexpr size_t kCap = (size_t)64 << 20;
  Ctx *ctx;
  RedBatch b;
  size_t used = 0;
  explicit Reducer(Ctx *c) : ctx(c) { b.count = 0; }
  // arena space for the partials of `jobs` messages (added next)
  void *take(size_t bytes, int jobs = 1) {
    bytes = (bytes + 255) & ~(size_t)255;
    if (b.count + jobs > RED_BATCH ||
        (used > 0 && (used + bytes > kCap || used + bytes > ctx->arena_bytes)))
      flush();
    // the arena grows only while no partials are pending in it
    char *a = (char *)ctx->get_arena(used == 0 ? std::max(bytes, kCap) : used + bytes);
    void *ptr = a + used;
    used += bytes;
    return ptr;
  }
  void add(const void *parts, bool f32, int nch, int n, double2 *M) {
    b.j[b.count++] = RedBatchJob{parts, M, nch, n, f32 ? 1 : 0};
  }
  void flush() {
    launch_reduce_batch(ctx->stream, b);
    b.count = 0;
    used = 0;
  }
};

// M = sum over all bonds but k of (V x_{j != k} diag E_in(j))[.. i ..] conj V[.. j ..]:
// every incoming message on the other bonds is diagonal, so the mode products
// are row weights folded into the message kernel.  k == z gives the reduced
// matrix over the physical index (measure.h:54-84).
// rho[p, p'] = sum_l w[l] T[l + L p] conj V[l + L p'] for a physical index of
// dimension 2 (the last axis of T and V, every bond before it contracted)
template <typename S>
static void rho_message(Ctx *ctx, const void *T, const void *V, const double2 *w, int64_t L,
                        double2 *M, Reducer *red) {
  constexpr int kMaxBlocks = 1024;
  const size_t pb = (size_t)kMaxBlocks * 4 * 16 + 64;
  double2 *parts = (double2 *)(red ? red->take(pb) : ctx->get_scratch(pb));
  const int n = rho_partials<S>(ctx->stream, T, V, w, L, parts, kMaxBlocks);
  if (red)
    red->add(parts, false, n, 4, M);
  else
    launch_reduce_partials(ctx->stream, parts, n, 4, M);
}

// Weight arguments of the bonds [j0, j1) of site s (empty: nb == 0) and the
// vector length (times `unit`).
static WeightArgs weight_args(const Peps *p, int s, int j0, int j1, int64_t unit, int64_t *n_out) {
  WeightArgs a{};
  int64_t n = unit;
  for (int j = j0; j < j1; ++j) {
    a.dims[a.nb] = p->bond_dim[p->bonds[s][j]];
    a.diag_off[a.nb] = p->E_off[p->in_msg(s, j)];
    n *= a.dims[a.nb];
    ++a.nb;
  }
  *n_out = n;
  return a;
}

// The message on bond k of site s from given row weights wL (bonds before k)
// and wR (bonds after k, times the physical index); nullptr = all ones.
template <typename S>
static void weighted_message_w(Ctx *ctx, Peps *p, int s, int k, double2 *M, const double2 *wL,
                               const double2 *wR, Reducer *red) {
  const int z = p->z(s);
  int32_t sh[HBP_MAX_ORDER];
  p->shape(s, sh);
  int64_t L = 1, R = 1;
  for (int d = 0; d < k; ++d) L *= sh[d];
  for (int d = k + 1; d <= z; ++d) R *= sh[d];
  const int D = sh[k];
  const void *Vs = site_data(ctx, p, s);
  if (k == z && D == 2 && !wR) {  // the reduced matrix over the physical index
    rho_message<S>(ctx, Vs, Vs, wL, L, M, red);
    return;
  }
  const int64_t nch = message_num_chunks(L * R);
  const size_t pb = (size_t)nch * D * D * 16 + 64;
  double2 *parts = (double2 *)(red ? red->take(pb) : ctx->get_scratch(pb));
  const int pr = ctx->prof_begin(HBP_PROF_MESSAGE, 8.0 * L * R * D * D,
                                 2.0 * sizeof(S) * 2.0 * L * R * D);
  const int n = message_partials<S>(ctx->stream, Vs, Vs, parts, L, D, D, R, L * D,
                                    L * D, nch, 1, wL, wR);
  if (red)
    red->add(parts, false, n, D * D, M);
  else
    launch_reduce_partials(ctx->stream, parts, n, D * D, M);
  ctx->prof_end(pr);
}

template <typename S>
static void weighted_message(Ctx *ctx, Peps *p, int s, int k, double2 *M,
                             Reducer *red = nullptr) {
  const int z = p->z(s);
  int64_t nl = 0, nr = 0;
  double2 *wL = diag_weights(ctx, p, s, 0, k, 1, &nl);
  double2 *wR = k < z ? diag_weights(ctx, p, s, k + 1, z, p->phys[s], &nr) : nullptr;
  weighted_message_w<S>(ctx, p, s, k, M, wL, wR, red);
  if (wL) ctx->release(wL, (size_t)nl * 16);
  if (wR) ctx->release(wR, (size_t)nr * 16);
}

// The separate weighted messages of site s on the bonds ks (diagonal
// environment), with all their row weights built in one launch.
template <typename S>
static void weighted_messages_site(Ctx *ctx, Peps *p, int s, const std::vector<int> &ks,
                                   double2 *raw, const std::vector<int64_t> &raw_off,
                                   Reducer *red) {
  const int z = p->z(s);
  std::vector<WeightArgs> args;
  std::vector<int64_t> len, off;
  std::vector<int> slot(2 * ks.size(), -1);
  int64_t total = 0;
  for (size_t i = 0; i < ks.size(); ++i) {
    const int k = ks[i];
    for (int side = 0; side < 2; ++side) {
      int64_t n = 0;
      const WeightArgs a = side == 0 ? weight_args(p, s, 0, k, 1, &n)
                                     : weight_args(p, s, k + 1, z, p->phys[s], &n);
      if (a.nb == 0) continue;
      slot[2 * i + side] = (int)args.size();
      args.push_back(a);
      len.push_back(n);
      off.push_back(total);
      total += n;
    }
  }
  double2 *wbuf = total ? (double2 *)ctx->alloc((size_t)total * 16) : nullptr;
  WeightBatch wb;
  wb.count = 0;
  for (size_t v = 0; v < args.size(); ++v) {
    wb.a[wb.count] = args[v];
    wb.n[wb.count] = len[v];
    wb.w[wb.count] = wbuf + off[v];
    if (++wb.count == WEIGHT_BATCH) {
      launch_build_weights_batch(ctx->stream, p->E, wb);
      wb.count = 0;
    }
  }
  launch_build_weights_batch(ctx->stream, p->E, wb);
  for (size_t i = 0; i < ks.size(); ++i) {
    const int k = ks[i];
    const double2 *wL = slot[2 * i] >= 0 ? wbuf + off[slot[2 * i]] : nullptr;
    const double2 *wR = slot[2 * i + 1] >= 0 ? wbuf + off[slot[2 * i + 1]] : nullptr;
    weighted_message_w<S>(ctx, p, s, k, raw + raw_off[p->out_msg(s, k)], wL, wR, red);
  }
  if (wbuf) ctx->release(wbuf, (size_t)total * 16);
}

// every message entering site s except on bond `skip` is diagonal
static bool diag_env(const Peps *p, int s, int skip) {
  for (int k = 0; k < p->z(s); ++k)
    if (k != skip && !p->E_diag[p->in_msg(s, k)]) return false;
  return true;
}

template <typename S>
static void mode_prod(Ctx *ctx, const void *T, const double2 *E, void *Out, const int32_t *sh,
                      int order, int k, int Dout, int e_trans) {
  int64_t L = 1, R = 1;
  for (int d = 0; d < k; ++d) L *= sh[d];
  for (int d = k + 1; d < order; ++d) R *= sh[d];
  const double es = sizeof(S) * 2.0;
  const int pr = ctx->prof_begin(HBP_PROF_MODE_PRODUCT, 8.0 * L * R * sh[k] * Dout,
                                 es * L * R * (sh[k] + Dout) + 16.0 * sh[k] * Dout);
  launch_mode_product<S>(ctx->stream, T, E, Out, L, sh[k], Dout, R, e_trans, 1.0, nullptr,
                         nullptr);
  ctx->prof_end(pr);
}

// raw M[i, j] = sum T[.. i ..] conj V[.. j ..] over all axes but k (both same shape)
template <typename S>
static void message_raw(Ctx *ctx, const void *T, const void *V, const int32_t *sh, int order,
                        int k, double2 *M, Reducer *red = nullptr) {
  int64_t L = 1, R = 1;
  for (int d = 0; d < k; ++d) L *= sh[d];
  for (int d = k + 1; d < order; ++d) R *= sh[d];
  const int D = sh[k];
  if (R == 1 && D == 2) {  // the reduced matrix over a last (physical) index
    rho_message<S>(ctx, T, V, nullptr, L, M, red);
    return;
  }
  const int64_t nch = message_num_chunks(L * R);
  const size_t pb = (size_t)nch * D * D * 16 + 64;
  double2 *parts = (double2 *)(red ? red->take(pb) : ctx->get_scratch(pb));
  const int pr = ctx->prof_begin(HBP_PROF_MESSAGE, 8.0 * L * R * D * D,
                                 2.0 * sizeof(S) * 2.0 * L * R * D);
  const int n = message_partials<S>(ctx->stream, T, V, parts, L, D, D, R, L * D, L * D, nch, 1);
  if (red)
    red->add(parts, false, n, D * D, M);
  else
    launch_reduce_partials(ctx->stream, parts, n, D * D, M);
  ctx->prof_end(pr);
}

// Merged form of T = P x_j E followed by the message on bond k (k, j adjacent):
// the product stays in registers inside the message kernel.
static bool merge_ok(const Peps *p, const int32_t *sh, int k, int j) {
  if (p->sweep_kernel != HBP_SWEEP_MERGED && p->sweep_kernel != HBP_SWEEP_AUTO) return false;
  return (k - j == 1 || j - k == 1) && sh[k] > 32 && sh[k] <= 64 && sh[j] > 32 && sh[j] <= 64;
}

// P, V viewed as (L, i, a, R) (k_first) or (L, a, i, R)
template <typename S>
static void product_message(Ctx *ctx, const void *P, const double2 *E, const void *V, int64_t L,
                            int64_t R, int Dk, int Dj, int k_first, double2 *M,
                            Reducer *red = nullptr) {
  constexpr int kMaxBlocks = 512;
  using C = std::conditional_t<sizeof(S) == 4, float2, double2>;
  const size_t pb = (size_t)kMaxBlocks * Dk * Dk * sizeof(C) + 64;
  C *parts = (C *)(red ? red->take(pb) : ctx->get_scratch(pb));
  const double n = (double)L * R * Dk * Dj;
  const int pr = ctx->prof_begin(HBP_PROF_PRODUCT_MESSAGE, 8.0 * n * (Dk + Dj),
                                 2.0 * sizeof(S) * 2.0 * n + 16.0 * Dj * Dj);
  const int nb = product_message_partials<S>(ctx->stream, P, V, E, parts, L, R, Dk, Dj, k_first,
                                             kMaxBlocks);
  if (red)
    red->add(parts, sizeof(C) == 8, nb, Dk * Dk, M);
  else
    launch_reduce_partials(ctx->stream, parts, nb, Dk * Dk, M);
  ctx->prof_end(pr);
}

// Both messages of the target pair (t, t + 1) of site s from the half-product
// P (the site with every bond absorbed but the pair's): for k in the pair,
// M_k = <P x_j E_j, V> over all bonds but k, j the other one.  P and V have
// the pair in front: (D_t, D_{t+1}, R) views.
template <typename S>
static void merged_pair(Ctx *ctx, Peps *p, int s, const void *P, const void *V,
                        const int32_t *sh, int t, int64_t R, double2 *raw,
                        const std::vector<int64_t> &raw_off, Reducer *red) {
  for (int k = t; k <= t + 1; ++k) {
    const int j = k == t ? t + 1 : t;
    product_message<S>(ctx, P, p->E + p->E_off[p->in_msg(s, j)], V, 1, R, sh[k], sh[j],
                       k == t ? 1 : 0, raw + raw_off[p->out_msg(s, k)], red);
  }
}

// T = V_s with every incoming message absorbed except on bond `skip` (-1: none),
// returns the buffer holding it (V_s itself when nothing is absorbed).
template <typename S>
static const void *absorb_env(Ctx *ctx, Peps *p, int s, int skip, int32_t *sh) {
  const int order = p->z(s) + 1;
  p->shape(s, sh);
  const void *cur = site_data(ctx, p, s);
  int which = 0;
  for (int k = 0; k < p->z(s); ++k) {
    if (k == skip) continue;
    const int mi = p->in_msg(s, k);
    void *dst = p->buf[which];
    mode_prod<S>(ctx, cur, p->E + p->E_off[mi], dst, sh, order, k, sh[k], 0);
    cur = dst;
    which ^= 1;
  }
  return cur;
}

// Whether the pair-front copies of the sites fit: decided once per set of
// sites from the bytes all of them need and the free device memory, keeping
// max(16 GiB, 15 %) of the device free.  HBP_PAIR_CACHE=0 in the environment
// turns the copies off (each sweep then rearranges into a work buffer).
static void plan_pair_cache(Ctx *ctx, Peps *p) {
  if ((int)p->Vpair.size() != p->n_sites) {
    p->Vpair.assign(p->n_sites, nullptr);
    p->Vpair_bytes.assign(p->n_sites, 0);
  }
  static const bool allowed = [] {
    const char *e = getenv("HBP_PAIR_CACHE");
    return !(e && e[0] == '0');
  }();
  if (!allowed || p->pair_cache_policy == 0) {
    p->pair_cache = 0;
    return;
  }
  if (p->pair_cache_policy == 1) {
    p->pair_cache = 1;
    return;
  }
  int32_t sh[HBP_MAX_ORDER];
  size_t need = 0;
  for (int s = 0; s < p->n_sites; ++s) {
    const int z = p->z(s), h = z / 2;
    if (z < 3 || z - h != 2 || p->pending[s]) continue;
    p->shape(s, sh);
    if (merge_ok(p, sh, h, h + 1)) need += p->site_size(s) * p->elem_bytes();
  }
  size_t tot = 0;
  const size_t fr = ctx->free_bytes(&tot);
  const size_t keep = std::max<size_t>((size_t)16 << 30, tot / 100 * 15);
  p->pair_cache = need + keep < fr ? 1 : 0;
}

// Site s (data Vs) with its bond pair (t, t + 1) moved to the front:
// (D_t, D_{t+1}, bonds before t, bonds after t + 1, phys).  Vs itself when no
// bond precedes the pair; else the cached copy (built here on first use) or a
// rearrangement into work buffer 2.
template <typename S>
static const void *pair_front(Ctx *ctx, Peps *p, int s, const void *Vs, const int32_t *sh,
                              int order, int t) {
  int64_t Lb = 1, R = 1;
  for (int d = 0; d < t; ++d) Lb *= sh[d];
  for (int d = t + 2; d < order; ++d) R *= sh[d];
  if (Lb == 1) return Vs;
  HBP_REQUIRE(p->pair_cache >= 0 || !graph_capturing(), HBP_ERR_UNSUPPORTED,
              "pair-front copies planned while a graph is recorded");
  if (p->pair_cache < 0) plan_pair_cache(ctx, p);
  if (p->Vpair[s]) return p->Vpair[s];
  void *dst = p->buf[2];
  if (p->pair_cache == 1 && !p->pending[s]) {
    HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED,
                "pair-front copy built while a graph is recorded");
    const size_t vb = p->site_size(s) * p->elem_bytes();
    p->Vpair[s] = ctx->alloc(vb);
    p->Vpair_bytes[s] = vb;
    dst = p->Vpair[s];
  }
  launch_swap_blocks(ctx->stream, Vs, dst, Lb, (int64_t)sh[t] * sh[t + 1], R, (int)(2 * sizeof(S)));
  return dst;
}

// Both messages of the pair (t, t + 1) of site s under a diagonal environment
// (first sweep after AbsorbTPO / Truncation) with one pass over the site:
// pair_gram_partials on the pair-front tensor.  Needs t == 0 or t + 2 == z, so
// that the remaining bonds form one block.
template <typename S>
static void pair_messages_diag(Ctx *ctx, Peps *p, int s, const int32_t *sh, int order, int t,
                               double2 *Mx, double2 *My, Reducer *red) {
  const int z = p->z(s);
  const void *X = pair_front<S>(ctx, p, s, site_data(ctx, p, s), sh, order, t);
  int64_t nw = 0;
  double2 *wo = t == 0 ? diag_weights(ctx, p, s, 2, z, p->phys[s], &nw)
                       : diag_weights(ctx, p, s, 0, t, p->phys[s], &nw);
  const int Dx = sh[t], Dy = sh[t + 1];
  const int64_t nO = (int64_t)(p->site_size(s) / ((size_t)Dx * Dy));
  constexpr int kMaxBlocks = 512;
  using C = std::conditional_t<sizeof(S) == 4, float2, double2>;
  const size_t pb = (size_t)kMaxBlocks * (Dx * Dx + Dy * Dy) * sizeof(C) + 64;
  C *px = (C *)(red ? red->take(pb, 2) : ctx->get_scratch(pb));
  C *py = px + (size_t)kMaxBlocks * Dx * Dx;
  const double n = (double)nO * Dx * Dy;
  const int pr = ctx->prof_begin(HBP_PROF_MESSAGE, 8.0 * n * (Dx + Dy), sizeof(S) * 2.0 * n);
  const int nb = pair_gram_partials<S>(ctx->stream, X, nO, Dx, Dy,
                                       p->E + p->E_off[p->in_msg(s, t)],
                                       p->E + p->E_off[p->in_msg(s, t + 1)], wo, px, py,
                                       kMaxBlocks);
  if (red) {
    red->add(px, sizeof(C) == 8, nb, Dx * Dx, Mx);
    red->add(py, sizeof(C) == 8, nb, Dy * Dy, My);
  } else {
    launch_reduce_partials(ctx->stream, px, nb, Dx * Dx, Mx);
    launch_reduce_partials(ctx->stream, py, nb, Dy * Dy, My);
  }
  ctx->prof_end(pr);
  if (wo) ctx->release(wo, (size_t)nw * 16);
}

// Sites the small-site pass takes: degree 1..4, every bond <= SMALL_D_MAX.
static bool small_site(const Peps *p, int s) {
  const int z = p->z(s);
  if (z < 1 || z > 4) return false;
  for (int k = 0; k < z; ++k)
    if (p->bond_dim[p->bonds[s][k]] > SMALL_D_MAX) return false;
  return true;
}

// The absorbed tensor of every pending small site in p->sabs (one launch of
// k_absorb_batch, the same arithmetic per element as materialize(), so the
// pass reads bitwise what eager absorption would have stored).
static void materialize_small_pending(Ctx *ctx, Peps *p) {
  if (p->sabs || !p->sabs_off.empty()) return;
  p->sabs_off.assign(p->n_sites, -1);
  std::vector<AbsorbJob> aj;
  size_t total = 0;
  for (int s = 0; s < p->n_sites; ++s) {
    if (!p->pending[s] || !small_site(p, s)) continue;
    p->sabs_off[s] = (int64_t)total;
    total += (p->site_size(s) * p->elem_bytes() + 255) & ~(size_t)255;
    aj.push_back(AbsorbJob{p->V[s], (const double2 *)p->pend_O[s], nullptr, p->site_size(s),
                           p->pend_args[s]});
  }
  if (aj.empty()) return;
  p->sabs = ctx->alloc(total);
  p->sabs_bytes = total;
  size_t i = 0;
  for (int s = 0; s < p->n_sites; ++s)
    if (p->sabs_off[s] >= 0) aj[i++].Vn = (char *)p->sabs + p->sabs_off[s];
  absorb_batch(ctx, p, aj);
}

// The sites of a fused sweep whose bonds are all <= SMALL_D_MAX (degree
// 1..4; `only` / `colour`: the checkerboard's colour filter) in one launch of
// the batched small-site kernel (hbp_batched.hip: one work item per outgoing
// message, FP64 sums in a fixed order) instead of a dozen launches each; they
// are marked in `done`.  Lazily absorbed sites are read from their materialised copies, so
// lazy and eager absorption give bitwise the same messages.
// HBP_SWEEP_PER_SITE keeps every site on its own launches.
template <typename S>
static void small_sites_sweep(Ctx *ctx, Peps *p, double2 *raw, const std::vector<int64_t> &raw_off,
                              const std::vector<char> *only, char colour, std::vector<char> *done) {
  if (p->sweep_kernel == HBP_SWEEP_PER_SITE) return;
  materialize_small_pending(ctx, p);
  std::vector<SmallSiteJob> jobs;
  std::vector<int> sites;
  int64_t slab = 1;
  for (int s = 0; s < p->n_sites; ++s) {
    if (only && (*only)[s] != colour) continue;
    if (!small_site(p, s)) continue;
    const int z = p->z(s);
    SmallSiteJob j{};
    j.V = p->V[s];
    if (p->pending[s]) {
      HBP_REQUIRE(p->sabs && p->sabs_off[s] >= 0, HBP_ERR_LOGIC,
                  "small-site sweep: pending site not materialised");
      j.V = (char *)p->sabs + p->sabs_off[s];
    }
    j.z = z;
    for (int k = 0; k < z; ++k) {
      j.dims[k] = p->bond_dim[p->bonds[s][k]];
      j.in_off[k] = p->E_off[p->in_msg(s, k)];
      j.out_off[k] = raw_off[p->out_msg(s, k)];
    }
    j.dims[z] = p->phys[s];
    jobs.push_back(j);
    sites.push_back(s);
    slab = std::max<int64_t>(slab, (int64_t)p->site_size(s));
  }
  if (jobs.empty()) return;
  const int slot = only ? 1 + colour : 0;
  const size_t bytes = jobs.size() * sizeof(SmallSiteJob);
  std::vector<SmallSiteJob> &host = p->stab_host[slot];
  if (host.size() != jobs.size() || std::memcmp(host.data(), jobs.data(), bytes) != 0) {
    if (p->stab[slot] && p->stab_bytes[slot] != bytes) {
      ctx->release(p->stab[slot], p->stab_bytes[slot]);
      p->stab[slot] = nullptr;
    }
    if (!p->stab[slot]) {
      p->stab[slot] = (SmallSiteJob *)ctx->alloc(bytes);
      p->stab_bytes[slot] = bytes;
    }
    ctx->upload(p->stab[slot], jobs.data(), bytes);
    host = jobs;
  }
  const int nj = (int)jobs.size();
  const size_t scratch_bytes =
      (size_t)small_sweep_grid(SMALL_GRID_VARIANTS - 1, SMALL_SWEEP_ITEMS * nj) * 2 *
      (size_t)slab * 2 * sizeof(S);
  void *scratch = ctx->alloc(scratch_bytes);
  launch_sweep_small_tuned<S>(ctx->stream, p->stab[slot], nj, scratch, slab, p->E, raw);
  ctx->release(scratch, scratch_bytes);
  done->assign(p->n_sites, 0);
  for (int s : sites) (*done)[s] = 1;
}

template <typename S>
static void sweep_raw(Ctx *ctx, Peps *p, double2 *raw, const std::vector<int64_t> &raw_off,
                      const std::vector<char> *only = nullptr, char colour = 0) {
  size_t maxb = 16;
  for (int s = 0; s < p->n_sites; ++s) maxb = std::max(maxb, p->site_size(s) * p->elem_bytes());
  ensure_work(ctx, p, maxb);
  // HBP_REDUCE_BATCH=0 in the environment reduces each message on its own
  // (for tests that compare the two bitwise)
  static const bool batch = [] {
    const char *e = getenv("HBP_REDUCE_BATCH");
    return !(e && e[0] == '0');
  }();
  Reducer reducer(ctx);
  Reducer *red = batch ? &reducer : nullptr;
  std::vector<char> small;
  small_sites_sweep<S>(ctx, p, raw, raw_off, only, colour, &small);
  int32_t sh[HBP_MAX_ORDER];
  for (int s = 0; s < p->n_sites; ++s) {
    if (only && (*only)[s] != colour) continue;
    if (!small.empty() && small[s]) continue;
    const int z = p->z(s), order = z + 1;
    p->shape(s, sh);
    if (diag_env(p, s, -1)) {
      // first sweep after AbsorbTPO / Truncation: E = diag(L) (x) I on every bond;
      // pairs that the merged sweep takes get both messages from one pass
      bool done[HBP_MAX_ORDER] = {false};
      if (z >= 3) {
        const int pairs[2] = {z == 4 ? 0 : -1, z / 2};
        for (int t : pairs) {
          if (t < 0 || !(t == 0 || t + 2 == z) || !merge_ok(p, sh, t, t + 1)) continue;
          pair_messages_diag<S>(ctx, p, s, sh, order, t, raw + raw_off[p->out_msg(s, t)],
                                raw + raw_off[p->out_msg(s, t + 1)], red);
          done[t] = done[t + 1] = true;
        }
      }
      std::vector<int> ks;
      for (int k = 0; k < z; ++k)
        if (!done[k]) ks.push_back(k);
      weighted_messages_site<S>(ctx, p, s, ks, raw, raw_off, red);
      continue;
    }
    if (z <= 2) {
      for (int k = 0; k < z; ++k) {
        const void *T = absorb_env<S>(ctx, p, s, k, sh);
        message_raw<S>(ctx, T, site_data(ctx, p, s), sh, order, k, raw + raw_off[p->out_msg(s, k)],
                       red);
      }
      continue;
    }
    // Shared partial products: with the bonds split in halves H1 = [0, h) and
    // H2 = [h, z), A = V x_{H1} serves every target in H2 and B = V x_{H2}
    // every target in H1 (8 instead of 12 mode products at z = 4).  Mode
    // products on different bonds commute; only rounding changes.
    const int h = z / 2;
    for (int half = 0; half < 2; ++half) {
      const int a0 = half == 0 ? 0 : h, a1 = half == 0 ? h : z;   // absorbed first
      const int t0 = half == 0 ? h : 0, t1 = half == 0 ? z : h;   // targets
      // A = V x_{[a0, a1)} into buf[0] (ping-pong with buf[1] when needed)
      const void *Vs = site_data(ctx, p, s);
      const void *cur = Vs;
      int nprod = a1 - a0, w = (nprod % 2 == 1) ? 0 : 1;
      if (t1 - t0 == 2 && merge_ok(p, sh, t0, t0 + 1)) {
        // merged messages want the target pair in front (whole contiguous
        // D_t x D_{t+1} slices): when bonds precede it, take V as (pair,
        // leading bonds, rest) and absorb the other bonds in that layout
        int64_t Lb = 1, R = 1;
        for (int d = 0; d < t0; ++d) Lb *= sh[d];
        for (int d = t0 + 2; d < order; ++d) R *= sh[d];
        int32_t shp[HBP_MAX_ORDER];
        int pos[HBP_MAX_ORDER];
        for (int d = 0; d < order; ++d) {
          pos[d] = d < t0 ? d + 2 : d <= t0 + 1 ? d - t0 : d;
          shp[pos[d]] = sh[d];
        }
        Vs = pair_front<S>(ctx, p, s, Vs, sh, order, t0);
        cur = Vs;
        for (int j = a0; j < a1; ++j) {
          void *dst = p->buf[w];
          mode_prod<S>(ctx, cur, p->E + p->E_off[p->in_msg(s, j)], dst, shp, order, pos[j], sh[j],
                       0);
          cur = dst;
          w ^= 1;
        }
        merged_pair<S>(ctx, p, s, cur, Vs, sh, t0, Lb * R, raw, raw_off, red);
        continue;
      }
      for (int j = a0; j < a1; ++j) {
        void *dst = p->buf[w];
        mode_prod<S>(ctx, cur, p->E + p->E_off[p->in_msg(s, j)], dst, sh, order, j, sh[j], 0);
        cur = dst;
        w ^= 1;
      }
      // cur == buf[0] (or V when nothing was absorbed); targets use buf[1], buf[2]
      for (int k = t0; k < t1; ++k) {
        const void *T = cur;
        int u = 1;
        for (int j = t0; j < t1; ++j) {
          if (j == k) continue;
          void *dst = p->buf[u];
          mode_prod<S>(ctx, T, p->E + p->E_off[p->in_msg(s, j)], dst, sh, order, j, sh[j], 0);
          T = dst;
          u = u == 1 ? 2 : 1;
        }
        message_raw<S>(ctx, T, Vs, sh, order, k, raw + raw_off[p->out_msg(s, k)], red);
      }
    }
  }
  if (red) red->flush();
}

// The reference's op sequence for every message, one TCI call at a time
// (bp.h:128-175): T = copy(V); for every bond k: (k != target) T = T x_k E_in,
// then normalize(T) as a norm pass and a scale pass; C = conj(copy(V));
// M = <T, C> over all bonds but the target; normalize(M).
template <typename S>
static void sweep_raw_plain(Ctx *ctx, Peps *p, double2 *raw, const std::vector<int64_t> &raw_off) {
  size_t maxb = 16;
  for (int s = 0; s < p->n_sites; ++s) maxb = std::max(maxb, p->site_size(s) * p->elem_bytes());
  ensure_work(ctx, p, maxb);
  double *dv = (double *)ctx->alloc(sizeof(double));
  int32_t sh[HBP_MAX_ORDER];
  for (int s = 0; s < p->n_sites; ++s) {
    const int z = p->z(s);
    p->shape(s, sh);
    const size_t n = p->site_size(s), bytes = n * p->elem_bytes();
    for (int k = 0; k < z; ++k) {
      HBP_HIP(hipMemcpyAsync(p->buf[0], site_data(ctx, p, s), bytes, hipMemcpyDeviceToDevice,
                             ctx->stream));
      int cur = 0;
      for (int j = 0; j < z; ++j) {
        if (j != k) {
          mode_prod<S>(ctx, p->buf[cur], p->E + p->E_off[p->in_msg(s, j)], p->buf[cur ^ 1], sh,
                       z + 1, j, sh[j], 0);
          cur ^= 1;
        }
        device_norm2(ctx, p->buf[cur], n, p->site_dtype(), dv);
        launch_scale(ctx->stream, p->buf[cur], n, p->site_dtype(), 1.0, 0.0, dv);
      }
      HBP_HIP(hipMemcpyAsync(p->buf[2], site_data(ctx, p, s), bytes, hipMemcpyDeviceToDevice,
                             ctx->stream));
      launch_conj(ctx->stream, p->buf[2], n, p->site_dtype());
      int64_t L = 1, R = 1;
      for (int d = 0; d < k; ++d) L *= sh[d];
      for (int d = k + 1; d <= z; ++d) R *= sh[d];
      const int D = sh[k];
      const int64_t nch = message_num_chunks(L * R);
      double2 *parts = (double2 *)ctx->get_scratch((size_t)nch * D * D * 16 + 64);
      const int pr = ctx->prof_begin(HBP_PROF_MESSAGE, 8.0 * L * R * D * D,
                                     2.0 * sizeof(S) * 2.0 * L * R * D);
      const int nc = message_partials<S>(ctx->stream, p->buf[cur], p->buf[2], parts, L, D, D, R,
                                         L * D, L * D, nch, 0);
      double2 *M = raw + raw_off[p->out_msg(s, k)];
      launch_reduce_partials(ctx->stream, parts, nc, D * D, M);
      ctx->prof_end(pr);
      device_norm2(ctx, M, (size_t)D * D, HBP_C128, dv);
      launch_scale(ctx->stream, M, (size_t)D * D, HBP_C128, 1.0, 0.0, dv);
    }
  }
  ctx->release(dv, sizeof(double));
}

// Two-colouring of the lattice graph (BFS); false when it has an odd cycle.
static bool colour_sites(Peps *p) {
  if (p->bipartite >= 0) return p->bipartite == 1;
  std::vector<std::vector<int>> nb(p->n_sites);
  for (int m = 0; m < p->n_edges; ++m) {
    nb[p->ea[m]].push_back(p->eb[m]);
    nb[p->eb[m]].push_back(p->ea[m]);
  }
  std::vector<int> c(p->n_sites, -1);
  bool ok = true;
  for (int s0 = 0; s0 < p->n_sites && ok; ++s0) {
    if (c[s0] >= 0) continue;
    c[s0] = 0;
    std::vector<int> q{s0};
    for (size_t i = 0; i < q.size() && ok; ++i)
      for (int t : nb[q[i]]) {
        if (c[t] < 0) {
          c[t] = 1 - c[q[i]];
          q.push_back(t);
        } else if (c[t] == c[q[i]]) {
          ok = false;
        }
      }
  }
  p->bipartite = ok ? 1 : 0;
  if (ok) p->colour.assign(c.begin(), c.end());
  return ok;
}

// One checkerboard BP sweep: the messages leaving colour-0 sites from E, then
// those leaving colour-1 sites from E with the colour-0 ones already replaced.
// On a bipartite lattice a message leaving a colour-c site depends only on
// messages leaving colour 1 - c sites, so the Jacobi sweep advances two
// independent sequences by one step each while this sweep advances one of
// them by two: the fixed points are the same (both schedules stop on a
// message set that one more update leaves within the tolerance).  newE gets
// the new messages; the residual compares them with E like the Jacobi form.
template <typename S>
static void sweep_checkerboard(Ctx *ctx, Peps *p, double2 *raw, double2 *newE, double *dist,
                               bool with_residual) {
  const int nm = 2 * p->n_edges;
  double2 *E_old = p->E;
  double2 *mix = (double2 *)ctx->alloc(std::max<size_t>(p->E_bytes, 16));
  HBP_HIP(hipMemcpyAsync(mix, E_old, p->E_bytes, hipMemcpyDeviceToDevice, ctx->stream));
  const std::vector<char> diag_saved = p->E_diag;
  for (int c = 0; c < 2; ++c) {
    std::vector<MsgJob> jobs;
    std::vector<int> ms;
    for (int m = 0; m < nm; ++m) {
      const int src = m < p->n_edges ? p->ea[m] : p->eb[m - p->n_edges];  // E[m] leaves ea
      if (p->colour[src] == c) ms.push_back(m);
    }
    if (ms.empty()) continue;
    // colour 1 reads the colour-0 updates (dense from here on) through p->E
    if (c == 1) p->E = mix;
    sweep_raw<S>(ctx, p, raw, p->E_off, &p->colour, (char)c);
    p->E = E_old;
    for (int m : ms) jobs.push_back(MsgJob{p->E_off[m], p->E_off[m], p->E_off[m], p->msg_dim(m)});
    MsgJob *dj = (MsgJob *)ctx->alloc(jobs.size() * sizeof(MsgJob));
    ctx->upload(dj, jobs.data(), jobs.size() * sizeof(MsgJob));
    double *d = (double *)ctx->alloc(jobs.size() * sizeof(double));
    hipLaunchKernelGGL(k_finalize_messages, dim3((unsigned)jobs.size()), dim3(256), 0,
                       ctx->stream, raw, newE, with_residual ? E_old : nullptr, dj, d,
                       p->phase_gauge ? 1 : 0, c == 0 ? mix : nullptr);
    HBP_HIP(hipGetLastError());
    if (with_residual) {
      // scatter the colour's distances to their message positions (tiny)
      std::vector<int> pos(ms.begin(), ms.end());
      int *dpos = (int *)ctx->alloc(pos.size() * sizeof(int));
      ctx->upload(dpos, pos.data(), pos.size() * sizeof(int));
      hipLaunchKernelGGL(k_scatter_dist, dim3(1), dim3(256), 0, ctx->stream, d, dpos,
                         (int)pos.size(), dist);
      HBP_HIP(hipGetLastError());
      ctx->release(dpos, pos.size() * sizeof(int));
    }
    ctx->release(d, jobs.size() * sizeof(double));
    ctx->release(dj, jobs.size() * sizeof(MsgJob));
    if (c == 0)
      for (int m : ms) p->E_diag[m] = 0;
  }
  p->E_diag = diag_saved;
  ctx->release(mix, std::max<size_t>(p->E_bytes, 16));
}

// one BP sweep: newE = normalise(BP(E)); returns residual on device if dist != null.
// dj_pre (Jacobi schedule only): the message table (MsgJob per message,
// offsets p->E_off) already on the device, so that the sweep needs no upload
static void sweep(Ctx *ctx, Peps *p, double2 *newE, bool with_residual, double *dev_res,
                  const MsgJob *dj_pre = nullptr) {
  const int nm = 2 * p->n_edges;
  if (p->mode == HBP_MODE_FUSED && p->schedule == HBP_SCHEDULE_CHECKERBOARD && colour_sites(p)) {
    double2 *raw = (double2 *)ctx->alloc(std::max<size_t>(p->E_bytes, 16));
    double *dist = (double *)ctx->alloc(nm * sizeof(double));
    if (p->precision == HBP_PREC_FP64)
      sweep_checkerboard<double>(ctx, p, raw, newE, dist, with_residual);
    else
      sweep_checkerboard<float>(ctx, p, raw, newE, dist, with_residual);
    if (with_residual) {
      hipLaunchKernelGGL(k_residual, dim3(1), dim3(256), 0, ctx->stream, dist, nm,
                         1.0 / (2.0 * p->n_edges), dev_res);
      HBP_HIP(hipGetLastError());
    }
    ctx->release(dist, nm * sizeof(double));
    ctx->release(raw, std::max<size_t>(p->E_bytes, 16));
    return;
  }
  // job table first: its upload waits on the stream, which is then still short
  const MsgJob *dj = dj_pre;
  MsgJob *dj_own = nullptr;
  if (!dj) {
    std::vector<MsgJob> jobs(nm);
    for (int m = 0; m < nm; ++m)
      jobs[m] = MsgJob{p->E_off[m], p->E_off[m], p->E_off[m], p->msg_dim(m)};
    dj_own = (MsgJob *)ctx->alloc(nm * sizeof(MsgJob));
    ctx->upload(dj_own, jobs.data(), nm * sizeof(MsgJob));
    dj = dj_own;
  }
  double2 *raw = (double2 *)ctx->alloc(std::max<size_t>(p->E_bytes, 16));
  if (p->mode == HBP_MODE_PLAIN) {
    if (p->precision == HBP_PREC_FP64)
      sweep_raw_plain<double>(ctx, p, raw, p->E_off);
    else
      sweep_raw_plain<float>(ctx, p, raw, p->E_off);
  } else {
    // AUTO runs the per-site launches with merged product-message kernels where
    // the bonds allow (HBP_SWEEP_GROUPED is rejected by hbp_peps_set_sweep_kernel)
    if (p->precision == HBP_PREC_FP64)
      sweep_raw<double>(ctx, p, raw, p->E_off);
    else
      sweep_raw<float>(ctx, p, raw, p->E_off);
  }
  double *dist = (double *)ctx->alloc(nm * sizeof(double));
  hipLaunchKernelGGL(k_finalize_messages, dim3(nm), dim3(256), 0, ctx->stream, raw, newE,
                     with_residual ? p->E : nullptr, dj, dist,
                     p->mode == HBP_MODE_FUSED && p->phase_gauge ? 1 : 0, nullptr);
  HBP_HIP(hipGetLastError());
  if (with_residual) {
    hipLaunchKernelGGL(k_residual, dim3(1), dim3(256), 0, ctx->stream, dist, nm,
                       1.0 / (2.0 * p->n_edges), dev_res);
    HBP_HIP(hipGetLastError());
  }
  ctx->release(dist, nm * sizeof(double));
  if (dj_own) ctx->release(dj_own, nm * sizeof(MsgJob));
  ctx->release(raw, std::max<size_t>(p->E_bytes, 16));
}

// ----------------------------------------------------------- truncation
// Sites of a truncation with every bond <= SMALL_D_MAX and degree <= 4, of
// any number of PEPS of one precision: all their projectors and
// normalisations run in one launch of the small-site kernel (per-site
// launches of a few microseconds each were most of a truncation's time at
// these sizes); lazily absorbed ones are materialised first, in one launch.
struct SmallProjBatch {
  std::vector<SmallProjJob> jobs;
  std::vector<std::pair<Peps *, int>> site;  // (PEPS, site) of each job
  std::vector<size_t> bytes;                 // of each new tensor
  std::vector<AbsorbJob> absorb;
  std::vector<std::pair<void *, size_t>> tmp;
  int64_t slab = 1;
  Peps *any = nullptr;
};

// Adds the small sites of p to b and marks them in `small`.
static void collect_small_projections(Ctx *ctx, Peps *p, const double2 *proj,
                                      const std::vector<EdgeJob> &ej,
                                      const std::vector<int> &keep, std::vector<char> &small,
                                      SmallProjBatch &b) {
  small.assign(p->n_sites, 0);
  for (int s = 0; s < p->n_sites; ++s) {
    const int z = p->z(s);
    if (z > 4) continue;
    // (a projector never widens a bond: keep <= the bond's dim, so the
    // kernel's ping-pong buffers hold every intermediate)
    bool ok = true;
    for (int k = 0; k < z; ++k) {
      const int m = p->bonds[s][k];
      ok = ok && p->bond_dim[m] <= SMALL_D_MAX && keep[m] <= p->bond_dim[m] && keep[m] >= 1;
    }
    if (!ok) continue;
    SmallProjJob j{};
    j.V = p->V[s];
    if (p->pending[s]) {
      const size_t vb = p->site_size(s) * p->elem_bytes();
      void *t = ctx->alloc(vb);
      b.absorb.push_back(AbsorbJob{p->V[s], (const double2 *)p->pend_O[s], t, p->site_size(s),
                                   p->pend_args[s]});
      b.tmp.emplace_back(t, vb);
      j.V = t;
    }
    j.z = z;
    size_t n_old = p->phys[s], n_new = p->phys[s];
    for (int k = 0; k < z; ++k) {
      const int m = p->bonds[s][k];
      j.dims[k] = p->bond_dim[m];
      j.keep[k] = keep[m];
      j.trans[k] = p->ea[m] == s ? 0 : 1;
      j.P[k] = proj + (p->ea[m] == s ? ej[m].pa_off : ej[m].pb_off);
      n_old *= (size_t)p->bond_dim[m];
      n_new *= (size_t)keep[m];
    }
    j.dims[z] = p->phys[s];
    const size_t bytes = n_new * p->elem_bytes();
    j.out = ctx->alloc(bytes);
    b.slab = std::max<int64_t>(b.slab, (int64_t)n_old);
    b.jobs.push_back(j);
    b.site.emplace_back(p, s);
    b.bytes.push_back(bytes);
    b.any = p;
    small[s] = 1;
  }
}

static void run_small_projections(Ctx *ctx, SmallProjBatch &b) {
  if (b.jobs.empty()) return;
  absorb_batch(ctx, b.any, b.absorb);
  const int nj = (int)b.jobs.size();
  const int grid = small_sweep_grid(0, nj);
  const size_t scratch_bytes = (size_t)grid * 2 * (size_t)b.slab * b.any->elem_bytes();
  void *scratch = ctx->alloc(scratch_bytes);
  SmallProjJob *dj = (SmallProjJob *)ctx->alloc(nj * sizeof(SmallProjJob));
  ctx->upload(dj, b.jobs.data(), nj * sizeof(SmallProjJob));
  if (b.any->precision == HBP_PREC_FP64)
    launch_project_small<double>(ctx->stream, dj, nj, scratch, b.slab, grid);
  else
    launch_project_small<float>(ctx->stream, dj, nj, scratch, b.slab, grid);
  ctx->release(dj, nj * sizeof(SmallProjJob));
  ctx->release(scratch, scratch_bytes);
  for (auto &t : b.tmp) ctx->release(t.first, t.second);
  for (int i = 0; i < nj; ++i)
    set_resident(ctx, b.site[i].first, b.site[i].second, b.jobs[i].out, b.bytes[i]);
}

// The other sites: projectors bond by bond (the last one writes the new site
// tensor), then the normalisation (truncation.h:156-217).
template <typename S>
static void apply_projectors_large(Ctx *ctx, Peps *p, const double2 *proj,
                                   const std::vector<EdgeJob> &ej, const std::vector<int> &keep,
                                   const std::vector<char> &small,
                                   const std::vector<int> &old_dim) {
  for (int s = 0; s < p->n_sites; ++s) {
    if (small[s]) continue;
    const int z = p->z(s);
    int32_t sh[HBP_MAX_ORDER];
    for (int k = 0; k < z; ++k) sh[k] = old_dim[p->bonds[s][k]];
    sh[z] = p->phys[s];
    size_t maxn = p->phys[s];
    for (int k = 0; k < z; ++k) maxn *= (size_t)sh[k];
    ensure_work(ctx, p, std::max(p->buf_bytes, maxn * p->elem_bytes()));
    size_t n = p->phys[s];
    for (int k = 0; k < z; ++k) n *= (size_t)keep[p->bonds[s][k]];
    const size_t bytes = n * p->elem_bytes();
    void *nv = ctx->alloc(bytes);
    const void *cur = site_data(ctx, p, s);
    int which = 0;
    for (int k = 0; k < z; ++k) {
      const int m = p->bonds[s][k];
      const int kp = keep[m];
      void *dst = k + 1 == z ? nv : p->buf[which];
      if (p->ea[m] == s)
        mode_prod<S>(ctx, cur, proj + ej[m].pa_off, dst, sh, z + 1, k, kp, 0);
      else
        mode_prod<S>(ctx, cur, proj + ej[m].pb_off, dst, sh, z + 1, k, kp, 1);
      sh[k] = kp;
      cur = dst;
      which ^= 1;
    }
    if (z == 0) HBP_HIP(hipMemcpyAsync(nv, cur, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    double *dv = (double *)ctx->alloc(sizeof(double));
    device_norm2(ctx, nv, n, p->site_dtype(), dv);
    launch_scale(ctx->stream, nv, n, p->site_dtype(), 1.0, 0.0, dv);
    ctx->release(dv, sizeof(double));
    set_resident(ctx, p, s, nv, bytes);
  }
}

// One PEPS of a (batched) truncation and where its part of the batch lies.
struct TruncMember {
  Peps *p = nullptr;
  int chi = 1;
  double sv_min = 0.0;
  int32_t *dims_out = nullptr;
  double *errs_out = nullptr;
  int64_t e_base = 0;  // its messages' offset in the batch's message buffer
  int64_t e_len = 0;   // and their elements
  int m_base = 0;      // its first message in the batch
  int e_first = 0;     // its first edge in the batch
  std::vector<EdgeJob> ej;
};

// Truncation (truncation.h:17-222) of every PEPS of `mem` at once: one
// batched SVD of all their messages, one root launch, one T = Ra^T Rb launch,
// one batched SVD of all T, one projector launch, one read of the kept
// dimensions, the small sites' projectors in one launch per precision, the
// others site by site, and all new diag(Lambda) messages in one launch.
// Every matrix, edge and site is computed by the same kernel code as for a
// PEPS on its own, so each PEPS's result does not depend on the batch.
static void truncate_many(Ctx *ctx, std::vector<TruncMember> &mem) {
  // The batched SVD keeps a launch's matrices in LDS when its largest fits
  // (bonds <= 64): members whose own bonds fit come first and get that form,
  // the others a launch of their own, so that every PEPS's matrices take the
  // form they take when it is truncated alone.
  auto own_max = [](const Peps *p) {
    int d = 1;
    for (int m = 0; m < p->n_edges; ++m) d = std::max(d, p->bond_dim[m]);
    return d;
  };
  auto lds_form = [&](const Peps *p) {
    const int64_t d = own_max(p);
    return 2 * d * d <= SVD_LDS_ELEMS;
  };
  std::stable_partition(mem.begin(), mem.end(), [&](const TruncMember &t) { return lds_form(t.p); });
  int nm_lds = 0, ne_lds = 0, maxD_lds = 1, maxD_glob = 1;
  for (auto &t : mem) {
    if (lds_form(t.p)) {
      nm_lds += 2 * t.p->n_edges;
      ne_lds += t.p->n_edges;
      maxD_lds = std::max(maxD_lds, own_max(t.p));
    } else {
      maxD_glob = std::max(maxD_glob, own_max(t.p));
    }
  }
  int64_t tot = 0;
  int nm_all = 0, ne_all = 0, maxD = 1;
  for (auto &t : mem) {
    Peps *p = t.p;
    t.e_base = tot;
    t.e_len = 0;
    for (int m = 0; m < 2 * p->n_edges; ++m) t.e_len += (int64_t)p->msg_dim(m) * p->msg_dim(m);
    tot += t.e_len;
    t.m_base = nm_all;
    t.e_first = ne_all;
    nm_all += 2 * p->n_edges;
    ne_all += p->n_edges;
    for (int m = 0; m < p->n_edges; ++m) maxD = std::max(maxD, p->bond_dim[m]);
  }
  HBP_REQUIRE(maxD <= 512, HBP_ERR_UNSUPPORTED, "truncation: bond > 512");
  const int64_t totc = std::max<int64_t>(tot, 1);
  // the messages of every member in one buffer (a lone PEPS's own)
  double2 *Ecat = nullptr;
  const double2 *Eall = mem.size() == 1 ? mem[0].p->E : nullptr;
  if (!Eall) {
    Ecat = (double2 *)ctx->alloc(totc * 16);
    for (auto &t : mem)
      if (t.e_len > 0)
        HBP_HIP(hipMemcpyAsync(Ecat + t.e_base, t.p->E, (size_t)t.e_len * 16,
                               hipMemcpyDeviceToDevice, ctx->stream));
    Eall = Ecat;
  }
  double2 *U = (double2 *)ctx->alloc(totc * 16);
  double2 *Vd = (double2 *)ctx->alloc(totc * 16);
  double2 *W = (double2 *)ctx->alloc(2 * totc * 16);
  double *sv = (double *)ctx->alloc((size_t)std::max(nm_all, 1) * maxD * 8);
  double2 *roots = (double2 *)ctx->alloc(2 * totc * 16);
  // 1.-4. job tables: SVD of every message, its roots, the edges (T, its
  // SVD, projectors, new messages); offsets into the shared buffers
  std::vector<SvdJob> sj, tj;
  std::vector<RootJob> rj;
  std::vector<EdgeJob> ej_all;
  int64_t tt = 0, poff = 0;
  for (auto &t : mem) {
    Peps *p = t.p;
    const int ne = p->n_edges;
    for (int m = 0; m < 2 * ne; ++m) {
      const int D = p->msg_dim(m);
      const int64_t o = t.e_base + p->E_off[m];
      const int64_t so = (int64_t)(t.m_base + m) * maxD;
      sj.push_back(SvdJob{o, o, so, o, 2 * o, D, D});
      rj.push_back(RootJob{o, so, o, 2 * o, 2 * o + (int64_t)D * D, D, t.sv_min});
    }
    t.ej.assign(ne, EdgeJob{});
    for (int m = 0; m < ne; ++m) {
      const int D = p->bond_dim[m];
      EdgeJob j{};
      j.D = D;
      j.chi = t.chi;
      j.s_min = t.sv_min;
      j.ra_off = rj[t.m_base + m].r_off;       // M_{a->b} = E[m]
      j.sa_off = rj[t.m_base + m].si_off;
      j.rb_off = rj[t.m_base + m + ne].r_off;  // M_{b->a} = E[m + ne]
      j.sb_off = rj[t.m_base + m + ne].si_off;
      j.t_off = tt;
      j.u_off = tt;
      j.vd_off = tt;
      j.s_off = (int64_t)(t.e_first + m) * maxD;
      j.pa_off = poff;
      j.pb_off = poff + (int64_t)D * D;
      j.e_off = poff + 2 * (int64_t)D * D;
      poff += 3 * (int64_t)D * D;
      tt += (int64_t)D * D;
      t.ej[m] = j;
      ej_all.push_back(j);
      tj.push_back(SvdJob{j.t_off, j.u_off, j.s_off, j.vd_off, 2 * j.t_off, D, D});
    }
  }
  // one upload: [message SVDs | roots | edges | T SVDs]
  const size_t b_sj = sj.size() * sizeof(SvdJob), b_rj = rj.size() * sizeof(RootJob),
               b_ej = ej_all.size() * sizeof(EdgeJob), b_tj = tj.size() * sizeof(SvdJob);
  const size_t tab_bytes = std::max<size_t>(b_sj + b_rj + b_ej + b_tj, 16);
  std::vector<char> tab(tab_bytes);
  std::memcpy(tab.data(), sj.data(), b_sj);
  std::memcpy(tab.data() + b_sj, rj.data(), b_rj);
  std::memcpy(tab.data() + b_sj + b_rj, ej_all.data(), b_ej);
  std::memcpy(tab.data() + b_sj + b_rj + b_ej, tj.data(), b_tj);
  char *dtab = (char *)ctx->alloc(tab_bytes);
  ctx->upload(dtab, tab.data(), tab_bytes);
  const SvdJob *dsj = (const SvdJob *)dtab;
  const RootJob *drj = (const RootJob *)(dtab + b_sj);
  const EdgeJob *dej = (const EdgeJob *)(dtab + b_sj + b_rj);
  const SvdJob *dtj = (const SvdJob *)(dtab + b_sj + b_rj + b_ej);
  double2 *T = (double2 *)ctx->alloc(std::max<int64_t>(tt, 1) * 16);
  double2 *proj = (double2 *)ctx->alloc(std::max<int64_t>(poff, 1) * 16);
  int *dkeep = (int *)ctx->alloc(std::max(ne_all, 1) * sizeof(int));
  double *derr = (double *)ctx->alloc(std::max(ne_all, 1) * sizeof(double));
  std::vector<int> keep_all(ne_all);
  std::vector<double> err_all(ne_all);
  auto svd_both = [&](const double2 *A, const SvdJob *jobs, int n_all, int n_lds) {
    if (n_lds > 0) launch_batched_svd(ctx->stream, A, U, sv, Vd, W, jobs, n_lds, maxD_lds, maxD_lds);
    if (n_all > n_lds)
      launch_batched_svd(ctx->stream, A, U, sv, Vd, W, jobs + n_lds, n_all - n_lds, maxD_glob,
                         maxD_glob);
  };
  if (nm_all > 0) {
    svd_both(Eall, dsj, nm_all, nm_lds);
    hipLaunchKernelGGL(k_roots, dim3(nm_all), dim3(256), 0, ctx->stream, U, sv, Vd, roots, drj);
    HBP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_rr, dim3(ne_all), dim3(256), 0, ctx->stream, roots, T, dej);
    HBP_HIP(hipGetLastError());
    svd_both(T, dtj, ne_all, ne_lds);
    // 5. selection and projectors
    hipLaunchKernelGGL(k_projectors, dim3(ne_all), dim3(256), 0, ctx->stream, roots, U, sv, Vd,
                       proj, dej, dkeep, derr);
    HBP_HIP(hipGetLastError());
    HBP_HIP(hipMemcpyAsync(keep_all.data(), dkeep, ne_all * sizeof(int), hipMemcpyDeviceToHost,
                           ctx->stream));
    HBP_HIP(hipMemcpyAsync(err_all.data(), derr, ne_all * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->sync();
  }
  // 6. projectors into the sites: the small sites of all members in one
  // launch per precision, the others site by site
  SmallProjBatch sb[2];
  std::vector<std::vector<char>> small(mem.size());
  std::vector<std::vector<int>> keeps(mem.size()), old_dims(mem.size());
  for (size_t i = 0; i < mem.size(); ++i) {
    TruncMember &t = mem[i];
    keeps[i].assign(keep_all.begin() + t.e_first, keep_all.begin() + t.e_first + t.p->n_edges);
    old_dims[i] = t.p->bond_dim;
    collect_small_projections(ctx, t.p, proj, t.ej, keeps[i], small[i],
                              sb[t.p->precision == HBP_PREC_MIXED ? 1 : 0]);
  }
  run_small_projections(ctx, sb[0]);
  run_small_projections(ctx, sb[1]);
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    if (p->precision == HBP_PREC_FP64)
      apply_projectors_large<double>(ctx, p, proj, mem[i].ej, keeps[i], small[i], old_dims[i]);
    else
      apply_projectors_large<float>(ctx, p, proj, mem[i].ej, keeps[i], small[i], old_dims[i]);
  }
  // 7. messages = diag(Lambda) in both directions, every member in one launch
  std::vector<CopyJob> cj;
  for (size_t i = 0; i < mem.size(); ++i) {
    TruncMember &t = mem[i];
    Peps *p = t.p;
    const int ne = p->n_edges;
    for (int m = 0; m < ne; ++m) p->bond_dim[m] = keeps[i][m];
    double2 *nE;
    size_t nbytes;
    std::vector<int64_t> noff;
    layout_messages(ctx, p, &nE, &nbytes, &noff);
    for (int m = 0; m < ne; ++m)
      cj.push_back(CopyJob{t.ej[m].e_off, nE + noff[m], nE + noff[m + ne], keeps[i][m] * keeps[i][m]});
    ctx->release(p->E, p->E_bytes);
    p->E = nE;
    p->E_bytes = nbytes;
    p->E_off = noff;
    p->E_diag.assign(2 * ne, 1);
    if (t.dims_out)
      for (int m = 0; m < ne; ++m) t.dims_out[m] = keeps[i][m];
    if (t.errs_out)
      for (int m = 0; m < ne; ++m) t.errs_out[m] = err_all[t.e_first + m];
  }
  if (!cj.empty()) {
    CopyJob *dcj = (CopyJob *)ctx->alloc(cj.size() * sizeof(CopyJob));
    ctx->upload(dcj, cj.data(), cj.size() * sizeof(CopyJob));
    hipLaunchKernelGGL(k_copy_twice, dim3((unsigned)cj.size()), dim3(256), 0, ctx->stream, proj,
                       dcj);
    HBP_HIP(hipGetLastError());
    ctx->release(dcj, cj.size() * sizeof(CopyJob));
  }
  ctx->release(derr, std::max(ne_all, 1) * sizeof(double));
  ctx->release(dkeep, std::max(ne_all, 1) * sizeof(int));
  ctx->release(proj, std::max<int64_t>(poff, 1) * 16);
  ctx->release(T, std::max<int64_t>(tt, 1) * 16);
  ctx->release(dtab, tab_bytes);
  ctx->release(roots, 2 * totc * 16);
  ctx->release(sv, (size_t)std::max(nm_all, 1) * maxD * 8);
  ctx->release(W, 2 * totc * 16);
  ctx->release(Vd, totc * 16);
  ctx->release(U, totc * 16);
  if (Ecat) ctx->release(Ecat, totc * 16);
}

static void truncate(Ctx *ctx, Peps *p, int chi, double sv_min, int32_t *dims_out,
                     double *errs_out) {
  std::vector<TruncMember> mem(1);
  mem[0].p = p;
  mem[0].chi = chi;
  mem[0].sv_min = sv_min;
  mem[0].dims_out = dims_out;
  mem[0].errs_out = errs_out;
  truncate_many(ctx, mem);
}

// ------------------------------------------------------------- measure
template <typename S>
static void measure_sites_impl(Ctx *ctx, Peps *p, int n, const int *addr, const double *ops,
                               double *out) {
  // reduced matrices rho_s[p, p'] for each distinct requested site
  std::vector<int> uniq(addr, addr + n);
  std::sort(uniq.begin(), uniq.end());
  uniq.erase(std::unique(uniq.begin(), uniq.end()), uniq.end());
  int maxd = 1;
  for (int s : uniq) maxd = std::max(maxd, p->phys[s]);
  size_t maxb = 16;
  for (int s : uniq) maxb = std::max(maxb, p->site_size(s) * p->elem_bytes());
  ensure_work(ctx, p, maxb);
  double2 *rho = (double2 *)ctx->alloc(uniq.size() * maxd * maxd * 16);
  int32_t sh[HBP_MAX_ORDER];
  for (size_t u = 0; u < uniq.size(); ++u) {
    const int s = uniq[u];
    if (diag_env(p, s, -1)) {  // after Truncation every message is diag(L)
      weighted_message<S>(ctx, p, s, p->z(s), rho + u * maxd * maxd);
      continue;
    }
    const void *W = absorb_env<S>(ctx, p, s, -1, sh);
    // rho[p, p'] = sum_b W[b, p] conj V[b, p']: message over the physical axis
    message_raw<S>(ctx, W, site_data(ctx, p, s), sh, p->z(s) + 1, p->z(s), rho + u * maxd * maxd);
  }
  std::vector<int> idx(n);
  for (int t = 0; t < n; ++t) {
    idx[t] = (int)(std::lower_bound(uniq.begin(), uniq.end(), addr[t]) - uniq.begin());
    HBP_REQUIRE(p->phys[addr[t]] == maxd, HBP_ERR_UNSUPPORTED, "mixed physical dims in one call");
  }
  const int d = maxd;
  double2 *dops = (double2 *)ctx->alloc((size_t)n * d * d * 16);
  int *didx = (int *)ctx->alloc(n * sizeof(int));
  double2 *dout = (double2 *)ctx->alloc(n * 16);
  ctx->upload(dops, ops, (size_t)n * d * d * 16);
  ctx->upload(didx, idx.data(), n * sizeof(int));
  hipLaunchKernelGGL(k_site_ratio, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, rho, dops,
                     didx, d, n, dout);
  HBP_HIP(hipGetLastError());
  HBP_HIP(hipMemcpyAsync(out, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(dout, n * 16);
  ctx->release(didx, n * sizeof(int));
  ctx->release(dops, (size_t)n * d * d * 16);
  ctx->release(rho, uniq.size() * maxd * maxd * 16);
}

// AdagA[p, p', i, j] = sum A[.., i, .., p] conj V[.., j, .., p'] (measure.h:194-204)
// in one pass: the physical slices become further columns, so out is the
// (D d) x (D d) matrix with row i + D p and column j + D p'.
template <typename S>
static void edge_env(Ctx *ctx, Peps *p, int s, int kt, double2 *out) {
  int32_t sh[HBP_MAX_ORDER];
  const bool diag = diag_env(p, s, kt);
  const void *Vs = site_data(ctx, p, s);
  const void *A = diag ? Vs : absorb_env<S>(ctx, p, s, kt, sh);
  if (diag) p->shape(s, sh);
  const int z = p->z(s), d = p->phys[s];
  // diagonal environment: the absorbed bonds become row weights of A
  int64_t nl = 0, nr = 0;
  double2 *wL = diag ? diag_weights(ctx, p, s, 0, kt, 1, &nl) : nullptr;
  double2 *wR = diag ? diag_weights(ctx, p, s, kt + 1, z, 1, &nr) : nullptr;
  int64_t L = 1, R = 1;
  for (int k = 0; k < kt; ++k) L *= sh[k];
  for (int k = kt + 1; k < z; ++k) R *= sh[k];
  const int D = sh[kt], Dd = D * d;
  HBP_REQUIRE(Dd <= 128, HBP_ERR_UNSUPPORTED, "measure: bond x physical dim above 128");
  const int64_t slab = L * D * R;  // elements per physical index
  const int64_t nch = message_num_chunks(L * R);
  double2 *parts = (double2 *)ctx->get_scratch((size_t)nch * Dd * Dd * 16 + 64);
  const int pr = ctx->prof_begin(HBP_PROF_MESSAGE, 8.0 * L * R * Dd * Dd,
                                 2.0 * sizeof(S) * 2.0 * L * R * Dd);
  const int nc = message_partials<S>(ctx->stream, A, Vs, parts, L, Dd, Dd, R, L * D, L * D,
                                     nch, 1, wL, wR, D, slab);
  launch_reduce_partials(ctx->stream, parts, nc, Dd * Dd, out);
  ctx->prof_end(pr);
  if (wL) ctx->release(wL, (size_t)nl * 16);
  if (wR) ctx->release(wR, (size_t)nr * 16);
}

template <typename S>
static void measure_edges_impl(Ctx *ctx, Peps *p, int n, const int *pa, const int *pb,
                               const double *ops, double *out) {
  size_t maxb = 16;
  for (int s = 0; s < p->n_sites; ++s) maxb = std::max(maxb, p->site_size(s) * p->elem_bytes());
  ensure_work(ctx, p, maxb);
  std::vector<int64_t> off(n);
  std::vector<int> Dm(n);
  int64_t tot = 0;
  int d = 0;
  std::vector<int> ka(n), kb(n);
  for (int t = 0; t < n; ++t) {
    const int a = pa[t], b = pb[t];
    int m = -1;
    for (int k = 0; k < p->z(a); ++k) {
      const int e = p->bonds[a][k];
      if ((p->ea[e] == a && p->eb[e] == b) || (p->ea[e] == b && p->eb[e] == a)) {
        m = e;
        ka[t] = k;
      }
    }
    HBP_REQUIRE(m >= 0, HBP_ERR_INVALID_ARGUMENT, "measure: sites are not nearest neighbours");
    kb[t] = (int)(std::find(p->bonds[b].begin(), p->bonds[b].end(), m) - p->bonds[b].begin());
    HBP_REQUIRE(p->phys[a] == p->phys[b] && (d == 0 || d == p->phys[a]), HBP_ERR_UNSUPPORTED,
                "measure: mixed physical dims");
    d = p->phys[a];
    Dm[t] = p->bond_dim[m];
    off[t] = tot;
    tot += (int64_t)d * d * Dm[t] * Dm[t];
  }
  double2 *AA = (double2 *)ctx->alloc(std::max<int64_t>(tot, 1) * 16);
  double2 *BB = (double2 *)ctx->alloc(std::max<int64_t>(tot, 1) * 16);
  for (int t = 0; t < n; ++t) {
    edge_env<S>(ctx, p, pa[t], ka[t], AA + off[t]);
    edge_env<S>(ctx, p, pb[t], kb[t], BB + off[t]);
  }
  const size_t ob = (size_t)n * d * d * d * d * 16;
  double2 *dops = (double2 *)ctx->alloc(ob);
  int64_t *doff = (int64_t *)ctx->alloc(n * sizeof(int64_t));
  int *dD = (int *)ctx->alloc(n * sizeof(int));
  double2 *dout = (double2 *)ctx->alloc(n * 16);
  ctx->upload(dops, ops, ob);
  ctx->upload(doff, off.data(), n * sizeof(int64_t));
  ctx->upload(dD, Dm.data(), n * sizeof(int));
  hipLaunchKernelGGL(k_edge_ratio, dim3(n), dim3(256), 0, ctx->stream, AA, BB, dops, doff, dD, d,
                     dout);
  HBP_HIP(hipGetLastError());
  HBP_HIP(hipMemcpyAsync(out, dout, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(dout, n * 16);
  ctx->release(dD, n * sizeof(int));
  ctx->release(doff, n * sizeof(int64_t));
  ctx->release(dops, ob);
  ctx->release(BB, std::max<int64_t>(tot, 1) * 16);
  ctx->release(AA, std::max<int64_t>(tot, 1) * 16);
}

// Product operator O_1 ... O_k on a simple path s_1 - ... - s_k (an extension of
// measure.h, whose operators stop at two neighbouring sites).  The BP
// environment of the path is the incoming message on every bond that leaves
// it; along the path a transfer matrix X[b, b'] (ket, bra) carries the
// contraction from site to site:
//   X_1 = M_{b_1}(V_1 env O_1),  X_i = M_{b_i}((V_i env) x_{a_i} X_{i-1} x_p O_i)
//   value = sum X_{k-1} o M_{a_k}((V_k env) x_p O_k)
// where M_b(T)[i, j] = sum over all other axes of T[.. i ..] conj V[.. j ..].
// X is normalised after every site, the logarithm of the scale kept beside it.
// ops == null runs the same path with identities (the normalisation S).
template <typename S>
static void chain_value(Ctx *ctx, Peps *p, const std::vector<int> &path,
                        const std::vector<int> &kin, const std::vector<int> &kout,
                        const double2 *dops, double2 *X, double2 *M, double2 *val, double *lg) {
  const int k = (int)path.size();
  int32_t sh[HBP_MAX_ORDER];
  for (int i = 0; i < k; ++i) {
    const int s = path[i], z = p->z(s), order = z + 1, d = p->phys[s];
    p->shape(s, sh);
    const void *Vs = site_data(ctx, p, s);
    const void *cur = Vs;
    int which = 0;
    for (int b = 0; b < z; ++b) {
      if (b == kin[i] || b == kout[i]) continue;
      mode_prod<S>(ctx, cur, p->E + p->E_off[p->in_msg(s, b)], p->buf[which], sh, order, b,
                   sh[b], 0);
      cur = p->buf[which];
      which ^= 1;
    }
    if (kin[i] >= 0 && kout[i] >= 0) {  // the last site meets X in k_chain_close
      mode_prod<S>(ctx, cur, X, p->buf[which], sh, order, kin[i], sh[kin[i]], 0);
      cur = p->buf[which];
      which ^= 1;
    }
    if (dops) {
      mode_prod<S>(ctx, cur, dops + (size_t)i * d * d, p->buf[which], sh, order, z, d, 0);
      cur = p->buf[which];
      which ^= 1;
    }
    const int t = kout[i] >= 0 ? kout[i] : (kin[i] >= 0 ? kin[i] : z);
    message_raw<S>(ctx, cur, Vs, sh, order, t, M);
    const int n = sh[t] * sh[t];
    if (kout[i] >= 0) {
      HBP_HIP(hipMemcpyAsync(X, M, (size_t)n * 16, hipMemcpyDeviceToDevice, ctx->stream));
      hipLaunchKernelGGL(k_chain_norm, dim3(1), dim3(256), 0, ctx->stream, X, n, lg, i == 0);
    } else {
      hipLaunchKernelGGL(k_chain_close, dim3(1), dim3(256), 0, ctx->stream,
                         kin[i] >= 0 ? (const double2 *)X : nullptr, (const double2 *)M, n,
                         sh[t], val, lg, i == 0);
    }
    HBP_HIP(hipGetLastError());
  }
}

// bond index at site a of the edge {a, b}, -1 when they are not neighbours
static int bond_to(const Peps *p, int a, int b) {
  for (int k = 0; k < p->z(a); ++k) {
    const int e = p->bonds[a][k];
    if ((p->ea[e] == a && p->eb[e] == b) || (p->ea[e] == b && p->eb[e] == a)) return k;
  }
  return -1;
}

template <typename S>
static void measure_paths_impl(Ctx *ctx, Peps *p, int n, const int *offsets, const int *addr,
                               const double *ops, double *out) {
  std::vector<std::vector<int>> paths(n), kins(n), kouts(n);
  std::map<std::vector<int>, int> den_slot;
  std::vector<int> den(n);
  int d = 0, maxD = 1;
  size_t maxb = 16;
  for (int t = 0; t < n; ++t) {
    const int k = offsets[t + 1] - offsets[t];
    HBP_REQUIRE(k >= 1, HBP_ERR_INVALID_ARGUMENT, "measure: empty operator path");
    std::vector<int> &pt = paths[t];
    pt.assign(addr + offsets[t], addr + offsets[t + 1]);
    std::vector<int> sorted = pt;
    std::sort(sorted.begin(), sorted.end());
    HBP_REQUIRE(std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end(),
                HBP_ERR_INVALID_ARGUMENT, "measure: a path visits a site twice");
    kins[t].assign(k, -1);
    kouts[t].assign(k, -1);
    for (int i = 0; i < k; ++i) {
      const int s = pt[i];
      HBP_REQUIRE(d == 0 || p->phys[s] == d, HBP_ERR_UNSUPPORTED, "measure: mixed physical dims");
      d = p->phys[s];
      maxb = std::max(maxb, p->site_size(s) * p->elem_bytes());
      for (int b = 0; b < p->z(s); ++b) maxD = std::max(maxD, p->bond_dim[p->bonds[s][b]]);
      if (i + 1 < k) {
        kouts[t][i] = bond_to(p, s, pt[i + 1]);
        HBP_REQUIRE(kouts[t][i] >= 0, HBP_ERR_INVALID_ARGUMENT,
                    "measure: consecutive path sites are not neighbours");
      }
      if (i > 0) kins[t][i] = bond_to(p, s, pt[i - 1]);
    }
    auto it = den_slot.find(pt);
    if (it == den_slot.end()) it = den_slot.emplace(pt, n + (int)den_slot.size()).first;
    den[t] = it->second;
  }
  maxD = std::max(maxD, d);
  ensure_work(ctx, p, maxb);
  const int nslots = n + (int)den_slot.size();
  size_t nop = (size_t)(offsets[n] - offsets[0]) * d * d;
  double2 *dops = (double2 *)ctx->alloc(std::max<size_t>(nop, 1) * 16);
  ctx->upload(dops, ops, nop * 16);
  double2 *X = (double2 *)ctx->alloc((size_t)maxD * maxD * 16);
  double2 *M = (double2 *)ctx->alloc((size_t)maxD * maxD * 16);
  double2 *val = (double2 *)ctx->alloc((size_t)nslots * 16);
  double *lg = (double *)ctx->alloc((size_t)nslots * 8);
  // Paths whose sites all have bonds <= SMALL_D_MAX (resident, degree <= 4)
  // go into one launch of the small-bond chain kernel, one workgroup per
  // path; the others run site by site (chain_value).
  auto small_site = [&](int s) {
    if (p->z(s) > 4 || p->pending[s] || p->phys[s] > SMALL_D_MAX) return false;
    for (int b = 0; b < p->z(s); ++b)
      if (p->bond_dim[p->bonds[s][b]] > SMALL_D_MAX) return false;
    return true;
  };
  std::vector<SmallChainSite> csites;
  std::vector<SmallChainJob> cjobs;
  int64_t slab = 1;
  auto add_small = [&](const std::vector<int> &path, const std::vector<int> &kin,
                       const std::vector<int> &kout, const double2 *ops_dev, int slot) {
    for (int s : path)
      if (!small_site(s)) return false;
    cjobs.push_back(SmallChainJob{(int)csites.size(), (int)path.size(), slot});
    for (size_t i = 0; i < path.size(); ++i) {
      const int s = path[i];
      SmallChainSite c{};
      c.V = p->V[s];
      c.op = ops_dev ? ops_dev + i * d * d : nullptr;
      c.z = p->z(s);
      c.kin = kin[i];
      c.kout = kout[i];
      for (int b = 0; b < c.z; ++b) {
        c.dims[b] = p->bond_dim[p->bonds[s][b]];
        c.in_off[b] = p->E_off[p->in_msg(s, b)];
      }
      c.dims[c.z] = p->phys[s];
      slab = std::max<int64_t>(slab, (int64_t)p->site_size(s));
      csites.push_back(c);
    }
    return true;
  };
  for (const auto &kv : den_slot) {
    const int t = (int)(std::find(paths.begin(), paths.end(), kv.first) - paths.begin());
    if (!add_small(kv.first, kins[t], kouts[t], nullptr, kv.second))
      chain_value<S>(ctx, p, kv.first, kins[t], kouts[t], nullptr, X, M, val + kv.second,
                     lg + kv.second);
  }
  for (int t = 0; t < n; ++t) {
    const double2 *ot = dops + (size_t)(offsets[t] - offsets[0]) * d * d;
    if (!add_small(paths[t], kins[t], kouts[t], ot, t))
      chain_value<S>(ctx, p, paths[t], kins[t], kouts[t], ot, X, M, val + t, lg + t);
  }
  if (!cjobs.empty()) {
    const int nj = (int)cjobs.size();
    const int grid = small_sweep_grid(0, nj);
    const size_t scratch_bytes = (size_t)grid * 2 * (size_t)slab * p->elem_bytes();
    void *scratch = ctx->alloc(scratch_bytes);
    SmallChainSite *dcs = (SmallChainSite *)ctx->alloc(csites.size() * sizeof(SmallChainSite));
    SmallChainJob *dcj = (SmallChainJob *)ctx->alloc(nj * sizeof(SmallChainJob));
    ctx->upload(dcs, csites.data(), csites.size() * sizeof(SmallChainSite));
    ctx->upload(dcj, cjobs.data(), nj * sizeof(SmallChainJob));
    launch_chain_small<S>(ctx->stream, dcs, dcj, nj, p->E, scratch, slab, val, lg, grid);
    ctx->release(dcj, nj * sizeof(SmallChainJob));
    ctx->release(dcs, csites.size() * sizeof(SmallChainSite));
    ctx->release(scratch, scratch_bytes);
  }
  int *dden = (int *)ctx->alloc(n * sizeof(int));
  double2 *dout = (double2 *)ctx->alloc((size_t)n * 16);
  ctx->upload(dden, den.data(), n * sizeof(int));
  hipLaunchKernelGGL(k_chain_ratio, dim3((n + 127) / 128), dim3(128), 0, ctx->stream, val, lg,
                     dden, n, dout);
  HBP_HIP(hipGetLastError());
  HBP_HIP(hipMemcpyAsync(out, dout, (size_t)n * 16, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(dout, (size_t)n * 16);
  ctx->release(dden, n * sizeof(int));
  ctx->release(lg, (size_t)nslots * 8);
  ctx->release(val, (size_t)nslots * 16);
  ctx->release(M, (size_t)maxD * maxD * 16);
  ctx->release(X, (size_t)maxD * maxD * 16);
  ctx->release(dops, std::max<size_t>(nop, 1) * 16);
}

}  // namespace hbp

using namespace hbp;
struct hbp_peps : hbp::Peps {};

#define API_BEGIN try {
#define API_END                                   \
  }                                               \
  catch (const ::hbp::Error &e) {                 \
    set_last_error(e.what());                     \
    return e.code;                                \
  }                                               \
  catch (const std::bad_alloc &) {                \
    set_last_error("bad_alloc");                  \
    return HBP_ERR_BAD_ALLOC;                     \
  }                                               \
  catch (const std::exception &e) {               \
    set_last_error(e.what());                     \
    return HBP_ERR_LOGIC;                         \
  }                                               \
  return HBP_OK;

static int site_addr(const hbp_peps *p, int label) {
  auto it = std::lower_bound(p->site_label.begin(), p->site_label.end(), label);
  HBP_REQUIRE(it != p->site_label.end() && *it == label, HBP_ERR_OUT_OF_RANGE, "unknown site");
  return (int)(it - p->site_label.begin());
}

extern "C" {

hbp_status hbp_peps_create(hbp_ctx *ctx, int n_edges, const int32_t *edges,
                           const int32_t *phys_dims, int precision, hbp_peps **out) {
  API_BEGIN
  HBP_REQUIRE(n_edges >= 1, HBP_ERR_INVALID_ARGUMENT, "peps: no edges");
  HBP_REQUIRE(precision == HBP_PREC_FP64 || precision == HBP_PREC_MIXED, HBP_ERR_INVALID_ARGUMENT,
              "peps: unknown precision");
  hbp_peps *p = new hbp_peps();
  p->precision = precision;
  std::vector<int> labels;
  for (int m = 0; m < n_edges; ++m) {
    HBP_REQUIRE(edges[2 * m] != edges[2 * m + 1], HBP_ERR_INVALID_ARGUMENT, "peps: self loop");
    labels.push_back(edges[2 * m]);
    labels.push_back(edges[2 * m + 1]);
  }
  std::sort(labels.begin(), labels.end());
  labels.erase(std::unique(labels.begin(), labels.end()), labels.end());
  p->site_label = labels;
  p->n_sites = (int)labels.size();
  p->n_edges = n_edges;
  p->bonds.assign(p->n_sites, {});
  for (int m = 0; m < n_edges; ++m) {
    const int a = site_addr(p, edges[2 * m]), b = site_addr(p, edges[2 * m + 1]);
    p->ea.push_back(a);
    p->eb.push_back(b);
    p->bonds[a].push_back(m);
    p->bonds[b].push_back(m);
  }
  for (int s = 0; s < p->n_sites; ++s) {
    HBP_REQUIRE((int)p->bonds[s].size() + 1 <= HBP_MAX_ORDER, HBP_ERR_UNSUPPORTED,
                "peps: site degree above 15");
    p->phys.push_back(phys_dims ? phys_dims[s] : 2);
    HBP_REQUIRE(p->phys[s] >= 1 && p->phys[s] <= 16, HBP_ERR_UNSUPPORTED, "peps: physical dim");
  }
  p->bond_dim.assign(n_edges, 1);
  // |0>: bonds of dimension 1, V[0] = 1 (init.h:57-60)
  p->V.resize(p->n_sites);
  p->V_bytes.resize(p->n_sites);
  p->V_borrowed.assign(p->n_sites, 0);
  p->pending.assign(p->n_sites, 0);
  p->pend_O.assign(p->n_sites, nullptr);
  p->pend_O_bytes.assign(p->n_sites, 0);
  p->pend_args.resize(p->n_sites);
  for (int s = 0; s < p->n_sites; ++s) {
    const size_t bytes = (size_t)p->phys[s] * p->elem_bytes();
    p->V[s] = ctx->alloc(bytes);
    p->V_bytes[s] = bytes;
    launch_fill(ctx->stream, p->V[s], p->phys[s], p->site_dtype(), 0.0, 0.0);
    launch_fill(ctx->stream, p->V[s], 1, p->site_dtype(), 1.0, 0.0);
  }
  // 1x1 messages equal to one (init.h:72-78)
  layout_messages(ctx, p, &p->E, &p->E_bytes, &p->E_off);
  launch_fill(ctx->stream, p->E, 2 * n_edges, HBP_C128, 1.0, 0.0);
  p->E_diag.assign(2 * n_edges, 1);
  ctx->sync();
  *out = p;
  API_END
}

hbp_status hbp_peps_destroy(hbp_ctx *ctx, hbp_peps *p) {
  API_BEGIN
  if (!p) return HBP_OK;
  for (int s = 0; s < p->n_sites; ++s) {
    release_site(ctx, p, s);
    if (p->pend_O[s]) ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
  }
  if (p->abs_buf) ctx->release(p->abs_buf, p->abs_bytes);
  drop_pair_cache(ctx, p);
  for (int i = 0; i < 3; ++i)
    if (p->stab[i]) ctx->release(p->stab[i], p->stab_bytes[i]);
  ctx->release(p->E, p->E_bytes);
  for (int i = 0; i < 3; ++i)
    if (p->buf[i]) ctx->release(p->buf[i], p->buf_bytes);
  ctx->sync();
  delete p;
  API_END
}

hbp_status hbp_peps_num_sites(const hbp_peps *p, int *n_sites, int *n_edges) {
  API_BEGIN
  if (n_sites) *n_sites = p->n_sites;
  if (n_edges) *n_edges = p->n_edges;
  API_END
}

hbp_status hbp_peps_site_shape(const hbp_peps *p, int site, int *order, int32_t *shape) {
  API_BEGIN
  HBP_REQUIRE(site >= 0 && site < p->n_sites, HBP_ERR_OUT_OF_RANGE, "site out of range");
  *order = p->z(site) + 1;
  p->shape(site, shape);
  API_END
}

hbp_status hbp_peps_get_site(hbp_ctx *ctx, hbp_peps *p, int site, double *host) {
  API_BEGIN
  HBP_REQUIRE(site >= 0 && site < p->n_sites, HBP_ERR_OUT_OF_RANGE, "site out of range");
  const size_t n = p->site_size(site);
  double2 *tmp = (double2 *)ctx->alloc(n * 16);
  launch_convert(ctx->stream, site_data(ctx, p, site), p->site_dtype(), tmp, HBP_C128, n);
  HBP_HIP(hipMemcpyAsync(host, tmp, n * 16, hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(tmp, n * 16);
  API_END
}

hbp_status hbp_peps_set_site(hbp_ctx *ctx, hbp_peps *p, int site, int order,
                             const int32_t *shape, const double *host) {
  API_BEGIN
  HBP_REQUIRE(site >= 0 && site < p->n_sites, HBP_ERR_OUT_OF_RANGE, "site out of range");
  HBP_REQUIRE(order == p->z(site) + 1, HBP_ERR_INVALID_ARGUMENT, "set_site: order");
  HBP_REQUIRE(shape[order - 1] == p->phys[site], HBP_ERR_INVALID_ARGUMENT, "set_site: phys dim");
  for (int k = 0; k < p->z(site); ++k)
    HBP_REQUIRE(shape[k] == p->bond_dim[p->bonds[site][k]], HBP_ERR_INVALID_ARGUMENT,
                "set_site: bond dims must match the current edge dims");
  const size_t n = p->site_size(site);
  double2 *tmp = (double2 *)ctx->alloc(n * 16);
  ctx->upload(tmp, host, n * 16);
  const size_t bytes = n * p->elem_bytes();
  void *nv = ctx->alloc(bytes);
  launch_convert(ctx->stream, tmp, HBP_C128, nv, p->site_dtype(), n);
  set_resident(ctx, p, site, nv, bytes);
  ctx->sync();
  ctx->release(tmp, n * 16);
  API_END
}

hbp_status hbp_peps_get_message(hbp_ctx *ctx, hbp_peps *p, int m, int32_t *dim, double *host) {
  API_BEGIN
  HBP_REQUIRE(m >= 0 && m < 2 * p->n_edges, HBP_ERR_OUT_OF_RANGE, "message out of range");
  const int D = p->msg_dim(m);
  if (dim) *dim = D;
  if (host) {
    HBP_HIP(hipMemcpyAsync(host, p->E + p->E_off[m], (size_t)D * D * 16, hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->sync();
  }
  API_END
}

hbp_status hbp_peps_set_bond_dims(hbp_ctx *ctx, hbp_peps *p, const int32_t *bond_dims,
                                  const int32_t *phys_dims) {
  API_BEGIN
  for (int m = 0; m < p->n_edges; ++m)
    HBP_REQUIRE(bond_dims[m] >= 1 && bond_dims[m] <= 4096, HBP_ERR_INVALID_ARGUMENT,
                "set_bond_dims: bond dimension out of range");
  if (phys_dims)
    for (int s = 0; s < p->n_sites; ++s)
      HBP_REQUIRE(phys_dims[s] >= 1 && phys_dims[s] <= 16, HBP_ERR_INVALID_ARGUMENT,
                  "set_bond_dims: physical dimension out of range");
  // every site and message is replaced by zeros at the new dimensions
  for (int s = 0; s < p->n_sites; ++s) {
    if (p->pending[s]) {
      ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
      p->pend_O[s] = nullptr;
      p->pending[s] = 0;
    }
  }
  p->abs_site = -1;
  for (int m = 0; m < p->n_edges; ++m) p->bond_dim[m] = bond_dims[m];
  for (int s = 0; s < p->n_sites; ++s) {
    if (phys_dims) p->phys[s] = phys_dims[s];
    const size_t bytes = p->site_size(s) * p->elem_bytes();
    void *nv = ctx->alloc(bytes);
    launch_fill(ctx->stream, nv, p->site_size(s), p->site_dtype(), 0.0, 0.0);
    set_resident(ctx, p, s, nv, bytes);
  }
  double2 *nE;
  size_t nbytes;
  std::vector<int64_t> noff;
  layout_messages(ctx, p, &nE, &nbytes, &noff);
  launch_fill(ctx->stream, nE, nbytes / 16, HBP_C128, 0.0, 0.0);
  ctx->release(p->E, p->E_bytes);
  p->E = nE;
  p->E_bytes = nbytes;
  p->E_off = noff;
  p->E_diag.assign(2 * p->n_edges, 1);
  ctx->sync();
  API_END
}

hbp_status hbp_peps_set_message(hbp_ctx *ctx, hbp_peps *p, int m, int32_t dim,
                                const double *host) {
  API_BEGIN
  HBP_REQUIRE(m >= 0 && m < 2 * p->n_edges, HBP_ERR_OUT_OF_RANGE, "message out of range");
  HBP_REQUIRE(dim == p->msg_dim(m), HBP_ERR_INVALID_ARGUMENT, "set_message: dim");
  bool diag = true;
  for (int c = 0; c < dim && diag; ++c)
    for (int r = 0; r < dim; ++r)
      if (r != c && (host[2 * (r + (size_t)dim * c)] != 0.0 || host[2 * (r + (size_t)dim * c) + 1] != 0.0)) {
        diag = false;
        break;
      }
  p->E_diag[m] = diag ? 1 : 0;
  HBP_HIP(hipMemcpyAsync(p->E + p->E_off[m], host, (size_t)dim * dim * 16, hipMemcpyHostToDevice,
                         ctx->stream));
  ctx->sync();
  API_END
}

hbp_status hbp_peps_bind_sites(hbp_ctx *ctx, hbp_peps *p, const hbp_tensor *const *sites) {
  API_BEGIN
  HBP_REQUIRE(sites, HBP_ERR_INVALID_ARGUMENT, "bind_sites: null");
  const int ns = p->n_sites, ne = p->n_edges;
  std::vector<int> dims(ne, 0);
  for (int s = 0; s < ns; ++s) {
    const hbp_tensor *t = sites[s];
    HBP_REQUIRE(t && t->data, HBP_ERR_INVALID_ARGUMENT, "bind_sites: empty site tensor");
    HBP_REQUIRE(t->dtype == p->site_dtype(), HBP_ERR_INVALID_ARGUMENT,
                "bind_sites: element type differs from the PEPS precision's");
    HBP_REQUIRE(t->order == p->z(s) + 1 && t->shape[p->z(s)] == p->phys[s],
                HBP_ERR_INVALID_ARGUMENT, "bind_sites: site order or physical dim");
    for (int k = 0; k < p->z(s); ++k) {
      const int m = p->bonds[s][k];
      HBP_REQUIRE(dims[m] == 0 || dims[m] == t->shape[k], HBP_ERR_INVALID_ARGUMENT,
                  "bind_sites: bond dims differ at the two ends of an edge");
      dims[m] = t->shape[k];
    }
  }
  drop_pair_cache(ctx, p);
  p->abs_site = -1;
  for (int s = 0; s < ns; ++s) {
    if (p->pending[s]) {
      ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
      p->pend_O[s] = nullptr;
      p->pending[s] = 0;
    }
    if (p->V_borrowed[s] && p->V[s] == sites[s]->data) continue;
    release_site(ctx, p, s);
    p->V[s] = sites[s]->data;
    p->V_bytes[s] = sites[s]->bytes;
    p->V_borrowed[s] = 1;
  }
  if (dims != p->bond_dim) {
    p->bond_dim = dims;
    ctx->release(p->E, p->E_bytes);
    layout_messages(ctx, p, &p->E, &p->E_bytes, &p->E_off);
    launch_fill(ctx->stream, p->E, p->E_bytes / 16, HBP_C128, 0.0, 0.0);
  }
  p->E_diag.assign(2 * ne, 0);
  API_END
}

hbp_status hbp_peps_load_messages(hbp_ctx *ctx, hbp_peps *p, const hbp_tensor *const *msgs) {
  API_BEGIN
  HBP_REQUIRE(msgs, HBP_ERR_INVALID_ARGUMENT, "load_messages: null");
  for (int m = 0; m < 2 * p->n_edges; ++m) {
    const hbp_tensor *t = msgs[m];
    const int D = p->msg_dim(m);
    HBP_REQUIRE(t && t->data && t->dtype == HBP_C128 && t->order == 2 && t->shape[0] == D &&
                    t->shape[1] == D,
                HBP_ERR_INVALID_ARGUMENT,
                "load_messages: message " + std::to_string(m) + " is not a complex128 " +
                    std::to_string(D) + " x " + std::to_string(D) + " matrix");
  }
  for (int m = 0; m < 2 * p->n_edges; ++m) {
    const size_t D = (size_t)p->msg_dim(m);
    HBP_HIP(hipMemcpyAsync(p->E + p->E_off[m], msgs[m]->data, D * D * 16,
                           hipMemcpyDeviceToDevice, ctx->stream));
  }
  // dense unless known otherwise (the diagonal fast paths only save time)
  p->E_diag.assign(2 * p->n_edges, 0);
  API_END
}

hbp_status hbp_peps_store_messages(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *which,
                                   hbp_tensor *const *out) {
  API_BEGIN
  for (int i = 0; i < n; ++i) {
    const int m = which ? which[i] : i;
    HBP_REQUIRE(m >= 0 && m < 2 * p->n_edges, HBP_ERR_OUT_OF_RANGE, "store_messages: index");
    HBP_REQUIRE(out[i], HBP_ERR_INVALID_ARGUMENT, "store_messages: null tensor");
    const int32_t D = p->msg_dim(m);
    const int32_t sh[2] = {D, D};
    tensor_resize(ctx, out[i], HBP_C128, 2, sh);
    HBP_HIP(hipMemcpyAsync(out[i]->data, p->E + p->E_off[m], (size_t)D * D * 16,
                           hipMemcpyDeviceToDevice, ctx->stream));
  }
  API_END
}

hbp_status hbp_peps_store_sites(hbp_ctx *ctx, hbp_peps *p, hbp_tensor *const *out) {
  API_BEGIN
  HBP_REQUIRE(out, HBP_ERR_INVALID_ARGUMENT, "store_sites: null");
  drop_pair_cache(ctx, p);
  for (int s = 0; s < p->n_sites; ++s) {
    hbp_tensor *t = out[s];
    HBP_REQUIRE(t, HBP_ERR_INVALID_ARGUMENT, "store_sites: null tensor");
    if (p->V_borrowed[s] && !p->pending[s] && p->V[s] == t->data) continue;
    int32_t sh[HBP_MAX_ORDER];
    p->shape(s, sh);
    const size_t n = p->site_size(s), bytes = n * 16;
    // the new contents go to fresh memory first: the old tensor may still be
    // this site's (borrowed) source
    void *nv = ctx->alloc(bytes);
    if (p->precision == HBP_PREC_FP64) {
      if (p->pending[s]) materialize(ctx, p, s, nv);
      else HBP_HIP(hipMemcpyAsync(nv, p->V[s], bytes, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      launch_convert(ctx->stream, site_data(ctx, p, s), p->site_dtype(), nv, HBP_C128, n);
    }
    if (t->data) ctx->release(t->data, t->bytes);
    t->data = nv;
    t->bytes = bytes;
    t->dtype = HBP_C128;
    t->order = p->z(s) + 1;
    t->size = n;
    for (int d = 0; d < HBP_MAX_ORDER; ++d) t->shape[d] = d < t->order ? sh[d] : 0;
    touch(t);
    if (p->precision == HBP_PREC_FP64) {
      // the PEPS now reads the caller's tensor
      if (p->pending[s]) {
        ctx->release(p->pend_O[s], p->pend_O_bytes[s]);
        p->pend_O[s] = nullptr;
        p->pending[s] = 0;
      }
      release_site(ctx, p, s);
      p->V[s] = nv;
      p->V_bytes[s] = t->bytes;
      p->V_borrowed[s] = 1;
    }
  }
  p->abs_site = -1;
  API_END
}

// AbsorbTPO (apply.h:14-177) of several PEPS at once: the eager sites' operators of
// every PEPS in one upload, their sites up to kAbsorbBatchMax elements in one
// launch per precision (larger ones one launch each), every message
// expansion E <- E (x) I_t in one launch, one wait at the end (the callers'
// operator arrays may be freed on return).  Each site and message gets the
// arithmetic of a PEPS absorbed on its own.
struct AbsorbMember {
  Peps *p;
  const double *const *ops;
  const int32_t *tpo_shapes;
};
static void absorb_many(Ctx *ctx, const std::vector<AbsorbMember> &mem) {
  constexpr size_t kAbsorbBatchMax = (size_t)1 << 24;
  std::vector<std::vector<int>> tdim(mem.size());
  std::vector<char> lazy(mem.size(), 0);
  // checks, and the state before the absorption (a second absorption before
  // BP composes with the first: that one is stored first)
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    const int32_t *tpo_shapes = mem[i].tpo_shapes;
    const int ns = p->n_sites, ne = p->n_edges;
    std::vector<int> &t = tdim[i];
    t.assign(ne, 0);
    for (int s = 0; s < ns; ++s) {
      const int32_t *sh = tpo_shapes + (size_t)s * HBP_MAX_ORDER;
      const int z = p->z(s);
      HBP_REQUIRE(sh[z] == p->phys[s], HBP_ERR_INVALID_ARGUMENT, "absorb: d_in differs from phys");
      for (int k = 0; k < z; ++k) {
        const int m = p->bonds[s][k];
        HBP_REQUIRE(sh[k] >= 1, HBP_ERR_INVALID_ARGUMENT, "absorb: TPO bond < 1");
        HBP_REQUIRE(t[m] == 0 || t[m] == sh[k], HBP_ERR_INVALID_ARGUMENT,
                    "absorb: TPO bond dims differ across an edge");
        t[m] = sh[k];
      }
    }
  }
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    drop_pair_cache(ctx, p);
    for (int s = 0; s < p->n_sites; ++s) make_resident(ctx, p, s);
    // lazy: keep V_s and O_s, materialise per sweep; eager: store the absorbed
    // lattice (AUTO: eager while it takes at most half of the free device memory)
    if (p->mode == HBP_MODE_FUSED && p->absorb_mode != HBP_ABSORB_EAGER) {
      lazy[i] = 1;
      if (p->absorb_mode == HBP_ABSORB_AUTO) {
        size_t total = 0;
        for (int s = 0; s < p->n_sites; ++s) {
          const int32_t *osh = mem[i].tpo_shapes + (size_t)s * HBP_MAX_ORDER;
          size_t n = (size_t)osh[p->z(s) + 1];
          for (int k = 0; k < p->z(s); ++k) n *= (size_t)p->bond_dim[p->bonds[s][k]] * osh[k];
          total += n * p->elem_bytes();
        }
        // (memory the context pool holds in reserve counts as free)
        lazy[i] = total > ctx->free_bytes() / 2;
      }
    }
  }
  // the eager PEPS' operators in one upload (a pageable copy per site would
  // wait for the previous site's absorption each time)
  std::vector<std::vector<size_t>> op_off(mem.size());
  size_t op_total = 0;
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    op_off[i].assign(p->n_sites + 1, 0);
    for (int s = 0; s < p->n_sites; ++s) {
      const int32_t *osh = mem[i].tpo_shapes + (size_t)s * HBP_MAX_ORDER;
      size_t on = 1;
      for (int k = 0; k <= p->z(s) + 1; ++k) on *= (size_t)osh[k];
      op_off[i][s + 1] = op_off[i][s] + on;
    }
    if (lazy[i]) continue;
    for (int s = 0; s <= p->n_sites; ++s) op_off[i][s] += op_total;
    op_total = op_off[i][p->n_sites];
  }
  double2 *dAll = nullptr;
  if (op_total > 0) {
    std::vector<double2> packed(op_total);
    for (size_t i = 0; i < mem.size(); ++i) {
      if (lazy[i]) continue;
      for (int s = 0; s < mem[i].p->n_sites; ++s)
        std::memcpy(packed.data() + op_off[i][s], mem[i].ops[s],
                    (op_off[i][s + 1] - op_off[i][s]) * 16);
    }
    dAll = (double2 *)ctx->alloc(op_total * 16);
    ctx->upload(dAll, packed.data(), op_total * 16);
  }
  // eager sites up to kAbsorbBatchMax elements: one launch per precision
  std::vector<AbsorbJob> batch[2];
  std::vector<std::pair<Peps *, int>> batch_site[2];
  std::vector<size_t> batch_bytes[2];
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    for (int s = 0; s < p->n_sites; ++s) {
      const int32_t *osh = mem[i].tpo_shapes + (size_t)s * HBP_MAX_ORDER;
      const int z = p->z(s);
      AbsorbArgs a{};
      a.z = z;
      a.din = osh[z];
      a.dout = osh[z + 1];
      int64_t vs = 1, os = 1, ns_ = 1;
      for (int k = 0; k < z; ++k) {
        const int D = p->bond_dim[p->bonds[s][k]];
        a.vdim[k] = D;
        a.vstride[k] = vs;
        vs *= D;
        a.odim[k] = osh[k];
        a.ostride[k] = os;
        os *= osh[k];
        a.ndim[k] = D * osh[k];
        a.nstride[k] = ns_;
        ns_ *= D * osh[k];
      }
      a.vstride[z] = vs;
      a.ostride[z] = os;
      a.ostride[z + 1] = os * osh[z];
      a.ndim[z] = a.dout;
      a.fast = (uint64_t)ns_ * (uint64_t)a.dout < (uint64_t(1) << 32) ? 1 : 0;
      for (int k = 0; k < z; ++k) {
        fast_div_params((uint32_t)a.ndim[k], &a.nm[k], &a.ns[k]);
        fast_div_params((uint32_t)a.vdim[k], &a.vm[k], &a.vs[k]);
      }
      const size_t on = op_off[i][s + 1] - op_off[i][s];
      const size_t n = (size_t)ns_ * a.dout;
      p->phys[s] = a.dout;
      if (lazy[i]) {
        double2 *dO = (double2 *)ctx->alloc(on * 16);
        ctx->upload(dO, mem[i].ops[s], on * 16);
        p->pending[s] = 1;
        p->pend_O[s] = dO;
        p->pend_O_bytes[s] = on * 16;
        p->pend_args[s] = a;
        if (p->abs_site == s) p->abs_site = -1;
        continue;
      }
      const size_t bytes = n * p->elem_bytes();
      void *nv = ctx->alloc(bytes);
      const double2 *dO = dAll + op_off[i][s];
      if (n <= kAbsorbBatchMax) {
        // joins the launch over the small sites below (the old tensor is
        // released once that launch is enqueued)
        const int g = p->precision == HBP_PREC_MIXED ? 1 : 0;
        batch[g].push_back(AbsorbJob{p->V[s], dO, nv, n, a});
        batch_site[g].emplace_back(p, s);
        batch_bytes[g].push_back(bytes);
        continue;
      }
      if (p->precision == HBP_PREC_FP64)
        hipLaunchKernelGGL(k_absorb_site<double2>, dim3(gridn(n)), dim3(256), 0, ctx->stream,
                           (const double2 *)p->V[s], dO, (double2 *)nv, n, a);
      else
        hipLaunchKernelGGL(k_absorb_site<float2>, dim3(gridn(n)), dim3(256), 0, ctx->stream,
                           (const float2 *)p->V[s], dO, (float2 *)nv, n, a);
      HBP_HIP(hipGetLastError());
      release_site(ctx, p, s);
      p->V[s] = nv;
      p->V_bytes[s] = bytes;
    }
  }
  for (int g = 0; g < 2; ++g) {
    if (batch[g].empty()) continue;
    absorb_batch(ctx, batch_site[g][0].first, batch[g]);
    for (size_t j = 0; j < batch[g].size(); ++j) {
      Peps *p = batch_site[g][j].first;
      const int s = batch_site[g][j].second;
      release_site(ctx, p, s);
      p->V[s] = batch[g][j].Vn;
      p->V_bytes[s] = batch_bytes[g][j];
    }
  }
  if (dAll) ctx->release(dAll, op_total * 16);
  // E_e <- E_e (x) I_t for every message of every PEPS in one launch
  std::vector<ExpandJob> xj;
  std::vector<std::pair<double2 *, size_t>> old_E;
  for (size_t i = 0; i < mem.size(); ++i) {
    Peps *p = mem[i].p;
    const int ne = p->n_edges;
    std::vector<int> old = p->bond_dim;
    std::vector<int64_t> old_off = p->E_off;
    for (int m = 0; m < ne; ++m) p->bond_dim[m] = old[m] * tdim[i][m];
    double2 *nE;
    size_t nbytes;
    std::vector<int64_t> noff;
    layout_messages(ctx, p, &nE, &nbytes, &noff);
    for (int m = 0; m < 2 * ne; ++m)
      xj.push_back(ExpandJob{p->E + old_off[m], nE + noff[m], old[m % ne], tdim[i][m % ne]});
    old_E.emplace_back(p->E, p->E_bytes);
    p->E = nE;
    p->E_bytes = nbytes;
    p->E_off = noff;
  }
  if (!xj.empty()) {
    ExpandJob *dxj = (ExpandJob *)ctx->alloc(xj.size() * sizeof(ExpandJob));
    ctx->upload(dxj, xj.data(), xj.size() * sizeof(ExpandJob));
    hipLaunchKernelGGL(k_expand_messages, dim3((unsigned)xj.size()), dim3(256), 0, ctx->stream,
                       dxj);
    HBP_HIP(hipGetLastError());
    ctx->release(dxj, xj.size() * sizeof(ExpandJob));
  }
  for (auto &e : old_E) ctx->release(e.first, e.second);
  ctx->sync();  // the caller's operator arrays may be freed on return
}

hbp_status hbp_peps_absorb_tpo(hbp_ctx *ctx, hbp_peps *p, const double *const *ops,
                               const int32_t *tpo_shapes) {
  API_BEGIN
  absorb_many(ctx, {AbsorbMember{p, ops, tpo_shapes}});
  API_END
}

hbp_status hbp_peps_absorb_tpo_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps,
                                       const double *const *const *ops,
                                       const int32_t *const *tpo_shapes) {
  API_BEGIN
  HBP_REQUIRE(n >= 0 && (n == 0 || (ps && ops && tpo_shapes)), HBP_ERR_INVALID_ARGUMENT,
              "absorb batched: null argument");
  std::vector<AbsorbMember> mem;
  for (int i = 0; i < n; ++i) {
    HBP_REQUIRE(ps[i] && ops[i] && tpo_shapes[i], HBP_ERR_INVALID_ARGUMENT,
                "absorb batched: null PEPS or operator");
    for (int j = 0; j < i; ++j)
      HBP_REQUIRE(ps[i] != ps[j], HBP_ERR_INVALID_ARGUMENT, "absorb batched: a PEPS appears twice");
    mem.push_back(AbsorbMember{ps[i], ops[i], tpo_shapes[i]});
  }
  if (!mem.empty()) absorb_many(ctx, mem);
  API_END
}

hbp_status hbp_peps_set_mode(hbp_peps *p, int mode) {
  API_BEGIN
  HBP_REQUIRE(mode == HBP_MODE_FUSED || mode == HBP_MODE_PLAIN, HBP_ERR_INVALID_ARGUMENT,
              "peps: unknown mode");
  p->mode = mode;
  API_END
}

hbp_status hbp_peps_set_sweep_kernel(hbp_peps *p, int kind) {
  API_BEGIN
  HBP_REQUIRE(kind == HBP_SWEEP_AUTO || kind == HBP_SWEEP_PER_SITE || kind == HBP_SWEEP_GROUPED ||
                  kind == HBP_SWEEP_MERGED,
              HBP_ERR_INVALID_ARGUMENT, "peps: unknown sweep kernel");
  // the value stays reserved; its implementation was taken out
  HBP_REQUIRE(kind != HBP_SWEEP_GROUPED, HBP_ERR_UNSUPPORTED,
              "peps: the grouped sweep is disabled and removed (its first GPU run left the "
              "device unusable, cause not found); hbp_peps_bp_run_batched batches sweeps of "
              "bonds <= 16");
  p->sweep_kernel = kind;
  API_END
}

hbp_status hbp_peps_set_absorb_mode(hbp_peps *p, int mode) {
  API_BEGIN
  HBP_REQUIRE(mode == HBP_ABSORB_AUTO || mode == HBP_ABSORB_LAZY || mode == HBP_ABSORB_EAGER,
              HBP_ERR_INVALID_ARGUMENT, "peps: unknown absorb mode");
  p->absorb_mode = mode;
  API_END
}

hbp_status hbp_peps_set_phase_gauge(hbp_peps *p, int on) {
  API_BEGIN
  p->phase_gauge = on != 0;
  API_END
}

hbp_status hbp_peps_set_pair_cache(hbp_peps *p, int policy) {
  API_BEGIN
  HBP_REQUIRE(policy >= -1 && policy <= 1, HBP_ERR_INVALID_ARGUMENT, "peps: pair cache policy");
  p->pair_cache_policy = policy;
  p->pair_cache = -1;
  API_END
}

hbp_status hbp_peps_set_schedule(hbp_peps *p, int schedule) {
  API_BEGIN
  HBP_REQUIRE(schedule == HBP_SCHEDULE_JACOBI || schedule == HBP_SCHEDULE_CHECKERBOARD,
              HBP_ERR_INVALID_ARGUMENT, "peps: unknown BP schedule");
  p->schedule = schedule;
  API_END
}

hbp_status hbp_peps_set_residual_floor(hbp_peps *p, double floor) {
  API_BEGIN
  HBP_REQUIRE(floor >= 0.0, HBP_ERR_INVALID_ARGUMENT, "peps: negative residual floor");
  p->mixed_floor = floor;
  API_END
}

hbp_status hbp_peps_set_accel(hbp_peps *p, int depth, double start) {
  API_BEGIN
  HBP_REQUIRE(p, HBP_ERR_INVALID_ARGUMENT, "peps: null");
  HBP_REQUIRE(depth >= 0 && depth <= AA_MAX, HBP_ERR_INVALID_ARGUMENT,
              "peps: mixing depth outside 0.." + std::to_string(AA_MAX));
  HBP_REQUIRE(start > 0.0, HBP_ERR_INVALID_ARGUMENT, "peps: mixing start residual <= 0");
  p->aa_depth = depth;
  p->aa_start = start;
  API_END
}

hbp_status hbp_peps_bp_sweep(hbp_ctx *ctx, hbp_peps *p) {
  API_BEGIN
  double2 *nE = (double2 *)ctx->alloc(p->E_bytes);
  sweep(ctx, p, nE, false, nullptr);
  ctx->release(p->E, p->E_bytes);
  p->E = nE;
  p->E_diag.assign(2 * p->n_edges, 0);
  API_END
}

hbp_status hbp_peps_bp_condition(hbp_ctx *ctx, hbp_peps *p, double *residual) {
  API_BEGIN
  double2 *nE = (double2 *)ctx->alloc(p->E_bytes);
  double *dres = (double *)ctx->alloc(sizeof(double));
  sweep(ctx, p, nE, true, dres);
  HBP_HIP(hipMemcpyAsync(residual, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  ctx->sync();
  ctx->release(dres, sizeof(double));
  ctx->release(nE, p->E_bytes);
  API_END
}

// Solve (H + reg I) y = 1 by Gaussian elimination with partial pivoting and
// return a = y / sum(y), the minimiser of |sum_i a_i f_i| with sum_i a_i = 1
// for the Gram matrix H_ij = Re<f_i, f_j>; false when that is ill-posed.
static bool aa_coefficients(std::vector<std::vector<double>> H, std::vector<double> *a) {
  const int w = (int)H.size();
  double tr = 0.0;
  for (int i = 0; i < w; ++i) tr += H[i][i];
  if (!(tr > 0.0) || !std::isfinite(tr)) return false;
  const double reg = 1e-10 * tr / w;
  std::vector<double> y(w, 1.0);
  for (int i = 0; i < w; ++i) H[i][i] += reg;
  for (int c = 0; c < w; ++c) {
    int piv = c;
    for (int r = c + 1; r < w; ++r)
      if (std::fabs(H[r][c]) > std::fabs(H[piv][c])) piv = r;
    if (!(std::fabs(H[piv][c]) > 0.0)) return false;
    std::swap(H[c], H[piv]);
    std::swap(y[c], y[piv]);
    for (int r = c + 1; r < w; ++r) {
      const double q = H[r][c] / H[c][c];
      for (int k = c; k < w; ++k) H[r][k] -= q * H[c][k];
      y[r] -= q * y[c];
    }
  }
  for (int c = w - 1; c >= 0; --c) {
    double t = y[c];
    for (int k = c + 1; k < w; ++k) t -= H[c][k] * y[k];
    y[c] = t / H[c][c];
  }
  double sum = 0.0;
  for (double v : y) sum += v;
  if (!(std::fabs(sum) > 0.0) || !std::isfinite(sum)) return false;
  a->resize(w);
  double l1 = 0.0;
  for (int i = 0; i < w; ++i) {
    (*a)[i] = y[i] / sum;
    l1 += std::fabs((*a)[i]);
  }
  return std::isfinite(l1) && l1 < 1e3;
}

// The fused BP loop with Anderson mixing (type II, constrained form) of the
// message vector: the window keeps the last sweep outputs G_i = BP(E_i)
// and residual vectors f_i = G_i - E_i (at most aa_depth + 1), and the next
// iterate is E = normalise(sum_i a_i G_i) with the real a (Hermitian messages
// stay Hermitian) that minimise |sum_i a_i f_i| under sum_i a_i = 1.  Mixing
// starts once the residual is below aa_start; the window restarts from the
// newest pair when the residual rises above four times its smallest value
// since the window began or the coefficients are ill-posed.  Every iteration
// still computes G = BP(E) and the residual of E exactly as the plain loop
// does (bp.h:266-355), and the loop stops on the same test, keeping E
// (main.cc:138-158): a fixed point of one iteration is one of the other, so
// only the path there changes (and the number of sweeps).  Returns the
// number of iterations run.
static int bp_loop_mixed(Ctx *ctx, Peps *p, int n, double stop, int last, double *dres,
                         double *residuals, int *conv) {
  const int nm = 2 * p->n_edges;
  const size_t bytes = p->E_bytes;
  const int64_t len = (int64_t)(bytes / sizeof(double2));
  const int depth = std::min(p->aa_depth, AA_MAX);
  std::vector<MsgJob> jobs(nm);
  for (int m = 0; m < nm; ++m) jobs[m] = MsgJob{p->E_off[m], p->E_off[m], p->E_off[m], p->msg_dim(m)};
  MsgJob *dj = (MsgJob *)ctx->alloc(nm * sizeof(MsgJob));
  ctx->upload(dj, jobs.data(), nm * sizeof(MsgJob));
  const size_t part_bytes = (size_t)AA_BLOCKS * (AA_MAX + 2) * sizeof(double);
  double *part = (double *)ctx->alloc(part_bytes);
  double *ddots = (double *)ctx->alloc((AA_MAX + 2) * sizeof(double));
  std::vector<double2 *> wG, wf;             // the window, oldest first
  std::vector<std::vector<double>> H;        // H[i][j] = Re<f_i, f_j>
  double rmin = 0.0;
  auto drop_oldest = [&]() {
    ctx->release(wG.front(), bytes);
    ctx->release(wf.front(), bytes);
    wG.erase(wG.begin());
    wf.erase(wf.begin());
    H.erase(H.begin());
    for (auto &row : H) row.erase(row.begin());
  };
  int it = 0;
  *conv = 0;
  for (it = 1; it <= n; ++it) {
    double2 *G = (double2 *)ctx->alloc(bytes);
    sweep(ctx, p, G, true, dres);
    double2 *f = (double2 *)ctx->alloc(bytes);
    AAVecs hv{};
    hv.n = (int)wf.size();
    for (int i = 0; i < hv.n; ++i) hv.v[i] = wf[i];
    const int blocks = (int)std::min<int64_t>(AA_BLOCKS, std::max<int64_t>(1, (len + 255) / 256));
    hipLaunchKernelGGL(k_aa_diff_dots, dim3(blocks), dim3(256), 0, ctx->stream, G, p->E, f, len,
                       hv, part);
    HBP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_aa_sum, dim3(1), dim3(64), 0, ctx->stream, part, hv.n + 1, blocks, ddots);
    HBP_HIP(hipGetLastError());
    double r = 0.0;
    std::vector<double> dots(hv.n + 1);
    HBP_HIP(hipMemcpyAsync(&r, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HBP_HIP(hipMemcpyAsync(dots.data(), ddots, dots.size() * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->stream));
    ctx->sync();
    if (residuals) residuals[it - 1] = r;
    if (r < stop || (it == n && last)) {
      *conv = r < stop ? 1 : 0;
      ctx->release(G, bytes);
      ctx->release(f, bytes);
      break;
    }
    // window update
    const bool restart = r >= p->aa_start || (!wf.empty() && r > 4.0 * rmin);
    if (restart) {
      while (!wf.empty()) drop_oldest();
    } else {
      for (int i = 0; i < (int)wf.size(); ++i) H[i].push_back(dots[i]);
    }
    std::vector<double> row(wf.size() + 1);
    for (size_t i = 0; i < wf.size(); ++i) row[i] = dots[i];
    row.back() = dots[hv.n];
    H.push_back(row);
    wG.push_back(G);
    wf.push_back(f);
    rmin = wf.size() == 1 ? r : std::min(rmin, r);
    while ((int)wf.size() > depth + 1) drop_oldest();
    std::vector<double> a;
    if (wf.size() > 1 && !aa_coefficients(H, &a)) {
      while (wf.size() > 1) drop_oldest();
      rmin = r;
    }
    double2 *nE = (double2 *)ctx->alloc(bytes);
    if (wf.size() == 1) {
      // the plain iteration: E <- BP(E)
      HBP_HIP(hipMemcpyAsync(nE, G, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    } else {
      AAVecs gv{};
      gv.n = (int)wG.size();
      for (int i = 0; i < gv.n; ++i) {
        gv.v[i] = wG[i];
        gv.c[i] = a[i];
      }
      double2 *comb = (double2 *)ctx->alloc(bytes);
      hipLaunchKernelGGL(k_aa_combine, dim3(gridn((size_t)len)), dim3(256), 0, ctx->stream, gv,
                         comb, len);
      HBP_HIP(hipGetLastError());
      hipLaunchKernelGGL(k_finalize_messages, dim3(nm), dim3(256), 0, ctx->stream, comb, nE,
                         nullptr, dj, nullptr, p->phase_gauge ? 1 : 0, nullptr);
      HBP_HIP(hipGetLastError());
      ctx->release(comb, bytes);
    }
    ctx->release(p->E, bytes);
    p->E = nE;
    p->E_diag.assign(nm, 0);
  }
  while (!wf.empty()) drop_oldest();
  ctx->release(ddots, (AA_MAX + 2) * sizeof(double));
  ctx->release(part, part_bytes);
  ctx->release(dj, nm * sizeof(MsgJob));
  return std::min(it, n);
}

// Up to n iterations of the driver's BP loop (main.cc:121-159).  `first`
// starts a loop: the fused path then runs the loop's first sweep E_1 = BP(E_0).
// Each fused iteration computes G = BP(E) once: G gives the condition's
// residual against E, and becomes the next E unless the loop stops there.
// When n iterations end without convergence and `last` is 0, G is kept so
// that the next call continues the same sequence; with `last` the loop stops
// on E like the reference at max_bp_iterations.
static void bp_iterate(Ctx *ctx, Peps *p, int n, double tol, int first, int last, int *iters,
                       double *residuals, int *converged) {
  HBP_REQUIRE(n >= 1, HBP_ERR_INVALID_ARGUMENT, "bp: iteration count < 1");
  const double stop = p->precision == HBP_PREC_MIXED ? std::max(tol, p->mixed_floor) : tol;
  double *dres = (double *)ctx->alloc(sizeof(double));
  int it = 0, conv = 0;
  if (p->mode == HBP_MODE_PLAIN) {
    // every iteration is a sweep (E <- F) followed by a separate condition sweep
    for (it = 1; it <= n; ++it) {
      double2 *nE = (double2 *)ctx->alloc(p->E_bytes);
      sweep(ctx, p, nE, false, nullptr);
      ctx->release(p->E, p->E_bytes);
      p->E = nE;
      p->E_diag.assign(2 * p->n_edges, 0);
      double2 *G = (double2 *)ctx->alloc(p->E_bytes);
      sweep(ctx, p, G, true, dres);
      ctx->release(G, p->E_bytes);
      double r = 0.0;
      HBP_HIP(hipMemcpyAsync(&r, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
      ctx->sync();
      if (residuals) residuals[it - 1] = r;
      if (r < stop) {
        conv = 1;
        break;
      }
    }
  } else {
    if (first) {
      double2 *nE = (double2 *)ctx->alloc(p->E_bytes);
      sweep(ctx, p, nE, false, nullptr);
      ctx->release(p->E, p->E_bytes);
      p->E = nE;
      p->E_diag.assign(2 * p->n_edges, 0);
    }
    if (p->aa_depth > 0 && p->n_edges > 0) {
      it = bp_loop_mixed(ctx, p, n, stop, last, dres, residuals, &conv);
    } else {
      for (it = 1; it <= n; ++it) {
        double2 *G = (double2 *)ctx->alloc(p->E_bytes);
        sweep(ctx, p, G, true, dres);
        double r = 0.0;
        HBP_HIP(hipMemcpyAsync(&r, dres, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        ctx->sync();
        if (residuals) residuals[it - 1] = r;
        if (r < stop || (it == n && last)) {
          conv = r < stop ? 1 : 0;
          ctx->release(G, p->E_bytes);
          break;
        }
        ctx->release(p->E, p->E_bytes);
        p->E = G;
        p->E_diag.assign(2 * p->n_edges, 0);
      }
    }
  }
  if (iters) *iters = std::min(it, n);
  if (converged) *converged = conv;
  ctx->release(dres, sizeof(double));
}

}  // extern "C"

// ------------------------------------------------------ batched BP loops
// PEPS the batched small-bond sweep takes: fused mode with the reference's
// Jacobi schedule and no mixing, degree 1..4, every bond <= SMALL_D_MAX.
static bool small_eligible(const Peps *p) {
  if (p->mode != HBP_MODE_FUSED || p->schedule != HBP_SCHEDULE_JACOBI || p->aa_depth > 0)
    return false;
  for (int s = 0; s < p->n_sites; ++s)
    if (p->z(s) < 1 || p->z(s) > 4) return false;
  for (int m = 0; m < p->n_edges; ++m)
    if (p->bond_dim[m] > SMALL_D_MAX) return false;
  return true;
}

// PEPS whose own fused sweep (per-site launches) the batch records into its
// per-iteration graph: fused mode, Jacobi schedule, no mixing.
static bool graph_eligible(const Peps *p) {
  return p->mode == HBP_MODE_FUSED && p->schedule == HBP_SCHEDULE_JACOBI && p->aa_depth == 0 &&
         p->n_edges > 0;
}

namespace {
// one PEPS in a batched BP loop
struct BatchMember {
  Peps *p = nullptr;
  int kind = 0;            // HBP_BATCH_SMALL or HBP_BATCH_GRAPH
  double2 *G = nullptr;    // the other message set (E <-> G swap per iteration)
  double2 *raw = nullptr;  // raw messages of the small-bond sweep
  MsgJob *dj = nullptr;    // message table of the PEPS's own sweep
  int slot = 0;            // position of its residual (own sweep)
  int index = 0;           // position in the caller's arrays
  double stop = 0.0;
  int iters = 0, conv = 0;
  bool done = false;
  std::vector<double> history;  // residual of every iteration
};

// The work of one sweep of every member (E -> raw -> G), uploaded to the
// device: site jobs for the batched kernel, message jobs for the batched
// finalisation, each member's message range for its residual.
struct BatchTables {
  SmallSiteJob *sites = nullptr;
  MsgJobB *msgs = nullptr;
  int *first = nullptr;
  double *vol = nullptr;
  int nsites = 0, nmsgs = 0, nmem = 0;
  size_t bytes_sites = 0, bytes_msgs = 0, bytes_first = 0, bytes_vol = 0;
  void release(Ctx *ctx) {
    if (sites) ctx->release(sites, bytes_sites);
    if (msgs) ctx->release(msgs, bytes_msgs);
    if (first) ctx->release(first, bytes_first);
    if (vol) ctx->release(vol, bytes_vol);
    *this = BatchTables{};
  }
};

BatchTables build_tables(Ctx *ctx, const std::vector<BatchMember *> &mem, bool residual) {
  std::vector<SmallSiteJob> sj;
  std::vector<MsgJobB> mj;
  std::vector<int> first{0};
  std::vector<double> vol;
  for (BatchMember *b : mem) {
    Peps *p = b->p;
    for (int s = 0; s < p->n_sites; ++s) {
      SmallSiteJob j{};
      j.V = p->V[s];
      j.E = p->E;
      j.raw = b->raw;
      j.z = p->z(s);
      for (int k = 0; k < j.z; ++k) {
        j.dims[k] = p->bond_dim[p->bonds[s][k]];
        j.in_off[k] = p->E_off[p->in_msg(s, k)];
        j.out_off[k] = p->E_off[p->out_msg(s, k)];
      }
      j.dims[j.z] = p->phys[s];
      sj.push_back(j);
    }
    for (int m = 0; m < 2 * p->n_edges; ++m)
      mj.push_back(MsgJobB{b->raw + p->E_off[m], b->G + p->E_off[m],
                           residual ? p->E + p->E_off[m] : nullptr, p->msg_dim(m),
                           p->phase_gauge ? 1 : 0});
    first.push_back((int)mj.size());
    vol.push_back(1.0 / (2.0 * p->n_edges));
  }
  BatchTables t;
  t.nsites = (int)sj.size();
  t.nmsgs = (int)mj.size();
  t.nmem = (int)mem.size();
  t.bytes_sites = std::max<size_t>(sj.size(), 1) * sizeof(SmallSiteJob);
  t.bytes_msgs = std::max<size_t>(mj.size(), 1) * sizeof(MsgJobB);
  t.bytes_first = first.size() * sizeof(int);
  t.bytes_vol = std::max<size_t>(vol.size(), 1) * sizeof(double);
  t.sites = (SmallSiteJob *)ctx->alloc(t.bytes_sites);
  t.msgs = (MsgJobB *)ctx->alloc(t.bytes_msgs);
  t.first = (int *)ctx->alloc(t.bytes_first);
  t.vol = (double *)ctx->alloc(t.bytes_vol);
  ctx->upload(t.sites, sj.data(), sj.size() * sizeof(SmallSiteJob));
  ctx->upload(t.msgs, mj.data(), mj.size() * sizeof(MsgJobB));
  ctx->upload(t.first, first.data(), first.size() * sizeof(int));
  if (!vol.empty()) ctx->upload(t.vol, vol.data(), vol.size() * sizeof(double));
  return t;
}

// one sweep of every member from the tables: the batched site kernel with
// the grid `variant`, finalisation, residuals (when the tables carry the old
// messages); nothing here synchronises, so it can be captured
template <typename S>
void launch_batch_sweep(Ctx *ctx, const BatchTables &t, void *scratch, int64_t slab, int variant,
                        double *dist, double *res, bool residual) {
  launch_sweep_small<S>(ctx->stream, t.sites, t.nsites, scratch, slab,
                        small_sweep_grid(variant, SMALL_SWEEP_ITEMS * t.nsites));
  hipLaunchKernelGGL(k_finalize_batched, dim3((unsigned)t.nmsgs), dim3(256), 0, ctx->stream,
                     t.msgs, dist);
  HBP_HIP(hipGetLastError());
  if (residual) {
    hipLaunchKernelGGL(k_residual_batched, dim3((unsigned)t.nmem), dim3(256), 0, ctx->stream,
                       dist, t.first, t.vol, res);
    HBP_HIP(hipGetLastError());
  }
}

// The members of one precision that the small-bond kernel takes, with the
// kernel's scratch (two site-sized buffers per workgroup of the largest grid)
// and per-parity tables.  act[i]'s residual lands at res[res_off + i].
struct SmallGroup {
  bool fp32 = false;
  std::vector<BatchMember *> act;
  int64_t slab = 1;
  void *scratch = nullptr;
  size_t scratch_bytes = 0;
  double *dist = nullptr;
  size_t dist_bytes = 0;
  int variant = 0;
  int res_off = 0;
  BatchTables tabs[2];
  bool have[2] = {false, false};
  void sweep(Ctx *ctx, int parity, double *res) const {
    if (fp32)
      launch_batch_sweep<float>(ctx, tabs[parity], scratch, slab, variant, dist, res + res_off, true);
    else
      launch_batch_sweep<double>(ctx, tabs[parity], scratch, slab, variant, dist, res + res_off,
                                 true);
  }
};
}  // namespace

// The driver's BP loop (as bp_iterate with first = last = 1) for several
// PEPS in lockstep: every iteration advances every member still iterating by
// one sweep (E -> G) and computes its residual; a member leaves when its
// residual is below its stop value or at max_iter, keeping E like the
// reference (main.cc:138-158).  Members with bonds <= 16 run the persistent
// small-bond kernel, one launch per precision over all their sites; the
// others run their own fused sweeps, which read the device tables prepared
// here and are bitwise those of bp_iterate.  One iteration of all members is
// recorded into a HIP graph per message-set parity and replayed until a
// member leaves; the first iteration (and the loop's first sweep) runs
// unrecorded, so that the autotuner's trials, the work buffers, the
// pair-front copies and the block cache are settled before recording.  A
// recording that meets anything a graph cannot hold (graph_capturing) is
// dropped and the loop goes on without graphs.
static void bp_run_batched_impl(Ctx *ctx, std::vector<BatchMember> &mem, int max_iter) {
  SmallGroup grp[2];
  grp[1].fp32 = true;
  std::vector<BatchMember *> big;
  for (auto &b : mem) {
    Peps *p = b.p;
    if (b.kind == HBP_BATCH_SMALL) {
      SmallGroup &g = grp[p->precision == HBP_PREC_MIXED ? 1 : 0];
      drop_pair_cache(ctx, p);
      for (int s = 0; s < p->n_sites; ++s) make_resident(ctx, p, s);
      for (int s = 0; s < p->n_sites; ++s)
        g.slab = std::max<int64_t>(g.slab, (int64_t)p->site_size(s));
      b.raw = (double2 *)ctx->alloc(std::max<size_t>(p->E_bytes, 16));
      g.act.push_back(&b);
    } else {
      const int nm = 2 * p->n_edges;
      std::vector<MsgJob> jobs(nm);
      for (int m = 0; m < nm; ++m)
        jobs[m] = MsgJob{p->E_off[m], p->E_off[m], p->E_off[m], p->msg_dim(m)};
      b.dj = (MsgJob *)ctx->alloc(nm * sizeof(MsgJob));
      ctx->upload(b.dj, jobs.data(), nm * sizeof(MsgJob));
      big.push_back(&b);
    }
    b.G = (double2 *)ctx->alloc(std::max<size_t>(p->E_bytes, 16));
  }
  int nres = 0;
  for (auto &g : grp) {
    g.res_off = nres;
    nres += (int)g.act.size();
  }
  for (auto *b : big) b->slot = nres++;
  const size_t res_bytes = (size_t)std::max(nres, 1) * sizeof(double);
  double *res = (double *)ctx->alloc(res_bytes);
  for (auto &g : grp) {
    if (g.act.empty()) continue;
    int nsites = 0, nmsgs = 0;
    for (auto *b : g.act) {
      nsites += b->p->n_sites;
      nmsgs += 2 * b->p->n_edges;
    }
    const int grid_max = small_sweep_grid(SMALL_GRID_VARIANTS - 1, SMALL_SWEEP_ITEMS * nsites);
    g.scratch_bytes = (size_t)grid_max * 2 * (size_t)g.slab * 2 * (g.fp32 ? sizeof(float) : sizeof(double));
    g.scratch = ctx->alloc(g.scratch_bytes);
    g.dist_bytes = (size_t)std::max(nmsgs, 1) * sizeof(double);
    g.dist = (double *)ctx->alloc(g.dist_bytes);
  }
  auto swap_sets = [&](BatchMember &b) {
    std::swap(b.p->E, b.G);
    b.p->E_diag.assign(2 * b.p->n_edges, 0);
  };
  // the loop's first sweep E_1 = BP(E_0), no residual (main.cc:121-141)
  for (auto &g : grp) {
    if (g.act.empty()) continue;
    BatchTables t = build_tables(ctx, g.act, false);
    g.variant = g.fp32 ? launch_sweep_small_tuned<float>(ctx->stream, t.sites, t.nsites, g.scratch, g.slab)
                       : launch_sweep_small_tuned<double>(ctx->stream, t.sites, t.nsites, g.scratch, g.slab);
    hipLaunchKernelGGL(k_finalize_batched, dim3((unsigned)t.nmsgs), dim3(256), 0, ctx->stream,
                       t.msgs, g.dist);
    HBP_HIP(hipGetLastError());
    ctx->sync();
    t.release(ctx);
  }
  for (auto *b : big) sweep(ctx, b->p, b->G, false, nullptr, b->dj);
  for (auto &b : mem) swap_sets(b);

  std::vector<double> hres((size_t)std::max(nres, 1));
  hipGraphExec_t exec[2] = {nullptr, nullptr};
  bool graphs = true;                        // false once a recording failed
  const int unrecorded = big.empty() ? 0 : 1;  // iterations run before recording
  auto drop_graphs = [&]() {
    for (int q = 0; q < 2; ++q) {
      if (exec[q]) (void)hipGraphExecDestroy(exec[q]);
      exec[q] = nullptr;
      for (auto &g : grp)
        if (g.have[q]) {
          g.tabs[q].release(ctx);
          g.have[q] = false;
        }
    }
  };
  // one iteration of every active member (E -> G and residuals), enqueued
  auto enqueue = [&](int parity) {
    for (auto &g : grp)
      if (!g.act.empty()) g.sweep(ctx, parity, res);
    for (auto *b : big) sweep(ctx, b->p, b->G, true, res + b->slot, b->dj);
  };
  // Tables of both parities are built together, before either graph is
  // recorded: a recorded sweep takes its temporaries from the block cache and
  // gives them back at the end of the recording, so nothing may allocate
  // outside a recording while a graph that uses those blocks can still be
  // replayed (both graphs are dropped together when a member leaves).
  auto build_both = [&](int parity) {
    for (auto &g : grp) {
      if (g.act.empty() || g.have[0] || g.have[1]) continue;
      g.tabs[parity] = build_tables(ctx, g.act, true);
      for (auto *b : g.act) std::swap(b->p->E, b->G);  // the other parity's sets
      g.tabs[parity ^ 1] = build_tables(ctx, g.act, true);
      for (auto *b : g.act) std::swap(b->p->E, b->G);
      g.have[0] = g.have[1] = true;
    }
  };
  int parity = 0;
  for (int it = 1; it <= max_iter; ++it) {
    if (grp[0].act.empty() && grp[1].act.empty() && big.empty()) break;
    build_both(parity);
    bool replay = graphs && it > unrecorded;
    if (replay && !exec[parity]) {
      hipGraph_t gr = nullptr;
      try {
        HBP_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        graph_capturing() = true;
        enqueue(parity);
        graph_capturing() = false;
        HBP_HIP(hipStreamEndCapture(ctx->stream, &gr));
        HBP_HIP(hipGraphInstantiate(&exec[parity], gr, nullptr, nullptr, 0));
        (void)hipGraphDestroy(gr);
        gr = nullptr;
        ++ctx->graphs_recorded;
      } catch (const Error &) {
        // nothing of the recording ran: forget host state that it advanced
        // (the absorbed site held in the work buffer, pair-front copies whose
        // rearrangement was only recorded) and go on without graphs
        graph_capturing() = false;
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(ctx->stream, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        if (gr) (void)hipGraphDestroy(gr);
        if (exec[parity]) (void)hipGraphExecDestroy(exec[parity]);
        exec[parity] = nullptr;
        (void)hipGetLastError();
        for (auto *b : big) {
          b->p->abs_site = -1;
          drop_pair_cache(ctx, b->p);
        }
        graphs = replay = false;
        ++ctx->graphs_dropped;
      }
    }
    if (replay) {
      HBP_HIP(hipGraphLaunch(exec[parity], ctx->stream));
      ++ctx->graph_replays;
    } else
      enqueue(parity);
    HBP_HIP(hipMemcpyAsync(hres.data(), res, (size_t)nres * sizeof(double), hipMemcpyDeviceToHost,
                           ctx->stream));
    ctx->sync();
    bool left = false;
    // true while b iterates on (its sets swapped: E <- G)
    auto advance = [&](BatchMember *b, double r) {
      b->iters = it;
      b->history.push_back(r);
      b->conv = r < b->stop ? 1 : 0;
      b->done = b->conv || it == max_iter;
      if (b->done) {
        left = true;  // keeps E, like the reference at its stop (main.cc:138-158)
        return false;
      }
      swap_sets(*b);
      return true;
    };
    for (auto &g : grp) {
      std::vector<BatchMember *> next;
      for (size_t i = 0; i < g.act.size(); ++i)
        if (advance(g.act[i], hres[g.res_off + i])) next.push_back(g.act[i]);
      g.act.swap(next);
    }
    std::vector<BatchMember *> next;
    for (auto *b : big)
      if (advance(b, hres[b->slot])) next.push_back(b);
    big.swap(next);
    if (left) drop_graphs();
    parity ^= 1;
  }
  drop_graphs();
  ctx->release(res, res_bytes);
  for (auto &g : grp) {
    if (g.scratch) ctx->release(g.scratch, g.scratch_bytes);
    if (g.dist) ctx->release(g.dist, g.dist_bytes);
  }
  for (auto &b : mem) {
    ctx->release(b.G, std::max<size_t>(b.p->E_bytes, 16));
    if (b.raw) ctx->release(b.raw, std::max<size_t>(b.p->E_bytes, 16));
    if (b.dj) ctx->release(b.dj, 2 * b.p->n_edges * sizeof(MsgJob));
  }
}

extern "C" {

hbp_status hbp_peps_bp_run_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps, int max_iter,
                                   double tol, int *iters, double *residuals, int *converged,
                                   int *batched) {
  API_BEGIN
  HBP_REQUIRE(n >= 0 && (n == 0 || ps), HBP_ERR_INVALID_ARGUMENT, "bp batched: no PEPS");
  HBP_REQUIRE(max_iter >= 1, HBP_ERR_INVALID_ARGUMENT, "bp: iteration count < 1");
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j)
      HBP_REQUIRE(ps[i] != ps[j], HBP_ERR_INVALID_ARGUMENT, "bp batched: a PEPS appears twice");
  std::vector<BatchMember> mem;
  std::vector<int> alone;
  for (int i = 0; i < n; ++i) {
    Peps *p = ps[i];
    const int kind = small_eligible(p) ? HBP_BATCH_SMALL : graph_eligible(p) ? HBP_BATCH_GRAPH : 0;
    if (kind) {
      BatchMember b;
      b.p = p;
      b.kind = kind;
      b.index = i;
      b.stop = p->precision == HBP_PREC_MIXED ? std::max(tol, p->mixed_floor) : tol;
      mem.push_back(b);
    } else {
      alone.push_back(i);
    }
  }
  auto report = [&](int i, int it, int conv, const double *hist, int nh, int kind) {
    if (iters) iters[i] = it;
    if (converged) converged[i] = conv;
    if (batched) batched[i] = kind;
    if (residuals)
      for (int k = 0; k < nh && k < max_iter; ++k) residuals[(size_t)i * max_iter + k] = hist[k];
  };
  if (!mem.empty()) bp_run_batched_impl(ctx, mem, max_iter);
  for (auto &b : mem) report(b.index, b.iters, b.conv, b.history.data(), (int)b.history.size(), b.kind);
  for (int i : alone) {
    int it = 0, conv = 0;
    std::vector<double> h(max_iter, 0.0);
    bp_iterate(ctx, ps[i], max_iter, tol, 1, 1, &it, h.data(), &conv);
    report(i, it, conv, h.data(), it, 0);
  }
  API_END
}

hbp_status hbp_peps_bp_run(hbp_ctx *ctx, hbp_peps *p, int max_iter, double tol, int *iters,
                           double *residuals) {
  API_BEGIN
  bp_iterate(ctx, p, max_iter, tol, 1, 1, iters, residuals, nullptr);
  API_END
}

hbp_status hbp_peps_bp_iterate(hbp_ctx *ctx, hbp_peps *p, int n, double tol, int first,
                               int last, int *iters, double *residuals, int *converged) {
  API_BEGIN
  bp_iterate(ctx, p, n, tol, first, last, iters, residuals, converged);
  API_END
}

hbp_status hbp_peps_truncate(hbp_ctx *ctx, hbp_peps *p, int chi_max, double sv_min,
                             int32_t *bond_dims, double *trunc_errs) {
  API_BEGIN
  HBP_REQUIRE(chi_max >= 1, HBP_ERR_INVALID_ARGUMENT, "truncate: chi_max < 1");
  truncate(ctx, p, chi_max, sv_min, bond_dims, trunc_errs);
  API_END
}

hbp_status hbp_peps_truncate_batched(hbp_ctx *ctx, int n, hbp_peps *const *ps,
                                     const int32_t *chi_max, const double *sv_min,
                                     int32_t *bond_dims, double *trunc_errs) {
  API_BEGIN
  HBP_REQUIRE(n >= 0 && (n == 0 || (ps && chi_max && sv_min)), HBP_ERR_INVALID_ARGUMENT,
              "truncate batched: null argument");
  std::vector<TruncMember> mem(n);
  int64_t off = 0;
  for (int i = 0; i < n; ++i) {
    HBP_REQUIRE(ps[i], HBP_ERR_INVALID_ARGUMENT, "truncate batched: null PEPS");
    HBP_REQUIRE(chi_max[i] >= 1, HBP_ERR_INVALID_ARGUMENT, "truncate: chi_max < 1");
    for (int j = 0; j < i; ++j)
      HBP_REQUIRE(ps[i] != ps[j], HBP_ERR_INVALID_ARGUMENT, "truncate batched: a PEPS appears twice");
    mem[i].p = ps[i];
    mem[i].chi = chi_max[i];
    mem[i].sv_min = sv_min[i];
    mem[i].dims_out = bond_dims ? bond_dims + off : nullptr;
    mem[i].errs_out = trunc_errs ? trunc_errs + off : nullptr;
    off += ps[i]->n_edges;
  }
  if (n > 0) truncate_many(ctx, mem);
  API_END
}

hbp_status hbp_peps_measure_sites(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *sites,
                                  const double *ops, double *out) {
  API_BEGIN
  if (n <= 0) return HBP_OK;
  std::vector<int> addr(n);
  for (int t = 0; t < n; ++t) addr[t] = site_addr(p, sites[t]);
  if (p->precision == HBP_PREC_FP64)
    measure_sites_impl<double>(ctx, p, n, addr.data(), ops, out);
  else
    measure_sites_impl<float>(ctx, p, n, addr.data(), ops, out);
  API_END
}

hbp_status hbp_peps_measure_edges(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *pairs,
                                  const double *ops, double *out) {
  API_BEGIN
  if (n <= 0) return HBP_OK;
  std::vector<int> a(n), b(n);
  for (int t = 0; t < n; ++t) {
    a[t] = site_addr(p, pairs[2 * t]);
    b[t] = site_addr(p, pairs[2 * t + 1]);
  }
  if (p->precision == HBP_PREC_FP64)
    measure_edges_impl<double>(ctx, p, n, a.data(), b.data(), ops, out);
  else
    measure_edges_impl<float>(ctx, p, n, a.data(), b.data(), ops, out);
  API_END
}

hbp_status hbp_peps_measure_paths(hbp_ctx *ctx, hbp_peps *p, int n, const int32_t *offsets,
                                  const int32_t *sites, const double *ops, double *out) {
  API_BEGIN
  if (n <= 0) return HBP_OK;
  HBP_REQUIRE(offsets[0] == 0, HBP_ERR_INVALID_ARGUMENT, "measure_paths: offsets[0] != 0");
  for (int t = 0; t < n; ++t)
    HBP_REQUIRE(offsets[t + 1] > offsets[t] && offsets[t + 1] - offsets[t] < 64,
                HBP_ERR_INVALID_ARGUMENT, "measure_paths: path length outside 1..63");
  std::vector<int> addr(offsets[n]);
  for (int i = 0; i < offsets[n]; ++i) addr[i] = site_addr(p, sites[i]);
  if (p->precision == HBP_PREC_FP64)
    measure_paths_impl<double>(ctx, p, n, offsets, addr.data(), ops, out);
  else
    measure_paths_impl<float>(ctx, p, n, offsets, addr.data(), ops, out);
  API_END
}

}  // extern "C"
