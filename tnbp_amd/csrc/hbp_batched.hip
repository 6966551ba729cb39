// Batched BP sweep for small bond dimensions: every site of every PEPS of a
// batch in one persistent launch.
//
// At bond dimensions up to SMALL_D_MAX a site tensor has at most
// 16^4 * 2 elements and its messages are a few hundred numbers: the per-site
// launches of the fused sweep (a dozen kernels of a few microseconds per
// site) are bound by launch and host overhead, not by the device.  Here one
// workgroup computes all outgoing messages of a site, walking the work items
// (site, PEPS) of the whole batch with a fixed stride, so one launch covers a
// BP sweep of every lattice in the batch (bp.h:128-175 per message:
// T = V x_{l != k} E_in(l), M_k = <T, conj V> over all bonds but k), and the
// launch sequence of a BP iteration is captured once in a HIP graph and
// replayed (hbp_peps.hip).  The mode products and contractions are VALU work
// (FP64 and FP32 vector rates of gfx950 equal its MFMA FP64 rate; tensors of
// this size would leave MFMA tiles mostly empty); sums run in double in a
// fixed order, so a site's messages do not depend on the batch around it.
#include <hip/hip_runtime.h>

#include "hbp_internal.h"

namespace hbp {
namespace {

template <typename S> struct CT;
template <> struct CT<double> { using type = double2; };
template <> struct CT<float> { using type = float2; };

// Out[a, j, b] = sum_i In[a, i, b] E[i, j] over bond `l` (In: L x D x R)
// with E staged in LDS as complex double; every thread of the block takes
// outputs tid, tid + 256, ...  All reads of In happen-before the barrier at
// the caller's next step.
template <typename S>
__device__ void small_mode_product(const typename CT<S>::type *__restrict__ in,
                                   typename CT<S>::type *__restrict__ out, int64_t L, int D,
                                   int64_t R, const double2 *__restrict__ Es) {
  using C = typename CT<S>::type;
  const int64_t n = L * D * R;
  for (int64_t o = threadIdx.x; o < n; o += blockDim.x) {
    const int64_t a = o % L, t = o / L;
    const int j = (int)(t % D);
    const int64_t b = t / D;
    const C *src = in + a + L * D * b;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < D; ++i) {
      const C x = src[L * i];
      const double2 e = Es[i + D * j];
      re += (double)x.x * e.x - (double)x.y * e.y;
      im += (double)x.x * e.y + (double)x.y * e.x;
    }
    C y;
    y.x = (S)re;
    y.y = (S)im;
    out[o] = y;
  }
}

// M[i, j] = sum_{a, b} T[a, i, b] conj V[a, j, b] (L x D x R views) into
// raw (complex double, D x D column-major): thread (pair p, slice s) sums
// the rows q = s, s + ns, ... of pair p; the slices are added in order.
template <typename S>
__device__ void small_message(const typename CT<S>::type *__restrict__ T,
                              const typename CT<S>::type *__restrict__ V, double2 *__restrict__ M,
                              int64_t L, int D, int64_t R, double2 *red) {
  using C = typename CT<S>::type;
  const int np = D * D;
  const int ns = (int)blockDim.x / np;  // >= 1 for D <= 16
  const int tid = threadIdx.x;
  const int64_t Q = L * R;
  if (tid < np * ns) {
    const int p = tid % np, s = tid / np;
    const int i = p % D, j = p / D;
    double re = 0.0, im = 0.0;
    for (int64_t q = s; q < Q; q += ns) {
      const int64_t a = q % L, b = q / L;
      const C x = T[a + L * (i + (int64_t)D * b)];
      const C v = V[a + L * (j + (int64_t)D * b)];
      // x conj(v)
      re += (double)x.x * v.x + (double)x.y * v.y;
      im += (double)x.y * v.x - (double)x.x * v.y;
    }
    red[tid] = make_double2(re, im);
  }
  __syncthreads();
  if (tid < np) {
    double re = 0.0, im = 0.0;
    for (int s = 0; s < ns; ++s) {
      re += red[tid + np * s].x;
      im += red[tid + np * s].y;
    }
    M[tid] = make_double2(re, im);
  }
  __syncthreads();
}

template <typename S>
__global__ __launch_bounds__(256) void k_sweep_small(const SmallSiteJob *__restrict__ jobs,
                                                     int njobs, void *__restrict__ scratch,
                                                     int64_t slab_elems) {
  using C = typename CT<S>::type;
  __shared__ double2 Es[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double2 red[256];
  C *bufA = (C *)scratch + (int64_t)blockIdx.x * 2 * slab_elems;
  C *bufB = bufA + slab_elems;
  for (int w = blockIdx.x; w < njobs; w += gridDim.x) {
    const SmallSiteJob job = jobs[w];
    const int z = job.z;
    const C *V = (const C *)job.V;
    for (int k = 0; k < z; ++k) {
      const C *cur = V;
      C *nxt = bufA;
      for (int l = 0; l < z; ++l) {
        if (l == k) continue;
        int64_t L = 1, R = 1;
        for (int d = 0; d < l; ++d) L *= job.dims[d];
        for (int d = l + 1; d <= z; ++d) R *= job.dims[d];
        const int D = job.dims[l];
        const double2 *E = job.E + job.in_off[l];
        for (int e = threadIdx.x; e < D * D; e += blockDim.x) Es[e] = E[e];
        __syncthreads();
        small_mode_product<S>(cur, nxt, L, D, R, Es);
        __syncthreads();
        cur = nxt;
        nxt = nxt == bufA ? bufB : bufA;
      }
      int64_t L = 1, R = 1;
      for (int d = 0; d < k; ++d) L *= job.dims[d];
      for (int d = k + 1; d <= z; ++d) R *= job.dims[d];
      small_message<S>(cur, V, job.raw + job.out_off[k], L, job.dims[k], R, red);
    }
  }
}

}  // namespace

template <typename S>
void launch_sweep_small(hipStream_t st, const SmallSiteJob *jobs_dev, int njobs, void *scratch,
                        int64_t slab_elems, int grid) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_sweep_small<S>, dim3((unsigned)grid), dim3(256), 0, st, jobs_dev, njobs,
                     scratch, slab_elems);
  HBP_HIP(hipGetLastError());
}
template void launch_sweep_small<double>(hipStream_t, const SmallSiteJob *, int, void *, int64_t,
                                         int);
template void launch_sweep_small<float>(hipStream_t, const SmallSiteJob *, int, void *, int64_t,
                                        int);

}  // namespace hbp
