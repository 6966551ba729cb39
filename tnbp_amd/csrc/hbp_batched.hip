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

// Out[a, j, b] = sum_i In[a, i, b] E[i, j] over bond `l` (In: L x Din x R,
// Out: L x Dout x R) with E staged in LDS as complex double, E[i + Din j];
// every thread of the block takes outputs tid, tid + 256, ...  All reads of
// In happen-before the barrier at the caller's next step.
template <typename S>
__device__ void small_mode_product(const typename CT<S>::type *__restrict__ in,
                                   typename CT<S>::type *__restrict__ out, int64_t L, int Din,
                                   int Dout, int64_t R, const double2 *__restrict__ Es) {
  using C = typename CT<S>::type;
  const int64_t n = L * Dout * R;
  for (int64_t o = threadIdx.x; o < n; o += blockDim.x) {
    const int64_t a = o % L, t = o / L;
    const int j = (int)(t % Dout);
    const int64_t b = t / Dout;
    const C *src = in + a + L * Din * b;
    double re = 0.0, im = 0.0;
    for (int i = 0; i < Din; ++i) {
      const C x = src[L * i];
      const double2 e = Es[i + Din * j];
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

// E_base / raw_base (both null, or both set): the messages every job reads
// and the raw messages it writes, the jobs' E and raw then unused (the
// small sites of one PEPS inside its fused sweep)
template <typename S>
__global__ __launch_bounds__(256) void k_sweep_small(const SmallSiteJob *__restrict__ jobs,
                                                     int njobs, void *__restrict__ scratch,
                                                     int64_t slab_elems,
                                                     const double2 *__restrict__ E_base,
                                                     double2 *__restrict__ raw_base) {
  using C = typename CT<S>::type;
  __shared__ double2 Es[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double2 red[256];
  C *bufA = (C *)scratch + (int64_t)blockIdx.x * 2 * slab_elems;
  C *bufB = bufA + slab_elems;
  // work item w = (site w / SMALL_SWEEP_ITEMS, outgoing message w % ...):
  // each message is computed on its own from V (no shared partial
  // products), so the split only spreads a site over more workgroups
  const int nitems = SMALL_SWEEP_ITEMS * njobs;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SmallSiteJob job = jobs[w / SMALL_SWEEP_ITEMS];
    const int z = job.z;
    const int k = w % SMALL_SWEEP_ITEMS;
    if (k >= z) continue;
    const C *V = (const C *)job.V;
    {
      const C *cur = V;
      C *nxt = bufA;
      for (int l = 0; l < z; ++l) {
        if (l == k) continue;
        int64_t L = 1, R = 1;
        for (int d = 0; d < l; ++d) L *= job.dims[d];
        for (int d = l + 1; d <= z; ++d) R *= job.dims[d];
        const int D = job.dims[l];
        const double2 *E = (E_base ? E_base : job.E) + job.in_off[l];
        for (int e = threadIdx.x; e < D * D; e += blockDim.x) Es[e] = E[e];
        __syncthreads();
        small_mode_product<S>(cur, nxt, L, D, D, R, Es);
        __syncthreads();
        cur = nxt;
        nxt = nxt == bufA ? bufB : bufA;
      }
      int64_t L = 1, R = 1;
      for (int d = 0; d < k; ++d) L *= job.dims[d];
      for (int d = k + 1; d <= z; ++d) R *= job.dims[d];
      small_message<S>(cur, V, (raw_base ? raw_base : job.raw) + job.out_off[k], L, job.dims[k], R,
                       red);
    }
  }
}

// Projectors into the sites of a truncation, bonds <= SMALL_D_MAX: one
// workgroup per site (striding over the sites), the z mode products in order
// through the workgroup's ping-pong buffers, then the tensor's squared norm
// summed in FP64 in a fixed order (every thread its strided elements, then a
// fixed tree) and the normalised tensor written to the site's new storage.
template <typename S>
__global__ __launch_bounds__(256) void k_project_small(const SmallProjJob *__restrict__ jobs,
                                                       int njobs, void *__restrict__ scratch,
                                                       int64_t slab_elems) {
  using C = typename CT<S>::type;
  __shared__ double2 Es[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double red[256];
  C *bufA = (C *)scratch + (int64_t)blockIdx.x * 2 * slab_elems;
  C *bufB = bufA + slab_elems;
  const int tid = threadIdx.x;
  for (int w = blockIdx.x; w < njobs; w += gridDim.x) {
    const SmallProjJob job = jobs[w];
    const int z = job.z;
    int dims[5];
    for (int d = 0; d <= z; ++d) dims[d] = job.dims[d];
    const C *cur = (const C *)job.V;
    C *nxt = bufA;
    for (int k = 0; k < z; ++k) {
      int64_t L = 1, R = 1;
      for (int d = 0; d < k; ++d) L *= dims[d];
      for (int d = k + 1; d <= z; ++d) R *= dims[d];
      const int Din = dims[k], Dout = job.keep[k];
      const double2 *P = job.P[k];
      for (int e = tid; e < Din * Dout; e += blockDim.x) {
        const int i = e % Din, j = e / Din;
        Es[e] = job.trans[k] ? P[j + Dout * i] : P[i + Din * j];
      }
      __syncthreads();
      small_mode_product<S>(cur, nxt, L, Din, Dout, R, Es);
      __syncthreads();
      dims[k] = Dout;
      cur = nxt;
      nxt = nxt == bufA ? bufB : bufA;
    }
    int64_t n = 1;
    for (int d = 0; d <= z; ++d) n *= dims[d];
    double s2 = 0.0;
    for (int64_t o = tid; o < n; o += blockDim.x) {
      const C x = cur[o];
      s2 += (double)x.x * x.x + (double)x.y * x.y;
    }
    red[tid] = s2;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
      if (tid < h) red[tid] += red[tid + h];
      __syncthreads();
    }
    const double v = red[0];
    const double sc = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
    C *out = (C *)job.out;
    for (int64_t o = tid; o < n; o += blockDim.x) {
      const C x = cur[o];
      C y;
      y.x = (S)((double)x.x * sc);
      y.y = (S)((double)x.y * sc);
      out[o] = y;
    }
    __syncthreads();  // red[0] and the buffers are reused by the next site
  }
}

template <typename S>
__global__ __launch_bounds__(256) void k_chain_small(const SmallChainSite *__restrict__ sites,
                                                     const SmallChainJob *__restrict__ jobs,
                                                     int njobs, const double2 *__restrict__ E,
                                                     void *__restrict__ scratch,
                                                     int64_t slab_elems, double2 *__restrict__ val,
                                                     double *__restrict__ lg) {
  using C = typename CT<S>::type;
  __shared__ double2 Es[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double2 X[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double2 M[SMALL_D_MAX * SMALL_D_MAX];
  __shared__ double2 red[256];
  C *bufA = (C *)scratch + (int64_t)blockIdx.x * 2 * slab_elems;
  C *bufB = bufA + slab_elems;
  const int tid = threadIdx.x;
  for (int w = blockIdx.x; w < njobs; w += gridDim.x) {
    const SmallChainJob job = jobs[w];
    double lsum = 0.0;
    for (int i = 0; i < job.n; ++i) {
      const SmallChainSite st = sites[job.first + i];
      const int z = st.z;
      const C *V = (const C *)st.V;
      const C *cur = V;
      C *nxt = bufA;
      auto product = [&](int b, int D, const double2 *src, bool lds) {
        int64_t L = 1, R = 1;
        for (int d = 0; d < b; ++d) L *= st.dims[d];
        for (int d = b + 1; d <= z; ++d) R *= st.dims[d];
        if (!lds)
          for (int e = tid; e < D * D; e += blockDim.x) Es[e] = src[e];
        __syncthreads();
        small_mode_product<S>(cur, nxt, L, D, D, R, lds ? src : Es);
        __syncthreads();
        cur = nxt;
        nxt = nxt == bufA ? bufB : bufA;
      };
      for (int b = 0; b < z; ++b)
        if (b != st.kin && b != st.kout) product(b, st.dims[b], E + st.in_off[b], false);
      if (st.kin >= 0 && st.kout >= 0) product(st.kin, st.dims[st.kin], X, true);
      if (st.op) product(z, st.dims[z], st.op, false);
      const int t = st.kout >= 0 ? st.kout : (st.kin >= 0 ? st.kin : z);
      int64_t L = 1, R = 1;
      for (int d = 0; d < t; ++d) L *= st.dims[d];
      for (int d = t + 1; d <= z; ++d) R *= st.dims[d];
      const int D = st.dims[t], n = D * D;
      small_message<S>(cur, V, M, L, D, R, red);  // ends with a barrier
      if (st.kout >= 0) {
        // X <- M / |M|_F, lg += log |M|_F
        double s2 = 0.0;
        for (int e = tid; e < n; e += blockDim.x) s2 += M[e].x * M[e].x + M[e].y * M[e].y;
        red[tid].x = s2;
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
          if (tid < h) red[tid].x += red[tid + h].x;
          __syncthreads();
        }
        const double nrm = sqrt(red[0].x);
        const double inv = nrm > 0.0 ? 1.0 / nrm : 1.0;
        for (int e = tid; e < n; e += blockDim.x) X[e] = make_double2(M[e].x * inv, M[e].y * inv);
        lsum += nrm > 0.0 ? log(nrm) : 0.0;
        __syncthreads();
      } else {
        double re = 0.0, im = 0.0;
        if (st.kin >= 0) {
          for (int e = tid; e < n; e += blockDim.x) {
            const double2 a = X[e], b = M[e];
            re += a.x * b.x - a.y * b.y;
            im += a.x * b.y + a.y * b.x;
          }
        } else {
          for (int q = tid; q < D; q += blockDim.x) {
            re += M[q + D * q].x;
            im += M[q + D * q].y;
          }
        }
        red[tid] = make_double2(re, im);
        __syncthreads();
        for (int h = 128; h > 0; h >>= 1) {
          if (tid < h) {
            red[tid].x += red[tid + h].x;
            red[tid].y += red[tid + h].y;
          }
          __syncthreads();
        }
        if (tid == 0) {
          val[job.slot] = red[0];
          lg[job.slot] = lsum;
        }
        __syncthreads();
      }
    }
  }
}

}  // namespace

template <typename S>
void launch_chain_small(hipStream_t st, const SmallChainSite *sites, const SmallChainJob *jobs,
                        int njobs, const double2 *E, void *scratch, int64_t slab_elems,
                        double2 *val, double *lg, int grid) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_chain_small<S>, dim3((unsigned)grid), dim3(256), 0, st, sites, jobs, njobs,
                     E, scratch, slab_elems, val, lg);
  HBP_HIP(hipGetLastError());
}
template void launch_chain_small<double>(hipStream_t, const SmallChainSite *, const SmallChainJob *,
                                         int, const double2 *, void *, int64_t, double2 *,
                                         double *, int);
template void launch_chain_small<float>(hipStream_t, const SmallChainSite *, const SmallChainJob *,
                                        int, const double2 *, void *, int64_t, double2 *,
                                        double *, int);

template <typename S>
void launch_project_small(hipStream_t st, const SmallProjJob *jobs_dev, int njobs, void *scratch,
                          int64_t slab_elems, int grid) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_project_small<S>, dim3((unsigned)grid), dim3(256), 0, st, jobs_dev, njobs,
                     scratch, slab_elems);
  HBP_HIP(hipGetLastError());
}
template void launch_project_small<double>(hipStream_t, const SmallProjJob *, int, void *, int64_t,
                                           int);
template void launch_project_small<float>(hipStream_t, const SmallProjJob *, int, void *, int64_t,
                                          int);

template <typename S>
void launch_sweep_small(hipStream_t st, const SmallSiteJob *jobs_dev, int njobs, void *scratch,
                        int64_t slab_elems, int grid, const double2 *E_base, double2 *raw_base) {
  if (njobs <= 0) return;
  hipLaunchKernelGGL(k_sweep_small<S>, dim3((unsigned)grid), dim3(256), 0, st, jobs_dev, njobs,
                     scratch, slab_elems, E_base, raw_base);
  HBP_HIP(hipGetLastError());
}
template void launch_sweep_small<double>(hipStream_t, const SmallSiteJob *, int, void *, int64_t,
                                         int, const double2 *, double2 *);
template void launch_sweep_small<float>(hipStream_t, const SmallSiteJob *, int, void *, int64_t,
                                        int, const double2 *, double2 *);

}  // namespace hbp
