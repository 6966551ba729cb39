// Hand-written CDNA4 (gfx950) kernels of the BP-PEPS backend.
//
// The two kernels that carry the flops are
//   * mode product   Out[l, j, r] = sum_i T[l, i, r] E[i, j]         (bp.h:148/150)
//   * message        M[i, j]      = sum_{l,r} T[l, i, r] conj V[l, j, r] (bp.h:165)
// on column-major site tensors viewed as (L, D, R) around one bond.  Both run
// complex arithmetic as four real MFMA products (v_mfma_f64_16x16x4f64 for
// FP64, v_mfma_f32_16x16x4f32 for the mixed path) with the (re, im) planes
// split in registers.  The message kernel stages row tiles through LDS so
// that global loads are coalesced for every bond position, and splits the huge
// contracted dimension into fixed chunks whose FP64 partials are summed in a
// fixed order: results are bitwise reproducible run to run and independent of
// what else runs on the device.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>

#include <algorithm>
#include <cstdio>

#include "hbp_internal.h"

namespace hbp {
namespace {

template <typename S> struct Cplx;
template <> struct Cplx<double> { using type = double2; };
template <> struct Cplx<float> { using type = float2; };

template <typename S> struct Mfma;
// Operand layout of both 16x16x4 MFMAs: lane (li = lane % 16, lk = lane / 16)
// holds A[li][lk] and B[lk][li].  The output differs: register v of lane
// holds D[row(lk, v)][li] with row = 4 v + lk for f64 (four groups of one row)
// and row = 4 lk + v for f32 (one group of four rows).
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int v) { return 4 * v + lk; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  __device__ static inline v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static inline int row(int lk, int v) { return 4 * lk + v; }
};

// Complex multiply-accumulate of 16x16 blocks on the real 16x16x4 MFMA, the
// small matrix E as the A operand and the streamed tensor T as B.
//   FP64: four real products, Re += Er Tr - Ei Ti, Im += Ei Tr + Er Ti.
//   FP32: Gauss's three real products, P1 += Er Tr, P2 += Ei Ti,
//         P3 += (Er + Ei)(Tr + Ti); Re = P1 - P2, Im = (P3 - P1) - P2: a
//         quarter fewer MFMA cycles for two adds per step, with an error of
//         the same order as the four-product form's (a few FP32 ulps of
//         |E| |T| per term, against the 1e-5 budget of the mixed path).
// Every mode-product form of one precision goes through this, so they stay
// bitwise equal to each other.
template <typename S> struct CAcc;
template <> struct CAcc<double> {
  using M = Mfma<double>;
  struct B { double r, i, ni; };
  __device__ static inline B prep(double r, double i) { return B{r, i, -i}; }
  M::v4 re, im;
  __device__ inline void zero() {
    re = M::v4{0, 0, 0, 0};
    im = M::v4{0, 0, 0, 0};
  }
  // (-Ei) Ti = Ei (-Ti) exactly: the negation once per B operand
  __device__ inline void mac(double ar, double ai, const B &b) {
    re = M::op(ar, b.r, re);
    re = M::op(ai, b.ni, re);
    im = M::op(ai, b.r, im);
    im = M::op(ar, b.i, im);
  }
  __device__ inline double out_re(int v) const { return re[v]; }
  __device__ inline double out_im(int v) const { return im[v]; }
};
template <> struct CAcc<float> {
  using M = Mfma<float>;
  struct B { float r, i, s; };
  __device__ static inline B prep(float r, float i) { return B{r, i, r + i}; }
  M::v4 p1, p2, p3;
  __device__ inline void zero() {
    p1 = M::v4{0, 0, 0, 0};
    p2 = M::v4{0, 0, 0, 0};
    p3 = M::v4{0, 0, 0, 0};
  }
  __device__ inline void mac(float ar, float ai, const B &b) {
    const float as = ar + ai;
    p1 = M::op(ar, b.r, p1);
    p2 = M::op(ai, b.i, p2);
    p3 = M::op(as, b.s, p3);
  }
  __device__ inline float out_re(int v) const { return p1[v] - p2[v]; }
  __device__ inline float out_im(int v) const { return (p3[v] - p1[v]) - p2[v]; }
};

// Workgroups per CU the mode-product kernel asks the compiler to fit: two for
// FP64 (its LDS tile takes 66 KB), three for FP32 (33 KB; at four its widest
// variant spills registers), so that FP32 has more waves per SIMD to hide HBM
// latency.
template <typename S> struct Occ { static constexpr int v = sizeof(S) == 4 ? 3 : 2; };


__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block-wide sum in a fixed order (lane tree, then waves in order).
__device__ inline double block_sum(double v, double *sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wave] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) t += sh[w];
  __syncthreads();
  return t;
}

// ------------------------------------------------------------ mode product
constexpr int MP_KC = 64;


// One workgroup (4 waves) stages E through LDS (converted to S) and then walks
// a strided list of row tiles; each wave owns RB blocks of 16 rows and all
// 16*NCB output columns of a tile.  T is read straight into MFMA A-operand
// registers (lane (li, lk) holds row li, k lk of the 16x4 tile) with the next
// k-step's loads issued before the current step's MFMAs; addresses are
// clamped in bounds and out-of-range rows/k are zeroed after the load.
// Tiles b0, b0 + nb, ... of one mode product; er/ei/red are the caller's LDS.
template <typename S, int NCB, int RB, int PFN>
__device__ __forceinline__ void mode_product_tiles(
    const typename Cplx<S>::type *__restrict__ T, const double2 *__restrict__ E,
    typename Cplx<S>::type *__restrict__ Out, int64_t L, int Din, int Dout, int64_t Q,
    int e_trans, S scale, double *__restrict__ norm_out, int64_t b0, int64_t nb,
    S (*er)[16 * NCB + 1], S (*ei)[16 * NCB + 1], double *red) {
  using C = typename Cplx<S>::type;
  using M = Mfma<S>;
  constexpr int NC = 16 * NCB;
  constexpr int64_t ROWS = 4 * 16 * RB;  // rows per tile

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t ntiles = (Q + ROWS - 1) / ROWS;
  const bool single = Din <= MP_KC;

  auto stage = [&](int kc) {
    for (int idx = threadIdx.x; idx < MP_KC * NC; idx += blockDim.x) {
      const int kk = idx / NC, j = idx - kk * NC;
      const int k = kc + kk;
      double2 e = make_double2(0.0, 0.0);
      if (k < Din && j < Dout) e = e_trans ? E[j + (int64_t)Dout * k] : E[k + (int64_t)Din * j];
      er[kk][j] = (S)e.x;
      ei[kk][j] = (S)e.y;
    }
  };
  if (single) {
    stage(0);
    __syncthreads();
  }

  double nrm = 0.0;
  for (int64_t tile = b0; tile < ntiles; tile += nb) {
    const int64_t row0 = tile * ROWS + (int64_t)wave * 16 * RB;
    const C *abase[RB];
    bool avalid[RB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int64_t q = row0 + rb * 16 + li;
      avalid[rb] = q < Q;
      const int64_t qq = avalid[rb] ? q : 0;
      const int64_t l = qq % L, r = qq / L;
      abase[rb] = T + l + L * (int64_t)Din * r;
    }
    CAcc<S> acc[RB][NCB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) acc[rb][cb].zero();

    for (int kc = 0; kc < Din; kc += MP_KC) {
      if (!single) {
        __syncthreads();
        stage(kc);
        __syncthreads();
      }
      const int kend = min(MP_KC, Din - kc);
      // A operand of k-step k4: rows of this lane, column kc + k4 + lk
      auto load = [&](int k4, S *re, S *im) {
        const int k = kc + k4 + lk;
        const int kcl = k < Din ? k : Din - 1;
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) {
          const C a = abase[rb][L * (int64_t)kcl];
          const bool ok = avalid[rb] && k < Din;
          re[rb] = ok ? (S)a.x : (S)0;
          im[rb] = ok ? (S)a.y : (S)0;
        }
      };
      auto step = [&](int k4, const S *are, const S *aim) {
        typename CAcc<S>::B bt[RB];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) bt[rb] = CAcc<S>::prep(are[rb], aim[rb]);
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const S bre = er[k4 + lk][cb * 16 + li];
          const S bim = ei[k4 + lk][cb * 16 + li];
          // E block as the A operand, the T rows as B: the product comes out
          // transposed (lane li holds output row li), so that each store
          // instruction writes 16 consecutive rows of 4 output columns
#pragma unroll
          for (int rb = 0; rb < RB; ++rb) acc[rb][cb].mac(bre, bim, bt[rb]);
        }
      };
      // a ring of PF k-steps in flight: the loads for k-step s + PF are issued
      // right after k-step s has used its registers (k >= Din is zeroed, and
      // the LDS rows beyond kend are zero there too)
      // PFN = 4: four k-steps in flight; PFN = 0: the next k-step only
      if constexpr (PFN > 0) {
        constexpr int PF = PFN;
        S re[PF][RB], im[PF][RB];
#pragma unroll
        for (int p = 0; p < PF; ++p) load(4 * p, re[p], im[p]);
        for (int k4 = 0; k4 < kend; k4 += 4 * PF) {
#pragma unroll
          for (int p = 0; p < PF; ++p) {
            step(k4 + 4 * p, re[p], im[p]);
            load(k4 + 4 * p + 4 * PF, re[p], im[p]);
          }
        }
      } else {
        S re0[RB], im0[RB], re1[RB], im1[RB];
        load(0, re0, im0);
        for (int k4 = 0; k4 < kend; k4 += 8) {
          load(k4 + 4, re1, im1);
          step(k4, re0, im0);
          load(k4 + 8, re0, im0);
          step(k4 + 4, re1, im1);
        }
      }
    }

    // register v of lane (li, lk) holds output row q = row0 + 16 rb + li,
    // column j = 16 cb + row(lk, v)
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int64_t q = row0 + rb * 16 + li;
      if (q >= Q) continue;
      const int64_t l = q % L, r = q / L;
      C *ob = Out + l + L * (int64_t)Dout * r;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const int j = cb * 16 + M::row(lk, v);
          if (j >= Dout) continue;
          C o;
          o.x = acc[rb][cb].out_re(v) * scale;
          o.y = acc[rb][cb].out_im(v) * scale;
          ob[L * (int64_t)j] = o;
          nrm += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
        }
      }
    }
  }
  if (norm_out) {
    const double t = block_sum(nrm, red);
    if (threadIdx.x == 0) *norm_out = t;
  }
}

// Whole-tile prefetch for Din <= 64: the loads of a wave's next tile (all of
// its k-steps) are in flight while the current tile's MFMAs run, so HBM
// latency is hidden by one tile of work (256 MFMAs per wave at 64 columns)
// rather than by one k-step.  The same k order and the same step() as
// mode_product_tiles: outputs are bitwise equal.  FULL: Din = Dout = 64 and
// every tile whole (Q a multiple of the tile rows), so that no load or store
// needs a bounds test: with them the compiler issues the tile's loads one at a
// time, each followed by a wait.
template <typename S, int NCB, int RB, bool FULL>
__device__ __forceinline__ void mode_product_tiles_pf(
    const typename Cplx<S>::type *__restrict__ T, const double2 *__restrict__ E,
    typename Cplx<S>::type *__restrict__ Out, int64_t L, int Din_rt, int Dout_rt, int64_t Q,
    int e_trans, S scale, double *__restrict__ norm_out, int64_t b0, int64_t nb,
    S (*er)[16 * NCB + 1], S (*ei)[16 * NCB + 1], double *red) {
  using C = typename Cplx<S>::type;
  using M = Mfma<S>;
  constexpr int NC = 16 * NCB;
  constexpr int64_t ROWS = 4 * 16 * RB;
  constexpr int KS = MP_KC / 4;  // k-steps of a tile
  const int Din = FULL ? MP_KC : Din_rt, Dout = FULL ? NC : Dout_rt;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t ntiles = (Q + ROWS - 1) / ROWS;

  for (int idx = threadIdx.x; idx < MP_KC * NC; idx += blockDim.x) {
    const int k = idx / NC, j = idx - k * NC;
    double2 e = make_double2(0.0, 0.0);
    if (FULL || (k < Din && j < Dout))
      e = e_trans ? E[j + (int64_t)Dout * k] : E[k + (int64_t)Din * j];
    er[k][j] = (S)e.x;
    ei[k][j] = (S)e.y;
  }
  __syncthreads();

  auto load_tile = [&](int64_t tile, S (*re)[RB], S (*im)[RB]) {
    const int64_t row0 = tile * ROWS + (int64_t)wave * 16 * RB;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int64_t q = row0 + rb * 16 + li;
      const bool valid = FULL || q < Q;
      const int64_t qq = valid ? q : 0;
      const C *base = T + (qq % L) + L * (int64_t)Din * (qq / L) + L * (int64_t)lk;
#pragma unroll
      for (int s = 0; s < KS; ++s) {
        if (FULL) {
          const C a = base[L * (int64_t)(4 * s)];
          re[s][rb] = (S)a.x;
          im[s][rb] = (S)a.y;
        } else {
          const int k = 4 * s + lk;
          const bool ok = valid && k < Din;
          const C a = base[L * (int64_t)(ok ? 4 * s : -lk)];
          re[s][rb] = ok ? (S)a.x : (S)0;
          im[s][rb] = ok ? (S)a.y : (S)0;
        }
      }
    }
  };
  double nrm = 0.0;
  auto compute_store = [&](int64_t tile, S (*re)[RB], S (*im)[RB]) {
    CAcc<S> acc[RB][NCB];
#pragma unroll
    for (int rb = 0; rb < RB; ++rb)
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) acc[rb][cb].zero();
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      if (!FULL && 4 * s >= Din) break;
      typename CAcc<S>::B bt[RB];
#pragma unroll
      for (int rb = 0; rb < RB; ++rb) bt[rb] = CAcc<S>::prep(re[s][rb], im[s][rb]);
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const S bre = er[4 * s + lk][cb * 16 + li];
        const S bim = ei[4 * s + lk][cb * 16 + li];
#pragma unroll
        for (int rb = 0; rb < RB; ++rb) acc[rb][cb].mac(bre, bim, bt[rb]);
      }
    }
    const int64_t row0 = tile * ROWS + (int64_t)wave * 16 * RB;
#pragma unroll
    for (int rb = 0; rb < RB; ++rb) {
      const int64_t q = row0 + rb * 16 + li;
      if (!FULL && q >= Q) continue;
      C *ob = Out + (q % L) + L * (int64_t)Dout * (q / L);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
#pragma unroll
        for (int cb = 0; cb < NCB; ++cb) {
          const int j = cb * 16 + M::row(lk, v);
          if (!FULL && j >= Dout) continue;
          C o;
          o.x = acc[rb][cb].out_re(v) * scale;
          o.y = acc[rb][cb].out_im(v) * scale;
          ob[L * (int64_t)j] = o;
          if (norm_out) nrm += (double)o.x * (double)o.x + (double)o.y * (double)o.y;
        }
      }
    }
  };

  S reA[KS][RB], imA[KS][RB], reB[KS][RB], imB[KS][RB];
  int64_t tile = b0;
  if (tile < ntiles) load_tile(tile, reA, imA);
  // the scheduler is kept from interleaving the two tiles' phases (it would
  // keep both tiles' accumulators and addresses live and spill)
  for (; tile < ntiles; tile += 2 * nb) {
    const int64_t t1 = tile + nb;
    if (t1 < ntiles) load_tile(t1, reB, imB);
    __builtin_amdgcn_sched_barrier(0);
    compute_store(tile, reA, imA);
    __builtin_amdgcn_sched_barrier(0);
    if (t1 >= ntiles) break;
    if (t1 + nb < ntiles) load_tile(t1 + nb, reA, imA);
    __builtin_amdgcn_sched_barrier(0);
    compute_store(t1, reB, imB);
    __builtin_amdgcn_sched_barrier(0);
  }
  if (norm_out) {
    const double t = block_sum(nrm, red);
    if (threadIdx.x == 0) *norm_out = t;
  }
}

// The whole-tile prefetch of mode_product_tiles_pf for 64 -> 64 columns on
// whole tiles, with buffer loads and stores: one 32-bit per-lane offset
// serves all 16 k-steps (and all stores), the step moves in the scalar
// offset, so that a tile's addresses take 1 register instead of 32 and two
// tiles of operands fit at three workgroups per CU.  Needs T and Out below
// 2 GiB each (the host checks).  Same k order and step() as
// mode_product_tiles: outputs are bitwise equal.
typedef unsigned int hbp_u2 __attribute__((ext_vector_type(2)));
typedef unsigned int hbp_u4 __attribute__((ext_vector_type(4)));
// cache-policy operand 2: the non-temporal hint.  T is streamed once and the
// output (a whole interior half-product) is read back only by the next
// kernel, after more data than the caches hold; measured on a D = 64 sweep:
// 2.7-3.6 % less time for the FP32 mode products, the merged kernel that
// reads the output unchanged.  (The same hint on the general variants' global
// stores cost FP64 1.7 % and the small messages that read their outputs.)
template <typename S>
__device__ __forceinline__ typename Cplx<S>::type buf_load(__amdgpu_buffer_rsrc_t r, unsigned vo,
                                                          unsigned so) {
  if constexpr (sizeof(S) == 4)
    return __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, 2));
  else
    return __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, vo, so, 0));
}
template <typename S>
__device__ __forceinline__ void buf_store(typename Cplx<S>::type x, __amdgpu_buffer_rsrc_t r,
                                          unsigned vo, unsigned so) {
  if constexpr (sizeof(S) == 4)
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(hbp_u2, x), r, vo, so, 2);
  else
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(hbp_u4, x), r, vo, so, 2);
}

template <typename S>
__device__ __forceinline__ void mode_product_tiles_buf(
    const typename Cplx<S>::type *__restrict__ T, const double2 *__restrict__ E,
    typename Cplx<S>::type *__restrict__ Out, int64_t L, int64_t Q, S scale, int64_t b0,
    int64_t nb, S (*er)[65], S (*ei)[65]) {
  using C = typename Cplx<S>::type;
  constexpr int D = 64, KS = 16;
  constexpr unsigned ES = sizeof(C);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int64_t ntiles = Q / 64;
  const unsigned bytes = (unsigned)(Q * D * ES);
  const __amdgpu_buffer_rsrc_t rT =
      __builtin_amdgcn_make_buffer_rsrc((void *)T, (short)0, (int)bytes, 0x00020000);
  const __amdgpu_buffer_rsrc_t rO =
      __builtin_amdgcn_make_buffer_rsrc((void *)Out, (short)0, (int)bytes, 0x00020000);
  const unsigned Lb = (unsigned)L * ES;  // bytes between consecutive k (or j)

  for (int idx = threadIdx.x; idx < D * D; idx += blockDim.x) {
    const int k = idx / D, j = idx - k * D;
    const double2 e = E[k + (int64_t)D * j];
    er[k][j] = (S)e.x;
    ei[k][j] = (S)e.y;
  }
  __syncthreads();

  // byte offset of this lane's row of tile `tile`
  auto row_off = [&](int64_t tile) -> unsigned {
    const int64_t q = tile * 64 + wave * 16 + li;
    return (unsigned)(((q % L) + L * (int64_t)D * (q / L)) * ES);
  };
  auto load_tile = [&](int64_t tile, S *re, S *im) {
    const unsigned vo = row_off(tile) + Lb * (unsigned)lk;
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const C a = buf_load<S>(rT, vo, Lb * (unsigned)(4 * s));
      re[s] = a.x;
      im[s] = a.y;
    }
  };
  auto compute_store = [&](int64_t tile, const S *re, const S *im) {
    CAcc<S> acc[4];
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) acc[cb].zero();
#pragma unroll
    for (int s = 0; s < KS; ++s) {
      const typename CAcc<S>::B bt = CAcc<S>::prep(re[s], im[s]);
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        const S bre = er[4 * s + lk][cb * 16 + li];
        const S bim = ei[4 * s + lk][cb * 16 + li];
        acc[cb].mac(bre, bim, bt);
      }
    }
    // register v of lane (li, lk): column j = 16 cb + row(lk, v); the lane
    // part of row() goes into the per-lane offset, the rest into the scalar
    // one (the host takes this form for scale = 1 only)
    const unsigned vo = row_off(tile) + Lb * (unsigned)(sizeof(S) == 4 ? 4 * lk : lk);
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int cb = 0; cb < 4; ++cb) {
        C o;
        o.x = acc[cb].out_re(v);
        o.y = acc[cb].out_im(v);
        buf_store<S>(o, rO, vo, Lb * (unsigned)(cb * 16 + (sizeof(S) == 4 ? v : 4 * v)));
      }
    }
  };

  S reA[KS], imA[KS], reB[KS], imB[KS];
  int64_t tile = b0;
  if (tile < ntiles) load_tile(tile, reA, imA);
  for (; tile < ntiles; tile += 2 * nb) {
    const int64_t t1 = tile + nb;
    if (t1 < ntiles) load_tile(t1, reB, imB);
    __builtin_amdgcn_sched_barrier(0);
    compute_store(tile, reA, imA);
    __builtin_amdgcn_sched_barrier(0);
    if (t1 >= ntiles) break;
    if (t1 + nb < ntiles) load_tile(t1 + nb, reA, imA);
    __builtin_amdgcn_sched_barrier(0);
    compute_store(t1, reB, imB);
    __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename S, int NCB, int RB, int PFN, int OCC>
__global__ __launch_bounds__(256, OCC) void k_mode_product(
    const typename Cplx<S>::type *__restrict__ T, const double2 *__restrict__ E,
    typename Cplx<S>::type *__restrict__ Out, int64_t L, int Din, int Dout, int64_t Q,
    int e_trans, S scale, double *__restrict__ norm_partials) {
  __shared__ S er[MP_KC][16 * NCB + 1];
  __shared__ S ei[MP_KC][16 * NCB + 1];
  __shared__ double red[8];
  if constexpr (PFN == 18) {  // 64 -> 64 on whole tiles with buffer operations
    static_assert(NCB == 4 && RB == 1, "buffer form: 64 columns, 16 rows per wave");
    mode_product_tiles_buf<S>(T, E, Out, L, Q, scale, blockIdx.x, gridDim.x, er, ei);
  } else if constexpr (PFN == 16 || PFN == 17)  // 17: the FULL form
    mode_product_tiles_pf<S, NCB, RB, PFN == 17>(T, E, Out, L, Din, Dout, Q, e_trans, scale,
                                                 norm_partials ? norm_partials + blockIdx.x
                                                               : nullptr,
                                                 blockIdx.x, gridDim.x, er, ei, red);
  else
    mode_product_tiles<S, NCB, RB, PFN>(T, E, Out, L, Din, Dout, Q, e_trans, scale,
                                        norm_partials ? norm_partials + blockIdx.x : nullptr,
                                        blockIdx.x, gridDim.x, er, ei, red);
}

// --------------------------------------------------------------- message
constexpr int MSG_QS = MSG_ROWS_PER_STAGE;  // rows per LDS stage
constexpr int MSG_IB = MSG_OUT_ROWS;        // output rows (i) per block (4 waves x 16)

// One row chunk [qbeg, qend) and output rows [i0, i0 + 64) of a message;
// writes the D1 x D2 partial at `out`.  The LDS tiles are the caller's.
template <typename S, int NCB>
__device__ __forceinline__ void message_chunk(
    const typename Cplx<S>::type *__restrict__ T, const typename Cplx<S>::type *__restrict__ V,
    double2 *__restrict__ out, int64_t L, int D1, int D2, int64_t qbeg, int64_t qend, int i0,
    int64_t ldT_R, int64_t ldV_R, int conjV, const double2 *__restrict__ wL,
    const double2 *__restrict__ wR, int Da, int64_t Sb, S (*tre)[MSG_IB + 1],
    S (*tim)[MSG_IB + 1], S (*vre)[16 * NCB + 1], S (*vim)[16 * NCB + 1]) {
  using C = typename Cplx<S>::type;
  using M = Mfma<S>;
  using v4 = typename M::v4;
  constexpr int NC = 16 * NCB;
  // column c of a row sits at L (c % Da) + Sb (c / Da) from the row's base:
  // Da = D1 gives the plain (L, D, R) view; Da = D, Sb = slab stacks the
  // physical slices of a tensor as further columns (edge environments)
  auto coff = [&](int c) -> int64_t { return L * (int64_t)(c % Da) + Sb * (int64_t)(c / Da); };
  constexpr int NT = MSG_QS * MSG_IB / 256;  // T elements per thread per stage
  constexpr int NV = MSG_QS * NC / 256;      // V elements per thread per stage

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const bool kcontig = (L == 1);
  const int tid = threadIdx.x;

  v4 acc_re[NCB], acc_im[NCB];
#pragma unroll
  for (int cb = 0; cb < NCB; ++cb) {
    acc_re[cb] = v4{0, 0, 0, 0};
    acc_im[cb] = v4{0, 0, 0, 0};
  }

  // Each thread owns NT elements of the T tile and NV of the V tile.  For
  // L == 1 a row holds its columns contiguously and threads run along columns;
  // otherwise threads run along rows (contiguous l).  Out-of-range elements
  // load from a clamped address and are zeroed.
  C tr[NT], vr[NV];
  auto weight = [&](int64_t l, int64_t r) -> C {
    double2 w = make_double2(1.0, 0.0);
    if (wL) w = wL[l];
    if (wR) {
      const double2 b = wR[r];
      w = make_double2(w.x * b.x - w.y * b.y, w.x * b.y + w.y * b.x);
    }
    C c;
    c.x = (S)w.x;
    c.y = (S)w.y;
    return c;
  };
  auto fetch = [&](int64_t q0) {
    if (kcontig) {
#pragma unroll
      for (int it = 0; it < NT; ++it) {
        const int idx = tid + 256 * it;
        const int qq = idx / MSG_IB, c = idx - qq * MSG_IB;
        const int64_t q = q0 + qq;
        const int i = i0 + c;
        const bool ok = q < qend && i < D1;
        C t = T[coff(ok ? i : 0) + ldT_R * (ok ? q : qbeg)];
        if (ok && (wL || wR)) {
          const C w = weight(0, q);
          t = C{t.x * w.x - t.y * w.y, t.x * w.y + t.y * w.x};
        }
        tr[it] = ok ? t : C{0, 0};
      }
#pragma unroll
      for (int it = 0; it < NV; ++it) {
        const int idx = tid + 256 * it;
        const int qq = idx / NC, c = idx - qq * NC;
        const int64_t q = q0 + qq;
        const bool ok = q < qend && c < D2;
        const C v = V[coff(ok ? c : 0) + ldV_R * (ok ? q : qbeg)];
        vr[it] = ok ? v : C{0, 0};
      }
    } else {
      const int qq = tid % MSG_QS;
      const int64_t q = q0 + qq;
      const bool okq = q < qend;
      const int64_t qc = okq ? q : qbeg;
      const int64_t l = qc % L, r = qc / L;
      const int64_t bT = l + ldT_R * r, bV = l + ldV_R * r;
      C w = C{1, 0};
      if (wL || wR) w = weight(l, r);
#pragma unroll
      for (int it = 0; it < NT; ++it) {
        const int i = i0 + tid / MSG_QS + (256 / MSG_QS) * it;
        const bool ok = okq && i < D1;
        C t = T[bT + coff(ok ? i : 0)];
        t = C{t.x * w.x - t.y * w.y, t.x * w.y + t.y * w.x};
        tr[it] = ok ? t : C{0, 0};
      }
#pragma unroll
      for (int it = 0; it < NV; ++it) {
        const int c = tid / MSG_QS + (256 / MSG_QS) * it;
        const bool ok = okq && c < D2;
        const C v = V[bV + coff(ok ? c : 0)];
        vr[it] = ok ? v : C{0, 0};
      }
    }
  };
  auto commit = [&]() {
    if (kcontig) {
#pragma unroll
      for (int it = 0; it < NT; ++it) {
        const int idx = tid + 256 * it;
        const int qq = idx / MSG_IB, c = idx - qq * MSG_IB;
        tre[qq][c] = tr[it].x;
        tim[qq][c] = tr[it].y;
      }
#pragma unroll
      for (int it = 0; it < NV; ++it) {
        const int idx = tid + 256 * it;
        const int qq = idx / NC, c = idx - qq * NC;
        vre[qq][c] = vr[it].x;
        vim[qq][c] = vr[it].y;
      }
    } else {
      const int qq = tid % MSG_QS;
#pragma unroll
      for (int it = 0; it < NT; ++it) {
        const int c = tid / MSG_QS + (256 / MSG_QS) * it;
        tre[qq][c] = tr[it].x;
        tim[qq][c] = tr[it].y;
      }
#pragma unroll
      for (int it = 0; it < NV; ++it) {
        const int c = tid / MSG_QS + (256 / MSG_QS) * it;
        vre[qq][c] = vr[it].x;
        vim[qq][c] = vr[it].y;
      }
    }
  };

  if (qbeg < qend) fetch(qbeg);
  for (int64_t q0 = qbeg; q0 < qend; q0 += MSG_QS) {
    __syncthreads();
    commit();
    __syncthreads();
    if (q0 + MSG_QS < qend) fetch(q0 + MSG_QS);  // in flight during the MFMAs below
#pragma unroll 2
    for (int k4 = 0; k4 < MSG_QS; k4 += 4) {
      const S ar = tre[k4 + lk][wave * 16 + li];
      const S ai = tim[k4 + lk][wave * 16 + li];
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const S br = vre[k4 + lk][cb * 16 + li];
        const S bi = conjV ? vim[k4 + lk][cb * 16 + li] : -vim[k4 + lk][cb * 16 + li];
        const S nbi = -bi;
        // T conj(V) = (Tr Vr + Ti Vi) + i (Ti Vr - Tr Vi); without conj Vi -> -Vi.
        // V as the A operand and T as B: the partial comes out transposed
        // (lane li holds output row i), so stores run along i
        acc_re[cb] = M::op(br, ar, acc_re[cb]);
        acc_re[cb] = M::op(bi, ai, acc_re[cb]);
        acc_im[cb] = M::op(br, ai, acc_im[cb]);
        acc_im[cb] = M::op(nbi, ar, acc_im[cb]);
      }
    }
  }

  // register v of lane (li, lk): row i = i0 + 16 wave + li, column j = 16 cb + row(lk, v)
  const int i = i0 + wave * 16 + li;
  if (i < D1) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int cb = 0; cb < NCB; ++cb) {
        const int j = cb * 16 + M::row(lk, v);
        if (j >= D2) continue;
        out[i + (int64_t)D1 * j] = make_double2((double)acc_re[cb][v], (double)acc_im[cb][v]);
      }
    }
  }
}

template <typename S, int NCB, int OCC>
__global__ __launch_bounds__(256, OCC) void k_message_partials(
    const typename Cplx<S>::type *__restrict__ T, const typename Cplx<S>::type *__restrict__ V,
    double2 *__restrict__ partials, int64_t L, int D1, int D2, int64_t Q, int64_t chunk,
    int64_t ldT_R, int64_t ldV_R, int conjV, const double2 *__restrict__ wL,
    const double2 *__restrict__ wR, int Da, int64_t Sb) {
  __shared__ S tre[MSG_QS][MSG_IB + 1], tim[MSG_QS][MSG_IB + 1];
  __shared__ S vre[MSG_QS][16 * NCB + 1], vim[MSG_QS][16 * NCB + 1];
  const int64_t qbeg = (int64_t)blockIdx.x * chunk;
  message_chunk<S, NCB>(T, V, partials + (int64_t)blockIdx.x * D1 * D2, L, D1, D2, qbeg,
                        min(Q, qbeg + chunk), blockIdx.y * MSG_IB, ldT_R, ldV_R, conjV, wL, wR,
                        Da, Sb, tre, tim, vre, vim);
}

// ------------------------------------------------ physical-index message
// rho[p, p'] = sum_l w[l] T[l + L p] conj V[l + L p'] for the physical index
// (d = 2): the reduced matrix of a site measurement (measure.h:54-84), whose
// contraction runs over all bonds.  The message kernel's 64 x 16 MFMA tiles
// would be almost empty for a 2 x 2 result and its staged chunks
// latency-bound (176 us for an interior site at D = 32); this streams the rows
// instead, FP64 products and sums, each thread a fixed row set and the block
// sums in a fixed order, so the partials are reproducible.
template <typename S>
__global__ __launch_bounds__(256) void k_rho_partials(const typename Cplx<S>::type *__restrict__ T,
                                                      const typename Cplx<S>::type *__restrict__ V,
                                                      const double2 *__restrict__ w, int64_t L,
                                                      double2 *__restrict__ partials) {
  __shared__ double sh[8];
  double acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};  // (re, im) of rho[p + 2 p']
  for (int64_t l = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; l < L;
       l += (int64_t)gridDim.x * blockDim.x) {
    double tr[2], ti[2], vr[2], vi[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const auto t = T[l + L * q];
      const auto v = V[l + L * q];
      tr[q] = (double)t.x;
      ti[q] = (double)t.y;
      vr[q] = (double)v.x;
      vi[q] = (double)v.y;
    }
    if (w) {
      const double2 c = w[l];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double a = c.x * tr[q] - c.y * ti[q], b = c.x * ti[q] + c.y * tr[q];
        tr[q] = a;
        ti[q] = b;
      }
    }
#pragma unroll
    for (int pp = 0; pp < 2; ++pp)
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        // (tr + i ti)(vr - i vi)
        acc[2 * (q + 2 * pp)] += tr[q] * vr[pp] + ti[q] * vi[pp];
        acc[2 * (q + 2 * pp) + 1] += ti[q] * vr[pp] - tr[q] * vi[pp];
      }
  }
  double2 *out = partials + (int64_t)blockIdx.x * 4;
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    const double re = block_sum(acc[2 * e], sh);
    const double im = block_sum(acc[2 * e + 1], sh);
    if (threadIdx.x == 0) out[e] = make_double2(re, im);
  }
}

// ------------------------------------------------ product + message, merged
// For two adjacent bonds k (index i) and j (index a) of column-major site
// tensors P and V of the same shape, with o = (l, r) the remaining index,
//   M[i, i'] = sum_{o, a'} T_o[i, a'] conj V_o[i', a'],
//   T_o[i, a'] = sum_a P_o[i, a] E(a, a'),       E(a, a') = E[a + Dj a'],
// i.e. the last mode product of an outgoing message (bp.h:148/150) and the
// message itself (bp.h:165) in one pass: the product tile T_o never leaves
// registers, which saves writing and re-reading the whole product tensor.
// Per o the workgroup runs two 64 x 64 x 64 complex products: wave w owns rows
// i in [16 w, 16 w + 16).  The first computes T_o^T = E^T P_o^T (E from LDS
// as the A operand, the wave's rows of P_o straight from global memory as B),
// so that register v of lane (li, lk) holds T_o[16 w + li][16 cb + row(lk, v)]
// - exactly the B-operand layout of the second product
// M^T += conj(V_o) T_o^T, contracted over a' in the order the first product
// left it in registers; V_o is staged through LDS for all four waves.  Bonds
// up to 64 (zero padded).  Blocks walk o = n0, n0 + nb, ... and each writes
// one partial; the fixed-order reduction makes the result reproducible.
constexpr int PM_D = 64;
template <typename S> struct OccPM { static constexpr int v = sizeof(S) == 4 ? 2 : 1; };

// contraction index of step s (0..15) for lane group lk: each lane's four
// consecutive steps read four consecutive a (one 32-byte run when a is the
// contiguous index)
__device__ __forceinline__ int pm_a(int s, int lk) { return 16 * (s >> 2) + 4 * lk + (s & 3); }

// LAY: 0 general strides, 1 a contiguous (L = 1, j before k), 2 i contiguous
// (L = 1, k before j).  FULL: Dk = Dj = 64, no bounds masks.  Offsets inside
// one o slice are 32-bit byte offsets from the slice base (the host checks
// that a slice spans less than 2^31 bytes).
template <typename S, int LAY, bool FULL>
__global__ __launch_bounds__(256, OccPM<S>::v) void k_product_message(
    const typename Cplx<S>::type *__restrict__ P, const typename Cplx<S>::type *__restrict__ V,
    const double2 *__restrict__ E, typename Cplx<S>::type *__restrict__ partials, int64_t L,
    int64_t nO, int Dk_rt, int Dj_rt, int64_t si, int64_t sa_rt, int64_t sr) {
  using C = typename Cplx<S>::type;
  using M = Mfma<S>;
  using v4 = typename M::v4;
  __shared__ C Es[PM_D][PM_D];
  __shared__ C Vs[PM_D][PM_D + 1];
  constexpr bool AC = LAY == 1;
  const int Dk = FULL ? PM_D : Dk_rt, Dj = FULL ? PM_D : Dj_rt;
  const int64_t sa = AC ? 1 : sa_rt;
  if (LAY == 2) si = 1;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nb = gridDim.x, b = blockIdx.x;
  // workgroups go round-robin to the 8 XCDs: the blocks of one XCD take
  // consecutive o, so that cache lines shared by neighbouring o (l runs
  // fastest) are read while they sit in that XCD's L2
  const int64_t n0 = (nb % 8 == 0) ? (int64_t)(b % 8) * (nb / 8) + b / 8 : b;

  for (int idx = tid; idx < PM_D * PM_D; idx += 256) {
    const int a = idx & (PM_D - 1), a2 = idx / PM_D;
    double2 e = make_double2(0.0, 0.0);
    if (FULL || (a < Dj && a2 < Dj)) e = E[a + (int64_t)Dj * a2];
    C c;
    c.x = (S)e.x;
    c.y = (S)e.y;
    Es[a][a2] = c;
  }

  const int i = wave * 16 + li;
  const bool iok = FULL || i < Dk;
  // V staging runs along the index with the smaller stride
  const bool afirst = AC || sa < si;
  const uint32_t si_b = (uint32_t)(si * sizeof(C)), sa_b = (uint32_t)(sa * sizeof(C));
  C pc[16], vr[16];
  auto slice = [&](const C *X, int64_t o) {
    const int64_t l = o % L, r = o / L;
    return (const char *)(X + (l + sr * r));
  };
  auto load_p = [&](int64_t o) {
    const char *base = slice(P, o);
    const uint32_t ib = (uint32_t)(iok ? i : 0) * si_b;
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int a = pm_a(s, lk);
      const bool ok = FULL || (iok && a < Dj);
      const C x = *(const C *)(base + (ib + (uint32_t)(ok ? a : 0) * sa_b));
      pc[s] = ok ? x : C{0, 0};
    }
  };
  auto vidx = [&](int n, int &ii, int &aa) {
    const int idx = tid + 256 * n;
    ii = afirst ? idx / PM_D : idx & (PM_D - 1);
    aa = afirst ? idx & (PM_D - 1) : idx / PM_D;
  };
  auto load_v = [&](int64_t o) {
    const char *base = slice(V, o);
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      int ii, aa;
      vidx(n, ii, aa);
      const bool ok = FULL || (ii < Dk && aa < Dj);
      const C x = *(const C *)(base + ((uint32_t)(ok ? ii : 0) * si_b +
                                       (uint32_t)(ok ? aa : 0) * sa_b));
      vr[n] = ok ? x : C{0, 0};
    }
  };
  auto commit_v = [&]() {
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      int ii, aa;
      vidx(n, ii, aa);
      Vs[aa][ii] = vr[n];
    }
  };

  // FP64: m_re / m_im are Re and Im of M^T, four real products per complex
  // one.  FP32: Gauss's three products (as in CAcc<float>), m_re = Q1 =
  // sum Vr Tr, m_im = Q2 = sum Vi Ti, m_s = Q3 = sum (Vr - Vi)(Tr + Ti), with
  // Re = Q1 + Q2 and Im = (Q3 - Q1) + Q2 for conj(V) T.
  constexpr bool G3 = sizeof(S) == 4;
  v4 m_re[4], m_im[4], m_s[G3 ? 4 : 1];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    m_re[c] = v4{0, 0, 0, 0};
    m_im[c] = v4{0, 0, 0, 0};
    if constexpr (G3) m_s[c] = v4{0, 0, 0, 0};
  }
  int64_t o = n0;
  if (o < nO) {
    load_p(o);
    load_v(o);
  }
  for (; o < nO; o += nb) {
    const bool more = o + nb < nO;
    __syncthreads();  // E staged; every wave done with the previous V_o
    commit_v();
    __syncthreads();
    if (more) load_v(o + nb);  // in flight until the next commit
    // FP64: t_re / t_im are Re and Im of T_o^T.  FP32: P1 = sum Er Pr (t_re),
    // P2 = sum Ei Pi (t_im), P3 = sum (Er + Ei)(Pr + Pi) (t_s); Re = P1 - P2,
    // Im = (P3 - P1) - P2.
    v4 t_re[4], t_im[4], t_s[G3 ? 4 : 1];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      t_re[c] = v4{0, 0, 0, 0};
      t_im[c] = v4{0, 0, 0, 0};
      if constexpr (G3) t_s[c] = v4{0, 0, 0, 0};
    }
    // the two products at raised wave priority, so that the MFMA phase of one
    // workgroup is not held back by the other's loads and LDS staging on the
    // same SIMD (measured 0.5-1.5 % less time for this kernel)
    __builtin_amdgcn_s_setprio(1);
    // T_o^T = E^T P_o^T:  T = P E = (Pr Er - Pi Ei) + i (Pr Ei + Pi Er)
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int a = pm_a(s, lk);
      if constexpr (G3) {
        const S pr = pc[s].x, pi = pc[s].y, ps = pc[s].x + pc[s].y;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          const C e = Es[a][cb * 16 + li];
          t_re[cb] = M::op(e.x, pr, t_re[cb]);
          t_im[cb] = M::op(e.y, pi, t_im[cb]);
          t_s[cb] = M::op(e.x + e.y, ps, t_s[cb]);
        }
      } else {
        const S pr = pc[s].x, pi = pc[s].y, npi = -pc[s].y;
#pragma unroll
        for (int cb = 0; cb < 4; ++cb) {
          const C e = Es[a][cb * 16 + li];
          t_re[cb] = M::op(e.x, pr, t_re[cb]);
          t_im[cb] = M::op(e.x, pi, t_im[cb]);
          t_re[cb] = M::op(e.y, npi, t_re[cb]);
          t_im[cb] = M::op(e.y, pr, t_im[cb]);
        }
      }
    }
    if (more) load_p(o + nb);  // registers free again; in flight during the second product
    // M^T += conj(V_o) T_o^T over a' = 16 cb + row(lk, v):
    // conj(V) T = (Vr Tr + Vi Ti) + i (Vr Ti - Vi Tr)
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int a2 = cb * 16 + M::row(lk, v);
        if constexpr (G3) {
          const S p1 = t_re[cb][v], p2 = t_im[cb][v];
          const S tr = p1 - p2, ti = (t_s[cb][v] - p1) - p2, ts = tr + ti;
#pragma unroll
          for (int c2 = 0; c2 < 4; ++c2) {
            const C x = Vs[a2][c2 * 16 + li];
            m_re[c2] = M::op(x.x, tr, m_re[c2]);
            m_im[c2] = M::op(x.y, ti, m_im[c2]);
            m_s[c2] = M::op(x.x - x.y, ts, m_s[c2]);
          }
        } else {
          const S tr = t_re[cb][v], ti = t_im[cb][v], ntr = -t_re[cb][v];
#pragma unroll
          for (int c2 = 0; c2 < 4; ++c2) {
            const C x = Vs[a2][c2 * 16 + li];
            m_re[c2] = M::op(x.x, tr, m_re[c2]);
            m_im[c2] = M::op(x.x, ti, m_im[c2]);
            m_re[c2] = M::op(x.y, ti, m_re[c2]);
            m_im[c2] = M::op(x.y, ntr, m_im[c2]);
          }
        }
      }
    }
    __builtin_amdgcn_s_setprio(0);
  }

  // register v of lane (li, lk) of m[c2]: M[i][i'] with i' = 16 c2 + row(lk, v)
  // partials in the accumulators' own precision (float2 for FP32)
  C *out = partials + (int64_t)b * Dk * Dk;
  if (iok) {
#pragma unroll
    for (int v = 0; v < 4; ++v) {
#pragma unroll
      for (int c2 = 0; c2 < 4; ++c2) {
        const int i2 = c2 * 16 + M::row(lk, v);
        if (!FULL && i2 >= Dk) continue;
        C x;
        if constexpr (G3) {
          const S q1 = m_re[c2][v], q2 = m_im[c2][v];
          x.x = q1 + q2;
          x.y = (m_s[c2][v] - q1) + q2;
        } else {
          x.x = m_re[c2][v];
          x.y = m_im[c2][v];
        }
        out[i + (int64_t)Dk * i2] = x;
      }
    }
  }
}

// Both messages of a bond pair (x, y) at the front of a site tensor whose
// environment is diagonal (the first BP sweep after AbsorbTPO or Truncation,
// bp.h:147-165 with E = diag(Lambda) (x) I):
//   M_x[i, i'] = sum_{o, y} w_o[o] ey[y] X_o[i, y] conj X_o[i', y]
//   M_y[j, j'] = sum_{o, x} w_o[o] ex[x] X_o[x, j] conj X_o[x, j']
// with X_o the Dx x Dy slice o (column-major, slices contiguous), ex / ey the
// diagonals of the messages entering on x / y and w_o the product of the other
// bonds' diagonals.  Each slice is read from HBM once for both messages: it is
// staged in LDS (double buffered, one barrier per slice) and both products
// take their MFMA operands from that tile, in either orientation.  Partials
// per workgroup as in k_product_message.
template <typename S, bool FULL>
__global__ __launch_bounds__(256, OccPM<S>::v) void k_pair_gram(
    const typename Cplx<S>::type *__restrict__ X, int64_t nO, int Dx_rt, int Dy_rt,
    const double2 *__restrict__ Ex, const double2 *__restrict__ Ey,
    const double2 *__restrict__ wo, typename Cplx<S>::type *__restrict__ part_x,
    typename Cplx<S>::type *__restrict__ part_y) {
  using C = typename Cplx<S>::type;
  using M = Mfma<S>;
  using v4 = typename M::v4;
  __shared__ C Xs[2][PM_D][PM_D + 1];  // [buffer][y][x]
  const int Dx = FULL ? PM_D : Dx_rt, Dy = FULL ? PM_D : Dy_rt;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nb = gridDim.x, b = blockIdx.x;
  const int64_t n0 = (nb % 8 == 0) ? (int64_t)(b % 8) * (nb / 8) + b / 8 : b;
  const int64_t slice = (int64_t)Dx * Dy;

  // diagonal weights of this lane's contraction indices
  C wy[16], wx[16];
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const int a = pm_a(s, lk);
    double2 ey = make_double2(0.0, 0.0), ex = make_double2(0.0, 0.0);
    if (FULL || a < Dy) ey = Ey[(int64_t)a * (Dy + 1)];
    if (FULL || a < Dx) ex = Ex[(int64_t)a * (Dx + 1)];
    wy[s].x = (S)ey.x;
    wy[s].y = (S)ey.y;
    wx[s].x = (S)ex.x;
    wx[s].y = (S)ex.y;
  }
  C vr[16];
  auto load = [&](int64_t o) {
    const C *base = X + o * slice;
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int idx = tid + 256 * n;
      const int x = idx & (PM_D - 1), y = idx / PM_D;
      const bool ok = FULL || (x < Dx && y < Dy);
      const C v = base[ok ? x + Dx * y : 0];
      vr[n] = ok ? v : C{0, 0};
    }
  };
  auto cmul = [](C a, C b) {
    C r;
    r.x = a.x * b.x - a.y * b.y;
    r.y = a.x * b.y + a.y * b.x;
    return r;
  };

  v4 mx_re[4], mx_im[4], my_re[4], my_im[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    mx_re[c] = mx_im[c] = my_re[c] = my_im[c] = v4{0, 0, 0, 0};
  }
  int buf = 0;
  int64_t o = n0;
  if (o < nO) load(o);
  for (; o < nO; o += nb, buf ^= 1) {
#pragma unroll
    for (int n = 0; n < 16; ++n) {
      const int idx = tid + 256 * n;
      Xs[buf][idx / PM_D][idx & (PM_D - 1)] = vr[n];
    }
    // one barrier: the other buffer was last read before this point
    __syncthreads();
    if (o + nb < nO) load(o + nb);
    C w = C{1, 0};
    if (wo) {
      const double2 t = wo[o];
      w.x = (S)t.x;
      w.y = (S)t.y;
    }
    // M_x^T += conj(X) T over y, T[i][y] = w ey[y] X[i][y]:
    // conj(A) T = (Ar Tr + Ai Ti) + i (Ar Ti - Ai Tr)
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int y = pm_a(s, lk);
      const C t = cmul(cmul(w, wy[s]), Xs[buf][y][wave * 16 + li]);
      const S tr = t.x, ti = t.y, ntr = -t.x;
#pragma unroll
      for (int c2 = 0; c2 < 4; ++c2) {
        const C a = Xs[buf][y][c2 * 16 + li];
        mx_re[c2] = M::op(a.x, tr, mx_re[c2]);
        mx_im[c2] = M::op(a.x, ti, mx_im[c2]);
        mx_re[c2] = M::op(a.y, ti, mx_re[c2]);
        mx_im[c2] = M::op(a.y, ntr, mx_im[c2]);
      }
    }
    // M_y^T += conj(X^T) T' over x, T'[j][x] = w ex[x] X[x][j]
#pragma unroll
    for (int s = 0; s < 16; ++s) {
      const int x = pm_a(s, lk);
      const C t = cmul(cmul(w, wx[s]), Xs[buf][wave * 16 + li][x]);
      const S tr = t.x, ti = t.y, ntr = -t.x;
#pragma unroll
      for (int c2 = 0; c2 < 4; ++c2) {
        const C a = Xs[buf][c2 * 16 + li][x];
        my_re[c2] = M::op(a.x, tr, my_re[c2]);
        my_im[c2] = M::op(a.x, ti, my_im[c2]);
        my_re[c2] = M::op(a.y, ti, my_re[c2]);
        my_im[c2] = M::op(a.y, ntr, my_im[c2]);
      }
    }
  }

  // register v of lane (li, lk): M[16 w + li][16 c2 + row(lk, v)]
  const int i = wave * 16 + li;
  C *ox = part_x + (int64_t)b * Dx * Dx, *oy = part_y + (int64_t)b * Dy * Dy;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
#pragma unroll
    for (int c2 = 0; c2 < 4; ++c2) {
      const int i2 = c2 * 16 + M::row(lk, v);
      if (FULL || (i < Dx && i2 < Dx)) {
        C t;
        t.x = mx_re[c2][v];
        t.y = mx_im[c2][v];
        ox[i + (int64_t)Dx * i2] = t;
      }
      if (FULL || (i < Dy && i2 < Dy)) {
        C t;
        t.x = my_re[c2][v];
        t.y = my_im[c2][v];
        oy[i + (int64_t)Dy * i2] = t;
      }
    }
  }
}

// Y[c + C (l + L r)] = X[l + L (c + C r)]: R column-major L x C matrices
// transposed through 64 x 64 LDS tiles (reads along l, writes along c), so
// that a pair of bonds behind a block of L leading indices moves to the front.
template <typename T>
__global__ __launch_bounds__(256) void k_swap_blocks(const T *__restrict__ X, T *__restrict__ Y,
                                                    int64_t L, int64_t Cn, int64_t R) {
  __shared__ T tile[64][65];
  const int64_t tl = (L + 63) / 64, tc = (Cn + 63) / 64, ntiles = tl * tc * R;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int64_t r = t / (tl * tc), rem = t - r * tl * tc;
    const int64_t c0 = (rem / tl) * 64, l0 = (rem % tl) * 64;
    const T *xr = X + L * Cn * r;
    T *yr = Y + L * Cn * r;
    __syncthreads();
    for (int cc = ty; cc < 64; cc += 4) {
      const int64_t l = l0 + tx, c = c0 + cc;
      if (l < L && c < Cn) tile[cc][tx] = xr[l + L * c];
    }
    __syncthreads();
    for (int ll = ty; ll < 64; ll += 4) {
      const int64_t c = c0 + tx, l = l0 + ll;
      if (l < L && c < Cn) yr[c + Cn * l] = tile[tx][ll];
    }
  }
}

// w[i] = prod_b E_b[c_b, c_b] over the column-major decomposition of i into
// (dims[0], ..., dims[nb-1], unit); the trailing unit factor covers indices
// that carry no message (the physical bond).
__device__ __forceinline__ double2 weight_at(const double2 *__restrict__ E, const WeightArgs &a,
                                             int64_t i) {
  int64_t rem = i;
  double re = 1.0, im = 0.0;
  for (int b = 0; b < a.nb; ++b) {
    const int64_t c = rem % a.dims[b];
    rem /= a.dims[b];
    const double2 e = E[a.diag_off[b] + c * (a.dims[b] + 1)];
    const double r2 = re * e.x - im * e.y, i2 = re * e.y + im * e.x;
    re = r2;
    im = i2;
  }
  return make_double2(re, im);
}
__global__ void k_build_weights(const double2 *__restrict__ E, double2 *__restrict__ w, int64_t n,
                                WeightArgs a) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = weight_at(E, a, i);
}
// several weight vectors in one launch (blockIdx.y = vector)
__global__ void k_build_weights_batch(const double2 *__restrict__ E, WeightBatch b) {
  const int v = blockIdx.y;
  const int64_t n = b.n[v];
  double2 *w = b.w[v];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = weight_at(E, b.a[v], i);
}

// M[e] = sum_c p[c n + e].  A block owns 16 entries; its G thread groups
// each sum the chunks g, g + G, ... in order and the G group sums are added
// in a fixed order, so the result is bitwise reproducible.
constexpr int RED_E = RED_ENTRIES;
// Partials of FP32 accumulators may be stored as float2: the conversion to
// double is exact, so the sums equal those of double2 partials bitwise.
template <typename PT, int G>
__device__ __forceinline__ void reduce_entries(const PT *__restrict__ p, int nchunks,
                                               int n, double2 *__restrict__ M, int64_t blk) {
  __shared__ double sre[G][RED_E + 1], sim[G][RED_E + 1];
  const int el = threadIdx.x % RED_E, g = threadIdx.x / RED_E;
  const int64_t e = blk * RED_E + el;
  double re = 0.0, im = 0.0;
  if (e < n) {
#pragma unroll 4
    for (int c = g; c < nchunks; c += G) {
      const PT v = p[(int64_t)c * n + e];
      re += (double)v.x;
      im += (double)v.y;
    }
  }
  sre[g][el] = re;
  sim[g][el] = im;
  __syncthreads();
  if (g == 0 && e < n) {
    double tr = 0.0, ti = 0.0;
    for (int k = 0; k < G; ++k) {
      tr += sre[k][el];
      ti += sim[k][el];
    }
    M[e] = make_double2(tr, ti);
  }
}

// 64 groups of 16 threads: each thread sums at most nchunks / 64 partials
// (8 of the merged kernels' 512), so that enough loads are in flight
constexpr int RED_G = 64;
template <typename PT>
__global__ __launch_bounds__(RED_G * RED_E) void k_reduce_partials(const PT *__restrict__ p,
                                                                    int nchunks, int n,
                                                                    double2 *__restrict__ M) {
  reduce_entries<PT, RED_G>(p, nchunks, n, M, blockIdx.x);
}

// Several reductions in one launch (blockIdx.y = job), the jobs passed by
// value: the same reduce_entries as k_reduce_partials, so the sums are
// bitwise those of separate launches.
__global__ __launch_bounds__(RED_G * RED_E) void k_reduce_batch(RedBatch b) {
  const RedBatchJob &j = b.j[blockIdx.y];
  if ((int64_t)blockIdx.x * RED_E >= j.n) return;
  if (j.f32)
    reduce_entries<float2, RED_G>((const float2 *)j.partials, j.nch, j.n, j.M, blockIdx.x);
  else
    reduce_entries<double2, RED_G>((const double2 *)j.partials, j.nch, j.n, j.M, blockIdx.x);
}

// ------------------------------------------------------------ elementwise
template <typename C>
__global__ void k_norm2_partials(const C *__restrict__ x, size_t n, double *__restrict__ part) {
  __shared__ double sh[8];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const C v = x[i];
    s += (double)v.x * (double)v.x + (double)v.y * (double)v.y;
  }
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}
template <typename R>
__global__ void k_norm2_partials_real(const R *__restrict__ x, size_t n, double *__restrict__ part) {
  __shared__ double sh[8];
  double s = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    s += (double)x[i] * (double)x[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_sum_partials(const double *__restrict__ part, int n, double *__restrict__ out) {
  __shared__ double sh[8];
  double s = 0.0;
  // fixed assignment of partials to threads -> deterministic
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  const double t = block_sum(s, sh);
  if (threadIdx.x == 0) out[0] = t;
}

template <typename C, typename R>
__global__ void k_scale(C *__restrict__ x, size_t n, double re, double im,
                        const double *__restrict__ dev_norm2) {
  double sr = re, si = im;
  if (dev_norm2) {
    const double v = dev_norm2[0];
    sr = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
    si = 0.0;
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const C v = x[i];
    C o;
    o.x = (R)((double)v.x * sr - (double)v.y * si);
    o.y = (R)((double)v.x * si + (double)v.y * sr);
    x[i] = o;
  }
}
template <typename R>
__global__ void k_scale_real(R *__restrict__ x, size_t n, double re, const double *dev_norm2) {
  double sr = re;
  if (dev_norm2) {
    const double v = dev_norm2[0];
    sr = v > 0.0 ? 1.0 / sqrt(v) : 1.0;
  }
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i] = (R)((double)x[i] * sr);
}

template <typename C>
__global__ void k_conj(C *__restrict__ x, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    x[i].y = -x[i].y;
}

__global__ void k_fill_bytes(unsigned char *x, size_t n, int esz, double re, double im,
                             int dtype) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    switch (dtype) {
      case HBP_C128: reinterpret_cast<double2 *>(x)[i] = make_double2(re, im); break;
      case HBP_F64: reinterpret_cast<double *>(x)[i] = re; break;
      case HBP_C64: reinterpret_cast<float2 *>(x)[i] = make_float2((float)re, (float)im); break;
      case HBP_F32: reinterpret_cast<float *>(x)[i] = (float)re; break;
    }
  }
}

__device__ inline double2 load_any(const void *x, int dtype, size_t i) {
  switch (dtype) {
    case HBP_C128: return reinterpret_cast<const double2 *>(x)[i];
    case HBP_F64: return make_double2(reinterpret_cast<const double *>(x)[i], 0.0);
    case HBP_C64: {
      const float2 v = reinterpret_cast<const float2 *>(x)[i];
      return make_double2(v.x, v.y);
    }
    default: return make_double2(reinterpret_cast<const float *>(x)[i], 0.0);
  }
}
__device__ inline void store_any(void *x, int dtype, size_t i, double2 v) {
  switch (dtype) {
    case HBP_C128: reinterpret_cast<double2 *>(x)[i] = v; break;
    case HBP_F64: reinterpret_cast<double *>(x)[i] = v.x; break;
    case HBP_C64: reinterpret_cast<float2 *>(x)[i] = make_float2((float)v.x, (float)v.y); break;
    default: reinterpret_cast<float *>(x)[i] = (float)v.x; break;
  }
}

__global__ void k_convert(const void *x, int di, void *y, int dout, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    store_any(y, dout, i, load_any(x, di, i));
}

struct PermArgs {
  int order;
  int64_t out_dims[HBP_MAX_ORDER];
  int64_t in_strides[HBP_MAX_ORDER];  // stride of the input axis feeding output axis d
};

template <typename E>
__global__ void k_permute(const E *__restrict__ in, E *__restrict__ out, size_t n, PermArgs a) {
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    size_t rem = o;
    int64_t off = 0;
    for (int d = 0; d < a.order; ++d) {
      const int64_t c = rem % a.out_dims[d];
      rem /= a.out_dims[d];
      off += c * a.in_strides[d];
    }
    out[o] = in[off];
  }
}

__global__ void k_axpy(const double2 *__restrict__ x, double2 *__restrict__ y, size_t n, double re,
                       double im) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    const double2 v = x[i];
    double2 w = y[i];
    w.x += re * v.x - im * v.y;
    w.y += re * v.y + im * v.x;
    y[i] = w;
  }
}

struct TraceArgs {
  int order_out;
  int64_t out_dims[HBP_MAX_ORDER];
  int64_t out_strides_in[HBP_MAX_ORDER];
  int npairs;
  int64_t tr_dims[HBP_MAX_ORDER];
  int64_t tr_strides[HBP_MAX_ORDER];  // stride_i + stride_j of each pair
};

__global__ void k_trace(const double2 *__restrict__ in, double2 *__restrict__ out, size_t n,
                        TraceArgs a) {
  for (size_t o = (size_t)blockIdx.x * blockDim.x + threadIdx.x; o < n;
       o += (size_t)gridDim.x * blockDim.x) {
    size_t rem = o;
    int64_t base = 0;
    for (int d = 0; d < a.order_out; ++d) {
      const int64_t c = rem % a.out_dims[d];
      rem /= a.out_dims[d];
      base += c * a.out_strides_in[d];
    }
    int64_t total = 1;
    for (int p = 0; p < a.npairs; ++p) total *= a.tr_dims[p];
    double re = 0.0, im = 0.0;
    for (int64_t t = 0; t < total; ++t) {
      int64_t r2 = t, off = base;
      for (int p = 0; p < a.npairs; ++p) {
        const int64_t c = r2 % a.tr_dims[p];
        r2 /= a.tr_dims[p];
        off += c * a.tr_strides[p];
      }
      re += in[off].x;
      im += in[off].y;
    }
    out[o] = make_double2(re, im);
  }
}

// ----------------------------------------------------------- Jacobi SVD
__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// One block of SVD_THREADS threads per matrix.  Hestenes one-sided Jacobi:
// pairs of columns (p, q) are rotated until mutually orthogonal; the
// rotations are accumulated in V so that A V = [sigma_j u_j].  Round-robin
// pairing puts n'/2 disjoint pairs in each round.  A pair is taken by a group
// of G lanes (G from the matrix's own row count, svd_group_lanes), so every pair
// of a round (32 at n = 64) is rotated at once and a round costs one
// dot-product reduction over G lanes and one rotation; a matrix's arithmetic
// does not depend on the batch it is in.  The group sums are xor
// butterflies, which leave bitwise the same sum in every lane of the group
// (floating-point addition is commutative), so the whole group takes the
// same skip decision and the same rotation.
// FORM: where the working columns W (m x n) and V (n x n) live.
// SVD_FORM_LDS: both in the block's LDS (m n + n^2 <= SVD_LDS_ELEMS, up to
// 64 x 64); SVD_FORM_W_LDS: W in LDS and V in the job's work area (m n up to
// the device's LDS per workgroup less the static LDS: 96 x 96 on 160 KiB),
// so the dot products and W's rotation never wait on global memory and V's
// loads can be issued early; SVD_FORM_GLOBAL: both in the work area.  The
// same operations in the same order in every form, so the results are
// bitwise those of the global form.
constexpr int SVD_THREADS = 512;

// Lanes per column pair: about 4 to 6 rows of the m-row columns per lane.
// Fewer lanes per pair (more rows each) leave one wave per SIMD with the
// latency of its reduction and rotation chain exposed; measured slower at
// 64 and 96 rows.  It depends on the matrix alone, not on its batch.
__device__ inline int svd_group_lanes(int m) {
  return m <= 48 ? 8 : m <= 96 ? 16 : m <= 192 ? 32 : 64;
}

// Sum over the aligned group of G lanes (G a power of two dividing 64).
__device__ inline double group_sum(double v, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Four group sums at once: every step issues the four exchanges before
// waiting, so the sums cost one chain of log2(G) exchanges instead of four;
// each value's additions are those of group_sum, in the same order.
__device__ inline void group_sum4(double &a, double &b, double &c, double &d, int G) {
  for (int o = G >> 1; o > 0; o >>= 1) {
    const double ta = __shfl_xor(a, o), tb = __shfl_xor(b, o);
    const double tc = __shfl_xor(c, o), td = __shfl_xor(d, o);
    a += ta;
    b += tb;
    c += tc;
    d += td;
  }
}

constexpr int SVD_FORM_GLOBAL = 0, SVD_FORM_LDS = 1, SVD_FORM_W_LDS = 2;

template <int FORM>
__global__ __launch_bounds__(SVD_THREADS) void k_batched_svd(const double2 *__restrict__ Ain,
                                                              double2 *__restrict__ U,
                                                              double *__restrict__ s,
                                                              double2 *__restrict__ Vd,
                                                              double2 *__restrict__ work,
                                                              const SvdJob *__restrict__ jobs) {
  extern __shared__ double2 svd_lds[];
  const SvdJob job = jobs[blockIdx.x];
  const int m = job.m, n = job.n;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
  const int G = svd_group_lanes(m);
  const int grp = threadIdx.x / G, gl = threadIdx.x % G, ngrp = blockDim.x / G;
  double2 *W = FORM != SVD_FORM_GLOBAL ? svd_lds : work + job.work_off;  // m x n
  double2 *Vm = FORM == SVD_FORM_LDS ? W + (int64_t)m * n                 // n x n
                                     : work + job.work_off + (int64_t)m * n;
  const double2 *A = Ain + job.a_off;
  double fro = 0.0;
  for (int64_t e = threadIdx.x; e < (int64_t)m * n; e += blockDim.x) {
    const double2 x = A[e];
    W[e] = x;
    fro += x.x * x.x + x.y * x.y;
  }
  for (int64_t e = threadIdx.x; e < (int64_t)n * n; e += blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    Vm[e] = make_double2(i == j ? 1.0 : 0.0, 0.0);
  }
  __shared__ int rotated;
  __shared__ double fro_part[SVD_THREADS / 64];
  fro = wave_sum(fro);
  if (lane == 0) fro_part[wave] = fro;
  __syncthreads();
  double fro_all = 0.0;
  for (int w = 0; w < nw; ++w) fro_all += fro_part[w];
  const int np = (n + 1) & ~1;  // players incl. a dummy when n is odd
  const double tol = 1.0e-15;
  // A column whose squared norm is below 1e-30 ||A||_F^2 is rounding noise:
  // rotating it against another column changes nothing at the scale of A,
  // and the relative orthogonality test would keep such pairs rotating to
  // the sweep limit (rank-deficient inputs such as Ra^T Rb of two roots of
  // rank-2 messages), so such pairs count as converged.
  const double small2 = 1.0e-30 * fro_all;
  for (int sweep = 0; sweep < 60; ++sweep) {
    if (threadIdx.x == 0) rotated = 0;
    __syncthreads();
    for (int round = 0; round < np - 1; ++round) {
      for (int pi = grp; pi < np / 2; pi += ngrp) {
        int a = (pi == 0) ? 0 : ((pi - 1 + round) % (np - 1)) + 1;
        int b = ((np - 2 - pi + round) % (np - 1)) + 1;
        if (a >= n || b >= n) continue;
        if (a > b) { const int t = a; a = b; b = t; }
        double al = 0.0, be = 0.0, gr = 0.0, gi = 0.0;
        for (int i = gl; i < m; i += G) {
          const double2 x = W[i + (int64_t)m * a], y = W[i + (int64_t)m * b];
          al += x.x * x.x + x.y * x.y;
          be += y.x * y.x + y.y * y.y;
          gr += x.x * y.x + x.y * y.y;  // conj(x) y
          gi += x.x * y.y - x.y * y.x;
        }
        group_sum4(al, be, gr, gi, G);
        const double g2 = gr * gr + gi * gi, g = sqrt(g2);
        if (al <= small2 || be <= small2 || !(g > tol * sqrt(al * be))) continue;
        // The rotation zeroing the pair's inner product: tan = the root of
        // t^2 + 2 zeta t - 1 = 0 of smaller modulus, zeta = (be - al) / 2g.
        // With d = be - al, r = sqrt(d^2 + 4 g^2) and q = |d| + r:
        // c = q / sqrt(q^2 + 4 g^2), s = sign(d) 2g / sqrt(q^2 + 4 g^2)
        // (two square roots and one division in a row); where the squares
        // would overflow, the zeta form.
        const double d = be - al, sgn = d >= 0.0 ? 1.0 : -1.0;
        const double q = fabs(d) + sqrt(d * d + 4.0 * g2);
        const double den = sqrt(q * q + 4.0 * g2);
        double c, sn;
        if (den < 1.0e300) {
          const double inv = 1.0 / den;
          c = q * inv;
          sn = sgn * 2.0 * g * inv;
        } else {
          const double zeta = d / (2.0 * g);
          const double t = fabs(zeta) < 1.0e150 ? sgn / (fabs(zeta) + sqrt(1.0 + zeta * zeta))
                                                : 0.5 / zeta;
          c = 1.0 / sqrt(1.0 + t * t);
          sn = c * t;
        }
        const double ig = 1.0 / g;
        const double2 w = make_double2(gr * ig, -gi * ig);  // e^{-i phi}
        for (int i = gl; i < m; i += G) {
          const double2 x = W[i + (int64_t)m * a], y = cmul(w, W[i + (int64_t)m * b]);
          W[i + (int64_t)m * a] = make_double2(c * x.x - sn * y.x, c * x.y - sn * y.y);
          W[i + (int64_t)m * b] = make_double2(sn * x.x + c * y.x, sn * x.y + c * y.y);
        }
        for (int i = gl; i < n; i += G) {
          const double2 x = Vm[i + (int64_t)n * a], y = cmul(w, Vm[i + (int64_t)n * b]);
          Vm[i + (int64_t)n * a] = make_double2(c * x.x - sn * y.x, c * x.y - sn * y.y);
          Vm[i + (int64_t)n * b] = make_double2(sn * x.x + c * y.x, sn * x.y + c * y.y);
        }
        if (gl == 0) rotated = 1;
      }
      __syncthreads();
    }
    const int any = rotated;
    __syncthreads();
    if (!any) break;
  }
  // column norms, descending order (ties keep the lower index first)
  __shared__ double sig[512];
  __shared__ int ord[512];
  for (int j = grp; j < n; j += ngrp) {
    double al = 0.0;
    for (int i = gl; i < m; i += G) {
      const double2 x = W[i + (int64_t)m * j];
      al += x.x * x.x + x.y * x.y;
    }
    al = group_sum(al, G);
    if (gl == 0) sig[j] = sqrt(al);
  }
  __syncthreads();
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    int rank = 0;
    for (int t = 0; t < n; ++t)
      if (sig[t] > sig[j] || (sig[t] == sig[j] && t < j)) ++rank;
    ord[rank] = j;
  }
  __syncthreads();
  const int k = min(m, n);
  for (int c = grp; c < k; c += ngrp) {
    const int j = ord[c];
    const double sv = sig[j];
    const double inv = sv > 0.0 ? 1.0 / sv : 0.0;
    for (int i = gl; i < m; i += G) {
      const double2 x = W[i + (int64_t)m * j];
      U[job.u_off + i + (int64_t)m * c] = make_double2(x.x * inv, x.y * inv);
    }
    for (int i = gl; i < n; i += G)
      Vd[job.vd_off + c + (int64_t)k * i] = cconj(Vm[i + (int64_t)n * j]);
    if (gl == 0) s[job.s_off + c] = sv;
  }
}

// Writes one job record into device memory from the kernel arguments: lets a
// caller that may not allocate or copy from the host (hbp_svd_noalloc) place
// the record in memory it already owns.
__global__ void k_store_svd_job(SvdJob job, SvdJob *__restrict__ dst) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *dst = job;
}

// --------------------------------------------------------- batched GEMM
__device__ inline double2 gemm_elem(const double2 *X, int ld, int op, int r, int c) {
  // element (r, c) of op(X)
  switch (op) {
    case 0: return X[r + (int64_t)ld * c];
    case 1: return X[c + (int64_t)ld * r];
    case 2: return cconj(X[c + (int64_t)ld * r]);
    default: return cconj(X[r + (int64_t)ld * c]);
  }
}

__global__ __launch_bounds__(256) void k_batched_gemm(const double2 *__restrict__ A,
                                                       const double2 *__restrict__ B,
                                                       double2 *__restrict__ Cm,
                                                       const GemmJob *__restrict__ jobs) {
  const GemmJob jb = jobs[blockIdx.y];
  const int64_t total = (int64_t)jb.m * jb.n;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e % jb.m), j = (int)(e / jb.m);
    double re = 0.0, im = 0.0;
    for (int kk = 0; kk < jb.k; ++kk) {
      const double2 a = gemm_elem(A + jb.a_off, jb.lda, jb.opa, i, kk);
      const double2 b = gemm_elem(B + jb.b_off, jb.ldb, jb.opb, kk, j);
      re += a.x * b.x - a.y * b.y;
      im += a.x * b.y + a.y * b.x;
    }
    Cm[jb.c_off + i + (int64_t)jb.ldc * j] =
        make_double2(jb.alpha.x * re - jb.alpha.y * im, jb.alpha.x * im + jb.alpha.y * re);
  }
}

inline int grid_for(size_t n, int block = 256, int cap = 4096) {
  size_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > (size_t)cap) g = cap;
  return (int)g;
}

}  // namespace

// =================================================================== host
// compute units of the current device (cached per device)
static int device_cus() {
  static std::atomic<int> cus[64];
  int dev = 0;
  HBP_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) dev = 0;
  int n = cus[dev].load(std::memory_order_relaxed);
  if (n == 0) {
    HBP_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    if (n <= 0) n = 256;
    cus[dev].store(n, std::memory_order_relaxed);
  }
  return n;
}

template <typename S, int NCB, int RB, int PFN, int OCC>
static void mp_launch(hipStream_t st, const void *T, const double2 *E, void *Out, int64_t L,
                      int Din, int Dout, int64_t Q, int e_trans, double scale, double *parts,
                      int *np) {
  const int64_t rows_per_tile = 4 * 16 * RB;
  const int64_t ntiles = (Q + rows_per_tile - 1) / rows_per_tile;
  // persistent tiles: E is staged once per workgroup and reused across tiles
  // one round of resident workgroups (as many per CU as fit), each walking its
  // tiles: against 2048 workgroups (2.7 or 4 rounds, the last one partly
  // empty) 5 % less time for the FP32 buffer form and 3 % for FP64 at D = 64
  static std::atomic<int> fit{0};
  int per_cu = fit.load(std::memory_order_relaxed);
  if (per_cu == 0) {
    HBP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void *>(&k_mode_product<S, NCB, RB, PFN, OCC>), 256, 0));
    if (per_cu <= 0) per_cu = 1;
    fit.store(per_cu, std::memory_order_relaxed);
  }
  const int64_t nb = std::min<int64_t>(ntiles, (int64_t)per_cu * device_cus());
  using C = typename Cplx<S>::type;
  timed_launch(k_mode_product<S, NCB, RB, PFN, OCC>, dim3((unsigned)nb), dim3(256), st,
               (const C *)T, E, (C *)Out, L, Din, Dout, Q, e_trans, (S)scale, parts);
  HBP_HIP(hipGetLastError());
  if (np) *np = (int)nb;
}

// ---------------------------------------------------------------- autotuner
// Launch variants of one kernel that compute bitwise the same result (row
// blocks per wave, loads in flight, occupancy target) are timed on the first
// launch of each shape class and the fastest is kept for the process.  Off
// with HBP_AUTOTUNE=0 (then variant 0, the default, always runs).
namespace {
struct TuneKey {
  int kind, elem, ncb, qlog, aux;
  bool operator<(const TuneKey &o) const {
    return std::tie(kind, elem, ncb, qlog, aux) < std::tie(o.kind, o.elem, o.ncb, o.qlog, o.aux);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, int> g_tune;
bool tune_enabled() {
  static const bool on = [] {
    const char *e = getenv("HBP_AUTOTUNE");
    return !(e && e[0] == '0');
  }();
  return on;
}
int ilog2(int64_t x) {
  int r = 0;
  while (x > 1) {
    x >>= 1;
    ++r;
  }
  return r;
}
// run(v) launches variant v on `st`; returns the variant to use for `key`,
// timing all nv variants (twice each, best of two) the first time.  The
// chosen variant is launched last, so the output is its output.
// HBP_AUTOTUNE_LOG=1 prints every timed trial to stderr.
bool tune_log() {
  static const bool on = [] {
    const char *e = getenv("HBP_AUTOTUNE_LOG");
    return e && e[0] == '1';
  }();
  return on;
}
// HBP_FORCE_VARIANT=v runs variant v wherever a shape class has it (else 0):
// for tests that compare the variants' outputs bitwise.
int forced_variant() {
  static const int v = [] {
    const char *e = getenv("HBP_FORCE_VARIANT");
    return e && *e ? atoi(e) : -1;
  }();
  return v;
}
template <typename F>
void tuned_launch(hipStream_t st, const TuneKey &key, int nv, F run) {
  if (forced_variant() >= 0) {
    run(forced_variant() < nv ? forced_variant() : 0);
    return;
  }
  if (nv <= 1 || !tune_enabled()) {
    run(0);
    return;
  }
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune.find(key);
    if (it != g_tune.end()) {
      run(it->second);
      return;
    }
  }
  HBP_REQUIRE(!graph_capturing(), HBP_ERR_UNSUPPORTED,
              "autotuner trial of a new launch shape while a graph is recorded");
  hipEvent_t a, b;
  HBP_HIP(hipEventCreate(&a));
  HBP_HIP(hipEventCreate(&b));
  // Every variant is timed three times, in rounds over all variants, and
  // keeps the least of its times: a slow moment of the device (clocks
  // ramping up, a dispatch delayed on the host) then costs one timing of
  // each variant rather than every timing of one.
  constexpr int kRounds = 3;
  float least[32];
  HBP_REQUIRE(nv <= 32, HBP_ERR_LOGIC, "autotuner: too many variants");
  for (int v = 0; v < nv; ++v) least[v] = 1e30f;
  launch_tuning() = true;  // the trial launches do not take a profiled bracket's events
  for (int round = 0; round < kRounds; ++round) {
    for (int v = 0; v < nv; ++v) {
      HBP_HIP(hipEventRecord(a, st));
      run(v);
      HBP_HIP(hipEventRecord(b, st));
      HBP_HIP(hipEventSynchronize(b));
      float ms = 0.0f;
      HBP_HIP(hipEventElapsedTime(&ms, a, b));
      if (tune_log())
        fprintf(stderr, "[hbp autotune] kind=%d elem=%d ncb=%d log2Q=%d aux=%d variant %d: %.4f ms\n",
                key.kind, key.elem, key.ncb, key.qlog, key.aux, v, ms);
      least[v] = std::min(least[v], ms);
    }
  }
  launch_tuning() = false;
  int best = 0;
  for (int v = 1; v < nv; ++v)
    if (least[v] < least[best]) best = v;
  HBP_HIP(hipEventDestroy(a));
  HBP_HIP(hipEventDestroy(b));
  {
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune[key] = best;
  }
  run(best);
}
}  // namespace

int small_sweep_grid(int variant, int njobs) {
  static const int mult[SMALL_GRID_VARIANTS] = {1, 2, 4};
  const int g = device_cus() * mult[variant < 0 || variant >= SMALL_GRID_VARIANTS ? 0 : variant];
  return std::max(1, std::min(g, njobs));
}

template <typename S>
int launch_sweep_small_tuned(hipStream_t st, const SmallSiteJob *jobs_dev, int njobs,
                             void *scratch, int64_t slab_elems, const double2 *E_base,
                             double2 *raw_base) {
  const TuneKey key{2, (int)sizeof(S), 0, ilog2(std::max(1, njobs)), ilog2(std::max<int64_t>(1, slab_elems))};
  int chosen = 0;
  tuned_launch(st, key, SMALL_GRID_VARIANTS, [&](int v) {
    chosen = v;
    launch_sweep_small<S>(st, jobs_dev, njobs, scratch, slab_elems,
                          small_sweep_grid(v, SMALL_SWEEP_ITEMS * njobs), E_base, raw_base);
  });
  return chosen;
}
template int launch_sweep_small_tuned<double>(hipStream_t, const SmallSiteJob *, int, void *,
                                              int64_t, const double2 *, double2 *);
template int launch_sweep_small_tuned<float>(hipStream_t, const SmallSiteJob *, int, void *,
                                             int64_t, const double2 *, double2 *);

std::string autotune_report() {
  std::lock_guard<std::mutex> g(g_tune_mu);
  std::string out;
  for (const auto &kv : g_tune) {
    const TuneKey &k = kv.first;
    out += std::string(k.kind == 0 ? "mode_product" : k.kind == 1 ? "message" : "small_sweep") + " elem=" +
           std::to_string(k.elem) + " ncb=" + std::to_string(k.ncb) + " log2Q=" +
           std::to_string(k.qlog) + " aux=" + std::to_string(k.aux) + " -> variant " +
           std::to_string(kv.second) + "\n";
  }
  return out;
}

// Mode-product variants for up to 64 output columns: (rows per wave, loads in
// flight, occupancy) = (32, 1 step, base), (32, 4 steps, base), (16, 1 step,
// base + 1), (16, 4 steps, base + 1); base is 2 (FP64) or 3 (FP32).  For
// FP32 with Din <= 64 also (16, next tile, base) (in FP64 its registers
// spill), and for FP32 64 -> 64 on whole tiles (16, next tile, 2) without
// bounds tests.
template <typename S, int NCB>
static void mp_variant(int v, hipStream_t st, const void *T, const double2 *E, void *Out,
                       int64_t L, int Din, int Dout, int64_t Q, int e_trans, double scale,
                       double *parts, int *np) {
  constexpr int B = Occ<S>::v;
  switch (v) {
    case 0: mp_launch<S, NCB, 2, 0, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np); break;
    case 1: mp_launch<S, NCB, 2, 4, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np); break;
    case 2: mp_launch<S, NCB, 1, 0, B + 1>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np); break;
    case 3: mp_launch<S, NCB, 1, 4, B + 1>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np); break;
    case 4:
      if constexpr (sizeof(S) == 4)
        mp_launch<S, NCB, 1, 16, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np);
      else
        HBP_REQUIRE(false, HBP_ERR_LOGIC, "mode product: no such variant");
      break;
    case 5:
      // FP32, whole tiles of 64 -> 64 columns (the host checks the shape), two
      // workgroups per CU (two tiles of operands per wave in registers)
      if constexpr (NCB == 4 && sizeof(S) == 4)
        mp_launch<S, NCB, 1, 17, 2>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np);
      else
        HBP_REQUIRE(false, HBP_ERR_LOGIC, "mode product: no such variant");
      break;
    case 6:
      // FP32, whole tiles of 64 -> 64 columns with buffer operations (the host
      // checks shape and size), three workgroups per CU (168 registers, four of
      // them spilled with the three-product accumulators)
      if constexpr (NCB == 4 && sizeof(S) == 4)
        mp_launch<S, NCB, 1, 18, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np);
      else
        HBP_REQUIRE(false, HBP_ERR_LOGIC, "mode product: no such variant");
      break;
    default:
      // the same at two workgroups per CU: no spills, fewer waves in flight
      if constexpr (NCB == 4 && sizeof(S) == 4)
        mp_launch<S, NCB, 1, 18, 2>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, parts, np);
      else
        HBP_REQUIRE(false, HBP_ERR_LOGIC, "mode product: no such variant");
      break;
  }
}

template <typename S>
void launch_mode_product(hipStream_t st, const void *T, const double2 *E, void *Out, int64_t L,
                         int Din, int Dout, int64_t R, int e_trans, double scale,
                         double *norm_partials, int *n_partials) {
  HBP_REQUIRE(L >= 1 && R >= 1 && Din >= 1 && Dout >= 1, HBP_ERR_INVALID_ARGUMENT,
              "mode product: empty dimension");
  HBP_REQUIRE(Dout <= 128, HBP_ERR_UNSUPPORTED, "mode product: output bond > 128");
  const int64_t Q = L * R;
  const int ncb = (Dout + 15) / 16;
  constexpr int B = Occ<S>::v;
  // the whole-tile prefetch variants need the full k range in one LDS stage;
  // variant 5 (no bounds tests) whole 64-row tiles of 64 -> 64 columns.  The
  // shape class records which of them exist (the tuned choice is an index).
  const bool whole = Din == MP_KC && Dout == 64 && Q % 64 == 0 && !e_trans &&
                     !norm_partials && scale == 1.0;
  const bool full = whole && sizeof(S) == 4;
  // (FP32 only: the FP64 form, with 128-bit buffer operations, did not
  // reproduce the other variants bitwise - its messages differed by ~1e-5 -
  // and is not offered)
  const bool buf = whole && sizeof(S) == 4 && Q * 64 * (int64_t)(2 * sizeof(S)) < ((int64_t)1 << 31);
  int vars[8] = {0, 1, 2, 3, 0, 0, 0, 0}, nv = 4;
  if (Din <= MP_KC && sizeof(S) == 4) vars[nv++] = 4;
  if (full) vars[nv++] = 5;
  if (buf) {
    vars[nv++] = 6;
    vars[nv++] = 7;
  }
  const TuneKey key{0, (int)sizeof(S), ncb, ilog2(Q),
                    (L == 1 ? 1 : 0) + 2 * (Din > MP_KC ? 1 : 0) + 4 * (full ? 1 : 0) +
                        8 * (Din <= MP_KC ? 1 : 0) + 16 * (buf ? 1 : 0)};
  auto go = [&](auto ncb_tag, int v) {
    constexpr int N = decltype(ncb_tag)::value;
    mp_variant<S, N>(vars[v], st, T, E, Out, L, Din, Dout, Q, e_trans, scale, norm_partials,
                     n_partials);
  };
  switch (ncb) {
    case 1: tuned_launch(st, key, nv, [&](int v) { go(std::integral_constant<int, 1>{}, v); }); break;
    case 2: tuned_launch(st, key, nv, [&](int v) { go(std::integral_constant<int, 2>{}, v); }); break;
    case 3: tuned_launch(st, key, nv, [&](int v) { go(std::integral_constant<int, 3>{}, v); }); break;
    case 4: tuned_launch(st, key, nv, [&](int v) { go(std::integral_constant<int, 4>{}, v); }); break;
    case 5: mp_launch<S, 5, 1, 0, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, norm_partials, n_partials); break;
    case 6: mp_launch<S, 6, 1, 0, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, norm_partials, n_partials); break;
    case 7: mp_launch<S, 7, 1, 0, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, norm_partials, n_partials); break;
    default: mp_launch<S, 8, 1, 0, B>(st, T, E, Out, L, Din, Dout, Q, e_trans, scale, norm_partials, n_partials); break;
  }
}
template void launch_mode_product<double>(hipStream_t, const void *, const double2 *, void *,
                                          int64_t, int, int, int64_t, int, double, double *,
                                          int *);
template void launch_mode_product<float>(hipStream_t, const void *, const double2 *, void *,
                                         int64_t, int, int, int64_t, int, double, double *,
                                         int *);

int mode_product_col_blocks(int Dout) { return (Dout + 15) / 16; }
int64_t message_num_chunks(int64_t Q) {
  // enough workgroups to fill 256 CUs several times over: up to 1024 chunks,
  // at least MSG_QS rows each
  int64_t nch = (Q + MSG_QS - 1) / MSG_QS;
  if (nch < 1) nch = 1;
  if (nch > 1024) nch = 1024;
  return nch;
}

template <typename S, int NCB>
static void msg_launch(hipStream_t st, const void *T, const void *V, double2 *partials,
                       int64_t L, int D1, int D2, int64_t Q, int64_t nch, int64_t chunk,
                       int64_t ldT_R, int64_t ldV_R, int conjV, const double2 *wL,
                       const double2 *wR, int Da, int64_t Sb) {
  using C = typename Cplx<S>::type;
  dim3 grid((unsigned)nch, (unsigned)((D1 + MSG_IB - 1) / MSG_IB));
  // occupancy variants of the same code (FP32 up to 64 columns): 2 or 3
  // workgroups per CU; the result does not depend on the choice
  auto run = [&](int v) {
    if (v == 0 || sizeof(S) != 4 || NCB > 4)
      timed_launch(k_message_partials<S, NCB, 2>, grid, dim3(256), st, (const C *)T,
                   (const C *)V, partials, L, D1, D2, Q, chunk, ldT_R, ldV_R, conjV, wL, wR, Da,
                   Sb);
    else
      timed_launch(k_message_partials<S, NCB, 3>, grid, dim3(256), st, (const C *)T,
                   (const C *)V, partials, L, D1, D2, Q, chunk, ldT_R, ldV_R, conjV, wL, wR, Da,
                   Sb);
    HBP_HIP(hipGetLastError());
  };
  const TuneKey key{1, (int)sizeof(S), NCB, ilog2(Q), (L == 1 ? 1 : 0) + 2 * (wL || wR ? 1 : 0)};
  tuned_launch(st, key, (sizeof(S) == 4 && NCB <= 4) ? 2 : 1, run);
}

template <typename S>
int message_partials(hipStream_t st, const void *T, const void *V, double2 *partials,
                     int64_t L, int D1, int D2, int64_t R, int64_t ldT_R, int64_t ldV_R,
                     int64_t max_chunks, int conjV, const double2 *wL, const double2 *wR,
                     int Da, int64_t Sb) {
  if (Da <= 0) {
    Da = std::max(D1, D2);
    Sb = 0;
  }
  HBP_REQUIRE(D2 <= 128, HBP_ERR_UNSUPPORTED, "message: bond > 128");
  const int64_t Q = L * R;
  int64_t nch = message_num_chunks(Q);
  if (nch > max_chunks) nch = max_chunks;
  int64_t chunk = (Q + nch - 1) / nch;
  chunk = ((chunk + MSG_QS - 1) / MSG_QS) * MSG_QS;
  nch = (Q + chunk - 1) / chunk;
  const int ncb = (D2 + 15) / 16;
  switch (ncb) {
    case 1: msg_launch<S, 1>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 2: msg_launch<S, 2>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 3: msg_launch<S, 3>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 4: msg_launch<S, 4>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 5: msg_launch<S, 5>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 6: msg_launch<S, 6>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    case 7: msg_launch<S, 7>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
    default: msg_launch<S, 8>(st, T, V, partials, L, D1, D2, Q, nch, chunk, ldT_R, ldV_R, conjV, wL, wR, Da, Sb); break;
  }
  return (int)nch;
}
template int message_partials<double>(hipStream_t, const void *, const void *, double2 *,
                                      int64_t, int, int, int64_t, int64_t, int64_t, int64_t,
                                      int, const double2 *, const double2 *, int, int64_t);
template int message_partials<float>(hipStream_t, const void *, const void *, double2 *,
                                     int64_t, int, int, int64_t, int64_t, int64_t, int64_t,
                                     int, const double2 *, const double2 *, int, int64_t);

template <typename S>
int rho_partials(hipStream_t st, const void *T, const void *V, const double2 *w, int64_t L,
                 double2 *partials, int max_blocks) {
  using C = typename Cplx<S>::type;
  const int64_t nb = std::max<int64_t>(1, std::min<int64_t>(max_blocks, (L + 255) / 256));
  hipLaunchKernelGGL(k_rho_partials<S>, dim3((unsigned)nb), dim3(256), 0, st, (const C *)T,
                     (const C *)V, w, L, partials);
  HBP_HIP(hipGetLastError());
  return (int)nb;
}
template int rho_partials<float>(hipStream_t, const void *, const void *, const double2 *, int64_t,
                                 double2 *, int);
template int rho_partials<double>(hipStream_t, const void *, const void *, const double2 *,
                                  int64_t, double2 *, int);

template <typename S>
int product_message_partials(hipStream_t st, const void *P, const void *V, const double2 *E,
                             void *partials, int64_t L, int64_t R, int Dk, int Dj,
                             int k_first, int max_blocks) {
  HBP_REQUIRE(Dk >= 1 && Dk <= PM_D && Dj >= 1 && Dj <= PM_D, HBP_ERR_UNSUPPORTED,
              "merged product-message: bonds must be <= 64");
  const int64_t nO = L * R;
  HBP_REQUIRE(nO >= 1 && max_blocks >= 1, HBP_ERR_INVALID_ARGUMENT,
              "merged product-message: empty problem");
  // (l, i, a, r) for k before j, (l, a, i, r) otherwise
  const int64_t si = k_first ? L : L * Dj;
  const int64_t sa = k_first ? L * Dk : L;
  const int64_t sr = L * Dk * Dj;
  // one pass of workgroups (two per CU on 256 CUs), a multiple of the 8 XCDs
  int64_t nb = std::min<int64_t>(std::min<int64_t>(max_blocks, 512), nO);
  if (nb >= 8) nb -= nb % 8;
  using C = typename Cplx<S>::type;
  HBP_REQUIRE(sr * (int64_t)sizeof(C) < ((int64_t)1 << 31), HBP_ERR_UNSUPPORTED,
              "merged product-message: slice beyond 2 GiB");
  const int lay = sa == 1 ? 1 : si == 1 ? 2 : 0;
  const bool full = Dk == PM_D && Dj == PM_D;
  auto go = [&](auto kern) {
    timed_launch(kern, dim3((unsigned)nb), dim3(256), st, (const C *)P, (const C *)V, E,
                 (C *)partials, L, nO, Dk, Dj, si, sa, sr);
  };
  if (full) {
    if (lay == 1) go(k_product_message<S, 1, true>);
    else if (lay == 2) go(k_product_message<S, 2, true>);
    else go(k_product_message<S, 0, true>);
  } else {
    if (lay == 1) go(k_product_message<S, 1, false>);
    else if (lay == 2) go(k_product_message<S, 2, false>);
    else go(k_product_message<S, 0, false>);
  }
  HBP_HIP(hipGetLastError());
  return (int)nb;
}
template <typename S>
int pair_gram_partials(hipStream_t st, const void *X, int64_t nO, int Dx, int Dy,
                       const double2 *Ex, const double2 *Ey, const double2 *wo, void *part_x,
                       void *part_y, int max_blocks) {
  HBP_REQUIRE(Dx >= 1 && Dx <= PM_D && Dy >= 1 && Dy <= PM_D, HBP_ERR_UNSUPPORTED,
              "pair gram: bonds must be <= 64");
  HBP_REQUIRE(nO >= 1 && max_blocks >= 1, HBP_ERR_INVALID_ARGUMENT, "pair gram: empty problem");
  int64_t nb = std::min<int64_t>(std::min<int64_t>(max_blocks, 512), nO);
  if (nb >= 8) nb -= nb % 8;
  using C = typename Cplx<S>::type;
  auto go = [&](auto kern) {
    timed_launch(kern, dim3((unsigned)nb), dim3(256), st, (const C *)X, nO, Dx, Dy, Ex, Ey, wo,
                 (C *)part_x, (C *)part_y);
  };
  if (Dx == PM_D && Dy == PM_D) go(k_pair_gram<S, true>);
  else go(k_pair_gram<S, false>);
  HBP_HIP(hipGetLastError());
  return (int)nb;
}
template int pair_gram_partials<double>(hipStream_t, const void *, int64_t, int, int,
                                        const double2 *, const double2 *, const double2 *, void *,
                                        void *, int);
template int pair_gram_partials<float>(hipStream_t, const void *, int64_t, int, int,
                                       const double2 *, const double2 *, const double2 *, void *,
                                       void *, int);

void launch_swap_blocks(hipStream_t st, const void *X, void *Y, int64_t L, int64_t Cn,
                        int64_t R, int elem_bytes) {
  const int64_t ntiles = ((L + 63) / 64) * ((Cn + 63) / 64) * R;
  const unsigned grid = (unsigned)std::min<int64_t>(ntiles, 8192);
  if (elem_bytes == 16)
    hipLaunchKernelGGL(k_swap_blocks<double2>, dim3(grid), dim3(256), 0, st, (const double2 *)X,
                       (double2 *)Y, L, Cn, R);
  else
    hipLaunchKernelGGL(k_swap_blocks<float2>, dim3(grid), dim3(256), 0, st, (const float2 *)X,
                       (float2 *)Y, L, Cn, R);
  HBP_HIP(hipGetLastError());
}

template int product_message_partials<double>(hipStream_t, const void *, const void *,
                                              const double2 *, void *, int64_t, int64_t, int, int,
                                              int, int);
template int product_message_partials<float>(hipStream_t, const void *, const void *,
                                             const double2 *, void *, int64_t, int64_t, int, int,
                                             int, int);

void launch_build_weights(hipStream_t st, const double2 *E, double2 *w, int64_t n, int nb,
                          const int64_t *dims, const int64_t *diag_off) {
  WeightArgs a{};
  a.nb = nb;
  for (int b = 0; b < nb; ++b) {
    a.dims[b] = dims[b];
    a.diag_off[b] = diag_off[b];
  }
  hipLaunchKernelGGL(k_build_weights, dim3(grid_for((size_t)n)), dim3(256), 0, st, E, w, n, a);
  HBP_HIP(hipGetLastError());
}

void launch_build_weights_batch(hipStream_t st, const double2 *E, const WeightBatch &b) {
  if (b.count <= 0) return;
  int64_t nmax = 1;
  for (int i = 0; i < b.count; ++i) nmax = std::max(nmax, b.n[i]);
  hipLaunchKernelGGL(k_build_weights_batch, dim3(grid_for((size_t)nmax), b.count), dim3(256), 0,
                     st, E, b);
  HBP_HIP(hipGetLastError());
}

void launch_reduce_batch(hipStream_t st, const RedBatch &b) {
  if (b.count <= 0) return;
  int gx = 1;
  for (int i = 0; i < b.count; ++i) gx = std::max(gx, (b.j[i].n + RED_E - 1) / RED_E);
  hipLaunchKernelGGL(k_reduce_batch, dim3(gx, b.count), dim3(RED_G * RED_E), 0, st, b);
  HBP_HIP(hipGetLastError());
}

void launch_reduce_partials(hipStream_t st, const double2 *partials, int nchunks, int n,
                            double2 *M) {
  hipLaunchKernelGGL(k_reduce_partials<double2>, dim3((n + RED_E - 1) / RED_E),
                     dim3(RED_G * RED_E), 0, st,
                     partials, nchunks, n, M);
  HBP_HIP(hipGetLastError());
}

void launch_reduce_partials(hipStream_t st, const float2 *partials, int nchunks, int n,
                            double2 *M) {
  hipLaunchKernelGGL(k_reduce_partials<float2>, dim3((n + RED_E - 1) / RED_E),
                     dim3(RED_G * RED_E), 0, st,
                     partials, nchunks, n, M);
  HBP_HIP(hipGetLastError());
}

void launch_norm2_partials(hipStream_t st, const void *x, size_t n, int dtype, double *partials,
                           int *nblocks) {
  const int g = grid_for(n, 256, 1024);
  switch (dtype) {
    case HBP_C128: hipLaunchKernelGGL(k_norm2_partials<double2>, dim3(g), dim3(256), 0, st, (const double2 *)x, n, partials); break;
    case HBP_C64: hipLaunchKernelGGL(k_norm2_partials<float2>, dim3(g), dim3(256), 0, st, (const float2 *)x, n, partials); break;
    case HBP_F64: hipLaunchKernelGGL(k_norm2_partials_real<double>, dim3(g), dim3(256), 0, st, (const double *)x, n, partials); break;
    default: hipLaunchKernelGGL(k_norm2_partials_real<float>, dim3(g), dim3(256), 0, st, (const float *)x, n, partials); break;
  }
  HBP_HIP(hipGetLastError());
  *nblocks = g;
}

void launch_sum_partials(hipStream_t st, const double *partials, int n, double *out) {
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, st, partials, n, out);
  HBP_HIP(hipGetLastError());
}

void launch_scale(hipStream_t st, void *x, size_t n, int dtype, double re, double im,
                  const double *dev_norm2) {
  const int g = grid_for(n);
  switch (dtype) {
    case HBP_C128: hipLaunchKernelGGL((k_scale<double2, double>), dim3(g), dim3(256), 0, st, (double2 *)x, n, re, im, dev_norm2); break;
    case HBP_C64: hipLaunchKernelGGL((k_scale<float2, float>), dim3(g), dim3(256), 0, st, (float2 *)x, n, re, im, dev_norm2); break;
    case HBP_F64: hipLaunchKernelGGL(k_scale_real<double>, dim3(g), dim3(256), 0, st, (double *)x, n, re, dev_norm2); break;
    default: hipLaunchKernelGGL(k_scale_real<float>, dim3(g), dim3(256), 0, st, (float *)x, n, re, dev_norm2); break;
  }
  HBP_HIP(hipGetLastError());
}

void launch_conj(hipStream_t st, void *x, size_t n, int dtype) {
  const int g = grid_for(n);
  if (dtype == HBP_C128)
    hipLaunchKernelGGL(k_conj<double2>, dim3(g), dim3(256), 0, st, (double2 *)x, n);
  else if (dtype == HBP_C64)
    hipLaunchKernelGGL(k_conj<float2>, dim3(g), dim3(256), 0, st, (float2 *)x, n);
  else
    return;
  HBP_HIP(hipGetLastError());
}

void launch_fill(hipStream_t st, void *x, size_t n, int dtype, double re, double im) {
  hipLaunchKernelGGL(k_fill_bytes, dim3(grid_for(n)), dim3(256), 0, st, (unsigned char *)x, n,
                     (int)dtype_size(dtype), re, im, dtype);
  HBP_HIP(hipGetLastError());
}

void launch_convert(hipStream_t st, const void *x, int dtype_in, void *y, int dtype_out,
                    size_t n) {
  hipLaunchKernelGGL(k_convert, dim3(grid_for(n)), dim3(256), 0, st, x, dtype_in, y, dtype_out, n);
  HBP_HIP(hipGetLastError());
}

void launch_permute(hipStream_t st, const void *in, void *out, int elem_bytes, int order,
                    const int32_t *in_shape, const int32_t *perm) {
  PermArgs a{};
  a.order = order;
  int64_t strides[HBP_MAX_ORDER];
  int64_t acc = 1;
  size_t n = 1;
  for (int d = 0; d < order; ++d) {
    strides[d] = acc;
    acc *= in_shape[d];
  }
  for (int d = 0; d < order; ++d) {
    a.out_dims[d] = in_shape[perm[d]];
    a.in_strides[d] = strides[perm[d]];
    n *= (size_t)in_shape[perm[d]];
  }
  if (n == 0) return;
  const int g = grid_for(n);
  switch (elem_bytes) {
    case 16: hipLaunchKernelGGL(k_permute<double2>, dim3(g), dim3(256), 0, st, (const double2 *)in, (double2 *)out, n, a); break;
    case 8: hipLaunchKernelGGL(k_permute<double>, dim3(g), dim3(256), 0, st, (const double *)in, (double *)out, n, a); break;
    default: hipLaunchKernelGGL(k_permute<float>, dim3(g), dim3(256), 0, st, (const float *)in, (float *)out, n, a); break;
  }
  HBP_HIP(hipGetLastError());
}

void launch_axpy_c128(hipStream_t st, const double2 *x, double2 *y, size_t n, double re,
                      double im) {
  hipLaunchKernelGGL(k_axpy, dim3(grid_for(n)), dim3(256), 0, st, x, y, n, re, im);
  HBP_HIP(hipGetLastError());
}

void launch_trace(hipStream_t st, const double2 *in, double2 *out, int order,
                  const int32_t *shape, int npairs, const int32_t *pairs) {
  TraceArgs a{};
  int64_t strides[HBP_MAX_ORDER];
  int64_t acc = 1;
  for (int d = 0; d < order; ++d) {
    strides[d] = acc;
    acc *= shape[d];
  }
  bool traced[HBP_MAX_ORDER] = {false};
  a.npairs = npairs;
  for (int p = 0; p < npairs; ++p) {
    const int i = pairs[2 * p], j = pairs[2 * p + 1];
    traced[i] = traced[j] = true;
    a.tr_dims[p] = shape[i];
    a.tr_strides[p] = strides[i] + strides[j];
  }
  size_t n = 1;
  a.order_out = 0;
  for (int d = 0; d < order; ++d) {
    if (traced[d]) continue;
    a.out_dims[a.order_out] = shape[d];
    a.out_strides_in[a.order_out] = strides[d];
    ++a.order_out;
    n *= (size_t)shape[d];
  }
  hipLaunchKernelGGL(k_trace, dim3(grid_for(n)), dim3(256), 0, st, in, out, n, a);
  HBP_HIP(hipGetLastError());
}

// Complex128 elements of W the SVD_FORM_W_LDS launch can hold on the current
// device (0: that form is not used): the device's LDS per workgroup less the
// kernel's static LDS, once the kernel's dynamic-LDS limit has been raised.
static int64_t svd_w_lds_cap() {
  static std::once_flag once;
  static int64_t cap = 0;
  std::call_once(once, [] {
    int dev = 0, lds = 0;
    hipFuncAttributes fa{};
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess ||
        hipFuncGetAttributes(&fa, (const void *)k_batched_svd<SVD_FORM_W_LDS>) != hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    const int64_t bytes = ((int64_t)lds - (int64_t)fa.sharedSizeBytes) / 16 * 16;
    if (bytes <= (int64_t)SVD_LDS_ELEMS * 16) return;
    if (hipFuncSetAttribute((const void *)k_batched_svd<SVD_FORM_W_LDS>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) !=
        hipSuccess) {
      (void)hipGetLastError();
      return;
    }
    cap = bytes / 16;
  });
  return cap;
}

void launch_batched_svd(hipStream_t st, const double2 *A, double2 *U, double *s, double2 *Vd,
                        double2 *work, const SvdJob *jobs_dev, int batch, int max_n, int max_m) {
  HBP_REQUIRE(max_n <= 512, HBP_ERR_UNSUPPORTED, "svd: more than 512 columns");
  if (batch <= 0) return;
  const int64_t w_elems = (int64_t)max_m * max_n;
  const int64_t elems = w_elems + (int64_t)max_n * max_n;
  if (max_m > 0 && elems <= SVD_LDS_ELEMS) {
    const unsigned bytes = (unsigned)(elems * sizeof(double2));
    static std::once_flag once;
    std::call_once(once, [] {
      (void)hipFuncSetAttribute((const void *)k_batched_svd<SVD_FORM_LDS>,
                                hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)(SVD_LDS_ELEMS * sizeof(double2)));
    });
    hipLaunchKernelGGL(k_batched_svd<SVD_FORM_LDS>, dim3(batch), dim3(SVD_THREADS), bytes, st, A,
                       U, s, Vd, work, jobs_dev);
  } else if (max_m > 0 && w_elems <= svd_w_lds_cap()) {
    hipLaunchKernelGGL(k_batched_svd<SVD_FORM_W_LDS>, dim3(batch), dim3(SVD_THREADS),
                       (unsigned)(w_elems * sizeof(double2)), st, A, U, s, Vd, work, jobs_dev);
  } else {
    hipLaunchKernelGGL(k_batched_svd<SVD_FORM_GLOBAL>, dim3(batch), dim3(SVD_THREADS), 0, st, A,
                       U, s, Vd, work, jobs_dev);
  }
  HBP_HIP(hipGetLastError());
}

void launch_store_svd_job(hipStream_t st, const SvdJob &job, SvdJob *dst) {
  hipLaunchKernelGGL(k_store_svd_job, dim3(1), dim3(64), 0, st, job, dst);
  HBP_HIP(hipGetLastError());
}

void launch_batched_gemm(hipStream_t st, const double2 *A, const double2 *B, double2 *C,
                         const GemmJob *jobs_dev, int batch, int max_m, int max_n) {
  if (batch <= 0) return;
  const int64_t tot = (int64_t)max_m * max_n;
  int gx = (int)std::min<int64_t>((tot + 255) / 256, 64);
  if (gx < 1) gx = 1;
  hipLaunchKernelGGL(k_batched_gemm, dim3(gx, batch), dim3(256), 0, st, A, B, C, jobs_dev);
  HBP_HIP(hipGetLastError());
}

}  // namespace hbp
