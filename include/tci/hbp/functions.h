/*
 * tci/hbp/functions.h - the TCI free functions on hbp tensors.
 *
 * Names, argument order and meaning follow min-tci
 * (external/min-tci/include/tci/gqten/{getter,constructor,manipulator,linalg,
 * miscellaneous,io}.h); each function forwards to one hbp_* entry point.
 */
#ifndef TCI_HBP_FUNCTIONS_H
#define TCI_HBP_FUNCTIONS_H

#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iostream>
#include <numeric>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "tci/hbp/tensor.h"

namespace tci {

#define HBP_TCI_T template <typename TenT, std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>

// ------------------------------------------------------------ context
inline void create_context(hbp::Context &ctx) { ctx.get(); }
inline void destroy_context(hbp::Context &ctx) { ctx.h.reset(); }

// ------------------------------------------------------------ getters
HBP_TCI_T inline order_t<TenT> order(context_handle_t<TenT> &, const TenT &a) {
  return (order_t<TenT>)a.shape().size();
}
HBP_TCI_T inline shape_t<TenT> shape(context_handle_t<TenT> &, const TenT &a) { return a.shape(); }
HBP_TCI_T inline ten_size_t<TenT> size(context_handle_t<TenT> &, const TenT &a) { return a.size(); }

HBP_TCI_T inline elem_t<TenT> get_elem(context_handle_t<TenT> &ctx, const TenT &a,
                                       const elem_coors_t<TenT> &coors) {
  double re = 0.0, im = 0.0;
  std::vector<int32_t> c(coors.begin(), coors.end());
  if (c.empty()) c.push_back(0);
  hbp::check(hbp_get_elem(ctx.get(), a.handle(), c.data(), &re, &im));
  if constexpr (std::is_same_v<elem_t<TenT>, std::complex<double>> ||
                std::is_same_v<elem_t<TenT>, std::complex<float>>)
    return elem_t<TenT>(re, im);
  else
    return (elem_t<TenT>)re;
}
HBP_TCI_T inline void get_elem(context_handle_t<TenT> &ctx, const TenT &a,
                               const elem_coors_t<TenT> &coors, elem_t<TenT> &elem) {
  elem = get_elem(ctx, a, coors);
}

// -------------------------------------------------------- constructors
HBP_TCI_T inline void fill(context_handle_t<TenT> &ctx, const shape_t<TenT> &shape,
                           const elem_t<TenT> v, TenT &a) {
  a.create(ctx, shape);
  const std::complex<double> z(v);
  hbp::check(hbp_fill(ctx.get(), a.handle(), z.real(), z.imag()));
}
HBP_TCI_T inline TenT fill(context_handle_t<TenT> &ctx, const shape_t<TenT> &shape,
                           const elem_t<TenT> v) {
  TenT a;
  fill(ctx, shape, v, a);
  return a;
}
HBP_TCI_T inline TenT zeros(context_handle_t<TenT> &ctx, const shape_t<TenT> &shape) {
  return TenT(ctx, shape);
}

// element at column-major coordinate c is init[coors2idx(c)] (constructor.h:114-163)
template <typename TenT, typename RandomIt, typename Func,
          std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
inline void assign_from_range(context_handle_t<TenT> &ctx, const shape_t<TenT> &shape,
                              RandomIt init, Func &&coors2idx, TenT &a) {
  size_t n = 1;
  for (auto d : shape) n *= (size_t)d;
  std::vector<elem_t<TenT>> data(n);
  std::vector<bond_dim_t<TenT>> coor(shape.size());
  for (size_t i = 0; i < n; ++i) {
    size_t rem = i;
    for (size_t k = 0; k < shape.size(); ++k) {
      coor[k] = (bond_dim_t<TenT>)(rem % shape[k]);
      rem /= shape[k];
    }
    data[i] = *(init + (std::ptrdiff_t)std::invoke(coors2idx, coor));
  }
  a.from_host(ctx, shape, data);
}
template <typename TenT, typename RandomIt, typename Func,
          std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
inline TenT assign_from_range(context_handle_t<TenT> &ctx, const shape_t<TenT> &shape,
                              RandomIt init, Func &&coors2idx) {
  TenT a;
  assign_from_range(ctx, shape, init, std::forward<Func>(coors2idx), a);
  return a;
}

// (miscellaneous.h:38-66): first[coors2idx(c)] = element at c
template <typename TenT, typename RandomIt, typename Func,
          std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
inline void to_range(context_handle_t<TenT> &, const TenT &a, RandomIt first, Func &&coors2idx) {
  const auto shape = a.shape();
  const auto data = a.to_host();
  std::vector<bond_dim_t<TenT>> coor(shape.size());
  for (size_t i = 0; i < data.size(); ++i) {
    size_t rem = i;
    for (size_t k = 0; k < shape.size(); ++k) {
      coor[k] = (bond_dim_t<TenT>)(rem % shape[k]);
      rem /= shape[k];
    }
    *(first + (std::ptrdiff_t)std::invoke(coors2idx, coor)) = data[i];
  }
}

HBP_TCI_T inline void copy(context_handle_t<TenT> &, const TenT &orig, TenT &dist) { dist = orig; }
HBP_TCI_T inline TenT copy(context_handle_t<TenT> &, const TenT &orig) { return orig; }
HBP_TCI_T inline void move(context_handle_t<TenT> &, TenT &from, TenT &to) noexcept {
  to = std::move(from);
}
HBP_TCI_T inline TenT move(context_handle_t<TenT> &, TenT &from) noexcept { return std::move(from); }
HBP_TCI_T inline void clear(context_handle_t<TenT> &, TenT &a) { a = TenT(); }

// -------------------------------------------------------- manipulators
HBP_TCI_T inline void reshape(context_handle_t<TenT> &ctx, TenT &inout,
                              const shape_t<TenT> &new_shape) {
  hbp::check(hbp_tensor_reshape(ctx.get(), inout.handle(), (int)new_shape.size(),
                                new_shape.data()));
}
HBP_TCI_T inline void reshape(context_handle_t<TenT> &ctx, const TenT &in,
                              const shape_t<TenT> &new_shape, TenT &out) {
  out = in;
  reshape(ctx, out, new_shape);
}

HBP_TCI_T inline void transpose(context_handle_t<TenT> &ctx, TenT &inout,
                                const List<bond_idx_t<TenT>> &new_order) {
  std::vector<int32_t> o(new_order.begin(), new_order.end());
  hbp::check(hbp_transpose(ctx.get(), inout.handle(), o.data(), inout.handle()));
}
HBP_TCI_T inline void transpose(context_handle_t<TenT> &ctx, const TenT &in,
                                const List<bond_idx_t<TenT>> &new_order, TenT &out) {
  out = in;
  transpose(ctx, out, new_order);
}

HBP_TCI_T inline void cplx_conj(context_handle_t<TenT> &ctx, TenT &inout) {
  hbp::check(hbp_conj(ctx.get(), inout.handle()));
}
HBP_TCI_T inline void cplx_conj(context_handle_t<TenT> &ctx, const TenT &in, TenT &out) {
  out = in;
  cplx_conj(ctx, out);
}

// (manipulator.h:104-142): real -> complex
HBP_TCI_T inline void to_cplx(context_handle_t<TenT> &ctx, const TenT &in,
                              cplx_ten_t<TenT> &out) {
  const auto v = in.to_host();
  std::vector<cplx_t<TenT>> c(v.begin(), v.end());
  out.from_host(ctx, in.shape(), c);
}
HBP_TCI_T inline cplx_ten_t<TenT> to_cplx(context_handle_t<TenT> &ctx, const TenT &in) {
  cplx_ten_t<TenT> out;
  to_cplx(ctx, in, out);
  return out;
}

// (manipulator.h:578-607): f mutates every element in place, on the host
template <typename TenT, typename Func, std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
inline void for_each(context_handle_t<TenT> &ctx, TenT &inout, Func &&f) {
  auto v = inout.to_host();
  for (auto &e : v) f(e);
  inout.from_host(ctx, inout.shape(), v);
}

// ------------------------------------------------------------ linalg
HBP_TCI_T inline void diag(context_handle_t<TenT> &ctx, TenT &inout) {
  hbp::check(hbp_diag(ctx.get(), inout.handle(), inout.handle()));
}
HBP_TCI_T inline void diag(context_handle_t<TenT> &ctx, const TenT &in, TenT &out) {
  out = in;
  diag(ctx, out);
}

HBP_TCI_T inline real_t<TenT> norm(context_handle_t<TenT> &ctx, const TenT &a) {
  double v = 0.0;
  hbp::check(hbp_norm(ctx.get(), a.handle(), &v));
  return (real_t<TenT>)v;
}
HBP_TCI_T inline real_t<TenT> normalize(context_handle_t<TenT> &ctx, TenT &inout) {
  double v = 0.0;
  hbp::check(hbp_normalize(ctx.get(), inout.handle(), &v));
  return (real_t<TenT>)v;
}
HBP_TCI_T inline real_t<TenT> normalize(context_handle_t<TenT> &ctx, const TenT in, TenT &out) {
  out = in;
  return normalize(ctx, out);
}

HBP_TCI_T inline void scale(context_handle_t<TenT> &ctx, TenT &inout, const elem_t<TenT> s) {
  const std::complex<double> z(s);
  hbp::check(hbp_scale(ctx.get(), inout.handle(), z.real(), z.imag()));
}
HBP_TCI_T inline void scale(context_handle_t<TenT> &ctx, const TenT &in, const elem_t<TenT> s,
                            TenT &out) {
  out = in;
  scale(ctx, out, s);
}

HBP_TCI_T inline void trace(context_handle_t<TenT> &ctx, const TenT &in,
                            const List<Pair<bond_idx_t<TenT>, bond_idx_t<TenT>>> &pairs,
                            TenT &out) {
  std::vector<int32_t> flat;
  for (const auto &p : pairs) {
    flat.push_back((int32_t)p.first);
    flat.push_back((int32_t)p.second);
  }
  hbp::check(hbp_trace(ctx.get(), in.handle(), (int)pairs.size(), flat.data(), out.out(ctx)));
}
HBP_TCI_T inline void trace(context_handle_t<TenT> &ctx, TenT &inout,
                            const List<Pair<bond_idx_t<TenT>, bond_idx_t<TenT>>> &pairs) {
  TenT tmp;
  trace(ctx, inout, pairs, tmp);
  inout = std::move(tmp);
}

// (linalg.h:311-323): shared labels are summed, survivors ordered by bd_labs_c
HBP_TCI_T inline void contract(context_handle_t<TenT> &ctx, const TenT &a,
                               const List<bond_label_t<TenT>> &bd_labs_a, const TenT &b,
                               const List<bond_label_t<TenT>> &bd_labs_b, TenT &c,
                               const List<bond_label_t<TenT>> &bd_labs_c) {
  if ((int)bd_labs_a.size() != (int)a.shape().size() ||
      (int)bd_labs_b.size() != (int)b.shape().size())
    throw std::invalid_argument("contract: label count differs from tensor order");
  hbp::check(hbp_contract(ctx.get(), a.handle(), bd_labs_a.data(), b.handle(), bd_labs_b.data(),
                          c.out(ctx), bd_labs_c.data(), (int)bd_labs_c.size()));
}

HBP_TCI_T inline void linear_combine(context_handle_t<TenT> &ctx, const List<TenT> &ins,
                                     const List<elem_t<TenT>> &coefs, TenT &out) {
  std::vector<const hbp_tensor *> hs;
  std::vector<double> cf;
  for (size_t i = 0; i < ins.size(); ++i) {
    hs.push_back(ins[i].handle());
    const std::complex<double> z(coefs[i]);
    cf.push_back(z.real());
    cf.push_back(z.imag());
  }
  hbp::check(hbp_linear_combine(ctx.get(), (int)hs.size(), hs.data(), cf.data(), out.out(ctx)));
}
HBP_TCI_T inline TenT linear_combine(context_handle_t<TenT> &ctx, const List<TenT> &ins,
                                     const List<elem_t<TenT>> &coefs) {
  TenT out;
  linear_combine(ctx, ins, coefs, out);
  return out;
}
HBP_TCI_T inline void linear_combine(context_handle_t<TenT> &ctx, const List<TenT> &ins,
                                     TenT &out) {
  linear_combine(ctx, ins, List<elem_t<TenT>>(ins.size(), elem_t<TenT>(1)), out);
}
HBP_TCI_T inline TenT linear_combine(context_handle_t<TenT> &ctx, const List<TenT> &ins) {
  TenT out;
  linear_combine(ctx, ins, out);
  return out;
}

// (linalg.h:500-512)
HBP_TCI_T inline void svd(context_handle_t<TenT> &ctx, const TenT &a,
                          const order_t<TenT> &num_of_bonds_as_rows, TenT &u,
                          real_ten_t<TenT> &s_diag, TenT &v_dag) {
  hbp::check(hbp_svd(ctx.get(), a.handle(), (int)num_of_bonds_as_rows, u.out(ctx),
                     s_diag.out(ctx), v_dag.out(ctx)));
}

// (linalg.h:552-569)
HBP_TCI_T inline void trunc_svd(context_handle_t<TenT> &ctx, const TenT &a,
                                const order_t<TenT> &num_of_bonds_as_rows, TenT &u,
                                real_ten_t<TenT> &s_diag, TenT &v_dag, real_t<TenT> &trunc_err,
                                const bond_dim_t<TenT> chi_max, const real_t<TenT> s_min) {
  double err = 0.0;
  hbp::check(hbp_trunc_svd(ctx.get(), a.handle(), (int)num_of_bonds_as_rows, (int)chi_max,
                           (double)s_min, u.out(ctx), s_diag.out(ctx), v_dag.out(ctx), &err));
  trunc_err = (real_t<TenT>)err;
}
// (linalg.h:517-546): chi_max = min(m, n)
HBP_TCI_T inline void trunc_svd(context_handle_t<TenT> &ctx, const TenT &a,
                                const order_t<TenT> &num_of_bonds_as_rows, TenT &u,
                                real_ten_t<TenT> &s_diag, TenT &v_dag, real_t<TenT> &trunc_err,
                                const real_t<TenT> s_min) {
  const auto sh = a.shape();
  int64_t m = 1, n = 1;
  for (int k = 0; k < (int)sh.size(); ++k) (k < num_of_bonds_as_rows ? m : n) *= sh[k];
  trunc_svd(ctx, a, num_of_bonds_as_rows, u, s_diag, v_dag, trunc_err,
            (bond_dim_t<TenT>)std::min(m, n), s_min);
}

// ----------------------------------------------------------------- io
// Binary tensor format of this backend (the reference's gqten stream format
// is not available to match): "HBPT", uint32 version = 1, int32 dtype,
// int32 order, int32 shape[order], then the column-major elements.
HBP_TCI_T inline void save_stream(const TenT &a, std::ostream &out) {
  const uint32_t version = 1;
  const int32_t dtype = hbp::dtype_of<elem_t<TenT>>::value;
  const auto sh = a.shape();
  const int32_t order = (int32_t)sh.size();
  out.write("HBPT", 4);
  out.write((const char *)&version, 4);
  out.write((const char *)&dtype, 4);
  out.write((const char *)&order, 4);
  out.write((const char *)sh.data(), 4 * (size_t)order);
  const auto v = a.to_host();
  out.write((const char *)v.data(), (std::streamsize)(v.size() * sizeof(elem_t<TenT>)));
  if (!out) throw std::system_error(errno, std::generic_category(), "tensor write failed");
}

template <typename TenT, typename Storage, std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
void save(context_handle_t<TenT> &, const TenT &a, Storage &&strg) {
  using S = std::decay_t<Storage>;
  if constexpr (std::is_base_of_v<std::ostream, S>) {
    save_stream(a, strg);
  } else {
    std::filesystem::path p{std::forward<Storage>(strg)};
    std::ofstream ofs(p, std::ios::binary);
    if (!ofs)
      throw std::system_error(errno, std::generic_category(), "Failed to open file for write: " + p.string());
    save_stream(a, ofs);
  }
}

template <typename TenT, std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
TenT load_stream(context_handle_t<TenT> &ctx, std::istream &in) {
  char magic[4];
  uint32_t version = 0;
  int32_t dtype = -1, order = -1;
  in.read(magic, 4);
  in.read((char *)&version, 4);
  in.read((char *)&dtype, 4);
  in.read((char *)&order, 4);
  if (!in || std::memcmp(magic, "HBPT", 4) != 0 || version != 1 ||
      dtype != hbp::dtype_of<elem_t<TenT>>::value || order < 0 || order > HBP_MAX_ORDER)
    throw std::system_error(EINVAL, std::generic_category(), "not an HBPT v1 tensor of this type");
  std::vector<int32_t> sh(order);
  in.read((char *)sh.data(), 4 * (size_t)order);
  size_t n = 1;
  for (auto d : sh) n *= (size_t)d;
  std::vector<elem_t<TenT>> v(n);
  in.read((char *)v.data(), (std::streamsize)(n * sizeof(elem_t<TenT>)));
  if (!in) throw std::system_error(EIO, std::generic_category(), "truncated tensor stream");
  TenT a;
  a.from_host(ctx, sh, v);
  return a;
}

template <typename TenT, typename Storage, std::enable_if_t<is_hbp_tensor_v<TenT>, int> = 0>
TenT load(context_handle_t<TenT> &ctx, Storage &&strg) {
  using S = std::decay_t<Storage>;
  if constexpr (std::is_base_of_v<std::istream, S>) {
    return load_stream<TenT>(ctx, strg);
  } else {
    std::filesystem::path p{std::forward<Storage>(strg)};
    std::ifstream ifs(p, std::ios::binary);
    if (!ifs)
      throw std::system_error(errno, std::generic_category(), "Failed to open file for read: " + p.string());
    return load_stream<TenT>(ctx, ifs);
  }
}

#undef HBP_TCI_T

}  // namespace tci

#endif
