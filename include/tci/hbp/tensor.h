/*
 * tci/hbp/tensor.h - device tensor with value semantics over the hbp C-ABI.
 */
#ifndef TCI_HBP_TENSOR_H
#define TCI_HBP_TENSOR_H

#include <complex>
#include <cstdint>
#include <cstdlib>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <system_error>
#include <type_traits>
#include <vector>

#include "hbp.h"
#include "tci/traits.h"

namespace hbp {

// Turn an hbp_status into the exception min-tci would have thrown
// (std::invalid_argument / logic_error / bad_alloc / out_of_range).
inline void check(hbp_status s) {
  if (s == HBP_OK) return;
  const std::string msg = hbp_last_error();
  switch (s) {
    case HBP_ERR_INVALID_ARGUMENT: throw std::invalid_argument(msg);
    case HBP_ERR_BAD_ALLOC: throw std::bad_alloc();
    case HBP_ERR_OUT_OF_RANGE: throw std::out_of_range(msg);
    default: throw std::logic_error(msg);
  }
}

// One hbp context (device, stream, memory pool) per device per process; every
// context handle and tensor on that device shares it.
inline std::shared_ptr<hbp_ctx> device_context(int device) {
  static std::mutex mu;
  static std::map<int, std::weak_ptr<hbp_ctx>> live;
  std::lock_guard<std::mutex> lock(mu);
  if (auto p = live[device].lock()) return p;
  hbp_ctx *c = nullptr;
  check(hbp_ctx_create(device, &c));
  std::shared_ptr<hbp_ctx> p(c, [](hbp_ctx *x) { hbp_ctx_destroy(x); });
  live[device] = p;
  return p;
}

inline int default_device() {
  const char *e = std::getenv("HBP_DEVICE");
  return e ? std::atoi(e) : 0;
}

// tci::context_handle_t of the hbp backend.
struct Context {
  std::shared_ptr<hbp_ctx> h;
  hbp_ctx *get() {
    if (!h) h = device_context(default_device());
    return h.get();
  }
};

template <typename E> struct dtype_of;
template <> struct dtype_of<std::complex<double>> { static constexpr int value = HBP_C128; };
template <> struct dtype_of<double> { static constexpr int value = HBP_F64; };
template <> struct dtype_of<std::complex<float>> { static constexpr int value = HBP_C64; };
template <> struct dtype_of<float> { static constexpr int value = HBP_F32; };

// Column-major device tensor of element type E.  Copies are deep (tci::copy
// semantics, constructor.h:192-204), moves transfer the device buffer.
template <typename E>
class Tensor {
 public:
  Tensor() = default;
  Tensor(Context &ctx, const std::vector<int32_t> &shape) { create(ctx, shape); }
  Tensor(const Tensor &o) { *this = o; }
  Tensor(Tensor &&o) noexcept : ctx_(std::move(o.ctx_)), t_(o.t_) { o.t_ = nullptr; }
  ~Tensor() { reset(); }

  Tensor &operator=(const Tensor &o) {
    if (this == &o) return *this;
    if (!o.t_) {
      reset();
      return *this;
    }
    if (!t_) {
      ctx_ = o.ctx_;
      int32_t one = 1;
      check(hbp_tensor_create(ctx_.get(), dtype_of<E>::value, 1, &one, &t_));
    }
    check(hbp_tensor_copy(ctx_.get(), o.t_, t_));
    return *this;
  }
  Tensor &operator=(Tensor &&o) noexcept {
    if (this != &o) {
      reset();
      ctx_ = std::move(o.ctx_);
      t_ = o.t_;
      o.t_ = nullptr;
    }
    return *this;
  }

  void create(Context &ctx, const std::vector<int32_t> &shape) {
    reset();
    ctx.get();
    ctx_ = ctx.h;
    check(hbp_tensor_create(ctx_.get(), dtype_of<E>::value, (int)shape.size(), shape.data(), &t_));
  }
  // an empty output tensor receives a handle on first use
  hbp_tensor *out(Context &ctx) {
    if (!t_) create(ctx, {1});
    return t_;
  }
  hbp_tensor *handle() const {
    if (!t_) throw std::logic_error("tci: use of an empty tensor");
    return t_;
  }
  hbp_ctx *ctx() const { return ctx_.get(); }
  bool empty() const { return t_ == nullptr; }

  std::vector<int32_t> shape() const {
    if (!t_) return {};
    int order = 0;
    int32_t sh[HBP_MAX_ORDER];
    check(hbp_tensor_info(t_, nullptr, &order, sh, nullptr));
    return std::vector<int32_t>(sh, sh + order);
  }
  size_t size() const {
    if (!t_) return 0;
    size_t n = 0;
    check(hbp_tensor_info(t_, nullptr, nullptr, nullptr, &n));
    return n;
  }
  std::vector<E> to_host() const {
    std::vector<E> v(size());
    if (!v.empty()) check(hbp_tensor_download(ctx_.get(), handle(), v.data()));
    return v;
  }
  void from_host(Context &ctx, const std::vector<int32_t> &shape, const std::vector<E> &v) {
    create(ctx, shape);
    if (!v.empty()) check(hbp_tensor_upload(ctx_.get(), t_, v.data()));
  }

 private:
  void reset() {
    if (t_ && ctx_) hbp_tensor_destroy(ctx_.get(), t_);
    t_ = nullptr;
  }
  std::shared_ptr<hbp_ctx> ctx_;
  hbp_tensor *t_ = nullptr;
};

}  // namespace hbp

// The apps name their tensor type gqten::tensor<E> (apps/kicked_ising_lattice/type.h:8-14,
// apps/estimator/typedef.h:1-8); with this backend that name is the hbp tensor.
namespace gqten {
template <typename E>
using tensor = hbp::Tensor<E>;
}

namespace tci {

template <class T> struct is_hbp_tensor : std::false_type {};
template <class E> struct is_hbp_tensor<hbp::Tensor<E>> : std::true_type {};
template <class T> inline constexpr bool is_hbp_tensor_v = is_hbp_tensor<T>::value;

template <typename E, typename R, typename BondIdx>
struct hbp_traits {
  using ten_t = hbp::Tensor<E>;
  using order_t = int32_t;
  using shape_t = std::vector<int32_t>;
  using bond_dim_t = int32_t;
  using bond_idx_t = BondIdx;
  using bond_label_t = int32_t;
  using ten_size_t = size_t;
  using elem_t = E;
  using elem_coor_t = int32_t;
  using elem_coors_t = std::vector<int32_t>;
  using real_t = R;
  using real_ten_t = hbp::Tensor<R>;
  using cplx_t = std::complex<R>;
  using cplx_ten_t = hbp::Tensor<std::complex<R>>;
  using context_handle_t = hbp::Context;
};

template <> struct tensor_traits<hbp::Tensor<std::complex<double>>>
    : hbp_traits<std::complex<double>, double, int32_t> {};
template <> struct tensor_traits<hbp::Tensor<double>> : hbp_traits<double, double, size_t> {};
template <> struct tensor_traits<hbp::Tensor<std::complex<float>>>
    : hbp_traits<std::complex<float>, float, int32_t> {};
template <> struct tensor_traits<hbp::Tensor<float>> : hbp_traits<float, float, int32_t> {};

}  // namespace tci

#endif
