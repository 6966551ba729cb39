/*
 * tci/hbp/fused.h - tnbp's algorithm functions on the fused PEPS layer.
 *
 * Non-template overloads of tnbp::AbsorbTPO (apply.h:14-177),
 * tnbp::BeliefPropagation (bp.h:65-261), tnbp::BeliefPropagationCondition
 * (bp.h:266-355), tnbp::Truncation (truncation.h:17-222) and both
 * tnbp::Measure (measure.h:14-94, :99-300) for complex128 hbp tensors.  The
 * drivers' qualified calls (apps/kicked_ising_lattice/main.cc:121-231,
 * apps/estimator/main.cc:119-240) pick them over the templates by overload
 * resolution, so the drivers run on the fused kernels without an edit.
 *
 * They run one cached hbp_peps ("mirror") per lattice that reads the drivers'
 * own site tensors in place (hbp_peps_bind_sites) and copies the messages in
 * and out.  Tensor versions (hbp_tensor_version) tell when V or E changed.
 * A BP iteration of the drivers (every layer of BeliefPropagation, E <- F,
 * BeliefPropagationCondition) costs one sweep: the first layer call computes
 * every message of the sweep (Jacobi over all layers gives the same F as
 * layer by layer), later layers copy theirs, and the condition's recomputed
 * messages - exactly the next iteration's F - stay in the mirror for the next
 * BeliefPropagation on the same E.  The residual is the reference's (no
 * phase gauge), so iteration counts are the reference's.
 *
 * The overloads take the fused path for one MPI rank holding the whole
 * lattice (SiteIdx = every site, EdgeIdx = every edge) and hand every other
 * case to the templates.  Define HBP_TCI_NO_FUSED to compile the drivers on
 * the plain TCI path alone.
 */
#ifndef TCI_HBP_FUSED_H
#define TCI_HBP_FUSED_H

#if !defined(HBP_TCI_NO_FUSED) && __has_include(<mpi.h>)
#define HBP_TCI_FUSED 1

#include <mpi.h>

#include <algorithm>
#include <complex>
#include <cstdint>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "tci/hbp/tensor.h"

namespace hbp {
namespace fused {

using CT = Tensor<std::complex<double>>;
using Edges = std::vector<std::pair<int, int>>;

inline std::vector<uint64_t> versions(const std::vector<CT> &ts) {
  std::vector<uint64_t> v(ts.size(), 0);
  for (size_t i = 0; i < ts.size(); ++i)
    if (!ts[i].empty()) check(hbp_tensor_version(ts[i].handle(), &v[i]));
  return v;
}

inline std::vector<const hbp_tensor *> handles(const std::vector<CT> &ts) {
  std::vector<const hbp_tensor *> h(ts.size());
  for (size_t i = 0; i < ts.size(); ++i) h[i] = ts[i].handle();
  return h;
}

// The PEPS state of one lattice, reading the caller's V in place.
struct Mirror {
  std::shared_ptr<hbp_ctx> ctx;
  hbp_peps *p = nullptr;
  Edges I;
  std::vector<const hbp_tensor *> vh;  // bound site tensors and their versions
  std::vector<uint64_t> vv;
  std::vector<uint64_t> ev;  // versions of the E the mirror's messages came from
  bool swept = false;        // mirror messages = BP(E of versions ev)

  ~Mirror() { reset(); }
  void reset() {
    if (p && ctx) hbp_peps_destroy(ctx.get(), p);
    p = nullptr;
    vh.clear();
    vv.clear();
    ev.clear();
    swept = false;
  }
};

inline Mirror &mirror() {
  static Mirror m;
  return m;
}

// The mirror of lattice I on ctx with the sites of V bound.
inline Mirror &prepare(Context &cx, const Edges &I, const std::vector<CT> &V) {
  Mirror &M = mirror();
  hbp_ctx *c = cx.get();
  if (!M.p || M.ctx.get() != c || M.I != I) {
    M.reset();
    M.ctx = cx.h;
    M.I = I;
    std::vector<int32_t> flat;
    for (const auto &e : I) {
      flat.push_back(e.first);
      flat.push_back(e.second);
    }
    std::vector<int> sites;
    for (const auto &e : I) {
      sites.push_back(e.first);
      sites.push_back(e.second);
    }
    std::sort(sites.begin(), sites.end());
    sites.erase(std::unique(sites.begin(), sites.end()), sites.end());
    std::vector<int32_t> phys;
    for (size_t s = 0; s < sites.size(); ++s) phys.push_back(V[s].shape().back());
    check(hbp_peps_create(c, (int)I.size(), flat.data(), phys.data(), HBP_PREC_FP64, &M.p));
    check(hbp_peps_set_phase_gauge(M.p, 0));  // the reference's residual (bp.h:321-350)
  }
  const auto h = handles(V);
  const auto v = versions(V);
  if (h != M.vh || v != M.vv) {
    check(hbp_peps_bind_sites(c, M.p, h.data()));
    M.vh = h;
    M.vv = v;
    M.swept = false;
  }
  return M;
}

inline void load(Mirror &M, const std::vector<CT> &E) {
  const auto h = handles(E);
  check(hbp_peps_load_messages(M.ctx.get(), M.p, h.data()));
}

// one rank with the whole lattice: the case the PEPS layer covers
inline bool whole_lattice(MPI_Comm comm, const Edges &I, const std::vector<int> &SiteIdx,
                          const std::vector<int> &EdgeIdx, size_t n_v, size_t n_e) {
  int size = 1;
  MPI_Comm_size(comm, &size);
  if (size != 1 || EdgeIdx.size() != I.size() || n_e != 2 * I.size()) return false;
  for (size_t m = 0; m < EdgeIdx.size(); ++m)
    if (EdgeIdx[m] != (int)m) return false;
  std::vector<int> sites;
  for (const auto &e : I) {
    sites.push_back(e.first);
    sites.push_back(e.second);
  }
  std::sort(sites.begin(), sites.end());
  sites.erase(std::unique(sites.begin(), sites.end()), sites.end());
  return SiteIdx == sites && n_v == sites.size();
}

inline std::vector<std::complex<double>> host_ops(const std::vector<CT> &O, size_t per) {
  std::vector<std::complex<double>> buf;
  buf.reserve(O.size() * per);
  for (const auto &o : O) {
    const auto h = o.to_host();
    if (h.size() != per) throw std::invalid_argument("Measure: operator of unexpected size");
    buf.insert(buf.end(), h.begin(), h.end());
  }
  return buf;
}

}  // namespace fused
}  // namespace hbp

namespace tnbp {

// The templates (declared again here, defined in include/tnbp/ptns/function):
// the overloads below hand them the cases the PEPS layer does not cover.
template <typename TenT>
void AbsorbTPO(typename tci::tensor_traits<TenT>::context_handle_t &ctx,
               const std::vector<TenT> &O, const std::vector<std::pair<int, int>> &Edge,
               std::vector<TenT> &V, const std::vector<int> &SiteIdx,
               const std::map<int, int> &Site_To_MpiRank, std::vector<TenT> &E,
               const std::vector<int> &EdgeIdx, MPI_Comm comm);
template <typename TenT>
void BeliefPropagation(typename tci::tensor_traits<TenT>::context_handle_t &ctx,
                       const std::vector<std::pair<int, int>> &I, const std::vector<TenT> &V,
                       const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
                       const std::vector<std::pair<int, int>> &J, const std::vector<TenT> &E,
                       const std::vector<int> &EdgeIdx, MPI_Comm comm, std::vector<TenT> &F);
template <typename TenT>
void BeliefPropagationCondition(typename tci::tensor_traits<TenT>::context_handle_t &ctx,
                                const std::vector<std::pair<int, int>> &I,
                                const std::vector<TenT> &V, const std::vector<int> &SiteIdx,
                                const std::map<int, int> &Site_To_MpiRank,
                                const std::vector<TenT> &E, const std::vector<int> &EdgeIdx,
                                MPI_Comm comm, typename tci::tensor_traits<TenT>::real_t &result);
template <typename TenT>
void Truncation(typename tci::tensor_traits<TenT>::context_handle_t &ctx,
                const std::vector<std::pair<int, int>> &I, std::vector<TenT> &V,
                const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
                std::vector<TenT> &E, const std::vector<int> &EdgeIdx, MPI_Comm comm,
                typename tci::tensor_traits<TenT>::bond_dim_t max_dim,
                typename tci::tensor_traits<TenT>::real_t eps,
                typename tci::tensor_traits<TenT>::real_t err,
                std::vector<typename tci::tensor_traits<TenT>::bond_dim_t> &res_bond_dim,
                std::vector<typename tci::tensor_traits<TenT>::real_t> &res_trunc_err);
template <typename TenT>
std::vector<typename tci::tensor_traits<TenT>::elem_t> Measure(
    typename tci::tensor_traits<TenT>::context_handle_t &ctx,
    const std::vector<std::pair<int, int>> &I, const std::vector<TenT> &V,
    const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
    const std::vector<TenT> &E, const std::vector<int> EdgeIdx, MPI_Comm comm,
    const std::vector<int> &Site, const std::vector<TenT> &O);
template <typename TenT>
std::vector<typename tci::tensor_traits<TenT>::elem_t> Measure(
    typename tci::tensor_traits<TenT>::context_handle_t &ctx,
    const std::vector<std::pair<int, int>> &I, const std::vector<TenT> &V,
    const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
    const std::vector<TenT> &E, const std::vector<int> EdgeIdx, MPI_Comm comm,
    const std::vector<std::pair<int, int>> &Edge, const std::vector<TenT> &O);

using HbpCT = hbp::fused::CT;

// AbsorbTPO (apply.h:14-177): the operators absorbed lazily in the mirror,
// then every site materialised into the caller's V and E <- E (x) I.
inline void AbsorbTPO(hbp::Context &ctx, const std::vector<HbpCT> &O,
                      const std::vector<std::pair<int, int>> &Edge, std::vector<HbpCT> &V,
                      const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
                      std::vector<HbpCT> &E, const std::vector<int> &EdgeIdx, MPI_Comm comm) {
  namespace F = hbp::fused;
  if (!F::whole_lattice(comm, Edge, SiteIdx, EdgeIdx, V.size(), E.size())) {
    AbsorbTPO<HbpCT>(ctx, O, Edge, V, SiteIdx, Site_To_MpiRank, E, EdgeIdx, comm);
    return;
  }
  F::Mirror &M = F::prepare(ctx, Edge, V);
  F::load(M, E);
  std::vector<std::vector<std::complex<double>>> host(O.size());
  std::vector<const double *> ptrs(O.size());
  std::vector<int32_t> shapes(O.size() * HBP_MAX_ORDER, 0);
  for (size_t s = 0; s < O.size(); ++s) {
    host[s] = O[s].to_host();
    ptrs[s] = reinterpret_cast<const double *>(host[s].data());
    const auto sh = O[s].shape();
    for (size_t d = 0; d < sh.size() && d < (size_t)HBP_MAX_ORDER; ++d)
      shapes[s * HBP_MAX_ORDER + d] = sh[d];
  }
  hbp::check(hbp_peps_set_absorb_mode(M.p, HBP_ABSORB_LAZY));
  hbp::check(hbp_peps_absorb_tpo(ctx.get(), M.p, ptrs.data(), shapes.data()));
  std::vector<hbp_tensor *> vo(V.size());
  for (size_t s = 0; s < V.size(); ++s) vo[s] = V[s].out(ctx);
  hbp::check(hbp_peps_store_sites(ctx.get(), M.p, vo.data()));
  std::vector<hbp_tensor *> eo(E.size());
  for (size_t m = 0; m < E.size(); ++m) eo[m] = E[m].out(ctx);
  hbp::check(hbp_peps_store_messages(ctx.get(), M.p, (int)E.size(), nullptr, eo.data()));
  M.vh = F::handles(V);
  M.vv = F::versions(V);
  M.swept = false;
}

// BeliefPropagation (bp.h:65-261) for one layer J: the messages of the whole
// sweep are computed once per (V, E) and the layer's copied into F.
inline void BeliefPropagation(hbp::Context &ctx, const std::vector<std::pair<int, int>> &I,
                              const std::vector<HbpCT> &V, const std::vector<int> &SiteIdx,
                              const std::map<int, int> &Site_To_MpiRank,
                              const std::vector<std::pair<int, int>> &J,
                              const std::vector<HbpCT> &E, const std::vector<int> &EdgeIdx,
                              MPI_Comm comm, std::vector<HbpCT> &F) {
  namespace FS = hbp::fused;
  if (!FS::whole_lattice(comm, I, SiteIdx, EdgeIdx, V.size(), E.size())) {
    BeliefPropagation<HbpCT>(ctx, I, V, SiteIdx, Site_To_MpiRank, J, E, EdgeIdx, comm, F);
    return;
  }
  FS::Mirror &M = FS::prepare(ctx, I, V);
  const auto ev = FS::versions(E);
  if (!(M.swept && M.ev == ev)) {
    FS::load(M, E);
    hbp::check(hbp_peps_bp_sweep(ctx.get(), M.p));
    M.ev = ev;
    M.swept = true;
  }
  const int ne = (int)I.size();
  if (F.size() < (size_t)(2 * ne)) F.resize(2 * ne);
  std::vector<int32_t> which;
  std::vector<hbp_tensor *> out;
  for (const auto &e : J) {
    int m = -1;
    for (int k = 0; k < ne && m < 0; ++k)
      if ((I[k].first == e.first && I[k].second == e.second) ||
          (I[k].first == e.second && I[k].second == e.first))
        m = k;
    if (m < 0) throw std::invalid_argument("BeliefPropagation: layer edge not in the lattice");
    which.push_back(m);
    out.push_back(F[m].out(ctx));
    which.push_back(m + ne);
    out.push_back(F[m + ne].out(ctx));
  }
  hbp::check(hbp_peps_store_messages(ctx.get(), M.p, (int)which.size(), which.data(), out.data()));
}

// BeliefPropagationCondition (bp.h:266-355): G = BP(E) and the mean distance
// of normalised G and E; G stays in the mirror as BeliefPropagation's F for
// this E.
inline void BeliefPropagationCondition(hbp::Context &ctx, const std::vector<std::pair<int, int>> &I,
                                       const std::vector<HbpCT> &V, const std::vector<int> &SiteIdx,
                                       const std::map<int, int> &Site_To_MpiRank,
                                       const std::vector<HbpCT> &E, const std::vector<int> &EdgeIdx,
                                       MPI_Comm comm, double &result) {
  namespace FS = hbp::fused;
  if (!FS::whole_lattice(comm, I, SiteIdx, EdgeIdx, V.size(), E.size())) {
    BeliefPropagationCondition<HbpCT>(ctx, I, V, SiteIdx, Site_To_MpiRank, E, EdgeIdx, comm,
                                      result);
    return;
  }
  FS::Mirror &M = FS::prepare(ctx, I, V);
  FS::load(M, E);
  int it = 0, conv = 0;
  double r = 0.0;
  // one iteration that never stops early (tolerance 0): messages <- G
  hbp::check(hbp_peps_bp_iterate(ctx.get(), M.p, 1, 0.0, 0, 0, &it, &r, &conv));
  M.ev = FS::versions(E);
  M.swept = true;
  result = r;
}

// Truncation (truncation.h:17-222) in the mirror; the new sites and the
// diag(Lambda) messages are written back into V and E.
inline void Truncation(hbp::Context &ctx, const std::vector<std::pair<int, int>> &I,
                       std::vector<HbpCT> &V, const std::vector<int> &SiteIdx,
                       const std::map<int, int> &Site_To_MpiRank, std::vector<HbpCT> &E,
                       const std::vector<int> &EdgeIdx, MPI_Comm comm, int32_t max_dim, double eps,
                       double err, std::vector<int32_t> &res_bond_dim,
                       std::vector<double> &res_trunc_err) {
  namespace FS = hbp::fused;
  if (!FS::whole_lattice(comm, I, SiteIdx, EdgeIdx, V.size(), E.size())) {
    Truncation<HbpCT>(ctx, I, V, SiteIdx, Site_To_MpiRank, E, EdgeIdx, comm, max_dim, eps, err,
                      res_bond_dim, res_trunc_err);
    return;
  }
  FS::Mirror &M = FS::prepare(ctx, I, V);
  FS::load(M, E);
  res_bond_dim.assign(I.size(), 0);
  res_trunc_err.assign(I.size(), 0.0);
  hbp::check(hbp_peps_truncate(ctx.get(), M.p, max_dim, eps, res_bond_dim.data(),
                               res_trunc_err.data()));
  std::vector<hbp_tensor *> vo(V.size());
  for (size_t s = 0; s < V.size(); ++s) vo[s] = V[s].out(ctx);
  hbp::check(hbp_peps_store_sites(ctx.get(), M.p, vo.data()));
  std::vector<hbp_tensor *> eo(E.size());
  for (size_t m = 0; m < E.size(); ++m) eo[m] = E[m].out(ctx);
  hbp::check(hbp_peps_store_messages(ctx.get(), M.p, (int)E.size(), nullptr, eo.data()));
  M.vh = FS::handles(V);
  M.vv = FS::versions(V);
  M.swept = false;
}

// single-site Measure (measure.h:14-94)
inline std::vector<std::complex<double>> Measure(
    hbp::Context &ctx, const std::vector<std::pair<int, int>> &I, const std::vector<HbpCT> &V,
    const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
    const std::vector<HbpCT> &E, const std::vector<int> EdgeIdx, MPI_Comm comm,
    const std::vector<int> &Site, const std::vector<HbpCT> &O) {
  namespace FS = hbp::fused;
  if (!FS::whole_lattice(comm, I, SiteIdx, EdgeIdx, V.size(), E.size()))
    return Measure<HbpCT>(ctx, I, V, SiteIdx, Site_To_MpiRank, E, EdgeIdx, comm, Site, O);
  std::vector<std::complex<double>> res(Site.size(), 0.0);
  if (Site.empty()) return res;
  FS::Mirror &M = FS::prepare(ctx, I, V);
  FS::load(M, E);
  M.swept = false;
  const size_t d = (size_t)V[0].shape().back();
  const auto ops = FS::host_ops(O, d * d);
  std::vector<int32_t> sites(Site.begin(), Site.end());
  hbp::check(hbp_peps_measure_sites(ctx.get(), M.p, (int)Site.size(), sites.data(),
                                    reinterpret_cast<const double *>(ops.data()),
                                    reinterpret_cast<double *>(res.data())));
  return res;
}

// nearest-neighbour Measure (measure.h:99-300); the result has |I| entries
// (measure.h:123), or one per pair when there are more pairs than edges
inline std::vector<std::complex<double>> Measure(
    hbp::Context &ctx, const std::vector<std::pair<int, int>> &I, const std::vector<HbpCT> &V,
    const std::vector<int> &SiteIdx, const std::map<int, int> &Site_To_MpiRank,
    const std::vector<HbpCT> &E, const std::vector<int> EdgeIdx, MPI_Comm comm,
    const std::vector<std::pair<int, int>> &Edge, const std::vector<HbpCT> &O) {
  namespace FS = hbp::fused;
  if (!FS::whole_lattice(comm, I, SiteIdx, EdgeIdx, V.size(), E.size()))
    return Measure<HbpCT>(ctx, I, V, SiteIdx, Site_To_MpiRank, E, EdgeIdx, comm, Edge, O);
  std::vector<std::complex<double>> res(std::max(I.size(), Edge.size()), 0.0);
  if (Edge.empty()) return res;
  FS::Mirror &M = FS::prepare(ctx, I, V);
  FS::load(M, E);
  M.swept = false;
  const size_t d = (size_t)V[0].shape().back();
  const auto ops = FS::host_ops(O, d * d * d * d);
  std::vector<int32_t> pairs;
  for (const auto &e : Edge) {
    pairs.push_back(e.first);
    pairs.push_back(e.second);
  }
  hbp::check(hbp_peps_measure_edges(ctx.get(), M.p, (int)Edge.size(), pairs.data(),
                                    reinterpret_cast<const double *>(ops.data()),
                                    reinterpret_cast<double *>(res.data())));
  return res;
}

}  // namespace tnbp

#endif  // fused path available
#endif
