/*
 * tci/traits.h - the Tensor Computing Interface's trait names, for the hbp
 * backend.  The names and alias templates are the interface the tnbp
 * algorithms are written against (min-tci include/tci/traits.h), so they are
 * kept identical; every backend supplies its own tensor_traits<TenT>.
 */
#ifndef TCI_TRAITS_H
#define TCI_TRAITS_H

#include <map>
#include <utility>
#include <vector>

namespace tci {

template <typename TenT>
struct tensor_traits {};

template <typename T> using ten_t = typename tensor_traits<T>::ten_t;
template <typename T> using order_t = typename tensor_traits<T>::order_t;
template <typename T> using shape_t = typename tensor_traits<T>::shape_t;
template <typename T> using bond_dim_t = typename tensor_traits<T>::bond_dim_t;
template <typename T> using bond_idx_t = typename tensor_traits<T>::bond_idx_t;
template <typename T> using bond_label_t = typename tensor_traits<T>::bond_label_t;
template <typename T> using ten_size_t = typename tensor_traits<T>::ten_size_t;
template <typename T> using elem_t = typename tensor_traits<T>::elem_t;
template <typename T> using elem_coor_t = typename tensor_traits<T>::elem_coor_t;
template <typename T> using elem_coors_t = typename tensor_traits<T>::elem_coors_t;
template <typename T> using real_t = typename tensor_traits<T>::real_t;
template <typename T> using real_ten_t = typename tensor_traits<T>::real_ten_t;
template <typename T> using cplx_t = typename tensor_traits<T>::cplx_t;
template <typename T> using cplx_ten_t = typename tensor_traits<T>::cplx_ten_t;
template <typename T> using context_handle_t = typename tensor_traits<T>::context_handle_t;

template <typename E> using List = std::vector<E>;
template <typename A, typename B> using Pair = std::pair<A, B>;
template <typename K, typename V> using Map = std::map<K, V>;

}  // namespace tci

#endif
