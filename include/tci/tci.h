/*
 * tci/tci.h - Tensor Computing Interface backed by the MI355X hbp library.
 *
 * Drop-in for min-tci (external/min-tci/include/tci/tci.h): put this
 * directory's parent first on the include path and link libhbp.so; the tnbp
 * headers and the kicked_ising_lattice / estimator apps compile unchanged.
 * Tensors live in device memory; elementwise host callbacks (for_each,
 * assign_from_range, to_range) copy through host memory.
 */
#ifndef TCI_TCI_H
#define TCI_TCI_H

// standard headers min-tci's headers bring in, which the tnbp headers use
// without including them (e.g. assert in include/tnbp/lattice/honeycomb.h)
#include <algorithm>
#include <cassert>
#include <cctype>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <string_view>
#include <thread>
#include <unordered_map>
#include <unordered_set>

#include "tci/traits.h"
#include "tci/hbp/tensor.h"
#include "tci/hbp/functions.h"
// tnbp's algorithm functions on the fused PEPS layer (overloads; off with
// HBP_TCI_NO_FUSED or without mpi.h)
#include "tci/hbp/fused.h"

#endif
