// Front-end pieces of the reference, compiled from its own sources (they need
// no tensor engine): lattice bond lists (include/tnbp/lattice/square.h:14-34,
// oned.h:14-36), GreedyEdgeLayering and FindShortestPath
// (include/tnbp/framework/graph.h:57-170), the OpenQASM 2/3 parser
// (include/qasm/any.h, v2/parser.h, v3/parser.h) and the sparse-Pauli reader
// (include/pauli/sparse_pauli.h, pauli_string.h).  Each command prints JSON;
// tests compare this build's restatements against it.
//
// TEST INFRASTRUCTURE ONLY (the oracle's reference build): used by tests/ to
// pin the front end, never by the product.
//
//   tnbp_ref square LX LY | oned LX PX         bond list
//   tnbp_ref layers  < edges                    "u v" per line on stdin
//   tnbp_ref path SRC DST < edges
//   tnbp_ref qasm FILE                          parsed program
//   tnbp_ref pauli FILE                         sparse-Pauli terms
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "tnbp/framework/graph.h"
#include "tnbp/lattice/oned.h"
#include "tnbp/lattice/square.h"
#include "qasm/any.h"
#include "pauli/sparse_pauli.h"
#include "pauli/pauli_string.h"  // uses sparse_pauli.h's trim_inplace

namespace {

void print_edges(const std::vector<std::pair<int, int>> &e) {
  std::cout << "[";
  for (size_t i = 0; i < e.size(); ++i)
    std::cout << (i ? "," : "") << "[" << e[i].first << "," << e[i].second << "]";
  std::cout << "]";
}

std::vector<std::pair<int, int>> read_edges() {
  std::vector<std::pair<int, int>> e;
  int u, v;
  while (std::cin >> u >> v) e.emplace_back(u, v);
  return e;
}

std::string escape(const std::string &s) {
  std::string o;
  for (char c : s) {
    if (c == '"' || c == '\\') o += '\\';
    if (c == '\n') {
      o += "\\n";
      continue;
    }
    o += c;
  }
  return o;
}

}  // namespace

int main(int argc, char **argv) {
  std::cout.precision(17);
  if (argc < 2) return 2;
  const std::string cmd = argv[1];
  if (cmd == "square" && argc == 4) {
    print_edges(tnbp::bond_square_lattice(std::stoi(argv[2]), std::stoi(argv[3])));
  } else if (cmd == "oned" && argc == 4) {
    print_edges(tnbp::bond_oned_lattice(std::stoi(argv[2]), std::stoi(argv[3])));
  } else if (cmd == "layers") {
    const auto layers = tnbp::GreedyEdgeLayering<int>(read_edges());
    std::cout << "[";
    for (size_t l = 0; l < layers.size(); ++l) {
      std::cout << (l ? "," : "");
      print_edges(layers[l]);
    }
    std::cout << "]";
  } else if (cmd == "path" && argc == 4) {
    const auto p = tnbp::FindShortestPath<int>(read_edges(), std::stoi(argv[2]), std::stoi(argv[3]));
    std::cout << "[";
    for (size_t i = 0; i < p.size(); ++i) std::cout << (i ? "," : "") << p[i];
    std::cout << "]";
  } else if (cmd == "qasm" && argc == 3) {
    std::ifstream in(argv[2]);
    std::stringstream ss;
    ss << in.rdbuf();
    const std::string src = ss.str();
    try {
      const qasm::Program prog = qasm::parse_any(src);
      std::cout << "{\"qregs\":[";
      for (size_t i = 0; i < prog.qregs.size(); ++i)
        std::cout << (i ? "," : "") << "[\"" << escape(prog.qregs[i].name) << "\","
                  << prog.qregs[i].size << "]";
      std::cout << "],\"instructions\":[";
      for (size_t i = 0; i < prog.instructions.size(); ++i) {
        const auto &ins = prog.instructions[i];
        std::cout << (i ? "," : "") << "{\"op\":" << static_cast<int>(ins.op) << ",\"name\":\""
                  << escape(ins.name) << "\",\"params\":[";
        for (size_t k = 0; k < ins.params.size(); ++k) std::cout << (k ? "," : "") << ins.params[k];
        std::cout << "],\"qubits\":[";
        for (size_t k = 0; k < ins.qubits.size(); ++k)
          std::cout << (k ? "," : "") << "[\"" << escape(ins.qubits[k].reg) << "\","
                    << ins.qubits[k].index << "]";
        std::cout << "]}";
      }
      std::cout << "]}";
    } catch (const std::exception &e) {
      std::cout << "{\"error\":\"" << escape(e.what()) << "\"}";
    }
  } else if (cmd == "pauli" && argc == 3) {
    try {
      const auto terms = pauli::load_sparse_pauli_op<std::complex<double>>(argv[2]);
      std::cout << "[";
      for (size_t i = 0; i < terms.size(); ++i) {
        const auto ops = pauli::string_to_non_identity_localops(terms[i].pauli_string);
        std::cout << (i ? "," : "") << "{\"coeff\":[" << terms[i].coeff.real() << ","
                  << terms[i].coeff.imag() << "],\"label\":\"" << escape(terms[i].pauli_string)
                  << "\",\"ops\":[";
        for (size_t k = 0; k < ops.size(); ++k)
          std::cout << (k ? "," : "") << "[" << ops[k].qubit << ",\"" << ops[k].op << "\"]";
        std::cout << "]}";
      }
      std::cout << "]";
    } catch (const std::exception &e) {
      std::cout << "{\"error\":\"" << escape(e.what()) << "\"}";
    }
  } else {
    return 2;
  }
  std::cout << "\n";
  return 0;
}
