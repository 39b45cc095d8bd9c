// Binary CSR / right-hand-side files in the reference's on-disk layout
// (IO/cpp_read.hpp:331-442): int32 n, m, nz; int32 r[n+1]; int32 c[nz]; f64 v[nz]
// and for the rhs int32 n; f64 b[n]. The rhs name "ones"/"unity"/... builds b = A*1.
#pragma once
#include <cstdio>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include "deco.hpp"

namespace IO {

template <typename R, typename C, typename V>
void cpp_write_binary_matrix(const std::string& filename, const std::shared_ptr<CSR<R, C, V>>& A) {
  std::ofstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("Could not open file " + filename + " for writing.");
  f.write(reinterpret_cast<const char*>(&A->n), sizeof(C));
  f.write(reinterpret_cast<const char*>(&A->m), sizeof(C));
  f.write(reinterpret_cast<const char*>(&A->nz), sizeof(R));
  f.write(reinterpret_cast<const char*>(A->r), sizeof(R) * (A->n + 1));
  f.write(reinterpret_cast<const char*>(A->c), sizeof(C) * A->nz);
  f.write(reinterpret_cast<const char*>(A->v), sizeof(V) * A->nz);
}

template <typename R, typename C, typename V>
std::shared_ptr<CSR<R, C, V>> cpp_read_binary_matrix(const std::string& filename) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("Could not open file " + filename);
  C n = -1, m = -1;
  R nz = -1;
  f.read(reinterpret_cast<char*>(&n), sizeof(C));
  f.read(reinterpret_cast<char*>(&m), sizeof(C));
  f.read(reinterpret_cast<char*>(&nz), sizeof(R));
  if (!f || n < 0 || m < 0 || nz < 0) throw std::runtime_error("bad matrix header in " + filename);
  auto A = std::make_shared<CSR<R, C, V>>(n, m, nz);
  if (!A->r) A->r = (R*)std::malloc(sizeof(R) * (n + 1));
  f.read(reinterpret_cast<char*>(A->r), sizeof(R) * (n + 1));
  f.read(reinterpret_cast<char*>(A->c), sizeof(C) * nz);
  f.read(reinterpret_cast<char*>(A->v), sizeof(V) * nz);
  if (!f) throw std::runtime_error("truncated matrix file " + filename);
  return A;
}

template <typename C, typename V>
void cpp_write_binary_array(const std::string& filename, const C n, const V* v) {
  std::ofstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("Could not open file " + filename + " for writing.");
  f.write(reinterpret_cast<const char*>(&n), sizeof(C));
  f.write(reinterpret_cast<const char*>(v), sizeof(V) * n);
}

template <typename R = int, typename C = int, typename V = double>
V* cpp_read_binary_rhs(const std::string& filename, std::shared_ptr<CSR<R, C, V>> A) {
  const C n = A->n;
  V* v = (V*)std::malloc(sizeof(V) * n);
  if (filename == "unity" || filename == "unitary" || filename == "uno" || filename == "one" ||
      filename == "ones" || filename == "zero") {
    V* x = (V*)std::malloc(sizeof(V) * A->m);
    for (C i = 0; i < A->m; ++i) x[i] = (filename == "zero") ? 0.0 : 1.0;
    A->apply(x, v);
    std::free(x);
    return v;
  }
  std::ifstream f(filename, std::ios::binary);
  if (!f) {
    std::free(v);
    throw std::runtime_error("Could not open file " + filename);
  }
  C nn = -1;
  f.read(reinterpret_cast<char*>(&nn), sizeof(C));
  if (nn != n) {
    std::free(v);
    throw std::runtime_error("rhs length does not match the matrix in " + filename);
  }
  f.read(reinterpret_cast<char*>(v), sizeof(V) * n);
  if (!f) {
    std::free(v);
    throw std::runtime_error("truncated rhs file " + filename);
  }
  return v;
}

}  // namespace IO
