// CPU oracle of the Gauss-Newton normal equations.  TEST INFRASTRUCTURE ONLY.
//
// gauss_newton_reconstruct (src/Tomography_inverse.jl:489-494, final
// gradient :595) and laplace_posterior (:922-924) evaluate, with Julia's
// SparseArrays,
//   Ja = J[:, active]; normal = Matrix(transpose(Ja) * (W .* Ja)); grad = transpose(Ja) * (W .* r)
// The sparse-sparse product with a transposed left factor runs on copy(Ja')
// (Gustavson): column j of the result is built from the entries of column j
// of B = W .* Ja in ascending row order r, each adding Ja'[i, r] * B[r, j] =
// Ja[r, i] * (W[r] * Ja[r, j]) to row i.  The transposed sparse-dense product
// sums, per column j of Ja, Ja[r, j] * (W[r] * r[r]) in ascending row order,
// starting from zero.  J reaches the solver from assemble_forward_and_jacobian
// (src/Tomography.jl:806-818), which keeps only finite non-zero entries.
#pragma once
#include <cmath>
#include <utility>
#include <vector>

namespace orc {

// J in CSR form (rows ascending, columns ascending within a row); active[n_act]
// ascending column ids, or nullptr for every column.  normal (column-major
// n_act x n_act) and grad may be nullptr.
inline void normal_equations(int n_rows, int n_cols, const int* rp, const int* col,
                             const double* val, const double* W, const double* res, int n_act,
                             const int* active, double* normal, double* grad) {
  std::vector<int> pos(n_cols, -1);
  for (int a = 0; a < n_act; ++a) pos[active ? active[a] : a] = a;
  std::vector<std::vector<std::pair<int, double>>> by_col(n_act), by_row(n_rows);
  for (int r = 0; r < n_rows; ++r)
    for (int e = rp[r]; e < rp[r + 1]; ++e) {
      const double v = val[e];
      const int a = pos[col[e]];
      if (a < 0 || !std::isfinite(v) || v == 0.0) continue;
      by_col[a].push_back({r, v});   // rows ascending: r runs in order
      by_row[r].push_back({a, v});
    }
  // columns are independent (each keeps its row order): spread over threads
  if (normal) {
#pragma omp parallel
    {
      std::vector<double> x(n_act);
#pragma omp for schedule(dynamic, 8)
      for (int j = 0; j < n_act; ++j) {
        for (int i = 0; i < n_act; ++i) x[i] = 0.0;
        for (const auto& rv : by_col[j]) {
          const int r = rv.first;
          const double nz_b = W ? W[r] * rv.second : rv.second;   // (W .* Ja)[r, j]
          for (const auto& iv : by_row[r]) x[iv.first] += iv.second * nz_b;
        }
        for (int i = 0; i < n_act; ++i) normal[(size_t)j * n_act + i] = x[i];
      }
    }
  }
  if (grad) {
#pragma omp parallel for schedule(static)
    for (int j = 0; j < n_act; ++j) {
      double g = 0.0;
      for (const auto& rv : by_col[j]) {
        const int r = rv.first;
        g += rv.second * (W ? W[r] * res[r] : res[r]);
      }
      grad[j] = g;
    }
  }
}

}  // namespace orc
