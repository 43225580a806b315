// Gauss-Newton normal equations of the sparse tomography Jacobian.
//
// gauss_newton_reconstruct (src/Tomography_inverse.jl:489-494) forms
//   Ja = J[:, active];  normal = Matrix(Ja' * (W .* Ja));  grad = Ja' * (W .* r)
// and laplace_posterior (:922-924) the same product over every column.  The
// reference's sparse product builds column j of the result from the rows of
// column j in ascending order (Gustavson): C[i, j] += J[r, i] * (W[r] J[r, j]).
// The build does the same on the device:
//
//   1. one 64-lane block per row tags each entry with its active column (sentinel
//      n_active for dropped entries: zero, non-finite, inactive column) and
//      records the entry's row;
//   2. a stable radix sort of (active column, entry) pairs (dp_sort.hpp: no
//      memset, so the whole assembly can be captured in a graph) lists every
//      column's entries in ascending row order;
//   3. one 64-lane block per (output column j, tile of 1024 rows i) walks the
//      rows of column j in that order and, for each, adds the row's products
//      into an LDS accumulator (entries of one row have distinct columns, so
//      the lanes never collide); the block then writes its tile of column j
//      of the column-major result, zeros included (row bounds are fetched 64
//      rows at a time and the next row's entries are loaded ahead);
//   4. one lane per active column sums the gradient in the same order
//      (operands of 8 entries loaded ahead of their in-order additions).
//
// Every output is accumulated in the reference's order with separate
// multiply and add (no FP contraction), so results are deterministic and do
// not depend on the launch shape.  The dense result dominates the HBM
// traffic: 8 n_active^2 bytes written.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include "dp_sort.hpp"
#include <cmath>

namespace {

// Rows of the output column per block (8 KB of LDS).  Measured on a 12 869-column
// system: 0.91 ms per assembly with 4096-row tiles, 0.81 ms with 2048, 0.78 ms
// with 1024 (more resident blocks per CU outweigh the repeated column walks).
constexpr int NE_TILE = 1024;

__global__ void ne_fill_index(int n_cols, int32_t* __restrict__ act_index) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < n_cols) act_index[c] = -1;
}

// active == nullptr: every column is active (act_index[c] = c).
__global__ void ne_scatter_index(int n_active, int n_cols, const int32_t* __restrict__ active,
                                 int32_t* __restrict__ act_index) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n_active) return;
  const int c = active ? active[a] : a;
  if (c >= 0 && c < n_cols) act_index[c] = a;
}

__device__ __forceinline__ bool ne_keep(double v) { return isfinite(v) && v != 0.0; }

// One 64-lane block per row: the row's entries are handled by consecutive
// lanes (coalesced reads and writes).
__global__ __launch_bounds__(64) void ne_keys(int n_rows, int n_cols, int n_active,
                                              const int32_t* __restrict__ row_ptr,
                                              const int32_t* __restrict__ col,
                                              const double* __restrict__ val,
                                              const int32_t* __restrict__ act_index,
                                              int32_t* __restrict__ key, int32_t* __restrict__ ent,
                                              int32_t* __restrict__ row_of) {
  const int r = blockIdx.x;
  if (r >= n_rows) return;
  const int e1 = row_ptr[r + 1];
  for (int e = row_ptr[r] + (int)threadIdx.x; e < e1; e += 64) {
    const int c = col[e];
    int a = (c >= 0 && c < n_cols) ? act_index[c] : -1;
    if (a < 0 || !ne_keep(val[e])) a = n_active;
    key[e] = a;
    ent[e] = e;
    row_of[e] = r;
  }
}

// col_ptr[a] = first sorted position whose key is >= a (a = 0 .. n_active).
__global__ void ne_col_ptr(int n_active, int nnz, const int32_t* __restrict__ skey,
                           int32_t* __restrict__ col_ptr) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > n_active) return;
  int lo = 0, hi = nnz;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (skey[mid] < a) lo = mid + 1; else hi = mid;
  }
  col_ptr[a] = lo;
}

__device__ __forceinline__ int lane_of(int x, int k) { return __builtin_amdgcn_readlane(x, k); }
__device__ __forceinline__ double lane_of(double x, int k) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), k),
                          __builtin_amdgcn_readlane(__double2loint(x), k));
}

// One wave per (column j, tile).  The rows of column j are taken 64 at a time:
// each lane fetches one row's bounds and (W .* Ja)[r, j], then the wave walks
// those rows in order, reading each row's bounds from the owning lane.  The
// first 64 entries of the next row are loaded before the current row is added,
// so the global-load latency overlaps the LDS updates.  The block is a single
// wave: its LDS operations complete in issue order, so row k's additions
// precede row k+1's and every entry keeps the ascending-row order.
__global__ __launch_bounds__(64) void ne_normal_kernel(
    int n_active, const int32_t* __restrict__ row_ptr, const double* __restrict__ val,
    const double* __restrict__ weights, const int32_t* __restrict__ key,
    const int32_t* __restrict__ sent, const int32_t* __restrict__ row_of,
    const int32_t* __restrict__ col_ptr, double* __restrict__ normal) {
  __shared__ double acc[NE_TILE];
  const int j = blockIdx.x;
  const int i0 = blockIdx.y * NE_TILE;
  const int tile = min(NE_TILE, n_active - i0);
  const int lane = threadIdx.x;
  for (int q = lane; q < tile; q += 64) acc[q] = 0.0;
  __syncthreads();
  const int s_beg = col_ptr[j], s_end = col_ptr[j + 1];
  for (int s0 = s_beg; s0 < s_end; s0 += 64) {
    const int cnt = min(64, s_end - s0);
    int rb = 0, re = 0;
    double bj = 0.0;
    if (lane < cnt) {
      const int e = sent[s0 + lane];
      const int r = row_of[e];
      rb = row_ptr[r];
      re = row_ptr[r + 1];
      bj = weights ? weights[r] * val[e] : val[e];   // (W .* Ja)[r, j]
    }
    int b = lane_of(rb, 0), en = lane_of(re, 0);
    int kk = -1;
    double vv = 0.0;
    if (b + lane < en) { kk = key[b + lane]; vv = val[b + lane]; }
    for (int k = 0; k < cnt; ++k) {
      const double m = lane_of(bj, k);
      int nb = 0, nen = 0, nk = -1;
      double nv = 0.0;
      if (k + 1 < cnt) {
        nb = lane_of(rb, k + 1);
        nen = lane_of(re, k + 1);
        if (nb + lane < nen) { nk = key[nb + lane]; nv = val[nb + lane]; }
      }
      const int i = kk - i0;
      if (kk >= 0 && i >= 0 && i < tile) acc[i] += vv * m;
      for (int e3 = b + 64 + lane; e3 < en; e3 += 64) {   // rows longer than 64 entries
        const int i3 = key[e3] - i0;
        if (i3 >= 0 && i3 < tile) acc[i3] += val[e3] * m;
      }
      __builtin_amdgcn_wave_barrier();
      b = nb; en = nen; kk = nk; vv = nv;
    }
  }
  __syncthreads();
  double* out = normal + (size_t)j * n_active + i0;
  for (int q = lane; q < tile; q += 64) out[q] = acc[q];
}

__global__ void ne_grad_kernel(int n_active, const double* __restrict__ val,
                               const double* __restrict__ weights,
                               const double* __restrict__ residual,
                               const int32_t* __restrict__ sent, const int32_t* __restrict__ row_of,
                               const int32_t* __restrict__ col_ptr, double* __restrict__ grad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n_active) return;
  const int end = col_ptr[j + 1];
  double g = 0.0;
  int s = col_ptr[j];
  // Operands of 8 entries are loaded before they are added, so the dependent
  // loads of a long column overlap; the additions keep ascending-row order.
  constexpr int B = 8;
  for (; s + B <= end; s += B) {
    int e[B], r[B];
    double v[B], wr[B];
#pragma unroll
    for (int q = 0; q < B; ++q) e[q] = sent[s + q];
#pragma unroll
    for (int q = 0; q < B; ++q) {
      r[q] = row_of[e[q]];
      v[q] = val[e[q]];
    }
#pragma unroll
    for (int q = 0; q < B; ++q) wr[q] = weights ? weights[r[q]] * residual[r[q]] : residual[r[q]];
#pragma unroll
    for (int q = 0; q < B; ++q) g += v[q] * wr[q];
  }
  for (; s < end; ++s) {
    const int e = sent[s];
    const int r = row_of[e];
    const double wr = weights ? weights[r] * residual[r] : residual[r];   // (W .* r)[r]
    g += val[e] * wr;
  }
  grad[j] = g;
}

int check_csr(const char* fn, const dp_csr* J, int32_t n_active) {
  const std::string f(fn);
  if (!J) return dph::fail(DP_E_ARGUMENT, f + ": null Jacobian");
  if (J->n_rows < 0 || J->n_cols < 0 || J->nnz < 0)
    return dph::fail(DP_E_ARGUMENT, f + ": negative dimension");
  if (J->n_rows > 0 && !J->row_ptr) return dph::fail(DP_E_ARGUMENT, f + ": null row_ptr");
  if (J->nnz > 0 && (!J->col || !J->val))
    return dph::fail(DP_E_ARGUMENT, f + ": null column or value array");
  if (n_active < 0 || n_active > J->n_cols)
    return dph::fail(DP_E_ARGUMENT, f + ": n_active must lie in [0, n_cols]");
  return DP_OK;
}

// Device part shared by both entry points; all pointers are device pointers.
int run_normal(dp_ctx c, const dp_csr* J, const double* weights, const double* residual,
               int32_t n_active, const int32_t* active, double* normal, double* grad) {
  hipStream_t st = c->stream;
  const int nr = J->n_rows, nc = J->n_cols, nnz = J->nnz, na = n_active;
  int bits = 1;
  while ((1ll << bits) <= (long long)na) ++bits;   // keys 0 .. n_active
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_ai = take(4 * (size_t)nc), o_k = take(4 * (size_t)nnz),
               o_e = take(4 * (size_t)nnz), o_ro = take(4 * (size_t)nnz),
               o_p0 = take(8 * (size_t)nnz), o_p1 = take(8 * (size_t)nnz),
               o_h = take(4 * dps::sort_hist_ints(nnz)),
               o_cp = take(4 * (size_t)(na + 1));
  int rc = dph::ensure(c->scratch2, o);
  if (rc) return rc;
  char* b = (char*)c->scratch2.ptr;
  int32_t *ai = (int32_t*)(b + o_ai), *k = (int32_t*)(b + o_k), *e = (int32_t*)(b + o_e),
          *ro = (int32_t*)(b + o_ro), *cp = (int32_t*)(b + o_cp);
  const int32_t *sk = k, *se = e;
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(ne_fill_index, dim3((nc + 255) / 256), dim3(256), 0, st, nc, ai);
  hipLaunchKernelGGL(ne_scatter_index, dim3((na + 255) / 256), dim3(256), 0, st, na, nc, active,
                     ai);
  if (nr > 0)
    hipLaunchKernelGGL(ne_keys, dim3(nr), dim3(64), 0, st, nr, nc, na, J->row_ptr, J->col, J->val,
                       ai, k, e, ro);
  DP_HIP(hipGetLastError());
  if (nnz > 0)
    dps::radix_sort_pairs(k, e, nnz, bits, (int32_t*)(b + o_p0), (int32_t*)(b + o_p1),
                          (int32_t*)(b + o_h), st, &sk, &se);
  hipLaunchKernelGGL(ne_col_ptr, dim3((na + 1 + 255) / 256), dim3(256), 0, st, na, nnz, sk, cp);
  if (normal) {
    const dim3 grid(na, (na + NE_TILE - 1) / NE_TILE);
    hipLaunchKernelGGL(ne_normal_kernel, grid, dim3(64), 0, st, na, J->row_ptr, J->val, weights,
                       k, se, ro, cp, normal);
  }
  if (grad)
    hipLaunchKernelGGL(ne_grad_kernel, dim3((na + 255) / 256), dim3(256), 0, st, na, J->val,
                       weights, residual, se, ro, cp, grad);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  return DP_OK;
}

// r = pred - obs and the data cost 0.5 * sum(W r^2) (W = 1 without weights),
// one block: thread t sums rows t, t + 256, ... in ascending order, then a
// fixed tree, so the cost does not depend on timing.
constexpr int RES_THREADS = 256;
__global__ __launch_bounds__(RES_THREADS) void ne_residual_kernel(
    int64_t n, const double* __restrict__ pred, const double* __restrict__ obs,
    const double* __restrict__ w, double* __restrict__ r, double* __restrict__ cost) {
  __shared__ double part[RES_THREADS];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += RES_THREADS) {
    const double d = pred[i] - obs[i];
    if (r) r[i] = d;
    acc += w ? w[i] * (d * d) : d * d;
  }
  part[threadIdx.x] = acc;
  __syncthreads();
  for (int h = RES_THREADS / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && cost) *cost = 0.5 * part[0];
}
}  // namespace

extern "C" {

int dp_normal_equations_dev(dp_ctx c, const dp_csr* J, const double* weights,
                            const double* residual, int32_t n_active, const int32_t* active,
                            double* normal, double* grad) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_normal_equations_dev: null context");
  int rc = check_csr("dp_normal_equations_dev", J, n_active);
  if (rc) return rc;
  if (grad && J->n_rows > 0 && !residual)
    return dph::fail(DP_E_ARGUMENT, "dp_normal_equations_dev: the gradient needs the residual");
  if (n_active > 0 && !active && n_active != J->n_cols)
    return dph::fail(DP_E_ARGUMENT,
                     "dp_normal_equations_dev: active == NULL means every column (n_active = n_cols)");
  if (n_active == 0) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  return run_normal(c, J, weights, residual, n_active, active, normal, grad);
}

int dp_normal_equations(dp_ctx c, const dp_csr* J, const double* weights, const double* residual,
                        int32_t n_active, const int32_t* active, double* normal, double* grad) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: null context");
  int rc = check_csr("dp_normal_equations", J, n_active);
  if (rc) return rc;
  const int nr = J->n_rows, nc = J->n_cols, na = n_active;
  if (nr > 0 && J->row_ptr[0] != 0)
    return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: row_ptr[0] must be 0");
  for (int r = 0; r < nr; ++r)
    if (J->row_ptr[r + 1] < J->row_ptr[r])
      return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: row_ptr must be non-decreasing");
  const int nnz = J->nnz;
  if ((nr > 0 ? J->row_ptr[nr] : 0) != nnz)
    return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: row_ptr[n_rows] must equal nnz");
  for (int e = 0; e < nnz; ++e)
    if (J->col[e] < 0 || J->col[e] >= nc)
      return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: column index out of range");
  for (int r = 0; r < nr; ++r)   // the reference's sparse() sums duplicates; require none
    for (int e = J->row_ptr[r] + 1; e < J->row_ptr[r + 1]; ++e)
      if (J->col[e] <= J->col[e - 1])
        return dph::fail(DP_E_ARGUMENT,
                         "dp_normal_equations: columns of a row must be strictly ascending");
  if (active) {
    for (int a = 0; a < na; ++a)
      if (active[a] < 0 || active[a] >= nc || (a > 0 && active[a] <= active[a - 1]))
        return dph::fail(DP_E_ARGUMENT,
                         "dp_normal_equations: active columns must be ascending, distinct and in range");
  } else if (na != nc) {
    return dph::fail(DP_E_ARGUMENT,
                     "dp_normal_equations: active == NULL means every column (n_active = n_cols)");
  }
  if (grad && nr > 0 && !residual)
    return dph::fail(DP_E_ARGUMENT, "dp_normal_equations: the gradient needs the residual");
  if (na == 0) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_rp = take(4 * (size_t)(nr + 1)), o_col = take(4 * (size_t)nnz),
               o_val = take(8 * (size_t)nnz), o_w = take(8 * (size_t)nr),
               o_r = take(8 * (size_t)nr), o_act = take(4 * (size_t)na),
               o_g = take(8 * (size_t)na), o_n = take(normal ? 8 * (size_t)na * na : 0);
  rc = dph::ensure(c->staging, o);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  if (nr) DP_HIP(hipMemcpyAsync(b + o_rp, J->row_ptr, 4 * (size_t)(nr + 1), hipMemcpyHostToDevice, st));
  if (nnz) {
    DP_HIP(hipMemcpyAsync(b + o_col, J->col, 4 * (size_t)nnz, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_val, J->val, 8 * (size_t)nnz, hipMemcpyHostToDevice, st));
  }
  if (weights && nr) DP_HIP(hipMemcpyAsync(b + o_w, weights, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  if (grad && nr) DP_HIP(hipMemcpyAsync(b + o_r, residual, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  if (active) DP_HIP(hipMemcpyAsync(b + o_act, active, 4 * (size_t)na, hipMemcpyHostToDevice, st));
  dp_csr dj{nr, nc, nnz, (int32_t*)(b + o_rp), (int32_t*)(b + o_col), (double*)(b + o_val)};
  rc = run_normal(c, &dj, weights ? (double*)(b + o_w) : nullptr, (double*)(b + o_r), na,
                  active ? (int32_t*)(b + o_act) : nullptr, normal ? (double*)(b + o_n) : nullptr,
                  grad ? (double*)(b + o_g) : nullptr);
  if (rc) return rc;
  if (normal)
    DP_HIP(hipMemcpyAsync(normal, b + o_n, 8 * (size_t)na * na, hipMemcpyDeviceToHost, st));
  if (grad) DP_HIP(hipMemcpyAsync(grad, b + o_g, 8 * (size_t)na, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

int dp_residual_dev(dp_ctx c, int64_t n, const double* pred, const double* obs,
                    const double* weights, double* residual, double* cost) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_residual_dev: null context");
  if (n < 0 || (n > 0 && (!pred || !obs)))
    return dph::fail(DP_E_ARGUMENT, "dp_residual_dev: n must be >= 0 with pred and obs given");
  if (!residual && !cost) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  hipLaunchKernelGGL(ne_residual_kernel, dim3(1), dim3(RES_THREADS), 0, c->stream, n, pred, obs,
                     weights, residual, cost);
  DP_HIP(hipGetLastError());
  return DP_OK;
}

}  // extern "C"
