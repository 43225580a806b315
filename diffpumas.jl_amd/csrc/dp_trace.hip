// Ray -> voxel traversal (one ray per lane), bit-exact with the reference
// tracers.  Built with -ffp-contract=off: every product and sum is rounded as
// in the reference, IEEE division and floor are exact on gfx950.
//
// Two launches of the same kernel: the first writes per-ray segment counts,
// the second (after the caller's exclusive scan into row_ptr) writes the CSR
// segment arrays.  The write stream (12 B per ray·voxel) bounds the kernel.
#include "dp_internal.hpp"
#include "dp_device.hpp"

namespace {

struct CubeDev {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int nx, ny, nz;
  double dx, dy, dz;
};

// get_cell_indices, examples/diff_flux_cube.jl:132-144 (1-based)
__device__ __forceinline__ bool cube_cell(const CubeDev& g, double x, double y, double z, int& i,
                                          int& j, int& k) {
  if (x < g.x_min || x > g.x_max || y < g.y_min || y > g.y_max || z < g.z_min || z > g.z_max)
    return false;
  i = min(max((int)floor((x - g.x_min) / g.dx) + 1, 1), g.nx);
  j = min(max((int)floor((y - g.y_min) / g.dy) + 1, 1), g.ny);
  k = min(max((int)floor((z - g.z_min) / g.dz) + 1, 1), g.nz);
  return true;
}

// trace_cube_path, examples/diff_flux_cube.jl:304-364, with
// distance_to_boundary (:194-280) inlined.  The cell found at the top of an
// iteration is the cell distance_to_boundary looks up again at the same
// point, so it is computed once.
__global__ __launch_bounds__(64) void trace_cube_kernel(CubeDev g, int n_rays,
                                                         const double* __restrict__ start_xy,
                                                         const double* __restrict__ dir,
                                                         int32_t* __restrict__ counts,
                                                         const int32_t* __restrict__ row_ptr,
                                                         int32_t* __restrict__ cells,
                                                         double* __restrict__ lens,
                                                         double* __restrict__ z_exit_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  const double ux = dir[3 * r], uy = dir[3 * r + 1], uz = dir[3 * r + 2];
  double px = start_xy[2 * r], py = start_xy[2 * r + 1], pz = 0.0, z_exit = 0.0;
  const bool write = cells != nullptr;
  const int out = write ? row_ptr[r] : 0;
  const int cap = write ? row_ptr[r + 1] - out : 0;
  int n = 0;
  const double eps = 1e-10;
  for (int step = 0; step < 1000; ++step) {
    int i, j, k;
    if (!cube_cell(g, px, py, pz, i, j, k)) break;
    const int cell = (i - 1) + (j - 1) * g.nx + (k - 1) * g.nx * g.ny;
    const double x_lo = g.x_min + (i - 1) * g.dx, x_hi = g.x_min + i * g.dx;
    const double y_lo = g.y_min + (j - 1) * g.dy, y_hi = g.y_min + j * g.dy;
    const double z_lo = g.z_min + (k - 1) * g.dz, z_hi = g.z_min + k * g.dz;
    double t_min = __builtin_huge_val();
    bool exiting = false;
    if (fabs(ux) > eps) {
      if (ux > 0) { double t = (x_hi - px) / ux; if (t > eps && t < t_min) { t_min = t; exiting = (i + 1 > g.nx); } }
      else        { double t = (x_lo - px) / ux; if (t > eps && t < t_min) { t_min = t; exiting = (i - 1 < 1); } }
    }
    if (fabs(uy) > eps) {
      if (uy > 0) { double t = (y_hi - py) / uy; if (t > eps && t < t_min) { t_min = t; exiting = (j + 1 > g.ny); } }
      else        { double t = (y_lo - py) / uy; if (t > eps && t < t_min) { t_min = t; exiting = (j - 1 < 1); } }
    }
    if (fabs(uz) > eps) {
      if (uz > 0) { double t = (z_hi - pz) / uz; if (t > eps && t < t_min) { t_min = t; exiting = (k + 1 > g.nz); } }
      else        { double t = (z_lo - pz) / uz; if (t > eps && t < t_min) { t_min = t; exiting = (k - 1 < 1); } }
    }
    const bool exits = exiting || t_min == __builtin_huge_val();
    const double dist = t_min;
    if (dist <= 1e-10 || !isfinite(dist)) {
      px += ux * 0.01; py += uy * 0.01; pz += uz * 0.01;
      continue;
    }
    if (write && n < cap) {
      cells[out + n] = cell;
      lens[out + n] = dist;
    }
    ++n;
    px += ux * dist; py += uy * dist; pz += uz * dist;
    z_exit = pz;
    if (exits) break;
  }
  if (!write) {
    counts[r] = n;
    if (z_exit_out) z_exit_out[r] = z_exit;
  }
}

// searchsortedlast on an edge vector: count of edges <= x.
__device__ __forceinline__ int ssl(const double* __restrict__ e, int n, double x) {
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (e[mid] <= x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// trace_path(::VoxelVolume), examples/lvd_tomography.jl:738-779, with
// voxel_cell_indices (:583-594) and distance_to_voxel_boundary (:596-627).
__global__ __launch_bounds__(64) void trace_edges_kernel(
    const double* __restrict__ xe, int nxe, const double* __restrict__ ye, int nye,
    const double* __restrict__ ze, int nze, int n_rays, const double* __restrict__ origin,
    const double* __restrict__ dir, const double* __restrict__ surface,
    int32_t* __restrict__ counts, const int32_t* __restrict__ row_ptr, int32_t* __restrict__ cells,
    double* __restrict__ lens, double* __restrict__ traveled_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n_rays) return;
  const double POINT_EPS = 1e-4;
  const double ux = dir[3 * r], uy = dir[3 * r + 1], uz = dir[3 * r + 2];
  double x = origin[3 * r] + ux * POINT_EPS;
  double y = origin[3 * r + 1] + uy * POINT_EPS;
  double z = origin[3 * r + 2] + uz * POINT_EPS;
  const double sd = surface[r];
  double traveled = POINT_EPS;
  const int nx = nxe - 1, ny = nye - 1;
  const bool write = cells != nullptr;
  const int out = write ? row_ptr[r] : 0;
  const int cap = write ? row_ptr[r + 1] - out : 0;
  int n = 0;
  const double e = 1e-10;
  while (traveled < sd - POINT_EPS) {
    if (x < xe[0] || x > xe[nxe - 1] || y < ye[0] || y > ye[nye - 1] || z < ze[0] || z > ze[nze - 1])
      break;
    const int i = min(max(ssl(xe, nxe, x), 1), nxe - 1);
    const int j = min(max(ssl(ye, nye, y), 1), nye - 1);
    const int k = min(max(ssl(ze, nze, z), 1), nze - 1);
    const int cell = (i - 1) + (j - 1) * nx + (k - 1) * nx * ny;
    double t_min = __builtin_huge_val();
    if (fabs(ux) > e) { double t = ux > 0 ? (xe[i] - x) / ux : (xe[i - 1] - x) / ux; if (t > e) t_min = fmin(t_min, t); }
    if (fabs(uy) > e) { double t = uy > 0 ? (ye[j] - y) / uy : (ye[j - 1] - y) / uy; if (t > e) t_min = fmin(t_min, t); }
    if (fabs(uz) > e) { double t = uz > 0 ? (ze[k] - z) / uz : (ze[k - 1] - z) / uz; if (t > e) t_min = fmin(t_min, t); }
    const double dist = isfinite(t_min) ? t_min : 0.0;
    if (!isfinite(dist) || dist <= 0.0) break;
    const double remaining = sd - traveled;
    const double step = fmin(dist, remaining);
    if (step > 1e-8) {
      if (write && n < cap) {
        cells[out + n] = cell;
        lens[out + n] = step;
      }
      ++n;
      traveled += step;
    }
    if (dist >= remaining - 1e-6) break;
    x += ux * (step + POINT_EPS);
    y += uy * (step + POINT_EPS);
    z += uz * (step + POINT_EPS);
  }
  if (!write) {
    counts[r] = n;
    if (traveled_out) traveled_out[r] = traveled;
  }
}

int launch_cube(dp_ctx c, const dp_cube_grid* grid, int32_t n, const double* sxy, const double* dir,
                int32_t* counts, const int32_t* row_ptr, int32_t* cells, double* lens,
                double* z_exit) {
  CubeDev g{grid->x_min, grid->x_max, grid->y_min, grid->y_max, grid->z_min, grid->z_max,
            grid->nx, grid->ny, grid->nz, grid->dx, grid->dy, grid->dz};
  if (n == 0) return DP_OK;
  // 64-lane blocks: 16 384 rays still give one wave to every CU
  dim3 b(64), gr((n + 63) / 64);
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(trace_cube_kernel, gr, b, 0, c->stream, g, n, sxy, dir, counts, row_ptr,
                     cells, lens, z_exit);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  return DP_OK;
}

int check_cube(const dp_cube_grid* g, int32_t n) {
  if (!g) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: null grid");
  if (n < 0) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: negative ray count");
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0 || !(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0))
    return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: grid needs positive cell counts and sizes");
  if ((int64_t)g->nx * g->ny * g->nz > INT32_MAX)
    return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: more than 2^31-1 cells");
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_trace_cube_dev(dp_ctx c, const dp_cube_grid* grid, int32_t n, const double* sxy,
                      const double* dir, int32_t* counts, const int32_t* row_ptr, int32_t* cells,
                      double* lens, double* z_exit) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube_dev: null context");
  int rc = check_cube(grid, n);
  if (rc) return rc;
  if (!cells && !counts) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube_dev: counts required in the counting pass");
  if (cells && (!row_ptr || !lens)) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube_dev: row_ptr and lengths required");
  DP_HIP(hipSetDevice(c->device));
  return launch_cube(c, grid, n, sxy, dir, counts, row_ptr, cells, lens, z_exit);
}

int dp_trace_cube(dp_ctx c, const dp_cube_grid* grid, int32_t n, const double* sxy,
                  const double* dir, int32_t* counts, const int32_t* row_ptr, int32_t* cells,
                  double* lens, double* z_exit) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: null context");
  int rc = check_cube(grid, n);
  if (rc) return rc;
  if (!cells && !counts) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: counts required in the counting pass");
  if (cells && (!row_ptr || !lens)) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: row_ptr and lengths required");
  if (n == 0) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  const int64_t nseg = cells ? (int64_t)row_ptr[n] : 0;
  if (cells && nseg < 0) return dph::fail(DP_E_ARGUMENT, "dp_trace_cube: negative row_ptr[n]");
  // staging layout
  size_t o_sxy = 0, o_dir = dph::align_up(o_sxy + 16 * (size_t)n, 256);
  size_t o_cnt = dph::align_up(o_dir + 24 * (size_t)n, 256);
  size_t o_rp = dph::align_up(o_cnt + 4 * (size_t)n, 256);
  size_t o_ze = dph::align_up(o_rp + 4 * (size_t)(n + 1), 256);
  size_t o_cell = dph::align_up(o_ze + 8 * (size_t)n, 256);
  size_t o_len = dph::align_up(o_cell + 4 * (size_t)nseg, 256);
  size_t total = o_len + 8 * (size_t)nseg + 256;
  rc = dph::ensure(c->staging, total);
  if (rc) return rc;
  char* base = (char*)c->staging.ptr;
  DP_HIP(hipMemcpyAsync(base + o_sxy, sxy, 16 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(base + o_dir, dir, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  if (cells) DP_HIP(hipMemcpyAsync(base + o_rp, row_ptr, 4 * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
  rc = launch_cube(c, grid, n, (double*)(base + o_sxy), (double*)(base + o_dir),
                   cells ? nullptr : (int32_t*)(base + o_cnt), (int32_t*)(base + o_rp),
                   cells ? (int32_t*)(base + o_cell) : nullptr, (double*)(base + o_len),
                   (double*)(base + o_ze));
  if (rc) return rc;
  if (!cells) {
    DP_HIP(hipMemcpyAsync(counts, base + o_cnt, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (z_exit) DP_HIP(hipMemcpyAsync(z_exit, base + o_ze, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  } else if (nseg > 0) {
    DP_HIP(hipMemcpyAsync(cells, base + o_cell, 4 * (size_t)nseg, hipMemcpyDeviceToHost, c->stream));
    DP_HIP(hipMemcpyAsync(lens, base + o_len, 8 * (size_t)nseg, hipMemcpyDeviceToHost, c->stream));
  }
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

int dp_trace_edges(dp_ctx c, const double* xe, int32_t nxe, const double* ye, int32_t nye,
                   const double* ze, int32_t nze, int32_t n, const double* origin,
                   const double* dir, const double* surface, int32_t* counts,
                   const int32_t* row_ptr, int32_t* cells, double* lens, double* traveled) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_trace_edges: null context");
  if (nxe < 2 || nye < 2 || nze < 2 || !xe || !ye || !ze)
    return dph::fail(DP_E_ARGUMENT, "dp_trace_edges: each axis needs >= 2 edges");
  if (n < 0) return dph::fail(DP_E_ARGUMENT, "dp_trace_edges: negative ray count");
  if (!cells && !counts) return dph::fail(DP_E_ARGUMENT, "dp_trace_edges: counts required in the counting pass");
  if (cells && (!row_ptr || !lens)) return dph::fail(DP_E_ARGUMENT, "dp_trace_edges: row_ptr and lengths required");
  if (n == 0) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  const int64_t nseg = cells ? (int64_t)row_ptr[n] : 0;
  size_t o_x = 0, o_y = dph::align_up(8 * (size_t)nxe, 256);
  size_t o_z = dph::align_up(o_y + 8 * (size_t)nye, 256);
  size_t o_or = dph::align_up(o_z + 8 * (size_t)nze, 256);
  size_t o_dir = dph::align_up(o_or + 24 * (size_t)n, 256);
  size_t o_sd = dph::align_up(o_dir + 24 * (size_t)n, 256);
  size_t o_cnt = dph::align_up(o_sd + 8 * (size_t)n, 256);
  size_t o_rp = dph::align_up(o_cnt + 4 * (size_t)n, 256);
  size_t o_tr = dph::align_up(o_rp + 4 * (size_t)(n + 1), 256);
  size_t o_cell = dph::align_up(o_tr + 8 * (size_t)n, 256);
  size_t o_len = dph::align_up(o_cell + 4 * (size_t)nseg, 256);
  size_t total = o_len + 8 * (size_t)nseg + 256;
  int rc = dph::ensure(c->staging, total);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  DP_HIP(hipMemcpyAsync(b + o_x, xe, 8 * (size_t)nxe, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(b + o_y, ye, 8 * (size_t)nye, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(b + o_z, ze, 8 * (size_t)nze, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(b + o_or, origin, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(b + o_dir, dir, 24 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  DP_HIP(hipMemcpyAsync(b + o_sd, surface, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  if (cells) DP_HIP(hipMemcpyAsync(b + o_rp, row_ptr, 4 * (size_t)(n + 1), hipMemcpyHostToDevice, c->stream));
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(trace_edges_kernel, dim3((n + 63) / 64), dim3(64), 0, c->stream,
                     (double*)(b + o_x), nxe, (double*)(b + o_y), nye, (double*)(b + o_z), nze, n,
                     (double*)(b + o_or), (double*)(b + o_dir), (double*)(b + o_sd),
                     cells ? nullptr : (int32_t*)(b + o_cnt), (int32_t*)(b + o_rp),
                     cells ? (int32_t*)(b + o_cell) : nullptr, (double*)(b + o_len),
                     (double*)(b + o_tr));
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  if (!cells) {
    DP_HIP(hipMemcpyAsync(counts, b + o_cnt, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    if (traveled) DP_HIP(hipMemcpyAsync(traveled, b + o_tr, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  } else if (nseg > 0) {
    DP_HIP(hipMemcpyAsync(cells, b + o_cell, 4 * (size_t)nseg, hipMemcpyDeviceToHost, c->stream));
    DP_HIP(hipMemcpyAsync(lens, b + o_len, 8 * (size_t)nseg, hipMemcpyDeviceToHost, c->stream));
  }
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

}  // extern "C"
