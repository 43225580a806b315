// Ray -> voxel traversal, bit-exact with the reference tracers.  Built with
// -ffp-contract=off: every product and sum is rounded as in the reference,
// IEEE division and floor are exact on gfx950.
//
// Two launches of the same kernel: the first writes per-ray segment counts,
// the second (after the caller's exclusive scan into row_ptr) writes the CSR
// segment arrays (12 B per ray-voxel).
//
// Each ray is a quad of lanes: lanes 0, 1, 2 own the x, y and z axis (the
// coordinate, the direction component, the cell index along the axis and the
// distance to the axis' next face); lane 3 shadows lane 0.  A step's per-axis
// work (one division for the cell index, one for the face distance, or one
// edge search) runs on the three lanes at once, and DPP quad broadcasts hand
// the three indices and distances to every lane of the quad, which then picks
// the nearest face in the reference's x, y, z order.  A wave holds 16 rays, so
// 16 384 rays fill 1024 waves (one per SIMD); with one ray per lane they were
// 256 waves, one per CU, each issuing all six divisions of a step: 0.20 ms for
// the fill pass of 16 384 rays through 128 x 128 x 64 voxels.  Every loop runs
// over the whole wave (a finished ray's quad idles), so each broadcast reads
// active lanes.
#include "dp_internal.hpp"
#include "dp_device.hpp"

namespace {

constexpr int TRACE_RAYS_PER_WAVE = 16;

using dpk::quad_bcast;

struct CubeDev {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int nx, ny, nz;
  double dx, dy, dz;
};

// trace_cube_path, examples/diff_flux_cube.jl:304-364, with
// distance_to_boundary (:194-280) and get_cell_indices (:132-144) inlined.
// The cell found at the top of an iteration is the cell
// distance_to_boundary looks up again at the same point, so it is computed
// once.  Per axis (lane): in-range test, 1-based index
// clamp(floor((p - min) / d) + 1, 1, n), the face the direction points to and
// its distance; the quad then takes the smallest distance > eps in x, y, z
// order with strict comparisons, as the reference's three if-blocks do.
__global__ __launch_bounds__(64) void trace_cube_kernel(CubeDev g, int n_rays,
                                                         const double* __restrict__ start_xy,
                                                         const double* __restrict__ dir,
                                                         int32_t* __restrict__ counts,
                                                         const int32_t* __restrict__ row_ptr,
                                                         int32_t* __restrict__ cells,
                                                         double* __restrict__ lens,
                                                         double* __restrict__ z_exit_out) {
  const int lane = (int)threadIdx.x, q = lane & 3;
  const int ax = q < 3 ? q : 0;
  const int r = blockIdx.x * TRACE_RAYS_PER_WAVE + (lane >> 2);
  const bool live = r < n_rays;
  const double lo = ax == 0 ? g.x_min : ax == 1 ? g.y_min : g.z_min;
  const double hi = ax == 0 ? g.x_max : ax == 1 ? g.y_max : g.z_max;
  const double d = ax == 0 ? g.dx : ax == 1 ? g.dy : g.dz;
  const int na = ax == 0 ? g.nx : ax == 1 ? g.ny : g.nz;
  const double u = live ? dir[3 * r + ax] : 0.0;
  double p = live && ax < 2 ? start_xy[2 * r + ax] : 0.0;
  double p_seg = 0.0;  // the coordinate after the last segment (z: the exit height)
  const bool write = cells != nullptr;
  const int out = write && live ? row_ptr[r] : 0;
  const int cap = write && live ? row_ptr[r + 1] - out : 0;
  int n = 0;
  const double eps = 1e-10;
  bool done = !live;
  for (int step = 0; step < 1000 && __ballot(!done) != 0; ++step) {
    const bool inb = !(p < lo || p > hi);
    const int i = min(max((int)floor((p - lo) / d) + 1, 1), na);
    double t = __builtin_huge_val();
    bool ex = false;
    if (fabs(u) > eps) {
      if (u > 0) { t = (lo + i * d - p) / u; ex = i + 1 > na; }
      else       { t = (lo + (i - 1) * d - p) / u; ex = i - 1 < 1; }
    }
    const int code = (inb ? 1 : 0) | (ex ? 2 : 0);
    const int c0 = quad_bcast<0>(code), c1 = quad_bcast<1>(code), c2 = quad_bcast<2>(code);
    const int i0 = quad_bcast<0>(i), i1 = quad_bcast<1>(i), i2 = quad_bcast<2>(i);
    const double t0 = quad_bcast<0>(t), t1 = quad_bcast<1>(t), t2 = quad_bcast<2>(t);
    if (done) continue;
    if (!(c0 & c1 & c2 & 1)) { done = true; continue; }
    const int cell = (i0 - 1) + (i1 - 1) * g.nx + (i2 - 1) * g.nx * g.ny;
    double t_min = __builtin_huge_val();
    bool exiting = false;
    if (t0 > eps && t0 < t_min) { t_min = t0; exiting = c0 & 2; }
    if (t1 > eps && t1 < t_min) { t_min = t1; exiting = c1 & 2; }
    if (t2 > eps && t2 < t_min) { t_min = t2; exiting = c2 & 2; }
    const bool exits = exiting || t_min == __builtin_huge_val();
    const double dist = t_min;
    if (dist <= 1e-10 || !isfinite(dist)) {
      p += u * 0.01;
      continue;
    }
    if (write && n < cap && q == 0) {
      cells[out + n] = cell;
      lens[out + n] = dist;
    }
    ++n;
    p += u * dist;
    p_seg = p;
    if (exits) done = true;
  }
  if (!write && live) {
    if (q == 0) counts[r] = n;
    if (q == 2 && z_exit_out) z_exit_out[r] = p_seg;
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
// Per axis (lane): in-range test, the edge search, the face distance; the
// quad takes the smallest distance > e (fmin, order-free) and every lane
// keeps the same copy of the travelled length.
__global__ __launch_bounds__(64) void trace_edges_kernel(
    const double* __restrict__ xe, int nxe, const double* __restrict__ ye, int nye,
    const double* __restrict__ ze, int nze, int n_rays, const double* __restrict__ origin,
    const double* __restrict__ dir, const double* __restrict__ surface,
    int32_t* __restrict__ counts, const int32_t* __restrict__ row_ptr, int32_t* __restrict__ cells,
    double* __restrict__ lens, double* __restrict__ traveled_out) {
  const int lane = (int)threadIdx.x, q = lane & 3;
  const int ax = q < 3 ? q : 0;
  const int r = blockIdx.x * TRACE_RAYS_PER_WAVE + (lane >> 2);
  const bool live = r < n_rays;
  const double POINT_EPS = 1e-4;
  const double* E = ax == 0 ? xe : ax == 1 ? ye : ze;
  const int ne = ax == 0 ? nxe : ax == 1 ? nye : nze;
  const double u = live ? dir[3 * r + ax] : 0.0;
  double p = live ? origin[3 * r + ax] + u * POINT_EPS : 0.0;
  const double sd = live ? surface[r] : 0.0;
  double traveled = POINT_EPS;
  const int nx = nxe - 1, ny = nye - 1;
  const bool write = cells != nullptr;
  const int out = write && live ? row_ptr[r] : 0;
  const int cap = write && live ? row_ptr[r + 1] - out : 0;
  int n = 0;
  const double e = 1e-10;
  bool done = !live || !(traveled < sd - POINT_EPS);
  while (__ballot(!done) != 0) {
    const bool inb = !(p < E[0] || p > E[ne - 1]);
    const int i = min(max(ssl(E, ne, p), 1), ne - 1);
    double t = __builtin_huge_val();
    if (fabs(u) > e) {
      const double tt = u > 0 ? (E[i] - p) / u : (E[i - 1] - p) / u;
      if (tt > e) t = tt;
    }
    const int c0 = quad_bcast<0>((int)inb), c1 = quad_bcast<1>((int)inb), c2 = quad_bcast<2>((int)inb);
    const int i0 = quad_bcast<0>(i), i1 = quad_bcast<1>(i), i2 = quad_bcast<2>(i);
    const double t0 = quad_bcast<0>(t), t1 = quad_bcast<1>(t), t2 = quad_bcast<2>(t);
    if (done) continue;
    if (!(c0 & c1 & c2)) { done = true; continue; }
    const int cell = (i0 - 1) + (i1 - 1) * nx + (i2 - 1) * nx * ny;
    const double t_min = fmin(fmin(fmin(__builtin_huge_val(), t0), t1), t2);
    const double dist = isfinite(t_min) ? t_min : 0.0;
    if (!isfinite(dist) || dist <= 0.0) { done = true; continue; }
    const double remaining = sd - traveled;
    const double step = fmin(dist, remaining);
    if (step > 1e-8) {
      if (write && n < cap && q == 0) {
        cells[out + n] = cell;
        lens[out + n] = step;
      }
      ++n;
      traveled += step;
    }
    if (dist >= remaining - 1e-6) { done = true; continue; }
    p += u * (step + POINT_EPS);
    if (!(traveled < sd - POINT_EPS)) done = true;
  }
  if (!write && live && q == 0) {
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
  dim3 b(64), gr((n + TRACE_RAYS_PER_WAVE - 1) / TRACE_RAYS_PER_WAVE);
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
  hipLaunchKernelGGL(trace_edges_kernel, dim3((n + TRACE_RAYS_PER_WAVE - 1) / TRACE_RAYS_PER_WAVE), dim3(64), 0, c->stream,
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
