// Topography: where rays from the detector leave a gridded elevation map.
// ray_surface_exit_distance (examples/lvd_tomography.jl:219-252) over
// map_elevation (src/Turtle.jl:523-553) for many directions, one ray per lane.
//
// Each lane marches in fixed steps (up to max_distance / step bilinear
// lookups), then bisects the last bracket 40 times.  The map stays in HBM and
// is read through the caches (a 256 x 256 map is 512 KB, L2-resident); the
// per-ray output is 17 B.  Operation order follows the reference, so results
// are bitwise those of the host restatement.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <cmath>

namespace {

struct DemArgs {
  int nx, ny;
  double x0, y0, dx, dy;
  const double* data;  // z(ix, iy) = data[iy + ny * ix]
  double xy_scale, z_ref, step;
  long n_steps;
  int n;
  const double* dirs;
  double* dist;
  uint8_t* found;
  double* exit_z;
};

// map_elevation (bilinear; NaN off the map) minus the detector elevation.
__device__ __forceinline__ double surface_z(const DemArgs& a, double xm, double ym) {
  const double x = xm * a.xy_scale, y = ym * a.xy_scale;
  if (isnan(x) || isnan(y)) return __builtin_nan("");
  double hx = (x - a.x0) / a.dx;
  double hy = (y - a.y0) / a.dy;
  if (hx > a.nx - 1 || hx < 0 || hy > a.ny - 1 || hy < 0) return __builtin_nan("");
  int ix = (int)floor(hx);
  int iy = (int)floor(hy);
  if (ix == a.nx - 1) { ix -= 1; hx = 1.0; } else { hx -= ix; }
  if (iy == a.ny - 1) { iy -= 1; hy = 1.0; } else { hy -= iy; }
  const double* c0 = a.data + (size_t)a.ny * ix + iy;
  const double* c1 = c0 + a.ny;
  const double z00 = c0[0], z01 = c0[1], z10 = c1[0], z11 = c1[1];
  const double e = z00 * (1.0 - hx) * (1.0 - hy) + z01 * (1.0 - hx) * hy +
                   z10 * hx * (1.0 - hy) + z11 * hx * hy;
  return e - a.z_ref;
}

__global__ __launch_bounds__(64) void ray_exit_kernel(DemArgs a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double d0 = a.dirs[3 * i], d1 = a.dirs[3 * i + 1], d2 = a.dirs[3 * i + 2];
  double dist = __builtin_nan(""), ez = __builtin_nan("");
  uint8_t found = 0;
  if (d2 > 0.0) {
    double prev = 0.0;
    for (long k = 1; k <= a.n_steps; ++k) {
      const double s = (double)k * a.step;
      const double sz = surface_z(a, d0 * s, d1 * s);
      if (!isfinite(sz)) break;
      if (d2 * s > sz) {
        double lo = prev, hi = s;
        for (int it = 0; it < 40; ++it) {
          const double mid = 0.5 * (lo + hi);
          const double ms = surface_z(a, d0 * mid, d1 * mid);
          if (d2 * mid > ms) hi = mid; else lo = mid;
        }
        dist = 0.5 * (lo + hi);
        found = 1;
        ez = d2 * dist;
        break;
      }
      prev = s;
    }
  }
  a.dist[i] = dist;
  a.found[i] = found;
  a.exit_z[i] = ez;
}

int check_dem(const char* fn, const dp_elevation_map* m, int32_t n, const double* dirs, double step,
              double max_distance, const double* dist, const uint8_t* found, const double* ez) {
  const std::string f(fn);
  if (!m || (n > 0 && (!dirs || !dist || !found || !ez)) || !m->data)
    return dph::fail(DP_E_ARGUMENT, f + ": null argument");
  if (n < 0) return dph::fail(DP_E_ARGUMENT, f + ": negative ray count");
  if (m->nx < 2 || m->ny < 2 || !(m->dx > 0) || !(m->dy > 0))
    return dph::fail(DP_E_ARGUMENT, f + ": the map needs at least 2 x 2 nodes and positive spacings");
  if (!(step > 0) || !(max_distance >= 0) || max_distance / step > 1e8)
    return dph::fail(DP_E_ARGUMENT, f + ": need step > 0 and at most 1e8 steps");
  return DP_OK;
}

int launch_dem(dp_ctx c, const dp_elevation_map* m, double xy_scale, double z_ref, int32_t n,
               const double* dirs, double step, double max_distance, double* dist,
               uint8_t* found, double* ez, const double* data) {
  DemArgs a;
  a.nx = m->nx; a.ny = m->ny; a.x0 = m->x0; a.y0 = m->y0; a.dx = m->dx; a.dy = m->dy;
  a.data = data;
  a.xy_scale = xy_scale; a.z_ref = z_ref; a.step = step;
  a.n_steps = (long)floor(max_distance / step + 1e-9);
  a.n = n; a.dirs = dirs; a.dist = dist; a.found = found; a.exit_z = ez;
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(ray_exit_kernel, dim3((n + 63) / 64), dim3(64), 0, c->stream, a);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_ray_surface_exit_dev(dp_ctx c, const dp_elevation_map* m, double xy_scale, double z_ref,
                            int32_t n, const double* dirs, double step, double max_distance,
                            double* dist, uint8_t* found, double* exit_z) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_ray_surface_exit_dev: null context");
  int rc = check_dem("dp_ray_surface_exit_dev", m, n, dirs, step, max_distance, dist, found, exit_z);
  if (rc || n == 0) return rc;
  DP_HIP(hipSetDevice(c->device));
  return launch_dem(c, m, xy_scale, z_ref, n, dirs, step, max_distance, dist, found, exit_z,
                    m->data);
}

int dp_ray_surface_exit(dp_ctx c, const dp_elevation_map* m, double xy_scale, double z_ref,
                        int32_t n, const double* dirs, double step, double max_distance,
                        double* dist, uint8_t* found, double* exit_z) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_ray_surface_exit: null context");
  int rc = check_dem("dp_ray_surface_exit", m, n, dirs, step, max_distance, dist, found, exit_z);
  if (rc || n == 0) return rc;
  DP_HIP(hipSetDevice(c->device));
  const size_t nm = (size_t)m->nx * m->ny;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_map = take(8 * nm), o_dir = take(24 * (size_t)n), o_d = take(8 * (size_t)n),
               o_f = take((size_t)n), o_z = take(8 * (size_t)n);
  rc = dph::ensure(c->staging, o);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_map, m->data, 8 * nm, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_dir, dirs, 24 * (size_t)n, hipMemcpyHostToDevice, st));
  rc = launch_dem(c, m, xy_scale, z_ref, n, (const double*)(b + o_dir), step, max_distance,
                  (double*)(b + o_d), (uint8_t*)(b + o_f), (double*)(b + o_z),
                  (const double*)(b + o_map));
  if (rc) return rc;
  DP_HIP(hipMemcpyAsync(dist, b + o_d, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(found, b + o_f, (size_t)n, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(exit_z, b + o_z, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

}  // extern "C"
