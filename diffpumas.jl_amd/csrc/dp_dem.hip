// Topography: where rays from the detector leave a gridded elevation map.
// ray_surface_exit_distance (examples/lvd_tomography.jl:219-252) over
// map_elevation (src/Turtle.jl:523-553) for many directions.
//
// The reference marches in fixed steps (up to max_distance / step bilinear
// lookups) and bisects the last bracket 40 times: a chain of dependent
// lookups per ray.  Here a group of G lanes (8 by default, a wave holds 64 / G
// rays) shares each ray.  The march takes G points at once (point k is at
// k * step, as in the reference's loop) and a ballot picks the first that
// leaves the map or lies below the ray.  The bisection runs log2(G) levels per
// pass: the lanes evaluate the G - 1 midpoints of the next levels (each lane
// derives its node's bracket by the same halvings the bisection makes on the
// way there) and the group walks the levels with those answers.  Midpoints,
// tests and the final bracket are the reference's, so results are bitwise
// those of the host restatement, with 14 dependent rounds of lookups in place
// of 40 for G = 8.  Measured at 16384 rays on a 257 x 257 map: G = 8 takes
// 0.029 ms, G = 4 0.030 ms, G = 16 0.032 ms, G = 64 0.058 ms and one ray per
// lane 0.33 ms.  The map stays in HBM and is read through the caches (a 257 x
// 257 map is 0.5 MB, L2-resident); the per-ray output is 17 B.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <algorithm>
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

#ifndef DP_DEM_GROUP
#define DP_DEM_GROUP 8
#endif
// Lanes per ray: a power of two from 2 to 64; a wave takes 64 / G rays.
constexpr int DEM_G = DP_DEM_GROUP;
static_assert(DEM_G >= 2 && DEM_G <= 64 && (DEM_G & (DEM_G - 1)) == 0, "lanes per ray");
constexpr int DEM_LEVELS = __builtin_ctz(DEM_G);  // bisection levels per pass: G - 1 nodes

// Every loop below is uniform over the wave (a group that is done idles), so
// each ballot sees the whole wave; a group reads its own G bits.
__global__ __launch_bounds__(64) void ray_exit_kernel(DemArgs a) {
  constexpr int RPW = 64 / DEM_G;
  constexpr unsigned long long GMASK = DEM_G == 64 ? ~0ull : ((1ull << (DEM_G & 63)) - 1);
  const int lane = (int)threadIdx.x;
  const int sub = lane & (DEM_G - 1), shift = lane - sub;
  for (int base = blockIdx.x * RPW; base < a.n; base += gridDim.x * RPW) {
    const int i = base + lane / DEM_G;
    const bool live = i < a.n;
    const double d0 = live ? a.dirs[3 * i] : 0.0, d1 = live ? a.dirs[3 * i + 1] : 0.0,
                 d2 = live ? a.dirs[3 * i + 2] : 0.0;
    // march: the first k in 1..n_steps that leaves the map or crosses
    long kstar = 0;
    bool cross = false, done = !(d2 > 0.0);
    for (long k0 = 1; __ballot(!done) != 0; k0 += DEM_G) {
      const long k = k0 + sub;
      bool stop = false, cr = false;
      if (!done && k <= a.n_steps) {
        const double s = (double)k * a.step;
        const double sz = surface_z(a, d0 * s, d1 * s);
        if (!isfinite(sz)) stop = true;
        else if (d2 * s > sz) { stop = true; cr = true; }
      }
      const unsigned long long ms = (__ballot(stop) >> shift) & GMASK;
      const unsigned long long mc = (__ballot(cr) >> shift) & GMASK;
      if (!done) {
        if (ms) {
          const int l = __ffsll((long long)ms) - 1;
          kstar = k0 + l;
          cross = (mc >> l) & 1ull;
          done = true;
        } else if (k0 + DEM_G > a.n_steps) {
          done = true;
        }
      }
    }
    // the reference's bracket: the previous march point (0 before the first)
    double lo = (double)(kstar - 1) * a.step, hi = (double)kstar * a.step;
    if (__ballot(cross) != 0) {
      for (int left = 40; left > 0;) {
        const int L = left < DEM_LEVELS ? left : DEM_LEVELS;
        const int nodes = (1 << L) - 1;  // internal nodes of L levels, numbered 1.. breadth-first
        bool below = false;              // ray above the surface at the node's midpoint
        if (cross && sub < nodes) {
          const int node = sub + 1;
          const int dep = 31 - __builtin_clz((unsigned)node);
          double l = lo, h = hi;
          for (int b = dep - 1; b >= 0; --b) {
            const double mid = 0.5 * (l + h);
            if ((node >> b) & 1) l = mid; else h = mid;
          }
          const double mid = 0.5 * (l + h);
          const double msz = surface_z(a, d0 * mid, d1 * mid);
          below = d2 * mid > msz;
        }
        const unsigned long long bm = (__ballot(below) >> shift) & GMASK;
        int node = 1;
        for (int t = 0; t < L; ++t) {
          const double mid = 0.5 * (lo + hi);
          if ((bm >> (node - 1)) & 1ull) { hi = mid; node = 2 * node; }
          else { lo = mid; node = 2 * node + 1; }
        }
        left -= L;
      }
    }
    if (live && sub == 0) {
      const double dist = cross ? 0.5 * (lo + hi) : __builtin_nan("");
      a.dist[i] = dist;
      a.found[i] = cross ? 1 : 0;
      a.exit_z[i] = cross ? d2 * dist : __builtin_nan("");
    }
  }
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
  const int waves = (n + 64 / DEM_G - 1) / (64 / DEM_G);
  hipLaunchKernelGGL(ray_exit_kernel, dim3((unsigned)std::min(waves, 65536)), dim3(64), 0, c->stream, a);
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
