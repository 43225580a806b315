// Ray walk through a tetrahedral mesh: trace_path(::TetraVolume)
// (examples/lvd_tomography.jl:694-736) for many rays.
//
//   tet_locate_kernel  one wave per ray: find_tetrahedron_containing_point of
//                      the detector + POINT_EPS along the ray.  The lanes test
//                      64 tets at a time; the lowest set bit of the ballot is
//                      the first tet in index order, as the reference's scan.
//                      Rays that share an origin scan only that origin's
//                      candidate list: the tets, in index order, whose
//                      bounding box, widened beyond the point-in-tet
//                      tolerance and the POINT_EPS nudge, holds the origin.
//                      A tet outside it cannot contain the nudged point, so
//                      the first hit is the full scan's.
//   tet_trace_kernel   one ray per quad of lanes, one face per lane:
//                      nearest exit face (Moller-Trumbore,
//                      src/TriangleIntersect.jl:86-127, with the tolerant
//                      plane hit of :484-501 as fallback), step to it, nudge
//                      by POINT_EPS and continue in the neighbour across that
//                      face.  Launched twice: counts, then the CSR segments.
//
// Built with -ffp-contract=off; every product, sum, division and square root
// is rounded as in the host restatement, so cells and lengths are bitwise
// those of the oracle.  The mesh (points, tets, neighbours) stays in HBM and
// is read through the caches.  A walk is capped at 4 n_tets + 16 steps (a
// straight ray in a valid mesh enters each tet at most once); reaching the
// cap is reported as DP_E_EXHAUSTED.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <tuple>
#include <vector>

namespace {

struct TetDev {
  int np, nt;
  const double* pts;
  const int32_t* tets;
  const int32_t* nbr;
};

constexpr double TET_POINT_EPS = 1e-4;   // examples/lvd_tomography.jl:147
constexpr int TET_RAYS_PER_WAVE = 16;
using dpk::quad_bcast;

__device__ __forceinline__ double tdot(const double* a, const double* b) {
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ void tcross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// point_in_tetrahedron (atol 1e-8): Julia's det and `\` on M = [b-a c-a d-a]
// (triangular M: diagonal product and a triangular solve; otherwise LU with
// partial pivoting).
__device__ bool point_in_tet(const TetDev& m, int t, const double* p) {
  const double atol = 1e-8;
  const int32_t* v = m.tets + 4 * (size_t)t;
  const double* a = m.pts + 3 * (size_t)v[0];
  double A[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double* q = m.pts + 3 * (size_t)v[c + 1];
#pragma unroll
    for (int r = 0; r < 3; ++r) A[r][c] = q[r] - a[r];
  }
  double b[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const bool lower = A[0][1] == 0.0 && A[0][2] == 0.0 && A[1][2] == 0.0;
  const bool upper = A[1][0] == 0.0 && A[2][0] == 0.0 && A[2][1] == 0.0;
  if (lower || upper) {
    const double det = (A[0][0] * A[1][1]) * A[2][2];
    if (fabs(det) < atol) return false;
    if (lower && upper) {
#pragma unroll
      for (int k = 0; k < 3; ++k) b[k] = b[k] / A[k][k];
    } else if (lower) {
#pragma unroll
      for (int k = 0; k < 3; ++k)
        if (b[k] != 0.0) {
          b[k] /= A[k][k];
#pragma unroll
          for (int r = k + 1; r < 3; ++r) b[r] -= b[k] * A[r][k];
        }
    } else {
#pragma unroll
      for (int k = 2; k >= 0; --k)
        if (b[k] != 0.0) {
          b[k] /= A[k][k];
#pragma unroll
          for (int r = k - 1; r >= 0; --r) b[r] -= b[k] * A[r][k];
        }
    }
  } else {
    int piv[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      int pr = j;
#pragma unroll
      for (int r = j + 1; r < 3; ++r)
        if (fabs(A[r][j]) > fabs(A[pr][j])) pr = r;
      piv[j] = pr;
      if (A[pr][j] != 0.0) {
        if (pr != j) {
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const double tmp = A[j][c];
            A[j][c] = A[pr][c];
            A[pr][c] = tmp;
          }
        }
        const double rcp = 1.0 / A[j][j];
#pragma unroll
        for (int r = j + 1; r < 3; ++r) A[r][j] *= rcp;
      }
#pragma unroll
      for (int r = j + 1; r < 3; ++r)
#pragma unroll
        for (int c = j + 1; c < 3; ++c) A[r][c] -= A[r][j] * A[j][c];
    }
    double det = 1.0;
    int swaps = 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      det *= A[j][j];
      if (piv[j] != j) ++swaps;
    }
    det = det * ((swaps & 1) ? -1.0 : 1.0);
    if (fabs(det) < atol) return false;
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (piv[j] != j) {
        const double tmp = b[j];
        b[j] = b[piv[j]];
        b[piv[j]] = tmp;
      }
#pragma unroll
    for (int k = 0; k < 3; ++k)
      if (b[k] != 0.0)
#pragma unroll
        for (int r = k + 1; r < 3; ++r) b[r] -= b[k] * A[r][k];
#pragma unroll
    for (int k = 2; k >= 0; --k)
      if (b[k] != 0.0) {
        b[k] /= A[k][k];
#pragma unroll
        for (int r = k - 1; r >= 0; --r) b[r] -= b[k] * A[r][k];
      }
  }
  return b[0] >= -atol && b[1] >= -atol && b[2] >= -atol && (b[0] + b[1]) + b[2] <= 1.0 + atol;
}

// intersect(ray, triangle), EPSILON 1e-10.
__device__ bool moller_trumbore(const double* o, const double* d, const double* v0,
                                const double* v1, const double* v2, double& t, double* ip) {
  const double EPS = 1e-10;
  const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  double h[3];
  tcross(d, e2, h);
  const double a = tdot(e1, h);
  if (fabs(a) < EPS) return false;
  const double f = 1.0 / a;
  const double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
  const double u = f * tdot(s, h);
  if (u < 0.0 || u > 1.0) return false;
  double q[3];
  tcross(s, e1, q);
  const double v = f * tdot(d, q);
  if (v < 0.0 || u + v > 1.0) return false;
  const double tt = f * tdot(e2, q);
  if (!(tt > EPS)) return false;
  t = tt;
  for (int i = 0; i < 3; ++i) ip[i] = o[i] + tt * d[i];
  return true;
}

// tolerant_tetra_face_hit: plane hit with barycentric slack 1e-10.
__device__ bool tolerant_hit(const double* o, const double* d, const double* p1,
                             const double* p2, const double* p3, double& dist, double* ip) {
  const double e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double n[3];
  tcross(e1, e2, n);
  const double den = tdot(n, d);
  if (!(fabs(den) > 1e-12 * sqrt(tdot(n, n)))) return false;
  const double ao[3] = {p1[0] - o[0], p1[1] - o[1], p1[2] - o[2]};
  const double distance = tdot(n, ao) / den;
  if (!(distance > 1e-8)) return false;
  const double off[3] = {o[0] + distance * d[0] - p1[0], o[1] + distance * d[1] - p1[1],
                         o[2] + distance * d[2] - p1[2]};
  const double aa = tdot(e1, e1), ab = tdot(e1, e2), bb = tdot(e2, e2);
  const double det = aa * bb - ab * ab;
  if (!(det > 0)) return false;
  const double u = (bb * tdot(off, e1) - ab * tdot(off, e2)) / det;
  const double v = (aa * tdot(off, e2) - ab * tdot(off, e1)) / det;
  const double w3 = 1 - u - v;
  if (!(u >= -1e-10 && v >= -1e-10 && w3 >= -1e-10)) return false;
  dist = distance;
  for (int i = 0; i < 3; ++i) ip[i] = o[i] + distance * d[i];
  return true;
}

// nearest_tetra_face_intersection over faces (1,2,3), (1,2,4), (1,3,4), (2,3,4).
__device__ int nearest_face(const TetDev& m, int t, const double* p, const double* d,
                            double& best, double* bp) {
  const int32_t* v = m.tets + 4 * (size_t)t;
  const double* P[4] = {m.pts + 3 * (size_t)v[0], m.pts + 3 * (size_t)v[1],
                        m.pts + 3 * (size_t)v[2], m.pts + 3 * (size_t)v[3]};
  int bf = -1;
  best = __builtin_huge_val();
  bp[0] = p[0]; bp[1] = p[1]; bp[2] = p[2];
  for (int f = 0; f < 4; ++f) {
    const double* p1 = P[f < 3 ? 0 : 1];
    const double* p2 = P[f < 2 ? 1 : 2];
    const double* p3 = P[f == 0 ? 2 : 3];
    double dist = 0.0, ip[3];
    bool ok = moller_trumbore(p, d, p1, p2, p3, dist, ip);
    if (!ok) ok = tolerant_hit(p, d, p1, p2, p3, dist, ip);
    if (ok && dist > 1e-8 && dist < best) {
      bf = f;
      best = dist;
      bp[0] = ip[0]; bp[1] = ip[1]; bp[2] = ip[2];
    }
  }
  return bf;
}

// group[r] >= 0: scan cand[gptr[group[r]] .. gptr[group[r] + 1]) (tet indices
// in ascending order); -1 (or no groups): scan every tet.
__global__ __launch_bounds__(64) void tet_locate_kernel(TetDev m, int n,
                                                        const double* __restrict__ origin,
                                                        const double* __restrict__ dir,
                                                        const int32_t* __restrict__ group,
                                                        const int32_t* __restrict__ gptr,
                                                        const int32_t* __restrict__ cand,
                                                        int32_t* __restrict__ start) {
  for (int r = blockIdx.x; r < n; r += gridDim.x) {
    const double p[3] = {origin[3 * r] + dir[3 * r] * TET_POINT_EPS,
                         origin[3 * r + 1] + dir[3 * r + 1] * TET_POINT_EPS,
                         origin[3 * r + 2] + dir[3 * r + 2] * TET_POINT_EPS};
    int found = -1;
    const int g = group ? group[r] : -1;
    if (g >= 0) {
      const int c_beg = gptr[g], c_end = gptr[g + 1];
      for (int c0 = c_beg; c0 < c_end; c0 += 64) {
        const int j = c0 + (int)threadIdx.x;
        const int t = j < c_end ? cand[j] : -1;
        const bool in = t >= 0 && point_in_tet(m, t, p);
        const unsigned long long mask = __ballot(in);
        if (mask) {
          found = cand[c0 + __ffsll((unsigned long long)mask) - 1];
          break;
        }
      }
    } else {
      for (int c0 = 0; c0 < m.nt; c0 += 64) {
        const int t = c0 + (int)threadIdx.x;
        const bool in = t < m.nt && point_in_tet(m, t, p);
        const unsigned long long mask = __ballot(in);
        if (mask) {
          found = c0 + __ffsll((unsigned long long)mask) - 1;
          break;
        }
      }
    }
    if (threadIdx.x == 0) start[r] = found;
  }
}

// One ray per quad of lanes, 16 rays per wave: lane f tests face f of the
// current tet (Moller-Trumbore, then the tolerant plane hit) and loads the
// neighbour across it; DPP quad broadcasts give every lane of the quad the
// four results, and the quad takes the nearest face in face order with strict
// comparisons, as nearest_tetra_face_intersection does.  The hit point is
// p + dist d, the expression both face tests use.  Every loop runs over the
// whole wave (a finished ray's quad idles), so each broadcast reads active
// lanes.
__global__ __launch_bounds__(64) void tet_trace_kernel(
    TetDev m, int n, const double* __restrict__ origin, const double* __restrict__ dir,
    const double* __restrict__ surface, const int32_t* __restrict__ start,
    int32_t* __restrict__ counts, const int32_t* __restrict__ row_ptr,
    int32_t* __restrict__ cells, double* __restrict__ lens, double* __restrict__ traveled_out,
    int* __restrict__ flag) {
  const int lane = (int)threadIdx.x, f = lane & 3;
  const int r = blockIdx.x * TET_RAYS_PER_WAVE + (lane >> 2);
  const bool live = r < n;
  const bool write = cells != nullptr;
  const int out = write && live ? row_ptr[r] : 0;
  const int cap = write && live ? row_ptr[r + 1] - out : 0;
  double d[3] = {0.0, 0.0, 0.0}, cur[3] = {0.0, 0.0, 0.0};
  if (live) {
    for (int i = 0; i < 3; ++i) d[i] = dir[3 * r + i];
    for (int i = 0; i < 3; ++i) cur[i] = origin[3 * r + i] + d[i] * TET_POINT_EPS;
  }
  const double sd = live ? surface[r] : 0.0;
  int tet = live ? start[r] : -1;
  double traveled = tet < 0 ? 0.0 : TET_POINT_EPS;
  int ns = 0;
  const long max_steps = 4L * m.nt + 16;
  long it = 0;
  // face f: vertices (1,2,3), (1,2,4), (1,3,4), (2,3,4) of the tet, 1-based
  const int k1 = f < 3 ? 0 : 1, k2 = f < 2 ? 1 : 2, k3 = f == 0 ? 2 : 3;
  bool done = !(tet >= 0 && traveled < sd - TET_POINT_EPS);
  while (__ballot(!done) != 0) {
    bool ok = false;
    double dist = 0.0;
    int nb = -1;
    if (!done && ++it > max_steps) {
      if (f == 0) atomicOr(flag, 1);
      done = true;
    }
    if (!done) {
      const int32_t* v = m.tets + 4 * (size_t)tet;
      const double* p1 = m.pts + 3 * (size_t)v[k1];
      const double* p2 = m.pts + 3 * (size_t)v[k2];
      const double* p3 = m.pts + 3 * (size_t)v[k3];
      nb = m.nbr[4 * (size_t)tet + f];
      double ip[3];
      ok = moller_trumbore(cur, d, p1, p2, p3, dist, ip);
      if (!ok) ok = tolerant_hit(cur, d, p1, p2, p3, dist, ip);
      ok = ok && dist > 1e-8;
    }
    const int o0 = quad_bcast<0>((int)ok), o1 = quad_bcast<1>((int)ok),
              o2 = quad_bcast<2>((int)ok), o3 = quad_bcast<3>((int)ok);
    const double t0 = quad_bcast<0>(dist), t1 = quad_bcast<1>(dist), t2 = quad_bcast<2>(dist),
                 t3 = quad_bcast<3>(dist);
    const int n0 = quad_bcast<0>(nb), n1 = quad_bcast<1>(nb), n2 = quad_bcast<2>(nb),
              n3 = quad_bcast<3>(nb);
    if (done) continue;
    int bf = -1;
    double best = __builtin_huge_val();
    if (o0 && t0 < best) { bf = 0; best = t0; }
    if (o1 && t1 < best) { bf = 1; best = t1; }
    if (o2 && t2 < best) { bf = 2; best = t2; }
    if (o3 && t3 < best) { bf = 3; best = t3; }
    if (bf < 0 || !isfinite(best)) { done = true; continue; }
    const double remaining = sd - traveled;
    const double step = fmin(best, remaining);
    if (step > 1e-8) {
      if (write && ns < cap && f == 0) {
        cells[out + ns] = tet;
        lens[out + ns] = step;
      }
      ++ns;
      traveled += step;
    }
    if (best >= remaining - 1e-6) { done = true; continue; }
    const int nt = bf == 0 ? n0 : bf == 1 ? n1 : bf == 2 ? n2 : n3;
    if (nt < 0) {
      tet = -1;
      done = true;
    } else {
      for (int i = 0; i < 3; ++i) cur[i] = (cur[i] + best * d[i]) + d[i] * TET_POINT_EPS;
      tet = nt;
      if (!(traveled < sd - TET_POINT_EPS)) done = true;
    }
  }
  if (live && f == 0) {
    if (!write) counts[r] = ns;
    if (traveled_out) traveled_out[r] = traveled;
  }
}

int check_mesh(const char* fn, const dp_tet_mesh* m) {
  const std::string f(fn);
  if (!m || !m->points || !m->tets || !m->neighbors)
    return dph::fail(DP_E_ARGUMENT, f + ": null mesh array");
  if (m->n_points < 4 || m->n_tets < 1)
    return dph::fail(DP_E_ARGUMENT, f + ": the mesh needs at least 4 points and one tet");
  for (int64_t i = 0; i < 4 * (int64_t)m->n_tets; ++i) {
    if (m->tets[i] < 0 || m->tets[i] >= m->n_points)
      return dph::fail(DP_E_ARGUMENT, f + ": tet vertex index out of range");
    if (m->neighbors[i] < -1 || m->neighbors[i] >= m->n_tets)
      return dph::fail(DP_E_ARGUMENT, f + ": neighbour index out of range");
  }
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_tet_neighbors(int32_t n_tets, const int32_t* tets, int32_t* neighbors) {
  if (n_tets < 0 || (n_tets > 0 && (!tets || !neighbors)))
    return dph::fail(DP_E_ARGUMENT, "dp_tet_neighbors: bad arguments");
  static const int F[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
  std::vector<std::tuple<int32_t, int32_t, int32_t, int32_t, int32_t>> faces;
  faces.reserve(4 * (size_t)n_tets);
  for (int32_t t = 0; t < n_tets; ++t)
    for (int f = 0; f < 4; ++f) {
      int32_t a = tets[4 * (size_t)t + F[f][0]], b = tets[4 * (size_t)t + F[f][1]],
              c = tets[4 * (size_t)t + F[f][2]];
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
      faces.emplace_back(a, b, c, t, f);
    }
  std::sort(faces.begin(), faces.end());
  std::fill(neighbors, neighbors + 4 * (size_t)n_tets, -1);
  size_t i = 0;
  while (i < faces.size()) {
    size_t j = i + 1;
    while (j < faces.size() && std::get<0>(faces[j]) == std::get<0>(faces[i]) &&
           std::get<1>(faces[j]) == std::get<1>(faces[i]) &&
           std::get<2>(faces[j]) == std::get<2>(faces[i]))
      ++j;
    if (j - i == 2) {   // a face shared by exactly two tets, as the reference links them
      const auto& x = faces[i];
      const auto& y = faces[i + 1];
      neighbors[4 * (size_t)std::get<3>(x) + std::get<4>(x)] = std::get<3>(y);
      neighbors[4 * (size_t)std::get<3>(y) + std::get<4>(y)] = std::get<3>(x);
    }
    i = j;
  }
  return DP_OK;
}

int dp_trace_tetra(dp_ctx c, const dp_tet_mesh* mesh, int32_t n, const double* origin,
                   const double* dir, const double* surface_distance, int32_t* counts,
                   const int32_t* row_ptr, int32_t* cells, double* lengths, double* traveled) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_trace_tetra: null context");
  int rc = check_mesh("dp_trace_tetra", mesh);
  if (rc) return rc;
  if (n < 0) return dph::fail(DP_E_ARGUMENT, "dp_trace_tetra: negative ray count");
  if (n == 0) return DP_OK;
  if (!origin || !dir || !surface_distance || (!cells && !counts) || (cells && (!row_ptr || !lengths)))
    return dph::fail(DP_E_ARGUMENT, "dp_trace_tetra: null argument");
  int nseg = 0;
  if (cells) {
    if (row_ptr[0] != 0) return dph::fail(DP_E_ARGUMENT, "dp_trace_tetra: row_ptr[0] must be 0");
    for (int r = 0; r < n; ++r)
      if (row_ptr[r + 1] < row_ptr[r])
        return dph::fail(DP_E_ARGUMENT, "dp_trace_tetra: row_ptr must be non-decreasing");
    nseg = row_ptr[n];
  }
  DP_HIP(hipSetDevice(c->device));
  const size_t np = (size_t)mesh->n_points, nt = (size_t)mesh->n_tets;
  // Rays grouped by their exact origin; for at most MAX_GROUPS origins, each
  // origin's candidate tets (ascending index): widened bounding box holds
  // the origin.  Widening: the nudge radius POINT_EPS |dir| (largest of the
  // group), 1e-2 of the box's largest extent (the barycentric tolerance 1e-8
  // moves an accepted point at most 4e-8 extents outside the tet; the rest
  // covers the solve's rounding on poorly shaped tets), and 1e-6 m.
  constexpr int MAX_GROUPS = 64;
  std::vector<int32_t> group((size_t)n, -1), gptr(1, 0), cand;
  {
    std::vector<std::array<double, 3>> orig;
    std::vector<double> rad;
    bool ok = true;
    for (int r = 0; r < n && ok; ++r) {
      const double* q = origin + 3 * (size_t)r;
      int g = -1;
      for (size_t k = 0; k < orig.size(); ++k)
        if (std::memcmp(orig[k].data(), q, 24) == 0) { g = (int)k; break; }
      if (g < 0) {
        if ((int)orig.size() == MAX_GROUPS) { ok = false; break; }
        orig.push_back({q[0], q[1], q[2]});
        rad.push_back(0.0);
        g = (int)orig.size() - 1;
      }
      const double* d = dir + 3 * (size_t)r;
      const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (!std::isfinite(dn)) { ok = false; break; }
      rad[g] = std::max(rad[g], dn * TET_POINT_EPS);
      group[r] = g;
    }
    if (!ok) {
      std::fill(group.begin(), group.end(), -1);
      orig.clear();
    }
    std::vector<std::vector<int32_t>> per(orig.size());
    if (!orig.empty())
      for (size_t t = 0; t < nt; ++t) {
        double lo[3], hi[3];
        for (int a = 0; a < 3; ++a) { lo[a] = INFINITY; hi[a] = -INFINITY; }
        for (int k = 0; k < 4; ++k) {
          const double* v = mesh->points + 3 * (size_t)mesh->tets[4 * t + k];
          for (int a = 0; a < 3; ++a) { lo[a] = std::min(lo[a], v[a]); hi[a] = std::max(hi[a], v[a]); }
        }
        const double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
        for (size_t g = 0; g < orig.size(); ++g) {
          const double w = 1e-2 * ext + 1.01 * rad[g] + 1e-6;
          bool in = true;  // a box that is not finite keeps the tet
          if (std::isfinite(w))
            for (int a = 0; a < 3; ++a)
              if (!(orig[g][a] >= lo[a] - w && orig[g][a] <= hi[a] + w)) in = false;
          if (in) per[g].push_back((int32_t)t);
        }
      }
    for (size_t g = 0; g < per.size(); ++g) {
      cand.insert(cand.end(), per[g].begin(), per[g].end());
      gptr.push_back((int32_t)cand.size());
    }
  }
  const size_t n_groups = gptr.size() - 1;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_p = take(24 * np), o_t = take(16 * nt), o_nb = take(16 * nt),
               o_or = take(24 * (size_t)n), o_dir = take(24 * (size_t)n),
               o_sd = take(8 * (size_t)n), o_st = take(4 * (size_t)n), o_cnt = take(4 * (size_t)n),
               o_rp = take(4 * (size_t)(n + 1)), o_cell = take(4 * (size_t)nseg),
               o_len = take(8 * (size_t)nseg), o_tr = take(8 * (size_t)n), o_fl = take(4),
               o_grp = take(4 * (size_t)n), o_gptr = take(4 * gptr.size()),
               o_cand = take(4 * std::max<size_t>(cand.size(), 1));
  rc = dph::ensure(c->staging, o);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_p, mesh->points, 24 * np, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_t, mesh->tets, 16 * nt, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_nb, mesh->neighbors, 16 * nt, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_or, origin, 24 * (size_t)n, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_dir, dir, 24 * (size_t)n, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_sd, surface_distance, 8 * (size_t)n, hipMemcpyHostToDevice, st));
  if (cells)
    DP_HIP(hipMemcpyAsync(b + o_rp, row_ptr, 4 * (size_t)(n + 1), hipMemcpyHostToDevice, st));
  const int zero = 0;
  DP_HIP(hipMemcpyAsync(b + o_fl, &zero, 4, hipMemcpyHostToDevice, st));
  if (n_groups > 0) {
    DP_HIP(hipMemcpyAsync(b + o_grp, group.data(), 4 * (size_t)n, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_gptr, gptr.data(), 4 * gptr.size(), hipMemcpyHostToDevice, st));
    if (!cand.empty())
      DP_HIP(hipMemcpyAsync(b + o_cand, cand.data(), 4 * cand.size(), hipMemcpyHostToDevice, st));
  }
  TetDev m{mesh->n_points, mesh->n_tets, (const double*)(b + o_p), (const int32_t*)(b + o_t),
           (const int32_t*)(b + o_nb)};
  const double* d_or = (const double*)(b + o_or);
  const double* d_dir = (const double*)(b + o_dir);
  int32_t* d_st = (int32_t*)(b + o_st);
  const unsigned loc_grid = (unsigned)std::min(n, 65536);
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(tet_locate_kernel, dim3(loc_grid), dim3(64), 0, st, m, n, d_or, d_dir,
                     n_groups > 0 ? (const int32_t*)(b + o_grp) : nullptr,
                     (const int32_t*)(b + o_gptr), (const int32_t*)(b + o_cand), d_st);
  hipLaunchKernelGGL(tet_trace_kernel, dim3((n + TET_RAYS_PER_WAVE - 1) / TET_RAYS_PER_WAVE), dim3(64), 0, st, m, n, d_or, d_dir,
                     (const double*)(b + o_sd), (const int32_t*)d_st,
                     cells ? nullptr : (int32_t*)(b + o_cnt), (const int32_t*)(b + o_rp),
                     cells ? (int32_t*)(b + o_cell) : nullptr, (double*)(b + o_len),
                     (double*)(b + o_tr), (int*)(b + o_fl));
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  int flag = 0;
  if (!cells) {
    DP_HIP(hipMemcpyAsync(counts, b + o_cnt, 4 * (size_t)n, hipMemcpyDeviceToHost, st));
  } else if (nseg > 0) {
    DP_HIP(hipMemcpyAsync(cells, b + o_cell, 4 * (size_t)nseg, hipMemcpyDeviceToHost, st));
    DP_HIP(hipMemcpyAsync(lengths, b + o_len, 8 * (size_t)nseg, hipMemcpyDeviceToHost, st));
  }
  if (traveled) DP_HIP(hipMemcpyAsync(traveled, b + o_tr, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(&flag, b + o_fl, 4, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (flag)
    return dph::fail(DP_E_EXHAUSTED,
                     "dp_trace_tetra: a ray exceeded 4 n_tets + 16 steps (inconsistent mesh)");
  return DP_OK;
}

}  // extern "C"
