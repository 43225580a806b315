// CPU oracle of the ray walk through a tetrahedral mesh.  TEST INFRASTRUCTURE ONLY.
//
// trace_path(::TetraVolume) (examples/lvd_tomography.jl:694-736) with
// find_tetrahedron_containing_point and point_in_tetrahedron (:453-479),
// nearest_tetra_face_intersection (:503-524), tolerant_tetra_face_hit
// (:484-501), tetra_face_indices (:415-422), build_tetra_neighbors (:431-451)
// and the Moller-Trumbore intersect of src/TriangleIntersect.jl:86-127.
//
// Arithmetic restated as the reference writes it: dot products summed left
// to right (x1 y1 + x2 y2) + x3 y3; cross products as LinearAlgebra.cross;
// det and the 3x3 solve of point_in_tetrahedron as Julia dispatches them (a
// triangular matrix: product of the diagonal and a triangular solve; any
// other: LU with partial pivoting as reference LAPACK dgetf2/dgetrs order
// it); norm as the square root of the dot product.  The reference calls
// OpenBLAS for dot, norm, LU and the triangular solves, whose kernels may
// round the last bit differently: parity with the reference is to rounding,
// which only decides cases within ~1e-8 m of a face.  Mesh indices are
// 0-based here; a missing neighbour is -1.
#pragma once
#include <cmath>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

namespace orc {

struct TetMesh {
  int np, nt;
  const double* pts;   // 3 * np: x y z of point p at pts[3p..3p+2]
  const int* tets;     // 4 * nt: the four point indices of tet t
  const int* nbr;      // 4 * nt: neighbour across face f of tet t, -1 = none
};

inline double tdot(const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline void tcross(const double* a, const double* b, double* c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

// tetra_face_indices: faces (1,2,3), (1,2,4), (1,3,4), (2,3,4) of the tet.
constexpr int TET_FACE[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};

inline void tet_face(const TetMesh& m, int t, int f, const double*& p1, const double*& p2,
                     const double*& p3) {
  const int* v = m.tets + 4 * t;
  p1 = m.pts + 3 * v[TET_FACE[f][0]];
  p2 = m.pts + 3 * v[TET_FACE[f][1]];
  p3 = m.pts + 3 * v[TET_FACE[f][2]];
}

// build_tetra_neighbors: faces shared by exactly two tets.
inline void build_neighbors(int nt, const int* tets, int* nbr) {
  std::map<std::tuple<int, int, int>, std::vector<std::pair<int, int>>> faces;
  for (int t = 0; t < nt; ++t)
    for (int f = 0; f < 4; ++f) {
      int a = tets[4 * t + TET_FACE[f][0]], b = tets[4 * t + TET_FACE[f][1]],
          c = tets[4 * t + TET_FACE[f][2]];
      if (a > b) std::swap(a, b);
      if (b > c) std::swap(b, c);
      if (a > b) std::swap(a, b);
      faces[std::make_tuple(a, b, c)].push_back({t, f});
    }
  for (int i = 0; i < 4 * nt; ++i) nbr[i] = -1;
  for (const auto& kv : faces)
    if (kv.second.size() == 2) {
      const auto& e1 = kv.second[0];
      const auto& e2 = kv.second[1];
      nbr[4 * e1.first + e1.second] = e2.first;
      nbr[4 * e2.first + e2.second] = e1.first;
    }
}

// point_in_tetrahedron (atol 1e-8) on M = [b-a c-a d-a].  As Julia's det and
// `\` do, a triangular M takes the triangular path (det = product of the
// diagonal; forward or backward substitution as BLAS trsv orders it); any
// other M one LU factorisation with partial pivoting.
inline bool point_in_tet(const TetMesh& m, int t, const double* p) {
  const double atol = 1e-8;
  const int* v = m.tets + 4 * t;
  const double* a = m.pts + 3 * v[0];
  double A[3][3];
  for (int c = 0; c < 3; ++c) {
    const double* q = m.pts + 3 * v[c + 1];
    for (int r = 0; r < 3; ++r) A[r][c] = q[r] - a[r];
  }
  double b[3] = {p[0] - a[0], p[1] - a[1], p[2] - a[2]};
  const bool lower = A[0][1] == 0.0 && A[0][2] == 0.0 && A[1][2] == 0.0;
  const bool upper = A[1][0] == 0.0 && A[2][0] == 0.0 && A[2][1] == 0.0;
  if (lower || upper) {
    const double det = (A[0][0] * A[1][1]) * A[2][2];
    if (std::fabs(det) < atol) return false;
    if (lower && upper) {
      for (int k = 0; k < 3; ++k) b[k] = b[k] / A[k][k];
    } else if (lower) {
      for (int k = 0; k < 3; ++k)
        if (b[k] != 0.0) {
          b[k] /= A[k][k];
          for (int r = k + 1; r < 3; ++r) b[r] -= b[k] * A[r][k];
        }
    } else {
      for (int k = 2; k >= 0; --k)
        if (b[k] != 0.0) {
          b[k] /= A[k][k];
          for (int r = k - 1; r >= 0; --r) b[r] -= b[k] * A[r][k];
        }
    }
  } else {
    int piv[3];
    for (int j = 0; j < 3; ++j) {
      int pr = j;
      for (int r = j + 1; r < 3; ++r)
        if (std::fabs(A[r][j]) > std::fabs(A[pr][j])) pr = r;
      piv[j] = pr;
      if (A[pr][j] != 0.0) {
        if (pr != j)
          for (int c = 0; c < 3; ++c) std::swap(A[j][c], A[pr][c]);
        const double rcp = 1.0 / A[j][j];
        for (int r = j + 1; r < 3; ++r) A[r][j] *= rcp;
      }
      for (int r = j + 1; r < 3; ++r)
        for (int c = j + 1; c < 3; ++c) A[r][c] -= A[r][j] * A[j][c];
    }
    double det = 1.0;
    int swaps = 0;
    for (int j = 0; j < 3; ++j) {
      det *= A[j][j];
      if (piv[j] != j) ++swaps;
    }
    det = det * ((swaps & 1) ? -1.0 : 1.0);
    if (std::fabs(det) < atol) return false;
    for (int j = 0; j < 3; ++j)
      if (piv[j] != j) std::swap(b[j], b[piv[j]]);
    for (int k = 0; k < 3; ++k)
      if (b[k] != 0.0)
        for (int r = k + 1; r < 3; ++r) b[r] -= b[k] * A[r][k];
    for (int k = 2; k >= 0; --k)
      if (b[k] != 0.0) {
        b[k] /= A[k][k];
        for (int r = k - 1; r >= 0; --r) b[r] -= b[k] * A[r][k];
      }
  }
  return b[0] >= -atol && b[1] >= -atol && b[2] >= -atol && (b[0] + b[1]) + b[2] <= 1.0 + atol;
}

// find_tetrahedron_containing_point: the first tet in index order, or -1.
inline int find_tet(const TetMesh& m, const double* p) {
  for (int t = 0; t < m.nt; ++t)
    if (point_in_tet(m, t, p)) return t;
  return -1;
}

// intersect(ray, triangle): Moller-Trumbore with EPSILON 1e-10.
inline bool moller_trumbore(const double* o, const double* d, const double* v0, const double* v1,
                            const double* v2, double& t, double* ip) {
  const double EPS = 1e-10;
  const double e1[3] = {v1[0] - v0[0], v1[1] - v0[1], v1[2] - v0[2]};
  const double e2[3] = {v2[0] - v0[0], v2[1] - v0[1], v2[2] - v0[2]};
  double h[3];
  tcross(d, e2, h);
  const double a = tdot(e1, h);
  if (std::fabs(a) < EPS) return false;
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
inline bool tolerant_hit(const double* o, const double* d, const double* p1, const double* p2,
                         const double* p3, double& dist, double* ip) {
  const double e1[3] = {p2[0] - p1[0], p2[1] - p1[1], p2[2] - p1[2]};
  const double e2[3] = {p3[0] - p1[0], p3[1] - p1[1], p3[2] - p1[2]};
  double n[3];
  tcross(e1, e2, n);
  const double den = tdot(n, d);
  if (!(std::fabs(den) > 1e-12 * std::sqrt(tdot(n, n)))) return false;
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
  const double w3 = 1 - u - v;   // min(u, v, 1-u-v) >= -1e-10, false if any is NaN
  if (!(u >= -1e-10 && v >= -1e-10 && w3 >= -1e-10)) return false;
  dist = distance;
  for (int i = 0; i < 3; ++i) ip[i] = o[i] + distance * d[i];
  return true;
}

// nearest_tetra_face_intersection: (face 0..3 or -1, distance, point).
inline int nearest_face(const TetMesh& m, int t, const double* p, const double* d, double& best,
                        double* bp) {
  int bf = -1;
  best = INFINITY;
  bp[0] = p[0]; bp[1] = p[1]; bp[2] = p[2];
  for (int f = 0; f < 4; ++f) {
    const double *p1, *p2, *p3;
    tet_face(m, t, f, p1, p2, p3);
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

// trace_path(::TetraVolume) from `origin` (the detector) along the unit
// vector d up to the topography exit at surface_distance, starting in tet
// `start` (find_tet of origin + POINT_EPS d; -1: none).  Appends (tet, length)
// segments and returns the distance travelled.
inline double trace_tet(const TetMesh& m, const double* origin, const double* d,
                        double surface_distance, int start, std::vector<int>& cells,
                        std::vector<double>& lens) {
  const double POINT_EPS = 1e-4;
  double cur[3] = {origin[0] + d[0] * POINT_EPS, origin[1] + d[1] * POINT_EPS,
                   origin[2] + d[2] * POINT_EPS};
  int tet = start;
  double traveled = tet < 0 ? 0.0 : POINT_EPS;
  while (tet >= 0 && traveled < surface_distance - POINT_EPS) {
    double dist, ip[3];
    const int f = nearest_face(m, tet, cur, d, dist, ip);
    if (f < 0 || !std::isfinite(dist)) break;
    const double remaining = surface_distance - traveled;
    const double step = std::fmin(dist, remaining);
    if (step > 1e-8) {
      cells.push_back(tet);
      lens.push_back(step);
      traveled += step;
    }
    if (dist >= remaining - 1e-6) break;
    const int nt = m.nbr[4 * tet + f];
    if (nt < 0) {
      tet = -1;
    } else {
      cur[0] = ip[0] + d[0] * POINT_EPS;
      cur[1] = ip[1] + d[1] * POINT_EPS;
      cur[2] = ip[2] + d[2] * POINT_EPS;
      tet = nt;
    }
  }
  return traveled;
}

}  // namespace orc
