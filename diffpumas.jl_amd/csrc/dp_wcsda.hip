// Water-fraction (w) tomography CSDA: directional flux and its sparse
// Jacobian d flux / d w.
//
// Work decomposition: persistent blocks take one path (bin) at a time from a
// work queue; the block's lanes are the energy samples.  Every lane walks the
// same segment list, so the segment data are wave-uniform loads.  Rock,
// water, porous and air stopping-power tables are staged in LDS.
//
// Per sample (lane) the forward sweep runs the Heun sub-stepping of
// propagate_csda_cell_segment (src/Tomography.jl:444-473) on each cell with
// a two-direction dual (w, E_in), giving the segment partials
//   ew = dE_out/dw, ee = dE_out/dE_in, lw = dlog G/dw, le = dlog G/dE_in
// (G = the segment's weight factor), i.e. _cell_segment_partials of
// src/Tomography_rules.jl:33-48.  The dry-rock remainder and the atmosphere
// (:481-551) carry one energy tangent.  The reverse sweep chains them:
//   d flux/d w_s = f (lw_s + T ew_s),   T <- le_s + T ee_s
// with T = d log(flux)/dE at the exit of segment s and f the sample's flux
// term, which is the cotangent algebra of the rrules written for a
// multiplicative weight.  Per-segment sums over samples use a fixed wave/LDS
// tree, so the Jacobian is deterministic and independent of launch shape.
//
// With the MC-calibrated corrections on (CORR), each sample also accumulates
// V = sum Omega(E_mid) dX and H = sum (S_csda - S_mixed)(E_mid) dX over the
// rock column (_directional_flux_core, src/Tomography.jl:585-637); the tape
// gains the total derivatives of each segment's V and H terms in (w, E_in),
// and the reverse sweep adds A_V dV + A_H dH, A_V and A_H the derivatives of
// log csda_var_correction (:384-391) in V and H.  The geometric and material
// multipliers (:394-412, :560-580) are per bin; the material one adds
// M'(q) dq/dw_s times the bin's sum.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <cstring>

using namespace dpk;

namespace {

constexpr double MAX_REL_GAIN = 0.02;  // CSDA_MAX_RELATIVE_GAIN, src/Tomography.jl:90
constexpr double MAX_STEP_M = 60.0;    // CSDA_MAX_STEP_M, src/Tomography.jl:91
constexpr int ERR_CAP = dph::KERR_CAPACITY, ERR_CELL = dph::KERR_CELL;

// _csda_min / _csda_max (src/Tomography.jl:421-434): branch on primal values,
// the left operand wins ties.
template <class A, class B>
__device__ __forceinline__ A cmin(A a, B b) { return val(a) <= val(b) ? a : A(b); }
template <class A, class B>
__device__ __forceinline__ A cmax(A a, B b) { return val(a) >= val(b) ? a : A(b); }

struct WTab {
  const double *rE, *rLE, *rS;  // rock
  const double *wE, *wLE, *wS;  // water
  const double *pE, *pLE, *pS;  // porous (aliases rock when unused)
  const double *aE, *aLE, *aS;  // air
  const double *rM, *rO, *wM, *wO, *pM, *pO;  // MIXED stopping power, straggling (CORR)
  int K, pow2, shared;          // shared: rock/water/porous energy grids identical
  double mass, ctau;
};
struct WMat {
  double rho_r, rho_w, rho_p;
  int porous_on;  // porous material set and porous_top_thickness > 0
};

template <class R>
__device__ __forceinline__ R lerp_at(const double* ys, int i, R t) {
  return ys[i] + t * (ys[i + 1] - ys[i]);
}

// cell_stopping_power, src/Tomography.jl:304-314.  With a shared energy grid
// the three interpolations reuse one clamp test and one interval search
// (started from the guess g, the interval of the previous lookup), which
// returns the same values as three separate lookups.
template <class R>
__device__ __forceinline__ R cell_stop(const WTab& T, R w, bool porous, R e, int& g) {
  const int K = T.K;
  R le = qlog(e);
  R sr, sw, sp;
  if (T.shared) {
    if (val(e) <= T.rE[0]) {
      sr = R(T.rS[0]); sw = R(T.wS[0]); sp = R(T.pS[0]);
    } else if (val(e) >= T.rE[K - 1]) {
      sr = R(T.rS[K - 1]); sw = R(T.wS[K - 1]); sp = R(T.pS[K - 1]);
    } else {
      const int i = locate_idx(T.rLE, K, T.pow2, val(le), g);
      g = i;
      R t = qdiv(le - T.rLE[i], T.rLE[i + 1] - T.rLE[i]);
      sr = lerp_at(T.rS, i, t);
      sw = lerp_at(T.wS, i, t);
      sp = porous ? lerp_at(T.pS, i, t) : R(0.0);
    }
  } else {
    sr = interp_r(T.rE, T.rLE, T.rS, K, T.pow2, e, le);
    sw = interp_r(T.wE, T.wLE, T.wS, K, T.pow2, e, le);
    sp = porous ? interp_r(T.pE, T.pLE, T.pS, K, T.pow2, e, le) : R(0.0);
  }
  if (porous) return 0.5 * (1.0 - w) * sp + 0.5 * (1.0 - w) * sr + w * sw;
  return (1.0 - w) * sr + w * sw;
}

// cell_density, src/Tomography.jl:295-302 (harmonic in the mass fraction).
template <class R>
__device__ __forceinline__ R cell_density(const WMat& m, R w, bool porous) {
  if (porous)
    return qdiv(1.0, qdiv(0.5 * (1.0 - w), m.rho_p) + qdiv(0.5 * (1.0 - w), m.rho_r) + qdiv(w, m.rho_w));
  return qdiv(1.0, qdiv(1.0 - w, m.rho_r) + qdiv(w, m.rho_w));
}

// cell_straggling and cell_hardloss (src/Tomography.jl:329-352) at one
// energy: om = Omega(w), hl = sum of fractions x (S_csda - S_mixed).
template <class R>
__device__ __forceinline__ void cell_strag_hard(const WTab& T, R w, bool porous, R e, R& om,
                                                R& hl) {
  const int K = T.K;
  R le = log(e);
  R orr, ow, op, hr, hw, hp;
  if (T.shared) {
    if (val(e) <= T.rE[0] || val(e) >= T.rE[K - 1]) {
      const int j = val(e) <= T.rE[0] ? 0 : K - 1;
      orr = R(T.rO[j]); ow = R(T.wO[j]); op = R(T.pO[j]);
      hr = R(T.rS[j]) - R(T.rM[j]); hw = R(T.wS[j]) - R(T.wM[j]); hp = R(T.pS[j]) - R(T.pM[j]);
    } else {
      const int i = bracket(T.rLE, K, T.pow2, val(le));
      R t = (le - T.rLE[i]) / (T.rLE[i + 1] - T.rLE[i]);
      orr = lerp_at(T.rO, i, t); ow = lerp_at(T.wO, i, t);
      hr = lerp_at(T.rS, i, t) - lerp_at(T.rM, i, t);
      hw = lerp_at(T.wS, i, t) - lerp_at(T.wM, i, t);
      if (porous) {
        op = lerp_at(T.pO, i, t);
        hp = lerp_at(T.pS, i, t) - lerp_at(T.pM, i, t);
      } else {
        op = R(0.0); hp = R(0.0);
      }
    }
  } else {
    orr = interp_r(T.rE, T.rLE, T.rO, K, T.pow2, e, le);
    ow = interp_r(T.wE, T.wLE, T.wO, K, T.pow2, e, le);
    hr = interp_r(T.rE, T.rLE, T.rS, K, T.pow2, e, le) - interp_r(T.rE, T.rLE, T.rM, K, T.pow2, e, le);
    hw = interp_r(T.wE, T.wLE, T.wS, K, T.pow2, e, le) - interp_r(T.wE, T.wLE, T.wM, K, T.pow2, e, le);
    if (porous) {
      op = interp_r(T.pE, T.pLE, T.pO, K, T.pow2, e, le);
      hp = interp_r(T.pE, T.pLE, T.pS, K, T.pow2, e, le) - interp_r(T.pE, T.pLE, T.pM, K, T.pow2, e, le);
    } else {
      op = R(0.0); hp = R(0.0);
    }
  }
  if (porous) {
    om = 0.5 * (1.0 - w) * op + 0.5 * (1.0 - w) * orr + w * ow;
    hl = 0.5 * (1.0 - w) * hp + 0.5 * (1.0 - w) * hr + w * hw;
  } else {
    om = (1.0 - w) * orr + w * ow;
    hl = (1.0 - w) * hr + w * hw;
  }
}

// property_straggling and rock_hardloss (src/Tomography.jl:318-321) of the
// dry-rock remainder.
template <class R>
__device__ __forceinline__ void rock_strag_hard(const WTab& T, R e, R& om, R& hl) {
  R le = log(e);
  om = interp_r(T.rE, T.rLE, T.rO, T.K, T.pow2, e, le);
  hl = interp_r(T.rE, T.rLE, T.rS, T.K, T.pow2, e, le) - interp_r(T.rE, T.rLE, T.rM, T.K, T.pow2, e, le);
}

// CsdaCorrection / CsdaMaterialCorrection coefficients.
struct WCorr {
  int enabled, geom, mat;  // geom: some residual coefficient is non-zero
  double ks, kh, ra, rb, rc, rd, lin, quad;
};

// log csda_var_correction (src/Tomography.jl:384-391) with the spectral
// index of local_spectral_index (:367-372) at charge +1.
template <class R>
__device__ __forceinline__ R log_var_corr(R Es, double V, double H, double ct, const WCorr& C) {
  R e = cmax(Es, 1.0);
  const double eta = 1.0e-2;
  R fp = flux_gccly_r(ct, e * (1.0 + eta), 1.0);
  R fm = flux_gccly_r(ct, e * (1.0 - eta), 1.0);
  R b = -(log(fp) - log(fm)) / (2 * eta);
  R cs = 0.5 * b * (b + 1.0);
  return C.ks * cs * V / (e * e) - C.kh * (H >= 0.0 ? H : 0.0) / e;
}

// compute_decay_weight_from_path, src/Geometry.jl:1083-1107 (the beta floor
// branch is chosen on the primal value).
template <class R, class P>
__device__ __forceinline__ R decay_exponent(const WTab& T, P path, R eavg) {
  R g = 1.0 + eavg / T.mass;
  R b2 = 1.0 - 1.0 / (g * g);
  R beta = sqrt(val(b2) >= 1e-12 ? b2 : R(1e-12));
  R L = beta * g * T.ctau;
  return path / L;
}
template <class R, class P>
__device__ __forceinline__ R decay_r(const WTab& T, P path, R eavg) {
  return exp(-decay_exponent(T, path, eavg));
}

// interp_r with the interval searched from the guess g (updated); locate_idx
// returns the interval of the bisection, so values are those of interp_r.
template <class R>
__device__ __forceinline__ R interp_rg(const double* es, const double* les, const double* ys, int n,
                                       int pow2, R x, R lx, int& g) {
  if (val(x) <= es[0]) return R(ys[0]);
  if (val(x) >= es[n - 1]) return R(ys[n - 1]);
  const int i = locate_idx(les, n, pow2, val(lx), g);
  g = i;
  R t = (lx - les[i]) / (les[i + 1] - les[i]);
  return ys[i] + t * (ys[i + 1] - ys[i]);
}
template <class R>
__device__ __forceinline__ R air_stop(const WTab& T, R e, int& g) {
  return interp_rg(T.aE, T.aLE, T.aS, T.K, T.pow2, e, log(e), g);
}
// The atmosphere walk's lookups and decay with qdiv/qlog/qsqrt quotients
// (refined hardware reciprocals): about 500 sub-steps per node, each with
// some twenty quotients in the dual form.
template <class R>
__device__ __forceinline__ R air_stop_q(const WTab& T, R e, int& g) {
  const R lx = qlog(e);
  if (val(e) <= T.aE[0]) return R(T.aS[0]);
  if (val(e) >= T.aE[T.K - 1]) return R(T.aS[T.K - 1]);
  const int i = locate_idx(T.aLE, T.K, T.pow2, val(lx), g);
  g = i;
  R t = qdiv(lx - T.aLE[i], T.aLE[i + 1] - T.aLE[i]);
  return T.aS[i] + t * (T.aS[i + 1] - T.aS[i]);
}
template <class R, class P>
__device__ __forceinline__ R decay_exponent_q(const WTab& T, P path, R eavg) {
  R g = 1.0 + qdiv(eavg, T.mass);
  R b2 = 1.0 - qdiv(1.0, g * g);
  R beta = qsqrt(val(b2) >= 1e-12 ? b2 : R(1e-12));
  R L = beta * g * T.ctau;
  return qdiv(path, L);
}
template <class R>
__device__ __forceinline__ R rock_stop(const WTab& T, R e, int& g) {
  return interp_rg(T.rE, T.rLE, T.rS, T.K, T.pow2, e, log(e), g);
}
template <class R>
__device__ __forceinline__ R rock_stop_q(const WTab& T, R e, int& g) {
  const R lx = qlog(e);
  if (val(e) <= T.rE[0]) return R(T.rS[0]);
  if (val(e) >= T.rE[T.K - 1]) return R(T.rS[T.K - 1]);
  const int i = locate_idx(T.rLE, T.K, T.pow2, val(lx), g);
  g = i;
  R t = qdiv(lx - T.rLE[i], T.rLE[i + 1] - T.rLE[i]);
  return T.rS[i] + t * (T.rS[i + 1] - T.rS[i]);
}

// propagate_csda_cell_segment, src/Tomography.jl:444-473.  The stopping power
// at the end of a sub-step is the next sub-step's starting value (same
// function of the same energy), so it is evaluated once.
template <class R>
__device__ bool cell_segment(const WTab& T, const WMat& m, R w, bool porous, double distance,
                             R& energy, R& weight, int& g) {
  R density = cell_density(m, w, porous);
  R remaining = R(distance);
  R dedx0 = cell_stop(T, w, porous, energy, g);
  while (val(remaining) > 1e-8) {
    if (!isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); return false; }
    R grel = qdiv(MAX_REL_GAIN * cmax(energy, 1e-6), dedx0);
    R gstep = cmin(cmin(remaining * density, cmax(grel, 1e-6)), MAX_STEP_M * density);
    R ds = cmin(remaining, qdiv(gstep, cmax(density, 1e-12)));
    if (val(ds) <= 0.0) { weight = R(0.0); return false; }
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = cell_stop(T, w, porous, epred, g);
    R enext = energy + 0.5 * (dedx0 + dedx1) * gram;
    if (!isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); return false; }
    R dnext = cell_stop(T, w, porous, enext, g);
    weight = weight * qdiv(dnext, cmax(dedx0, 1e-30));
    weight = weight * exp(-decay_exponent_q(T, ds, 0.5 * (energy + enext)));
    energy = enext;
    remaining = remaining - ds;
    dedx0 = dnext;
  }
  return true;
}

// propagate_csda_uniform_segment, src/Tomography.jl:481-509 (rock remainder).
template <class R>
__device__ bool rock_segment(const WTab& T, double density, double distance, R& energy, R& weight,
                             int& g) {
  R remaining = R(distance);
  R dedx0 = rock_stop_q(T, energy, g);
  while (val(remaining) > 1e-8) {
    if (!isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); return false; }
    R grel = qdiv(MAX_REL_GAIN * cmax(energy, 1e-6), dedx0);
    R gstep = cmin(cmin(remaining * density, cmax(grel, 1e-6)), R(MAX_STEP_M * density));
    R ds = cmin(remaining, qdiv(gstep, fmax(density, 1e-12)));
    if (val(ds) <= 0.0) { weight = R(0.0); return false; }
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = rock_stop_q(T, epred, g);
    R enext = energy + 0.5 * (dedx0 + dedx1) * gram;
    if (!isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); return false; }
    R dnext = rock_stop_q(T, enext, g);
    weight = weight * qdiv(dnext, cmax(dedx0, 1e-30));
    weight = weight * exp(-decay_exponent_q(T, ds, 0.5 * (energy + enext)));
    energy = enext;
    remaining = remaining - ds;
    dedx0 = dnext;
  }
  return true;
}

// propagate_csda_air_segment, src/Tomography.jl:516-551: exponential
// atmosphere at the mid-step altitude (locals_air_at_altitude,
// src/Geometry.jl:95-106), 1 mm minimum step.
template <class R>
__device__ bool air_segment(const WTab& T, double z0, double cos_t, double distance, R& energy,
                            R& weight, int& g) {
  R remaining = R(distance);
  R traversed = R(0.0);
  R dedx0 = air_stop_q(T, energy, g);
  // the sub-steps' decay exponents are summed and applied once at the end
  // (one exponential per walk instead of one per sub-step)
  R xdecay = R(0.0);
  while (val(remaining) > 1e-8) {
    if (!isfinite(val(dedx0)) || val(dedx0) <= 0.0) { weight = R(0.0); return false; }
    R grel = qdiv(MAX_REL_GAIN * cmax(energy, 1e-6), dedx0);
    const R cap = cmin(remaining, MAX_STEP_M);
    R alt0 = z0 + cos_t * (traversed + 0.5 * cap);
    R dens0 = 1.205 * exp(alt0 * (-1.0 / 12e3));
    // ds = cmax(cmin(cap, grel / dens0), 1e-3); where that is cap itself (the
    // usual case: the 60 m cap binds), the mid-step altitude below is alt0
    // and its density dens0, so the exponential is not evaluated again
    const R lim = qdiv(grel, cmax(dens0, 1e-12));
    const bool capped = val(cap) <= val(lim);
    R ds = capped ? cap : lim;
    const bool floored = !(val(ds) >= 1e-3);
    ds = cmax(ds, 1e-3);
    R density = dens0;
    if (!capped || floored) {
      R alt = z0 + cos_t * (traversed + 0.5 * ds);
      density = 1.205 * exp(alt * (-1.0 / 12e3));
    }
    R gram = ds * density;
    R epred = energy + dedx0 * gram;
    R dedx1 = air_stop_q(T, epred, g);
    R enext = energy + 0.5 * (dedx0 + dedx1) * gram;
    if (!isfinite(val(enext)) || val(enext) >= 1e12) { energy = enext; weight = R(0.0); return false; }
    R dnext = air_stop_q(T, enext, g);
    weight = weight * qdiv(dnext, cmax(dedx0, 1e-30));
    xdecay = xdecay + decay_exponent_q(T, ds, 0.5 * (energy + enext));
    energy = enext;
    traversed = traversed + ds;
    remaining = remaining - ds;
    dedx0 = dnext;
  }
  weight = weight * exp(-xdecay);
  return true;
}

struct WArgs {
  const double* tab;
  int K, pow2, shared;
  int air_shared;  // the air table is on the rock's energy grid
  int rock, water, air, porous;
  WMat mat;
  WCorr corr;
  double det_elev, palt, mass, ctau;
  int n_rays, max_seg, n_cells;
  const int32_t* row_ptr;
  const int32_t* cell;
  const double* len;
  const double* cos_t;
  const double* remaining;
  const double* exit_z;
  const uint8_t* valid;
  const double* w;
  const uint8_t* shallow;
  int n_nodes;
  const double* ne;
  const double* nw;
  const double* nq;
  double* flux;
  double* jac;
  double* scratch;
  int* queue;
  int* err;
};

// Columns (of K doubles) of LDS table staging: (E, log E, S_csda) for rock,
// water, porous and air, plus (S_mixed, Omega) for rock, water and porous with
// the corrections.  Materials on the rock's energy grid (sh: water and porous;
// ash: air) use the rock's E and log E columns, and without a porous material
// (pa) the porous columns are the rock's, so a block stages fewer columns and
// more blocks fit in a CU's LDS.
__host__ __device__ constexpr int w_tab_cols(bool corr, bool sh, bool ash, bool pa) {
  return 3 + (sh ? 1 : 3) + (pa ? 0 : (sh ? 1 : 3)) + (ash ? 1 : 3) + (corr ? (pa ? 4 : 6) : 0);
}
// Tape doubles per segment and lane.
__host__ __device__ constexpr int w_tape(bool corr) { return corr ? 8 : 4; }

// Half-wave sum (lanes 0-31 and 32-63 separately).
__device__ __forceinline__ double half_sum(double v) {
#pragma unroll
  for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// PACK (at most 32 energy nodes): a 64-lane block carries two bins, one per
// half-wave; each half sums its nodes with half_sum, which equals the
// full-wave sum of the unpacked layout (whose upper half only adds zeros).
// The corrected Jacobian instantiation asks for 2 waves per SIMD (a register
// budget of 256 per lane; unbounded it takes 258 and runs 1 wave/SIMD).  The
// others keep the default bound, which measured faster for the plain Jacobian.
template <bool JAC, bool CORR, bool PACK>
__global__ __launch_bounds__(256, (JAC && CORR) ? 2 : 1) void wcsda_kernel(WArgs a) {
  HIP_DYNAMIC_SHARED(double, lds)
  const int K = a.K;
  const int bd = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int grp = PACK ? (tid >> 5) : 0;      // bin slot of this lane
  const int gt = PACK ? (tid & 31) : tid;     // lane within the slot
  const int gsize = PACK ? 32 : bd;           // lanes per bin
  const int nwv = PACK ? 2 : (bd >> 6);       // partial-sum rows in LDS
  const int rows = PACK ? 1 : nwv;            // rows per bin
  const int row0 = PACK ? grp : 0;            // first row of this lane's bin
  const int wrow = PACK ? grp : wave;         // row this lane's wave (or half) adds to
  const int wl = PACK ? gt : lane;            // lane index within that wave (or half)
  constexpr int NT = w_tape(CORR);
  const bool sh = a.shared != 0, ash = a.air_shared != 0, pa = a.porous < 0;
  double* wsum = lds + w_tab_cols(CORR, sh, ash, pa) * K; // [nwv][max_seg]
  double* wflux = wsum + (size_t)nwv * a.max_seg;         // [nwv]
  int* scell = (int*)(wflux + nwv) + grp * a.max_seg;     // [slots][max_seg]
  __shared__ int s_ray[2];
  __shared__ double s_G[2], s_M[2], s_q[2], s_mass[2];    // per-bin multipliers (CORR)
  // column of each (material, quantity); a material on the rock's grid points
  // its E and log E at the rock's columns, an absent porous material (staged
  // as rock) every column
  int ce[4], cle[4], cs[4], cm[3], co[3];
  {
    int nc = 0;
    for (int j = 0; j < 4; ++j) {
      const bool same = j == 2 && pa;
      const bool alias = same || (j > 0 && (j < 3 ? sh : ash));
      ce[j] = alias ? ce[0] : nc++;
      cle[j] = alias ? cle[0] : nc++;
      cs[j] = same ? cs[0] : nc++;
    }
    for (int j = 0; j < 3; ++j) {
      const bool same = j == 2 && pa;
      cm[j] = !CORR ? 0 : same ? cm[0] : nc++;
      co[j] = !CORR ? 0 : same ? co[0] : nc++;
    }
  }
  {
    const int mats[4] = {a.rock, a.water, a.porous >= 0 ? a.porous : a.rock, a.air};
    for (int j = 0; j < 4; ++j) {
      const double* t = a.tab + (size_t)mats[j] * 19 * K;
      const bool own = j == 0 || ce[j] != ce[0];
      const bool own_s = j == 0 || cs[j] != cs[0];
      for (int i = tid; i < K; i += bd) {
        if (own) {
          lds[(size_t)ce[j] * K + i] = t[i];           // energies
          lds[(size_t)cle[j] * K + i] = t[K + i];      // log energies
        }
        if (own_s) lds[(size_t)cs[j] * K + i] = t[2 * K + i];  // CSDA stopping power
      }
    }
    if (CORR) {
      for (int j = 0; j < (pa ? 2 : 3); ++j) {
        const double* t = a.tab + (size_t)mats[j] * 19 * K;
        for (int i = tid; i < K; i += bd) {
          lds[(size_t)cm[j] * K + i] = t[5 * K + i];   // MIXED stopping power
          lds[(size_t)co[j] * K + i] = t[7 * K + i];   // straggling
        }
      }
    }
  }
  WTab T;
  T.rE = lds + ce[0] * K; T.rLE = lds + cle[0] * K; T.rS = lds + cs[0] * K;
  T.wE = lds + ce[1] * K; T.wLE = lds + cle[1] * K; T.wS = lds + cs[1] * K;
  T.pE = lds + ce[2] * K; T.pLE = lds + cle[2] * K; T.pS = lds + cs[2] * K;
  T.aE = lds + ce[3] * K; T.aLE = lds + cle[3] * K; T.aS = lds + cs[3] * K;
  T.rM = lds + cm[0] * K; T.rO = lds + co[0] * K;
  T.wM = lds + cm[1] * K; T.wO = lds + co[1] * K;
  T.pM = lds + cm[2] * K; T.pO = lds + co[2] * K;
  T.K = K; T.pow2 = a.pow2; T.shared = a.shared; T.mass = a.mass; T.ctau = a.ctau;
  const WMat M = a.mat;
  const WCorr C = a.corr;
  double* sc = JAC ? a.scratch + (size_t)blockIdx.x * a.max_seg * NT * bd : nullptr;

  for (;;) {
    if (gt == 0) s_ray[grp] = atomicAdd(a.queue, 1);
    __syncthreads();
    const int r = s_ray[grp];
    if (PACK ? (s_ray[0] >= a.n_rays && s_ray[1] >= a.n_rays) : r >= a.n_rays) break;
    // every lane of the block meets the same barriers: a slot without a usable
    // bin carries live = false (ns = 0) through the iteration
    bool live = r < a.n_rays;
    int b0 = 0, ns = 0;
    if (live) {
      b0 = a.row_ptr[r];
      ns = a.row_ptr[r + 1] - b0;
      if (ns < 0 || ns > a.max_seg) {
        if (gt == 0) { atomicOr(a.err, ERR_CAP); a.flux[r] = __builtin_nan(""); }
        live = false;
        ns = 0;
      }
    }
    const double ct = live ? a.cos_t[r] : 1.0;
    if (live && !((a.valid ? a.valid[r] != 0 : true) && ct > 0.0)) {
      if (gt == 0) a.flux[r] = __builtin_nan("");
      if (JAC) for (int s = gt; s < ns; s += gsize) a.jac[b0 + s] = 0.0;
      live = false;
      ns = 0;
    }
    for (int s = gt; s < ns; s += gsize) {
      int c = a.cell[b0 + s];
      if (c < 0 || c >= a.n_cells) { atomicOr(a.err, ERR_CELL); c = -1; }
      scell[s] = c;
    }
    if (JAC) for (int s = wl; s < ns; s += (PACK ? 32 : 64)) wsum[wrow * a.max_seg + s] = 0.0;
    if (wl == 0) wflux[wrow] = 0.0;
    __syncthreads();

    const double rem = live ? a.remaining[r] : 0.0;
    const double zex = live ? a.exit_z[r] : 0.0;
    const double air_d = fmax(0.0, a.palt - zex) / fmax(ct, 1e-3);
    const double z0 = a.det_elev + zex;
    if (CORR && live && gt == 0) {
      // csda_geom_multiplier (path_rock_length) and _column_water_fraction,
      // summed in path order
      double L = rem, water = 0.0, mass = rem * M.rho_r;
      bool bad = false;
      for (int s = 0; s < ns; ++s) {
        const double d = a.len[b0 + s];
        L += d;
        const int c = scell[s];
        if (c < 0) { bad = true; continue; }
        const double wk = a.w[c];
        const double gx = d * cell_density(M, wk, M.porous_on && a.shallow[c] != 0);
        water += wk * gx;
        mass += gx;
      }
      double G = 1.0;
      if (C.enabled && C.geom) {
        const double lL = log(fmax(L, 1.0));
        G = exp(C.ra + C.rb * lL + C.rc * lL * lL + C.rd * ct);
      }
      const double q = bad ? 0.0 : water / fmax(mass, 1e-30);
      s_G[grp] = G;
      s_q[grp] = q;
      s_mass[grp] = mass;
      s_M[grp] = C.mat ? exp(C.lin * q + C.quad * q * q) : 1.0;
    }

    for (int n0 = 0; n0 < a.n_nodes; n0 += gsize) {
      const int n = n0 + gt;
      const bool active = live && n < a.n_nodes;
      double E = active ? a.ne[n] : a.ne[0];
      double W = 1.0, V = 0.0, H = 0.0;
      int gc = K / 2, ga = K / 2;  // interval guesses: cell/rock and air stopping power
      const bool acc = CORR && C.enabled;
      bool ok = active;  // lanes past the last sample walk nothing and add 0
      // ------------------------------------------------ forward sweep
      // The next segment's water fraction, porous flag and length are
      // loaded one segment ahead (the gathers from the field hide behind the
      // current segment's sub-steps).
      double wv_nx = 0.0, d_nx = 0.0;
      bool por_nx = false;
      if (ns > 0 && scell[0] >= 0) {
        wv_nx = a.w[scell[0]];
        por_nx = M.porous_on && a.shallow[scell[0]] != 0;
        d_nx = a.len[b0];
      }
      for (int s = 0; s < ns && ok; ++s) {
        const int c = scell[s];
        if (c < 0) { ok = false; break; }
        const double wv = wv_nx;
        const bool porous = por_nx;
        const double d = d_nx;
        if (s + 1 < ns && scell[s + 1] >= 0) {
          const int cn = scell[s + 1];
          wv_nx = a.w[cn];
          por_nx = M.porous_on && a.shallow[cn] != 0;
          d_nx = a.len[b0 + s + 1];
        }
        if (JAC) {
          const D2 wd(wv, 1.0, 0.0);
          D2 e(E, 0.0, 1.0), g(W);
          ok = cell_segment(T, M, wd, porous, d, e, g, gc);
          double* p = sc + (size_t)s * NT * bd + tid;
          p[0] = e.a;
          p[bd] = e.b;
          p[2 * bd] = g.a / g.v;
          p[3 * bd] = g.b / g.v;
          if (CORR) {
            double vw = 0.0, ve = 0.0, hw = 0.0, he = 0.0;
            if (acc && ok) {
              D2 er = 0.5 * (D2(E, 0.0, 1.0) + e);
              D2 om, hl;
              cell_strag_hard(T, wd, porous, er, om, hl);
              D2 gx = d * cell_density(M, wd, porous);
              D2 v = om * gx, h = hl * gx;
              V += v.v; H += h.v;
              vw = v.a; ve = v.b; hw = h.a; he = h.b;
            }
            p[4 * bd] = vw;
            p[5 * bd] = ve;
            p[6 * bd] = hw;
            p[7 * bd] = he;
          }
          E = e.v;
          W = g.v;
        } else {
          const double e_in = E;
          ok = cell_segment(T, M, wv, porous, d, E, W, gc);
          if (CORR && acc && ok) {
            const double er = 0.5 * (e_in + E);
            double om, hl;
            cell_strag_hard(T, wv, porous, er, om, hl);
            const double gx = d * cell_density(M, wv, porous);
            V += om * gx;
            H += hl * gx;
          }
        }
      }
      double rc = 1.0, rl = 0.0, ac = 1.0, al = 0.0, rv = 0.0, rh = 0.0;
      if (ok && rem > 1e-8) {
        const double gx = rem * M.rho_r;
        if (JAC) {
          DualD e(E, 1.0), g(W);
          ok = rock_segment(T, M.rho_r, rem, e, g, gc);
          rc = e.d; rl = g.d / g.v;
          if (CORR && acc && ok) {
            DualD er = 0.5 * (DualD(E, 1.0) + e);
            DualD om, hl;
            rock_strag_hard(T, er, om, hl);
            V += om.v * gx; H += hl.v * gx;
            rv = om.d * gx; rh = hl.d * gx;
          }
          E = e.v; W = g.v;
        } else {
          const double e_in = E;
          ok = rock_segment(T, M.rho_r, rem, E, W, gc);
          if (CORR && acc && ok) {
            const double er = 0.5 * (e_in + E);
            double om, hl;
            rock_strag_hard(T, er, om, hl);
            V += om * gx;
            H += hl * gx;
          }
        }
      }
      if (ok) {
        if (JAC) {
          DualD e(E, 1.0), g(W);
          ok = air_segment(T, z0, ct, air_d, e, g, ga);
          ac = e.d; al = g.d / g.v; E = e.v; W = g.v;
        } else {
          ok = air_segment(T, z0, ct, air_d, E, W, ga);
        }
      }
      double f = 0.0, Tad = 0.0, AV = 0.0, AH = 0.0;
      if (active && ok) {
        double cv = 1.0, AE = 0.0;
        if (CORR && acc) {
          if (JAC) {
            const DualD lc = log_var_corr(DualD(E, 1.0), V, H, ct, C);
            cv = exp(lc.v);
            AE = lc.d;
            const double e = fmax(E, 1.0);
            const double eta = 1.0e-2;
            const double bi = -(log(flux_gccly_r(ct, e * (1.0 + eta), 1.0)) -
                                log(flux_gccly_r(ct, e * (1.0 - eta), 1.0))) / (2 * eta);
            AV = C.ks * (0.5 * bi * (bi + 1.0)) / (e * e);
            AH = H >= 0.0 ? -C.kh / e : 0.0;
          } else {
            cv = exp(log_var_corr(E, V, H, ct, C));
          }
        }
        if (JAC) {
          Dd phi = flux_gccly_d(ct, E, a.nq[n]);
          f = 2.0 * a.nw[n] * (W * cv * phi.v);
          Tad = phi.d / phi.v + AE;
          Tad = al + Tad * ac;
          if (rem > 1e-8) Tad = rl + Tad * rc + AV * rv + AH * rh;
        } else {
          f = 2.0 * a.nw[n] * (W * cv * flux_gccly(ct, E, a.nq[n]));
        }
      }
      const double fs = PACK ? half_sum(f) : wave_sum(f);
      if (wl == 0) wflux[wrow] += fs;
      if (JAC) {
        // ---------------------------------------------- reverse sweep
        for (int s = ns - 1; s >= 0; --s) {
          const double* p = sc + (size_t)s * NT * bd + tid;
          double v = 0.0;
          if (f != 0.0) {
            if (CORR) {
              v = f * (p[2 * bd] + Tad * p[0] + AV * p[4 * bd] + AH * p[6 * bd]);
              Tad = p[3 * bd] + Tad * p[bd] + AV * p[5 * bd] + AH * p[7 * bd];
            } else {
              v = f * (p[2 * bd] + Tad * p[0]);
              Tad = p[3 * bd] + Tad * p[bd];
            }
          }
          const double vs = PACK ? half_sum(v) : wave_sum(v);
          if (wl == 0) wsum[wrow * a.max_seg + s] += vs;
        }
      }
    }
    __syncthreads();
    double fsum = 0.0;
    for (int j = 0; j < rows; ++j) fsum += wflux[row0 + j];
    if (JAC) {
      // total per segment, then repeated cells merged into the first occurrence
      for (int s = gt; s < ns; s += gsize) {
        double v = 0.0;
        for (int j = 0; j < rows; ++j) v += wsum[(row0 + j) * a.max_seg + s];
        if (CORR) {
          v = s_G[grp] * s_M[grp] * (v / a.n_nodes);
          const int c = scell[s];
          if (C.mat && c >= 0) {
            // d q / d w_s of the column water fraction, times M'(q)
            const double wk = a.w[c];
            const DualD gx = a.len[b0 + s] * cell_density(M, DualD(wk, 1.0),
                                                          M.porous_on && a.shallow[c] != 0);
            const double dwater = gx.v + wk * gx.d;
            const double dq = s_mass[grp] >= 1e-30 ? (dwater - s_q[grp] * gx.d) / s_mass[grp]
                                                   : dwater / 1e-30;
            const double dM = s_M[grp] * (C.lin + 2.0 * C.quad * s_q[grp]);
            v += s_G[grp] * dM * dq * (fsum / a.n_nodes);
          }
          wsum[row0 * a.max_seg + s] = v;
        } else {
          // the bin's first row holds the totals; each s is read then written by one thread
          wsum[row0 * a.max_seg + s] = v / a.n_nodes;
        }
      }
      __syncthreads();
      for (int s = gt; s < ns; s += gsize) {
        const int c = scell[s];
        bool first = true;
        for (int t = 0; t < s; ++t)
          if (scell[t] == c) { first = false; break; }
        double v = 0.0;
        if (first) {
          v = wsum[row0 * a.max_seg + s];
          for (int t = s + 1; t < ns; ++t)
            if (scell[t] == c) v += wsum[row0 * a.max_seg + t];
        }
        a.jac[b0 + s] = v;
      }
    }
    if (live && gt == 0)
      a.flux[r] = CORR ? s_G[grp] * s_M[grp] * fsum / a.n_nodes : fsum / a.n_nodes;
    __syncthreads();
  }
}

// Clears the work queue and, for the host entry point, its error word (the
// _dev entry points accumulate errors in the context's sticky word).
__global__ void reset_counters_w(int* queue, int* err, int reset_err) {
  if (threadIdx.x == 0) {
    *queue = 0;
    if (reset_err) *err = 0;
  }
}

size_t w_lds_bytes(int K, int nwv, int max_seg, bool corr, bool sh, bool ash, bool pa,
                   int slots) {
  return ((size_t)w_tab_cols(corr, sh, ash, pa) * K + (size_t)nwv * max_seg + nwv) * sizeof(double) +
         (size_t)slots * max_seg * sizeof(int);
}

int check_w(const char* fn, dp_tables t, const dp_material_config* m, const dp_site_config* site,
            const dp_directional_paths* p, int32_t n_cells, const double* w,
            const uint8_t* shallow, const dp_energy_samples* es, double* flux) {
  std::string f(fn);
  if (!t || !m || !site || !p || !es || !flux || !w || !shallow)
    return dph::fail(DP_E_ARGUMENT, f + ": null argument");
  if (p->n_rays < 0 || p->n_segments < 0 || p->max_segments < 0)
    return dph::fail(DP_E_ARGUMENT, f + ": negative sizes");
  if (es->n <= 0) return dph::fail(DP_E_ARGUMENT, f + ": need at least one energy sample");
  if (n_cells <= 0) return dph::fail(DP_E_ARGUMENT, f + ": need at least one cell");
  const int M = t->n_materials;
  if (m->rock_material < 0 || m->rock_material >= M || m->water_material < 0 ||
      m->water_material >= M || m->air_material < 0 || m->air_material >= M ||
      m->porous_material >= M)
    return dph::fail(DP_E_ARGUMENT, f + ": material index out of range");
  if (!(m->rock_density > 0) || !(m->water_density > 0) ||
      (m->porous_material >= 0 && m->porous_top_thickness > 0 && !(m->porous_density > 0)))
    return dph::fail(DP_E_ARGUMENT, f + ": densities must be positive");
  return DP_OK;
}

bool same_grid(dp_tables t, int a, int b) {
  const size_t K = t->n_knots;
  const double* x = t->host_data.data() + (size_t)a * 19 * K;
  const double* y = t->host_data.data() + (size_t)b * 19 * K;
  return std::memcmp(x, y, 2 * K * sizeof(double)) == 0;
}

int launch_w(dp_ctx c, dp_tables t, const dp_material_config* m, const dp_site_config* site,
             const dp_directional_paths* p, int32_t n_cells, const double* w,
             const uint8_t* shallow, const dp_energy_samples* es, double* flux, double* jac,
             int* queue, int* err, bool reset_err) {
  if (p->n_rays == 0) return DP_OK;
  const int bd = es->n >= 256 ? 256 : ((es->n + 63) / 64) * 64;
  const bool PK = es->n <= 32;   // two bins per 64-lane block
  const int nwv = PK ? 2 : bd / 64;
  const int max_seg = p->max_segments > 0 ? p->max_segments : 1;
  const bool J = jac != nullptr;
  const bool CR = c->corr.enabled || c->mcorr.enabled;
  const int porous = m->porous_material;
  const bool sh = same_grid(t, m->rock_material, m->water_material) &&
                  (porous < 0 || same_grid(t, m->rock_material, porous));
  const bool ash = same_grid(t, m->rock_material, m->air_material);
  const size_t lds = w_lds_bytes(t->n_knots, nwv, max_seg, CR, sh, ash, porous < 0, PK ? 2 : 1);
  if (lds > 160 * 1024)
    return dph::fail(DP_E_CAPACITY, "dp_csda_w_flux_jacobian: max_segments " +
                                        std::to_string(max_seg) + " exceeds the LDS budget");
  auto kern = PK ? (J ? (CR ? wcsda_kernel<true, true, true> : wcsda_kernel<true, false, true>)
                      : (CR ? wcsda_kernel<false, true, true> : wcsda_kernel<false, false, true>))
                 : (J ? (CR ? wcsda_kernel<true, true, false> : wcsda_kernel<true, false, false>)
                      : (CR ? wcsda_kernel<false, true, false> : wcsda_kernel<false, false, false>));
  int occ = 0;
  DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, bd, lds));
  if (occ < 1) occ = 1;
  int grid = c->n_cu * occ;
  const int bins_per_block = PK ? 2 : 1;
  if (grid > (p->n_rays + bins_per_block - 1) / bins_per_block)
    grid = (p->n_rays + bins_per_block - 1) / bins_per_block;
  WArgs a;
  a.scratch = nullptr;
  if (J) {
    int rc = dph::ensure(c->scratch, (size_t)grid * max_seg * w_tape(CR) * bd * sizeof(double));
    if (rc) return rc;
    a.scratch = (double*)c->scratch.ptr;
  }
  const dp_csda_correction& k = c->corr;
  a.corr.enabled = k.enabled != 0;
  a.corr.geom = k.resid_a != 0.0 || k.resid_b != 0.0 || k.resid_c != 0.0 || k.resid_d != 0.0;
  a.corr.mat = c->mcorr.enabled != 0;
  a.corr.ks = k.kappa_strag; a.corr.kh = k.kappa_hard;
  a.corr.ra = k.resid_a; a.corr.rb = k.resid_b; a.corr.rc = k.resid_c; a.corr.rd = k.resid_d;
  a.corr.lin = c->mcorr.linear; a.corr.quad = c->mcorr.quadratic;
  a.tab = t->data; a.K = t->n_knots; a.pow2 = t->pow2;
  a.rock = m->rock_material; a.water = m->water_material; a.air = m->air_material;
  a.porous = porous;
  a.shared = sh;
  a.air_shared = ash;
  a.mat.rho_r = m->rock_density; a.mat.rho_w = m->water_density; a.mat.rho_p = m->porous_density;
  a.mat.porous_on = porous >= 0 && m->porous_top_thickness > 0.0;
  a.det_elev = site->detector_elevation; a.palt = site->primary_altitude;
  a.mass = t->mass; a.ctau = t->ctau;
  a.n_rays = p->n_rays; a.max_seg = max_seg; a.n_cells = n_cells;
  a.row_ptr = p->row_ptr; a.cell = p->cell; a.len = p->length; a.cos_t = p->cos_theta;
  a.remaining = p->remaining_rock; a.exit_z = p->exit_z; a.valid = p->valid;
  a.w = w; a.shallow = shallow;
  a.n_nodes = es->n; a.ne = es->energy; a.nw = es->weight; a.nq = es->charge;
  a.flux = flux; a.jac = jac;
  a.queue = queue; a.err = err;
  hipLaunchKernelGGL(reset_counters_w, dim3(1), dim3(64), 0, c->stream, queue, err, reset_err ? 1 : 0);
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(bd), lds, c->stream, a);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  return DP_OK;
}

}  // namespace

extern "C" {

int dp_set_csda_correction(dp_ctx c, const dp_csda_correction* k) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_set_csda_correction: null context");
  c->corr = k ? *k : dp_csda_correction{};
  return DP_OK;
}

int dp_set_csda_material_correction(dp_ctx c, const dp_csda_material_correction* k) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_set_csda_material_correction: null context");
  if (k && !(std::isfinite(k->linear) && std::isfinite(k->quadratic)))
    return dph::fail(DP_E_ARGUMENT, "dp_set_csda_material_correction: nonfinite material coefficients");
  c->mcorr = k ? *k : dp_csda_material_correction{};
  return DP_OK;
}

int dp_csda_w_flux_jacobian_dev(dp_ctx c, dp_tables t, const dp_material_config* m,
                                const dp_site_config* site, const dp_directional_paths* p,
                                int32_t n_cells, const double* w, const uint8_t* shallow,
                                const dp_energy_samples* es, double* flux, double* jac) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian_dev: null context");
  int rc = check_w("dp_csda_w_flux_jacobian_dev", t, m, site, p, n_cells, w, shallow, es, flux);
  if (rc) return rc;
  DP_HIP(hipSetDevice(c->device));
  int* q = (int*)c->counter.ptr;
  return launch_w(c, t, m, site, p, n_cells, w, shallow, es, flux, jac, q + dph::COUNTER_QUEUE,
                  q + dph::COUNTER_STICKY, false);
}

int dp_csda_w_flux_jacobian(dp_ctx c, dp_tables t, const dp_material_config* m,
                            const dp_site_config* site, const dp_directional_paths* p,
                            int32_t n_cells, const double* w, const uint8_t* shallow,
                            const dp_energy_samples* es, double* flux, double* jac) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian: null context");
  int rc = check_w("dp_csda_w_flux_jacobian", t, m, site, p, n_cells, w, shallow, es, flux);
  if (rc) return rc;
  const int nr = p->n_rays, ns = p->n_segments, nn = es->n;
  if (nr == 0) return DP_OK;
  if (p->row_ptr[0] != 0 || p->row_ptr[nr] != ns)
    return dph::fail(DP_E_ARGUMENT,
                     "dp_csda_w_flux_jacobian: row_ptr must start at 0 and end at n_segments");
  int mx = 0;
  for (int r = 0; r < nr; ++r) {
    const int k = p->row_ptr[r + 1] - p->row_ptr[r];
    if (k < 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian: row_ptr must be non-decreasing");
    mx = k > mx ? k : mx;
  }
  for (int s = 0; s < ns; ++s)
    if (p->cell[s] < 0 || p->cell[s] >= n_cells)
      return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian: cell index out of range at segment " +
                                          std::to_string(s));
  for (int i = 0; i < nn; ++i)
    if (!(es->energy[i] > 0))
      return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian: sample energies must be positive");
  DP_HIP(hipSetDevice(c->device));
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t o_rp = take(4 * (size_t)(nr + 1)), o_cell = take(4 * (size_t)ns),
               o_len = take(8 * (size_t)ns), o_cos = take(8 * (size_t)nr),
               o_rem = take(8 * (size_t)nr), o_zex = take(8 * (size_t)nr),
               o_val = take((size_t)nr), o_w = take(8 * (size_t)n_cells),
               o_sh = take((size_t)n_cells), o_ne = take(8 * (size_t)nn),
               o_nw = take(8 * (size_t)nn), o_nq = take(8 * (size_t)nn),
               o_flux = take(8 * (size_t)nr), o_jac = take(8 * (size_t)ns), o_q = take(256);
  rc = dph::ensure(c->staging, o);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_rp, p->row_ptr, 4 * (size_t)(nr + 1), hipMemcpyHostToDevice, st));
  if (ns) {
    DP_HIP(hipMemcpyAsync(b + o_cell, p->cell, 4 * (size_t)ns, hipMemcpyHostToDevice, st));
    DP_HIP(hipMemcpyAsync(b + o_len, p->length, 8 * (size_t)ns, hipMemcpyHostToDevice, st));
  }
  DP_HIP(hipMemcpyAsync(b + o_cos, p->cos_theta, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_rem, p->remaining_rock, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_zex, p->exit_z, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  if (p->valid) DP_HIP(hipMemcpyAsync(b + o_val, p->valid, (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_w, w, 8 * (size_t)n_cells, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_sh, shallow, (size_t)n_cells, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_ne, es->energy, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_nw, es->weight, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_nq, es->charge, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  dp_directional_paths dp = *p;
  dp.max_segments = mx;
  dp.row_ptr = (int32_t*)(b + o_rp); dp.cell = (int32_t*)(b + o_cell);
  dp.length = (double*)(b + o_len); dp.cos_theta = (double*)(b + o_cos);
  dp.remaining_rock = (double*)(b + o_rem); dp.exit_z = (double*)(b + o_zex);
  dp.valid = p->valid ? (uint8_t*)(b + o_val) : nullptr;
  dp_energy_samples de = *es;
  de.energy = (double*)(b + o_ne); de.weight = (double*)(b + o_nw); de.charge = (double*)(b + o_nq);
  int* qerr = (int*)(b + o_q);
  rc = launch_w(c, t, m, site, &dp, n_cells, (double*)(b + o_w), (uint8_t*)(b + o_sh), &de,
                (double*)(b + o_flux), jac ? (double*)(b + o_jac) : nullptr, qerr, qerr + 1, true);
  if (rc) return rc;
  int herr[2] = {0, 0};
  DP_HIP(hipMemcpyAsync(flux, b + o_flux, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  if (jac && ns) DP_HIP(hipMemcpyAsync(jac, b + o_jac, 8 * (size_t)ns, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(herr, qerr, 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (herr[1] & ERR_CAP) return dph::fail(DP_E_CAPACITY, "dp_csda_w_flux_jacobian: a path exceeded max_segments");
  if (herr[1] & ERR_CELL) return dph::fail(DP_E_ARGUMENT, "dp_csda_w_flux_jacobian: cell index out of range");
  return DP_OK;
}

}  // extern "C"
