// Per-voxel-density CSDA flux and its exact reverse-mode derivative.
//
// Work decomposition: one persistent block per ray at a time, taken from a
// work queue; the block's lanes are the energy nodes of that ray (all nodes
// walk the same segment list, so segment data and densities are wave-uniform
// loads).  Rock and air tables (energy, log energy, range, stopping power)
// are staged in LDS.
//
// Forward sweep per lane: the CSDA segment map of compute_flux_from_segments
// (examples/diff_flux_cube.jl:386-413).  For segment v with E_in -> E_out
//   c_v = dE_out/dE_in,  r_v = dE_out/drho_v,
//   b_v = dlog w_v/dE_in, e_v = dlog w_v/dE_out      (w_v = S(E_out)/S(E_in) * decay)
// define the reverse recursion  a* = T + e_v;  d flux/d rho_v = flux a* r_v;
// T <- a* c_v + b_v  with T = d log(flux)/dE at the segment's exit.  The tail
// (remaining rock at rho[0], air, flux_gccly) is differentiated in
// registers.  The recursion runs in prefix form during the forward sweep:
// two tape values per segment and node, summed over nodes once the node's
// flux is known (see the kernel), so there is no backward sweep.  These are
// the derivatives
// of the executed discrete program (bracket slopes; zero where a lookup is
// clamped), i.e. what Zygote returns on the reference.  Per-segment sums over
// nodes use a fixed wave/LDS tree, so results are deterministic.
#include "dp_internal.hpp"
#include <map>
#include <mutex>
#include "dp_device.hpp"
#include "dp_sort.hpp"
#include <climits>
#include <type_traits>

using namespace dpk;

namespace {

struct CsdaArgs {
  const double* tab;
  int K, pow2;
  int rock, air;
  double mass, ctau;
  int n_rays, max_seg;
  const int32_t* row_ptr;
  const int32_t* cell;
  const double* len;
  const double* cos_theta;
  const double* z_exit;
  double z_max, palt, scale;
  int n_nodes;
  const double* ne;
  const double* nw;
  const double* nq;
  const double* dens;
  double* flux;
  double* jac;
  double* jac_rem;
  double* scratch;
  int* queue;
  int per_pop;     // rays a block takes from the queue at a time (tuned)
  int* err;
};

// F: the type of the segment arithmetic (double; float for the FP32 stepping
// variant).  The log energy, the state the walk carries, is FP64 in both.
template <typename F>
struct StepOutT {
  F E;          // exit energy
  double lE;    // its log
  F ratio;      // S(E_out)/S(E_in)
  double dexp;  // decay exponent of the segment: decay factor exp(-dexp)
  F c, b, e;    // partials (see header)
  F k;          // dE_out/dX (X = grammage)
  bool ok;
};

// Entry-side lookups of a segment: range and stopping power at the entry
// energy with their slopes, and 1/S.  A rock segment's exit computes them
// for the next rock segment (same tables, same energy), which then skips
// its own entry lookup.
struct Entry64 {
  double y0, y1, dy0, dy1, ry1;
  bool valid;
};
using StepOut = StepOutT<double>;

// Table views of one material.  FP64: energy, log energy, range, stopping
// power.  FP32 stepping: log energy and range in FP64 with the reciprocals of
// their knot spacings (the FP64 spine of the walk), stopping power in float.
struct Tab64 {
  const double *E, *LE, *R, *S;
  const double *iLE, *iR;   // knot-spacing reciprocals (staged tables), or null
};
struct Tab32 {
  const double *LE, *R, *iLE, *iR;
  const float* S;
};
constexpr double LOG_1E12 = 27.631021115928547;   // log(1e12): the "ok" energy bound

// One CSDA segment: range inversion at grammage X_in(E) + gram, then the
// stopping-power ratio and the decay factor at the mean energy.  The interval
// searches start from what the walk already knows: ie is the interval of E
// on the log-energy grid (-1: unknown or E clamped), jr the last interval
// found on the range grid; locate_idx returns the interval the bisection
// would.  On return ie is the interval of the exit energy, which is the next
// segment's entry energy on the same tables.  Quotients that share a divisor
// use one reciprocal (FP64 division costs ~10 quarter-rate instructions on
// gfx950): values agree with the division form to a few ulp.
__device__ __forceinline__ double inv_step(const double* inv, const double* x, int i) {
  return inv ? inv[i] : 1.0 / (x[i + 1] - x[i]);
}

// `en` carries the entry lookups from the previous rock segment (en.valid)
// and returns this segment's exit lookups for the next one.  The exit log
// energy is the interpolated value whose exponential is the exit energy
// (not re-rounded through log).  Reciprocals that a quantity shares are
// taken once; the decay factor is returned as its exponent so the walk
// multiplies one exponential per node.
__device__ __forceinline__ StepOut csda_segment_g(const Tab64& T, int K, int pow2, double E,
                                                  double lE, double gram, double path,
                                                  double inv_mass, double ctau, double inv_ctau,
                                                  int& ie, int& jr, Entry64& en) {
  const double *sE = T.E, *sLE = T.LE, *sR = T.R, *sS = T.S;
  StepOut o;
  if (!en.valid) {
    if (E <= sE[0]) {
      en.y0 = sR[0]; en.y1 = sS[0]; en.dy0 = en.dy1 = 0.0;
    } else if (E >= sE[K - 1]) {
      en.y0 = sR[K - 1]; en.y1 = sS[K - 1]; en.dy0 = en.dy1 = 0.0;
    } else {
      const int i = ie >= 0 ? ie : locate_idx(sLE, K, pow2, lE, jr);
      const double rdl = inv_step(T.iLE, sLE, i);
      const double t = (lE - sLE[i]) * rdl;
      const double d0 = sR[i + 1] - sR[i], d1 = sS[i + 1] - sS[i];
      en.y0 = sR[i] + t * d0;
      en.y1 = sS[i] + t * d1;
      const double re = rdl * rcp_nr(E);
      en.dy0 = d0 * re;
      en.dy1 = d1 * re;
    }
    en.ry1 = rcp_nr(en.y1);
  }
  const double y0 = en.y0, dy0 = en.dy0, dy1 = en.dy1, ry1 = en.ry1;
  const double Xf = y0 + gram;
  double Eo, kdy;
  if (Xf <= sR[0]) {
    Eo = sE[0]; kdy = 0.0; o.lE = sLE[0];
  } else if (Xf >= sR[K - 1]) {
    Eo = sE[K - 1]; kdy = 0.0; o.lE = sLE[K - 1];
  } else {
    const int j = locate_idx(sR, K, pow2, Xf, jr);
    jr = j;
    const double rdr = inv_step(T.iR, sR, j);
    const double t = (Xf - sR[j]) * rdr;
    const double dle = sLE[j + 1] - sLE[j];
    o.lE = sLE[j] + t * dle;
    Eo = exp(o.lE);
    kdy = Eo * (dle * rdr);
  }
  o.E = Eo;
  o.ok = !(Eo >= 1e12) && isfinite(Eo);
  // exit lookups, which are also the next rock segment's entry lookups
  if (Eo <= sE[0]) {
    en.y0 = sR[0]; en.y1 = sS[0]; en.dy0 = en.dy1 = 0.0; ie = -1;
  } else if (Eo >= sE[K - 1]) {
    en.y0 = sR[K - 1]; en.y1 = sS[K - 1]; en.dy0 = en.dy1 = 0.0; ie = -1;
  } else {
    const int i = locate_idx(sLE, K, pow2, o.lE, jr);
    ie = i;
    const double rdl = inv_step(T.iLE, sLE, i);
    const double t = (o.lE - sLE[i]) * rdl;
    const double d0 = sR[i + 1] - sR[i], d1 = sS[i + 1] - sS[i];
    en.y0 = sR[i] + t * d0;
    en.y1 = sS[i] + t * d1;
    const double re = rdl * rcp_nr(Eo);
    en.dy0 = d0 * re;
    en.dy1 = d1 * re;
  }
  const double rso = rcp_nr(en.y1);
  en.ry1 = rso;
  en.valid = true;
  // decay_weight with its log-derivative (compute_decay_weight_from_path,
  // src/Geometry.jl:1083-1107): beta^2 = 1 - 1/gamma^2 floored at 1e-12
  const double rg = rcp_nr(1.0 + 0.5 * (E + Eo) * inv_mass);   // 1/gamma
  const double b2 = 1.0 - rg * rg;
  double rbeta, dbg_dg;   // 1/beta, d(beta*gamma)/d gamma
  if (b2 >= 1e-12) {
    rbeta = rsqrt_nr(b2);
    dbg_dg = rbeta;
  } else {
    rbeta = 1e6;
    dbg_dg = 1e-6;
  }
  const double rL = rbeta * rg * inv_ctau;   // 1 / (beta gamma c tau)
  o.dexp = path * rL;
  const double dlog = path * rL * rL * (ctau * dbg_dg * inv_mass);
  o.ratio = en.y1 * ry1;
  o.k = kdy;
  o.c = kdy * dy0;
  o.b = -dy1 * ry1 + 0.5 * dlog;
  o.e = en.dy1 * rso + 0.5 * dlog;
  return o;
}

// The same segment map with FP32 arithmetic around an FP64 spine.  The walk's
// state is the log energy: the entry range R(E) = R_i + t (R_i+1 - R_i), the
// exit grammage X = R(E) + gram, its interval and the exit log energy
// LE_j + t' (LE_j+1 - LE_j) are FP64 (subtractions and products only, with
// the knot reciprocals precomputed), so rounding does not accumulate along
// the path; the exit log energy is carried as is (log(exp(x)) is not
// re-rounded).  Everything else (exp to the exit energy, stopping powers,
// decay, the partials) is float.
__device__ __forceinline__ StepOutT<float> csda_segment_g(const Tab32& T, int K, int pow2,
                                                         float E, double lE, double gram,
                                                         float path, float inv_mass, float ctau,
                                                         int& ie, int& jr) {
  StepOutT<float> o;
  double y0;
  float dy0, y1, dy1;
  if (lE <= T.LE[0]) {
    y0 = T.R[0]; y1 = T.S[0]; dy0 = dy1 = 0.f;
  } else if (lE >= T.LE[K - 1]) {
    y0 = T.R[K - 1]; y1 = T.S[K - 1]; dy0 = dy1 = 0.f;
  } else {
    const int i = ie >= 0 ? ie : locate_idx(T.LE, K, pow2, lE, jr);
    const double rdl = T.iLE[i];
    const double t = (lE - T.LE[i]) * rdl;
    const double d0 = T.R[i + 1] - T.R[i];
    const float d1 = T.S[i + 1] - T.S[i];
    y0 = T.R[i] + t * d0;
    y1 = T.S[i] + (float)t * d1;
    const float re = (float)rdl * rcpf_nr(E);
    dy0 = (float)d0 * re;
    dy1 = d1 * re;
  }
  const double Xf = y0 + gram;
  double lEo;
  float kq = 0.f;   // dlogE/dX at the exit
  if (Xf <= T.R[0]) {
    lEo = T.LE[0];
  } else if (Xf >= T.R[K - 1]) {
    lEo = T.LE[K - 1];
  } else {
    const int j = locate_idx(T.R, K, pow2, Xf, jr);
    jr = j;
    const double rdr = T.iR[j];
    const double t = (Xf - T.R[j]) * rdr;
    const double dle = T.LE[j + 1] - T.LE[j];
    lEo = T.LE[j] + t * dle;
    kq = (float)(dle * rdr);
  }
  const float Eo = expf((float)lEo);
  const float kdy = Eo * kq;
  o.E = Eo;
  o.lE = lEo;
  o.ok = !(lEo >= LOG_1E12) && isfinite(lEo);
  float so_y, so_dy;
  if (lEo <= T.LE[0]) {
    so_y = T.S[0]; so_dy = 0.f; ie = -1;
  } else if (lEo >= T.LE[K - 1]) {
    so_y = T.S[K - 1]; so_dy = 0.f; ie = -1;
  } else {
    const int i = locate_idx(T.LE, K, pow2, lEo, jr);
    ie = i;
    const double rdl = T.iLE[i];
    const float t = (float)((lEo - T.LE[i]) * rdl);
    const float dyy = T.S[i + 1] - T.S[i];
    so_y = T.S[i] + t * dyy;
    so_dy = dyy * ((float)rdl * rcpf_nr(Eo));
  }
  const float g = 1.f + 0.5f * (E + Eo) * inv_mass;
  const float b2 = 1.f - rcpf_nr(g * g);
  float beta, dbg_dg;
  if (b2 >= 1e-12f) {
    beta = sqrtf(b2);
    dbg_dg = rcpf_nr(beta);
  } else {
    beta = 1e-6f;
    dbg_dg = beta;
  }
  const float rL = rcpf_nr(beta * g * ctau);
  o.dexp = (double)(path * rL);
  const float dlog = path * rL * rL * (ctau * dbg_dg * inv_mass);
  const float ry1 = rcpf_nr(y1), rso = rcpf_nr(so_y);
  o.ratio = so_y * ry1;
  o.k = kdy;
  o.c = kdy * dy0;
  o.b = -dy1 * ry1 + 0.5f * dlog;
  o.e = so_dy * rso + 0.5f * dlog;
  return o;
}

// WAVES: the occupancy bound (waves per SIMD), a tuned knob.
// TLDS: the rock and air tables are staged in LDS; otherwise they are read
// through L1/L2 from the table block.  Both give the same values; the
// autotuner picks one per workload.  F: the type of the segment arithmetic.
// With F = float (the FP32 stepping variant) the segment maps, their
// partials and the tape values are float around the FP64 log-energy spine;
// the prefix products and sums, the node weights, the flux model and every
// sum over nodes stay FP64.
template <bool TLDS, typename F, int WAVES>
__global__ __launch_bounds__(256, WAVES) void csda_cube_kernel(CsdaArgs a) {
  constexpr bool F64 = sizeof(F) == sizeof(double);
  static_assert(TLDS || F64, "FP32 stepping stages its tables in LDS");
  using Tab = typename std::conditional<F64, Tab64, Tab32>::type;
  HIP_DYNAMIC_SHARED(double, lds)
  const int K = a.K;
  const int bd = blockDim.x, tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int nw = bd >> 6;
  const double* tr = a.tab + (size_t)a.rock * 19 * K;
  const double* ta = a.tab + (size_t)a.air * 19 * K;
  Tab rock, air;
  size_t tab_bytes = 0;
  if constexpr (!F64) {
    // [rock LE, R, 1/dLE, 1/dR][air LE, R, 1/dLE, 1/dR] doubles, then
    // [rock S][air S] floats
    double* tl = lds;
    float* ts = (float*)(lds + 8 * K);
    for (int i = tid; i < K; i += bd) {
      const bool in = i + 1 < K;
      tl[i] = tr[K + i]; tl[K + i] = tr[3 * K + i];
      tl[2 * K + i] = in ? 1.0 / (tr[K + i + 1] - tr[K + i]) : 0.0;
      tl[3 * K + i] = in ? 1.0 / (tr[3 * K + i + 1] - tr[3 * K + i]) : 0.0;
      tl[4 * K + i] = ta[K + i]; tl[5 * K + i] = ta[3 * K + i];
      tl[6 * K + i] = in ? 1.0 / (ta[K + i + 1] - ta[K + i]) : 0.0;
      tl[7 * K + i] = in ? 1.0 / (ta[3 * K + i + 1] - ta[3 * K + i]) : 0.0;
      ts[i] = (float)tr[2 * K + i];
      ts[K + i] = (float)ta[2 * K + i];
    }
    rock = {tl, tl + K, tl + 2 * K, tl + 3 * K, ts};
    air = {tl + 4 * K, tl + 5 * K, tl + 6 * K, tl + 7 * K, ts + K};
    tab_bytes = (size_t)K * 72;
  } else if constexpr (TLDS) {
    // [E, LE, R, S, 1/dLE, 1/dR] of rock, then of air
    double* tl = lds;
    for (int i = tid; i < K; i += bd) {
      const bool in = i + 1 < K;
      tl[i] = tr[0 * K + i]; tl[K + i] = tr[1 * K + i]; tl[2 * K + i] = tr[3 * K + i];
      tl[3 * K + i] = tr[2 * K + i];
      tl[4 * K + i] = in ? 1.0 / (tr[K + i + 1] - tr[K + i]) : 0.0;
      tl[5 * K + i] = in ? 1.0 / (tr[3 * K + i + 1] - tr[3 * K + i]) : 0.0;
      tl[6 * K + i] = ta[0 * K + i]; tl[7 * K + i] = ta[1 * K + i]; tl[8 * K + i] = ta[3 * K + i];
      tl[9 * K + i] = ta[2 * K + i];
      tl[10 * K + i] = in ? 1.0 / (ta[K + i + 1] - ta[K + i]) : 0.0;
      tl[11 * K + i] = in ? 1.0 / (ta[3 * K + i + 1] - ta[3 * K + i]) : 0.0;
    }
    rock = {tl, tl + K, tl + 2 * K, tl + 3 * K, tl + 4 * K, tl + 5 * K};
    air = {tl + 6 * K, tl + 7 * K, tl + 8 * K, tl + 9 * K, tl + 10 * K, tl + 11 * K};
    tab_bytes = (size_t)K * 96;
  } else {
    rock = {tr, tr + K, tr + 3 * K, tr + 2 * K, nullptr, nullptr};
    air = {ta, ta + K, ta + 3 * K, ta + 2 * K, nullptr, nullptr};
  }
  // the table bytes (a multiple of 8) keep the doubles after them aligned
  double* wsum = (double*)((char*)lds + tab_bytes);  // [nw][max_seg]
  double* wrem = wsum + (size_t)nw * a.max_seg;      // [nw]
  double* wflux = wrem + nw;                         // [nw]
  __shared__ int s_ray;

  F* sc = (F*)a.scratch + (size_t)blockIdx.x * a.max_seg * 2 * bd;
  const double rho0 = a.dens[0];
  const F inv_mass = F(1.0 / a.mass);
  const F ctau = F(a.ctau);
  const double inv_ctau = 1.0 / a.ctau;
  // one segment map: FP64 with the entry lookups carried in `en`, or FP32
  auto seg = [&](const Tab& T, F E, double lE, double gram, F path, int& ie, int& jr,
                 Entry64& en) -> StepOutT<F> {
    if constexpr (F64)
      return csda_segment_g(T, K, a.pow2, E, lE, gram, path, inv_mass, ctau, inv_ctau, ie, jr, en);
    else
      return csda_segment_g(T, K, a.pow2, E, lE, gram, path, inv_mass, ctau, ie, jr);
  };

  for (;;) {
    if (tid == 0) s_ray = atomicAdd(a.queue, a.per_pop);
    __syncthreads();
    const int r0 = s_ray;
    __syncthreads();   // s_ray is read before thread 0 writes it again
    if (r0 >= a.n_rays) break;
    const int r_end = r0 + a.per_pop < a.n_rays ? r0 + a.per_pop : a.n_rays;
    for (int r = r0; r < r_end; ++r) {
      const int b0 = a.row_ptr[r];
      const int ns = a.row_ptr[r + 1] - b0;
      if (ns < 0 || ns > a.max_seg) {
        if (tid == 0) {
          atomicOr(a.err, dph::KERR_CAPACITY);
          a.flux[r] = __builtin_nan("");
          if (a.jac_rem) a.jac_rem[r] = __builtin_nan("");
        }
        __syncthreads();
        continue;
      }
      for (int s = lane; s < ns; s += 64) wsum[wave * a.max_seg + s] = 0.0;
      if (lane == 0) { wrem[wave] = 0.0; wflux[wave] = 0.0; }

      const double cosr = a.cos_theta[r];
      const double zex = a.z_exit[r];
      const double cs = fmax(fabs(cosr), 0.01);
      const bool has_rem = zex < a.z_max - EPS64;
      const double rem = (a.z_max - zex) / cs;
      const double Xair = air_grammage(a.z_max, a.palt, cosr);
      const double path_air = (a.palt - a.z_max) / cs;

      for (int n0 = 0; n0 < a.n_nodes; n0 += bd) {
        const int n = n0 + tid;
        const bool active = n < a.n_nodes;
        const double E0 = active ? a.ne[n] : a.ne[0];
        F E = F(E0);
        double lE = log(E0);
        double W = 1.0, X = 0.0;   // stopping-power ratios; summed decay exponents
        bool ok = true;
        int ie = -1, jr = K / 2;   // interval guesses (rock tables)
        Entry64 en;
        en.valid = false;
        // Prefix form of the reverse recursion T_s = T_s+1 c_s + beta_s
        // (beta_s = e_s c_s + b_s).  With P_s = prod_{t<s} c_t and
        // Q_s = sum_{t<s} beta_t P_t, T_s+1 = (K - Q_s+1) / P_s+1 where
        // K = T_N P_N + Q_N, so d flux/d rho_s = fl r_s (T_s+1 + e_s)
        //   = fl alpha_s + (fl K) gamma_s,
        // alpha_s = r_s (e_s - Q_s+1 / P_s+1), gamma_s = r_s / P_s+1:
        // two tape values per segment and node, weighted by per-node end
        // values, and no sequential reverse sweep.  (Q_s+1 telescopes to
        // T_0 - T_s+1 P_s+1, and P is a ratio of stopping powers, so the two
        // terms stay of the size of T P.)  A segment with c_s = 0 (a clamped
        // lookup) decouples what lies before it: T_s = beta_s, so the
        // earlier segments of the current prefix are closed there with
        // K = beta_s P_s + Q_s and a new prefix starts after it.
        double P = 1.0, Q = 0.0;
        int base = 0;
        // ------------------------------------------------ forward sweep
        // The next segment's length and density are loaded one segment
        // ahead: the density gather (cell index, then a random 8 MiB field)
        // is a chain of two dependent loads that the segment map hides.
        double d_nx = 0.0, rho_nx = 0.0;
        if (ns > 0) {
          d_nx = a.len[b0];
          rho_nx = a.dens[a.cell[b0]];
        }
        for (int s = 0; s < ns; ++s) {
          const double d = d_nx;
          const double rho = rho_nx;
          if (s + 1 < ns) {
            d_nx = a.len[b0 + s + 1];
            rho_nx = a.dens[a.cell[b0 + s + 1]];
          }
          StepOutT<F> o = seg(rock, E, lE, d * rho, F(d), ie, jr, en);
          ok = ok && o.ok;
          W = W * o.ratio;
          X += o.dexp;
          {
            // the prefix recursion in FP64 for both stepping types; the tape
            // holds alpha, gamma as F
            F* p = sc + (size_t)s * 2 * bd + tid;
            const double c = o.c, e = o.e;
            const double rr = (double)o.k * d;
            const double beta = e * c + (double)o.b;
            if (c != 0.0) {
              const double P1 = P * c;
              const double Q1 = Q + beta * P;
              const double iP1 = rcp_nr(P1);
              p[0] = F(rr * (e - Q1 * iP1));
              p[bd] = F(rr * iP1);
              P = P1;
              Q = Q1;
            } else {
              const double Kc = beta * P + Q;
              for (int t = base; t < s; ++t) {
                F* q = sc + (size_t)t * 2 * bd + tid;
                q[0] = F((double)q[0] + Kc * (double)q[bd]);
                q[bd] = F(0);
              }
              p[0] = F(rr * e);
              p[bd] = F(rr);
              P = 1.0;
              Q = 0.0;
              base = s + 1;
            }
          }
          E = o.E;
          lE = o.lE;
        }
        StepOutT<F> orem;
        orem.c = orem.b = orem.e = orem.k = F(0);
        if (has_rem) {
          orem = seg(rock, E, lE, rem * rho0, F(rem), ie, jr, en);
          ok = ok && orem.ok;
          W = W * orem.ratio;
          X += orem.dexp;
          E = orem.E;
          lE = orem.lE;
        }
        int ie_air = -1, jr_air = jr;   // air tables: a fresh entry search
        Entry64 en_air;
        en_air.valid = false;
        StepOutT<F> oair = seg(air, E, lE, Xair, F(path_air), ie_air, jr_air, en_air);
        ok = ok && oair.ok;
        W = W * oair.ratio;
        X += oair.dexp;
        W = W * exp(-X);
        double Es;   // surface energy
        if constexpr (F64) Es = oair.E; else Es = exp(oair.lE);
        Dd phi = flux_gccly_d(cosr, Es, active ? a.nq[n] : 0.0);
        const double fl = (active && ok) ? 2.0 * a.nw[n] * (W * phi.v) : 0.0;
        // d log(tail)/dE at the exit of the cell list
        F T = F(0);
        if (fl != 0.0) {
          F as = F(phi.d / phi.v) + oair.e;
          T = as * oair.c + oair.b;
        }
        double fs = wave_sum(fl);
        if (lane == 0) wflux[wave] += fs;
        if (has_rem) {
          F as = T + orem.e;
          double g = fl != 0.0 ? fl * (double)as * ((double)orem.k * rem) : 0.0;
          T = as * orem.c + orem.b;
          double gs = wave_sum(g);
          if (lane == 0) wrem[wave] += gs;
        }
        if (a.jac) {
          // --------------------------------- per-segment sums over nodes
          const double flK = fl != 0.0 ? fl * ((double)T * P + Q) : 0.0;
          for (int s = 0; s < ns; ++s) {
            const F* p = sc + (size_t)s * 2 * bd + tid;
            const double g = fl != 0.0 ? fl * (double)p[0] + flK * (double)p[bd] : 0.0;
            const double gs = wave_sum(g);
            if (lane == 0) wsum[wave * a.max_seg + s] += gs;
          }
        }
      }
      __syncthreads();
      if (a.jac) {
        for (int s = tid; s < ns; s += bd) {
          double v = 0.0;
          for (int w = 0; w < nw; ++w) v += wsum[w * a.max_seg + s];
          a.jac[b0 + s] = a.scale * v;
        }
      }
      if (tid == 0) {
        double f = 0.0, g = 0.0;
        for (int w = 0; w < nw; ++w) { f += wflux[w]; g += wrem[w]; }
        a.flux[r] = a.scale * f;
        if (a.jac_rem) a.jac_rem[r] = a.scale * g;
      }
      __syncthreads();
    }  // rays of this pop
  }
}

// Clears the work queue and, for the host entry point, its error word (the
// _dev entry points accumulate errors in the context's sticky word).
__global__ void reset_counters(int* queue, int* err, int reset_err) {
  if (threadIdx.x == 0) {
    *queue = 0;
    if (reset_err) *err = 0;
  }
}

// ------------------------------------------------------------------ VJP
__global__ void vjp_zero(int n, double* __restrict__ g) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) g[i] = 0.0;
}

// One 64-lane block per ray: the ray's segments are written by consecutive
// lanes (coalesced).
__global__ __launch_bounds__(64) void vjp_keys(int n_rays, int n_seg,
                                               const int32_t* __restrict__ row_ptr,
                                               const int32_t* __restrict__ cell,
                                               int32_t* __restrict__ keys,
                                               int32_t* __restrict__ vals,
                                               int32_t* __restrict__ ray_of) {
  const int r = blockIdx.x;
  if (r >= n_rays) return;
  const int s1 = row_ptr[r + 1];
  for (int s = row_ptr[r] + (int)threadIdx.x; s < s1; s += 64) {
    keys[s] = cell[s];
    vals[s] = s;
    ray_of[s] = r;
  }
}

// Gradient sums over the cell-sorted segment list (keys = cells, ascending)
// in a fixed shape.  vjp_chunk: one block per chunk of VJP_CHUNK positions
// gathers the terms fbar[ray] * jac[segment] in parallel and scans them by
// key (segmented Hillis-Steele in LDS).  A run of one cell that lies inside
// the chunk is written to grad; the first run's partial goes to head[chunk]
// and the last run's to tail[chunk].  vjp_join: for each run that crosses a
// chunk boundary, the chunk where it starts adds tail and the following
// heads in chunk order.  Cells crossed by no segment keep the zero written by
// vjp_zero.  Cells next to a detector are crossed by thousands of rays: the
// terms of such a run are summed by a block, not walked by one lane.
constexpr int VJP_CHUNK = 256;

__global__ __launch_bounds__(VJP_CHUNK) void vjp_chunk(int n_seg, int n_cells,
                                                       const int32_t* __restrict__ skeys,
                                                       const int32_t* __restrict__ svals,
                                                       const int32_t* __restrict__ ray_of,
                                                       const double* __restrict__ jac,
                                                       const double* __restrict__ fbar,
                                                       double* __restrict__ grad,
                                                       double* __restrict__ head,
                                                       double* __restrict__ tail) {
  __shared__ double sv[VJP_CHUNK];
  __shared__ int sk[VJP_CHUNK];
  const int b = blockIdx.x, t = threadIdx.x;
  const int p = b * VJP_CHUNK + t;
  const bool live = p < n_seg;
  const int key = live ? skeys[p] : INT_MAX;
  double v = 0.0;
  if (live) {
    const int sg = svals[p];
    v = fbar[ray_of[sg]] * jac[sg];
  }
  sk[t] = key;
  sv[t] = v;
  __syncthreads();
  for (int d = 1; d < VJP_CHUNK; d <<= 1) {
    const bool same = t >= d && sk[t - d] == key;
    const double add = same ? sv[t - d] : 0.0;
    __syncthreads();
    if (same) sv[t] += add;
    __syncthreads();
  }
  if (!live) return;
  const int last = min(VJP_CHUNK, n_seg - b * VJP_CHUNK) - 1;
  if (t != last && sk[t + 1] == key) return;   // not the last position of its run here
  const double run = sv[t];
  const bool from_prev = key == sk[0] && b > 0 && skeys[b * VJP_CHUNK - 1] == key;
  const bool to_next = t == last && p + 1 < n_seg && skeys[p + 1] == key;
  if (key == sk[0]) head[b] = run;
  if (t == last) tail[b] = run;
  if (!from_prev && !to_next && key >= 0 && key < n_cells) grad[key] = run;
}

__global__ void vjp_join(int n_seg, int n_chunks, int n_cells, const int32_t* __restrict__ skeys,
                         const double* __restrict__ head, const double* __restrict__ tail,
                         double* __restrict__ grad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= n_chunks) return;
  const int start = b * VJP_CHUNK;
  const int last = min(start + VJP_CHUNK, n_seg) - 1;
  const int c = skeys[last];
  if (last + 1 >= n_seg || skeys[last + 1] != c) return;   // the last run ends in this chunk
  if (skeys[start] == c && b > 0 && skeys[start - 1] == c) return;   // it started earlier
  double sum = tail[b];
  for (int j = b + 1; j < n_chunks; ++j) {
    sum += head[j];
    const int jl = min((j + 1) * VJP_CHUNK, n_seg) - 1;
    if (skeys[jl] != c || jl + 1 >= n_seg || skeys[jl + 1] != c) break;   // the run ends in chunk j
  }
  if (c >= 0 && c < n_cells) grad[c] = sum;
}

// The remaining-rock derivatives all belong to cell 0 (the remainder is
// transported at rho[0]): a fixed-shape two-stage tree sums
// fbar[r] jac_rem[r] over the rays and the result is added to grad[0], so
// one cell does not serialise n_rays terms.
constexpr int VJP_REM_THREADS = 256, VJP_REM_PER = 16;
__device__ __forceinline__ double block_sum_fixed(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int off = VJP_REM_THREADS / 2; off > 0; off >>= 1) {
    if (t < off) sh[t] += sh[t + off];
    __syncthreads();
  }
  return sh[0];
}
__global__ __launch_bounds__(VJP_REM_THREADS) void vjp_rem_partial(int n_rays,
                                                                    const double* __restrict__ jac_rem,
                                                                    const double* __restrict__ fbar,
                                                                    double* __restrict__ part) {
  __shared__ double sh[VJP_REM_THREADS];
  const int base = blockIdx.x * VJP_REM_THREADS * VJP_REM_PER;
  double v = 0.0;
  for (int k = 0; k < VJP_REM_PER; ++k) {
    const int r = base + k * VJP_REM_THREADS + (int)threadIdx.x;
    if (r < n_rays) v += fbar[r] * jac_rem[r];
  }
  const double tot = block_sum_fixed(v, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}
__global__ __launch_bounds__(VJP_REM_THREADS) void vjp_rem_final(int n_part,
                                                                  const double* __restrict__ part,
                                                                  double* __restrict__ grad) {
  __shared__ double sh[VJP_REM_THREADS];
  double v = 0.0;
  for (int i = threadIdx.x; i < n_part; i += VJP_REM_THREADS) v += part[i];
  const double tot = block_sum_fixed(v, sh);
  if (threadIdx.x == 0) grad[0] = grad[0] + tot;
}

int check_paths(const dp_paths* p, const dp_csda_nodes* nd, int32_t n_cells, dp_tables t) {
  if (!p || !nd || !t) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: null argument");
  if (p->n_rays < 0 || p->n_segments < 0 || p->max_segments < 0)
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: negative sizes");
  if (nd->n_nodes <= 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: need at least one energy node");
  if (n_cells <= 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: need at least one cell");
  if (nd->rock_material < 0 || nd->rock_material >= t->n_materials || nd->air_material < 0 ||
      nd->air_material >= t->n_materials)
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: material index out of range");
  if (!(nd->primary_altitude > nd->z_max))
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: primary altitude must lie above z_max");
  return DP_OK;
}

// Tables: FP64 12 K doubles (with the knot reciprocals); FP32 stepping 8 K doubles (log energy, range and
// their knot reciprocals) and 2 K floats (stopping powers).
size_t lds_bytes(bool tlds, int K, int nw, int max_seg, size_t fsize = sizeof(double)) {
  const size_t tab = !tlds ? 0 : (size_t)K * (fsize == sizeof(float) ? 72 : 96);
  return tab + ((size_t)nw * max_seg + 2 * nw) * sizeof(double);
}

// Launch knobs of csda_cube_kernel: k0 = tables in LDS (bit 0: 1) or read
// through L2 (0), plus 2 for the 3-wave occupancy bound (otherwise 4 waves
// per SIMD); k1 = rays per queue pop; k2 = grid multiplier over one resident
// wave of blocks.  The first candidate is the default: measured fastest at
// the differentiation workload (FP64 7.83 ms at 3 waves with twice the
// resident grid, 8.11 ms with one, 12.4 ms at 4 waves where it spills; FP32
// stepping 7.0-7.1 ms at 4 waves, 7.8 ms at 3).  FP32 stepping always
// stages its tables in LDS.
const std::vector<dph::TuneCfg>& csda_candidates(bool fp32) {
  static const std::vector<dph::TuneCfg> v = {
      {3, 1, 2}, {3, 1, 1}, {1, 1, 1}, {2, 1, 1}, {3, 4, 1}, {1, 1, 2}, {2, 1, 2}};
  static const std::vector<dph::TuneCfg> v32 = {{1, 1, 2}, {1, 1, 1}, {3, 1, 1}, {1, 4, 1},
                                                {3, 1, 2}};
  return fp32 ? v32 : v;
}

using CsdaKernel = void (*)(CsdaArgs);
CsdaKernel csda_kernel_for(bool f32, int k0) {
  const bool tlds = (k0 & 1) != 0, w3 = (k0 & 2) != 0;
  if (f32) return w3 ? csda_cube_kernel<true, float, 3> : csda_cube_kernel<true, float, 4>;
  if (tlds) return w3 ? csda_cube_kernel<true, double, 3> : csda_cube_kernel<true, double, 4>;
  return w3 ? csda_cube_kernel<false, double, 3> : csda_cube_kernel<false, double, 4>;
}

int launch_csda(dp_ctx c, dp_tables t, const dp_paths* p, const dp_csda_nodes* nd,
                const double* dens, double* flux, double* jac, double* jac_rem, int* queue,
                int* err, bool reset_err) {
  if (p->n_rays == 0) return DP_OK;
  const int bd = nd->n_nodes >= 256 ? 256 : ((nd->n_nodes + 63) / 64) * 64;
  const int nw = bd / 64;
  const int K = t->n_knots;
  const int max_seg = p->max_segments > 0 ? p->max_segments : 1;
  const bool f32 = c->fp32_stepping;
  const size_t fsz = f32 ? sizeof(float) : sizeof(double);
  if (lds_bytes(f32, K, nw, max_seg, fsz) > 160 * 1024)
    return dph::fail(DP_E_CAPACITY, "dp_csda_cube: max_segments " + std::to_string(max_seg) +
                                        " exceeds the LDS budget");
  // Per candidate: LDS, grid; the scratch covers the largest grid.
  const auto& all = csda_candidates(f32);
  std::vector<dph::TuneCfg> cands;
  std::vector<int> grids;
  int grid_max = 1;
  for (const auto& cfg : all) {
    const size_t lds = lds_bytes((cfg.k0 & 1) != 0, K, nw, max_seg, fsz);
    if (lds > 160 * 1024) continue;
    int occ = 0;
    DP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, csda_kernel_for(f32, cfg.k0), bd,
                                                        lds));
    if (occ < 1) occ = 1;
    long g = (long)c->n_cu * occ * cfg.k2;
    const long pops = (p->n_rays + cfg.k1 - 1) / cfg.k1;
    if (g > pops) g = pops;
    cands.push_back(cfg);
    grids.push_back((int)g);
    if (g > grid_max) grid_max = (int)g;
  }
  // tape: alpha, gamma per segment and node (double, or float for FP32 stepping)
  size_t scratch = (size_t)grid_max * max_seg * 2 * bd * fsz;
  int rc = dph::ensure(c->scratch, scratch);
  if (rc) return rc;
  CsdaArgs a;
  a.tab = t->data; a.K = K; a.pow2 = t->pow2;
  a.rock = nd->rock_material; a.air = nd->air_material;
  a.mass = t->mass; a.ctau = t->ctau;
  a.n_rays = p->n_rays; a.max_seg = max_seg;
  a.row_ptr = p->row_ptr; a.cell = p->cell; a.len = p->length;
  a.cos_theta = p->cos_theta; a.z_exit = p->z_exit;
  a.z_max = nd->z_max; a.palt = nd->primary_altitude; a.scale = nd->scale;
  a.n_nodes = nd->n_nodes; a.ne = nd->node_energy; a.nw = nd->node_weight; a.nq = nd->node_charge;
  a.dens = dens; a.flux = flux; a.jac = jac; a.jac_rem = jac_rem;
  a.scratch = (double*)c->scratch.ptr;
  a.queue = queue; a.err = err;
  auto launch = [&](const dph::TuneCfg& cfg) -> int {
    int i = 0;
    while (i < (int)cands.size() && !(cands[i].k0 == cfg.k0 && cands[i].k1 == cfg.k1 &&
                                      cands[i].k2 == cfg.k2))
      ++i;
    if (i == (int)cands.size()) i = 0;
    CsdaArgs b = a;
    b.per_pop = cands[i].k1;
    const size_t lds = lds_bytes((cands[i].k0 & 1) != 0, K, nw, max_seg, fsz);
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, c->stream, queue, err,
                       reset_err ? 1 : 0);
    dph::record(c, c->ev0);
    hipLaunchKernelGGL(csda_kernel_for(f32, cands[i].k0), dim3(grids[i]), dim3(bd), lds,
                       c->stream, b);
    dph::record(c, c->ev1);
    DP_HIP(hipGetLastError());
    return DP_OK;
  };
  const std::string key = std::string(f32 ? "csda_cube fp32" : "csda_cube") + " K=" + std::to_string(K) + " nodes=" +
                          std::to_string(nd->n_nodes) + " rays=" + std::to_string(p->n_rays) +
                          " max_seg=" + std::to_string(max_seg) + " jac=" + (jac ? "1" : "0");
  dph::TuneCfg cfg;
  rc = dph::tuned(c, key, cands, launch, &cfg);
  if (rc) return rc;
  return launch(cfg);
}

// Byte offsets of the VJP's scratch: cell-sorted keys/values and their sort
// buffers, sort histograms, remaining-rock partials, chunk head/tail sums.
struct VjpLayout {
  size_t k, v, ro, p0, p1, h, part, hd, tl, total;
};
VjpLayout vjp_layout(int nr, int ns) {
  VjpLayout L;
  size_t o = 0;
  auto take = [&](size_t bytes) { size_t at = o; o = dph::align_up(o + bytes, 256); return at; };
  const size_t ne = (size_t)(ns > 0 ? ns : 0);
  L.k = take(4 * ne); L.v = take(4 * ne); L.ro = take(4 * ne);
  L.p0 = take(8 * ne); L.p1 = take(8 * ne);
  L.h = take(4 * dps::sort_hist_ints(ns > 0 ? ns : 0));
  const int n_part = (nr + VJP_REM_THREADS * VJP_REM_PER - 1) / (VJP_REM_THREADS * VJP_REM_PER);
  L.part = take(8 * (size_t)(n_part > 0 ? n_part : 0));
  const int n_chunks = (ns + VJP_CHUNK - 1) / VJP_CHUNK;
  L.hd = take(8 * (size_t)(n_chunks > 0 ? n_chunks : 0) + 8);
  L.tl = take(8 * (size_t)(n_chunks > 0 ? n_chunks : 0) + 8);
  L.total = o;
  return L;
}
size_t vjp_scratch_bytes(int nr, int ns) { return vjp_layout(nr, ns).total; }

// The VJP's kernels on the scratch at b (vjp_scratch_bytes(nr, ns) bytes; no
// part of it needs to start at zero: every kernel writes what it later reads).
int vjp_launch(dp_ctx c, const dp_paths* p, const double* jac, const double* jac_rem,
               const double* fbar, int32_t n_cells, double* grad, char* b) {
  const int nr = p->n_rays, ns = p->n_segments;
  if (nr == 0) {   // a fill kernel, not a memset node, so the call can sit in a captured graph
    hipLaunchKernelGGL(vjp_zero, dim3((n_cells + 255) / 256), dim3(256), 0, c->stream, n_cells,
                       grad);
    DP_HIP(hipGetLastError());
    return DP_OK;
  }
  int bits = 1;
  while ((1ll << bits) < (long long)n_cells) ++bits;
  const VjpLayout Lo = vjp_layout(nr, ns);
  const int n_part = (nr + VJP_REM_THREADS * VJP_REM_PER - 1) / (VJP_REM_THREADS * VJP_REM_PER);
  const int n_chunks = (ns + VJP_CHUNK - 1) / VJP_CHUNK;
  int32_t *k = (int32_t*)(b + Lo.k), *v = (int32_t*)(b + Lo.v), *ro = (int32_t*)(b + Lo.ro);
  hipLaunchKernelGGL(vjp_keys, dim3(nr), dim3(64), 0, c->stream, nr, ns, p->row_ptr, p->cell, k, v,
                     ro);
  DP_HIP(hipGetLastError());
  const int32_t *sk = k, *sv = v;
  if (ns > 0)
    dps::radix_sort_pairs(k, v, ns, bits, (int32_t*)(b + Lo.p0), (int32_t*)(b + Lo.p1),
                          (int32_t*)(b + Lo.h), c->stream, &sk, &sv);
  hipLaunchKernelGGL(vjp_zero, dim3((n_cells + 255) / 256), dim3(256), 0, c->stream, n_cells,
                     grad);
  if (ns > 0) {
    double* head = (double*)(b + Lo.hd);
    double* tail = (double*)(b + Lo.tl);
    hipLaunchKernelGGL(vjp_chunk, dim3(n_chunks), dim3(VJP_CHUNK), 0, c->stream, ns, n_cells, sk,
                       sv, ro, jac, fbar, grad, head, tail);
    hipLaunchKernelGGL(vjp_join, dim3((n_chunks + 255) / 256), dim3(256), 0, c->stream, ns,
                       n_chunks, n_cells, sk, head, tail, grad);
  }
  if (jac_rem) {
    double* part = (double*)(b + Lo.part);
    hipLaunchKernelGGL(vjp_rem_partial, dim3(n_part), dim3(VJP_REM_THREADS), 0, c->stream, nr,
                       jac_rem, fbar, part);
    hipLaunchKernelGGL(vjp_rem_final, dim3(1), dim3(VJP_REM_THREADS), 0, c->stream, n_part, part,
                       grad);
  }
  DP_HIP(hipGetLastError());
  return DP_OK;
}

// Scratch of dp_csda_cube_vjp_dev_noalloc: one copy per GPU, for the whole
// process.  Host code never uses this name as an address: it uses the
// per-device address that hipGetSymbolAddress returns (dph::noalloc_scratch).
__device__ __align__(256) char scratch[DP_VJP_NOALLOC_SCRATCH_BYTES];

}  // namespace

namespace dph {
// The device address of `scratch` on `device`, looked up once per device
// (with that device current) and kept for the process.
int noalloc_scratch(int device, void** out) {
  static std::mutex mu;
  static std::map<int, void*> addr;
  std::lock_guard<std::mutex> lk(mu);
  auto it = addr.find(device);
  if (it == addr.end()) {
    DP_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(DP_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    it = addr.emplace(device, p).first;
  }
  *out = it->second;
  return DP_OK;
}
}  // namespace dph

extern "C" {

int dp_csda_cube_dev(dp_ctx c, dp_tables t, const dp_paths* p, const dp_csda_nodes* nd,
                     const double* dens, int32_t n_cells, double* flux, double* jac,
                     double* jac_rem) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_dev: null context");
  int rc = check_paths(p, nd, n_cells, t);
  if (rc) return rc;
  DP_HIP(hipSetDevice(c->device));
  int* q = (int*)c->counter.ptr;
  return launch_csda(c, t, p, nd, dens, flux, jac, jac_rem, q + dph::COUNTER_QUEUE,
                     q + dph::COUNTER_STICKY, false);
}

int dp_csda_cube(dp_ctx c, dp_tables t, const dp_paths* p, const dp_csda_nodes* nd,
                 const double* dens, int32_t n_cells, double* flux, double* jac,
                 double* jac_rem) {
  if (!c) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: null context");
  int rc = check_paths(p, nd, n_cells, t);
  if (rc) return rc;
  const int nr = p->n_rays, ns = p->n_segments, nn = nd->n_nodes;
  if (nr == 0) return DP_OK;
  if (p->row_ptr[0] != 0 || p->row_ptr[nr] != ns)
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: row_ptr must start at 0 and end at n_segments");
  int mx = 0;
  for (int r = 0; r < nr; ++r) {
    int k = p->row_ptr[r + 1] - p->row_ptr[r];
    if (k < 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: row_ptr must be non-decreasing");
    mx = k > mx ? k : mx;
  }
  for (int s = 0; s < ns; ++s)
    if (p->cell[s] < 0 || p->cell[s] >= n_cells)
      return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: cell index out of range at segment " + std::to_string(s));
  for (int i = 0; i < nn; ++i)
    if (!(nd->node_energy[i] > 0))
      return dph::fail(DP_E_ARGUMENT, "dp_csda_cube: node energies must be positive");
  DP_HIP(hipSetDevice(c->device));
  size_t o_rp = 0;
  size_t o_cell = dph::align_up(o_rp + 4 * (size_t)(nr + 1), 256);
  size_t o_len = dph::align_up(o_cell + 4 * (size_t)ns, 256);
  size_t o_cos = dph::align_up(o_len + 8 * (size_t)ns, 256);
  size_t o_zex = dph::align_up(o_cos + 8 * (size_t)nr, 256);
  size_t o_ne = dph::align_up(o_zex + 8 * (size_t)nr, 256);
  size_t o_nw = dph::align_up(o_ne + 8 * (size_t)nn, 256);
  size_t o_nq = dph::align_up(o_nw + 8 * (size_t)nn, 256);
  size_t o_den = dph::align_up(o_nq + 8 * (size_t)nn, 256);
  size_t o_flux = dph::align_up(o_den + 8 * (size_t)n_cells, 256);
  size_t o_jac = dph::align_up(o_flux + 8 * (size_t)nr, 256);
  size_t o_jr = dph::align_up(o_jac + 8 * (size_t)ns, 256);
  size_t o_q = dph::align_up(o_jr + 8 * (size_t)nr, 256);
  size_t total = o_q + 256;
  rc = dph::ensure(c->staging, total);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_rp, p->row_ptr, 4 * (size_t)(nr + 1), hipMemcpyHostToDevice, st));
  if (ns) DP_HIP(hipMemcpyAsync(b + o_cell, p->cell, 4 * (size_t)ns, hipMemcpyHostToDevice, st));
  if (ns) DP_HIP(hipMemcpyAsync(b + o_len, p->length, 8 * (size_t)ns, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_cos, p->cos_theta, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_zex, p->z_exit, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_ne, nd->node_energy, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_nw, nd->node_weight, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_nq, nd->node_charge, 8 * (size_t)nn, hipMemcpyHostToDevice, st));
  DP_HIP(hipMemcpyAsync(b + o_den, dens, 8 * (size_t)n_cells, hipMemcpyHostToDevice, st));
  dp_paths dp = *p;
  dp.max_segments = mx;
  dp.row_ptr = (int32_t*)(b + o_rp); dp.cell = (int32_t*)(b + o_cell); dp.length = (double*)(b + o_len);
  dp.cos_theta = (double*)(b + o_cos); dp.z_exit = (double*)(b + o_zex);
  dp_csda_nodes dn = *nd;
  dn.node_energy = (double*)(b + o_ne); dn.node_weight = (double*)(b + o_nw); dn.node_charge = (double*)(b + o_nq);
  int* qerr = (int*)(b + o_q);
  rc = launch_csda(c, t, &dp, &dn, (double*)(b + o_den), (double*)(b + o_flux),
                   jac ? (double*)(b + o_jac) : nullptr, jac_rem ? (double*)(b + o_jr) : nullptr, qerr,
                   qerr + 1, true);
  if (rc) return rc;
  int herr[2] = {0, 0};
  DP_HIP(hipMemcpyAsync(flux, b + o_flux, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  if (jac && ns) DP_HIP(hipMemcpyAsync(jac, b + o_jac, 8 * (size_t)ns, hipMemcpyDeviceToHost, st));
  if (jac_rem) DP_HIP(hipMemcpyAsync(jac_rem, b + o_jr, 8 * (size_t)nr, hipMemcpyDeviceToHost, st));
  DP_HIP(hipMemcpyAsync(herr, qerr, 8, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  if (herr[1]) return dph::fail(DP_E_CAPACITY, "dp_csda_cube: a ray exceeded max_segments");
  return DP_OK;
}

int dp_csda_cube_vjp_dev(dp_ctx c, const dp_paths* p, const double* jac, const double* jac_rem,
                         const double* fbar, int32_t n_cells, double* grad) {
  if (!c || !p || !jac || !fbar || !grad) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp: null argument");
  if (n_cells <= 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp: need at least one cell");
  DP_HIP(hipSetDevice(c->device));
  const size_t need = vjp_scratch_bytes(p->n_rays, p->n_segments);
  int rc = dph::ensure(c->scratch2, need);
  if (rc) return rc;
  return vjp_launch(c, p, jac, jac_rem, fbar, n_cells, grad, (char*)c->scratch2.ptr);
}

int dp_csda_cube_vjp_dev_noalloc(dp_ctx c, const dp_paths* p, const double* jac,
                                 const double* jac_rem, const double* fbar, int32_t n_cells,
                                 double* grad) {
  if (!c || !p || !jac || !fbar || !grad)
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp_dev_noalloc: null argument");
  if (n_cells <= 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp_dev_noalloc: need at least one cell");
  if (p->n_rays < 0 || p->n_segments < 0)
    return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp_dev_noalloc: negative sizes");
  const size_t need = vjp_scratch_bytes(p->n_rays, p->n_segments);
  if (need > DP_VJP_NOALLOC_SCRATCH_BYTES)
    return dph::fail(DP_E_CAPACITY, "dp_csda_cube_vjp_dev_noalloc: " + std::to_string(p->n_segments) +
                                        " segments and " + std::to_string(p->n_rays) + " rays need " +
                                        std::to_string(need) + " bytes of scratch; the static buffer holds " +
                                        std::to_string((size_t)DP_VJP_NOALLOC_SCRATCH_BYTES) +
                                        " (about 2.2e6 segments)");
  if (!c->noalloc_scratch)
    return dph::fail(DP_E_HIP, "dp_csda_cube_vjp_dev_noalloc: the context has no static scratch address");
  DP_HIP(hipSetDevice(c->device));
  return vjp_launch(c, p, jac, jac_rem, fbar, n_cells, grad, (char*)c->noalloc_scratch);
}

int dp_csda_cube_vjp(dp_ctx c, const dp_paths* p, const double* jac, const double* jac_rem,
                     const double* fbar, int32_t n_cells, double* grad) {
  if (!c || !p || !jac || !fbar || !grad) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp: null argument");
  if (n_cells <= 0) return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp: need at least one cell");
  const int nr = p->n_rays, ns = p->n_segments;
  for (int s = 0; s < ns; ++s)
    if (p->cell[s] < 0 || p->cell[s] >= n_cells)
      return dph::fail(DP_E_ARGUMENT, "dp_csda_cube_vjp: cell index out of range");
  DP_HIP(hipSetDevice(c->device));
  size_t o_rp = 0, o_cell = dph::align_up(4 * (size_t)(nr + 1), 256);
  size_t o_j = dph::align_up(o_cell + 4 * (size_t)ns, 256);
  size_t o_jr = dph::align_up(o_j + 8 * (size_t)ns, 256);
  size_t o_fb = dph::align_up(o_jr + 8 * (size_t)nr, 256);
  size_t o_g = dph::align_up(o_fb + 8 * (size_t)nr, 256);
  int rc = dph::ensure(c->staging, o_g + 8 * (size_t)n_cells + 256);
  if (rc) return rc;
  char* b = (char*)c->staging.ptr;
  hipStream_t st = c->stream;
  DP_HIP(hipMemcpyAsync(b + o_rp, p->row_ptr, 4 * (size_t)(nr + 1), hipMemcpyHostToDevice, st));
  if (ns) DP_HIP(hipMemcpyAsync(b + o_cell, p->cell, 4 * (size_t)ns, hipMemcpyHostToDevice, st));
  if (ns) DP_HIP(hipMemcpyAsync(b + o_j, jac, 8 * (size_t)ns, hipMemcpyHostToDevice, st));
  if (jac_rem && nr) DP_HIP(hipMemcpyAsync(b + o_jr, jac_rem, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  if (nr) DP_HIP(hipMemcpyAsync(b + o_fb, fbar, 8 * (size_t)nr, hipMemcpyHostToDevice, st));
  dp_paths dp = *p;
  dp.row_ptr = (int32_t*)(b + o_rp);
  dp.cell = (int32_t*)(b + o_cell);
  rc = dp_csda_cube_vjp_dev(c, &dp, (double*)(b + o_j), jac_rem ? (double*)(b + o_jr) : nullptr,
                            (double*)(b + o_fb), n_cells, (double*)(b + o_g));
  if (rc) return rc;
  DP_HIP(hipMemcpyAsync(grad, b + o_g, 8 * (size_t)n_cells, hipMemcpyDeviceToHost, st));
  DP_HIP(hipStreamSynchronize(st));
  return DP_OK;
}

}  // extern "C"
