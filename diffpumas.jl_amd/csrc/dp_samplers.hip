// Batched samplers of the reference's exported per-call API that the current
// transport path no longer calls: sample_mixture_material
// (src/Physics.jl:1331-1365) and the legacy discrete-energy-loss sampler
// sample_del_event (src/Straggling.jl:217-317).
//
// One sample per lane, grid-stride.  Sample i draws its uniforms, in the
// order the reference calls rand(rng), from Philox keyed (seed, i), so a
// batch is reproducible and independent of the launch shape.  Both are
// table-lookup bound (one to three log-energy brackets per draw); the tables
// are read through L1/L2.
#include "dp_internal.hpp"
#include "dp_device.hpp"
#include <cmath>

using namespace dpk;

namespace {

constexpr int SMP_THREADS = 256;
constexpr int MIX_MAX = 8;   // components of a mixture in one call

struct TabView {
  const double* tab;
  int K, pow2;
  __device__ __forceinline__ const double* a(int m, int arr) const {
    return tab + ((size_t)m * DP_T_NARRAYS + arr) * K;
  }
  // interpolate_table / interpolate_table_fast (src/Physics.jl:620-708):
  // clamped, linear in log energy
  __device__ __forceinline__ double at(int m, int arr, double e, double le) const {
    return interp_e(a(m, DP_T_ENERGY), a(m, DP_T_LOG_ENERGY), a(m, arr), K, pow2, e, le).y;
  }
};

struct MixArgs {
  TabView T;
  int n_mix;
  int mats[MIX_MAX];
  double frac[MIX_MAX];
  int64_t n;
  const double* energy;
  uint64_t seed;
  int32_t* out;
};

// sample_mixture_material: material j with probability f_j sigma_j(E) /
// sum_k f_k sigma_k(E) (sigma: the tabulated hard cross-section); by f_j
// alone when that sum is not positive.  A single material is returned
// without a draw.
__global__ __launch_bounds__(SMP_THREADS) void mixture_material_kernel(MixArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    if (a.n_mix == 1) {
      a.out[i] = a.mats[0];
      continue;
    }
    const double e = a.energy[i], le = log(e);
    double xs[MIX_MAX];
    double tot = 0.0;
    for (int j = 0; j < a.n_mix; ++j) {
      xs[j] = a.frac[j] * a.T.at(a.mats[j], DP_T_CROSS_SECTION, e, le);
      tot += xs[j];
    }
    Philox rng;
    rng.init(a.seed, (uint64_t)i);
    int pick = a.mats[a.n_mix - 1];
    double cum = 0.0;
    if (tot <= 0.0) {
      const double u = rng.uniform();
      for (int j = 0; j < a.n_mix; ++j) {
        cum += a.frac[j];
        if (u <= cum) { pick = a.mats[j]; break; }
      }
    } else {
      const double u = rng.uniform() * tot;
      for (int j = 0; j < a.n_mix; ++j) {
        cum += xs[j];
        if (u <= cum) { pick = a.mats[j]; break; }
      }
    }
    a.out[i] = pick;
  }
}

struct DelArgs {
  TabView T;
  int m;
  double cutoff;
  bool backward;
  int64_t n;
  const double *ki, *kf, *dX, *Xtot, *ratio;
  uint64_t seed;
  uint8_t* occurred;
  double* X;
  double* k;
  int32_t* process;
};

// compute_del_cross_section (src/Straggling.jl:186-199)
__device__ __forceinline__ double del_xs(const TabView& T, int m, double k) {
  if (k < 1e-3) return 0.0;
  return fmax(T.at(m, DP_T_CROSS_SECTION, k, log(k)), 0.0);
}

// sample_del_event: an exponential interaction grammage against the larger
// end-point cross-section, the energy there from the MIXED range table and
// the fluctuation ratio, rejection on the local cross-section, a transfer
// nu from the 1/nu^2 law on [cutoff, 1] (BMC_ALPHA = 2), the energy after
// the event (k/(1-nu) backward, k(1-nu) forward) and the process drawn by
// the brems : pair : photonuclear stopping powers at the event energy.
__global__ __launch_bounds__(SMP_THREADS) void del_event_kernel(DelArgs a) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const double sgn = a.backward ? -1.0 : 1.0;
  const int m = a.m;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += stride) {
    const double ki = a.ki[i];
    bool occ = false;
    double Xd = 0.0, kd = 0.0;
    int proc = 0;
    const double xs = fmax(del_xs(a.T, m, ki), del_xs(a.T, m, a.kf[i]));
    Philox rng;
    rng.init(a.seed, (uint64_t)i);
    do {
      if (xs <= 0.0) break;
      const double zeta = rng.uniform();
      if (zeta <= 1e-12 || zeta >= 1.0 - 1e-12) break;
      const double X = -log(zeta) / xs;
      if (X >= a.dX[i]) break;
      const Lookup kh = inverse_range(a.T.a(m, DP_T_ENERGY), a.T.a(m, DP_T_LOG_ENERGY),
                                      a.T.a(m, DP_T_MIXED_RANGE), a.T.K, a.T.pow2,
                                      a.Xtot[i] - sgn * X);
      const double kat = ki - sgn * a.ratio[i] * fabs(ki - kh.y);
      if (kat <= 0.0) break;
      const double r = del_xs(a.T, m, kat) / xs;
      if (rng.uniform() > r) break;
      // nu from P(nu) ~ nu^-2 on [cutoff, 1]: the inverse CDF with alpha = 2
      const double u = rng.uniform();
      const double ct = pow(a.cutoff, -1.0);
      double nu = pow(u * (1.0 - ct) + ct, -1.0);
      nu = fmin(fmax(nu, a.cutoff), 1.0);
      kd = a.backward ? kat / (1.0 - nu) : kat * (1.0 - nu);
      const double lk = log(kat);
      const double b = a.T.at(m, DP_T_BREMS, kat, lk);
      const double p = a.T.at(m, DP_T_PAIR, kat, lk);
      const double ph = a.T.at(m, DP_T_PHOTO, kat, lk);
      const double tot = b + p + ph;
      proc = 1;
      if (tot > 0.0) {
        const double rp = rng.uniform() * tot;
        proc = rp < b ? 1 : (rp < b + p ? 2 : 3);
      }
      occ = true;
      Xd = X;
    } while (false);
    a.occurred[i] = occ ? 1 : 0;
    a.X[i] = occ ? Xd : 0.0;
    a.k[i] = occ ? kd : 0.0;
    a.process[i] = proc;
  }
}

int grid_for(dp_ctx c, int64_t n) {
  int64_t g = (n + SMP_THREADS - 1) / SMP_THREADS;
  const int64_t cap = (int64_t)c->n_cu * 8;
  if (g > cap) g = cap;
  return (int)(g < 1 ? 1 : g);
}

}  // namespace

extern "C" {

int dp_sample_mixture_material(dp_ctx c, dp_tables t, int32_t n_components,
                               const int32_t* materials, const double* fractions, int64_t n,
                               const double* energies, uint64_t seed, int32_t* out) {
  if (!c || !t || !materials || !fractions || (n > 0 && (!energies || !out)))
    return dph::fail(DP_E_ARGUMENT, "dp_sample_mixture_material: null argument");
  if (n_components < 1 || n_components > MIX_MAX)
    return dph::fail(DP_E_ARGUMENT, "dp_sample_mixture_material: 1 to " +
                                        std::to_string(MIX_MAX) + " components");
  if (n < 0) return dph::fail(DP_E_ARGUMENT, "dp_sample_mixture_material: negative count");
  MixArgs a;
  a.T = {t->data, t->n_knots, t->pow2};
  a.n_mix = n_components;
  // MaterialMixture normalises positive fraction sums (src/Types.jl:360-369)
  double tot = 0.0;
  for (int j = 0; j < n_components; ++j) {
    if (materials[j] < 0 || materials[j] >= t->n_materials)
      return dph::fail(DP_E_ARGUMENT, "dp_sample_mixture_material: material index out of range");
    tot += fractions[j];
  }
  for (int j = 0; j < n_components; ++j) {
    a.mats[j] = materials[j];
    a.frac[j] = tot > 0 ? fractions[j] / tot : fractions[j];
  }
  if (n == 0) return DP_OK;
  for (int64_t i = 0; i < n; ++i)
    if (!(energies[i] > 0))
      return dph::fail(DP_E_ARGUMENT, "dp_sample_mixture_material: energies must be positive");
  DP_HIP(hipSetDevice(c->device));
  const size_t o_out = dph::align_up(8 * (size_t)n, 256);
  int rc = dph::ensure(c->scratch, o_out + 4 * (size_t)n);
  if (rc) return rc;
  char* b = (char*)c->scratch.ptr;
  DP_HIP(hipMemcpyAsync(b, energies, 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  a.n = n;
  a.energy = (const double*)b;
  a.seed = seed;
  a.out = (int32_t*)(b + o_out);
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(mixture_material_kernel, dim3(grid_for(c, n)), dim3(SMP_THREADS), 0,
                     c->stream, a);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  DP_HIP(hipMemcpyAsync(out, a.out, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

int dp_sample_del_event(dp_ctx c, dp_tables t, int32_t material, int64_t n, const double* ki,
                        const double* kf, const double* dX, const double* Xtot,
                        const double* ratio, int32_t backward, uint64_t seed, uint8_t* occurred,
                        double* X_del, double* k_del, int32_t* process) {
  if (!c || !t || (n > 0 && (!ki || !kf || !dX || !Xtot || !ratio || !occurred || !X_del ||
                             !k_del || !process)))
    return dph::fail(DP_E_ARGUMENT, "dp_sample_del_event: null argument");
  if (material < 0 || material >= t->n_materials)
    return dph::fail(DP_E_ARGUMENT, "dp_sample_del_event: material index out of range");
  if (n < 0) return dph::fail(DP_E_ARGUMENT, "dp_sample_del_event: negative count");
  if (n == 0) return DP_OK;
  DP_HIP(hipSetDevice(c->device));
  const size_t col = dph::align_up(8 * (size_t)n, 256);
  // inputs ki, kf, dX, Xtot, ratio; outputs X, k (doubles), process (int),
  // occurred (byte)
  int rc = dph::ensure(c->scratch, 7 * col + dph::align_up(4 * (size_t)n, 256) + (size_t)n);
  if (rc) return rc;
  char* b = (char*)c->scratch.ptr;
  const double* in[5] = {ki, kf, dX, Xtot, ratio};
  for (int j = 0; j < 5; ++j)
    DP_HIP(hipMemcpyAsync(b + j * col, in[j], 8 * (size_t)n, hipMemcpyHostToDevice, c->stream));
  DelArgs a;
  a.T = {t->data, t->n_knots, t->pow2};
  a.m = material;
  a.cutoff = t->cutoff;
  a.backward = backward != 0;
  a.n = n;
  a.ki = (const double*)b; a.kf = (const double*)(b + col); a.dX = (const double*)(b + 2 * col);
  a.Xtot = (const double*)(b + 3 * col); a.ratio = (const double*)(b + 4 * col);
  a.seed = seed;
  a.X = (double*)(b + 5 * col);
  a.k = (double*)(b + 6 * col);
  a.process = (int32_t*)(b + 7 * col);
  a.occurred = (uint8_t*)(b + 7 * col + dph::align_up(4 * (size_t)n, 256));
  dph::record(c, c->ev0);
  hipLaunchKernelGGL(del_event_kernel, dim3(grid_for(c, n)), dim3(SMP_THREADS), 0, c->stream, a);
  dph::record(c, c->ev1);
  DP_HIP(hipGetLastError());
  DP_HIP(hipMemcpyAsync(X_del, a.X, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(k_del, a.k, 8 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(process, a.process, 4 * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipMemcpyAsync(occurred, a.occurred, (size_t)n, hipMemcpyDeviceToHost, c->stream));
  DP_HIP(hipStreamSynchronize(c->stream));
  (void)hipEventElapsedTime(&c->last_ms, c->ev0, c->ev1);
  return DP_OK;
}

}  // extern "C"
