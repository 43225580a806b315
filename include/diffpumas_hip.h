/*
 * libdiffpumas_hip — MI355X (gfx950) kernels for DiffPumas' backward muon
 * transport and tomography forward model, behind a plain C ABI.
 *
 * The reference (grinisrit/diffpumas.jl) is pure Julia with no FFI seam; each
 * entry point below replaces the Julia function named in its comment and is
 * what a Julia `ccall` binding binds (one `ccall` per entry point, with the
 * argument order and types declared here).
 *
 * Conventions
 *   - Every function returns int: 0 = DP_OK, negative = error class.  The
 *     message of the last failure on the calling thread is dp_last_error().
 *   - Material, cell and ray indices are 0-based (the Julia side adds 1).
 *   - Host-pointer entry points copy inputs to device memory, run, copy
 *     results back and synchronise before returning.  The *_dev variants take
 *     device pointers and enqueue on the context's stream without
 *     synchronising.
 *   - A dp_ctx is not thread-safe; use one context per host thread.
 *   - Results do not depend on launch shape, and repeated calls with the same
 *     inputs return bitwise identical outputs.
 */
#ifndef DIFFPUMAS_HIP_H
#define DIFFPUMAS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
#define DP_OK 0
#define DP_E_ARGUMENT (-1)  /* ArgumentError (bad mode, density, charge, size) */
#define DP_E_DOMAIN (-2)    /* DomainError (energy outside a prepared table)   */
#define DP_E_EXHAUSTED (-3) /* ErrorException (event/step caps, zero progress) */
#define DP_E_HIP (-4)       /* HIP runtime failure                             */
#define DP_E_CAPACITY (-5)  /* output or scratch capacity exceeded             */

/* ------------------------------------------------------ material tables
 * MaterialTable (src/Physics.jl:64-102): 19 arrays on one knot grid, in the
 * order below, stored [material][array][knot].                             */
enum dp_table_array {
  DP_T_ENERGY = 0, DP_T_LOG_ENERGY, DP_T_CSDA_STOP, DP_T_CSDA_RANGE, DP_T_CSDA_TIME,
  DP_T_MIXED_STOP, DP_T_MIXED_RANGE, DP_T_STRAGGLING, DP_T_IONIZATION, DP_T_BREMS,
  DP_T_PAIR, DP_T_PHOTO, DP_T_CROSS_SECTION, DP_T_TRANSPORT_PATH, DP_T_ELASTIC_PATH,
  DP_T_ELASTIC_CUTOFF, DP_T_MAGNETIC_ROTATION, DP_T_SCREENING, DP_T_SPIN_FACTOR,
  DP_T_NARRAYS
};

typedef struct dp_ctx_s* dp_ctx;
typedef struct dp_tables_s* dp_tables;

typedef struct {
  int32_t n_materials;
  int32_t n_knots;          /* common knot grid (create_physics uses one grid) */
  const double* data;       /* n_materials * DP_T_NARRAYS * n_knots            */
  const double* scalars;    /* n_materials * 4: density kg/m3, I GeV, Z/A, 0   */
  double mass;              /* particle mass, GeV                               */
  double ctau;              /* decay length, m                                  */
  /* Element recipes (needed by stochastic transport: hard-event DCS weights
   * and targets).  n_elements[m] elements per material, then Z, A, I (GeV)
   * and mass fraction concatenated in material order.  NULL = none.       */
  const int32_t* n_elements;
  const double* element_Z;
  const double* element_A;
  const double* element_I;
  const double* element_fraction;
  double cutoff;            /* PhysicsSettings.cutoff (0.05)        */
  double elastic_ratio;     /* PhysicsSettings.elastic_ratio (0.05) */
} dp_table_desc;

/* ------------------------------------------------------------- lifecycle */
int dp_create(dp_ctx* out, int device);
int dp_destroy(dp_ctx ctx);
const char* dp_last_error(void);
/* Use `stream` (a hipStream_t) for all work of this context; NULL selects
 * the context's own stream. */
int dp_set_stream(dp_ctx ctx, void* stream);
int dp_synchronize(dp_ctx ctx);
/* Device-side errors of the _dev entry points (no counterpart in the
 * reference, where the same conditions throw at once).  A _dev call cannot
 * report what its kernels find, so they set bits in a sticky word of the
 * context instead: a path longer than the max_segments it was sized with
 * (that ray's flux is NaN), or a cell index outside the field (the segment
 * is skipped).  dp_check_errors synchronises the context's stream, reads and
 * clears the word, and returns DP_E_ARGUMENT (cell index) or DP_E_CAPACITY
 * (path length) if any bit was set since the last check, else DP_OK.  Call
 * it after a chain of _dev calls or graph launches; host entry points check
 * their own kernels and leave the word alone.                             */
int dp_check_errors(dp_ctx ctx);
/* Launch autotuning (no counterpart in the reference, whose kernels are
 * compiled by Julia per call with fixed launch shapes).  With it on, the
 * first call of a tuned kernel for a workload shape (kernel, sizes) times
 * each candidate launch configuration once on the context's stream and keeps
 * the fastest for that shape in the context; later calls, and calls while a
 * capture is open, use the kept choice (or the default if none).  Results do
 * not depend on the choice: every candidate computes the same values
 * bit for bit.  dp_clear_tuning forgets the choices; dp_tuning_report writes
 * one line per shape ("key | best k0,k1,k2 ms | tried cfg:ms ...") into buf
 * (NUL-terminated, truncated to size) and returns the full length;
 * dp_set_tuning pins the knobs for a key as the report spells it (a
 * configuration outside a kernel's candidate list selects its default).   */
int dp_set_autotune(dp_ctx ctx, int32_t on);
/* FP32 stepping for the per-voxel-density CSDA flux and adjoint (dp_csda_cube
 * and every entry point built on it) of this context: the segment maps'
 * exponentials, stopping powers, decay factors, their partials and the
 * adjoint tape in float, around an FP64 spine (the walk's log
 * energy, the range lookup and the exit-grammage inversion, which take
 * subtractions and products only); node weights, the flux model, every sum
 * over nodes and the outputs stay FP64.  The reference computes in Float64
 * only (examples/diff_flux_cube.jl:374-468); this variant trades relative
 * errors of a few 1e-6 (bar 1e-5) against the FP64 program for speed.  Off
 * by default.                                                             */
int dp_set_fp32_stepping(dp_ctx ctx, int32_t on);
int dp_set_tuning(dp_ctx ctx, const char* key, int32_t k0, int32_t k1, int32_t k2);
int dp_clear_tuning(dp_ctx ctx);
int64_t dp_tuning_report(dp_ctx ctx, char* buf, int64_t size);
/* HIP graphs of device work (no counterpart in the reference: Julia re-runs
 * the calls).  Between dp_capture_begin and dp_capture_end the context's
 * stream is captured instead of run; only _dev entry points may be called
 * in between, each already called once with the same sizes so that no
 * buffer grows during the capture (a growth returns DP_E_ARGUMENT at once; a
 * host entry point makes the capture fail, reported by dp_capture_end).  A
 * graph holds the addresses of the context's device buffers: once a later
 * call on the context reallocates one of them (a larger problem than at
 * capture), dp_graph_launch refuses the graph with DP_E_ARGUMENT, and it
 * must be captured again; a graph is launched on the context that captured
 * it.  dp_graph_launch replays the
 * captured work on the context's stream with the device buffers' current
 * contents (e.g. the per-iteration forward model + Jacobian + normal
 * equations of a Gauss-Newton loop, gauss_newton_reconstruct,
 * src/Tomography_inverse.jl:448-603).  Kernel timings are not recorded
 * inside a graph.                                                          */
typedef struct dp_graph_s* dp_graph;
int dp_capture_begin(dp_ctx ctx);
int dp_capture_end(dp_ctx ctx, dp_graph* out);
int dp_graph_launch(dp_ctx ctx, dp_graph graph);
int dp_graph_destroy(dp_graph graph);
/* Number of compute units and wavefronts the autotuned launches assume. */
int dp_device_info(dp_ctx ctx, int32_t* n_cu, int32_t* wave_size);

/* Replaces: the PhysicsTables object handed to every transport call
 * (create_physics, src/Physics.jl:534-599).  Copies the arrays to HBM. */
int dp_upload_tables(dp_ctx ctx, const dp_table_desc* desc, dp_tables* out);
int dp_free_tables(dp_ctx ctx, dp_tables tables);

/* ------------------------------------------------------ table building
 * Host-side construction of the MaterialTable arrays (setup, not the hot
 * path).  Element recipe of a material as read from a PUMAS MDF file
 * (src/Loader.jl:115-205): Z, A (g/mol), I (GeV), mass fractions.        */
typedef struct {
  int32_t n_elements;
  const double* Z;
  const double* A;
  const double* I;
  const double* fraction;
  double I_material;         /* GeV */
  double density;            /* kg/m3 */
  double ZoA;
} dp_material_recipe;

/* Columns of a PUMAS dE/dx table in GeV and kg/m2 units
 * (load_dedx_file, src/DEDXLoader.jl:100-191). */
typedef struct {
  int32_t n;
  const double* kinetic_energy;
  const double* ionization;
  const double* bremsstrahlung;
  const double* pair_production;
  const double* photonuclear;
  const double* stopping_power;
  const double* csda_range;
  double ZoA;                /* header <Z/A> of the file (used by the Coulomb tables) */
} dp_dedx_columns;

/* PhysicsSettings (src/Physics.jl:39-57) plus the particle mass. */
typedef struct {
  double cutoff;             /* relative soft/hard cutoff (0.05) */
  double elastic_ratio;      /* 0.05 */
  double mass;               /* GeV */
} dp_physics_settings;

/* Replaces: create_material_table_from_dedx (src/Physics.jl:146-273).
 * out[DP_T_NARRAYS * n] receives the 19 arrays on the dE/dx knots. */
int dp_build_material_table(const dp_material_recipe* recipe, const dp_dedx_columns* dedx,
                            const dp_physics_settings* settings, double* out);
/* Replaces: create_composite_table (src/Physics.jl:421-525); `flat` is the
 * flattened element recipe (flatten_to_base_material, src/Materials.jl). */
int dp_build_composite_table(int32_t n_components, const double* const* tables,
                             const double* fractions, int32_t n_knots,
                             const dp_material_recipe* flat, const dp_physics_settings* settings,
                             double* out);
/* Replaces: create_material_table_empirical (src/Physics.jl:280-408), the
 * table create_physics builds when a material has no PUMAS dE/dx file:
 * out[DP_T_NARRAYS * n] on the given increasing kinetic energies (GeV), with
 * the electronic stopping power of the Fano model and the radiative ones
 * from the DCS integrals.                                                  */
int dp_build_material_table_empirical(const dp_material_recipe* recipe, const double* energies,
                                      int32_t n, const dp_physics_settings* settings,
                                      double* out);
/* electronic_stopping_power (src/Materials.jl:486-520; GeV m2/kg) and the
 * Fano density effect delta(gamma) (electronic_density_effect, :442-480) of
 * a material, with its atomic shells built as atomic_shell_create (:378-436).
 * NaN for an invalid recipe.                                                */
double dp_electronic_stopping_power(const dp_material_recipe* recipe, double mass, double K);
double dp_density_effect(const dp_material_recipe* recipe, double gamma);
/* Per-atom dσ/dq (m2/GeV) of process 0 brems, 1 pair, 2 photonuclear,
 * 3 ionisation (src/Materials.jl:534-926), and its moment integral per kg
 * of element (compute_dcs_integral, src/Materials.jl:939-1007). */
double dp_dcs(int32_t process, double Z, double A, double I, double mass, double K, double q);
double dp_dcs_integral(int32_t process, double Z, double A, double I, double mass, double K,
                       double xlow, double xhigh, int32_t mode, int32_t npoints);

/* ---------------------------------------------------------- traversal
 * Uniform cube grid, CubeGeometry (examples/diff_flux_cube.jl:85-129). */
typedef struct {
  double x_min, x_max, y_min, y_max, z_min, z_max;
  int32_t nx, ny, nz;
  double dx, dy, dz;
} dp_cube_grid;

/* Replaces: trace_cube_path (examples/diff_flux_cube.jl:304-364), one ray
 * per entry.  start_xy[2n], dir[3n] (unit, upward-going).  Two-phase: with
 * cells == NULL only counts[n] (segments per ray) and z_exit[n] are written;
 * otherwise row_ptr[n+1] (from the counts) locates each ray's segments.    */
int dp_trace_cube(dp_ctx ctx, const dp_cube_grid* grid, int32_t n_rays,
                  const double* start_xy, const double* dir, int32_t* counts,
                  const int32_t* row_ptr, int32_t* cells, double* lengths, double* z_exit);
int dp_trace_cube_dev(dp_ctx ctx, const dp_cube_grid* grid, int32_t n_rays,
                      const double* start_xy, const double* dir, int32_t* counts,
                      const int32_t* row_ptr, int32_t* cells, double* lengths, double* z_exit);

/* Replaces: trace_path(::VoxelVolume) (examples/lvd_tomography.jl:738-779)
 * on a rectilinear grid given by its edge vectors.  origin[3n] is the
 * detector, surface_distance[n] the distance to the topography exit.      */
int dp_trace_edges(dp_ctx ctx, const double* x_edges, int32_t nxe, const double* y_edges,
                   int32_t nye, const double* z_edges, int32_t nze, int32_t n_rays,
                   const double* origin, const double* dir, const double* surface_distance,
                   int32_t* counts, const int32_t* row_ptr, int32_t* cells, double* lengths,
                   double* traveled);

/* ------------------------------------------- per-voxel-density CSDA
 * Ray paths in CSR form (rows = rays).                                     */
typedef struct {
  int32_t n_rays;
  int32_t n_segments;
  int32_t max_segments;      /* max over rays of row_ptr[r+1]-row_ptr[r] */
  const int32_t* row_ptr;    /* n_rays + 1 */
  const int32_t* cell;       /* n_segments, 0-based */
  const double* length;      /* n_segments, m */
  const double* cos_theta;   /* n_rays */
  const double* z_exit;      /* n_rays */
} dp_paths;

typedef struct {
  int32_t rock_material;
  int32_t air_material;
  double z_max;              /* top of the voxel block (air interface), m */
  double primary_altitude;   /* m (src/Geometry.jl:36 default 3e4)        */
  double scale;              /* multiplies every ray flux (e.g. 1/N)      */
  int32_t n_nodes;           /* energy nodes shared by all rays           */
  const double* node_energy; /* GeV */
  const double* node_weight; /* integration weight (sum uses 2*w)        */
  const double* node_charge; /* -1, 0 (charge-summed) or +1               */
} dp_csda_nodes;

/* Replaces: compute_flux_from_segments (examples/diff_flux_cube.jl:374-468)
 * summed as in compute_flux_gradient_cube (:531-541), per ray:
 *   flux[r] = scale * sum_n 2 w_n flux(ray r, E_n, q_n)
 * and, if jac != NULL, its reverse-mode derivative:
 *   jac[s]     = d flux[r] / d rho[cell[s]] through segment s of ray r
 *   jac_rem[r] = d flux[r] / d rho[0] through the remaining-rock segment.
 * density[n_cells] is the per-voxel density, kg/m3.                       */
int dp_csda_cube(dp_ctx ctx, dp_tables tables, const dp_paths* paths, const dp_csda_nodes* nodes,
                 const double* density, int32_t n_cells, double* flux, double* jac,
                 double* jac_rem);
int dp_csda_cube_dev(dp_ctx ctx, dp_tables tables, const dp_paths* paths,
                     const dp_csda_nodes* nodes, const double* density, int32_t n_cells,
                     double* flux, double* jac, double* jac_rem);

/* Replaces: the pullback of Zygote.gradient in compute_flux_gradient_cube
 * (examples/diff_flux_cube.jl:547): grad[c] = sum_r flux_bar[r] d flux[r]/d rho[c]
 * from the per-segment Jacobian, summed per cell in segment order (a stable
 * cell-sorted transpose of the paths is built on each call), so the result
 * is deterministic and independent of launch shape.                        */
int dp_csda_cube_vjp(dp_ctx ctx, const dp_paths* paths, const double* jac, const double* jac_rem,
                     const double* flux_bar, int32_t n_cells, double* grad);
int dp_csda_cube_vjp_dev(dp_ctx ctx, const dp_paths* paths, const double* jac,
                         const double* jac_rem, const double* flux_bar, int32_t n_cells,
                         double* grad);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): dp_csda_cube_vjp_dev with the same arguments, results
 * and kernels, whose scratch (the cell-sorted transpose and its sort buffers,
 * chunk partial sums) is a static device array of
 * DP_VJP_NOALLOC_SCRATCH_BYTES held by every GPU the library runs on until the
 * process ends, instead of a context buffer.  It calls no hipMalloc, hipFree
 * or allocating library routine.  Problems needing more scratch (about 29
 * bytes per segment plus 1 per ray: up to about 2.2e6 segments) return
 * DP_E_CAPACITY before anything is launched.  All calls on one GPU share that
 * GPU's array, so calls that overlap on one GPU (several streams or threads)
 * race; serialise them.                                                     */
#define DP_VJP_NOALLOC_SCRATCH_BYTES (64u * 1024u * 1024u)
int dp_csda_cube_vjp_dev_noalloc(dp_ctx ctx, const dp_paths* paths, const double* jac,
                                 const double* jac_rem, const double* flux_bar, int32_t n_cells,
                                 double* grad);

/* -------------------------------------------- flat backward Monte Carlo
 * TwoLayerGeometry (src/Geometry.jl:58-78), materials 0-based, porous < 0
 * disables the porous layer.                                               */
typedef struct {
  double rock_thickness;
  double rock_density;
  int32_t rock_material;
  int32_t air_material;
  int32_t porous_material;
  double porous_density;
  double porous_thickness;
  double primary_altitude;
  /* Optional runtime mixture for the rock layer (MaterialMixture,
   * src/Types.jl:357-391): rock_n_components > 0 replaces rock_material. */
  int32_t rock_n_components;
  int32_t rock_components[4];
  double rock_fractions[4];
} dp_flat_geometry;

typedef struct {
  double elevation;          /* degrees; cos/sin of the zenith from host   */
  double cos_theta;          /* cos((90 - elevation) * pi / 180)            */
  double sin_theta;
  double energy_min, energy_max;
  double energy_threshold;   /* source energy maximum (<= table max)        */
  double mode_switch;        /* energy_threshold_low, GeV (default 100)     */
  int64_t n_events;
  uint64_t seed;             /* Philox key; event e uses counter (e, block) */
  int32_t straggling;
  int32_t scattering;
  int32_t compute_gradient;
} dp_mc_config;

typedef struct {
  double flux, sigma;
  double gradient_flux, gradient;  /* NaN when compute_gradient == 0 */
  double source_energy_max;
  int64_t n_events;
} dp_mc_result;

/* Replaces: _compute_flux_estimate (src/Geometry.jl:818-896) and, with
 * compute_gradient, _sampled_csda_flux_gradient (:933-945) — the body of
 * run_backward_mc (:1286-1338).  straggling selects straggled stepping below
 * mode_switch and mixed stepping above (transport_backward_through_geometry,
 * :534-672); scattering adds EHS, DEL deflections and soft multiple
 * scattering.  If event_weights != NULL it receives each event's
 * contribution (n_events doubles, host memory). */
int dp_mc_flat_backward(dp_ctx ctx, dp_tables tables, const dp_flat_geometry* geo,
                        const dp_mc_config* cfg, dp_mc_result* out);
int dp_mc_flat_backward_events(dp_ctx ctx, dp_tables tables, const dp_flat_geometry* geo,
                               const dp_mc_config* cfg, dp_mc_result* out,
                               double* event_weights);

/* Replaces: compute_flux_csda_direct (src/Geometry.jl:1126-1209), and with
 * dflux_drho compute_flux_gradient_csda (:1247-1262), for n samples: a rock
 * slab of material `rock` (rock_density[i], rock_thickness[i] m) under an
 * exponential atmosphere of material `air` up to primary_altitude, muon
 * energy[i] GeV at the detector, elevation[i] degrees, charge[i].
 * flux[i] per GeV m^2 s sr; dflux_drho[i] = d flux / d rock_density (may be
 * NULL).  Samples whose energy leaves the tables give 0. */
int dp_csda_direct(dp_ctx ctx, dp_tables tables, int32_t rock, int32_t air,
                   double primary_altitude, int64_t n, const double* rock_density,
                   const double* rock_thickness, const double* elevation, const double* energy,
                   const double* charge, double* flux, double* dflux_drho);
int dp_csda_direct_dev(dp_ctx ctx, dp_tables tables, int32_t rock, int32_t air,
                       double primary_altitude, int64_t n, const double* rock_density,
                       const double* rock_thickness, const double* elevation,
                       const double* energy, const double* charge, double* flux,
                       double* dflux_drho);

/* ------------------------------------------------------------ topography
 * ElevationMap (src/Turtle.jl:469-480): nx x ny nodes from (x0, y0) with
 * spacings dx, dy in map units; z(ix, iy) = data[iy + ny * ix] (the storage
 * of Julia's Matrix{Float64} of size (ny, nx)).                            */
typedef struct {
  int32_t nx, ny;
  double x0, y0, dx, dy;
  const double* data;
} dp_elevation_map;

/* Replaces: ray_surface_exit_distance (examples/lvd_tomography.jl:219-252)
 * with local_surface_height (:214-217) and map_elevation (bilinear,
 * src/Turtle.jl:523-553) for n_rays unit directions dirs[3n] from the
 * detector.  Map coordinates are the local metres times xy_scale (1e-3 for a
 * map in km); heights are taken relative to z_ref (the detector elevation).
 * distance[i] (m), found[i], exit_z[i] = dirs[3i+2] * distance[i]; NaN and 0
 * when the ray points down or leaves the map first.  step 40 m and
 * max_distance 25 km are the reference defaults. */
int dp_ray_surface_exit(dp_ctx ctx, const dp_elevation_map* map, double xy_scale, double z_ref,
                        int32_t n_rays, const double* dirs, double step, double max_distance,
                        double* distance, uint8_t* found, double* exit_z);
/* Device pointers (map->data, dirs and the outputs); no synchronisation. */
int dp_ray_surface_exit_dev(dp_ctx ctx, const dp_elevation_map* map, double xy_scale,
                            double z_ref, int32_t n_rays, const double* dirs, double step,
                            double max_distance, double* distance, uint8_t* found,
                            double* exit_z);

/* ------------------------------- water-fraction tomography CSDA (w-model)
 * MaterialConfig (src/Tomography.jl:179-188), materials 0-based;
 * porous_material < 0 disables the shallow porous recipe.                  */
typedef struct {
  int32_t rock_material, water_material, air_material, porous_material;
  double rock_density, water_density, porous_density, porous_top_thickness;
} dp_material_config;

/* SiteConfig (src/Tomography.jl:198-201). */
typedef struct {
  double detector_elevation;
  double primary_altitude;
} dp_site_config;

/* DirectionalPath (src/Tomography.jl:214-223) for many bins in CSR form. */
typedef struct {
  int32_t n_rays;
  int32_t n_segments;
  int32_t max_segments;        /* max over rays of row_ptr[r+1]-row_ptr[r]  */
  const int32_t* row_ptr;      /* n_rays + 1 */
  const int32_t* cell;         /* 0-based cell of each CellSegment           */
  const double* length;        /* CellSegment distance, m                    */
  const double* cos_theta;     /* cosd(zenith) per ray                       */
  const double* remaining_rock;/* remaining_rock_distance per ray, m         */
  const double* exit_z;        /* surface_exit_z per ray, m                  */
  const uint8_t* valid;        /* per ray; NULL = all valid                  */
} dp_directional_paths;

/* CsdaCorrection (src/Tomography.jl:112-137).  Per energy node the flux is
 * multiplied by exp(kappa_strag b(b+1)/2 V/E^2 - kappa_hard max(H,0)/E)
 * (csda_var_correction, :384-391), E the surface energy floored at 1 GeV,
 * b the local spectral index of the surface flux (:367-372), V and H the
 * straggling and hard-loss integrals along the rock column; each bin by
 * exp(resid_a + resid_b logL + resid_c logL^2 + resid_d cos theta)
 * (csda_geom_multiplier, :405-412), L the rock path length in m.          */
typedef struct {
  int32_t enabled;
  double kappa_strag;
  double kappa_hard;
  double resid_a, resid_b, resid_c, resid_d;
} dp_csda_correction;

/* CsdaMaterialCorrection (src/Tomography.jl:164-166): each bin's flux is
 * multiplied by exp(linear q + quadratic q^2), q the water mass fraction of
 * the whole rock column (_column_water_fraction, :560-574).               */
typedef struct {
  int32_t enabled;
  double linear;
  double quadratic;
} dp_csda_material_correction;

/* Replaces: set_csda_correction! (src/Tomography.jl:149-155) and
 * set_csda_material_correction! (:168-172).  The context keeps the values
 * for its later w-model calls; NULL switches the correction off.
 * Non-finite material coefficients return DP_E_ARGUMENT (ArgumentError). */
int dp_set_csda_correction(dp_ctx ctx, const dp_csda_correction* corr);
int dp_set_csda_material_correction(dp_ctx ctx, const dp_csda_material_correction* corr);

/* Vector{EnergySample} (src/Tomography.jl:203-207). */
typedef struct {
  int32_t n;
  const double* energy;
  const double* weight;
  const double* charge;
} dp_energy_samples;

/* Replaces: compute_directional_flux_csda (src/Tomography.jl:667-726) for
 * every bin and, when jac != NULL, directional_flux_and_grad_csda /
 * assemble_forward_and_jacobian (:736-820) through the segment pullbacks of
 * src/Tomography_rules.jl:63-106 (exact derivatives of the executed Heun
 * sub-stepping).  water_fraction[n_cells] and shallow[n_cells] (0/1) are per
 * cell.  flux[r] is NaN for an invalid path or cos_theta <= 0.  jac has one
 * slot per path segment: d flux[r] / d w[cell] sits at the cell's first
 * occurrence on the path (the column order of _path_layout, :640-659), later
 * occurrences of the same cell hold 0.  The MC-calibrated corrections are
 * those last set on the context (dp_set_csda_correction,
 * dp_set_csda_material_correction); both are off by default, as in the
 * reference (:139, :167), and their derivatives enter jac when on. */
int dp_csda_w_flux_jacobian(dp_ctx ctx, dp_tables tables, const dp_material_config* mat,
                            const dp_site_config* site, const dp_directional_paths* paths,
                            int32_t n_cells, const double* water_fraction,
                            const uint8_t* shallow, const dp_energy_samples* samples,
                            double* flux, double* jac);
int dp_csda_w_flux_jacobian_dev(dp_ctx ctx, dp_tables tables, const dp_material_config* mat,
                                const dp_site_config* site, const dp_directional_paths* paths,
                                int32_t n_cells, const double* water_fraction,
                                const uint8_t* shallow, const dp_energy_samples* samples,
                                double* flux, double* jac);

/* ------------------------------------- fixed-path (tomography) backward MC
 * MaterialMixture (src/Types.jl:357-391): up to 4 components, mass
 * fractions (normalised by the library).                                   */
typedef struct {
  int32_t n_components;
  int32_t components[4];
  double fractions[4];
} dp_mixture;

/* Replaces: compute_directional_flux_mc (src/Tomography.jl:979-1041) for
 * every path in one persistent launch, with propagate_mc_uniform_segment /
 * propagate_mc_air_segment (:876-969).  Cell c is transported as
 * media[cell_medium[c]] at cell_density[c] (build_cell_properties_for_mc,
 * :826-861, or one rock medium with per-cell densities); the dry remainder
 * uses mat->rock_material at mat->rock_density and the atmosphere
 * mat->air_material at the altitude of the event (site).  Samples must carry
 * charges +/-1.  Event s of path r draws from Philox keyed (seed,
 * r * samples->n + s).  flux[r], sigma[r] (sigma from (E[w2]-E[w]^2)/(n-1))
 * are NaN for an invalid path or cos_theta <= 0.  event_weights, if not
 * NULL, receives the n_rays * samples->n contributions 2 w_s flux_s.
 * straggling = 0 runs the deterministic CSDA steps.                        */
int dp_mc_fixed_paths(dp_ctx ctx, dp_tables tables, const dp_material_config* mat,
                      const dp_site_config* site, const dp_directional_paths* paths,
                      int32_t n_cells, int32_t n_media, const dp_mixture* media,
                      const int32_t* cell_medium, const double* cell_density,
                      const dp_energy_samples* samples, uint64_t seed, int32_t straggling,
                      double energy_threshold_low, double* flux, double* sigma,
                      double* event_weights);

/* Replaces, in one persistent launch: compute_directional_flux_mc
 * (src/Tomography.jl:979-1041) of every path for the MC samples, as
 * dp_mc_fixed_paths with straggling on (flux, sigma: the same values bit for
 * bit), and the deterministic CSDA flux of every path over the quadrature
 * `nodes` (charge-summed nodes allowed): the same fixed-path stepping with
 * straggling off, csda_flux[r] = sum_n 2 w_n flux_n / n_nodes (NaN for an
 * invalid path).  The work queue holds both kinds of items (a path's chunks
 * of samples, then of nodes), so uneven path lengths and the two workloads
 * balance across the chip together.                                         */
int dp_mc_fixed_paths_csda(dp_ctx ctx, dp_tables tables, const dp_material_config* mat,
                           const dp_site_config* site, const dp_directional_paths* paths,
                           int32_t n_cells, int32_t n_media, const dp_mixture* media,
                           const int32_t* cell_medium, const double* cell_density,
                           const dp_energy_samples* samples, const dp_energy_samples* nodes,
                           uint64_t seed, double thr_low, double* flux, double* sigma,
                           double* csda_flux);

/* ---------------------------------------------- tetrahedral traversal
 * TetraVolume (examples/lvd_tomography.jl:172-179) as plain arrays, 0-based.
 * Face f of a tet is the triangle without its vertex 3 - f, i.e. the faces
 * (1,2,3), (1,2,4), (1,3,4), (2,3,4) of tetra_face_indices (:415-422).     */
typedef struct {
  int32_t n_points, n_tets;
  const double* points;      /* 3 * n_points: x y z                        */
  const int32_t* tets;       /* 4 * n_tets: vertex indices                 */
  const int32_t* neighbors;  /* 4 * n_tets: tet across face f, -1 = none   */
} dp_tet_mesh;

/* Replaces: build_tetra_neighbors (examples/lvd_tomography.jl:431-451): a face
 * shared by exactly two tets links them.  Host only (no device work).      */
int dp_tet_neighbors(int32_t n_tets, const int32_t* tets, int32_t* neighbors);

/* Replaces: trace_path(::TetraVolume) (examples/lvd_tomography.jl:694-736),
 * one ray per entry: from origin[3n] (the detector) along the unit dir[3n]
 * up to surface_distance[n] (ray_surface_exit_distance).  The start tet is
 * the first in index order that contains origin + 1e-4 dir.  Two-phase as
 * dp_trace_edges: with cells == NULL only counts[n] (segments per ray) and
 * traveled[n]; otherwise the segments (tet, length) at row_ptr.  A ray that
 * needs more than 4 n_tets + 16 steps (an inconsistent mesh) is an error.  */
int dp_trace_tetra(dp_ctx ctx, const dp_tet_mesh* mesh, int32_t n_rays, const double* origin,
                   const double* dir, const double* surface_distance, int32_t* counts,
                   const int32_t* row_ptr, int32_t* cells, double* lengths, double* traveled);

/* ------------------------------------- Gauss-Newton normal equations
 * Sparse matrix in CSR form, rows = bins, columns = cells (0-based).       */
typedef struct {
  int32_t n_rows, n_cols;
  int32_t nnz;               /* = row_ptr[n_rows]                          */
  const int32_t* row_ptr;    /* n_rows + 1, row_ptr[0] = 0                 */
  const int32_t* col;        /* nnz; strictly ascending within a row       */
  const double* val;         /* nnz                                        */
} dp_csr;

/* Replaces: the normal-equation assembly of gauss_newton_reconstruct
 * (src/Tomography_inverse.jl:489-494, final gradient :595) and the Fisher
 * information of laplace_posterior (:922-924):
 *   Ja = J[:, active]; normal = Matrix(Ja' * (W .* Ja)); grad = Ja' * (W .* r)
 * (weights == NULL: Matrix(Ja' * Ja) and Ja' * r).  Entries that are zero or
 * not finite are skipped, as assemble_forward_and_jacobian
 * (src/Tomography.jl:806-818) drops them.  active[n_active]: ascending,
 * distinct columns; NULL with n_active = n_cols means every column.
 * normal (n_active x n_active, column-major as Julia's Matrix) or grad may be
 * NULL.  Every entry is summed over rows in ascending order, the order of the
 * reference's sparse product, so results are deterministic.  The _dev form
 * takes device pointers (J's arrays, weights, residual, active, outputs),
 * trusts J's structure and returns without synchronising; there the columns
 * of a row need not be sorted (the w-model's segment-aligned Jacobian, whose
 * repeated cells carry zeros, can be passed as is) but a column may hold at
 * most one kept entry per row.                                              */
int dp_normal_equations(dp_ctx ctx, const dp_csr* J, const double* weights,
                        const double* residual, int32_t n_active, const int32_t* active,
                        double* normal, double* grad);
int dp_normal_equations_dev(dp_ctx ctx, const dp_csr* J, const double* weights,
                            const double* residual, int32_t n_active, const int32_t* active,
                            double* normal, double* grad);
/* Residual and data cost of a Gauss-Newton iterate on the device: the lines
 * r = pred .- obs and data_cost(r) = 0.5*sum(W .* r.^2) (0.5*dot(r, r)
 * without weights) of gauss_newton_reconstruct (src/Tomography_inverse.jl:
 * 478-482), so that a captured iteration (w-model flux and Jacobian, this,
 * the normal equations) needs no host round trip; no separate function of
 * the reference.  Device pointers: pred, obs, weights (NULL: all 1) of n
 * rows; residual (n) and cost (one double) may each be NULL.  The cost is
 * summed in a fixed order (deterministic); non-finite rows propagate into
 * it, as in the reference's objective.  Returns without synchronising.    */
int dp_residual_dev(dp_ctx ctx, int64_t n, const double* pred, const double* obs,
                    const double* weights, double* residual, double* cost);

/* Batched forms of two exported samplers the current transport path no
 * longer calls.  Sample i draws its uniforms, in the reference's order, from
 * Philox keyed (seed, i).  Host pointers; return after synchronising.
 *
 * Replaces: sample_mixture_material (src/Physics.jl:1331-1365): for each
 * energy, the material of an interaction in the mixture (1-8 components,
 * fractions normalised as MaterialMixture does), drawn by fraction x
 * tabulated hard cross-section (by fraction alone when that total is 0).  */
int dp_sample_mixture_material(dp_ctx ctx, dp_tables tables, int32_t n_components,
                               const int32_t* materials, const double* fractions, int64_t n,
                               const double* energies, uint64_t seed, int32_t* out);
/* Replaces: sample_del_event (src/Straggling.jl:217-317), the legacy
 * discrete-energy-loss sampler over a step of grammage dX ending at Xtot:
 * occurred (0/1), the interaction grammage X_del, the energy after the event
 * k_del and the process (1 brems, 2 pair, 3 photonuclear; 0 when none), with
 * the tables' cutoff as the lower end of the 1/nu^2 transfer law.          */
int dp_sample_del_event(dp_ctx ctx, dp_tables tables, int32_t material, int64_t n,
                        const double* ki, const double* kf, const double* dX, const double* Xtot,
                        const double* ratio, int32_t backward, uint64_t seed, uint8_t* occurred,
                        double* X_del, double* k_del, int32_t* process);

/* Last kernel duration in ms, timed with HIP events on the context's stream
 * around the dominant kernel of the previous call (0 if none). */
double dp_last_kernel_ms(dp_ctx ctx);

#ifdef __cplusplus
}
#endif
#endif /* DIFFPUMAS_HIP_H */
