"""ctypes binding of libdiffpumas_hip.so (the C ABI in include/diffpumas_hip.h).

The library must be built (diffpumas.jl_amd/build_lib.py); there is no
fallback path: every call goes to the HIP kernels, and a missing library or
GPU raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DIFFPUMAS_HIP_LIB selects another build of the library (e.g. a tuning variant).
LIB_PATH = os.environ.get("DIFFPUMAS_HIP_LIB") or os.path.join(os.path.dirname(_HERE),
                                                                "libdiffpumas_hip.so")

DP_OK = 0
DP_E_ARGUMENT = -1
DP_E_DOMAIN = -2
DP_E_EXHAUSTED = -3
DP_E_HIP = -4
DP_E_CAPACITY = -5

T_NARRAYS = 19
(T_ENERGY, T_LOG_ENERGY, T_CSDA_STOP, T_CSDA_RANGE, T_CSDA_TIME, T_MIXED_STOP, T_MIXED_RANGE,
 T_STRAGGLING, T_IONIZATION, T_BREMS, T_PAIR, T_PHOTO, T_CROSS_SECTION, T_TRANSPORT_PATH,
 T_ELASTIC_PATH, T_ELASTIC_CUTOFF, T_MAGNETIC_ROTATION, T_SCREENING, T_SPIN_FACTOR) = range(19)


class DomainError(ValueError):
    """Julia DomainError: an energy outside a prepared table."""


class ArgumentError(ValueError):
    """Julia ArgumentError: invalid mode, density, charge or size."""


class TransportError(RuntimeError):
    """Julia ErrorException from the transport (event/step caps)."""


class HipError(RuntimeError):
    """HIP runtime failure."""


class CapacityError(RuntimeError):
    """An output or scratch capacity was exceeded."""


_ERRORS = {DP_E_ARGUMENT: ArgumentError, DP_E_DOMAIN: DomainError,
           DP_E_EXHAUSTED: TransportError, DP_E_HIP: HipError, DP_E_CAPACITY: CapacityError}

dbl_p = C.POINTER(C.c_double)
i32_p = C.POINTER(C.c_int32)


class TableDesc(C.Structure):
    _fields_ = [("n_materials", C.c_int32), ("n_knots", C.c_int32), ("data", dbl_p),
                ("scalars", dbl_p), ("mass", C.c_double), ("ctau", C.c_double),
                ("n_elements", i32_p), ("element_Z", dbl_p), ("element_A", dbl_p),
                ("element_I", dbl_p), ("element_fraction", dbl_p), ("cutoff", C.c_double),
                ("elastic_ratio", C.c_double)]


class CubeGrid(C.Structure):
    _fields_ = [("x_min", C.c_double), ("x_max", C.c_double), ("y_min", C.c_double),
                ("y_max", C.c_double), ("z_min", C.c_double), ("z_max", C.c_double),
                ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double)]


class Paths(C.Structure):
    _fields_ = [("n_rays", C.c_int32), ("n_segments", C.c_int32), ("max_segments", C.c_int32),
                ("row_ptr", i32_p), ("cell", i32_p), ("length", dbl_p),
                ("cos_theta", dbl_p), ("z_exit", dbl_p)]


class CsdaNodes(C.Structure):
    _fields_ = [("rock_material", C.c_int32), ("air_material", C.c_int32),
                ("z_max", C.c_double), ("primary_altitude", C.c_double), ("scale", C.c_double),
                ("n_nodes", C.c_int32), ("node_energy", dbl_p), ("node_weight", dbl_p),
                ("node_charge", dbl_p)]


class FlatGeometry(C.Structure):
    _fields_ = [("rock_thickness", C.c_double), ("rock_density", C.c_double),
                ("rock_material", C.c_int32), ("air_material", C.c_int32),
                ("porous_material", C.c_int32), ("porous_density", C.c_double),
                ("porous_thickness", C.c_double), ("primary_altitude", C.c_double),
                ("rock_n_components", C.c_int32), ("rock_components", C.c_int32 * 4),
                ("rock_fractions", C.c_double * 4)]


class McConfig(C.Structure):
    _fields_ = [("elevation", C.c_double), ("cos_theta", C.c_double), ("sin_theta", C.c_double),
                ("energy_min", C.c_double), ("energy_max", C.c_double),
                ("energy_threshold", C.c_double), ("mode_switch", C.c_double),
                ("n_events", C.c_int64), ("seed", C.c_uint64), ("straggling", C.c_int32),
                ("scattering", C.c_int32), ("compute_gradient", C.c_int32)]


class McResult(C.Structure):
    _fields_ = [("flux", C.c_double), ("sigma", C.c_double), ("gradient_flux", C.c_double),
                ("gradient", C.c_double), ("source_energy_max", C.c_double),
                ("n_events", C.c_int64)]


class MaterialRecipe(C.Structure):
    _fields_ = [("n_elements", C.c_int32), ("Z", dbl_p), ("A", dbl_p), ("I", dbl_p),
                ("fraction", dbl_p), ("I_material", C.c_double), ("density", C.c_double),
                ("ZoA", C.c_double)]


class DedxColumns(C.Structure):
    _fields_ = [("n", C.c_int32), ("kinetic_energy", dbl_p), ("ionization", dbl_p),
                ("bremsstrahlung", dbl_p), ("pair_production", dbl_p), ("photonuclear", dbl_p),
                ("stopping_power", dbl_p), ("csda_range", dbl_p), ("ZoA", C.c_double)]


class PhysicsSettings(C.Structure):
    _fields_ = [("cutoff", C.c_double), ("elastic_ratio", C.c_double), ("mass", C.c_double)]


class MaterialConfig(C.Structure):
    _fields_ = [("rock_material", C.c_int32), ("water_material", C.c_int32),
                ("air_material", C.c_int32), ("porous_material", C.c_int32),
                ("rock_density", C.c_double), ("water_density", C.c_double),
                ("porous_density", C.c_double), ("porous_top_thickness", C.c_double)]


class SiteConfig(C.Structure):
    _fields_ = [("detector_elevation", C.c_double), ("primary_altitude", C.c_double)]


class DirectionalPaths(C.Structure):
    _fields_ = [("n_rays", C.c_int32), ("n_segments", C.c_int32), ("max_segments", C.c_int32),
                ("row_ptr", i32_p), ("cell", i32_p), ("length", dbl_p), ("cos_theta", dbl_p),
                ("remaining_rock", dbl_p), ("exit_z", dbl_p), ("valid", C.POINTER(C.c_uint8))]


class ElevationMap(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("x0", C.c_double), ("y0", C.c_double),
                ("dx", C.c_double), ("dy", C.c_double), ("data", C.c_void_p)]


class CsdaCorrection(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("kappa_strag", C.c_double), ("kappa_hard", C.c_double),
                ("resid_a", C.c_double), ("resid_b", C.c_double), ("resid_c", C.c_double),
                ("resid_d", C.c_double)]


class CsdaMaterialCorrection(C.Structure):
    _fields_ = [("enabled", C.c_int32), ("linear", C.c_double), ("quadratic", C.c_double)]


class EnergySamples(C.Structure):
    _fields_ = [("n", C.c_int32), ("energy", dbl_p), ("weight", dbl_p), ("charge", dbl_p)]


class Mixture(C.Structure):
    _fields_ = [("n_components", C.c_int32), ("components", C.c_int32 * 4),
                ("fractions", C.c_double * 4)]


class TetMesh(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("n_tets", C.c_int32), ("points", dbl_p),
                ("tets", i32_p), ("neighbors", i32_p)]


class Csr(C.Structure):
    _fields_ = [("n_rows", C.c_int32), ("n_cols", C.c_int32), ("nnz", C.c_int32),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("val", C.c_void_p)]


# Exported symbols and their argument types.
_SIGS = {
    "dp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "dp_destroy": (C.c_int, [C.c_void_p]),
    "dp_last_error": (C.c_char_p, []),
    "dp_set_stream": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dp_synchronize": (C.c_int, [C.c_void_p]),
    "dp_check_errors": (C.c_int, [C.c_void_p]),
    "dp_set_autotune": (C.c_int, [C.c_void_p, C.c_int32]),
    "dp_set_fp32_stepping": (C.c_int, [C.c_void_p, C.c_int32]),
    "dp_clear_tuning": (C.c_int, [C.c_void_p]),
    "dp_set_tuning": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32]),
    "dp_tuning_report": (C.c_int64, [C.c_void_p, C.c_char_p, C.c_int64]),
    "dp_device_info": (C.c_int, [C.c_void_p, i32_p, i32_p]),
    "dp_capture_begin": (C.c_int, [C.c_void_p]),
    "dp_capture_end": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "dp_graph_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dp_graph_destroy": (C.c_int, [C.c_void_p]),
    "dp_upload_tables": (C.c_int, [C.c_void_p, C.POINTER(TableDesc), C.POINTER(C.c_void_p)]),
    "dp_free_tables": (C.c_int, [C.c_void_p, C.c_void_p]),
    "dp_trace_cube": (C.c_int, [C.c_void_p, C.POINTER(CubeGrid), C.c_int32, dbl_p, dbl_p, i32_p,
                                i32_p, i32_p, dbl_p, dbl_p]),
    "dp_trace_cube_dev": (C.c_int, [C.c_void_p, C.POINTER(CubeGrid), C.c_int32, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p]),
    "dp_trace_edges": (C.c_int, [C.c_void_p, dbl_p, C.c_int32, dbl_p, C.c_int32, dbl_p,
                                 C.c_int32, C.c_int32, dbl_p, dbl_p, dbl_p, i32_p, i32_p, i32_p,
                                 dbl_p, dbl_p]),
    "dp_csda_cube": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Paths), C.POINTER(CsdaNodes),
                               dbl_p, C.c_int32, dbl_p, dbl_p, dbl_p]),
    "dp_csda_cube_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Paths),
                                   C.POINTER(CsdaNodes), C.c_void_p, C.c_int32, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    "dp_csda_cube_vjp": (C.c_int, [C.c_void_p, C.POINTER(Paths), dbl_p, dbl_p, dbl_p,
                                   C.c_int32, dbl_p]),
    "dp_csda_cube_vjp_dev": (C.c_int, [C.c_void_p, C.POINTER(Paths), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int32, C.c_void_p]),
    "dp_csda_cube_vjp_dev_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Paths), C.c_void_p, C.c_void_p,
                                       C.c_void_p, C.c_int32, C.c_void_p]),
    "dp_mc_flat_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FlatGeometry),
                                      C.POINTER(McConfig), C.POINTER(McResult)]),
    "dp_mc_flat_backward_events": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(FlatGeometry),
                                             C.POINTER(McConfig), C.POINTER(McResult), dbl_p]),
    "dp_last_kernel_ms": (C.c_double, [C.c_void_p]),
    "dp_mc_fixed_paths": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(MaterialConfig),
                                    C.POINTER(SiteConfig), C.POINTER(DirectionalPaths), C.c_int32,
                                    C.c_int32, C.POINTER(Mixture), i32_p, dbl_p,
                                    C.POINTER(EnergySamples), C.c_uint64, C.c_int32, C.c_double,
                                    dbl_p, dbl_p, dbl_p]),
    "dp_mc_fixed_paths_csda": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(MaterialConfig),
                                         C.POINTER(SiteConfig), C.POINTER(DirectionalPaths),
                                         C.c_int32, C.c_int32, C.POINTER(Mixture), i32_p, dbl_p,
                                         C.POINTER(EnergySamples), C.POINTER(EnergySamples),
                                         C.c_uint64, C.c_double, dbl_p, dbl_p, dbl_p]),
    "dp_ray_surface_exit": (C.c_int, [C.c_void_p, C.POINTER(ElevationMap), C.c_double,
                                      C.c_double, C.c_int32, C.POINTER(C.c_double), C.c_double,
                                      C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_uint8),
                                      C.POINTER(C.c_double)]),
    "dp_ray_surface_exit_dev": (C.c_int, [C.c_void_p, C.POINTER(ElevationMap), C.c_double,
                                          C.c_double, C.c_int32] + [C.c_void_p, C.c_double,
                                          C.c_double] + [C.c_void_p] * 3),
    "dp_csda_direct": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                 C.c_int64] + [C.POINTER(C.c_double)] * 7),
    "dp_csda_direct_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int32, C.c_double,
                                     C.c_int64] + [C.c_void_p] * 7),
    "dp_set_csda_correction": (C.c_int, [C.c_void_p, C.POINTER(CsdaCorrection)]),
    "dp_set_csda_material_correction": (C.c_int, [C.c_void_p, C.POINTER(CsdaMaterialCorrection)]),
    "dp_csda_w_flux_jacobian": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(MaterialConfig),
                                          C.POINTER(SiteConfig), C.POINTER(DirectionalPaths),
                                          C.c_int32, dbl_p, C.POINTER(C.c_uint8),
                                          C.POINTER(EnergySamples), dbl_p, dbl_p]),
    "dp_csda_w_flux_jacobian_dev": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(MaterialConfig),
                                              C.POINTER(SiteConfig), C.POINTER(DirectionalPaths),
                                              C.c_int32, C.c_void_p, C.c_void_p,
                                              C.POINTER(EnergySamples), C.c_void_p, C.c_void_p]),
    "dp_tet_neighbors": (C.c_int, [C.c_int32, i32_p, i32_p]),
    "dp_trace_tetra": (C.c_int, [C.c_void_p, C.POINTER(TetMesh), C.c_int32, dbl_p, dbl_p, dbl_p,
                                 i32_p, i32_p, i32_p, dbl_p, dbl_p]),
    "dp_normal_equations": (C.c_int, [C.c_void_p, C.POINTER(Csr)] + [C.c_void_p] * 2 +
                            [C.c_int32] + [C.c_void_p] * 3),
    "dp_normal_equations_dev": (C.c_int, [C.c_void_p, C.POINTER(Csr)] + [C.c_void_p] * 2 +
                                [C.c_int32] + [C.c_void_p] * 3),
    "dp_residual_dev": (C.c_int, [C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p]),
    "dp_build_material_table_empirical": (C.c_int, [C.POINTER(MaterialRecipe), dbl_p, C.c_int32,
                                                    C.POINTER(PhysicsSettings), dbl_p]),
    "dp_electronic_stopping_power": (C.c_double, [C.POINTER(MaterialRecipe), C.c_double,
                                                  C.c_double]),
    "dp_density_effect": (C.c_double, [C.POINTER(MaterialRecipe), C.c_double]),
    "dp_sample_mixture_material": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, i32_p, dbl_p,
                                             C.c_int64, dbl_p, C.c_uint64, i32_p]),
    "dp_sample_del_event": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int32, C.c_int64] + [dbl_p] * 5 +
                            [C.c_int32, C.c_uint64, C.POINTER(C.c_uint8), dbl_p, dbl_p, i32_p]),
    "dp_build_material_table": (C.c_int, [C.POINTER(MaterialRecipe), C.POINTER(DedxColumns),
                                          C.POINTER(PhysicsSettings), dbl_p]),
    "dp_build_composite_table": (C.c_int, [C.c_int32, C.POINTER(dbl_p), dbl_p, C.c_int32,
                                           C.POINTER(MaterialRecipe), C.POINTER(PhysicsSettings),
                                           dbl_p]),
    "dp_dcs": (C.c_double, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                            C.c_double, C.c_double]),
    "dp_dcs_integral": (C.c_double, [C.c_int32, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_double, C.c_int32, C.c_int32]),
}

_lib = None


def load():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with diffpumas.jl_amd/build_lib.py")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def exported_symbols():
    return sorted(_SIGS)


def check(rc):
    if rc != DP_OK:
        msg = load().dp_last_error().decode(errors="replace")
        raise _ERRORS.get(rc, RuntimeError)(msg)


def ptr(a, ctype=C.c_double):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ctype))


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class Context:
    """One HIP device context (stream, scratch, timing events)."""

    def __init__(self, device=0):
        lib = load()
        h = C.c_void_p()
        check(lib.dp_create(C.byref(h), int(device)))
        self.h = h
        self.lib = lib

    def close(self):
        if self.h:
            self.lib.dp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_handle):
        check(self.lib.dp_set_stream(self.h, C.c_void_p(stream_handle)))

    def synchronize(self):
        check(self.lib.dp_synchronize(self.h))

    def check_errors(self):
        """Raise what the kernels of earlier _dev calls (or graph replays)
        found: CapacityError for a path longer than its max_segments,
        ArgumentError for a cell index outside the field.  Clears the
        context's sticky error word."""
        check(self.lib.dp_check_errors(self.h))

    def set_autotune(self, on=True):
        """Time each candidate launch configuration of a tuned kernel on its
        first call per workload shape and keep the fastest (results are the
        same for every candidate)."""
        check(self.lib.dp_set_autotune(self.h, 1 if on else 0))

    def set_fp32_stepping(self, on=True):
        """FP32 segment arithmetic around an FP64 log-energy spine for the
        density CSDA flux and adjoint of this context (relative errors of a
        few 1e-6 against the FP64 program)."""
        check(self.lib.dp_set_fp32_stepping(self.h, 1 if on else 0))

    def clear_tuning(self):
        check(self.lib.dp_clear_tuning(self.h))

    def set_tuning(self, key, k0, k1=0, k2=0):
        """Pin the launch knobs of one (kernel, shape) key."""
        check(self.lib.dp_set_tuning(self.h, key.encode(), k0, k1, k2))

    def tuning(self):
        """Kept choices as {key: {"best": (k0, k1, k2), "ms": t, "tried":
        [((k0, k1, k2), ms), ...]}}."""
        out = {}
        for line in self.tuning_report().splitlines():
            key, best, tried = (x.strip() for x in line.split(" | "))
            _, knobs, ms, _ = (best + " x").split(" ", 3)
            tr = []
            for t in tried.split()[1:]:
                k, m = t.split(":")
                tr.append((tuple(int(x) for x in k.split(",")), float(m)))
            out[key] = {"best": tuple(int(x) for x in knobs.split(",")), "ms": float(ms),
                        "tried": tr}
        return out

    def tuning_report(self):
        """Kept choices, one line per (kernel, shape): key, best knobs and
        time, and every candidate's time."""
        n = int(self.lib.dp_tuning_report(self.h, None, 0))
        if n < 0:
            check(n)
        buf = C.create_string_buffer(n + 1)
        self.lib.dp_tuning_report(self.h, buf, n + 1)
        return buf.value.decode()

    def last_kernel_ms(self):
        return float(self.lib.dp_last_kernel_ms(self.h))

    def device_info(self):
        n = C.c_int32()
        w = C.c_int32()
        check(self.lib.dp_device_info(self.h, C.byref(n), C.byref(w)))
        return n.value, w.value

    def capture(self, work):
        """Run work() (calls of _dev entry points on this context) under
        stream capture and return the resulting Graph; nothing runs until
        Graph.launch().  Call the same work once eagerly first so that every
        buffer already has its size."""
        check(self.lib.dp_capture_begin(self.h))
        try:
            work()
        finally:
            g = C.c_void_p()
            rc = self.lib.dp_capture_end(self.h, C.byref(g))
        check(rc)
        return Graph(self, g)


class Graph:
    """A captured HIP graph of device work, replayed on its context's stream."""

    def __init__(self, ctx, handle):
        self.ctx = ctx
        self.h = handle

    def launch(self):
        check(self.ctx.lib.dp_graph_launch(self.ctx.h, self.h))

    def close(self):
        if self.h:
            self.ctx.lib.dp_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx
