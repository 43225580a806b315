"""diffpumas_hip: MI355X (gfx950) kernels for DiffPumas' backward muon
transport and tomography forward model, with a host mirror of the
reference's Julia API (grinisrit/diffpumas.jl).

All compute goes through libdiffpumas_hip.so; there is no CPU fallback.
"""
from ._lib import (ArgumentError, CapacityError, Context, DomainError, HipError,
                   TransportError, default_context, load)
from .physics import (PhysicsTables, create_physics, create_physics_csda, load_dedx_file, load_mdf,
                      MUON_MASS, MUON_C_TAU, builtin_materials, create_energy_grid,
                      electronic_stopping_power, electronic_density_effect)
from .cube import (CubeGeometry, RayPaths, compute_flux_gradient_cube, csda_cube_flux,
                   csda_cube_vjp, ray_direction, trace_cube_paths, trace_edge_paths)
from .topography import ElevationMap, ray_surface_exit_distance
from .inverse import make_csda_operator, normal_equations
from .calibration import calibrate_csda_correction, path_rock_length
from .tetra import TetraVolume, build_tetra_neighbors, trace_tetra_paths, trace_tetra_rays
from .geometry import (PRIMARY_ALTITUDE, compute_flux, run_backward_mc, compute_flux_csda_direct,
                       compute_flux_differentiable_csda, compute_flux_gradient_csda)
from .tomography import (DirectionalPaths, EnergySamples, MaterialConfig, SiteConfig,
                         assemble_forward_and_jacobian, cell_density,
                         compute_directional_flux_csda, csda_energy_quadrature,
                         direction_vector, directional_flux_and_grad_csda,
                         flux_and_segment_jacobian, sample_energy_set, trace_voxel_paths,
                         mc_cell_mixture_and_density, build_cell_properties_for_mc,
                         compute_directional_flux_mc, compute_directional_flux_mc_and_csda,
                         CsdaCorrection, CsdaMaterialCorrection,
                         set_csda_correction, get_csda_correction,
                         set_csda_material_correction, get_csda_material_correction)

from .sampling import sample_del_event, sample_mixture_material

__all__ = [
    "ArgumentError", "CapacityError", "Context", "DomainError", "HipError", "TransportError",
    "default_context", "load", "PhysicsTables", "create_physics", "create_physics_csda", "load_dedx_file",
    "load_mdf", "MUON_MASS", "MUON_C_TAU", "CubeGeometry", "RayPaths",
    "compute_flux_gradient_cube", "csda_cube_flux", "csda_cube_vjp", "ray_direction",
    "trace_cube_paths", "trace_edge_paths", "PRIMARY_ALTITUDE", "compute_flux",
    "run_backward_mc", "ElevationMap", "ray_surface_exit_distance", "compute_flux_csda_direct", "compute_flux_differentiable_csda",
    "compute_flux_gradient_csda", "DirectionalPaths", "EnergySamples", "MaterialConfig", "SiteConfig",
    "assemble_forward_and_jacobian", "cell_density", "compute_directional_flux_csda",
    "csda_energy_quadrature", "direction_vector", "directional_flux_and_grad_csda",
    "CsdaCorrection", "CsdaMaterialCorrection", "set_csda_correction", "get_csda_correction",
    "set_csda_material_correction", "get_csda_material_correction",
    "flux_and_segment_jacobian", "sample_energy_set", "trace_voxel_paths",
    "mc_cell_mixture_and_density", "build_cell_properties_for_mc", "compute_directional_flux_mc",
    "compute_directional_flux_mc_and_csda",
    "make_csda_operator", "normal_equations", "calibrate_csda_correction", "path_rock_length",
    "TetraVolume", "build_tetra_neighbors", "trace_tetra_paths", "trace_tetra_rays",
    "sample_del_event", "sample_mixture_material", "builtin_materials", "create_energy_grid",
    "electronic_stopping_power", "electronic_density_effect",
]
