"""The stochastic MC kernels of the library, compiled for the host (build_emu)
and run with the lanes of each block as threads, compared with the oracle
event by event.  Run as a child process with DIFFPUMAS_HIP_LIB pointing at the
host build; argument: a label printed with the results.

Cases (no scattering: the wave-cooperative routine needs lanes in step, which
the thread emulation does not model; coop_emu.cpp checks it in lockstep):
  * flat straggled + mixed MC over the 300 m rock/water/air mixture column
    (hard-event weights through the block's pool);
  * the same with a porous top layer;
  * fixed-path MC over a water-fraction field (path_mc_kernel).
Every event must equal the oracle's to 1e-9 relative, with no event zero on
one side only: the trajectories are the same step by step (energies,
positions, draws), and weights differ only where hard-event factors enter
them: the pool multiplies them in at the end of the round instead of inline,
and the library's DCS models regroup a few products that the oracle's
restatement keeps in the reference's order (about 1e-12 relative per hard
event).  Exit status 1 on a failure."""
import sys

import numpy as np

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.tomography import MaterialConfig, SiteConfig


def compare(name, got, ref):
    got = np.asarray(got).ravel()
    ref = np.asarray(ref).ravel()
    same = (np.abs(got - ref) <= 1e-9 * np.abs(ref)) | ((got == 0) & (ref == 0))
    one_zero = int(((got == 0) != (ref == 0)).sum())
    print(f"{name}: {same.sum()}/{same.size} events agree, {int((got == ref).sum())} bitwise, "
          f"{np.count_nonzero(ref)} non-zero, zero on one side only {one_zero}", flush=True)
    for i in np.flatnonzero(~same)[:5]:
        print(f"    event {i}: emulated {got[i]!r} oracle {ref[i]!r}", flush=True)
    return bool(same.all()) and one_zero == 0


def main():
    label = sys.argv[1] if len(sys.argv) > 1 else ""
    ctx = dp.Context(0)
    ph3 = dp.create_physics(("StandardRock", "Water", "Air"))
    ids = [ph3.get_material_index(m) for m in ("StandardRock", "Water", "Air")]
    mix = list(zip(ids, (0.85, 0.14, 0.01)))
    ok = True
    n = 640
    r = dp.run_backward_mc(ph3, rock_thickness=300.0, rock_density=2400.0, elevation=30.0,
                           energy_min=1.0, energy_max=1e4, n_samples=n, seed=2, straggling=True,
                           scattering=False, rock_mixture=mix, return_events=True, ctx=ctx)
    _, _, ev = O.mc_flat(ph3, ids, [0.85, 0.14, 0.01], 2400.0, 300.0, ids[2], 30.0, 1.0, 1e4, n,
                         2, scattering=False, per_event=True)
    ok &= compare(f"{label} flat mixture column", r["events"], ev)

    water = ids[1]
    r = dp.run_backward_mc(ph3, rock_thickness=120.0, rock_density=2650.0, elevation=50.0,
                           energy_min=2.0, energy_max=2e3, n_samples=320, seed=8,
                           scattering=False, porous_material=water, porous_density=1000.0,
                           porous_thickness=40.0, return_events=True, ctx=ctx)
    _, _, ev = O.mc_flat(ph3, [ids[0]], [1.0], 2650.0, 120.0, ids[2], 50.0, 2.0, 2e3, 320, 8,
                         scattering=False, porous=water, porous_rho=1000.0, porous_thick=40.0,
                         per_event=True)
    ok &= compare(f"{label} porous top layer", r["events"], ev)

    rng = np.random.default_rng(1)
    nx, ny, nz, size = 6, 6, 4, 10.0
    xe = np.linspace(-nx * size / 2, nx * size / 2, nx + 1)
    ye = np.linspace(-ny * size / 2, ny * size / 2, ny + 1)
    ze = np.linspace(0.0, nz * size, nz + 1)
    nd = 4
    zen = rng.uniform(5.0, 60.0, nd)
    az = rng.uniform(0.0, 360.0, nd)
    cz = np.cos(np.radians(zen))
    surface = (nz * size + rng.uniform(10.0, 60.0, nd)) / cz
    paths = dp.trace_voxel_paths(xe, ye, ze, zen, az, surface, surface * cz, ctx=ctx)
    n_cells = nx * ny * nz
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    w = np.random.default_rng(2).uniform(0.0, 0.4, n_cells)
    w[::5] = 0.0
    media, cm, cd = dp.build_cell_properties_for_mc(np.zeros(n_cells, bool), np.round(w, 1), m)
    s = dp.sample_energy_set(96, 1.0, 1e4, 5)
    _, _, evp = dp.compute_directional_flux_mc(ph3, m, SiteConfig(600.0, 3e4), paths, media, cm,
                                               cd, s, 11, return_events=True, ctx=ctx)
    op = dict(row_ptr=paths.row_ptr, cells=paths.cells, lengths=paths.lengths,
              remaining=paths.remaining_rock, exit_z=paths.exit_z, cos_theta=paths.cos_theta,
              valid=paths.valid)
    _, _, evo = O.mc_paths(ph3, media, cm, cd, 0, 2, 2650.0, (600.0, 3e4), op,
                           (s.energy, s.weight, s.charge), 11, per_event=True)
    ok &= compare(f"{label} fixed-path MC", evp, evo)
    ctx.close()
    print("ALL OK" if ok else "FAILED", flush=True)
    sys.exit(0 if ok else 1)


if __name__ == "__main__":
    main()
