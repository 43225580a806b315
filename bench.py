#!/usr/bin/env python3
"""Benchmark of the DiffPumas hot path on one MI355X.

Legs
  adjoint : deterministic CSDA flux + reverse-mode d flux/d rho over a
            128x128x64 voxel field of 10 m cells, densities U[2000, 3000]
            kg/m3 (seeded), 16384 rays (8 detectors x 2048 directions,
            elevation 15-75 deg), 256 log-spaced energy nodes 1-1e4 GeV
            (charge-summed, i.e. mu+ and mu-).  One step = forward + per-segment
            Jacobian + dense vector-Jacobian product.  Unit: rays*voxels/s
            (sum of segments over rays per step).
  mc      : straggled + mixed backward MC, single direction: 300 m column of a
            runtime StandardRock/Water/Air mixture (mass fractions 0.85, 0.14,
            0.01) at 2400 kg/m3, elevation 30 deg, E 1-1e4 GeV log-uniform,
            flux_gccly sky flux, Coulomb scattering and DEL/EHS hard events
            on, straggled below 100 GeV and mixed above, FP64 weights,
            Philox4x32-10 keyed by (seed, event).  Unit: events/s.  This is
            the headline value.

  wmodel  : water-fraction CSDA flux + sparse Jacobian (tomography forward
            model) over 4096 bins of a 64x64x32 grid, 32 energy nodes.
            Unit: rays*voxels/s.
  pathmc  : fixed-path straggled MC over all 4096 bins x 1e4 events of a
            lognormal density field with a cavity (4.1e7 events), plus each
            bin's deterministic CSDA flux over 256 charge-summed nodes, all in
            one persistent launch.  Unit: MC events/s (kernel time of the
            launch, CSDA items included).
  direct  : direct CSDA flux + density tangent for 2^24 samples.  Unit: samples/s.
  wmodel_corrected : the wmodel leg with the MC-calibrated corrections on.
  normal  : Gauss-Newton normal equations of the wmodel Jacobian over its
            12 869 free cells (dense result).  Unit: GB/s of the dense matrix.
  trace   : voxel traversal of the adjoint leg's 16384 rays.  Unit: rays*voxels/s.
  dem     : ray exit over a 257x257 elevation map, 16384 rays.  Unit: rays/s.
  tetra   : ray walk through a 331 776-tet mesh, 16384 rays.  Unit: rays*tets/s.

Inputs are resident in HBM before the timed region.  Timing: W warm-up
steps, then K steps between device synchronisations; kernel time of the
dominant kernel from HIP events on the stream it runs on.

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "diffpumas.jl_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

HBM_PEAK_GBS = 8000.0
# Measured HBM bytes per unit of work (FETCH_SIZE + WRITE_SIZE of one dispatch,
# scaled by what the counters read for a 1 GiB device copy), from the counter
# runs summarised in profiles/; `traffic` = this x the units of one launch.
MEASURED_BYTES_PER_UNIT = {
    "mc_flat_stochastic_kernel": 270402.0,    # per event
    "csda_cube_kernel": 8847.8,               # per ray-voxel (all 256 nodes)
    "csda_cube_kernel_fp32": 4701.2,          # per ray-voxel (FP32 stepping)
    "wcsda_kernel": 2121.7,                   # per ray-voxel (all 32 nodes)
    "wcsda_kernel_corrected": 4299.2,         # per ray-voxel (all 32 nodes)
    "path_mc_kernel": 255027.0,               # per event
    "csda_direct_kernel": 56.0,               # per sample
    "trace_cube_kernel": 13.3,                # per ray-voxel (fill launch)
    "ne_normal_kernel": 11.0,                 # per entry of the dense normal matrix
    "ray_exit_kernel": 137.8,                 # per ray
    "tet_trace_kernel": 63.2,                 # per ray-tet (fill launch)
}
FP64_PEAK_TFLOPS = 78.6
# VALU issue ceiling of an MI355X: 256 CUs x 4 SIMDs, 16 lanes each, so one
# wave64 VALU instruction per SIMD every 4 cycles at 2.4 GHz (the rate behind
# the 78.6 TFLOP/s FP64 vector peak: 16 FP64 FMA lanes per SIMD per cycle).
VALU_PEAK_WAVE_INSTS = 256 * 4 * 2.4e9 / 4
# Measured VALU wave instructions per unit of work (SQ_INSTS_VALU of one
# dispatch of the bench workload / its units, from the counter runs summarised
# in profiles/).  The kernels are FP64 VALU bound rather than HBM bound; the
# `valu` object prices the live kernel time against VALU_PEAK_WAVE_INSTS.
MEASURED_VALU_PER_UNIT = {
    "mc_flat_stochastic_kernel": 26415.9,     # per event
    "csda_cube_kernel": 938.5,                # per ray-voxel (all 256 nodes)
    "csda_cube_kernel_fp32": 1050.9,          # per ray-voxel (FP32 stepping)
    "wcsda_kernel": 6814.6,                   # per ray-voxel (all 32 nodes)
    "wcsda_kernel_corrected": 7065.9,         # per ray-voxel (all 32 nodes)
    "path_mc_kernel": 14097.0,                # per event
    "csda_direct_kernel": 36.02,              # per sample
    "trace_cube_kernel": 11.97,               # per ray-voxel (fill launch)
    "ray_exit_kernel": 318.4,                 # per ray
    "tet_trace_kernel": 30.35,                # per ray-tet (fill launch)
}


def valu_roofline(kernel, units, kern_ms):
    per = MEASURED_VALU_PER_UNIT.get(kernel)
    if per is None or kern_ms <= 0:
        return None
    achieved = per * units / (kern_ms * 1e-3)
    return {"bound": "valu", "achieved": achieved / 1e9, "peak": VALU_PEAK_WAVE_INSTS / 1e9,
            "unit": "G wave64 VALU instructions/s", "frac": achieved / VALU_PEAK_WAVE_INSTS,
            "instructions": per * units}


def host_threads():
    """Threads of the OpenMP CPU baselines (the oracle uses OMP_NUM_THREADS, default all cores)."""
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def adjoint_workload(dp, physics, ctx, torch):
    from workloads import cube_rays, energy_nodes, uniform_densities
    geo = dp.CubeGeometry(640.0, 128, 128, 64, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=1280.0)
    sxy, dirs, cos_t = cube_rays(geo, 8, 2048, seed=2024)
    paths = dp.trace_cube_paths(geo, sxy, dirs, cos_t, ctx=ctx)
    rho = uniform_densities(geo.n_cells, seed=7)
    ne, nw, nq = energy_nodes(256, 1.0, 1e4)
    return geo, paths, rho, ne, nw, nq


def run_adjoint(dp, L, physics, ctx, torch, steps, warmup, fp32=False):
    """fp32=True runs the FP32 stepping variant (dp_set_fp32_stepping) on the
    same workload."""
    import ctypes as C
    geo, paths, rho, ne, nw, nq = adjoint_workload(dp, physics, ctx, torch)
    ctx.set_fp32_stepping(fp32)
    dev = torch.device("cuda:0")
    T = lambda a, dt: torch.as_tensor(a, dtype=dt).to(dev)
    rp = T(paths.row_ptr, torch.int32)
    cell = T(paths.cell, torch.int32)
    length = T(paths.length, torch.float64)
    cos_t = T(paths.cos_theta, torch.float64)
    z_exit = T(paths.z_exit, torch.float64)
    d_rho = T(rho, torch.float64)
    d_ne, d_nw, d_nq = T(ne, torch.float64), T(nw, torch.float64), T(nq, torch.float64)
    flux = torch.zeros(paths.n_rays, dtype=torch.float64, device=dev)
    jac = torch.zeros(max(paths.n_segments, 1), dtype=torch.float64, device=dev)
    jrem = torch.zeros(paths.n_rays, dtype=torch.float64, device=dev)
    fbar = torch.ones(paths.n_rays, dtype=torch.float64, device=dev)
    grad = torch.zeros(geo.n_cells, dtype=torch.float64, device=dev)
    cp = L.Paths(paths.n_rays, paths.n_segments, paths.max_segments,
                 C.cast(rp.data_ptr(), L.i32_p), C.cast(cell.data_ptr(), L.i32_p),
                 C.cast(length.data_ptr(), L.dbl_p), C.cast(cos_t.data_ptr(), L.dbl_p),
                 C.cast(z_exit.data_ptr(), L.dbl_p))
    nodes = L.CsdaNodes(geo.rock_material, geo.air_material, geo.z_max, 3e4, 1.0, len(ne),
                        C.cast(d_ne.data_ptr(), L.dbl_p), C.cast(d_nw.data_ptr(), L.dbl_p),
                        C.cast(d_nq.data_ptr(), L.dbl_p))
    h = physics.device_handle(ctx)
    # a dedicated (non-null) stream shared by the library and the timing events
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)
    lib = ctx.lib

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        L.check(lib.dp_csda_cube_dev(ctx.h, h, C.byref(cp), C.byref(nodes), d_rho.data_ptr(),
                                     geo.n_cells, flux.data_ptr(), jac.data_ptr(),
                                     jrem.data_ptr()))
        if ev is not None:
            ev[1].record(stream)
        L.check(lib.dp_csda_cube_vjp_dev(ctx.h, C.byref(cp), jac.data_ptr(), jrem.data_ptr(),
                                         fbar.data_ptr(), geo.n_cells, grad.data_ptr()))

    log(f"adjoint: rays={paths.n_rays} segments={paths.n_segments} max_seg={paths.max_segments}")
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            step(evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / steps
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / steps
    ctx.set_stream(None)
    ctx.set_fp32_stepping(False)
    kname = "csda_cube_kernel_fp32" if fp32 else "csda_cube_kernel"
    units = paths.n_segments
    # algorithmic bytes of the forward+Jacobian kernel: 28 B per ray*voxel
    # (cell 4, length 8, density gather 8, Jacobian write 8) + 8 B per ray (flux)
    alg_bytes = 28.0 * units + 8.0 * paths.n_rays
    achieved = alg_bytes / (kern_ms * 1e-3) / 1e9
    flux_h = flux.cpu().numpy()
    return {
        "value": units / (ms_step * 1e-3), "ms_per_step": ms_step, "kernel_ms": kern_ms,
        "rays": paths.n_rays, "segments": units, "nodes": len(ne),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (MEASURED_BYTES_PER_UNIT[kname] * units
                                 if kname in MEASURED_BYTES_PER_UNIT else None)},
        "valu": valu_roofline(kname, units, kern_ms),
        "flux_sum": float(flux_h.sum()), "grad_norm": float(grad.norm().item()),
        "_problem": (geo, paths, rho, ne, nw, nq),
        # the node arrays stay referenced: `nodes` holds their device addresses
        "_device": dict(rp=rp, cell=cell, length=length, cos_t=cos_t, z_exit=z_exit, rho=d_rho,
                        ne=d_ne, nw=d_nw, nq=d_nq, nodes=nodes, h=h, flux=flux.clone(),
                        jac=jac.clone(), jrem=jrem.clone()),
    }


def run_adjoint_per_ray(dp, L, ctx, torch, adj):
    """The plain per-ray GPU launch loop the batched adjoint is measured
    against: the same forward + per-segment Jacobian, one dp_csda_cube_dev
    launch per ray (device-resident per-ray views of the same arrays), the
    shape of the reference's loop over rays (src/Tomography.jl:773-820).
    Every ray's flux and Jacobian must equal the batched launch bit for bit."""
    import ctypes as C
    geo, paths = adj["_problem"][0], adj["_problem"][1]
    d = adj["_device"]
    dev = torch.device("cuda:0")
    nr = paths.n_rays
    rp_h = paths.row_ptr.astype("int64")
    lens = rp_h[1:] - rp_h[:-1]
    rp1 = torch.zeros((nr, 2), dtype=torch.int32, device=dev)
    rp1[:, 1] = torch.as_tensor(lens, dtype=torch.int32, device=dev)
    flux = torch.zeros(nr, dtype=torch.float64, device=dev)
    jac = torch.zeros(max(paths.n_segments, 1), dtype=torch.float64, device=dev)
    jrem = torch.zeros(nr, dtype=torch.float64, device=dev)
    views = []
    for r in range(nr):
        a, n = int(rp_h[r]), int(lens[r])
        cp = L.Paths(1, n, max(n, 1), C.cast(rp1.data_ptr() + 8 * r, L.i32_p),
                     C.cast(d["cell"].data_ptr() + 4 * a, L.i32_p),
                     C.cast(d["length"].data_ptr() + 8 * a, L.dbl_p),
                     C.cast(d["cos_t"].data_ptr() + 8 * r, L.dbl_p),
                     C.cast(d["z_exit"].data_ptr() + 8 * r, L.dbl_p))
        views.append((cp, flux.data_ptr() + 8 * r, jac.data_ptr() + 8 * a, jrem.data_ptr() + 8 * r))
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)
    lib = ctx.lib

    def loop():
        for cp, fp, jp, jr in views:
            L.check(lib.dp_csda_cube_dev(ctx.h, d["h"], C.byref(cp), C.byref(d["nodes"]),
                                         d["rho"].data_ptr(), geo.n_cells, fp, jp, jr))

    with torch.cuda.stream(stream):
        loop()   # warm-up
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record(stream)
        loop()
        e1.record(stream)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) * 1e3
    ctx.set_stream(None)
    dev_ms = e0.elapsed_time(e1)
    same = bool(torch.equal(flux, d["flux"]) and torch.equal(jac, d["jac"]) and
                torch.equal(jrem, d["jrem"]))
    return {"rays": nr, "launches": nr, "device_ms": dev_ms, "wall_ms": wall,
            "value": paths.n_segments / (dev_ms * 1e-3), "unit": "rays*voxels/s",
            "bitwise_equal_to_batched": same,
            "speedup_of_batched": dev_ms / adj["kernel_ms"]}


def mc_kwargs(physics, n_events, ctx):
    ids = [physics.get_material_index(n) for n in ("StandardRock", "Water", "Air")]
    return dict(rock_thickness=300.0, rock_density=2400.0, elevation=30.0, energy_min=1.0,
                energy_max=1e4, n_samples=n_events, seed=1, straggling=True, scattering=True,
                rock_mixture=[(ids[0], 0.85), (ids[1], 0.14), (ids[2], 0.01)], ctx=ctx)


def run_mc(dp, physics, ctx, torch, n_events, steps, warmup):
    kw = mc_kwargs(physics, n_events, ctx)
    for _ in range(warmup):
        dp.run_backward_mc(physics, **kw)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    kms = 0.0
    for _ in range(steps):
        r = dp.run_backward_mc(physics, **kw)
        kms += ctx.last_kernel_ms()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / steps
    kern = kms / steps
    # algorithmic HBM bytes: the per-event contribution written (8 B) and
    # re-read by the reduction (8 B) = 16 B per event
    achieved = 16.0 * n_events / (kern * 1e-3) / 1e9
    return {"value": n_events / (ms * 1e-3), "ms_per_step": ms, "kernel_ms": kern,
            "events": n_events, "flux": r["flux"], "sigma": r["sigma"],
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["mc_flat_stochastic_kernel"] * n_events},
            "valu": valu_roofline("mc_flat_stochastic_kernel", n_events, kern)}


def cpu_mc_baseline(physics, budget_s=15.0):
    """The oracle's stochastic MC (same workload) on a bounded sample of
    events, all host cores."""
    import _oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    ids = [physics.get_material_index(n) for n in ("StandardRock", "Water", "Air")]
    n = 2000
    while True:
        t0 = time.perf_counter()
        O.mc_flat(physics, ids, [0.85, 0.14, 0.01], 2400.0, 300.0, ids[2], 30.0, 1.0, 1e4, n, 1,
                  threads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4:
            return {"value": n / dt, "unit": "events/s", "cores": threads, "kind": "port",
                    "sample": f"first {n} events of the same workload (seed 1), {dt:.2f} s"}
        n *= 2


def cpu_adjoint_baseline(physics, problem, budget_s=12.0):
    """The oracle's reverse-mode CPU restatement on a bounded sample of rays,
    all host cores (OpenMP)."""
    import numpy as np
    import _oracle as O
    geo, paths, rho, ne, nw, nq = problem
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = 16
    while True:
        idx = np.arange(0, paths.n_rays, max(1, paths.n_rays // n))[:n]
        rp = [0]
        cells, lens = [], []
        for r in idx:
            a, b = paths.row_ptr[r], paths.row_ptr[r + 1]
            cells.append(paths.cell[a:b]); lens.append(paths.length[a:b])
            rp.append(rp[-1] + (b - a))
        cells = np.concatenate(cells); lens = np.concatenate(lens)
        t0 = time.perf_counter()
        O.cube_csda_reverse(physics, geo.rock_material, geo.air_material, np.array(rp), cells,
                            lens, paths.cos_theta[idx], paths.z_exit[idx], geo.z_max, rho, ne, nw,
                            nq, threads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= paths.n_rays:
            return {"value": len(cells) / dt, "unit": "rays*voxels/s", "cores": threads,
                    "kind": "port",
                    "sample": f"{len(idx)} of {paths.n_rays} rays, all 256 nodes, "
                              f"{len(cells)} ray*voxels, {dt:.2f} s"}
        n *= 2


def tomo_workload(dp, n_dirs_per_det=1024, seed=31):
    """64x64x32 voxels of 10 m, 4 detectors 1 m above the floor of the block, directions
    with elevation U[15, 75] deg and azimuth U[0, 360); flat surface 50 m above
    the block.  Returns (edges, DirectionalPaths)."""
    import numpy as np
    nx, ny, nz, size = 64, 64, 32, 10.0
    xe = np.linspace(0.0, nx * size, nx + 1)
    ye = np.linspace(0.0, ny * size, ny + 1)
    ze = np.linspace(0.0, nz * size, nz + 1)
    rng = np.random.default_rng(seed)
    dets = [(160.0, 160.0, 1.0), (480.0, 160.0, 1.0), (160.0, 480.0, 1.0),
            (480.0, 480.0, 1.0)]
    zen, az, org = [], [], []
    for d in dets:
        zen.append(90.0 - rng.uniform(15.0, 75.0, n_dirs_per_det))
        az.append(rng.uniform(0.0, 360.0, n_dirs_per_det))
        org.append(np.tile(d, (n_dirs_per_det, 1)))
    zen, az, org = np.concatenate(zen), np.concatenate(az), np.concatenate(org)
    height = nz * size + 50.0 - org[:, 2]
    cz = np.cos(np.radians(zen))
    surface = height / cz
    paths = dp.trace_voxel_paths(xe, ye, ze, zen, az, surface, height, detectors=org)
    return (nx, ny, nz), paths


def lognormal_cavity(shape, seed=5):
    """rho = 2650 exp(0.1 N(0,1) - 0.005) with an ellipsoidal 1800 kg/m3 cavity."""
    import numpy as np
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    rho = 2650.0 * np.exp(0.1 * rng.standard_normal(nx * ny * nz) - 0.005)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    inside = (((i - 40) / 8.0) ** 2 + ((j - 28) / 6.0) ** 2 + ((k - 14) / 5.0) ** 2) <= 1.0
    rho[inside.ravel()] = 1800.0
    return rho


def _timed(torch, stream, step, steps, warmup):
    """W untimed steps, then K steps between device synchronisations on `stream`;
    returns (wall ms per step, mean ms between the events each step records
    around its dominant launch)."""
    with torch.cuda.stream(stream):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            step(evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, sum(a.elapsed_time(b) for a, b in evs) / steps


def run_trace(dp, L, physics, ctx, torch, steps, warmup):
    """trace_cube_path for the adjoint workload's 16384 rays (128x128x64 voxels
    of 10 m, 8 detectors x 2048 directions): count launch, device prefix sum of
    the counts, fill launch.  The fill launch is the dominant kernel: 12 B per
    ray*voxel written (cell 4, length 8) + 48 B per ray read (start 16,
    direction 24, row offset 8)."""
    import ctypes as C
    from workloads import cube_rays
    geo = dp.CubeGeometry(640.0, 128, 128, 64, physics.get_material_index("StandardRock"),
                          physics.get_material_index("Air"), width=1280.0)
    sxy, dirs, _ = cube_rays(geo, 8, 2048, seed=2024)
    n = len(dirs)
    dev = torch.device("cuda:0")
    t_sxy = torch.as_tensor(sxy, dtype=torch.float64).to(dev)
    t_dir = torch.as_tensor(dirs, dtype=torch.float64).to(dev)
    counts = torch.zeros(n, dtype=torch.int32, device=dev)
    z_exit = torch.zeros(n, dtype=torch.float64, device=dev)
    row_ptr = torch.zeros(n + 1, dtype=torch.int32, device=dev)
    grid = geo.c_grid()
    lib = ctx.lib
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def count():
        L.check(lib.dp_trace_cube_dev(ctx.h, C.byref(grid), n, t_sxy.data_ptr(), t_dir.data_ptr(),
                                      counts.data_ptr(), None, None, None, z_exit.data_ptr()))

    with torch.cuda.stream(stream):
        count()
        n_seg = int(counts.sum().item())
    cells = torch.zeros(max(n_seg, 1), dtype=torch.int32, device=dev)
    lens = torch.zeros(max(n_seg, 1), dtype=torch.float64, device=dev)

    def step(ev=None):
        count()
        row_ptr[1:].copy_(torch.cumsum(counts, 0, dtype=torch.int32))
        if ev is not None:
            ev[0].record(stream)
        L.check(lib.dp_trace_cube_dev(ctx.h, C.byref(grid), n, t_sxy.data_ptr(), t_dir.data_ptr(),
                                      None, row_ptr.data_ptr(), cells.data_ptr(), lens.data_ptr(),
                                      None))
        if ev is not None:
            ev[1].record(stream)

    ms_step, kern_ms = _timed(torch, stream, step, steps, warmup)
    ctx.set_stream(None)
    alg = 12.0 * n_seg + 48.0 * n
    achieved = alg / (kern_ms * 1e-3) / 1e9
    return {"value": n_seg / (ms_step * 1e-3), "unit": "rays*voxels/s", "ms_per_step": ms_step,
            "kernel_ms": kern_ms, "rays": n, "segments": n_seg,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["trace_cube_kernel"] * n_seg},
            "valu": valu_roofline("trace_cube_kernel", n_seg, kern_ms),
            "workload": "trace_cube_path, 128x128x64 voxels of 10 m, 16384 rays (8 det x 2048 "
                        "dirs, elevation 15-75 deg): count + prefix sum + fill",
            "_problem": (geo, sxy, dirs)}


def run_tetra(dp, L, ctx, steps, warmup):
    """trace_path(::TetraVolume) over a jittered Kuhn mesh of 48x48x24 cubes of
    10 m (331 776 tets), 4 detectors near the floor x 4096 directions (zenith
    U[0, 75] deg), topography 150-400 m away.  Host entry point: the mesh and
    rays are staged each call; `kernel_ms` is the device time of the locate and
    walk launches (count + fill), `value` ray*tets per second of kernel time."""
    import numpy as np
    from workloads import kuhn_mesh
    pts, tets = kuhn_mesh(48, 48, 24, 10.0, origin=(-240.0, -240.0, -5.0), jitter=0.1, seed=8)
    vol = dp.TetraVolume(pts, tets)
    rng = np.random.default_rng(21)
    dets = [(-60.0, -60.0, 0.0), (60.0, -60.0, 0.0), (-60.0, 60.0, 0.0), (60.0, 60.0, 0.0)]
    o, d = [], []
    for det in dets:
        zen = np.radians(rng.uniform(0.0, 75.0, 4096))
        az = np.radians(rng.uniform(0.0, 360.0, 4096))
        d.append(np.column_stack([np.sin(zen) * np.sin(az), np.sin(zen) * np.cos(az), np.cos(zen)]))
        o.append(np.tile(det, (4096, 1)))
    o, d = np.concatenate(o), np.concatenate(d)
    surf = rng.uniform(150.0, 400.0, len(o))
    for _ in range(warmup):
        dp.trace_tetra_rays(vol, o, d, surf, ctx=ctx)
    kms, t0 = 0.0, time.perf_counter()
    for _ in range(steps):
        rp, cells, lens, tr = dp.trace_tetra_rays(vol, o, d, surf, ctx=ctx)
        kms += ctx.last_kernel_ms()
    wall = (time.perf_counter() - t0) * 1e3 / steps
    kern = kms / steps     # the fill launch pair (locate + walk) of the last call per step
    ns = int(rp[-1])
    alg = 12.0 * ns + 56.0 * len(o)
    achieved = alg / (kern * 1e-3) / 1e9
    return {"value": ns / (kern * 1e-3), "unit": "rays*tets/s", "kernel_ms": kern,
            "wall_ms_incl_staging": wall, "rays": len(o), "segments": ns, "tets": len(tets),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["tet_trace_kernel"] * ns},
            "valu": valu_roofline("tet_trace_kernel", ns, kern),
            "workload": "trace_path(::TetraVolume), jittered Kuhn mesh 48x48x24 cubes of 10 m "
                        "(331776 tets), 4 detectors x 4096 directions",
            "_problem": (pts, tets, vol.neighbors, o, d, surf)}


def cpu_tetra_baseline(problem, budget_s=6.0):
    """The oracle's tetrahedral walk (start-tet search included) on a prefix of the rays, all host cores."""
    import _oracle as O
    pts, tets, nbr, o, d, surf = problem
    n = 64
    while True:
        t0 = time.perf_counter()
        rp = O.trace_tetra(pts, tets, nbr, o[:n], d[:n], surf[:n])[0]
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= len(o):
            return {"value": int(rp[-1]) / dt, "unit": "rays*tets/s", "cores": host_threads(), "kind": "port",
                    "sample": f"first {n} rays, {dt:.2f} s"}
        n *= 2


def run_dem(dp, L, ctx, torch, steps, warmup, n_dirs=16384):
    """ray_surface_exit_distance over a 257x257 diamond-square terrain (20 m
    nodes, 400 m relief on a 300 m dome above 1200 m), detector at the map
    centre at 963 m, n_dirs upward directions (zenith U[0, 75] deg).  41 B per
    ray (direction in, distance, flag and exit height out); the map (0.5 MB)
    stays in L2."""
    import ctypes as C
    import numpy as np
    from workloads import diamond_square
    n = 257
    d = 20.0
    xs = (np.arange(n) - (n - 1) / 2) * d
    X, Y = np.meshgrid(xs, xs)
    z = 1200.0 + diamond_square(8, 400.0, seed=17) + 300.0 * np.exp(-(X ** 2 + Y ** 2) / 1.5e6)
    emap = dp.ElevationMap(n, n, xs[0] * 1e-3, xs[0] * 1e-3, d * 1e-3, d * 1e-3, z)
    cmap, keep = emap.c()
    rng = np.random.default_rng(23)
    zen = np.radians(rng.uniform(0.0, 75.0, n_dirs))
    az = np.radians(rng.uniform(0.0, 360.0, n_dirs))
    dirs = np.stack([np.sin(zen) * np.sin(az), np.sin(zen) * np.cos(az), np.cos(zen)], axis=1)
    dev = torch.device("cuda:0")
    t_map = torch.as_tensor(keep.T.copy().ravel(), dtype=torch.float64).to(dev)
    cmap.data = t_map.data_ptr()
    t_dir = torch.as_tensor(dirs, dtype=torch.float64).to(dev)
    dist = torch.zeros(n_dirs, dtype=torch.float64, device=dev)
    found = torch.zeros(n_dirs, dtype=torch.uint8, device=dev)
    ez = torch.zeros(n_dirs, dtype=torch.float64, device=dev)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        L.check(ctx.lib.dp_ray_surface_exit_dev(ctx.h, C.byref(cmap), 1e-3, 963.0, n_dirs,
                                                t_dir.data_ptr(), 40.0, 25_000.0,
                                                dist.data_ptr(), found.data_ptr(), ez.data_ptr()))
        if ev is not None:
            ev[1].record(stream)

    ms_step, kern_ms = _timed(torch, stream, step, steps, warmup)
    ctx.set_stream(None)
    achieved = 41.0 * n_dirs / (kern_ms * 1e-3) / 1e9
    return {"value": n_dirs / (ms_step * 1e-3), "unit": "rays/s", "ms_per_step": ms_step,
            "kernel_ms": kern_ms, "rays": n_dirs, "found": float(found.float().mean().item()),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["ray_exit_kernel"] * n_dirs},
            "valu": valu_roofline("ray_exit_kernel", n_dirs, kern_ms),
            "workload": "ray_surface_exit_distance, 257x257 diamond-square DEM (20 m nodes), "
                        "16384 directions, 40 m march + 40 bisections",
            "_problem": (dict(nx=n, ny=n, x0=emap.x0, y0=emap.y0, dx=emap.dx, dy=emap.dy,
                              data=z), dirs)}


def run_wmodel(dp, L, physics, ctx, torch, steps, warmup, n_nodes=32, corrections=False):
    """w-model flux + sparse Jacobian over all tomography bins, device-resident.
    With corrections=True the MC-calibrated CsdaCorrection and
    CsdaMaterialCorrection are on (coefficients of the order a calibration
    gives), which runs the corrected kernel instantiation."""
    import ctypes as C
    import numpy as np
    shape, paths = tomo_workload(dp)
    n_cells = shape[0] * shape[1] * shape[2]
    w = np.random.default_rng(9).uniform(0.0, 0.3, n_cells)
    s = dp.csda_energy_quadrature(n_nodes, 1.0, 1e4)
    m = dp.MaterialConfig(physics.get_material_index("StandardRock"),
                          physics.get_material_index("Water"),
                          physics.get_material_index("Air"), -1, 2650.0, 1000.0)
    dev = torch.device("cuda:0")
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
    t_rp, t_cell, t_len = T(paths.row_ptr, torch.int32), T(paths.cells, torch.int32), T(paths.lengths, torch.float64)
    t_cos, t_rem, t_ez = T(paths.cos_theta, torch.float64), T(paths.remaining_rock, torch.float64), T(paths.exit_z, torch.float64)
    t_val = T(paths.valid, torch.uint8)
    t_w, t_sh = T(w, torch.float64), torch.zeros(n_cells, dtype=torch.uint8, device=dev)
    t_e, t_wt, t_q = T(s.energy, torch.float64), T(s.weight, torch.float64), T(s.charge, torch.float64)
    flux = torch.zeros(paths.n_rays, dtype=torch.float64, device=dev)
    jac = torch.zeros(max(paths.n_segments, 1), dtype=torch.float64, device=dev)
    dptr = lambda t, ty: C.cast(t.data_ptr(), ty)
    cp = L.DirectionalPaths(paths.n_rays, paths.n_segments, paths.max_segments(),
                            dptr(t_rp, L.i32_p), dptr(t_cell, L.i32_p), dptr(t_len, L.dbl_p),
                            dptr(t_cos, L.dbl_p), dptr(t_rem, L.dbl_p), dptr(t_ez, L.dbl_p),
                            dptr(t_val, C.POINTER(C.c_uint8)))
    es = L.EnergySamples(len(s.energy), dptr(t_e, L.dbl_p), dptr(t_wt, L.dbl_p), dptr(t_q, L.dbl_p))
    mc = m.c()
    site = L.SiteConfig(1000.0, 3e4)
    h = physics.device_handle(ctx)
    if corrections:
        L.check(ctx.lib.dp_set_csda_correction(ctx.h, C.byref(L.CsdaCorrection(
            1, 1.1, 0.6, 0.05, -0.02, 0.003, 0.1))))
        L.check(ctx.lib.dp_set_csda_material_correction(ctx.h, C.byref(
            L.CsdaMaterialCorrection(1, -0.8, 1.7))))
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        L.check(ctx.lib.dp_csda_w_flux_jacobian_dev(ctx.h, h, C.byref(mc), C.byref(site),
                                                    C.byref(cp), n_cells, t_w.data_ptr(),
                                                    t_sh.data_ptr(), C.byref(es),
                                                    flux.data_ptr(), jac.data_ptr()))
        if ev is not None:
            ev[1].record(stream)

    log(f"wmodel: rays={paths.n_rays} segments={paths.n_segments} max_seg={paths.max_segments()}")
    ms_step, kern_ms = _timed(torch, stream, step, steps, warmup)
    ctx.set_stream(None)
    if corrections:
        L.check(ctx.lib.dp_set_csda_correction(ctx.h, C.byref(L.CsdaCorrection(
            0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0))))
        L.check(ctx.lib.dp_set_csda_material_correction(ctx.h, C.byref(
            L.CsdaMaterialCorrection(0, 0.0, 0.0))))
    units = paths.n_segments
    alg = 29.0 * units + 8.0 * paths.n_rays
    achieved = alg / (kern_ms * 1e-3) / 1e9
    return {"value": units / (ms_step * 1e-3), "unit": "rays*voxels/s", "ms_per_step": ms_step,
            "kernel_ms": kern_ms, "rays": paths.n_rays, "segments": units, "nodes": n_nodes,
            "ray_nodes_per_s": paths.n_rays * n_nodes / (ms_step * 1e-3),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT[
                             "wcsda_kernel_corrected" if corrections else "wcsda_kernel"] * units},
            "valu": valu_roofline("wcsda_kernel_corrected" if corrections else "wcsda_kernel",
                                  units, kern_ms),
            "_problem": (paths, w, s, m)}


def run_inverse(dp, L, ctx, torch, steps, warmup, n_nodes=32, max_free=3000):
    """One model evaluation of the Gauss-Newton inverse loop at its full size
    (gauss_newton_reconstruct, src/Tomography_inverse.jl:448-603): 257x257
    diamond-square DEM, 96x96x48 voxels of 10 m, 4 detectors x 4096
    directions, water fractions w (0.03 with two buried anomalies, w0 = 0.1 on
    the free cells), Poisson-weighted residual.  The device work of one
    evaluation -- w-model flux and Jacobian, residual and data cost, normal
    matrix and gradient over the best-covered free cells -- is captured once
    as a HIP graph; a step is one replay.  Units: ray*voxels per evaluation."""
    import ctypes as C
    import numpy as np
    from workloads import inverse_workload
    ph = dp.create_physics_csda(("StandardRock", "Water", "Air"))
    paths, (xe, ye, ze), n_cells = inverse_workload(dp, ctx)
    m = dp.MaterialConfig(0, 1, 2, -1, ph.scalars[0, 0], ph.scalars[1, 0])
    nodes = dp.csda_energy_quadrature(n_nodes, 1.0, 1e4)
    site = dp.SiteConfig(963.0, 3e4)
    xc, yc, zc = (0.5 * (e[1:] + e[:-1]) for e in (xe, ye, ze))
    Zc, Yc, Xc = np.meshgrid(zc, yc, xc, indexing="ij")
    w_true = np.full(n_cells, 0.03)
    for ax, ay, az_, rad, wa in ((-150.0, -150.0, 1008.0, 30.0, 0.4),
                                 (150.0, 150.0, 1023.0, 35.0, 0.35)):
        w_true[((Xc - ax) ** 2 + (Yc - ay) ** 2 + (Zc - az_) ** 2).ravel() <= rad ** 2] = wa
    f_true, _ = dp.flux_and_segment_jacobian(ph, np.zeros(n_cells, np.uint8), m, site, paths,
                                             w_true, nodes, ctx=ctx)
    expo = 5e7
    obs = np.random.default_rng(13).poisson(f_true * expo) / expo
    W = expo / np.maximum(obs, 1.0 / expo)
    rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    pairs = np.unique(np.stack([paths.cells.astype(np.int64), rows]), axis=1)
    cross = np.bincount(pairs[0], minlength=n_cells)
    mc = 16
    while (cross >= mc).sum() > max_free:
        mc += 1
    active = np.flatnonzero(cross >= mc).astype(np.int32)
    na = len(active)
    w0 = w_true.copy()
    w0[active] = 0.1
    dev = torch.device("cuda:0")
    T = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt).to(dev)
    keep = dict(rp=T(paths.row_ptr, torch.int32), cell=T(paths.cells, torch.int32),
                len=T(paths.lengths, torch.float64), cos=T(paths.cos_theta, torch.float64),
                rem=T(paths.remaining_rock, torch.float64), ez=T(paths.exit_z, torch.float64),
                val=T(paths.valid, torch.uint8), w=T(w0, torch.float64),
                sh=torch.zeros(n_cells, dtype=torch.uint8, device=dev),
                e=T(nodes.energy, torch.float64), wt=T(nodes.weight, torch.float64),
                q=T(nodes.charge, torch.float64), obs=T(obs, torch.float64), W=T(W, torch.float64),
                act=T(active, torch.int32))
    nr, ns = paths.n_rays, paths.n_segments
    z = lambda n: torch.zeros(max(n, 1), dtype=torch.float64, device=dev)
    flux, jac, res, cost, N, g = z(nr), z(ns), z(nr), z(1), z(na * na), z(na)
    dptr = lambda t, ty: C.cast(t.data_ptr(), ty)
    k = keep
    cp = L.DirectionalPaths(nr, ns, paths.max_segments(), dptr(k["rp"], L.i32_p),
                            dptr(k["cell"], L.i32_p), dptr(k["len"], L.dbl_p),
                            dptr(k["cos"], L.dbl_p), dptr(k["rem"], L.dbl_p),
                            dptr(k["ez"], L.dbl_p), dptr(k["val"], C.POINTER(C.c_uint8)))
    es = L.EnergySamples(len(nodes.energy), dptr(k["e"], L.dbl_p), dptr(k["wt"], L.dbl_p),
                         dptr(k["q"], L.dbl_p))
    mcfg, sc = m.c(), L.SiteConfig(963.0, 3e4)
    csr = L.Csr(nr, n_cells, ns, k["rp"].data_ptr(), k["cell"].data_ptr(), jac.data_ptr())
    h = ph.device_handle(ctx)
    lib = ctx.lib
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def work():
        L.check(lib.dp_csda_w_flux_jacobian_dev(ctx.h, h, C.byref(mcfg), C.byref(sc), C.byref(cp),
                                                n_cells, k["w"].data_ptr(), k["sh"].data_ptr(),
                                                C.byref(es), flux.data_ptr(), jac.data_ptr()))
        L.check(lib.dp_residual_dev(ctx.h, nr, flux.data_ptr(), k["obs"].data_ptr(),
                                    k["W"].data_ptr(), res.data_ptr(), cost.data_ptr()))
        L.check(lib.dp_normal_equations_dev(ctx.h, C.byref(csr), k["W"].data_ptr(),
                                            res.data_ptr(), na, k["act"].data_ptr(),
                                            N.data_ptr(), g.data_ptr()))

    work()
    ctx.synchronize()
    ctx.check_errors()
    graph = ctx.capture(work)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        graph.launch()
        if ev is not None:
            ev[1].record(stream)

    log(f"inverse: {nr} bins, {ns} ray*voxels, {n_cells} cells, {na} free cells")
    ms_step, graph_ms = _timed(torch, stream, step, steps, warmup)
    c0 = float(cost.cpu()[0])
    graph.close()
    ctx.set_stream(None)
    ctx.check_errors()
    return {"value": ns / (graph_ms * 1e-3), "unit": "rays*voxels/s (per model evaluation)",
            "ms_per_step": ms_step, "graph_ms": graph_ms, "bins": nr, "segments": ns,
            "cells": n_cells, "free_cells": na, "nodes": n_nodes, "cost_at_w0": c0,
            "workload": "Gauss-Newton model evaluation (w-model flux + Jacobian, residual + cost, "
                        "normal equations) replayed from one HIP graph: 257x257 diamond-square "
                        "DEM, 96x96x48 voxels of 10 m, 4 detectors x 4096 directions, "
                        f"{na} free cells"}


def run_normal(dp, L, physics, ctx, torch, steps, warmup, min_crossings=4):
    """Gauss-Newton normal equations Ja' (W .* Ja) and Ja' (W .* r) of the
    w-model Jacobian of the tomography workload (4096 bins, 32 nodes), free
    cells = those crossed by at least min_crossings rays, W = 1/obs with obs
    the flux perturbed by 5 %.  The Jacobian is passed in its segment-aligned
    device form.  The dense result dominates: 8 n_free^2 bytes written."""
    import ctypes as C
    import numpy as np
    shape, paths = tomo_workload(dp)
    n_cells = shape[0] * shape[1] * shape[2]
    w = np.random.default_rng(9).uniform(0.0, 0.3, n_cells)
    s = dp.csda_energy_quadrature(32, 1.0, 1e4)
    m = dp.MaterialConfig(physics.get_material_index("StandardRock"),
                          physics.get_material_index("Water"),
                          physics.get_material_index("Air"), -1, 2650.0, 1000.0)
    pred, jac = dp.flux_and_segment_jacobian(physics, np.zeros(n_cells, np.uint8), m,
                                             dp.SiteConfig(1000.0, 3e4), paths, w, s, ctx=ctx)
    rng = np.random.default_rng(44)
    ok = np.isfinite(pred)
    pred = np.where(ok, pred, 0.0)
    obs = pred * (1.0 + 0.05 * rng.standard_normal(len(pred)))
    W = np.where(ok, 1.0 / np.maximum(obs, 1e-30), 0.0)
    res = pred - obs
    kept = np.isfinite(jac) & (jac != 0.0)
    crossings = np.bincount(paths.cells[kept], minlength=n_cells)
    active = np.flatnonzero(crossings >= min_crossings).astype(np.int32)
    na = len(active)
    dev = torch.device("cuda:0")
    T = lambda a, t: torch.as_tensor(np.ascontiguousarray(a), dtype=t).to(dev)
    t_rp, t_col = T(paths.row_ptr, torch.int32), T(paths.cells, torch.int32)
    t_val, t_W, t_r = T(jac, torch.float64), T(W, torch.float64), T(res, torch.float64)
    t_act = T(active, torch.int32)
    normal = torch.empty(na * na, dtype=torch.float64, device=dev)
    grad = torch.empty(na, dtype=torch.float64, device=dev)
    csr = L.Csr(paths.n_rays, n_cells, paths.n_segments, t_rp.data_ptr(), t_col.data_ptr(),
                t_val.data_ptr())
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        L.check(ctx.lib.dp_normal_equations_dev(ctx.h, C.byref(csr), t_W.data_ptr(),
                                                t_r.data_ptr(), na, t_act.data_ptr(),
                                                normal.data_ptr(), grad.data_ptr()))
        if ev is not None:
            ev[1].record(stream)

    ms_step, kern_ms = _timed(torch, stream, step, steps, warmup)
    ctx.set_stream(None)
    alg = 8.0 * na * na + 8.0 * na + 16.0 * paths.n_segments
    achieved = alg / (kern_ms * 1e-3) / 1e9
    return {"value": alg / (ms_step * 1e-3) / 1e9, "unit": "GB/s of normal matrix",
            "ms_per_step": ms_step, "kernel_ms": kern_ms, "n_free": na,
            "bins": paths.n_rays, "entries": int(kept.sum()),
            "grad_norm": float(grad.norm().item()),
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["ne_normal_kernel"] * na * na},
            "workload": f"normal equations of the w-model Jacobian (64x64x32, 4096 bins, 32 "
                        f"nodes), {na} free cells (>= {min_crossings} crossings), weighted",
            "_problem": (paths, jac, W, res, active, n_cells)}


def cpu_normal_baseline(problem):
    """The oracle's normal equations on the same system, all host cores (OpenMP over columns)."""
    import numpy as np
    import scipy.sparse as sp
    import _oracle as O
    paths, jac, W, res, active, n_cells = problem
    kept = np.isfinite(jac) & (jac != 0.0)
    rows = np.repeat(np.arange(paths.n_rays), np.diff(paths.row_ptr))
    J = sp.csr_matrix((jac[kept], (rows[kept], paths.cells[kept])), shape=(paths.n_rays, n_cells))
    na = len(active)
    t0 = time.perf_counter()
    O.normal_equations(J, res, W, active)
    dt = time.perf_counter() - t0
    return {"value": 8.0 * na * na / dt / 1e9, "unit": "GB/s of normal matrix", "cores": host_threads(),
            "kind": "port", "sample": f"the full system ({na} free cells), {dt:.2f} s"}


def run_direct(dp, L, physics, ctx, torch, steps, warmup, n=1 << 24):
    """Direct CSDA flux + density tangent (compute_flux_gradient_csda) for n
    device-resident samples: 56 B per sample (5 inputs, 2 outputs)."""
    import numpy as np
    rng = np.random.default_rng(12)
    dev = torch.device("cuda:0")
    T = lambda a: torch.as_tensor(a, dtype=torch.float64).to(dev)
    rho, thick = T(rng.uniform(2000.0, 3000.0, n)), T(rng.uniform(10.0, 800.0, n))
    elev = T(rng.uniform(10.0, 90.0, n))
    energy = T(np.exp(rng.uniform(0.0, np.log(1e4), n)))
    charge = T(np.where(rng.random(n) < 0.5, -1.0, 1.0))
    flux = torch.zeros(n, dtype=torch.float64, device=dev)
    dflux = torch.zeros(n, dtype=torch.float64, device=dev)
    rock, air = physics.get_material_index("StandardRock"), physics.get_material_index("Air")
    h = physics.device_handle(ctx)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    ctx.set_stream(stream.cuda_stream)

    def step(ev=None):
        if ev is not None:
            ev[0].record(stream)
        L.check(ctx.lib.dp_csda_direct_dev(ctx.h, h, rock, air, 3e4, n, rho.data_ptr(),
                                           thick.data_ptr(), elev.data_ptr(), energy.data_ptr(),
                                           charge.data_ptr(), flux.data_ptr(), dflux.data_ptr()))
        if ev is not None:
            ev[1].record(stream)

    with torch.cuda.stream(stream):
        for _ in range(warmup):
            step()
        torch.cuda.synchronize()
        evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
               for _ in range(steps)]
        t0 = time.perf_counter()
        for k in range(steps):
            step(evs[k])
        torch.cuda.synchronize()
        t1 = time.perf_counter()
    ctx.set_stream(None)
    ms_step = (t1 - t0) * 1e3 / steps
    kern_ms = sum(a.elapsed_time(b) for a, b in evs) / steps
    achieved = 56.0 * n / (kern_ms * 1e-3) / 1e9
    host = [x.cpu().numpy() for x in (rho, thick, elev, energy, charge)]
    return {"value": n / (ms_step * 1e-3), "unit": "samples/s", "ms_per_step": ms_step,
            "kernel_ms": kern_ms, "samples": n,
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                         "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                         "traffic": MEASURED_BYTES_PER_UNIT["csda_direct_kernel"] * n},
            "valu": valu_roofline("csda_direct_kernel", n, kern_ms),
            "_problem": (rock, air, host)}


def cpu_direct_baseline(physics, problem, budget_s=6.0):
    """The oracle's direct CSDA flux + tangent, all host cores, on a prefix of the samples."""
    import _oracle as O
    rock, air, host = problem
    n = 1 << 14
    while True:
        t0 = time.perf_counter()
        O.csda_direct(physics, rock, air, *(x[:n] for x in host))
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= len(host[0]):
            return {"value": n / dt, "unit": "samples/s", "cores": host_threads(), "kind": "port",
                    "sample": f"first {n} samples, {dt:.2f} s"}
        n *= 2


def run_pathmc(dp, physics, ctx, n_rays=4096, n_events=10000, steps=2, warmup=1):
    """Fixed-path straggled MC over the lognormal-cavity density field: n_rays
    bins of the tomography workload (all 4096 by default), n_events per bin,
    one launch per step; `steps` timed launches after `warmup`."""
    import numpy as np
    shape, paths_all = tomo_workload(dp)
    idx = np.linspace(0, paths_all.n_rays - 1, n_rays).astype(int)
    rp = [0]
    cells, lens = [], []
    for r in idx:
        a, b = paths_all.row_ptr[r], paths_all.row_ptr[r + 1]
        cells.append(paths_all.cells[a:b]); lens.append(paths_all.lengths[a:b])
        rp.append(rp[-1] + (b - a))
    paths = dp.DirectionalPaths(rp, np.concatenate(cells), np.concatenate(lens),
                                paths_all.zenith[idx], paths_all.azimuth[idx],
                                paths_all.remaining_rock[idx], paths_all.exit_z[idx])
    rho = lognormal_cavity(shape)
    s = dp.sample_energy_set(n_events, 1.0, 1e4, 17)
    m = dp.MaterialConfig(physics.get_material_index("StandardRock"),
                          physics.get_material_index("Water"),
                          physics.get_material_index("Air"), -1, 2650.0, 1000.0)
    site = dp.SiteConfig(1000.0, 3e4)
    nodes = dp.csda_energy_quadrature(256, 1.0, 1e4)
    args = (physics, m, site, paths, [((m.rock_material,), (1.0,))],
            np.zeros(len(rho), np.int32), rho, s, nodes, 1)
    for _ in range(warmup):
        dp.compute_directional_flux_mc_and_csda(*args, ctx=ctx)
    kms, t0 = 0.0, time.perf_counter()
    for _ in range(steps):
        f, sg, cf = dp.compute_directional_flux_mc_and_csda(*args, ctx=ctx)
        kms += ctx.last_kernel_ms()
    t1 = time.perf_counter()
    kern = kms / steps
    ev = n_rays * n_events
    return {"value": ev / (kern * 1e-3), "unit": "events/s", "kernel_ms": kern,
            "wall_ms_incl_staging": (t1 - t0) * 1e3 / steps, "rays": n_rays,
            "events_per_ray": n_events, "steps": steps,
            "segments": paths.n_segments, "mean_flux": float(np.nanmean(f)),
            "csda_nodes_per_ray": len(nodes.energy), "mean_csda_flux": float(np.nanmean(cf)),
            "roofline": {"bound": "hbm", "achieved": (16.0 * ev + 12.0 * paths.n_segments) /
                         (kern * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": (16.0 * ev + 12.0 * paths.n_segments) / (kern * 1e-3) / 1e9 /
                         HBM_PEAK_GBS, "traffic": MEASURED_BYTES_PER_UNIT["path_mc_kernel"] * ev},
            "valu": valu_roofline("path_mc_kernel", ev, kern),
            "_problem": (paths, rho, s, m, site)}


def run_pathmc_per_ray(dp, physics, ctx, pm, n_sub=64):
    """The plain per-ray GPU launch loop for the fixed-path MC: one launch per
    ray with its 1e4 events and 256 CSDA nodes (the reference's loop over
    bins), on n_sub rays spread over the 4096; device time of the launches
    (staging excluded), scaled to all rays."""
    import numpy as np
    paths, rho, s, m, site = pm["_problem"]
    nodes = dp.csda_energy_quadrature(256, 1.0, 1e4)
    idx = np.linspace(0, paths.n_rays - 1, n_sub).astype(int)
    kms = 0.0
    t0 = time.perf_counter()
    for r in idx:
        a, b = paths.row_ptr[r], paths.row_ptr[r + 1]
        one = dp.DirectionalPaths([0, b - a], paths.cells[a:b], paths.lengths[a:b],
                                  paths.zenith[r:r + 1], paths.azimuth[r:r + 1],
                                  paths.remaining_rock[r:r + 1], paths.exit_z[r:r + 1])
        dp.compute_directional_flux_mc_and_csda(physics, m, site, one,
                                                [((m.rock_material,), (1.0,))],
                                                np.zeros(len(rho), np.int32), rho, s, nodes, 1,
                                                ctx=ctx)
        kms += ctx.last_kernel_ms()
    wall = (time.perf_counter() - t0) * 1e3
    scale = paths.n_rays / n_sub
    return {"rays_timed": n_sub, "rays": paths.n_rays, "events_per_ray": len(s.energy),
            "device_ms_all_rays": kms * scale, "wall_ms_all_rays_incl_staging": wall * scale,
            "value": paths.n_rays * len(s.energy) / (kms * scale * 1e-3), "unit": "events/s",
            "speedup_of_batched": kms * scale / pm["kernel_ms"]}


def cpu_wmodel_baseline(physics, problem, budget_s=10.0, corr=None):
    """The oracle's w-model flux + reverse Jacobian on a bounded sample of bins
    (corrections on when corr is given)."""
    import numpy as np
    import _oracle as O
    paths, w, s, m = problem
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    mc = O.matcfg_arrays(m.rock_material, m.water_material, m.air_material, m.porous_material,
                         m.rock_density, m.water_density)
    n = 8
    while True:
        idx = np.arange(0, paths.n_rays, max(1, paths.n_rays // n))[:n]
        rp = [0]
        cells, lens = [], []
        for r in idx:
            a, b = paths.row_ptr[r], paths.row_ptr[r + 1]
            cells.append(paths.cells[a:b]); lens.append(paths.lengths[a:b])
            rp.append(rp[-1] + (b - a))
        sub = dict(row_ptr=np.array(rp, np.int32), cells=np.concatenate(cells),
                   lengths=np.concatenate(lens), remaining=paths.remaining_rock[idx],
                   exit_z=paths.exit_z[idx], cos_theta=paths.cos_theta[idx],
                   valid=paths.valid[idx])
        t0 = time.perf_counter()
        O.w_jacobian(physics, mc, (1000.0, 3e4), sub, w, np.zeros(len(w), np.uint8), s.energy,
                     s.weight, s.charge, threads=threads, corr=corr)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= paths.n_rays:
            return {"value": rp[-1] / dt, "unit": "rays*voxels/s", "cores": threads,
                    "kind": "port", "sample": f"{len(idx)} of {paths.n_rays} bins, "
                                              f"{len(s.energy)} nodes, {dt:.2f} s"}
        n *= 2


def _wcorr_oracle():
    """The corrected leg's coefficients in the oracle's layout."""
    import _oracle as O
    return O.corr_array(enabled=True, kappa_strag=1.1, kappa_hard=0.6, resid_a=0.05,
                        resid_b=-0.02, resid_c=0.003, resid_d=0.1, material=(-0.8, 1.7))


def cpu_trace_baseline(problem):
    """The oracle's trace_cube_path over all rays of the trace leg, all host cores."""
    import _oracle as O
    geo, sxy, dirs = problem
    t0 = time.perf_counter()
    rp, cells, lens, ze = O.trace_cube(geo, sxy, dirs)
    dt = time.perf_counter() - t0
    return {"value": len(cells) / dt, "unit": "rays*voxels/s", "cores": host_threads(), "kind": "port",
            "sample": f"all {len(dirs)} rays, {len(cells)} ray*voxels, {dt:.2f} s"}


def cpu_dem_baseline(problem, budget_s=6.0):
    """The oracle's ray_surface_exit_distance on a prefix of the directions, all host cores."""
    import _oracle as O
    dem, dirs = problem
    n = 256
    while True:
        t0 = time.perf_counter()
        O.ray_surface_exit(dem, dirs[:n], z_ref=963.0)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= len(dirs):
            return {"value": n / dt, "unit": "rays/s", "cores": host_threads(), "kind": "port",
                    "sample": f"first {n} directions, {dt:.2f} s"}
        n *= 2


def cpu_pathmc_baseline(physics, problem, budget_s=10.0):
    import numpy as np
    import _oracle as O
    paths, rho, s, m, site = problem
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n = 256
    while True:
        t0 = time.perf_counter()
        sub = dict(row_ptr=paths.row_ptr, cells=paths.cells,
                   lengths=paths.lengths, remaining=paths.remaining_rock, exit_z=paths.exit_z,
                   cos_theta=paths.cos_theta, valid=paths.valid)
        O.mc_paths(physics, [([m.rock_material], [1.0])], np.zeros(len(rho), np.int32), rho,
                   m.rock_material, m.air_material, m.rock_density,
                   (site.detector_elevation, site.primary_altitude), sub,
                   (s.energy[:n], s.weight[:n], s.charge[:n]), 1, threads=threads)
        dt = time.perf_counter() - t0
        if dt > budget_s / 4 or n >= len(s.energy):
            return {"value": paths.n_rays * n / dt, "unit": "events/s", "cores": threads,
                    "kind": "port", "sample": f"{paths.n_rays} bins x {n} events, {dt:.2f} s"}
        n *= 2


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mc-events", type=int, default=10_000_000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="run every kernel at its default launch shape (profiling runs)")
    ap.add_argument("--no-per-ray", action="store_true",
                    help="skip the per-ray launch-loop comparators (counter runs)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a 1 GiB device copy (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(events, rays and bins are independent and would shard by range, one "
              "process per GPU).", file=sys.stderr)
        sys.exit(2)
    import torch
    import diffpumas_hip as dp
    from diffpumas_hip import _lib as L
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    ctx = dp.Context(0)
    # The first (warm-up) call of each tuned kernel times its candidate launch
    # configurations; the timed steps run the fastest.
    tune = not a.no_autotune
    ctx.set_autotune(tune)
    physics = dp.create_physics(("StandardRock", "Water", "Air"))
    mc = run_mc(dp, physics, ctx, torch, a.mc_events, a.steps, a.warmup)
    log(f"mc(straggled): {mc['value']:.4g} events/s, kernel {mc['kernel_ms']:.3f} ms, "
        f"flux {mc['flux']:.6g} +- {mc['sigma']:.3g}")
    adj = run_adjoint(dp, L, physics, ctx, torch, a.steps, a.warmup)
    log(f"adjoint: {adj['value']:.4g} rays*voxels/s, {adj['ms_per_step']:.3f} ms/step, "
        f"kernel {adj['kernel_ms']:.3f} ms")
    ctx.set_autotune(False)   # the per-ray comparators run the default launch shape
    per_ray_adj = None if a.no_per_ray else run_adjoint_per_ray(dp, L, ctx, torch, adj)
    ctx.set_autotune(tune)
    a32 = run_adjoint(dp, L, physics, ctx, torch, a.steps, a.warmup, fp32=True)
    f64d, f32d = adj["_device"], a32.pop("_device")
    a32.pop("_problem")
    fe = ((f32d["flux"] - f64d["flux"]).abs() / f64d["flux"].abs()).max().item()
    ge = ((f32d["jac"] - f64d["jac"]).abs().max() / f64d["jac"].abs().max()).item()
    adj32 = {k: a32[k] for k in ("value", "ms_per_step", "kernel_ms", "roofline", "valu")}
    adj32.update(max_rel_err_flux_vs_fp64=fe, max_err_jacobian_vs_fp64_rel_to_max=ge,
                 speedup_vs_fp64=adj["kernel_ms"] / a32["kernel_ms"])
    log(f"adjoint fp32 stepping: {a32['value']:.4g} rays*voxels/s, kernel {a32['kernel_ms']:.3f} ms, "
        f"flux err vs fp64 {fe:.2g}")
    if per_ray_adj:
        log(f"adjoint per-ray loop: {per_ray_adj['device_ms']:.1f} ms device for "
            f"{per_ray_adj['launches']} launches, batched is "
            f"{per_ray_adj['speedup_of_batched']:.0f}x faster, bitwise equal "
            f"{per_ray_adj['bitwise_equal_to_batched']}")
    adj.pop("_device")
    problem = adj.pop("_problem")
    wm = run_wmodel(dp, L, physics, ctx, torch, a.steps, a.warmup)
    log(f"wmodel: {wm['value']:.4g} rays*voxels/s, {wm['ms_per_step']:.3f} ms/step")
    wprob = wm.pop("_problem")
    pm = run_pathmc(dp, physics, ctx)
    log(f"pathmc: {pm['value']:.4g} events/s, kernel {pm['kernel_ms']:.3f} ms")
    ctx.set_autotune(False)
    per_ray_pm = None if a.no_per_ray else run_pathmc_per_ray(dp, physics, ctx, pm)
    ctx.set_autotune(tune)
    if per_ray_pm:
        log(f"pathmc per-ray loop: {per_ray_pm['device_ms_all_rays']:.1f} ms device (scaled), "
            f"batched is {per_ray_pm['speedup_of_batched']:.1f}x faster")
    pm["per_ray_loop"] = per_ray_pm
    pprob = pm.pop("_problem")
    dr = run_direct(dp, L, physics, ctx, torch, a.steps, a.warmup)
    log(f"direct: {dr['value']:.4g} samples/s, kernel {dr['kernel_ms']:.3f} ms, "
        f"{dr['roofline']['achieved']:.1f} GB/s")
    dprob = dr.pop("_problem")
    wc = run_wmodel(dp, L, physics, ctx, torch, a.steps, a.warmup, corrections=True)
    wcprob = wc.pop("_problem")
    log(f"wmodel+corrections: {wc['value']:.4g} rays*voxels/s, {wc['ms_per_step']:.3f} ms/step")
    inv = run_inverse(dp, L, ctx, torch, a.steps, a.warmup)
    log(f"inverse: {inv['value']:.4g} rays*voxels/s per evaluation, graph {inv['graph_ms']:.3f} ms")
    ne = run_normal(dp, L, physics, ctx, torch, a.steps, a.warmup)
    log(f"normal: {ne['n_free']} free cells, {ne['kernel_ms']:.3f} ms, "
        f"{ne['roofline']['achieved']:.1f} GB/s")
    neprob = ne.pop("_problem")
    tr = run_trace(dp, L, physics, ctx, torch, a.steps, a.warmup)
    log(f"trace: {tr['value']:.4g} rays*voxels/s, fill kernel {tr['kernel_ms']:.4f} ms, "
        f"{tr['roofline']['achieved']:.1f} GB/s")
    tprob = tr.pop("_problem")
    dm = run_dem(dp, L, ctx, torch, a.steps, a.warmup)
    log(f"dem: {dm['value']:.4g} rays/s, kernel {dm['kernel_ms']:.4f} ms, found {dm['found']:.3f}")
    demprob = dm.pop("_problem")
    tt = run_tetra(dp, L, ctx, a.steps, a.warmup)
    log(f"tetra: {tt['value']:.4g} rays*tets/s, kernel {tt['kernel_ms']:.3f} ms")
    ttprob = tt.pop("_problem")
    cpu = None if a.no_cpu else cpu_mc_baseline(physics)
    cpu_adj = None if a.no_cpu else cpu_adjoint_baseline(physics, problem)
    wm["cpu_baseline"] = None if a.no_cpu else cpu_wmodel_baseline(physics, wprob)
    pm["cpu_baseline"] = None if a.no_cpu else cpu_pathmc_baseline(physics, pprob)
    dr["cpu_baseline"] = None if a.no_cpu else cpu_direct_baseline(physics, dprob)
    wc["cpu_baseline"] = None if a.no_cpu else cpu_wmodel_baseline(physics, wcprob,
                                                                   corr=_wcorr_oracle())
    tr["cpu_baseline"] = None if a.no_cpu else cpu_trace_baseline(tprob)
    ne["cpu_baseline"] = None if a.no_cpu else cpu_normal_baseline(neprob)
    dm["cpu_baseline"] = None if a.no_cpu else cpu_dem_baseline(demprob)
    tt["cpu_baseline"] = None if a.no_cpu else cpu_tetra_baseline(ttprob)
    out = {
        "metric": "straggled backward-MC events/s (+ adjoint rays*voxels/s), 1 GPU, % HBM roofline",
        "value": mc["value"], "unit": "events/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": mc["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "straggled+mixed backward MC, 300 m rock/water/air mixture "
                               "(0.85/0.14/0.01) at 2400 kg/m3, elevation 30 deg, E 1-1e4 GeV, "
                               "scattering + DEL/EHS on",
                   "events": mc["events"]},
        "roofline": mc["roofline"],
        "valu": mc["valu"],
        "cpu_baseline": cpu,
        "kernel_ms": mc["kernel_ms"],
        "mc": {k: mc[k] for k in ("events", "flux", "sigma", "kernel_ms", "ms_per_step")},
        "adjoint_rays_voxels_per_s": adj["value"],
        "adjoint": {"workload": "csda forward+adjoint, 128x128x64 voxels of 10 m, rho U[2000,3000], "
                                "16384 rays (8 det x 2048 dirs), 256 nodes 1-1e4 GeV",
                    "rays": adj["rays"], "segments": adj["segments"], "nodes": adj["nodes"],
                    "ms_per_step": adj["ms_per_step"], "kernel_ms": adj["kernel_ms"],
                    "roofline": adj["roofline"], "valu": adj["valu"], "cpu_baseline": cpu_adj,
                    "per_ray_loop": per_ray_adj,
                    "fp32_stepping": adj32},
        "wmodel": dict(wm, workload="water-fraction CSDA flux + sparse Jacobian, 64x64x32 "
                                    "voxels of 10 m, w U[0, 0.3], 4 detectors x 1024 "
                                    "directions, 32 nodes 1-1e4 GeV"),
        "pathmc": dict(pm, workload="fixed-path straggled MC, 64x64x32 lognormal density "
                                    "field with a 1800 kg/m3 cavity, 4096 bins x (1e4 events + the "
                                    "CSDA flux over 256 nodes) in one persistent launch"),
        "direct": dict(dr, workload="direct CSDA flux + d/d rho (compute_flux_gradient_csda), "
                                    "2^24 samples: rock 10-800 m at 2000-3000 kg/m3, "
                                    "elevation 10-90 deg, E 1-1e4 GeV"),
        "wmodel_corrected": dict(wc, workload="the wmodel leg with CsdaCorrection (kappa 1.1/0.6, "
                                              "residual 0.05/-0.02/0.003/0.1) and "
                                              "CsdaMaterialCorrection (-0.8, 1.7) on"),
        "trace": tr,
        "dem": dm,
        "normal": ne,
        "tetra": tt,
        "inverse": inv,
        "autotune": {k: {"best": list(v["best"]), "ms": v["ms"],
                         "tried": [[list(c), m] for c, m in v["tried"]]}
                     for k, v in ctx.tuning().items()},
    }
    if a.calib_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["calib_copy_bytes"] = 2 * src.numel() * 8
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
