"""calibrate_csda_correction (src/Tomography.jl:1043-1179) on the GPU forward
model: (1) closed-form recovery of kappa_hard = 22 from fluxes generated with
it; (2) the same fit on the GPU and on the oracle's fluxes gives the same
coefficients; (3) the production loop: fixed-path MC reference fluxes from
dp_mc_fixed_paths, then the fit brings the corrected CSDA flux closer to them.
"""
import math

import numpy as np
import pytest

import _oracle as O
import diffpumas_hip as dp
from diffpumas_hip.tomography import MaterialConfig, SiteConfig
from test_gpu_pathmc import _traced

pytestmark = pytest.mark.gpu


def _say(m):
    print(m, flush=True)


@pytest.fixture(autouse=True)
def _reset():
    yield
    dp.set_csda_correction(enabled=False)
    dp.set_csda_material_correction(enabled=False)


def _oracle_flux(ph, m, site, w, sh, s):
    def f(sub):
        c = dp.get_csda_correction()
        corr = O.corr_array(enabled=c.enabled, kappa_strag=c.kappa_strag,
                            kappa_hard=c.kappa_hard, resid_a=c.resid_a, resid_b=c.resid_b,
                            resid_c=c.resid_c, resid_d=c.resid_d)
        return O.w_flux(ph, O.matcfg_arrays(0, 1, 2, -1, m.rock_density, m.water_density),
                        (site.detector_elevation, site.primary_altitude),
                        dict(row_ptr=sub.row_ptr, cells=sub.cells, lengths=sub.lengths,
                             remaining=sub.remaining_rock, exit_z=sub.exit_z,
                             cos_theta=sub.cos_theta, valid=sub.valid),
                        w, sh, s.energy, s.weight, s.charge, corr=corr)
    return f


def test_calibration_on_gpu(ctx):
    ph = dp.create_physics(("StandardRock", "Water", "Air"))
    paths, n_cells = _traced(ctx, 24, 7)
    m = MaterialConfig(0, 1, 2, -1, 2650.0, 1000.0)
    rng = np.random.default_rng(12)
    w = np.round(rng.uniform(0.0, 0.3, n_cells), 1)
    sh = np.zeros(n_cells, np.uint8)
    site = SiteConfig(600.0, 3e4)
    q = dp.csda_energy_quadrature(32, 1.0, 1e4)
    bins = np.arange(paths.n_rays)
    _say(f"calibration: {paths.n_rays} bins, {paths.n_segments} segments")
    # (1) closed form
    dp.set_csda_correction(enabled=True, kappa_strag=1.0, kappa_hard=22.0)
    mc = dp.compute_directional_flux_csda(ph, sh, m, site, paths, w, q, ctx=ctx)
    dp.set_csda_correction(enabled=False)
    ks, kh, st = dp.calibrate_csda_correction(ph, sh, m, site, paths, w, q, mc, bins, ctx=ctx)
    assert kh == 22.0 and st["mse_log"] == 0.0 and st["rel_after"] < 1e-14
    # (2) production loop: MC reference fluxes, then the fit on GPU and on the oracle
    media, cm, cd = dp.build_cell_properties_for_mc(np.zeros(n_cells, bool), w, m)
    s = dp.sample_energy_set(4000, 1.0, 1e4, 5)
    dp.set_csda_correction(enabled=False)
    mcf, sg = dp.compute_directional_flux_mc(ph, m, site, paths, media, cm, cd, s, 3, ctx=ctx)
    _say(f"calibration: MC reference rel. sigma {np.nanmedian(sg / mcf):.3g}")
    ks, kh, st = dp.calibrate_csda_correction(ph, sh, m, site, paths, w, q, mcf, bins, ctx=ctx)
    _say(f"calibration: kappa_hard {kh}, rel {st['rel_before']:.4f} -> {st['rel_after']:.4f}")
    assert st["rel_after"] <= st["rel_before"]
    dp.set_csda_correction(enabled=False)
    ko, kho, so = dp.calibrate_csda_correction(ph, sh, m, site, paths, w, q, mcf, bins,
                                               _flux=_oracle_flux(ph, m, site, w, sh, q))
    assert kho == kh
    for k in ("resid_a", "resid_b", "resid_c", "resid_d", "rel_before", "rel_after", "mse_log"):
        assert so[k] == pytest.approx(st[k], rel=1e-8, abs=1e-12), k
