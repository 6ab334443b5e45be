"""Grid densities on the GPU (hc_build_density) against the host construction
of the same formulas: flared disks with their mass normalisation, the Ulrich
envelope (the streamline cubic), the polynomial cavity, the clump multiplier,
cell volumes and (rho V)^(1/4).  The two differ only in the last bits of the
elementary functions and of the disk-mass sums.  The clump point counts
(hc_fractal_counts) equal the host's exactly."""
import time

import numpy as np
import pytest

from hochunk3d_amd import build_model, config1, config2
from hochunk3d_amd.batch import parameter_grid

pytestmark = pytest.mark.gpu

CASES = {
    "64x64x32 disk+envelope": config1(),
    "40x41x24 fractal, even theta walls": config2(nr=40, ntheta=41, nphi=24),
    "grid model 196x176x1": parameter_grid(12, max_cells=200_000)[10].cfg,
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_densities_match_host(name):
    cfg = CASES[name]
    h = build_model(cfg)
    d = build_model(cfg, device=0)
    for k in ("densarr", "massarr", "vcell"):
        a, b = getattr(d, k), getattr(h, k)
        assert a.shape == b.shape
        np.testing.assert_array_equal(a == 0.0, b == 0.0, err_msg=k)
        np.testing.assert_allclose(a, b, rtol=1e-11, atol=0.0, err_msg=k)
    assert d.fmass8 == h.fmass8 and d.z1_8 == h.z1_8 and d.bflare8 == h.bflare8
    # the disks carry their mass
    for id_, fm in enumerate(h.fmass8):
        if fm > 0:
            m = (d.densarr[..., id_] * d.vcell).sum()
            assert abs(m / (fm * cfg.massd * 1.989e33) - 1.0) < 1e-12


def test_device_build_time_of_the_8m_cell_model():
    cfg = config2()
    build_model(config1(), device=0)          # first use of the device
    t0 = time.perf_counter()
    d = build_model(cfg, device=0)
    td = time.perf_counter() - t0
    t0 = time.perf_counter()
    h = build_model(cfg)
    th = time.perf_counter() - t0
    print(f"200x199x200 model: host {th:.2f} s, with device densities {td:.2f} s")
    np.testing.assert_allclose(d.densarr, h.densarr, rtol=1e-11, atol=0.0)
    assert td < th


@pytest.mark.parametrize("kappa_v,albedo_max", [(200.0, 0.6), (35.0, 0.9)])
def test_device_emissivity_tables_match_host(kappa_v, albedo_max):
    """setemiss on the GPU (hc_build_emissivity) against the host quadrature:
    the same 8-point rule per bin, the same cumulative sums."""
    from hochunk3d_amd.dust import dustinit, synthetic_dust_table
    tab = synthetic_dust_table(kappa_v=kappa_v, albedo_max=albedo_max)
    t0 = time.perf_counter()
    h = dustinit(tab)
    th = time.perf_counter() - t0
    dustinit(tab, device=0)
    t0 = time.perf_counter()
    d = dustinit(tab, device=0)
    td = time.perf_counter() - t0
    print(f"emissivity tables: host {th:.3f} s, device {td:.3f} s")
    np.testing.assert_allclose(d.kappap, h.kappap, rtol=1e-11, atol=0.0)
    np.testing.assert_allclose(d.kapBint, h.kapBint, rtol=1e-10, atol=1e-300)
    # a normalised cumulative table: non-decreasing, 1 at the last node
    kb = d.kapBint[:, :, 0]
    assert np.all(np.diff(kb, axis=1) >= 0.0) and np.all(kb[:, -1] == 1.0)


@pytest.mark.parametrize("shape,n,levels", [((40, 41, 24), 8, 3), ((40, 41, 24), 32, 4),
                                            ((30, 31, 1), 16, 3)])
def test_device_clump_counts_equal_host(shape, n, levels):
    """hc_fractal_counts jumps NumPy's PCG64 stream to each point's draws:
    the counts per cell equal the host construction's exactly."""
    from hochunk3d_amd.fractal import fractal_counts
    nr, nt, nph = shape
    m = build_model(config2(nr=nr, ntheta=nt, nphi=nph, fractal=None))
    kw = dict(n=n, levels=levels, fl=3.792, seed=11)
    h = fractal_counts(m.rarr, m.thetarr, m.phiarr, **kw)
    t0 = time.perf_counter()
    d = fractal_counts(m.rarr, m.thetarr, m.phiarr, **kw, device=0)
    print(f"{n}^{levels + 1} points: device {time.perf_counter() - t0:.3f} s")
    np.testing.assert_array_equal(d, h)
    assert 0 < h.sum() <= n ** (levels + 1)      # points beyond rmax are dropped
