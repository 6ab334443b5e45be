"""The parameter-grid generator: ranges, determinism and the cell cap."""
import numpy as np
import pytest

from hochunk3d_amd.batch import parameter_grid


def test_grid_ranges_and_determinism():
    g = parameter_grid(256)
    assert len(g) == 256
    nr = np.array([m.cfg.nr for m in g])
    nth = np.array([m.cfg.ntheta for m in g])
    nph = np.array([m.cfg.nphi for m in g])
    md = np.array([m.cfg.massd for m in g])
    rate = np.array([m.cfg.rate for m in g])
    cav = np.array([m.cfg.cavity_deg for m in g])
    ts = np.array([m.cfg.tstar for m in g])
    npk = np.array([m.packets for m in g])
    assert nr.min() >= 50 and nr.max() <= 250
    assert nth.min() >= 60 and nth.max() <= 200
    assert nph.min() >= 1 and nph.max() <= 128
    assert md.min() >= 1e-4 and md.max() <= 0.1
    assert rate.min() >= 1e-7 and rate.max() <= 1e-4
    assert cav.min() >= 5 and cav.max() <= 40
    assert ts.min() >= 3000 and ts.max() <= 10000
    assert npk.min() >= 1_000_000 and npk.max() <= 20_000_000
    # log-uniform packet counts: about half below the geometric midpoint
    assert 0.35 < np.mean(npk < np.sqrt(1e6 * 2e7)) < 0.65
    g2 = parameter_grid(256)
    assert [m.cfg for m in g2] == [m.cfg for m in g]


def test_cell_cap_and_packet_scale():
    g = parameter_grid(20, packet_scale=1e-3, max_cells=100_000)
    assert all(m.cfg.nr * m.cfg.ntheta * m.cfg.nphi <= 100_000 for m in g)
    assert all(1_000 <= m.packets <= 20_000 for m in g)


def test_temperature_comparison_flags_real_differences():
    from hochunk3d_amd.testing import assert_temperatures_close
    rng = np.random.default_rng(1)
    t = rng.uniform(0.01, 0.1, size=(20, 20, 10, 4))
    assert_temperatures_close(t * (1 + 1e-6), t, True)
    few = t.copy()
    few.flat[:5] *= 1.003                       # a few cells of diverged histories
    assert_temperatures_close(few, t, False)
    with pytest.raises(AssertionError):
        assert_temperatures_close(few, t, True)
    noisy = t * (1 + 0.01 * rng.standard_normal(t.shape))   # other packets: MC noise
    with pytest.raises(AssertionError):
        assert_temperatures_close(noisy, t, False)
