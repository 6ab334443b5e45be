"""GPU parity of the filter-convolved peel images (peeloff_3d.f90:179-211):
the HIP pass's image_b tallies against the FP64 oracle's on the same packets,
to 1e-9 relative (the GPU sums them with atomics in another order), and the
image files written from them."""
import os

import numpy as np
import pytest

from hochunk3d_amd import build_model, config1
from hochunk3d_amd.filters import synthetic_filters
from hochunk3d_amd.testing import compare_traces, rel_close, synthetic_temperature

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def model():
    m = build_model(config1(nximg=41, nfreq=100))
    m.tdust[...] = synthetic_temperature(m)
    return m


@pytest.fixture(scope="module")
def filters():
    # 0.55 um (scattered starlight) to 70 um (thermal dust emission)
    return synthetic_filters((0.55, 1.25, 2.2, 10.0, 70.0))


def test_filter_images_match_oracle(model, filters):
    from hochunk3d_amd._lib import HcError, Transport
    from oracle.pyoracle import Oracle
    g = Transport(model, device=0)
    try:
        g.set_filters(filters)
        c = Oracle(model)
        c.set_filters(filters)
        n = 3000
        rg = g.run(n, ipeel=1, trace=True, seg0=4242)
        rc = c.run(n, ipeel=1, trace=True, seg0=4242)
        compare_traces(rg.trace, rc["trace"])
        assert rg.filt is not None
        for k, a in zip(("image_b_i", "image_b_q", "image_b_u", "image_b_v"), rg.filt):
            assert a.shape == rc[k].shape
            rel_close(a, rc[k], rtol=1e-9, atol=1e-14)
        nz = (rc["image_b_i"] > 0).sum(axis=(0, 1, 2))
        print("nonzero pixels per filter", nz.tolist())
        assert (nz > 0).sum() >= 3
        # a second peeled pass starts from zero
        rg2 = g.run(500, ipeel=1, seg0=99)
        rc2 = c.run(500, ipeel=1, seg0=99)
        rel_close(rg2.filt[0], rc2["image_b_i"], rtol=1e-9, atol=1e-14)
        # a Lucy pass leaves the last peeled pass's images
        g.run(500, ipeel=0, seg0=7)
        rel_close(g.filter_images()[0], rc2["image_b_i"], rtol=1e-9, atol=1e-14)
        g.set_filters(None)
        with pytest.raises(HcError):
            g.filter_images()
        assert g.run(200, ipeel=1).filt is None
    finally:
        g.close()


def test_filter_image_files_from_a_gpu_pass(model, filters, tmp_path):
    from hochunk3d_amd._lib import Transport
    from hochunk3d_amd.output import write_run_outputs
    g = Transport(model, device=0)
    try:
        g.set_filters(filters)
        n = 2000
        r = g.run(n, ipeel=1, seg0=1)
        paths = write_run_outputs(str(tmp_path), model, r, n, model.tdust, model.tdust,
                                  filter_names=filters.names)
    finally:
        g.close()
    img = [p for p in paths if os.path.basename(p).startswith("e_")]
    assert len(img) == model.cfg.npeel * filters.nfilt * 4
    nx = model.cfg.nximg
    rows = open(img[0]).read().split("\n")[:-1]
    assert len(rows) == nx and all(len(x) == 12 * nx for x in rows)
    vals = np.array([[float(x[12 * i:12 * i + 11]) for i in range(nx)] for x in rows])
    assert np.isfinite(vals).all() and (vals >= 0).all()
