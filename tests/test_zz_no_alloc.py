"""hc_build_density_noalloc (the entry point that allocates nothing) against
hc_build_density on the same model: densities, (rho V)^(1/4) and cell
volumes from device arrays the caller owns, two calls in a row, the second
on a smaller grid, so the library's device scratch array (the disk-mass
sums) is reused and must be cleared by the call itself.

The two entry points share their kernels; the disk-mass sums are FP64
atomics whose order varies from launch to launch, so the normalised disk
densities agree to rounding (rtol 1e-12), everything else exactly."""
import numpy as np
import pytest

from hochunk3d_amd import build_model, config1, config2
from hochunk3d_amd.model import density_fields

pytestmark = pytest.mark.gpu


def _run_noalloc(model, mult=None):
    import torch

    from hochunk3d_amd._lib import build_density_noalloc

    dev = torch.device("cuda:0")
    f = density_fields(model, mult)
    t = {k: torch.from_numpy(np.ascontiguousarray(f[k], dtype=np.float64)).to(dev)
         for k in ("rarr", "thetarr", "phiarr")}
    if mult is not None:
        # the multiplier is (nrg-1, ntg-1, npg-1) column-major
        t["fractal"] = torch.from_numpy(np.asfortranarray(mult).ravel(order="F").copy()).to(dev)
    nw = model.nrg * model.ntg * model.npg
    # outputs start as garbage: the call must write every element
    dens = torch.full((8 * nw,), float("nan"), dtype=torch.float64, device=dev)
    mass = torch.full((8 * nw,), float("nan"), dtype=torch.float64, device=dev)
    vol = torch.full((nw,), float("nan"), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream(dev)
    build_density_noalloc(f, {k: v.data_ptr() for k, v in t.items()}, dens.data_ptr(),
                          mass.data_ptr(), vol.data_ptr(), device=0,
                          stream=stream.cuda_stream)
    stream.synchronize()
    shp = (model.nrg, model.ntg, model.npg)
    return (dens.cpu().numpy().reshape(shp + (8,), order="F"),
            mass.cpu().numpy().reshape(shp + (8,), order="F"),
            vol.cpu().numpy().reshape(shp, order="F"))


def _check(model, got):
    from hochunk3d_amd._lib import build_density

    f = density_fields(model)
    ref = build_density(f, model.nrg, model.ntg, model.npg, device=0)
    for name, a, b in zip(("densarr", "massarr", "vcell"), got, ref):
        assert a.shape == b.shape, name
        assert np.all(np.isfinite(a)), name
        np.testing.assert_array_equal(a == 0.0, b == 0.0, err_msg=name)
        np.testing.assert_allclose(a, b, rtol=1e-12, atol=0.0, err_msg=name)
    # the disks carry their mass
    for id_, fm in enumerate(model.fmass8):
        if fm > 0:
            m = (got[0][..., id_] * got[2]).sum()
            assert abs(m / (fm * model.cfg.massd * 1.989e33) - 1.0) < 1e-12


def test_noalloc_density_matches_allocating_entry_point_twice():
    big = build_model(config1(nr=48, ntheta=49, nphi=16))
    small = build_model(config2(nr=20, ntheta=21, nphi=8, fractal=None))
    print("noalloc: 48x49x16 then 20x21x8", flush=True)
    _check(big, _run_noalloc(big))
    _check(small, _run_noalloc(small))


def test_noalloc_rejects_more_disks_than_its_scratch_holds():
    from hochunk3d_amd._lib import HcError, build_density_noalloc

    m = build_model(config1(nr=8, ntheta=9, nphi=4))
    f = density_fields(m)
    f["ndisk"] = 9
    with pytest.raises(HcError):
        build_density_noalloc(f, {"rarr": 1, "thetarr": 1, "phiarr": 1}, 1, 1, 1)
