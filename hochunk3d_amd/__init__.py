"""MI355X-native photon-packet transport for Hochunk3d models.

The transport pass runs in hand-written HIP kernels (hochunk3d_amd/csrc)
behind the C ABI of include/hc.h; this package holds the host side: model
setup (grid, densities, dust tables), the ctypes binding and the Lucy
iteration driver.
"""
from .model import ModelConfig, Model, build_model, config1, config2, config5  # noqa: F401
from .dust import dustinit, synthetic_dust_table  # noqa: F401

__all__ = ["ModelConfig", "Model", "build_model", "config1", "config2", "config5", "dustinit",
           "synthetic_dust_table"]
