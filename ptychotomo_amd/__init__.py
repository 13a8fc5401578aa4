"""MI355X-native ptycho-tomography ADMM (drop-in for nikitinvv/ptychotomo's
``Solver``, ``radonusfft``, ``ptychofft``, ``probe`` and ``scanner3``)."""
from .objects import (probe, scanner3, gaussian_probe, jittered_raster, ellipsoid_phantom,
                      make_theta, wavenumber)

__all__ = ["probe", "scanner3", "gaussian_probe", "jittered_raster", "ellipsoid_phantom",
           "make_theta", "wavenumber", "radonusfft", "ptychofft", "Solver"]


def __getattr__(name):
    # the HIP-backed classes load the shared library on first use
    if name == "radonusfft":
        from .radonusfft import radonusfft
        return radonusfft
    if name == "ptychofft":
        from .ptychofft import ptychofft
        return ptychofft
    if name == "Solver":
        from .solver import Solver
        return Solver
    raise AttributeError(name)
