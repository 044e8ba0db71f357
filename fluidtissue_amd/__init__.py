"""fluidtissue_amd: MI355X-native level-set reinitialization for fluidtissue.

The compute path is libftreinit.so (hand-written gfx950 HIP kernels behind
the C ABI in include/ftreinit.h).  This package holds its build script and a
thin host-side mirror of the reference's reinit interface.
"""
from .reinit import (EuclideanDistance, Interface, Mesh, Reinit, device_count, kuhn_mesh,
                     make_opts, phi_field)

__all__ = ["EuclideanDistance", "Interface", "Mesh", "Reinit", "device_count", "kuhn_mesh",
           "make_opts", "phi_field"]
