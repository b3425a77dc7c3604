"""MI355X-native DGSEM right-hand side + RK4(3)5[3R+]N update of PLENS (gfx950 HIP kernels).

The native library libplens_gpu.so (built in-tree by plens_amd.build) holds the kernels, the C ABI
of include/plens_gpu.h and the host setup of include/plens_host.h.
"""
from . import _lib  # noqa: F401
from .batch import GpuBatch  # noqa: F401
from .cases import CONFIGS, Mesh, batch_instance, cylinder_flare, nose_cone, periodic_cube, shock_tube, varying_inflow_tube, wall_flow, with_fluxes, with_gradient_variables, with_precision  # noqa: F401
from .solver import GpuPlens, PositivityError  # noqa: F401
