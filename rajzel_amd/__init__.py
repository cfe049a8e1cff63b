"""rajzel_amd -- MI355X (gfx950) implementation of RAJZEL's adjoint-state
first-arrival traveltime tomography gradient path.

The product is the HIP library ``librajzel_amd.so`` behind the C ABI of
``include/rajzel_amd.h``; :mod:`rajzel_amd.native` binds it for Python and
:mod:`rajzel_amd.torch_fatt` exposes it as a ``torch.autograd.Function``.
"""
from . import workloads  # noqa: F401

__all__ = ["native", "workloads"]
