"""MI355X build of the HDSDP dual-scaling interior-point hot path.

The solver runs in ``hdsdp_amd/lib/libhdsdp_amd.so`` (C host + HIP kernels for
gfx950).  Python is a thin driver: ``instances`` builds seeded problems,
``solver`` calls the C API, ``adjoint`` exposes gradients to PyTorch.
"""
from . import instances  # noqa: F401

__all__ = ["instances"]
