"""insilico_amd: the finite-element hot path of thrueberg/inSilico
(element assembly -> CSR -> Jacobi-PCG) on AMD MI355X (gfx950).

The compute path is the HIP library ``libinsilico_hip.so`` behind the C ABI
declared in ``include/insilico_hip.h``; this package holds its ctypes
binding (``abi``), the host-side problem setup (``problems``), the Newton
driver (``newton``) and the build step (``build``).
"""
from . import abi, problems  # noqa: F401

__all__ = ["abi", "problems"]
