"""MI355X (gfx950) Godunov hydro engine behind PeleC's construct_hydro_source
-> pc_umdrv -> pc_umeth_3D path.  The compute lives in libpc_hydro.so (HIP,
C ABI in include/pc_hydro.h); this package is the host-side mirror of the
reference interface plus synthetic workloads."""
from . import workloads  # noqa: F401
from .ctypes_types import make_params  # noqa: F401

EOS_GAMMALAW = 1
EOS_LIDRYER = 9

BC_INTERIOR, BC_INFLOW, BC_OUTFLOW, BC_SYMMETRY, BC_SLIPWALL, BC_NOSLIPWALL, \
    BC_USERBC = range(7)




def __getattr__(name):
    # the device classes load libpc_hydro.so on first use
    if name in ("HydroContext", "DeviceLevel"):
        from . import hydro
        return getattr(hydro, name)
    raise AttributeError(name)
