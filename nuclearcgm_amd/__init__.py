"""nuclearcgm_amd: MI355X-native self-consistent Skyrme HF(+BCS/HFB) on a 3D
lattice.  The compute path is libnhf_hip.so (hand-written gfx950 HIP kernels
behind the C ABI of include/nhf.h); this package holds the host-side mirror of
the reference's input/driver interface and the ctypes binding."""
from . import config  # noqa: F401
