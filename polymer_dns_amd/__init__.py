"""MI355X-native Kim-Moin-Moser channel time step (drop-in for the
Polymer_DNS FFTW_DNS time-step path).  HIP kernels for gfx950 behind the C
ABI in include/kmm.h; this package holds the host-side mirror of the
reference's solver interface, restart/config/geometry I/O and the seeded
initial-condition tool."""
from .kmm import Channel, KmmConfig, KmmError, CFLError, load_library, make_config, read_config  # noqa: F401
