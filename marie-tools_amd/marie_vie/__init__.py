"""MI355X-native VIE operator for MARIE (host side of libvie_hip.so).

The FFT N/K operators, their adjoints, the full VIE operator and a device
GMRES run in hand-written HIP kernels for gfx950; this package holds the
ctypes binding and the host-side mirror of MARIE's operator handles.
"""
from ._lib import VieError, load as load_library  # noqa: F401
from .operator import VieOperator  # noqa: F401
