"""scico_amd: MI355X-native (gfx950 HIP) tomographic hot path of lanl/scico.

The host side mirrors scico's operator / functional / optimizer interface;
all numerical work runs in hand-written HIP kernels reached through the C
ABI of ``libscico_amd.so`` (``include/scico_amd.h``).
"""

__version__ = "0.1.0"
