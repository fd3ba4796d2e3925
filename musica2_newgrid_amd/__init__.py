"""MI355X-native explicit RK time advance of MUSICAA (LouenasZm/musica2_newGrid).

The product is libmusica_gpu.so (HIP kernels for gfx950 behind the C-ABI in
include/musica_gpu.h); this package holds its Python front end and the
host-side case set-up.
"""
from .cases import Case  # noqa: F401
