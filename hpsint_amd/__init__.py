"""hpsint_amd: MI355X-native hot path of hpsint's sintering operator.

HIP kernels for gfx950 behind a C ABI (include/hpsint_amd.h, built into
hpsint_amd/libhpsint_amd.so); this package is the host-side mirror of the
reference's operator interface used by the tests and the benchmark.
"""
from .operator import (CartesianMesh, FreeEnergyParams, MobilityParams,  # noqa: F401
                       Preconditioner, SinteringOperatorGeneric, TimeIntegratorData)
