"""particlebase_amd -- MI355X-native short-range particle engine.

The drop-in boundary is the C ABI in include/pbamd.h (libpbamd.so, HIP
kernels for gfx950).  The C++ class surface of the reference
(cudaParticleMD, cudaParticleVV, cudaParticleLF, cudaCutoffBlock,
cudaSortedPositions, cudaParticleSPH_NS, cudaParticleDEM, cudaCalcGR) lives in
particlebase_amd/cpp/ over that ABI; this Python package is the ctypes
binding (the counterpart of the reference's SWIG module) plus the synthetic
workload generators used by the benchmark and tests.
"""
from ._lib import (Context, System, PBError, PB_FP64, PB_MIXED, LIB_PATH, load,  # noqa: F401
                   EXPORTS, PB_OK, PB_ERR_ARG, PB_ERR_HIP, PB_ERR_STATE, PB_ERR_CELLS, PB_ERR_NLIST,
                   PB_ERR_SKIN, PB_ERR_DT)
