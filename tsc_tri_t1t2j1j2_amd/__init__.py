"""MI355X-native hot path of the SU(2) two-site DMRG of
hyx15903/TSC_Tri_t1t2J1J2: superblock H.psi, Lanczos, density matrix /
eigensolve / truncation and U^H O U, as hand-written HIP kernels for gfx950
behind the C-ABI in include/tsc_hotpath.h (libtsc_hotpath.so).
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
