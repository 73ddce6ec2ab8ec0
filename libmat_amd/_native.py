"""Loader for the in-tree native library ``libmat_amd/_lib/librpd_amd.so``.

The library holds the HIP kernels, the C-ABI of include/rpd_capi.h and the
host-side input preparation of include/rpd_prep.h.  Loading fails loudly when
the library has not been built: there is no Python fallback for any of it.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "librpd_amd.so")

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def lib():
    """Return the loaded ctypes library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryMissing(
                f"{LIB_PATH} is not built; run `python -c 'import "
                f"__graft_entry__ as g; g.build()'` or `python -m "
                f"libmat_amd.build`")
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    c = ctypes
    i32, i64, u64, f64 = c.c_int32, c.c_int64, c.c_uint64, c.c_double
    vp = c.c_void_p
    L.rpd_rng_uniform.argtypes = [u64, u64, u64]
    L.rpd_rng_uniform.restype = f64
    L.rpd_synth_grid_mesh.argtypes = [i32, f64, u64, vp, vp]
    L.rpd_synth_grid_mesh.restype = c.c_int
    L.rpd_synth_spheres.argtypes = [i32, i32, f64, f64, u64, vp]
    L.rpd_synth_spheres.restype = c.c_int
    L.rpd_tet_adjacency.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp]
    L.rpd_tet_adjacency.restype = c.c_int
    L.rpd_power_neighbors.argtypes = [i32, vp, vp, i32, c.POINTER(vp)]
    L.rpd_power_neighbors.restype = c.c_int
    L.rpd_nbr_total.argtypes = [vp]
    L.rpd_nbr_total.restype = i64
    L.rpd_nbr_max_degree.argtypes = [vp]
    L.rpd_nbr_max_degree.restype = i32
    L.rpd_nbr_get.argtypes = [vp, vp, vp, vp]
    L.rpd_nbr_get.restype = c.c_int
    L.rpd_nbr_free.argtypes = [vp]
    L.rpd_nbr_free.restype = None
    if hasattr(L, "rpd_ctx_create"):
        from . import rpd as _rpd
        _rpd._declare_capi(L)
