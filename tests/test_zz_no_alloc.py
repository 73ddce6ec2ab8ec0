"""GPU: the allocation-free entry point rpd_run_noalloc against rpd_run on
the same instances, through the C-ABI: two calls in a row (the second on a
smaller instance), each bit-identical to rpd_run's cell stream; the
context's resident instance and last results are left as they were;
output buffers that are too small and an instance over the scratch are
refused with their error codes.  Output buffers are device memory the
caller (this test) allocates itself through the HIP runtime."""
import ctypes

import numpy as np
import pytest

from libmat_amd import prep
from libmat_amd.rpd import CELL_DTYPE, RPDResult

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """The caller's own device memory (hipMalloc of the HIP runtime the
    library uses), zeroed, with a copy back to the host."""
    _hip = None

    def __init__(self, nbytes):
        if DeviceBuffer._hip is None:
            h = ctypes.CDLL("libamdhip64.so")
            h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.c_size_t]
            h.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int,
                                    ctypes.c_size_t]
            h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.c_int]
            h.hipFree.argtypes = [ctypes.c_void_p]
            DeviceBuffer._hip = h
        self.n = max(int(nbytes), 8)
        self.p = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.p), self.n) == 0
        assert self._hip.hipMemset(self.p, 0, self.n) == 0

    def data_ptr(self):
        return self.p.value

    def host(self, dtype):
        out = np.empty(self.n // np.dtype(dtype).itemsize, dtype)
        assert self._hip.hipMemcpy(out.ctypes.data, self.p, out.nbytes,
                                   2) == 0  # device to host
        return out

    def __del__(self):
        if self._hip is not None and self.p:
            self._hip.hipFree(self.p)


def _same(a, b):
    for k in ("tet", "seed", "nb_v", "nb_p", "nb_e", "status", "word_off",
              "area_off"):
        assert np.array_equal(a.cells[k], b.cells[k]), k
    for f in ("words", "geom", "areas"):
        assert np.array_equal(getattr(a, f), getattr(b, f)), f


def _buffers(n_cells, n_words, n_areas):
    return (DeviceBuffer(max(n_cells, 1) * 32),
            DeviceBuffer(max(n_words, 1) * 4),
            DeviceBuffer(max(n_cells, 1) * 32),
            DeviceBuffer(max(n_areas, 1) * 8))


def _out(bufs, n_cells, n_words, n_areas):
    c, w, g, a = bufs
    return (c.data_ptr(), n_cells, w.data_ptr(), n_words, g.data_ptr(),
            a.data_ptr(), n_areas)


def test_noalloc_matches_rpd_run(ctx):
    verts, tets = prep.grid_mesh(6, 0.1, 4)
    small = prep.build_instance(verts, tets, prep.spheres(200, 2, 4))
    keep = prep.cube10_instance(3)
    ctx.upload(keep)
    ctx.run(cells_per_wave=8)
    kept = ctx.download()
    for inst in (prep.cube10_instance(0), small):
        # the reference result from rpd_run, on another context
        from libmat_amd.rpd import Context
        with Context(0) as ref_ctx:
            ref_ctx.upload(inst)
            st = ref_ctx.run(cells_per_wave=8)
            ref = ref_ctx.download()
        n_c, n_w, n_a = st["n_cells"], st["n_words"], st["n_areas"]
        bufs = _buffers(n_c + 7, n_w + 11, n_a + 5)
        rc, st2, err = ctx.run_noalloc(
            inst, _out(bufs, n_c + 7, n_w + 11, n_a + 5))
        assert rc == 0, err
        assert (st2["n_cells"], st2["n_words"], st2["n_areas"]) == \
            (n_c, n_w, n_a)
        c = bufs[0].host(np.uint8).view(CELL_DTYPE)
        w = bufs[1].host(np.uint32)
        g = bufs[2].host(np.float64)
        a = bufs[3].host(np.float64)
        res = RPDResult(c[:n_c].copy(), w[:n_w].copy(), g[:4 * n_c].copy(),
                        a[:n_a].copy())
        _same(res, ref)
        print(f"[noalloc] {inst.n_tet} tets: {n_c} cells identical to "
              f"rpd_run, scratch used {st2['scratch_bytes'] / 2**20:.1f} MiB",
              flush=True)
        # the context's own instance and results are untouched
        _same(ctx.download(), kept)
    # too small an output buffer: refused with -5
    bufs = _buffers(4, 4, 4)
    rc, _, err = ctx.run_noalloc(small, _out(bufs, 4, 4, 4))
    print(f"[noalloc] small buffers: rc {rc}: {err}", flush=True)
    assert rc == -5 and "too small" in err
    _same(ctx.download(), kept)


def test_noalloc_refuses_an_instance_over_the_scratch(ctx):
    big = prep.cube64_instance(0, n_threads=16, n=48, n_spheres=20000)
    bufs = _buffers(16, 16, 16)
    rc, _, err = ctx.run_noalloc(big, _out(bufs, 16, 16, 16))
    print(f"[noalloc] {big.n_tet} tets: rc {rc}: {err}", flush=True)
    assert rc == -4 and "scratch" in err
