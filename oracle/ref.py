"""ORACLE cross-check (test infrastructure only): ctypes binding of
oracle/_ref/libref_psm.so, built from the reference's own CPU sources by
oracle/ref.mk (only where the reference is mounted)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "_ref", "libref_psm.so")
_lib = None


def available():
    return os.path.exists(LIB)


def lib():
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB)
        vp = ctypes.c_void_p
        L.ref_det4_sign.argtypes = [vp]
        L.ref_det4_sign.restype = ctypes.c_int
        L.ref_rt_neighbors.argtypes = [ctypes.c_int32, vp, vp, vp,
                                       ctypes.c_int64]
        L.ref_rt_neighbors.restype = ctypes.c_int64
        L.ref_vbw_cell.argtypes = [vp, ctypes.c_int32, vp, vp]
        L.ref_vbw_cell.restype = ctypes.c_int
        L.ref_init()
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def det4_sign(m):
    m = np.ascontiguousarray(m, np.float64).reshape(16)
    return lib().ref_det4_sign(_p(m))


def rt_neighbors(xyzw):
    xyzw = np.ascontiguousarray(xyzw, np.float64)
    n = xyzw.shape[0]
    off = np.zeros(n + 1, np.int64)
    cap = 64 * n + 64
    ids = np.zeros(cap, np.int32)
    tot = lib().ref_rt_neighbors(n, _p(xyzw), _p(off), _p(ids), cap)
    assert tot >= 0
    return off, ids[:tot]


def vbw_cell(box6, planes):
    planes = np.ascontiguousarray(planes, np.float64).reshape(-1, 4)
    box6 = np.ascontiguousarray(box6, np.float64)
    out = np.zeros(4 + planes.shape[0])
    lib().ref_vbw_cell(_p(box6), planes.shape[0], _p(planes), _p(out))
    return out[0], out[1:4], out[4:]


def _vbw_worker(conn):
    while True:
        job = conn.recv()
        if job is None:
            return
        box6, planes = job
        conn.send(vbw_cell(box6, planes))


def vbw_cells_guarded(jobs, timeout=2.0):
    """VBW cells for a list of (box6, planes) jobs in a spawned worker.

    VBW decides plane sides with non-exact floating point, and on nearly
    degenerate planes its triangle lists can loop forever; a job that does
    not finish within `timeout` seconds is skipped (result None) and the
    worker is restarted."""
    import multiprocessing as mp
    ctx = mp.get_context("spawn")
    out = [None] * len(jobs)
    q = 0
    while q < len(jobs):
        a, b = ctx.Pipe()
        p = ctx.Process(target=_vbw_worker, args=(b,), daemon=True)
        p.start()
        try:
            while q < len(jobs):
                a.send(jobs[q])
                if not a.poll(timeout + (5.0 if q == 0 else 0.0)):
                    q += 1  # skip the hanging job
                    break
                out[q] = a.recv()
                q += 1
            else:
                a.send(None)
        finally:
            p.kill()
            p.join()
    return out
