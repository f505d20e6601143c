"""Python access to the request loop of the library (include/rsh.h, rsh_server_*).

Requests are ray queries (as Engine.trace) against scenes resident on the
device; the library admits them under an HBM budget and merges concurrent
requests of one scene into shared batches.  Host arrays handed to submit()
are kept referenced here until wait() returns.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import RshError, config_desc, lib, scene_desc

u32, u64, PV = C.c_uint32, C.c_uint64, C.c_void_p


class RequestInfo(C.Structure):
    _fields_ = [("queue_ms", C.c_double), ("latency_ms", C.c_double), ("batch_requests", u32)]


class ServerInfo(C.Structure):
    _fields_ = [("batches", u64), ("requests", u64), ("rays", u64), ("busy_ms", C.c_double),
                ("pending", u32), ("admitted_bytes", u64)]


_declared = False


def _declare(L):
    global _declared
    if _declared:
        return
    from . import FrameConfigDesc, SceneDesc
    L.rsh_server_create.restype = C.c_int
    L.rsh_server_create.argtypes = [C.POINTER(PV), C.c_int, u64, u32]
    L.rsh_server_add_scene.restype = C.c_int
    L.rsh_server_add_scene.argtypes = [PV, C.POINTER(SceneDesc), C.POINTER(FrameConfigDesc), C.POINTER(u32)]
    L.rsh_server_submit.restype = C.c_int
    L.rsh_server_submit.argtypes = [PV, u32, u32, PV, PV, u32, PV, C.POINTER(u64)]
    L.rsh_server_wait.restype = C.c_int
    L.rsh_server_wait.argtypes = [PV, u64, C.POINTER(RequestInfo)]
    L.rsh_server_stats.restype = C.c_int
    L.rsh_server_stats.argtypes = [PV, C.POINTER(ServerInfo)]
    L.rsh_server_last_error.restype = C.c_char_p
    L.rsh_server_last_error.argtypes = [PV]
    L.rsh_server_destroy.argtypes = [PV]
    L.rsh_server_set_slice.restype = C.c_int
    L.rsh_server_set_slice.argtypes = [PV, u32]
    L.rsh_host_alloc.restype = PV
    L.rsh_host_alloc.argtypes = [C.c_size_t]
    L.rsh_host_free.argtypes = [PV]
    _declared = True


def pinned(shape, dtype):
    """A numpy array in page-locked host memory (freed with the array's owner)."""
    L = lib()
    _declare(L)
    dtype = np.dtype(dtype)
    nbytes = int(np.prod(shape)) * dtype.itemsize
    p = L.rsh_host_alloc(max(nbytes, 1))
    if not p:
        raise MemoryError(f"pinned allocation of {nbytes} bytes failed")
    buf = (C.c_char * max(nbytes, 1)).from_address(p)
    arr = np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)
    _Pinned.keep[id(arr)] = p
    return arr


class _Pinned:
    keep = {}


class Server:
    def __init__(self, device=0, hbm_budget=64 << 30, max_batch_rays=32 << 20):
        self.L = lib()
        _declare(self.L)
        self.h = PV()
        rc = self.L.rsh_server_create(C.byref(self.h), device, hbm_budget, max_batch_rays)
        if rc != 0:
            raise RshError(f"server_create failed ({rc})")
        self._scenes = []
        self._live = {}

    def _check(self, rc, what):
        if rc != 0:
            raise RshError(f"{what} failed ({rc}): {self.L.rsh_server_last_error(self.h).decode()}")

    def set_slice(self, slice_rays):
        """Largest slice of one request per batch (rsh_server_set_slice)."""
        self._check(self.L.rsh_server_set_slice(self.h, int(slice_rays)), "set_slice")

    def add_scene(self, scene, cfg) -> int:
        sid = u32()
        sd = scene_desc(scene)
        self._check(self.L.rsh_server_add_scene(self.h, C.byref(sd), C.byref(config_desc(cfg)), C.byref(sid)),
                    "add_scene")
        self._scenes.append((scene, sd))
        return sid.value

    def submit(self, scene_id, rays, kind, tmax=None, out=None) -> int:
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 6)
        n = rays.shape[0]
        if tmax is not None:
            tmax = np.ascontiguousarray(tmax, np.float32)
        if out is None:
            out = np.empty(n, np.uint32 if kind == 1 else np.uint64)
        rid = u64()
        self._check(self.L.rsh_server_submit(self.h, scene_id, kind, rays.ctypes.data,
                                             tmax.ctypes.data if tmax is not None else None, n, out.ctypes.data,
                                             C.byref(rid)), "submit")
        self._live[rid.value] = (rays, tmax, out)
        return rid.value

    def wait(self, rid):
        """(result array, RequestInfo)."""
        info = RequestInfo()
        rc = self.L.rsh_server_wait(self.h, rid, C.byref(info))
        arrays = self._live.pop(rid, None)
        self._check(rc, "wait")
        return arrays[2], info

    def stats(self) -> ServerInfo:
        info = ServerInfo()
        self._check(self.L.rsh_server_stats(self.h, C.byref(info)), "stats")
        return info

    def close(self):
        if self.h:
            self.L.rsh_server_destroy(self.h)
            self.h = PV()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
