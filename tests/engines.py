"""Uniform handles on the three implementations of the Orkan C API the tests compare:

  HostEngine   a library whose state data lives in host memory: the CPU oracle
               (oracle/lib) or the reference library compiled from its sources
               (oracle/_ref).  Data is read and written through the data pointer.
  DeviceEngine the HIP library of this package (orkan_amd/lib): state data is in
               HBM and moves through orkan_state_upload / orkan_state_download.
"""
import ctypes as C
import os

import numpy as np

from orkan_amd import _ffi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def oracle_path(log_tile):
    return os.path.join(ROOT, "oracle", "lib", f"liborkan_oracle_t{log_tile}.so")


def ref_path(log_tile):
    return os.path.join(ROOT, "oracle", "_ref", f"liborkan_ref_t{log_tile}.so")


class Engine:
    log_tile = 5

    def __init__(self, api, log_tile):
        self.api = api
        self.log_tile = log_tile

    # every public entry point is reachable as an attribute
    def __getattr__(self, name):
        return getattr(self.api, name)

    def new(self, kind, n, data=None):
        st = _ffi.StateT(kind, None, 0)
        self.api.state_init(C.byref(st), n, None)
        assert st.data, "allocation failed"
        if data is not None:
            self.write(st, np.ascontiguousarray(data, dtype=np.complex128))
        return st

    def plus(self, kind, n):
        st = _ffi.StateT(kind, None, 0)
        self.api.state_plus(C.byref(st), n)
        return st

    def free(self, st):
        self.api.state_free(C.byref(st))

    def length(self, st):
        return int(self.api.state_len(C.byref(st)))

    def gate(self, st, name, *args):
        getattr(self.api, name)(C.byref(st), *args)

    def channel(self, st, sop, q):
        sop = np.ascontiguousarray(sop, dtype=np.complex128).reshape(16)
        so = _ffi.SuperopT(1, sop.ctypes.data)
        self.api.channel_1q(C.byref(st), C.byref(so), q)

    def channel2(self, st, sop, q1, q2):
        sop = np.ascontiguousarray(sop, dtype=np.complex128).reshape(256)
        so = _ffi.SuperopT(2, sop.ctypes.data)
        self.api.channel_2q(C.byref(st), C.byref(so), q1, q2)

    def mean(self, st, obs):
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        return float(self.api.mean(C.byref(st), _ffi.dptr(obs)))

    def matel(self, bra, ket, obs):
        obs = np.ascontiguousarray(obs, dtype=np.float64)
        return self.api.matel_diag(C.byref(bra), C.byref(ket), _ffi.dptr(obs)).value()

    def exp_diag(self, st, diag, t):
        diag = np.ascontiguousarray(diag, dtype=np.float64)
        self.api.exp_diag(C.byref(st), _ffi.dptr(diag), float(t))

    def get(self, st, r, c=0):
        return self.api.state_get(C.byref(st), r, c).value()

    def set(self, st, r, c, v):
        self.api.state_set(C.byref(st), r, c, _ffi.Cplx.of(v))


class HostEngine(Engine):
    def __init__(self, path, log_tile):
        super().__init__(_ffi.bind(path), log_tile)
        self.libc = C.CDLL(None)

    def write(self, st, arr):
        n = self.length(st)
        assert arr.size == n
        C.memmove(st.data, arr.ctypes.data, n * 16)

    def read(self, st):
        n = self.length(st)
        out = np.empty(n, dtype=np.complex128)
        C.memmove(out.ctypes.data, st.data, n * 16)
        return out


class DeviceEngine(Engine):
    def __init__(self, path, log_tile):
        import orkan_amd
        from orkan_amd import batch, surrogate
        extra = dict(orkan_amd.EXTRA)
        extra.update(batch.EXTRA)
        extra.update(surrogate.EXTRA)
        super().__init__(_ffi.bind(path, extra), log_tile)

    def write(self, st, arr):
        assert arr.size == self.length(st)
        rc = self.api.orkan_state_upload(C.byref(st), arr.ctypes.data)
        assert rc == 0

    def read(self, st):
        out = np.empty(self.length(st), dtype=np.complex128)
        rc = self.api.orkan_state_download(C.byref(st), out.ctypes.data)
        assert rc == 0
        return out


def oracle(log_tile=5):
    return HostEngine(oracle_path(log_tile), log_tile)


def reference(log_tile=5):
    p = ref_path(log_tile)
    return HostEngine(p, log_tile) if os.path.exists(p) else None
