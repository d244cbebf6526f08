"""ctypes description of the Orkan C API (include/*.h).

`bind(path)` loads any shared library that exports the Orkan API and returns
an `Api` whose attributes are the typed C entry points.  The same ABI is
exported by the HIP library of this package (device-resident states) and by
any host implementation of the reference interface.

ABI facts (reference include/state.h:31-44, include/channel.h:28-45,
include/q_types.h:46-48): `state_t` is {int enum; cplx_t*; unsigned char},
`cplx_t` is C99 `double _Complex`, which the x86-64 SysV ABI passes and
returns exactly like a struct of two doubles.
"""
import ctypes as C

PURE, MIXED_PACKED, MIXED_TILED = 0, 1, 2


class Cplx(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]

    def value(self):
        return complex(self.re, self.im)

    @classmethod
    def of(cls, z):
        z = complex(z)
        return cls(z.real, z.imag)


class StateT(C.Structure):
    _fields_ = [("type", C.c_int), ("data", C.c_void_p), ("qubits", C.c_ubyte)]


class KrausT(C.Structure):
    _fields_ = [("n_qubits", C.c_ubyte), ("n_terms", C.c_uint64), ("data", C.c_void_p)]


class SuperopT(C.Structure):
    _fields_ = [("n_qubits", C.c_ubyte), ("data", C.c_void_p)]


QOP = C.CFUNCTYPE(None, C.POINTER(StateT))


class AnsatzChannelT(C.Structure):
    _fields_ = [("n_ops", C.c_uint64), ("ops", C.POINTER(QOP)), ("weights", C.c_void_p)]


PS = C.POINTER(StateT)
Q = C.c_ubyte
D = C.c_double
DP = C.POINTER(C.c_double)
VP = C.c_void_p

# name -> (restype, argtypes); every entry point of include/{state,gate,channel,meas,circ}.h
API = {
    "state_len": (C.c_uint64, [PS]),
    "state_init": (None, [PS, Q, C.POINTER(VP)]),
    "state_free": (None, [PS]),
    "state_plus": (None, [PS, Q]),
    "state_cp": (StateT, [PS]),
    "state_print": (None, [PS]),
    "state_get": (Cplx, [PS, C.c_uint64, C.c_uint64]),
    "state_set": (None, [PS, C.c_uint64, C.c_uint64, Cplx]),
    "x": (None, [PS, Q]), "y": (None, [PS, Q]), "z": (None, [PS, Q]),
    "h": (None, [PS, Q]), "s": (None, [PS, Q]), "sdg": (None, [PS, Q]),
    "t": (None, [PS, Q]), "tdg": (None, [PS, Q]), "hy": (None, [PS, Q]),
    "rx": (None, [PS, Q, D]), "ry": (None, [PS, Q, D]),
    "rz": (None, [PS, Q, D]), "p": (None, [PS, Q, D]),
    "cx": (None, [PS, Q, Q]), "cy": (None, [PS, Q, Q]), "cz": (None, [PS, Q, Q]),
    "swap_gate": (None, [PS, Q, Q]),
    "ccx": (None, [PS, Q, Q, Q]),
    "kraus_to_superop": (SuperopT, [C.POINTER(KrausT)]),
    "channel_1q": (None, [PS, C.POINTER(SuperopT), Q]),
    "channel_2q": (None, [PS, C.POINTER(SuperopT), Q, Q]),
    "sample_matrix": (C.c_int, [PS, C.c_void_p, C.POINTER(AnsatzChannelT), C.c_uint64, C.c_uint64, C.c_size_t,
                                C.c_void_p]),
    "mean": (D, [PS, DP]),
    "matel_diag": (Cplx, [PS, PS, DP]),
    "exp_diag": (None, [PS, DP, D]),
}

# Backend symbols the reference exports without declaring them in its headers
# (src/gate/tiled/gate_tiled.h:32, src/gate/tiled/gate_tiled_2q.c:3,
# src/channel/channel.c:57-58); its tests and profilers link against them.
BACKEND = {
    "single_from_mat": (None, [PS, Q, VP]),
    "two_from_mat": (None, [PS, Q, Q, VP]),
    "channel_packed_1q": (None, [PS, VP, Q]),
    "channel_tiled_1q": (None, [PS, VP, Q]),
}

GATES_1Q = ["x", "y", "z", "h", "s", "sdg", "t", "tdg", "hy"]
GATES_ROT = ["rx", "ry", "rz", "p"]
GATES_2Q = ["cx", "cy", "cz", "swap_gate"]


class Api:
    """Typed entry points of one loaded library."""

    def __init__(self, path, extra=None):
        self.path = str(path)
        self.lib = C.CDLL(self.path, mode=C.RTLD_LOCAL)
        table = dict(API)
        table.update(BACKEND)
        if extra:
            table.update(extra)
        for name, (res, args) in table.items():
            try:
                fn = getattr(self.lib, name)
            except AttributeError:
                continue
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)

    def has(self, name):
        return hasattr(self, name) and callable(getattr(self, name))


def bind(path, extra=None):
    return Api(path, extra)


def dptr(arr):
    """Pointer to the first double of a C-contiguous float64 numpy array."""
    return arr.ctypes.data_as(DP)
