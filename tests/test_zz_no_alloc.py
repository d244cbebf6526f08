"""orkan_exp_diag_noalloc, the allocation-free exp_diag, called through the C ABI and checked against
exp_diag itself (which shares its kernels but allocates its density-matrix phase scratch): two calls in
a row on each layout and storage precision, the second on a smaller state, results bit-identical; and
its argument checks, which reject before any device work."""
import ctypes as C

import numpy as np
import pytest

import orkan_amd
from orkan_amd import ORKAN_FP32, StateT
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, random_psi, random_rho, to_layout

pytestmark = pytest.mark.gpu

QS_OK, QS_ERR_NULL, QS_ERR_QUBIT, QS_ERR_TYPE, QS_ERR_PARAM = 0, -1, -3, -4, -7
MAX_QUBITS = 17


def _new(lib, kind, n, data, fp32):
    st = StateT(kind, None, 0)
    lib.state_init(C.byref(st), n, None)
    assert st.data
    assert lib.orkan_state_upload(C.byref(st), data.ctypes.data) == 0
    if not fp32:
        return st
    st32 = lib.orkan_state_convert(C.byref(st), 1)
    assert st32.data
    lib.state_free(C.byref(st))
    return st32


def _read(lib, st, fp32):
    out = np.empty(int(lib.state_len(C.byref(st))), np.complex64 if fp32 else np.complex128)
    assert lib.orkan_state_download(C.byref(st), out.ctypes.data) == 0
    return out


@pytest.mark.parametrize("fp32", [False, True], ids=["c128", "c64"])
@pytest.mark.parametrize("kind,sizes", [(MIXED_TILED, (11, 6)), (MIXED_PACKED, (11, 6)), (PURE, (20, 9))],
                         ids=["tiled", "packed", "pure"])
def test_noalloc_matches_exp_diag(kind, sizes, fp32):
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(0x5EED + kind + 7 * fp32)
    for call, n in enumerate(sizes):       # two calls in a row, the second on a smaller state
        print(f"noalloc exp_diag: kind {kind}, n {n}, fp32 {fp32}, call {call}", flush=True)
        data = random_psi(n, rng) if kind == PURE else to_layout(random_rho(n, rng), kind, 5)
        data = np.ascontiguousarray(data, np.complex128)
        diag = rng.uniform(-2.0, 2.0, 1 << n)
        ddev = lib.orkan_dev_alloc(1 << n)
        assert ddev
        assert lib.orkan_dev_upload(ddev, diag.ctypes.data, 1 << n) == 0
        t = float(rng.uniform(0.2, 1.5))
        a, b = _new(lib, kind, n, data, fp32), _new(lib, kind, n, data, fp32)
        lib.exp_diag(C.byref(a), C.cast(ddev, C.POINTER(C.c_double)), t)
        assert lib.orkan_exp_diag_noalloc(C.byref(b), ddev, t) == QS_OK
        ra, rb = _read(lib, a, fp32), _read(lib, b, fp32)
        assert np.array_equal(ra.view(np.uint8), rb.view(np.uint8)), float(np.abs(ra - rb).max())
        assert not np.array_equal(ra, data.astype(ra.dtype))   # the phases were applied
        for st in (a, b):
            lib.state_free(C.byref(st))
        lib.orkan_dev_free(ddev)


def test_noalloc_argument_checks():
    lib = orkan_amd.load(5)
    n = 6
    rng = np.random.default_rng(11)
    data = np.ascontiguousarray(to_layout(random_rho(n, rng), MIXED_TILED, 5), np.complex128)
    st = _new(lib, MIXED_TILED, n, data, False)
    diag = rng.uniform(-1, 1, 1 << n)
    ddev = lib.orkan_dev_alloc(1 << n)
    assert lib.orkan_dev_upload(ddev, diag.ctypes.data, 1 << n) == 0
    # t == 0 leaves the state untouched
    assert lib.orkan_exp_diag_noalloc(C.byref(st), ddev, 0.0) == QS_OK
    assert np.array_equal(_read(lib, st, False), data)
    # a host diagonal is refused (it would need a staging copy)
    assert lib.orkan_exp_diag_noalloc(C.byref(st), diag.ctypes.data, 0.5) == QS_ERR_PARAM
    assert lib.orkan_exp_diag_noalloc(None, ddev, 0.5) == QS_ERR_NULL
    assert lib.orkan_exp_diag_noalloc(C.byref(st), None, 0.5) == QS_ERR_NULL
    assert np.array_equal(_read(lib, st, False), data)
    # a density matrix above the scratch maximum is refused before any device work: the state below
    # claims 18 qubits over a 6-qubit buffer that is never touched
    big = StateT(MIXED_TILED, st.data, MAX_QUBITS + 1)
    assert lib.orkan_exp_diag_noalloc(C.byref(big), ddev, 0.5) == QS_ERR_QUBIT
    bad = StateT(7, st.data, n)
    assert lib.orkan_exp_diag_noalloc(C.byref(bad), ddev, 0.5) == QS_ERR_TYPE
    assert np.array_equal(_read(lib, st, False), data)
    lib.state_free(C.byref(st))
    lib.orkan_dev_free(ddev)
