"""orkan_layer_1q: every qubit's 2x2 matrix applied in LDS-tiled sweeps, against per-qubit
application (numpy on the same state, and the library's own rx / ry / rz calls one by one).
The sweeps were emulated on the CPU first (tile index bijection, each qubit once)."""
import ctypes as C

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_TILED, PURE

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    return engines.DeviceEngine(orkan_amd.lib_path(5), 5)


def _unitaries(rng, n):
    return [np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0] for _ in range(n)]


def _colmajor(mats):
    return np.ascontiguousarray(np.concatenate([m.T.reshape(4) for m in mats]), np.complex128)


def _apply(psi, mats):
    n = len(mats)
    out = psi.reshape([2] * n)
    for q, m in enumerate(mats):   # axis n - 1 - q is qubit q
        out = np.moveaxis(np.tensordot(m, out, axes=([1], [n - 1 - q])), 0, n - 1 - q)
    return out.reshape(-1)


@pytest.mark.parametrize("n", [1, 3, 11, 12, 13, 17, 20, 21])
def test_layer_matches_per_qubit_application(eng, n):
    rng = np.random.default_rng(100 + n)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    mats = _unitaries(rng, n)
    st = eng.new(PURE, n, psi)
    m = _colmajor(mats)
    assert eng.api.orkan_layer_1q(C.byref(st), m.ctypes.data) == 0
    got = eng.read(st)
    eng.free(st)
    assert np.abs(got - _apply(psi, mats)).max() <= 1e-12


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
def test_layer_tile_shapes(eng, variant):
    """Every tile shape of tuning hook 17 (2^12 or 2^11 amplitudes, 256 or 512 threads) on sizes whose sweeps
    take the register-resident path and the generic LDS path (a last sweep of 1-2 qubits)."""
    api = eng.api
    try:
        assert api.orkan_set_launch_shape(17, variant, 0) == 0
        for n in (13, 17, 22):
            rng = np.random.default_rng(300 + 10 * variant + n)
            psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
            psi /= np.linalg.norm(psi)
            mats = _unitaries(rng, n)
            st = eng.new(PURE, n, psi)
            m = _colmajor(mats)
            assert api.orkan_layer_1q(C.byref(st), m.ctypes.data) == 0
            got = eng.read(st)
            eng.free(st)
            assert np.abs(got - _apply(psi, mats)).max() <= 1e-12, (variant, n)
    finally:
        assert api.orkan_set_launch_shape(14, 0, 0) == 0


def test_layer_fp32_storage(eng):
    n = 16
    rng = np.random.default_rng(7)
    psi = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    mats = _unitaries(rng, n)
    st = eng.new(PURE, n, psi)
    s32 = eng.api.orkan_state_convert(C.byref(st), 1)
    eng.free(st)
    m = _colmajor(mats)
    assert eng.api.orkan_layer_1q(C.byref(s32), m.ctypes.data) == 0
    back = eng.api.orkan_state_convert(C.byref(s32), 0)
    eng.free(s32)
    got = eng.read(back)
    eng.free(back)
    # complex64 storage: each of the 2 sweeps rounds to ~6e-8 relative
    assert np.abs(got - _apply(psi, mats)).max() <= 1e-6 * np.abs(psi).max()


def _rot(name, th):
    c, s = np.cos(th / 2), np.sin(th / 2)
    if name == "rx":
        return np.array([[c, -1j * s], [-1j * s, c]])
    if name == "ry":
        return np.array([[c, -s], [s, c]])
    return np.array([[np.exp(-0.5j * th), 0], [0, np.exp(0.5j * th)]])


def test_fused_rotations_match_gate_calls_at_a_gib(eng):
    """A brick layer's rx, ry, rz on every qubit of a 2^26-amplitude state (1 GiB: the nontemporal
    forms of both paths), as one fused layer (U = RZ RY RX per qubit) and as 78 gate calls."""
    n = 26
    rng = np.random.default_rng(11)
    ang = rng.uniform(0, 2 * np.pi, (n, 3))
    api = eng.api
    a = eng.plus(PURE, n)
    for q in range(n):
        api.rx(C.byref(a), q, float(ang[q, 0]))
        api.ry(C.byref(a), q, float(ang[q, 1]))
        api.rz(C.byref(a), q, float(ang[q, 2]))
    b = eng.plus(PURE, n)
    mats = [_rot("rz", ang[q, 2]) @ _rot("ry", ang[q, 1]) @ _rot("rx", ang[q, 0]) for q in range(n)]
    m = _colmajor(mats)
    assert api.orkan_layer_1q(C.byref(b), m.ctypes.data) == 0
    ga, gb = eng.read(a), eng.read(b)
    eng.free(a)
    eng.free(b)
    assert np.abs(ga - gb).max() <= 1e-12


def test_layer_errors(eng):
    api = eng.api
    m = _colmajor(_unitaries(np.random.default_rng(1), 4))
    assert api.orkan_layer_1q(None, m.ctypes.data) == -1          # QS_ERR_NULL
    st = eng.plus(PURE, 4)
    assert api.orkan_layer_1q(C.byref(st), None) == -1
    eng.free(st)
    dm = eng.plus(MIXED_TILED, 4)
    assert api.orkan_layer_1q(C.byref(dm), m.ctypes.data) == -4   # QS_ERR_TYPE: pure states only
    eng.free(dm)
