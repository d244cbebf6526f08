"""The CPU oracle against the reference library compiled from its own sources
(oracle/_ref, built by oracle/Makefile when /root/reference is present).

Every gate at every qubit position / ordered pair / ordered triple, the channel,
mean, matel_diag, exp_diag, state_plus and Hermitian get/set, on seeded random
inputs, at both tile sizes the reference tests use (8x8 and 32x32), to the
reference's own tolerance 1e-12 (test/include/test.h:40).
"""
import itertools

import numpy as np
import pytest

import engines
from linalg_np import (MIXED_PACKED, MIXED_TILED, PURE, from_layout, gate_full, kraus_amp_damp,
                       kraus_depolarizing, random_psi, random_rho, superop, to_layout)

TOL = 1e-12
THETAS = [0.0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, -np.pi / 3]   # test/include/test_gate.h:217
KINDS = [PURE, MIXED_PACKED, MIXED_TILED]


def _pair(log_tile):
    ref = engines.reference(log_tile)
    if ref is None:
        pytest.skip("reference library not built (oracle/_ref)")
    return engines.oracle(log_tile), ref


def _load(eng, kind, n, rng_seed):
    rng = np.random.default_rng(rng_seed)
    if kind == PURE:
        data = random_psi(n, rng)
    else:
        data = to_layout(random_rho(n, rng), kind, eng.log_tile)
    return eng.new(kind, n, data)


def _check(o, r, so, sr):
    a, b = o.read(so), r.read(sr)
    assert a.shape == b.shape
    err = np.abs(a - b).max() if a.size else 0.0
    assert err <= TOL, err


def _ns(kind, log_tile):
    return [1, 2, 3, 4] if kind != MIXED_TILED else list(range(1, log_tile + 3))


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", KINDS)
def test_1q_gates(log_tile, kind):
    o, r = _pair(log_tile)
    for n in _ns(kind, log_tile):
        for q in range(n):
            for name in ["x", "y", "z", "h", "s", "sdg", "t", "tdg", "hy"]:
                so, sr = _load(o, kind, n, 7 + n), _load(r, kind, n, 7 + n)
                o.gate(so, name, q); r.gate(sr, name, q)
                _check(o, r, so, sr)
                o.free(so); r.free(sr)
            for name in ["rx", "ry", "rz", "p"]:
                for th in THETAS:
                    so, sr = _load(o, kind, n, 11 + n), _load(r, kind, n, 11 + n)
                    o.gate(so, name, q, th); r.gate(sr, name, q, th)
                    _check(o, r, so, sr)
                    o.free(so); r.free(sr)


def ref_swap_defect(kind, name, qs, log_tile):
    """The reference's swap_tiled, case 3 (both qubits >= LOG_TILE_DIM,
    src/gate/tiled/gate_tiled_swap.c:304-474), pairs the wrong tiles: its output
    is not U rho U^dagger there, which is what its own test demands
    (test/gate/test_gate.c:100 with SWAPMAT) and what its packed and pure swaps
    compute.  The oracle follows the definition; at those positions the
    comparison is against brute force instead."""
    return kind == MIXED_TILED and name == "swap_gate" and min(qs) >= log_tile


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", KINDS)
def test_2q_3q_gates(log_tile, kind):
    o, r = _pair(log_tile)
    defects = candidates = 0
    for n in _ns(kind, log_tile):
        for a, b in itertools.permutations(range(n), 2):
            for name in ["cx", "cy", "cz", "swap_gate"]:
                so, sr = _load(o, kind, n, 3 * n), _load(r, kind, n, 3 * n)
                o.gate(so, name, a, b); r.gate(sr, name, a, b)
                if ref_swap_defect(kind, name, (a, b), log_tile):
                    rho = random_rho(n, np.random.default_rng(3 * n))
                    G = gate_full(name, (a, b), n)
                    want = G @ rho @ G.conj().T
                    assert np.abs(from_layout(o.read(so), kind, n, log_tile) - want).max() <= TOL
                    defects += np.abs(from_layout(r.read(sr), kind, n, log_tile) - want).max() > TOL
                    candidates += 1
                else:
                    _check(o, r, so, sr)
                o.free(so); r.free(sr)
        for a, b, c in itertools.permutations(range(n), 3):
            so, sr = _load(o, kind, n, 5 * n), _load(r, kind, n, 5 * n)
            o.gate(so, "ccx", a, b, c); r.gate(sr, "ccx", a, b, c)
            _check(o, r, so, sr)
            o.free(so); r.free(sr)
    # the documented divergence stays pinned: the reference deviates from U rho U^dagger on every
    # such case (a fixed reference would fail here, and the exception above would then be removed)
    assert defects == candidates
    assert candidates == (0 if kind != MIXED_TILED else 2)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_channel(log_tile, kind):
    o, r = _pair(log_tile)
    rng = np.random.default_rng(5)
    sops = [superop(kraus_depolarizing(0.1)), superop(kraus_amp_damp(0.05)),
            superop(kraus_depolarizing(1.0)), superop(kraus_amp_damp(1.0)),
            # a random Hermiticity-preserving map (sum of K (x) conj K over random K).  The
            # layouts store only a Hermitian matrix, so a map that does not keep rho
            # Hermitian has no layout-independent result; channels always keep it.
            superop(list(rng.standard_normal((3, 2, 2)) + 1j * rng.standard_normal((3, 2, 2))))]
    for n in _ns(kind, log_tile):
        for q in range(n):
            for S in sops:
                so, sr = _load(o, kind, n, 21 + n), _load(r, kind, n, 21 + n)
                o.channel(so, S, q); r.channel(sr, S, q)
                _check(o, r, so, sr)
                o.free(so); r.free(sr)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", KINDS)
def test_mean_exp_diag(log_tile, kind):
    o, r = _pair(log_tile)
    for n in _ns(kind, log_tile) + [9]:
        rng = np.random.default_rng(100 + n)
        obs = rng.uniform(-1, 1, 1 << n)
        so, sr = _load(o, kind, n, 31 + n), _load(r, kind, n, 31 + n)
        assert abs(o.mean(so, obs) - r.mean(sr, obs)) <= TOL
        for t in [0.0, 0.7, -2.5]:
            o.exp_diag(so, obs, t); r.exp_diag(sr, obs, t)
            _check(o, r, so, sr)
        o.free(so); r.free(sr)


@pytest.mark.parametrize("log_tile", [3, 5])
def test_matel_diag(log_tile):
    o, r = _pair(log_tile)
    for n in [1, 3, 6, 13]:
        rng = np.random.default_rng(n)
        obs = rng.uniform(-1, 1, 1 << n)
        b0, k0 = _load(o, PURE, n, 1), _load(o, PURE, n, 2)
        b1, k1 = _load(r, PURE, n, 1), _load(r, PURE, n, 2)
        assert abs(o.matel(b0, k0, obs) - r.matel(b1, k1, obs)) <= TOL
        for s in (b0, k0):
            o.free(s)
        for s in (b1, k1):
            r.free(s)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", KINDS)
def test_plus_get_set(log_tile, kind):
    o, r = _pair(log_tile)
    for n in range(0, 7):
        so, sr = o.plus(kind, n), r.plus(kind, n)
        assert o.length(so) == r.length(sr)
        _check(o, r, so, sr)
        D = 1 << n
        rng = np.random.default_rng(n)
        for _ in range(20):
            row = int(rng.integers(D))
            col = 0 if kind == PURE else int(rng.integers(D))
            v = complex(rng.standard_normal(), rng.standard_normal())
            if kind != PURE and row == col:
                v = v.real
            o.set(so, row, col, v); r.set(sr, row, col, v)
            assert o.get(so, row, col) == r.get(sr, row, col)
        _check(o, r, so, sr)
        o.free(so); r.free(sr)


@pytest.mark.parametrize("log_tile", [3, 5])
def test_generic_matrices_tiled(log_tile):
    """single_from_mat / two_from_mat, exported tiled backends of the reference."""
    import ctypes as C
    o, r = _pair(log_tile)
    rng = np.random.default_rng(77)
    for n in range(2, log_tile + 3):
        for q in range(n):
            U = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
            m = np.ascontiguousarray(U.T.reshape(-1))            # column-major
            so, sr = _load(o, MIXED_TILED, n, 41), _load(r, MIXED_TILED, n, 41)
            o.api.single_from_mat(C.byref(so), q, m.ctypes.data)
            r.api.single_from_mat(C.byref(sr), q, m.ctypes.data)
            _check(o, r, so, sr)
            o.free(so); r.free(sr)
        for a, b in itertools.permutations(range(n), 2):
            U = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
            m = np.ascontiguousarray(U.T.reshape(-1))
            so, sr = _load(o, MIXED_TILED, n, 43), _load(r, MIXED_TILED, n, 43)
            o.api.two_from_mat(C.byref(so), a, b, m.ctypes.data)
            r.api.two_from_mat(C.byref(sr), a, b, m.ctypes.data)
            _check(o, r, so, sr)
            o.free(so); r.free(sr)


def test_kraus_to_superop():
    """the oracle's, the compiled reference's and the shipped library's kraus_to_superop
    (orkan_amd/csrc/ork_channel.c; host code, no device call) against sum_a kron(K_a, conj K_a)
    (reference src/channel/channel.c:21-50)"""
    import ctypes as C
    import orkan_amd
    from orkan_amd import _ffi

    class Shipped:
        api = orkan_amd.load(5)
        libc = C.CDLL(None)

    o, r = _pair(5)
    rng = np.random.default_rng(3)
    for nq in (1, 2, 3):
        N = 1 << nq
        ks = rng.standard_normal((3, N, N)) + 1j * rng.standard_normal((3, N, N))
        flat = np.ascontiguousarray(ks.reshape(-1))
        kr = _ffi.KrausT(nq, 3, flat.ctypes.data)
        outs = []
        for eng in (o, r, Shipped):
            so = eng.api.kraus_to_superop(C.byref(kr))
            assert so.n_qubits == nq
            arr = np.ctypeslib.as_array(C.cast(so.data, C.POINTER(C.c_double)), shape=(2 * N ** 4,)).copy()
            outs.append(arr.view(np.complex128))
            eng.libc.free(C.c_void_p(so.data))
        want = sum(np.kron(K, K.conj()) for K in ks).reshape(-1)
        for got in outs:
            assert np.abs(got - want).max() <= TOL
        assert np.abs(outs[2] - outs[1]).max() <= TOL
