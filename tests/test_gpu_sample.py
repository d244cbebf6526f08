"""sample_matrix (specified in the reference's docs/SAMPLE_MATRIX.md, not implemented there)
against brute force: every leaf |psi_a> built independently with dense numpy matrices and
S_ab = <psi_a| diag(obs) |psi_b>; Hermiticity and real diagonal exactly; the same result bit for
bit whatever the leaf block size the workspace allows; the prefix left in iota."""
import ctypes as C

import numpy as np
import pytest

import orkan_amd
from linalg_np import PURE, gate_full, random_psi
from orkan_amd import _ffi

pytestmark = pytest.mark.gpu


def _ops(lib, specs):
    """ctypes qop_t callbacks applying this library's gates, and their dense matrices"""
    cbs, mats = [], []
    for name, qs, th in specs:
        fn = getattr(lib, name)
        args = list(qs) + ([th] if th is not None else [])
        cbs.append(_ffi.QOP(lambda st, fn=fn, args=args: fn(st, *args)))
        mats.append(gate_full(name, qs, N_Q, th or 0.0))
    return cbs, mats


N_Q = 5


def _ansatz(lib, rng, sizes, lstar):
    names = [("h", 1, False), ("rx", 1, True), ("ry", 1, True), ("rz", 1, True), ("cx", 2, False),
             ("cz", 2, False), ("s", 1, False), ("t", 1, False)]
    chans, keep = [], []
    for l, K in enumerate(sizes):
        specs = []
        for _ in range(K):
            name, k, ang = names[rng.integers(len(names))]
            qs = tuple(int(q) for q in rng.choice(N_Q, size=k, replace=False))
            specs.append((name, qs, float(rng.uniform(0, 2 * np.pi)) if ang else None))
        cbs, mats = _ops(lib, specs)
        w = (rng.standard_normal(K) + 1j * rng.standard_normal(K)) / np.sqrt(K)
        arr = (_ffi.QOP * K)(*cbs)
        wa = np.ascontiguousarray(w, np.complex128)
        keep += [cbs, arr, wa]
        chans.append((arr, None if l == lstar else wa, mats, w))
    table = (_ffi.AnsatzChannelT * len(sizes))(*[
        _ffi.AnsatzChannelT(len(m), a, None if wa is None else wa.ctypes.data) for a, wa, m, _ in chans])
    return table, chans, keep


def _brute(psi0, obs, chans, lstar):
    psi = psi0.copy()
    for l in range(lstar):
        _, _, mats, w = chans[l]
        psi = sum(wm * (U @ psi) for U, wm in zip(mats, w))
    leaves = []
    for U in chans[lstar][2]:
        v = U @ psi
        for l in range(lstar + 1, len(chans)):
            _, _, mats, w = chans[l]
            v = sum(wm * (M @ v) for M, wm in zip(mats, w))
        leaves.append(v)
    Phi = np.stack(leaves, 1)
    return Phi.conj().T @ (obs[:, None] * Phi), psi


def _run(lib, psi0, obs, table, L, lstar, K, budget):
    st = _ffi.StateT(PURE, None, 0)
    lib.state_init(C.byref(st), N_Q, None)
    assert lib.orkan_state_upload(C.byref(st), psi0.ctypes.data) == 0
    S = np.zeros(K * K, np.complex128)
    rc = lib.sample_matrix(C.byref(st), obs.ctypes.data, table, L, lstar, budget, S.ctypes.data)
    assert rc == 0
    pre = np.empty(1 << N_Q, np.complex128)
    assert lib.orkan_state_download(C.byref(st), pre.ctypes.data) == 0
    lib.state_free(C.byref(st))
    return S.reshape(K, K, order="F"), pre


@pytest.mark.parametrize("sizes,lstar", [((4,), 0), ((3, 6, 2), 1), ((2, 3, 5), 2), ((70, 2), 0)])
def test_sample_matrix_against_brute_force(sizes, lstar):
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(sum(sizes) + lstar)
    table, chans, keep = _ansatz(lib, rng, sizes, lstar)
    psi0 = np.ascontiguousarray(random_psi(N_Q, rng))
    obs = np.ascontiguousarray(rng.uniform(-1, 1, 1 << N_Q))
    want, pre_want = _brute(psi0, obs, chans, lstar)
    K = sizes[lstar]
    vec = 16 << N_Q
    S_big, pre = _run(lib, psi0, obs, table, len(sizes), lstar, K, 200 * vec)
    S_small, _ = _run(lib, psi0, obs, table, len(sizes), lstar, K, 4 * vec)   # leaf block of 1
    scale = max(1.0, np.abs(want).max())
    assert np.abs(S_big - want).max() <= 1e-12 * scale
    assert np.array_equal(S_big, S_small)                      # independent of the block size
    assert np.array_equal(S_big, S_big.conj().T)               # Hermitian by construction
    assert np.all(np.diag(S_big).imag == 0.0)
    assert np.abs(pre - pre_want).max() <= 1e-12                # iota consumed: holds the prefix
    del keep
