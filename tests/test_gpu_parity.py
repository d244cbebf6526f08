"""Parity of the HIP library with the CPU oracle on the GPU (gfx950).

Same seeded inputs into both; every gate at every position / ordered pair /
ordered triple, the channel, mean, matel_diag, exp_diag, state_plus and
get/set, at both tile sizes (8x8 and 32x32) and on all three layouts.  The bar
is the reference's own tolerance, 1e-12 absolute per entry
(test/include/test.h:40), on the raw storage arrays (which includes the upper
triangle of diagonal tiles).
"""
import numpy as np
import pytest

import engines
import known_answers as ka
import orkan_amd
from linalg_np import (MIXED_PACKED, MIXED_TILED, PURE, kraus_amp_damp, kraus_depolarizing, random_psi,
                       random_rho, superop, to_layout)

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module", params=[5, 3], ids=["tile32", "tile8"])
def pair(request):
    lt = request.param
    return engines.DeviceEngine(orkan_amd.lib_path(lt), lt), engines.oracle(lt)


def _data(kind, n, seed, lt):
    rng = np.random.default_rng(seed)
    if kind == PURE:
        return random_psi(n, rng)
    return to_layout(random_rho(n, rng), kind, lt)


def _sizes(kind, lt):
    if kind == PURE:
        return [1, 2, 3, 5, 8]
    if kind == MIXED_PACKED:
        return [1, 2, 3, 6, 7]
    return [1, 2, lt, lt + 1, lt + 2]


def _same(dev, orc, sd, so, what):
    a, b = dev.read(sd), orc.read(so)
    err = float(np.abs(a - b).max()) if a.size else 0.0
    assert err <= TOL, (what, err)


@pytest.mark.parametrize("kind", [PURE, MIXED_PACKED, MIXED_TILED])
def test_every_gate_every_position(pair, kind):
    dev, orc = pair
    for n in _sizes(kind, dev.log_tile):
        data = _data(kind, n, 100 + n, dev.log_tile)
        sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
        rng = np.random.default_rng(n)
        for name, qs, th in ka.gate_cases(n, rng, limit=400):
            args = list(qs) + ([th] if th is not None else [])
            dev.gate(sd, name, *args)
            orc.gate(so, name, *args)
            _same(dev, orc, sd, so, (n, name, qs, th))
        dev.free(sd); orc.free(so)


@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_channels(pair, kind):
    dev, orc = pair
    rng = np.random.default_rng(1)
    sops = [superop(kraus_depolarizing(0.1)), superop(kraus_amp_damp(0.05)), superop(kraus_depolarizing(0.02)),
            superop(list(rng.standard_normal((3, 2, 2)) + 1j * rng.standard_normal((3, 2, 2))))]
    for n in _sizes(kind, dev.log_tile):
        data = _data(kind, n, 200 + n, dev.log_tile)
        sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
        for S in sops:
            for q in range(n):
                dev.channel(sd, S, q)
                orc.channel(so, S, q)
                _same(dev, orc, sd, so, (n, q))
        dev.free(sd); orc.free(so)


@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_channel_2q(pair, kind):
    """channel_2q (no reference counterpart; the oracle's is pinned to sum_a K_a rho K_a^dagger)
    on every ordered qubit pair."""
    dev, orc = pair
    rng = np.random.default_rng(3)
    sops = [superop(ks) for ks in ka.kraus_2q_sets(rng)]
    for n in [2, 3] + _sizes(kind, dev.log_tile)[3:]:
        data = _data(kind, n, 300 + n, dev.log_tile)
        sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
        for S in sops:
            for q1 in range(n):
                for q2 in range(n):
                    if q1 != q2:
                        dev.channel2(sd, S, q1, q2)
                        orc.channel2(so, S, q1, q2)
                        _same(dev, orc, sd, so, (n, q1, q2))
        dev.free(sd); orc.free(so)


@pytest.mark.parametrize("kind", [PURE, MIXED_PACKED, MIXED_TILED])
def test_mean_exp_diag_plus(pair, kind):
    dev, orc = pair
    for n in _sizes(kind, dev.log_tile) + [9]:
        rng = np.random.default_rng(300 + n)
        obs = rng.uniform(-1, 1, 1 << n)
        data = _data(kind, n, 300 + n, dev.log_tile)
        sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
        assert abs(dev.mean(sd, obs) - orc.mean(so, obs)) <= TOL
        for t in (0.0, 0.7, -2.5):
            dev.exp_diag(sd, obs, t)
            orc.exp_diag(so, obs, t)
            _same(dev, orc, sd, so, ("exp_diag", n, t))
        dev.free(sd); orc.free(so)
        pd, po = dev.plus(kind, n), orc.plus(kind, n)
        _same(dev, orc, pd, po, ("plus", n))
        D = 1 << n
        for _ in range(6):
            r = int(rng.integers(D)); c = 0 if kind == PURE else int(rng.integers(D))
            v = complex(rng.standard_normal(), rng.standard_normal()) if r != c or kind == PURE else 0.25
            dev.set(pd, r, c, v); orc.set(po, r, c, v)
            assert dev.get(pd, r, c) == orc.get(po, r, c)
        _same(dev, orc, pd, po, ("set", n))
        dev.free(pd); orc.free(po)


def test_matel_diag(pair):
    dev, orc = pair
    for n in (1, 4, 11):
        rng = np.random.default_rng(n)
        obs = rng.uniform(-1, 1, 1 << n)
        a, b = random_psi(n, rng), random_psi(n, rng)
        bd, kd = dev.new(PURE, n, a), dev.new(PURE, n, b)
        bo, ko = orc.new(PURE, n, a), orc.new(PURE, n, b)
        assert abs(dev.matel(bd, kd, obs) - orc.matel(bo, ko, obs)) <= TOL
        for s in (bd, kd):
            dev.free(s)
        for s in (bo, ko):
            orc.free(s)


@pytest.mark.parametrize("kind", [PURE, MIXED_PACKED, MIXED_TILED])
def test_known_answers_on_device(pair, kind):
    dev, _ = pair
    for n in (1, 2, 3, dev.log_tile + 1):
        ka.check_mean_closed_forms(dev, kind, n)
        ka.check_exp_diag_closed_forms(dev, kind, n)
        if kind != PURE:
            ka.check_channel_invariants(dev, kind, n, seed=n)
    if kind == PURE:
        ka.check_matel_closed_forms(dev)


def test_generic_matrices(pair):
    import ctypes as C
    dev, orc = pair
    rng = np.random.default_rng(7)
    for kind in (MIXED_PACKED, MIXED_TILED):
        for n in (2, dev.log_tile + 1):
            data = _data(kind, n, 400 + n, dev.log_tile)
            sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
            for q in range(n):
                U = np.linalg.qr(rng.standard_normal((2, 2)) + 1j * rng.standard_normal((2, 2)))[0]
                m = np.ascontiguousarray(U.T.reshape(-1))
                dev.api.single_from_mat(C.byref(sd), q, m.ctypes.data)
                orc.api.single_from_mat(C.byref(so), q, m.ctypes.data)
                _same(dev, orc, sd, so, ("single_from_mat", q))
            for a in range(n):
                for b in range(n):
                    if a == b:
                        continue
                    U = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
                    m = np.ascontiguousarray(U.T.reshape(-1))
                    dev.api.two_from_mat(C.byref(sd), a, b, m.ctypes.data)
                    orc.api.two_from_mat(C.byref(so), a, b, m.ctypes.data)
                    _same(dev, orc, sd, so, ("two_from_mat", a, b))
            dev.free(sd); orc.free(so)


def test_state_init_takes_ownership_of_host_buffer(pair):
    """state_init(s, n, &buf) with a malloc'd host buffer: the data lands in the state,
    the library frees the buffer and nulls the caller's pointer (src/state/state.c:60-99)."""
    import ctypes as C
    from orkan_amd import _ffi
    dev, orc = pair
    libc = C.CDLL(None)
    libc.malloc.restype = C.c_void_p
    libc.malloc.argtypes = [C.c_size_t]
    for kind, n in ((PURE, 5), (MIXED_TILED, 4), (MIXED_PACKED, 3)):
        data = _data(kind, n, 7, dev.log_tile)
        buf = C.c_void_p(libc.malloc(data.nbytes))
        C.memmove(buf.value, data.ctypes.data, data.nbytes)
        st = _ffi.StateT(kind, None, 0)
        dev.api.state_init(C.byref(st), n, C.byref(buf))
        assert buf.value is None
        assert np.array_equal(dev.read(st), data)
        dev.free(st)
