"""Nontemporal (streaming) loads and stores change only the cache policy of an
access, never a value: every kernel family that has them gives bit-identical
results with them on and off.

Statevector kernels use them on states of 512 MiB and more (2^25 amplitudes in
complex128, 2^26 in complex64), where the tuned shapes switch them on; the hook
(kernel 13) turns them off per family and sets that size.  The streamed
density-matrix kernels and the phase kernels (cz, exp_diag) take them from
bit 2 of their launch-shape flags (kernels 4-7, 9, 15, 16, 18) on states of at least that
size, which the test lowers to one byte for a 10-qubit density matrix."""
import ctypes as C

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, to_layout
from benchmarks import workloads as wl

pytestmark = pytest.mark.gpu
NT_OFF_HOOK, RESET_HOOK = 13, 14


@pytest.fixture(scope="module")
def eng():
    e = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    yield e
    assert e.api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0


def _sv_run(eng, n, fp32, off, psi, obs_dev, shape=None):
    api = eng.api
    assert api.orkan_set_launch_shape(NT_OFF_HOOK, off, 0) == 0
    if shape is not None:
        assert api.orkan_set_launch_shape(12, *shape) == 0
    st = eng.new(PURE, n, psi)
    if fp32:
        st32 = api.orkan_state_convert(C.byref(st), 1)
        eng.free(st)
        st = st32
    ps = C.byref(st)
    obs_p = C.cast(obs_dev, C.POINTER(C.c_double))
    for q in (0, 2, 3, 10, n - 1):
        api.rx(ps, q, 0.3 + 0.1 * q)
        api.ry(ps, q, 0.7)
        api.rz(ps, q, 1.1)
        api.h(ps, q)
    for c, t in ((0, 1), (3, 4), (10, 2), (n - 1, 5)):
        api.cx(ps, c, t)
    api.cz(ps, 4, 9)
    api.exp_diag(ps, obs_p, 0.4)
    m = api.mean(ps, obs_p)
    other = api.state_cp(ps)
    api.rx(C.byref(other), 7, 0.9)
    me = api.matel_diag(ps, C.byref(other), obs_p)
    eng.free(other)
    if fp32:
        back = api.orkan_state_convert(ps, 0)
        eng.free(st)
        st = back
    out = eng.read(st)
    eng.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    return out, m, (me.re, me.im)


@pytest.mark.parametrize("fp32", [False, True])
def test_statevector_streaming_accesses_change_no_value(eng, fp32):
    n = 26 if fp32 else 25            # 512 MiB either way
    g = wl.rng(41)
    psi = (g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)) / np.sqrt(2 << n)
    obs = g.uniform(-1, 1, 1 << n)
    obs_dev = eng.api.orkan_dev_alloc(1 << n)
    assert eng.api.orkan_dev_upload(obs_dev, obs.ctypes.data, 1 << n) == 0
    try:
        a, ma, ea = _sv_run(eng, n, fp32, 0, psi, obs_dev)
        b, mb, eb = _sv_run(eng, n, fp32, 15, psi, obs_dev)
        # every target nontemporal: the lane-exchange kernel for targets whose pair shares a line
        c, mc, ec = _sv_run(eng, n, fp32, 0, psi, obs_dev, shape=(16 | 2 | 1, 53248))
    finally:
        eng.api.orkan_dev_free(obs_dev)
    assert np.array_equal(a, b) and np.array_equal(c, b)
    assert ma == mb == mc and ea == eb == ec


@pytest.fixture(scope="module")
def eng3():
    e = engines.DeviceEngine(orkan_amd.lib_path(3), 3)
    yield e
    assert e.api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0


@pytest.mark.parametrize("lt", [5, 3])
@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED])
@pytest.mark.parametrize("fp32", [False, True])
def test_density_matrix_streaming_accesses_change_no_value(eng, eng3, lt, kind, fp32):
    """Both tile sizes: 32 x 32 and 8 x 8 (where the lane-exchange kernel pairs lanes within rows of
    8 columns)."""
    eng = eng if lt == 5 else eng3
    api = eng.api
    n = 10
    rho = to_layout(wl.low_rank_rho(n, 16, wl.rng(43)), kind, lt)
    S = wl.superop(wl.kraus_amp_damp(0.05))
    diag = wl.rng(44).uniform(-1, 1, 1 << n)

    def run(mode):
        # mode 0: cached accesses; 1: nontemporal; 2: nontemporal with bit 6 (packed superoperator: only
        # on the lines a workgroup owns alone; phase kernels: whole-line writes)
        nt, b6 = min(mode, 1), 64 * (mode == 2)
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(NT_OFF_HOOK, 0, 1) == 0
        for k6, k7 in ((6, 7), (15, 16)):   # complex128 and complex64 shapes
            # with nt: the lane-exchange kernel (bit 3) for in-tile targets whose quad columns share lines
            assert api.orkan_set_launch_shape(k6, 1 | 12 * nt, 65536) == 0
            assert api.orkan_set_launch_shape(k7, 4 * nt, 81920) == 0
        assert api.orkan_set_launch_shape(9, 3 | 4 * nt | b6, 81920) == 0
        assert api.orkan_set_launch_shape(18, 3 | 4 * nt | b6, 81920) == 0
        assert api.orkan_set_launch_shape(4, 3 | 4 * nt | b6, 0) == 0
        assert api.orkan_set_launch_shape(5, 4 * nt | b6, 0) == 0
        st = eng.new(kind, n, rho)
        if fp32:
            st32 = api.orkan_state_convert(C.byref(st), 1)
            eng.free(st)
            st = st32
        for q in range(n):
            eng.gate(st, "h", q)
            eng.gate(st, "ry", q, 0.3 + 0.1 * q)
            eng.channel(st, S, q)
        for q in range(n - 1):
            eng.gate(st, "cz", q, q + 1)
        eng.exp_diag(st, diag, 0.7)
        if fp32:
            back = api.orkan_state_convert(C.byref(st), 0)
            eng.free(st)
            st = back
        out = eng.read(st)
        eng.free(st)
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        return out

    base = run(0)
    assert np.array_equal(run(1), base)
    assert np.array_equal(run(2), base)


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED])
@pytest.mark.parametrize("fp32", [False, True])
def test_split_tile_quads_change_no_value(eng, kind, fp32):
    """The streamed superoperator kernels with 2 or 4 workgroups per tile quad (tiled, target above
    the tile) or per 32 x 32 block of quads (packed) -- bits 3-4 of the kernel-7 / 16 / 9 / 18 shape
    flags -- form the same products for every entry as with one: bit-identical."""
    api = eng.api
    n = 9
    rho = to_layout(wl.low_rank_rho(n, 16, wl.rng(45)), kind, 5)
    S = wl.superop(wl.kraus_depolarizing(0.05))
    first = 5 if kind == MIXED_TILED else 0

    def run(lp, nt):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(NT_OFF_HOOK, 0, 1) == 0
        assert api.orkan_set_launch_shape(7, 8 * lp + 4 * nt, 81920) == 0
        assert api.orkan_set_launch_shape(16, 8 * lp + 4 * nt, 81920) == 0
        assert api.orkan_set_launch_shape(9, 8 * lp + 4 * nt + 3, 81920) == 0
        assert api.orkan_set_launch_shape(18, 8 * lp + 4 * nt + 3, 81920) == 0
        st = eng.new(kind, n, rho)
        if fp32:
            st32 = api.orkan_state_convert(C.byref(st), 1)
            eng.free(st)
            st = st32
        for q in range(first, n):
            eng.gate(st, "h", q)
            eng.gate(st, "ry", q, 0.2 + 0.1 * q)
            eng.channel(st, S, q)
        if fp32:
            back = api.orkan_state_convert(C.byref(st), 0)
            eng.free(st)
            st = back
        out = eng.read(st)
        eng.free(st)
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        return out

    ref = run(0, 0)
    for lp in (1, 2):
        for nt in (0, 1):
            assert np.array_equal(run(lp, nt), ref), (lp, nt)


@pytest.mark.parametrize("fp32", [False, True])
def test_exp_diag_shapes_change_no_value(eng, fp32):
    """exp_diag with 1, 2 or 4 amplitudes per thread and any residency cap (tuning hook kernel 19)
    forms the same product for every amplitude: bit-identical, on a size that is no multiple of a
    workgroup's amplitudes (2^7 with 4 per thread)."""
    api = eng.api
    g = wl.rng(47)
    outs = []
    for n in (7, 15):
        psi = (g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)) / np.sqrt(2 << n)
        diag = g.uniform(-3, 3, 1 << n)
        res = []
        for lu, lds in ((0, 0), (1, 32768), (2, 32768), (2, 0)):
            assert api.orkan_set_launch_shape(19, 4 * fp32 + lu, lds) == 0
            st = eng.new(PURE, n, psi)
            if fp32:
                st32 = api.orkan_state_convert(C.byref(st), 1)
                eng.free(st)
                st = st32
            eng.exp_diag(st, diag, 0.9)
            if fp32:
                back = api.orkan_state_convert(C.byref(st), 0)
                eng.free(st)
                st = back
            res.append(eng.read(st))
            eng.free(st)
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert all(np.array_equal(r, res[0]) for r in res[1:])
        outs.append(res[0])
    assert len(outs) == 2
