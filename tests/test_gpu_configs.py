"""The benchmark configurations, run through the HIP library and checked
against the CPU oracle (and against the compiled reference where it is
present) on the same seeded inputs.

config 1: 10-qubit tiled density matrix rho = A A^dagger / tr, A a seeded
          1024 x 1024 complex Gaussian; H(0), CX(0,1), RZ(2, 0.5),
          Toffoli(0,1,2), depolarizing p=0.1 on every qubit; mean of a seeded
          uniform [-1, 1] diagonal.  Every entry at 1e-12.
config 2: 28-qubit statevector brick layer (full size) and a 20-qubit
          multi-layer run, at 1e-12 per amplitude.
config 3: 14-qubit density matrix layer (full size) on both layouts at 1e-12.
FP32 storage: max amplitude error, trace and Hermiticity drift against FP64.
"""
import ctypes as C
import os

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, from_layout, tiled_diag_mirror_error, to_layout
from benchmarks import workloads as wl

pytestmark = pytest.mark.gpu
TOL = 1e-12


@pytest.fixture(scope="module")
def dev():
    return engines.DeviceEngine(orkan_amd.lib_path(5), 5)


@pytest.fixture(scope="module")
def orc():
    return engines.oracle(5)


def config1_inputs():
    g = wl.rng(1)
    n = 10
    A = g.standard_normal((1 << n, 1 << n)) + 1j * g.standard_normal((1 << n, 1 << n))
    rho = A @ A.conj().T
    rho /= np.trace(rho).real
    obs = g.uniform(-1, 1, 1 << n)
    return n, rho, obs


def run_config1(eng, n, rho, obs):
    st = eng.new(MIXED_TILED, n, to_layout(rho, MIXED_TILED, 5))
    eng.gate(st, "h", 0)
    eng.gate(st, "cx", 0, 1)
    eng.gate(st, "rz", 2, 0.5)
    eng.gate(st, "ccx", 0, 1, 2)
    S = wl.superop(wl.kraus_depolarizing(0.1))
    for q in range(n):
        eng.channel(st, S, q)
    m = eng.mean(st, obs)
    data = eng.read(st)
    eng.free(st)
    return data, m


def test_config1_exact(dev, orc):
    n, rho, obs = config1_inputs()
    d_dev, m_dev = run_config1(dev, n, rho, obs)
    d_orc, m_orc = run_config1(orc, n, rho, obs)
    assert np.abs(d_dev - d_orc).max() <= TOL
    assert abs(m_dev - m_orc) <= TOL
    full = from_layout(d_dev, MIXED_TILED, n, 5)          # all 1,048,576 logical entries
    assert np.abs(full - from_layout(d_orc, MIXED_TILED, n, 5)).max() <= TOL
    ref = engines.reference(5)
    if ref is not None:
        d_ref, m_ref = run_config1(ref, n, rho, obs)
        assert np.abs(d_dev - d_ref).max() <= TOL
        assert abs(m_dev - m_ref) <= TOL


def _apply_sv(eng, st, ops):
    for op in ops:
        eng.gate(st, *op)


def test_config2_full_size_layer(dev, orc):
    n = wl.SV_QUBITS
    layers, g = wl.brick_circuit(n, 1)
    sd, so = dev.plus(PURE, n), orc.plus(PURE, n)
    _apply_sv(dev, sd, layers[0])
    print("config2: device layer done", flush=True)
    _apply_sv(orc, so, layers[0])
    print("config2: oracle layer done", flush=True)
    a, b = dev.read(sd), orc.read(so)
    assert np.abs(a - b).max() <= TOL
    assert abs(np.vdot(a, a).real - 1.0) <= 1e-10
    obs = wl.sv_observable(n, g)
    assert abs(dev.mean(sd, obs) - orc.mean(so, obs)) <= TOL
    dev.free(sd); orc.free(so)


def test_config2_many_layers(dev, orc):
    n = 20
    layers, g = wl.brick_circuit(n, 8)
    sd, so = dev.plus(PURE, n), orc.plus(PURE, n)
    for ops in layers:
        _apply_sv(dev, sd, ops)
        _apply_sv(orc, so, ops)
    assert np.abs(dev.read(sd) - orc.read(so)).max() <= TOL
    obs = wl.sv_observable(n, g)
    assert abs(dev.mean(sd, obs) - orc.mean(so, obs)) <= TOL
    dev.free(sd); orc.free(so)


def _dm_ops(eng, st, ops):
    Sad = wl.superop(wl.kraus_amp_damp(0.05))
    Sdp = wl.superop(wl.kraus_depolarizing(0.02))
    for op in ops:
        if op[0] == "ampdamp":
            eng.channel(st, Sad, op[1])
        elif op[0] == "depol":
            eng.channel(st, Sdp, op[1])
        else:
            eng.gate(st, *op)


@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_config3_full_size_layer(dev, orc, kind):
    n = wl.DM_QUBITS
    g = wl.rng(3)
    rho = wl.low_rank_rho(n, 16, g)
    thetas = g.uniform(0, 2 * np.pi, size=(2, n))
    data = to_layout(rho, kind, 5)
    del rho
    sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
    del data
    ops = wl.dm_layer(n, 0, thetas[0])
    _dm_ops(dev, sd, ops)
    print(f"config3 kind={kind}: device layer done", flush=True)
    for i in range(0, len(ops), 8):
        _dm_ops(orc, so, ops[i:i + 8])
        print(f"config3 kind={kind}: oracle {min(i + 8, len(ops))}/{len(ops)} ops", flush=True)
    a, b = dev.read(sd), orc.read(so)
    assert np.abs(a - b).max() <= TOL
    dev.free(sd); orc.free(so)


@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_config3_layers_small(dev, orc, kind):
    n = 9
    g = wl.rng(3)
    rho = wl.low_rank_rho(n, 64, g)
    thetas = g.uniform(0, 2 * np.pi, size=(10, n))
    data = to_layout(rho, kind, 5)
    sd, so = dev.new(kind, n, data), orc.new(kind, n, data)
    for l in range(10):
        ops = wl.dm_layer(n, l, thetas[l])
        _dm_ops(dev, sd, ops)
        _dm_ops(orc, so, ops)
        assert np.abs(dev.read(sd) - orc.read(so)).max() <= TOL, l
    dev.free(sd); orc.free(so)


# ---- FP32 storage ------------------------------------------------------------------------------
def test_fp32_statevector_error(dev):
    n = 20
    layers, _ = wl.brick_circuit(n, 6)
    s64 = dev.plus(PURE, n)
    s32 = dev.plus(PURE | orkan_amd.ORKAN_FP32, n)
    for ops in layers:
        _apply_sv(dev, s64, ops)
        _apply_sv(dev, s32, ops)
    a = dev.read(s64)
    conv = dev.api.orkan_state_convert(C.byref(s32), 0)
    b = dev.read(conv)
    err = float(np.abs(a - b).max())
    assert err < 1e-6, err
    assert abs(np.vdot(b, b).real - 1.0) < 1e-5
    for s in (s64, s32, conv):
        dev.free(s)


@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_fp32_density_matrix_drift(dev, kind):
    n = 9
    g = wl.rng(3)
    rho = wl.low_rank_rho(n, 64, g)
    thetas = g.uniform(0, 2 * np.pi, size=(10, n))
    d = dev.new(kind, n, to_layout(rho, kind, 5))
    f = dev.api.orkan_state_convert(C.byref(d), 1)
    for l in range(10):
        ops = wl.dm_layer(n, l, thetas[l])
        _dm_ops(dev, d, ops)
        _dm_ops(dev, f, ops)
    back = dev.api.orkan_state_convert(C.byref(f), 0)
    a, b = dev.read(d), dev.read(back)
    assert np.abs(a - b).max() < 1e-6
    full = from_layout(b, kind, n, 5)
    trace_drift = abs(np.trace(full).real - 1.0)
    herm_diag = float(np.abs(np.diag(full).imag).max())
    assert trace_drift < 1e-5 and herm_diag < 1e-6
    if kind == MIXED_TILED:
        assert tiled_diag_mirror_error(b, n, 5) < 1e-6
    for s in (d, f, back):
        dev.free(s)


# ---- streamed tiled superoperator vs the LDS tile-group kernel ---------------------------------
@pytest.mark.parametrize("kind,lt,n", [(MIXED_TILED, 5, 8), (MIXED_TILED, 5, 9), (MIXED_TILED, 3, 7),
                                       (MIXED_PACKED, 5, 6), (MIXED_PACKED, 5, 9)])
@pytest.mark.parametrize("fp32", [False, True])
def test_streamed_superop_matches_tile_groups(kind, lt, n, fp32):
    """The streamed kernels (tiled: target inside a tile; target above it, with
    and without a mirrored tile, and diagonal tile quads; packed: direct,
    mirrored and diagonal quads) against the tile-group kernel.  Both form the same products in the same order; where a stored
    entry is the mirror of an upper-triangle entry, the tile-group kernel may
    write it as the conjugate of that entry's value instead, which differs by
    rounding only.  One gate from the same input agrees to 1e-15 (FP64) and a
    whole sequence of gates to 1e-13."""
    eng = engines.DeviceEngine(orkan_amd.lib_path(lt), lt)
    api = eng.api
    g = wl.rng(17 + n)
    rho = to_layout(wl.low_rank_rho(n, 16, g), kind, lt)
    S = wl.superop(wl.kraus_amp_damp(0.05))
    ops = [(lambda st, q: eng.gate(st, "h", q)), (lambda st, q: eng.gate(st, "ry", q, 0.3 + 0.1 * q)),
           (lambda st, q: eng.channel(st, S, q))]

    def run(on, seq):
        assert api.orkan_set_launch_shape(8, 0, on) == 0
        st = eng.new(kind, n, rho)
        if fp32:
            st32 = api.orkan_state_convert(C.byref(st), 1)
            eng.free(st)
            st = st32
        for op, q in seq:
            op(st, q)
        if fp32:
            back = api.orkan_state_convert(C.byref(st), 0)
            eng.free(st)
            st = back
        out = eng.read(st)
        eng.free(st)
        return out

    try:
        for q in range(n):
            for op in ops:
                a, b = run(1, [(op, q)]), run(0, [(op, q)])
                assert np.abs(a - b).max() <= (1e-6 if fp32 else 1e-15), q
        seq = [(op, q) for q in range(n) for op in ops]
        a, b = run(1, seq), run(0, seq)
        assert np.abs(a - b).max() <= (1e-6 if fp32 else 1e-13)
    finally:
        assert api.orkan_set_launch_shape(8, 0, 1) == 0
