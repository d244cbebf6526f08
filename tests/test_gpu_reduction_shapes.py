"""mean / matel_diag under every first-stage shape of their two-stage reduction (tuning hook kernel
10: 4 or 8 loads in flight per thread, at most 2048 >> k workgroups, two residency caps): each
value against the CPU oracle to 1e-12, and repeated calls under one shape bit-identical (the
partials are summed in a fixed order)."""
import ctypes as C

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import PURE
from benchmarks import workloads as wl

pytestmark = pytest.mark.gpu
RED_HOOK, RESET_HOOK = 10, 14


@pytest.mark.parametrize("fp32", [False, True])
@pytest.mark.parametrize("n", [12, 21])
def test_reduction_shapes_match_oracle(n, fp32):
    """complex64 states (bit 4 of the hook flags selects their shape) hold the rounded amplitudes:
    their values are checked against the oracle on the same rounded vectors."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    ora = engines.oracle(5)
    g = wl.rng(60 + n)
    psi = g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)
    phi = g.standard_normal(1 << n) + 1j * g.standard_normal(1 << n)
    psi /= np.linalg.norm(psi)
    phi /= np.linalg.norm(phi)
    obs = g.uniform(-1, 1, 1 << n)
    if fp32:
        psi = psi.astype(np.complex64).astype(np.complex128)
        phi = phi.astype(np.complex64).astype(np.complex128)
    oa, ob = ora.new(PURE, n, psi), ora.new(PURE, n, phi)
    m_ref, e_ref = ora.mean(oa, obs), ora.matel(oa, ob, obs)
    ora.free(oa)
    ora.free(ob)
    # closed form of the oracle's values on the same vectors
    assert abs(m_ref - np.sum(obs * np.abs(psi) ** 2)) <= 1e-13
    assert abs(e_ref - np.vdot(psi, obs * phi)) <= 1e-13
    a, b = dev.new(PURE, n, psi), dev.new(PURE, n, phi)
    if fp32:
        a32, b32 = dev.api.orkan_state_convert(C.byref(a), 1), dev.api.orkan_state_convert(C.byref(b), 1)
        dev.free(a)
        dev.free(b)
        a, b = a32, b32
    try:
        for flags in range(10):
            for lds in (0, 98304):
                assert dev.api.orkan_set_launch_shape(RED_HOOK, 16 * fp32 + flags, lds) == 0
                m1, m2 = dev.mean(a, obs), dev.mean(a, obs)
                e1, e2 = dev.matel(a, b, obs), dev.matel(a, b, obs)
                assert m1 == m2 and e1 == e2, (flags, lds)
                assert abs(m1 - m_ref) <= 1e-12, (flags, lds)
                assert abs(e1 - e_ref) <= 1e-12, (flags, lds)
    finally:
        assert dev.api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        dev.free(a)
        dev.free(b)
