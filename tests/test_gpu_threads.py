"""Host threads driving different states at once (the reference's threading
contract: calls are synchronous and carry no global mutable state, so
different states may be driven from different host threads,
include/state.h and the module docs).  Each thread runs a small circuit with
gates, a channel, exp_diag and mean on its own state; every result must equal
the same circuit run alone."""
import ctypes as C
import threading

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, random_psi, random_rho, superop, kraus_amp_damp, to_layout

pytestmark = pytest.mark.gpu


def _circuit(eng, kind, n, seed):
    rng = np.random.default_rng(seed)
    data = random_psi(n, rng) if kind == PURE else to_layout(random_rho(n, rng), kind, eng.log_tile)
    diag = rng.uniform(-1, 1, 1 << n)
    st = eng.new(kind, n, data)
    S = superop(kraus_amp_damp(0.1))
    means = []
    for layer in range(4):
        for q in range(n):
            eng.gate(st, "ry", q, 0.1 * (q + 1) + layer)
        for q in range(0, n - 1, 2):
            eng.gate(st, "cx", q, q + 1)
        if kind != PURE:
            eng.channel(st, S, layer % n)
        eng.exp_diag(st, diag, 0.2 + layer)
        means.append(eng.mean(st, diag))
    out = eng.read(st)
    eng.free(st)
    return out, means


def test_threads_on_separate_states():
    eng = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    jobs = [(PURE, 14, 1), (MIXED_PACKED, 7, 2), (MIXED_TILED, 7, 3), (PURE, 12, 4), (MIXED_TILED, 6, 5),
            (MIXED_PACKED, 6, 6)]
    alone = [_circuit(eng, *j) for j in jobs]
    results = [None] * len(jobs)
    errors = []

    def work(i):
        try:
            for _ in range(3):
                results[i] = _circuit(eng, *jobs[i])
        except Exception as e:   # pragma: no cover - reported below
            errors.append(e)

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(jobs))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=240)
    assert not errors and all(not t.is_alive() for t in threads)
    for (a_state, a_means), (b_state, b_means) in zip(alone, results):
        assert np.array_equal(a_state, b_state)
        assert a_means == b_means
