"""Pins the CPU oracle to the mathematics, independently of the compiled
reference: brute-force U rho U^dagger for every gate, sum_k K rho K^dagger for
channels, and the closed forms the reference's own tests assert."""
import numpy as np
import pytest

import engines
import known_answers as ka
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", ka.KINDS)
def test_gates_bruteforce(log_tile, kind):
    eng = engines.oracle(log_tile)
    for n in (1, 2, 3, 4):
        ka.check_gates_bruteforce(eng, kind, n, seed=n)
    if kind == MIXED_TILED:
        for n in (log_tile + 1, log_tile + 2):
            ka.check_gates_bruteforce(eng, kind, n, limit=120, seed=n)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_channels_bruteforce(log_tile, kind):
    eng = engines.oracle(log_tile)
    for n in (1, 2, 3, log_tile + 1):
        ka.check_channels_bruteforce(eng, kind, n, seed=n)
        ka.check_channel_invariants(eng, kind, n, seed=n)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", [MIXED_PACKED, MIXED_TILED])
def test_channel_2q_bruteforce(log_tile, kind):
    """channel_2q has no reference counterpart: pinned to sum_a K_a rho K_a^dagger."""
    eng = engines.oracle(log_tile)
    for n in (2, 3, log_tile + 1):
        ka.check_channel2_bruteforce(eng, kind, n, seed=n)


@pytest.mark.parametrize("log_tile", [3, 5])
@pytest.mark.parametrize("kind", ka.KINDS)
def test_closed_forms(log_tile, kind):
    eng = engines.oracle(log_tile)
    for n in (1, 2, 3, 4, log_tile + 1):
        ka.check_mean_closed_forms(eng, kind, n)
        ka.check_exp_diag_closed_forms(eng, kind, n)
    if kind == PURE:
        ka.check_matel_closed_forms(eng)


def test_mean_consistency_across_layouts():
    """test_mean_consistency / test_exp_diag_consistency (test/meas/test_meas.c:71,
    test/circ/test_circ.c:75): one logical state, three layouts, same results."""
    from linalg_np import random_psi, to_layout
    eng = engines.oracle(3)
    rng = np.random.default_rng(9)
    n = 5
    psi = random_psi(n, rng)
    rho = np.outer(psi, psi.conj())
    obs = rng.uniform(-1, 1, 1 << n)
    vals, phased = [], []
    for kind in (PURE, MIXED_PACKED, MIXED_TILED):
        st = eng.new(kind, n, psi if kind == PURE else to_layout(rho, kind, 3))
        vals.append(eng.mean(st, obs))
        eng.exp_diag(st, obs, 0.8)
        full = ka.logical(eng, st, kind, n)
        phased.append(np.outer(full, full.conj()) if kind == PURE else full)
        eng.free(st)
    assert max(vals) - min(vals) <= 1e-12
    assert np.abs(phased[0] - phased[1]).max() <= 1e-12
    assert np.abs(phased[0] - phased[2]).max() <= 1e-12
