"""Known-answer checks, written once and run against any engine (the CPU oracle
here, the HIP library in the GPU tests).

Closed forms are the ones the reference's own tests assert
(docs/MEAS_MODULE.md:140-191, docs/CIRCUIT_MODULE.md:123-171,
test/channel/test_channel_superop.c, test/channel/test_channel_packed.c:228,363);
gates are checked against brute-force U rho U^dagger / U psi with full 2^n
matrices (the reference's test oracle, test/utility/linalg.c:27-110).
"""
import itertools

import numpy as np

from linalg_np import (MIXED_PACKED, MIXED_TILED, PURE, from_layout, gate_full, kraus_amp_damp,
                       kraus_depolarizing, random_psi, random_rho, superop, to_layout)

TOL = 1e-12
KINDS = (PURE, MIXED_PACKED, MIXED_TILED)
THETAS = [0.0, np.pi / 4, np.pi / 2, np.pi, 3 * np.pi / 2, -np.pi / 3]


def load(eng, kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == PURE:
        v = random_psi(n, rng)
        return eng.new(kind, n, v), v
    rho = random_rho(n, rng)
    return eng.new(kind, n, to_layout(rho, kind, eng.log_tile)), rho


def logical(eng, st, kind, n):
    return from_layout(eng.read(st), kind, n, eng.log_tile)


def expect_gate(ref, name, qs, n, theta, kind):
    G = gate_full(name, qs, n, theta)
    return G @ ref if kind == PURE else G @ ref @ G.conj().T


def gate_cases(n, rng, limit=None):
    """(name, qubits, theta) for every gate; positions sampled when `limit` is set."""
    cases = []
    for q in range(n):
        for name in ("x", "y", "z", "h", "s", "sdg", "t", "tdg", "hy"):
            cases.append((name, (q,), None))
        for name in ("rx", "ry", "rz", "p"):
            for th in THETAS:
                cases.append((name, (q,), th))
    for a, b in itertools.permutations(range(n), 2):
        for name in ("cx", "cy", "cz", "swap_gate"):
            cases.append((name, (a, b), None))
    for a, b, c in itertools.permutations(range(n), 3):
        cases.append(("ccx", (a, b, c), None))
    if limit is not None and len(cases) > limit:
        idx = rng.choice(len(cases), size=limit, replace=False)
        cases = [cases[i] for i in sorted(idx)]
    return cases


def check_gates_bruteforce(eng, kind, n, limit=None, seed=0):
    rng = np.random.default_rng(seed)
    worst = 0.0
    for name, qs, th in gate_cases(n, rng, limit):
        st, ref = load(eng, kind, n, seed + 17)
        args = list(qs) + ([th] if th is not None else [])
        eng.gate(st, name, *args)
        want = expect_gate(ref, name, qs, n, th or 0.0, kind)
        got = logical(eng, st, kind, n)
        err = float(np.abs(got - want).max())
        eng.free(st)
        assert err <= TOL, (name, qs, th, kind, n, err)
        worst = max(worst, err)
    return worst


def check_channels_bruteforce(eng, kind, n, seed=0):
    rng = np.random.default_rng(seed)
    krs = [kraus_depolarizing(0.1), kraus_depolarizing(0.02), kraus_amp_damp(0.05),
           kraus_amp_damp(1.0), kraus_depolarizing(1.0),
           list(rng.standard_normal((3, 2, 2)) + 1j * rng.standard_normal((3, 2, 2)))]
    for ks in krs:
        S = superop(ks)
        for q in range(n):
            st, rho = load(eng, kind, n, seed + q)
            eng.channel(st, S, q)
            want = sum(gate_full_k(K, q, n) @ rho @ gate_full_k(K, q, n).conj().T for K in ks)
            got = logical(eng, st, kind, n)
            eng.free(st)
            assert np.abs(got - want).max() <= TOL


def kraus_2q_sets(rng):
    """2-qubit Kraus sets (local index x_q1 + 2 x_q2): two-qubit depolarizing, amplitude damping
    on both qubits (kron(K_q2, K_q1)), and a random trace-preserving set."""
    from linalg_np import mat1
    P = [np.eye(2), mat1("x"), mat1("y"), mat1("z")]
    pr = 0.07
    dep = [np.sqrt(1 - 15 * pr / 16) * np.eye(4)] + [np.sqrt(pr / 16) * np.kron(a, b)
                                                     for i, a in enumerate(P) for j, b in enumerate(P) if i or j]
    ad = [np.kron(a, b) for a in kraus_amp_damp(0.05) for b in kraus_amp_damp(0.2)]
    rnd = rng.standard_normal((3, 4, 4)) + 1j * rng.standard_normal((3, 4, 4))
    # made trace preserving (K_a <- K_a M^{-1/2}, M = sum_a K_a^dagger K_a) so that repeated
    # application keeps entries of order one
    M = sum(K.conj().T @ K for K in rnd)
    w, V = np.linalg.eigh(M)
    rnd = [K @ (V @ np.diag(w ** -0.5) @ V.conj().T) for K in rnd]
    return [dep, ad, rnd]


def check_channel2_bruteforce(eng, kind, n, seed=0):
    """channel_2q against sum_a K_a rho K_a^dagger on dense matrices, every ordered qubit pair."""
    from linalg_np import embed
    rng = np.random.default_rng(seed)
    for ks in kraus_2q_sets(rng):
        S = superop(ks)
        for q1 in range(n):
            for q2 in range(n):
                if q1 == q2:
                    continue
                st, rho = load(eng, kind, n, seed + 7 * q1 + q2)
                eng.channel2(st, S, q1, q2)
                F = [embed(np.asarray(K, complex), [q1, q2], n) for K in ks]
                want = sum(f @ rho @ f.conj().T for f in F)
                got = logical(eng, st, kind, n)
                eng.free(st)
                assert np.abs(got - want).max() <= TOL * max(1.0, np.abs(want).max()), (q1, q2)


def gate_full_k(K, q, n):
    from linalg_np import embed
    return embed(np.asarray(K, complex), [q], n)


def check_mean_closed_forms(eng, kind, n):
    D = 1 << n
    st = eng.plus(kind, n)
    assert abs(eng.mean(st, np.ones(D)) - 1.0) <= TOL
    assert abs(eng.mean(st, np.arange(D, dtype=float)) - (D - 1) / 2) <= TOL
    assert abs(eng.mean(st, -np.ones(D)) + 1.0) <= TOL
    eng.free(st)
    for k in (0, D - 1, D // 3):
        if kind == PURE:
            v = np.zeros(D, complex); v[k] = 1
            st = eng.new(kind, n, v)
        else:
            r = np.zeros((D, D), complex); r[k, k] = 1
            st = eng.new(kind, n, to_layout(r, kind, eng.log_tile))
        assert abs(eng.mean(st, np.arange(D, dtype=float)) - k) <= TOL
        eng.free(st)
    if n == 2:
        st = eng.plus(kind, n)
        assert abs(eng.mean(st, np.array([-3.0, 1.0, -1.0, 3.0]))) <= TOL
        eng.free(st)


def check_matel_closed_forms(eng):
    s2 = 1 / np.sqrt(2)
    bra = eng.new(PURE, 1, np.array([s2, 1j * s2]))
    ket = eng.new(PURE, 1, np.array([s2, s2]))
    assert abs(eng.matel(bra, ket, np.array([0.0, 1.0])) - (-0.5j)) <= TOL
    eng.free(bra); eng.free(ket)
    bra = eng.new(PURE, 2, np.full(4, 0.5, complex))
    ket = eng.new(PURE, 2, np.full(4, 0.5j, complex))
    assert abs(eng.matel(bra, ket, np.arange(4.0)) - 1.5j) <= TOL
    eng.free(bra); eng.free(ket)
    b0 = eng.new(PURE, 3, np.eye(8, dtype=complex)[0])
    b1 = eng.new(PURE, 3, np.eye(8, dtype=complex)[1])
    assert abs(eng.matel(b0, b1, np.arange(8.0))) <= TOL
    eng.free(b0); eng.free(b1)


def check_exp_diag_closed_forms(eng, kind, n):
    D = 1 << n
    x = np.arange(D, dtype=float)
    st = eng.plus(kind, n)
    eng.exp_diag(st, x, np.pi / 4)
    got = logical(eng, st, kind, n)
    if kind == PURE:
        want = np.exp(-1j * x * np.pi / 4) / np.sqrt(D)
    else:
        want = np.exp(-1j * (x[:, None] - x[None, :]) * np.pi / 4) / D
    assert np.abs(got - want).max() <= TOL
    # t then -t is the identity; t twice equals 2t
    eng.exp_diag(st, x, -np.pi / 4)
    base = eng.plus(kind, n)
    assert np.abs(eng.read(st) - eng.read(base)).max() <= TOL
    eng.exp_diag(st, x, 0.3); eng.exp_diag(st, x, 0.3)
    eng.exp_diag(base, x, 0.6)
    assert np.abs(eng.read(st) - eng.read(base)).max() <= TOL
    eng.free(st); eng.free(base)
    if n == 1 and kind == MIXED_TILED:
        st = eng.plus(kind, 1)
        eng.exp_diag(st, np.array([0.0, 1.0]), np.pi / 2)
        assert abs(eng.get(st, 1, 0) - (-0.5j)) <= TOL
        eng.free(st)


def check_channel_invariants(eng, kind, n, seed=0):
    """trace preservation; full depolarisation of every qubit gives I/D;
    amplitude damping with gamma=1 on every qubit gives |0><0|."""
    st, rho = load(eng, kind, n, seed)
    Sdep = superop(kraus_depolarizing(1.0))
    for q in range(n):
        eng.channel(st, Sdep, q)
    got = logical(eng, st, kind, n)
    assert np.abs(got - np.eye(1 << n) / (1 << n)).max() <= TOL
    eng.free(st)
    st, rho = load(eng, kind, n, seed + 1)
    Sad = superop(kraus_amp_damp(1.0))
    for q in range(n):
        eng.channel(st, Sad, q)
    got = logical(eng, st, kind, n)
    want = np.zeros_like(got); want[0, 0] = 1
    assert np.abs(got - want).max() <= TOL
    eng.free(st)
    st, rho = load(eng, kind, n, seed + 2)
    eng.channel(st, superop(kraus_amp_damp(0.3)), n - 1)
    assert abs(np.trace(logical(eng, st, kind, n)) - 1.0) <= TOL
    eng.free(st)
