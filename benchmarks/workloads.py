"""Synthetic workloads of the benchmark configurations, generated on the host
once so that the same arrays feed the GPU path and any CPU baseline.

Random numbers: numpy's PCG64 generator seeded with 0x4F524B414E000000 + k for
configuration k (k = 1..5), angles uniform in [0, 2 pi), observables uniform in
[-1, 1], Gaussian matrices with i.i.d. N(0, 1) real and imaginary parts.
"""
import numpy as np

SEED_BASE = 0x4F524B414E000000


def rng(config):
    return np.random.default_rng(SEED_BASE + config)


# ---- configuration 2: 28-qubit statevector brick circuit ---------------------------------------
SV_QUBITS = 28
SV_LAYERS = 40


def brick_layer(n, layer, angles):
    """Gate list of one brick layer: rx, ry, rz on every qubit (angles: 3n values),
    then cx on (0,1),(2,3),... in even layers and (1,2),(3,4),... in odd layers."""
    ops = []
    for q in range(n):
        ops.append(("rx", q, float(angles[3 * q])))
        ops.append(("ry", q, float(angles[3 * q + 1])))
        ops.append(("rz", q, float(angles[3 * q + 2])))
    start = 0 if layer % 2 == 0 else 1
    for q in range(start, n - 1, 2):
        ops.append(("cx", q, q + 1))
    return ops


def brick_circuit(n=SV_QUBITS, layers=SV_LAYERS, config=2):
    g = rng(config)
    angles = g.uniform(0.0, 2 * np.pi, size=(layers, 3 * n))
    return [brick_layer(n, l, angles[l]) for l in range(layers)], g


def sv_observable(n, g):
    return g.uniform(-1.0, 1.0, size=1 << n)


# algorithmic HBM bytes per gate on a statevector of 2^n complex128 amplitudes:
# every amplitude the gate can change is read once and written once
def sv_gate_bytes(name, n, elem=16):
    N = 1 << n
    if name in ("rx", "ry", "rz", "h", "hy", "x", "y"):
        return 2 * elem * N
    if name in ("p", "z", "s", "sdg", "t", "tdg", "cx", "cy", "swap_gate"):
        return elem * N
    if name in ("cz", "ccx"):
        return elem * N // 2
    raise KeyError(name)


# ---- configuration 3: 14-qubit density matrix ----------------------------------------------
DM_QUBITS = 14
DM_LAYERS = 10


def dm_layer(n, layer, thetas):
    """h on all qubits; ry(theta) on all; cz on alternating neighbours; amplitude
    damping (0.05) and depolarizing (0.02) on all qubits."""
    ops = [("h", q) for q in range(n)]
    ops += [("ry", q, float(thetas[q])) for q in range(n)]
    start = 0 if layer % 2 == 0 else 1
    ops += [("cz", q, q + 1) for q in range(start, n - 1, 2)]
    ops += [("ampdamp", q) for q in range(n)]
    ops += [("depol", q) for q in range(n)]
    return ops


def low_rank_rho(n, r, g):
    """rho = A A^dagger / tr(A A^dagger) with A a D x r complex Gaussian."""
    D = 1 << n
    A = g.standard_normal((D, r)) + 1j * g.standard_normal((D, r))
    rho = A @ A.conj().T
    return rho / np.trace(rho).real


def kraus_depolarizing(p):
    """{sqrt(1-3p/4) I, sqrt(p/4) X, sqrt(p/4) Y, sqrt(p/4) Z}: the reference test
    suite's convention (test/channel/test_channel_tiled.c:38-46)."""
    a, b = np.sqrt(1 - 3 * p / 4), np.sqrt(p / 4)
    X = np.array([[0, 1], [1, 0]], complex)
    Y = np.array([[0, -1j], [1j, 0]], complex)
    Z = np.array([[1, 0], [0, -1]], complex)
    return [a * np.eye(2, dtype=complex), b * X, b * Y, b * Z]


def kraus_amp_damp(gm):
    return [np.array([[1, 0], [0, np.sqrt(1 - gm)]], complex), np.array([[0, np.sqrt(gm)], [0, 0]], complex)]


def superop(kraus):
    """S[(i,j),(k,l)] = sum_a K_a[i,k] conj(K_a[j,l]), row-major (reference
    src/channel/channel.c:21-50)."""
    return sum(np.kron(K, K.conj()) for K in kraus)


def dm_stored_len(n, kind, log_tile=5):
    D = 1 << n
    if kind == 1:
        return D * (D + 1) // 2
    td = 1 << log_tile
    nt = (D + td - 1) // td
    return nt * (nt + 1) // 2 * td * td


def dm_gate_bytes(name, L, elem=16):
    """algorithmic bytes on L stored entries"""
    if name in ("h", "ry", "rx", "x", "y", "hy", "ampdamp", "depol", "channel"):
        return 2 * elem * L
    if name in ("rz", "p", "z", "s", "sdg", "t", "tdg"):
        return elem * L
    if name == "cz":
        return (3 * elem * L) // 4
    if name in ("cx", "cy", "swap_gate"):
        return (3 * elem * L) // 2
    if name == "ccx":
        return (7 * elem * L) // 8
    raise KeyError(name)
