"""numpy helpers for the tests: storage layouts, full gate matrices, random states.

Layouts (reference docs/STATE_MODULE.md:114-157, src/internal/index.h:57-88):
  packed -- lower triangle, column-major: (r, c), r >= c, at c(2D-c+1)/2 + (r-c)
  tiled  -- lower tiles (tr >= tc) in row-major tile order, each TILE_DIM x TILE_DIM
            row-major; diagonal tiles hold both triangles; n < LOG_TILE_DIM uses
            one zero-padded tile with row stride TILE_DIM.
Gate matrices follow the reference's own test matrices (test/utility/gatemat.c:267-315)
and gate comments (src/gate/gate_pure.c:148-267).
"""
import numpy as np

PURE, MIXED_PACKED, MIXED_TILED = 0, 1, 2
S2 = 1.0 / np.sqrt(2.0)


# ---------------------------------------------------------------------------- layouts
def state_len(kind, n, log_tile=5):
    D = 1 << n
    if kind == PURE:
        return D
    if kind == MIXED_PACKED:
        return 1 if n == 0 else (D // 2) * (D + 1)
    td = 1 << log_tile
    nt = (D + td - 1) // td
    return nt * (nt + 1) // 2 * td * td


def to_layout(rho, kind, log_tile=5):
    rho = np.asarray(rho, dtype=np.complex128)
    if kind == PURE:
        return rho.copy()
    D = rho.shape[0]
    if kind == MIXED_PACKED:
        i, j = np.triu_indices(D)
        return np.ascontiguousarray(rho.T[i, j])
    td = 1 << log_tile
    if D < td:
        out = np.zeros((td, td), dtype=np.complex128)
        out[:D, :D] = rho
        return out.ravel()
    nt = D // td
    tiles = rho.reshape(nt, td, nt, td).transpose(0, 2, 1, 3)
    tr, tc = np.tril_indices(nt)
    return np.ascontiguousarray(tiles[tr, tc].reshape(-1))


def from_layout(data, kind, n, log_tile=5):
    """Full D x D matrix (Hermitian completion of the stored lower triangle)."""
    data = np.asarray(data)
    D = 1 << n
    if kind == PURE:
        return data.copy()
    if kind == MIXED_PACKED:
        low = np.zeros((D, D), dtype=np.complex128)
        i, j = np.triu_indices(D)
        low.T[i, j] = data
    else:
        td = 1 << log_tile
        if D < td:
            low = data.reshape(td, td)[:D, :D].copy()
        else:
            nt = D // td
            tiles = np.zeros((nt, nt, td, td), dtype=np.complex128)
            tr, tc = np.tril_indices(nt)
            tiles[tr, tc] = data.reshape(-1, td, td)
            low = tiles.transpose(0, 2, 1, 3).reshape(D, D)
        low = np.tril(low)
    full = low + np.tril(low, -1).conj().T
    return full


def tiled_diag_mirror_error(data, n, log_tile=5):
    """max |t[l,c] - conj(t[c,l])| over the diagonal tiles of a tiled state."""
    td = 1 << log_tile
    D = 1 << n
    if D < td:
        t = data.reshape(td, td)[:D, :D]
        return float(np.abs(t - t.conj().T).max()) if D else 0.0
    nt = D // td
    worst = 0.0
    for k in range(nt):
        off = (k * (k + 1) // 2 + k) * td * td
        t = data[off:off + td * td].reshape(td, td)
        worst = max(worst, float(np.abs(t - t.conj().T).max()))
    return worst


# ---------------------------------------------------------------------------- gates
def mat1(name, theta=0.0):
    c, s = np.cos(theta / 2), np.sin(theta / 2)
    return {
        "x": np.array([[0, 1], [1, 0]], complex),
        "y": np.array([[0, -1j], [1j, 0]], complex),
        "z": np.array([[1, 0], [0, -1]], complex),
        "h": S2 * np.array([[1, 1], [1, -1]], complex),
        "hy": S2 * np.array([[1, -1], [1, 1]], complex),
        "s": np.array([[1, 0], [0, 1j]], complex),
        "sdg": np.array([[1, 0], [0, -1j]], complex),
        "t": np.array([[1, 0], [0, np.exp(1j * np.pi / 4)]], complex),
        "tdg": np.array([[1, 0], [0, np.exp(-1j * np.pi / 4)]], complex),
        "rx": np.array([[c, -1j * s], [-1j * s, c]], complex),
        "ry": np.array([[c, -s], [s, c]], complex),
        "rz": np.array([[np.exp(-1j * theta / 2), 0], [0, np.exp(1j * theta / 2)]], complex),
        "p": np.array([[1, 0], [0, np.exp(1j * theta)]], complex),
    }[name]


def local_matrix(name, theta=0.0):
    """(matrix, k): local unitary; bit j of the local index <-> the j-th gate argument."""
    if name in ("cx", "cy", "cz"):
        g = {"cx": mat1("x"), "cy": mat1("y"), "cz": mat1("z")}[name]
        U = np.eye(4, dtype=complex)
        # local index a = ctrl_bit + 2*target_bit; act on target when ctrl = 1 (a = 1, 3)
        for tb_out in range(2):
            for tb_in in range(2):
                U[1 + 2 * tb_out, 1 + 2 * tb_in] = g[tb_out, tb_in]
        return U, 2
    if name == "swap_gate":
        U = np.eye(4, dtype=complex)[[0, 2, 1, 3]]
        return U, 2
    if name == "ccx":
        U = np.eye(8, dtype=complex)
        U[[3, 7]] = U[[7, 3]]
        return U, 3
    return mat1(name, theta), 1


def embed(U, qs, n):
    """Full 2^n matrix of local unitary U acting on qubits qs (bit j <-> qs[j])."""
    D = 1 << n
    k = len(qs)
    mask = 0
    for q in qs:
        mask |= 1 << q
    x = np.arange(D)
    loc = np.zeros(D, dtype=np.int64)
    for j, q in enumerate(qs):
        loc |= ((x >> q) & 1) << j
    rest = x & ~mask
    full = np.zeros((D, D), dtype=complex)
    for a in range(1 << k):
        spread = 0
        for j, q in enumerate(qs):
            if a >> j & 1:
                spread |= 1 << q
        rows = rest | spread
        full[rows, x] = U[a, loc]
    return full


def gate_full(name, qs, n, theta=0.0):
    U, _ = local_matrix(name, theta)
    return embed(U, list(qs), n)


# ---------------------------------------------------------------------------- states
def random_rho(n, rng, rank=None):
    D = 1 << n
    r = D if rank is None else rank
    A = rng.standard_normal((D, r)) + 1j * rng.standard_normal((D, r))
    rho = A @ A.conj().T
    return rho / np.trace(rho).real


def random_psi(n, rng):
    v = rng.standard_normal(1 << n) + 1j * rng.standard_normal(1 << n)
    return v / np.linalg.norm(v)


def kraus_depolarizing(pr):
    """Reference test-suite form {sqrt(1-3p/4) I, sqrt(p/4) X, Y, Z}
    (test/channel/test_channel_tiled.c:38-46)."""
    a, b = np.sqrt(1 - 3 * pr / 4), np.sqrt(pr / 4)
    return [a * np.eye(2), b * mat1("x"), b * mat1("y"), b * mat1("z")]


def kraus_amp_damp(g):
    """K0=[[1,0],[0,sqrt(1-g)]], K1=[[0,sqrt(g)],[0,0]] (test/channel/test_channel_tiled.c:48-55)."""
    return [np.array([[1, 0], [0, np.sqrt(1 - g)]], complex), np.array([[0, np.sqrt(g)], [0, 0]], complex)]


def superop(kraus):
    """S[(i,j),(k,l)] = sum_a K[i,k] conj(K[j,l]) (row-major, vec(rho) row-major)."""
    return sum(np.kron(K, K.conj()) for K in kraus)
