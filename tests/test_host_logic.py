"""Host-side logic that needs no GPU: workload generators, byte accounting,
surrogate features, MaxCut diagonals, layout helpers."""
import numpy as np

from linalg_np import MIXED_PACKED, MIXED_TILED, from_layout, random_rho, state_len, to_layout
from benchmarks import workloads as wl
from orkan_amd import batch, surrogate
from benchmarks.dm import build_rho_layout


def test_brick_layer_structure():
    layers, _ = wl.brick_circuit(28, 2)
    even, odd = layers
    assert len(even) == 84 + 14 and len(odd) == 84 + 13
    assert [op[0] for op in even[:3]] == ["rx", "ry", "rz"]
    assert [op[1:] for op in even if op[0] == "cx"][:2] == [(0, 1), (2, 3)]
    assert [op[1:] for op in odd if op[0] == "cx"][:2] == [(1, 2), (3, 4)]
    assert all(0 <= op[2] < 2 * np.pi for op in even if op[0] != "cx")


def test_byte_accounting():
    N = 1 << 28
    assert wl.sv_gate_bytes("rx", 28) == 8_589_934_592 == 32 * N
    assert wl.sv_gate_bytes("cx", 28) == 16 * N
    assert wl.sv_gate_bytes("cz", 28) == 8 * N
    L = wl.dm_stored_len(14, MIXED_TILED)
    assert L == 134_479_872 and wl.dm_stored_len(14, MIXED_PACKED) == 134_225_920
    assert wl.dm_gate_bytes("h", L) == 32 * L and wl.dm_gate_bytes("cz", L) == 12 * L
    assert wl.dm_gate_bytes("cx", L) == 24 * L and wl.dm_gate_bytes("ccx", L) == 14 * L


def test_dm_layer_structure():
    ops = wl.dm_layer(14, 1, np.zeros(14))
    kinds = [o[0] for o in ops]
    assert kinds.count("h") == 14 and kinds.count("ry") == 14 and kinds.count("cz") == 6
    assert kinds.count("ampdamp") == 14 and kinds.count("depol") == 14


def test_low_rank_layout_builder_matches_dense():
    g = np.random.default_rng(1)
    for kind in (MIXED_PACKED, MIXED_TILED):
        n = 7
        data = build_rho_layout(n, kind, np.random.default_rng(5), rank=4)
        A = np.random.default_rng(5).standard_normal((1 << n, 4)) + 1j * np.random.default_rng(5).standard_normal((1 << n, 4))
        full = from_layout(data, kind, n, 5)
        assert abs(np.trace(full).real - 1.0) < 1e-12
        assert np.allclose(full, full.conj().T)
        assert data.size == state_len(kind, n, 5)
        assert np.linalg.matrix_rank(full, tol=1e-10) == 4
    del g


def test_maxcut_diag_bruteforce():
    rng = np.random.default_rng(2)
    i = batch.maxcut_instances(1, rng, n_lo=6, n_hi=6)[0]
    d = batch.maxcut_diag(6, i["edges"], i["w"])
    for x in (0, 5, 37, 63):
        bits = [(x >> k) & 1 for k in range(6)]
        want = sum(w for (a, b), w in zip(i["edges"], i["w"]) if bits[a] != bits[b])
        assert abs(d[x] - want) < 1e-12
    assert d[0] == 0.0 and abs(d.max() - d[::-1].max()) < 1e-12   # C(x) = C(~x)


def test_instance_generator_ranges():
    inst = batch.maxcut_instances(400, wl.rng(4))
    ns = [i["n"] for i in inst]
    ps = [i["p"] for i in inst]
    assert min(ns) == 10 and max(ns) == 20 and min(ps) == 1 and max(ps) == 8
    for i in inst[:50]:
        assert i["edges"].shape[0] == i["w"].size
        assert np.all(i["edges"][:, 0] < i["edges"][:, 1])
        assert np.all((0 <= i["gamma"]) & (i["gamma"] <= np.pi)) and i["gamma"].size == i["p"]


def test_surrogate_features():
    w = np.array([0.3, 0.9, 0.1])
    gf = surrogate.graph_features(w)
    assert gf.shape == (120,) and np.allclose(gf[:3], [0.9, 0.3, 0.1]) and not gf[3:].any()
    X = surrogate.features(gf, np.full((2, 4), np.pi / 2), np.full((2, 4), np.pi))
    assert X.shape == (2, 128) and np.allclose(X[:, 120:124], 0.5) and np.allclose(X[:, 124:], 1.0)


def test_layout_roundtrip():
    rng = np.random.default_rng(3)
    for lt in (3, 5):
        for kind in (MIXED_PACKED, MIXED_TILED):
            for n in (1, 2, 4, 6):
                rho = random_rho(n, rng)
                assert np.allclose(from_layout(to_layout(rho, kind, lt), kind, n, lt), rho)


def test_state_file_errors_without_device(tmp_path):
    """orkan_state_load rejects missing, foreign, truncated and inconsistent files before it
    touches the device (QS_ERR_FILE = -5, QS_ERR_FORMAT = -6, QS_ERR_NULL = -1)."""
    import ctypes as C
    import struct
    import orkan_amd
    lib = orkan_amd.load(5)
    st = orkan_amd.StateT(0, None, 0)
    assert lib.orkan_state_load(C.byref(st), str(tmp_path / "missing.bin").encode()) == -5
    assert lib.orkan_state_load(None, b"x") == -1
    assert lib.orkan_state_save(None, b"x") == -1

    def header(magic=b"ORKAN-ST", version=1, typ=0, qubits=3, lt=5, length=8, eb=16, bom=0x01020304):
        return magic + struct.pack("<IIIIQII", version, typ, qubits, lt, length, eb, bom) + bytes(24)

    cases = {
        "short.bin": b"ORKAN-ST" + bytes(10),
        "magic.bin": header(magic=b"NOTORKAN") + bytes(8 * 16),
        "version.bin": header(version=2) + bytes(8 * 16),
        "bom.bin": header(bom=0x04030201) + bytes(8 * 16),
        "type.bin": header(typ=7) + bytes(8 * 16),
        "elem.bin": header(eb=8) + bytes(8 * 16),
        "length.bin": header(length=9) + bytes(9 * 16),
        "tile.bin": header(typ=2, qubits=5, lt=3, length=state_len(2, 5, 5)) + bytes(state_len(2, 5, 5) * 16),
        "truncated.bin": header() + bytes(7 * 16),
        "long.bin": header() + bytes(9 * 16),
        # lengths whose byte count wraps 64 bits (2^60 x 16 = 2^64 -> 0), or whose packed length does
        # (D (D + 1) / 2 at D = 2^62): rejected before any allocation
        "wrap_pure.bin": header(qubits=60, length=1 << 60) + bytes(0),
        "wrap_packed.bin": header(typ=1, qubits=62, length=((1 << 61) * ((1 << 62) + 1)) % (1 << 64)),
        "wrap_fp32.bin": header(typ=0x100, qubits=61, length=1 << 61, eb=8),
        "huge.bin": header(qubits=57, length=1 << 57) + bytes(16),
    }
    for name, blob in cases.items():
        p = tmp_path / name
        p.write_bytes(blob)
        assert lib.orkan_state_load(C.byref(st), str(p).encode()) == -6, name
        assert st.data is None and st.qubits == 0, name


def test_sample_matrix_validation_without_device():
    """sample_matrix rejects bad arguments before it touches the device (QS_ERR_NULL -1,
    QS_ERR_TYPE -4, QS_ERR_PARAM -7)."""
    import ctypes as C
    import orkan_amd
    from orkan_amd import _ffi
    lib = orkan_amd.load(5)
    fake = C.c_void_p(0x1000)
    st = _ffi.StateT(0, fake.value, 3)
    obs = (C.c_double * 8)()
    noop = _ffi.QOP(lambda s: None)
    ops = (_ffi.QOP * 2)(noop, noop)
    w = (C.c_double * 4)(1, 0, 1, 0)
    ch = (_ffi.AnsatzChannelT * 2)(_ffi.AnsatzChannelT(2, ops, C.addressof(w)), _ffi.AnsatzChannelT(2, ops, None))
    S = (C.c_double * 8)()
    big = 1 << 20
    call = lambda s, o, c, L, ls, b, out: lib.sample_matrix(s, o, c, L, ls, b, out)
    assert call(None, obs, ch, 2, 1, big, S) == -1
    assert call(C.byref(st), None, ch, 2, 1, big, S) == -1
    assert call(C.byref(st), obs, ch, 2, 1, big, None) == -1
    st_mixed = _ffi.StateT(1, fake.value, 3)
    assert call(C.byref(st_mixed), obs, ch, 2, 1, big, S) == -4
    assert call(C.byref(st), obs, ch, 2, 2, big, S) == -7            # lstar out of range
    assert call(C.byref(st), obs, ch, 2, 0, big, S) == -7            # channel 1 has no weights
    assert call(C.byref(st), obs, ch, 2, 1, 3 * 16 * 8, S) == -7     # workspace below 4 vectors
    empty = (_ffi.AnsatzChannelT * 1)(_ffi.AnsatzChannelT(0, ops, None))
    assert call(C.byref(st), obs, empty, 1, 0, big, S) == -7


def test_launch_shape_hook_ranges_without_device():
    """The tuning hook only stores host-side settings: valid ids and values are taken, every value
    outside the documented ranges is refused, and kernel 14 restores the defaults."""
    import orkan_amd
    lib = orkan_amd.load(5)
    ok = [(0, 3, 0), (7, 4 | 16, 49152), (9, 3 | 8, 81920), (10, 0, 81920), (10, 9, 81920), (10, 1, 0), (10, 16 | 9, 0), (11, 1, 0), (11, 2, 0), (15, 5, 0),
          (16, 4 | 8, 49152), (17, 3, 0), (17, 0, 0), (18, 32 | 16 | 2, 81920), (19, 2, 49152), (19, 0, 0), (19, 4 | 2, 32768), (20, 0, 53248),
          (12, 16 | 2 | 4, 53248), (12, 0, 0), (13, 15, 0), (13, 0, 1), (14, 0, 0)]
    ok += [(23, 0, 0), (24, 63, 0), (23, 2, 0), (25, 1, 0), (9, 32 | 16 | 2, 81920), (7, 32 | 4, 49152), (5, 64 | 32 | 4, 0), (4, 64, 65536), (22, 2, 0), (22, 33, 49152), (2, 64 | 3, 0),
           (21, 1, 0), (21, 255, 81920), (22, 128 | 36, 0), (22, 128, 0)]
    bad = [(-1, 0, 0), (21, 0, 0), (22, 256, 0), (22, 128 | 24, 0), (3, 128, 0), (26, 0, 0), (25, 64, 0), (23, -1, 0), (24, 64, 0), (23, 0, 4096), (21, 256, 0), (21, 4, 0), (20, 0, -1), (18, 24, 0), (19, 3, 0), (19, 7, 0), (19, 8, 0), (19, 0, -1), (17, 4, 0), (9, 128, 0), (7, 24, 0), (0, -1, 0), (0, 128, 0), (7, 24, 0), (0, 0, -1), (0, 0, 160 * 1024 + 1),
           (10, 0, -1), (10, 10, 0), (10, 26, 0), (10, 32, 0), (10, -1, 0), (11, 3, 0), (11, -1, 0), (12, 32, 0), (12, 16 | 12, 0), (13, 16, 0), (13, 0, -1)]
    try:
        for k, f, l in ok:
            assert lib.orkan_set_launch_shape(k, f, l) == 0, (k, f, l)
        for k, f, l in bad:
            assert lib.orkan_set_launch_shape(k, f, l) != 0, (k, f, l)
    finally:
        assert lib.orkan_set_launch_shape(14, 0, 0) == 0
