"""Permutation gates move amplitudes exactly, so every kernel that can run one gives bit-identical
results.  Statevector: the lane-exchange kernel (every gate qubit below 6), the paired complex64
kernel (no gate qubit at bit 0, two amplitudes per 16 B access) and the per-group kernel, under every
launch shape of tuning hook 21, with nontemporal accesses forced on through hook 13.  Density matrix:
x and y run through the streamed superoperator kernels (their superoperators hold 0, +-1, +-i only),
cx / cy / ccx through the same kernels on their target as a controlled permutation of each quad
(ccx with every qubit at or above the tile excepted), swap and that ccx through the whole-tile or
in-tile permutation kernels where their qubits allow, swap across the tile edge through virtual tiles
below the tile, everything else through the tile-group kernel,
which leaves unchanged entries unread and unwritten.  Each path is checked against the oracle and
bit for bit against the tile-group kernel (hooks 8 and 22 switch the others off)."""
import ctypes as C

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, random_psi, random_rho, to_layout

pytestmark = pytest.mark.gpu
NT_HOOK, RESET_HOOK, STREAM_HOOK, PERM_HOOK, PERMT_HOOK = 13, 14, 8, 21, 22
GROUP_PACKED_HOOK, GROUP_TILED_HOOK = 2, 3

GATES = [("cx", (0, 1)), ("cx", (1, 0)), ("cx", (1, 2)), ("cx", (2, 3)), ("cx", (3, 4)), ("cx", (5, 0)),
         ("cx", (0, 9)), ("cx", (9, 10)), ("cy", (0, 1)), ("cy", (2, 1)), ("cy", (4, 7)), ("swap_gate", (0, 1)),
         ("swap_gate", (1, 5)), ("swap_gate", (6, 8)), ("ccx", (0, 1, 2)), ("ccx", (2, 0, 1)), ("ccx", (1, 3, 4)),
         ("ccx", (5, 9, 2)), ("ccx", (7, 8, 9))]


@pytest.fixture(scope="module")
def eng():
    e = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    yield e
    assert e.api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0


def _sv(eng, n, psi, fp32):
    api = eng.api
    st = eng.new(PURE, n, psi)
    if fp32:
        s32 = api.orkan_state_convert(C.byref(st), 1)
        eng.free(st)
        st = s32
    for name, qs in GATES:
        getattr(api, name)(C.byref(st), *qs)
    out = np.empty(1 << n, np.complex64 if fp32 else np.complex128)
    assert api.orkan_state_download(C.byref(st), out.ctypes.data) == 0
    eng.free(st)
    return out


@pytest.mark.parametrize("fp32", [False, True], ids=["c128", "c64"])
def test_sv_permutation_kernels_bit_identical(eng, fp32):
    api = eng.api
    n = 12
    psi = random_psi(n, np.random.default_rng(31))
    f32 = 64 if fp32 else 0
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    # the per-group kernels only: lane kernel and paired kernel off
    assert api.orkan_set_launch_shape(PERM_HOOK, 1 | 8 | 32 | f32, 0) == 0
    want = _sv(eng, n, psi, fp32)
    # the oracle agrees with the per-group kernels
    orc = engines.oracle(5)
    so = orc.new(PURE, n, psi)
    for name, qs in GATES:
        orc.gate(so, name, *qs)
    ref = orc.read(so)
    orc.free(so)
    assert float(np.abs(want.astype(np.complex128) - (ref.astype(np.complex64) if fp32 else ref)).max()) <= 1e-12
    shapes = []
    for lu in (1, 2, 3):
        for wl in (0, 4):
            for extra in (0, 16):
                shapes.append((lu | wl | extra | f32, 0))
    shapes += [(2 | 4 | f32, 40960), (2 | 4 | 8 | f32, 0)]
    if fp32:
        shapes += [(3 | 4 | 64 | 128, 26624), (3 | 4 | 64 | 32, 0)]
    for nt_min in (0, 1):   # 1: nontemporal accesses on every state size
        assert api.orkan_set_launch_shape(NT_HOOK, 0, nt_min) == 0
        for flags, lds in shapes:
            assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
            assert api.orkan_set_launch_shape(NT_HOOK, 0, nt_min) == 0
            assert api.orkan_set_launch_shape(PERM_HOOK, flags, lds) == 0
            got = _sv(eng, n, psi, fp32)
            assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), (flags, lds, nt_min)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
@pytest.mark.parametrize("log_tile", [5, 3])
def test_dm_permutations_against_oracle(kind, log_tile):
    """Every permutation gate on many positions against the oracle through the default dispatch, and
    the same sequence with every gate sent to the tile-group kernel (hooks 8 and 22), its loads one
    after the other (the default) and all in flight at once (group shape flag bit 6): all three
    bit-identical."""
    n = 8
    dev = engines.DeviceEngine(orkan_amd.lib_path(log_tile), log_tile)
    api = dev.api
    orc = engines.oracle(log_tile)
    rng = np.random.default_rng(77 + kind + log_tile)
    rho = to_layout(random_rho(n, rng), kind, log_tile)
    ops = [("x", (q,)) for q in range(n)] + [("y", (q,)) for q in range(n)]
    ops += [(g, (a, b)) for g in ("cx", "cy", "swap_gate") for a in range(n) for b in range(n) if a != b and (a + 2 * b) % 3 == 0]
    ops += [("ccx", (a, b, c)) for (a, b, c) in ((0, 1, 2), (2, 1, 0), (7, 0, 4), (3, 6, 5), (5, 7, 6), (1, 4, 7))]
    so = orc.new(kind, n, rho)
    for name, qs in ops:
        orc.gate(so, name, *qs)
    want = orc.read(so)
    orc.free(so)
    outs = []
    for hoist in (None, 0, 64):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        if hoist is not None:
            assert api.orkan_set_launch_shape(STREAM_HOOK, 0, 0) == 0
            assert api.orkan_set_launch_shape(PERMT_HOOK, 32 | 2, 0) == 0
            if kind == MIXED_PACKED:
                assert api.orkan_set_launch_shape(GROUP_PACKED_HOOK, 1 | 2 | hoist, 0) == 0
            else:
                assert api.orkan_set_launch_shape(GROUP_TILED_HOOK, 1 | hoist, 0) == 0
        sd = dev.new(kind, n, rho)
        for name, qs in ops:
            dev.gate(sd, name, *qs)
        outs.append(dev.read(sd))
        dev.free(sd)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert float(np.abs(outs[0] - want).max()) <= 1e-12
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[1], outs[2])


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
def test_dm_x_y_streamed_equal_tile_groups(eng, kind):
    api = eng.api
    n = 9
    rho = to_layout(random_rho(n, np.random.default_rng(5)), kind, 5)
    outs = []
    for streamed in (1, 0):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(STREAM_HOOK, 0, streamed) == 0
        st = eng.new(kind, n, rho)
        for q in range(n):
            api.x(C.byref(st), q)
            api.y(C.byref(st), (q + 3) % n)
        outs.append(eng.read(st))
        eng.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    # equal entry by entry (the superoperator's zero products may only flip the sign of a zero)
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
@pytest.mark.parametrize("log_tile", [5, 3])
def test_dm_tile_permutations_equal_tile_groups(kind, log_tile):
    """Gates whose qubits all lie at or above the tile (32 x 32 blocks for packed) move whole tiles
    (dm_perm_tiles_kernel; packed cx / cy / ccx stream unless hook 22 flag bit 6 is set); hook 22 with
    flag bit 1 sends them to the tile-group kernel instead.  All move entries exactly: bit-identical,
    and equal to the oracle."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(log_tile), log_tile)
    api = dev.api
    orc = engines.oracle(log_tile)
    T = log_tile if kind == MIXED_TILED else 5
    n = T + 3
    hi = list(range(T, n))
    ops = [(g, (a, b)) for g in ("cx", "cy", "swap_gate") for a in hi for b in hi if a != b]
    ops += [("ccx", (a, b, c)) for a in hi for b in hi for c in hi if len({a, b, c}) == 3]
    rho = to_layout(random_rho(n, np.random.default_rng(90 + kind + log_tile)), kind, log_tile)
    outs = []
    for flags in (0, 64, 2):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(PERMT_HOOK, 32 | flags, 0) == 0
        st = dev.new(kind, n, rho)
        for name, qs in ops:
            dev.gate(st, name, *qs)
        outs.append(dev.read(st))
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    so = orc.new(kind, n, rho)
    for name, qs in ops:
        orc.gate(so, name, *qs)
    assert float(np.abs(outs[0] - orc.read(so)).max()) <= 1e-12
    orc.free(so)


@pytest.mark.parametrize("log_tile", [5, 3])
def test_dm_tiled_streaming_accesses_change_no_value(log_tile):
    """Nontemporal accesses of the tile-group kernel (tiled layout, virtual tiles of whole lines) and of
    the whole-tile permutations change the cache policy only: with hook 13 lowering the size limit to one
    byte they are used on a 9-qubit state, and the results equal those with them switched off (bit 2 of
    the kernel 3 and 22 shapes) bit for bit."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(log_tile), log_tile)
    api = dev.api
    n = 9
    rng = np.random.default_rng(123 + log_tile)
    rho = to_layout(random_rho(n, rng), MIXED_TILED, log_tile)
    U = np.linalg.qr(rng.standard_normal((4, 4)) + 1j * rng.standard_normal((4, 4)))[0]
    Um = np.ascontiguousarray(U.T.reshape(-1), np.complex128)   # column-major, as two_from_mat takes it
    ops = [("cx", (0, 1)), ("cx", (7, 8)), ("ccx", (6, 7, 8)), ("swap_gate", (2, 7)), ("cy", (8, 3)),
           ("ccx", (0, 5, 8)), ("swap_gate", (5, 6))]
    outs = []
    for nt in (4, 0):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(NT_HOOK, 0, 1) == 0
        assert api.orkan_set_launch_shape(GROUP_TILED_HOOK, 1 | nt, 0) == 0
        assert api.orkan_set_launch_shape(PERMT_HOOK, 32 | nt, 0) == 0
        st = dev.new(MIXED_TILED, n, rho)
        for name, qs in ops:
            dev.gate(st, name, *qs)
        for q1, q2 in ((0, 6), (7, 2), (4, 5)):
            api.two_from_mat(C.byref(st), q1, q2, Um.ctypes.data)
        outs.append(dev.read(st))
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
@pytest.mark.parametrize("n", [5, 7])
def test_dm_intile_permutations_equal_tile_groups(kind, n):
    """Gates whose qubits all lie below 5 permute each 32 x 32 tile (packed: block) within itself
    (dm_perm_intile_kernel); hook 22 with flag bit 1 sends them to the tile-group kernel instead.  Both
    move entries exactly: bit-identical, and equal to the oracle (n = 5 is a single tile)."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    api = dev.api
    orc = engines.oracle(5)
    lo = list(range(5))
    ops = [(g, (a, b)) for g in ("cx", "cy", "swap_gate") for a in lo for b in lo if a != b]
    ops += [("ccx", (a, b, c)) for a in lo for b in lo for c in lo if len({a, b, c}) == 3 and (a + b + c) % 2 == 0]
    rho = to_layout(random_rho(n, np.random.default_rng(200 + kind + n)), kind, 5)
    outs = []
    for off in (0, 2):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(PERMT_HOOK, 32 | 4 | off, 0) == 0
        st = dev.new(kind, n, rho)
        for name, qs in ops:
            dev.gate(st, name, *qs)
        outs.append(dev.read(st))
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert np.array_equal(outs[0], outs[1])
    so = orc.new(kind, n, rho)
    for name, qs in ops:
        orc.gate(so, name, *qs)
    assert float(np.abs(outs[0] - orc.read(so)).max()) <= 1e-12
    orc.free(so)


@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
@pytest.mark.parametrize("log_tile", [5, 3])
def test_dm_straddling_controlled_gates(kind, log_tile):
    """cx, cy and ccx whose qubits straddle the tile (32 x 32 blocks for packed) run through the streamed
    1-qubit kernels on their target as a controlled permutation of each quad; hook 8 (lds_bytes 0) sends
    them to the tile-group kernel instead.  Both move entries exactly: bit-identical, and equal to the
    oracle, with targets inside the tile (lane-exchange and quad kernels) and above it (tile quads)."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(log_tile), log_tile)
    api = dev.api
    orc = engines.oracle(log_tile)
    T = log_tile if kind == MIXED_TILED else 5
    n = T + 3
    ops = []
    for lo in range(T):
        for hi in range(T, n):
            ops += [("cx", (lo, hi)), ("cx", (hi, lo)), ("cy", (lo, hi)), ("cy", (hi, lo))]
    ops += [("ccx", (0, T, 1)), ("ccx", (T - 1, T + 1, T)), ("ccx", (T, 1, T + 2)), ("ccx", (0, T - 1, T + 1)),
            ("ccx", (T + 2, T, 0)), ("ccx", (1, 2, T + 1)), ("swap_gate", (T - 1, T))]
    rho = to_layout(random_rho(n, np.random.default_rng(300 + kind + log_tile)), kind, log_tile)
    outs = []
    for streamed in (1, 0):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(STREAM_HOOK, 0, streamed) == 0
        st = dev.new(kind, n, rho)
        for name, qs in ops:
            dev.gate(st, name, *qs)
        outs.append(dev.read(st))
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert np.array_equal(outs[0], outs[1])
    so = orc.new(kind, n, rho)
    for name, qs in ops:
        orc.gate(so, name, *qs)
    assert float(np.abs(outs[0] - orc.read(so)).max()) <= 1e-12
    orc.free(so)


SKIP_HOOKS = (23, 24, 25)   # packed cx / cy, tiled, packed ccx


@pytest.mark.parametrize("fp32", [False, True], ids=["c128", "c64"])
@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
@pytest.mark.parametrize("log_tile", [5, 3])
def test_dm_controlled_gates_skip_unchanged_quads(fp32, kind, log_tile):
    """Where every control qubit lies above the bits of a 128 B line, the streamed kernels leave the
    quads of cx / cy / ccx that keep their value unread and unwritten (hooks 23 / 24 / 25: the margin
    above the line bits for packed cx / cy, tiled, packed ccx; 63 never).  Skipping, never skipping (ccx with every qubit at or above the tile then
    moves whole tiles) and the permutation kernels (hook 8 off: whole-tile, in-tile and tile-group
    kernels) give bit-identical results on every
    ordered pair and a spread of triples, in complex128 and complex64 storage; complex128 also equals
    the oracle."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(log_tile), log_tile)
    api = dev.api
    n = 8 if (kind == MIXED_PACKED or log_tile == 5) else 7
    ops = [(g, (a, b)) for g in ("cx", "cy") for a in range(n) for b in range(n) if a != b]
    ops += [("ccx", t) for t in ((0, 1, 2), (3, 4, 5), (5, 6, 7 % n), (6, 4, 3), (4, 5, 0), (n - 1, n - 2, n - 3),
                                 (2, n - 1, 4), (n - 2, 3, n - 1))]
    rho = to_layout(random_rho(n, np.random.default_rng(400 + kind + log_tile + 7 * fp32)), kind, log_tile)
    outs = []
    for mode in ("skip", "never", "groups"):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        for h in SKIP_HOOKS:   # skip wherever the controls lie above the line bits, or never
            assert api.orkan_set_launch_shape(h, 63 if mode == "never" else 0, 0) == 0
        if mode == "groups":
            assert api.orkan_set_launch_shape(STREAM_HOOK, 0, 0) == 0
        st = dev.new(kind, n, rho)
        if fp32:
            s32 = api.orkan_state_convert(C.byref(st), 1)
            dev.free(st)
            st = s32
        for name, qs in ops:
            dev.gate(st, name, *qs)
        out = np.empty(dev.length(st), np.complex64 if fp32 else np.complex128)
        assert api.orkan_state_download(C.byref(st), out.ctypes.data) == 0
        outs.append(out)
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])
    if not fp32:
        orc = engines.oracle(log_tile)
        so = orc.new(kind, n, rho)
        for name, qs in ops:
            orc.gate(so, name, *qs)
        assert float(np.abs(outs[0] - orc.read(so)).max()) <= 1e-12
        orc.free(so)


@pytest.mark.parametrize("fp32", [False, True], ids=["c128", "c64"])
@pytest.mark.parametrize("kind", [MIXED_TILED, MIXED_PACKED], ids=["tiled", "packed"])
def test_dm_subtile_permutations(fp32, kind):
    """A gate whose lowest qubit V lies below the 32 x 32 tile and whose highest at or above it moves
    virtual tiles of 2^V x 2^V entries (V 3 or 4 for complex128, 4 for complex64): swap by default, and cx /
    cy / ccx once the streamed kernels are switched off (hook 8), cy with its phases.  On the tiled layout
    the upper virtual tiles of diagonal storage tiles are rewritten from their lower mirrors.  The results
    equal, bit for bit, those of the tile-group kernel (hook 22 flag bit 7 sends these gates there), with
    nontemporal accesses and without, with one and with 16 iterations per workgroup (hook 22 bits 3-4), and
    in complex128 the oracle."""
    dev = engines.DeviceEngine(orkan_amd.lib_path(5), 5)
    api = dev.api
    n = 10
    ops = []
    for a in ((4,) if fp32 else (3, 4)):
        for b in (5, 6, 9):
            c = 8 if b != 8 else 7
            ops += [("swap_gate", (a, b)), ("swap_gate", (b, a)), ("cx", (a, b)), ("cx", (b, a)), ("cy", (b, a)),
                    ("cy", (a, b)), ("ccx", (b, a, c)), ("ccx", (a, c, b)), ("ccx", (c, b, a))]
    rho = to_layout(random_rho(n, np.random.default_rng(600 + kind + 7 * fp32)), kind, 5)
    outs = []
    # virtual tiles with and without nontemporal accesses, 16 iterations per workgroup; tile groups
    for flags in (32 | 4, 32, 32 | 4 | 16, 32 | 4 | 128):
        assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(STREAM_HOOK, 0, 0) == 0
        assert api.orkan_set_launch_shape(PERMT_HOOK, flags, 0) == 0
        st = dev.new(kind, n, rho)
        if fp32:
            s32 = api.orkan_state_convert(C.byref(st), 1)
            dev.free(st)
            st = s32
        for name, qs in ops:
            dev.gate(st, name, *qs)
        out = np.empty(dev.length(st), np.complex64 if fp32 else np.complex128)
        assert api.orkan_state_download(C.byref(st), out.ctypes.data) == 0
        outs.append(out)
        dev.free(st)
    assert api.orkan_set_launch_shape(RESET_HOOK, 0, 0) == 0
    assert all(np.array_equal(outs[0], o) for o in outs[1:])
    if not fp32:
        orc = engines.oracle(5)
        so = orc.new(kind, n, rho)
        for name, qs in ops:
            orc.gate(so, name, *qs)
        assert float(np.abs(outs[0] - orc.read(so)).max()) <= 1e-12
        orc.free(so)
