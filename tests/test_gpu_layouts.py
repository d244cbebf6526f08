"""Layout conversion on the device and state files (extensions without a reference
counterpart).  Conversions move or conjugate entries, so they are checked bit for bit
against numpy relayouts of the same logical matrix (the projector of a pure state to
rounding, one product per entry); files round-trip bit for bit in both precisions, and
rejected files leave the state unchanged."""
import ctypes as C
import os

import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import MIXED_PACKED, MIXED_TILED, PURE, from_layout, random_psi, random_rho, to_layout

pytestmark = pytest.mark.gpu
F32 = orkan_amd.ORKAN_FP32


def _eng(lt):
    return engines.DeviceEngine(orkan_amd.lib_path(lt), lt)


def _src(eng, kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == PURE:
        psi = random_psi(n, rng)
        return eng.new(PURE, n, psi), np.outer(psi, psi.conj())
    rho = random_rho(n, rng)
    return eng.new(kind, n, to_layout(rho, kind, eng.log_tile)), rho


@pytest.mark.parametrize("lt", [5, 3])
@pytest.mark.parametrize("src_kind", [PURE, MIXED_PACKED, MIXED_TILED])
@pytest.mark.parametrize("dst_kind", [MIXED_PACKED, MIXED_TILED])
def test_relayout_matches_numpy(lt, src_kind, dst_kind):
    eng = _eng(lt)
    for n in sorted({1, 2, lt - 1, lt, lt + 1, lt + 2}):
        st, rho = _src(eng, src_kind, n, 10 * n + src_kind)
        out = orkan_amd.StateT(0, None, 0)
        assert eng.api.orkan_state_relayout(C.byref(st), C.byref(out), dst_kind) == 0
        assert out.type == dst_kind and out.qubits == n
        got = eng.read(out)
        # a packed store holds the lower triangle only: its logical matrix is lower + conj(lower)^T,
        # which differs from a numerically Hermitian rho in the last bit of the upper entries
        lower = np.tril(rho) + np.tril(rho, -1).conj().T
        want = to_layout(lower if src_kind == MIXED_PACKED else rho, dst_kind, lt)
        if src_kind == PURE:
            assert np.abs(got - want).max() <= 1e-15, n
        else:
            assert np.array_equal(got, want), n
        # and back to the source layout
        if src_kind != PURE:
            back = orkan_amd.StateT(0, None, 0)
            assert eng.api.orkan_state_relayout(C.byref(out), C.byref(back), src_kind) == 0
            keep = rho if (src_kind == MIXED_TILED and dst_kind == MIXED_TILED) else lower
            assert np.array_equal(eng.read(back), to_layout(keep, src_kind, lt)), n
            eng.free(back)
        eng.free(out)
        eng.free(st)


def test_relayout_fp32_and_errors():
    eng = _eng(5)
    n = 7
    st, rho = _src(eng, MIXED_TILED, n, 3)
    st32 = eng.api.orkan_state_convert(C.byref(st), 1)
    out = orkan_amd.StateT(0, None, 0)
    assert eng.api.orkan_state_relayout(C.byref(st32), C.byref(out), MIXED_PACKED) == 0
    assert out.type == MIXED_PACKED | F32
    back = eng.api.orkan_state_convert(C.byref(out), 0)
    s64 = eng.api.orkan_state_convert(C.byref(st32), 0)
    want = to_layout(from_layout(eng.read(s64), MIXED_TILED, n, 5), MIXED_PACKED, 5)
    assert np.array_equal(eng.read(back), want)
    eng.free(s64)
    bad = orkan_amd.StateT(0, None, 0)
    assert eng.api.orkan_state_relayout(C.byref(st), C.byref(bad), PURE) == -4
    assert eng.api.orkan_state_relayout(None, C.byref(bad), MIXED_PACKED) == -1
    for s in (st, st32, out, back):
        eng.free(s)


@pytest.mark.parametrize("kind", [PURE, MIXED_PACKED, MIXED_TILED])
@pytest.mark.parametrize("fp32", [False, True])
def test_state_file_round_trip(tmp_path, kind, fp32):
    eng = _eng(5)
    for n in (1, 4, 7):
        st, _ = _src(eng, kind, n, 5 + n)
        if fp32:
            s32 = eng.api.orkan_state_convert(C.byref(st), 1)
            eng.free(st)
            st = s32
        raw = np.empty(eng.length(st) * (8 if fp32 else 16), np.uint8)
        assert eng.api.orkan_state_download(C.byref(st), raw.ctypes.data) == 0
        path = str(tmp_path / f"s{kind}_{n}_{int(fp32)}.orkan").encode()
        assert eng.api.orkan_state_save(C.byref(st), path) == 0
        assert os.path.getsize(path) == 64 + raw.size
        got = eng.plus(PURE, 2)                       # any valid state: load replaces it
        assert eng.api.orkan_state_load(C.byref(got), path) == 0
        assert got.type == st.type and got.qubits == st.qubits
        back = np.empty_like(raw)
        assert eng.api.orkan_state_download(C.byref(got), back.ctypes.data) == 0
        assert np.array_equal(raw, back)
        # a tiled file of another tile size is refused and the state is left alone
        if kind == MIXED_TILED:
            e3 = _eng(3)
            keep = e3.plus(PURE, 3)
            before = e3.read(keep)
            assert e3.api.orkan_state_load(C.byref(keep), path) == -6
            assert keep.type == PURE and keep.qubits == 3 and np.array_equal(e3.read(keep), before)
            e3.free(keep)
        eng.free(got)
        eng.free(st)
