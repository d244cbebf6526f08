"""The C-ABI library: it loads without a GPU, exports every function the
headers in include/ declare, and fails fast with the reference's messages on
invalid arguments (validation happens before any device call)."""
import os
import re
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for f in os.listdir(INC):
        if not f.endswith(".h"):
            continue
        text = open(os.path.join(INC, f)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"//[^\n]*", "", text)
        # expand the backend macro families
        for fam, suffixes in (("ORKAN_BACKEND_1Q", ("_pure", "_packed", "_tiled")),
                              ("ORKAN_BACKEND_ROT", ("_pure", "_packed", "_tiled")),
                              ("ORKAN_BACKEND_2Q", ("_pure", "_packed", "_tiled"))):
            for m in re.finditer(fam + r"\((\w+)\)", text):
                if "#define" in text[max(0, m.start() - 10):m.start()]:
                    continue
                for s in suffixes:
                    names.add(m.group(1) + s)
        text = re.sub(r"#define(?:[^\n]*\\\n)*[^\n]*", "", text)
        for m in re.finditer(r"\b([A-Za-z_]\w*)\s*\(([^;{}()]*)\)\s*;", text):
            name = m.group(1)
            if name in ("if", "while", "sizeof", "return") or name.isupper():
                continue
            names.add(name)
    return names


def exported(path):
    out = subprocess.run(["nm", "-D", "--defined-only", path], stdout=subprocess.PIPE, text=True, check=True).stdout
    return {ln.split()[-1] for ln in out.splitlines() if " T " in ln}


@pytest.mark.parametrize("log_tile", [5, 3])
def test_library_exports_header_symbols(log_tile):
    import orkan_amd
    path = orkan_amd.lib_path(log_tile)
    assert os.path.exists(path), "build the library first (python -m orkan_amd.build)"
    want = declared_functions()
    assert {"state_plus", "h", "ccx", "channel_1q", "mean", "exp_diag", "two_from_mat",
            "x_tiled", "swap_packed", "orkan_state_upload"} <= want
    missing = sorted(want - exported(path))
    assert not missing, missing
    lib = orkan_amd.load(log_tile)     # loads with no GPU present
    assert lib.has("state_len")


def test_state_len_without_device():
    """state_len is pure arithmetic on (type, qubits) (src/state/state.c:43-57)."""
    import ctypes as C
    import orkan_amd
    from linalg_np import state_len
    for lt in (5, 3):
        lib = orkan_amd.load(lt)
        for kind in (0, 1, 2):
            for n in range(0, 15):
                st = orkan_amd.StateT(kind, None, n)
                assert lib.state_len(C.byref(st)) == state_len(kind, n, lt)
                st32 = orkan_amd.StateT(kind | orkan_amd.ORKAN_FP32, None, n)
                assert lib.state_len(C.byref(st32)) == state_len(kind, n, lt)


DEATH = r'''
import ctypes as C, sys
sys.path.insert(0, {root!r})
import orkan_amd
from orkan_amd import _ffi
lib = orkan_amd.load(5)
fake = C.c_void_p(0x1000)
st = _ffi.StateT({kind}, fake.value, {n})
obs = (C.c_double * 16)()
{call}
print("returned")
'''

CASES = [
    (0, 2, "lib.h(C.byref(st), 5)", "orkan: gate: h: target qubit out of range"),
    (0, 2, "lib.rx(C.byref(st), 2, 0.5)", "orkan: gate: rx: target qubit out of range"),
    (1, 3, "lib.cx(C.byref(st), 1, 1)", "orkan: gate: cx: control and target must differ"),
    (2, 3, "lib.cz(C.byref(st), 7, 1)", "orkan: gate: cz: control qubit out of range"),
    (0, 3, "lib.swap_gate(C.byref(st), 0, 9)", "orkan: gate: swap: q2 qubit out of range"),
    (0, 3, "lib.ccx(C.byref(st), 0, 0, 1)", "orkan: gate: ccx: ctrl1 and ctrl2 must differ"),
    (0, 3, "lib.ccx(C.byref(st), 0, 1, 1)", "orkan: gate: ccx: ctrl2 and target must differ"),
    (0, 3, "st.data = None; lib.x(C.byref(st), 0)", "orkan: gate: x: null state or data pointer"),
    (0, 3, "lib.h(None, 0)", "orkan: gate: h: null state or data pointer"),
    (0, 2, "so=_ffi.SuperopT(1, C.addressof(obs)); lib.channel_1q(C.byref(st), C.byref(so), 0)",
     "orkan: channel: channel_1q: not available for pure states"),
    (1, 2, "so=_ffi.SuperopT(2, C.addressof(obs)); lib.channel_1q(C.byref(st), C.byref(so), 0)",
     "orkan: channel: channel_1q: superoperator is not 1-local"),
    (1, 2, "so=_ffi.SuperopT(1, None); lib.channel_1q(C.byref(st), C.byref(so), 0)",
     "orkan: channel: channel_1q: null superoperator or data pointer"),
    (0, 3, "so=_ffi.SuperopT(2, C.addressof(obs)); lib.channel_2q(C.byref(st), C.byref(so), 0, 1)",
     "orkan: channel: channel_2q: not available for pure states"),
    (1, 3, "so=_ffi.SuperopT(1, C.addressof(obs)); lib.channel_2q(C.byref(st), C.byref(so), 0, 1)",
     "orkan: channel: channel_2q: superoperator is not 2-local"),
    (2, 3, "so=_ffi.SuperopT(2, C.addressof(obs)); lib.channel_2q(C.byref(st), C.byref(so), 1, 1)",
     "orkan: channel: channel_2q: qubits must differ"),
    (2, 3, "so=_ffi.SuperopT(2, C.addressof(obs)); lib.channel_2q(C.byref(st), C.byref(so), 0, 5)",
     "orkan: channel: channel_2q: qubit out of range"),
    (1, 2, "lib.mean(C.byref(st), None)", "orkan: meas: mean: null observable pointer"),
    (1, 2, "lib.matel_diag(C.byref(st), C.byref(st), obs)", "orkan: meas: matel_diag: bra must be PURE (mixed not supported)"),
    (0, 2, "st2=_ffi.StateT(0, fake.value, 3); lib.matel_diag(C.byref(st), C.byref(st2), obs)",
     "orkan: meas: matel_diag: qubit count mismatch"),
    (0, 2, "lib.exp_diag(C.byref(st), None, 1.0)", "orkan: circ: exp_diag: null diagonal pointer"),
    (7, 2, "lib.h(C.byref(st), 0)", "orkan: gate: h: unknown state type"),
    # the internal tiled backends the reference exports; here validated before any device call
    (0, 3, "lib.single_from_mat(C.byref(st), 0, C.addressof(obs))",
     "orkan: gate: single_from_mat: state must be MIXED_PACKED or MIXED_TILED"),
    (2, 3, "lib.single_from_mat(C.byref(st), 3, C.addressof(obs))",
     "orkan: gate: single_from_mat: target qubit out of range"),
    (1, 3, "lib.single_from_mat(C.byref(st), 0, None)", "orkan: gate: single_from_mat: null matrix pointer"),
    (2, 3, "st.data = None; lib.single_from_mat(C.byref(st), 0, C.addressof(obs))",
     "orkan: gate: single_from_mat: null state or data pointer"),
    (0, 3, "lib.two_from_mat(C.byref(st), 0, 1, C.addressof(obs))",
     "orkan: gate: two_from_mat: state must be MIXED_PACKED or MIXED_TILED"),
    (2, 3, "lib.two_from_mat(C.byref(st), 0, 3, C.addressof(obs))", "orkan: gate: two_from_mat: q2 qubit out of range"),
    (1, 3, "lib.two_from_mat(C.byref(st), 4, 0, C.addressof(obs))", "orkan: gate: two_from_mat: q1 qubit out of range"),
    (2, 3, "lib.two_from_mat(C.byref(st), 1, 1, C.addressof(obs))", "orkan: gate: two_from_mat: q1 and q2 must differ"),
    (1, 3, "lib.two_from_mat(C.byref(st), 0, 1, None)", "orkan: gate: two_from_mat: null matrix pointer"),
]


@pytest.mark.parametrize("kind,n,call,msg", CASES)
def test_validation_exits_like_reference(kind, n, call, msg):
    code = DEATH.format(root=ROOT, kind=kind, n=n, call=call)
    r = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 1, (r.returncode, r.stdout, r.stderr)
    assert msg in r.stderr
    assert "returned" not in r.stdout


def test_exp_diag_zero_t_is_a_no_op_without_device():
    code = DEATH.format(root=ROOT, kind=0, n=2, call="lib.exp_diag(C.byref(st), obs, 0.0)")
    r = subprocess.run([sys.executable, "-c", code], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True,
                       timeout=120)
    assert r.returncode == 0 and "returned" in r.stdout
