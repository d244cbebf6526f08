"""A C program written in the reference's usage style (tests/c_caller/quickstart.c: designated
initialisers, cplx_t returned by value from state_get / matel_diag, a caller buffer handed to
state_init, kraus_to_superop memory freed by the caller), compiled once per library that exports
the API (tests/c_caller/Makefile, run by the build).  The drop-in claim is that the same source,
linked against this repository's HIP library, prints what it prints linked against the oracle and
the compiled reference, to the reference's tolerance 1e-12 (test/include/test.h:40)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
BIN = os.path.join(HERE, "c_caller")
TOL = 1e-12


def run(name):
    path = os.path.join(BIN, name)
    if not os.path.exists(path):
        pytest.skip(f"{name} not built (make -C tests/c_caller)")
    print(f"running {name}", flush=True)
    r = subprocess.run([path], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = {}
    for line in r.stdout.splitlines():
        f = line.split()
        out[f[0]] = f[1:]
    return out


def compare(a, b):
    assert a.keys() == b.keys()
    worst = 0.0
    for k in a:
        try:
            va = [float(x) for x in a[k]]
            vb = [float(x) for x in b[k]]
        except ValueError:
            assert a[k] == b[k], k
            continue
        assert len(va) == len(vb), k
        worst = max([worst] + [abs(x - y) for x, y in zip(va, vb)])
    return worst


def test_quickstart_oracle_matches_reference():
    o = run("quickstart_oracle")
    assert o["buf_after_init"] == ["null"] and o["tiled_freed"] == ["null", "0"] and "done" in o
    # the quick start's amplitudes: h(0) takes qubit 0 of |+>^3 to |0>, so cx(0, 1) does nothing and
    # rz(2, 0.5) leaves 0.5 e^(-+i/4) on the even indices
    for i in range(8):
        re, im = (float(x) for x in o[f"amp[{i}]"])
        if i & 1:
            assert abs(re) + abs(im) <= TOL
        else:
            sign = -1 if i < 4 else 1
            assert abs(complex(re, im) - 0.5 * complex(0.96891242171064473, sign * 0.24740395925452294)) <= TOL
    ref = os.path.join(BIN, "quickstart_ref")
    if os.path.exists(ref):
        assert compare(o, run("quickstart_ref")) <= TOL


@pytest.mark.gpu
def test_quickstart_on_the_gpu_library():
    g, o = run("quickstart_orkan"), run("quickstart_oracle")
    assert g["buf_after_init"] == ["null"] and "done" in g
    assert compare(g, o) <= TOL
