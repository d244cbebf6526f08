"""Batched QAOA launcher against per-instance runs of the reference call
sequence on the CPU oracle (state_plus, exp_diag, rx(q, 2 beta), mean) at
1e-12, plus the device-side MaxCut diagonal against numpy."""
import numpy as np
import pytest

import engines
import orkan_amd
from linalg_np import PURE
from orkan_amd import batch

pytestmark = pytest.mark.gpu


def oracle_eval(orc, n, diag, p, gamma, beta):
    st = orc.plus(PURE, n)
    for l in range(p):
        orc.exp_diag(st, diag, gamma[l])
        for q in range(n):
            orc.gate(st, "rx", q, 2 * beta[l])
    m = orc.mean(st, diag)
    orc.free(st)
    return m


@pytest.mark.parametrize("config", [None, (11, 256), (12, 1024), (13, 512)])
def test_batch_matches_per_instance_runs(config):
    lib = orkan_amd.load(5)
    orc = engines.oracle(5)
    rng = np.random.default_rng(44)
    inst = batch.maxcut_instances(48, rng, n_lo=2, n_hi=18)
    diags = [batch.maxcut_diag(i["n"], i["edges"], i["w"]) for i in inst]
    b = batch.QaoaBatch(lib, [i["n"] for i in inst], diags, max_evals=64, max_p=8)
    if config:
        b.set_config(*config)
    got = b.eval(list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst],
                 [i["beta"] for i in inst])
    for k, i in enumerate(inst):
        want = oracle_eval(orc, i["n"], diags[k], i["p"], i["gamma"], i["beta"])
        assert abs(got[k] - want) <= 1e-12, (k, i["n"], i["p"], got[k], want)
    # repeated launches of the captured graph give identical results
    again = b.eval(list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst],
                   [i["beta"] for i in inst])
    assert np.array_equal(got, again)
    b.close()


def test_device_maxcut_diag():
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(3)
    for n in (3, 11, 17):
        i = batch.maxcut_instances(1, rng, n_lo=n, n_hi=n)[0]
        out = lib.orkan_dev_alloc(1 << n)
        e = np.ascontiguousarray(i["edges"])
        ei = np.ascontiguousarray(e[:, 0]); ej = np.ascontiguousarray(e[:, 1])
        w = np.ascontiguousarray(i["w"])
        rc = lib.orkan_maxcut_diag_dev(n, len(w), ei.ctypes.data_as(batch.IP), ej.ctypes.data_as(batch.IP),
                                       w.ctypes.data_as(batch.DP), out)
        assert rc == 0
        host = np.empty(1 << n)
        assert lib.orkan_copy_to_host(host.ctypes.data, out, host.nbytes) == 0
        assert np.abs(host - batch.maxcut_diag(n, i["edges"], i["w"])).max() <= 1e-12
        lib.orkan_dev_free(out)


def test_autotuner_picks_a_valid_config():
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(5)
    inst = batch.maxcut_instances(32, rng, n_lo=10, n_hi=16)
    diags = [batch.maxcut_diag(i["n"], i["edges"], i["w"]) for i in inst]
    b = batch.QaoaBatch(lib, [i["n"] for i in inst], diags, max_evals=32, max_p=8)
    args = (list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst], [i["beta"] for i in inst])
    ref = b.eval(*args)
    ms, cfg = b.autotune(*args, reps=2)
    assert ms > 0 and cfg["lt"] in (11, 12, 13)
    assert np.abs(b.eval(*args) - ref).max() <= 1e-12
    b.close()


def test_largest_states_n19_n20():
    """States of 19 and 20 qubits (the top of the 10..20 range of the 4096-instance workload) stream
    through HBM in p + 1 LDS-tiled sweeps; each against its per-instance oracle run."""
    lib = orkan_amd.load(5)
    orc = engines.oracle(5)
    rng = np.random.default_rng(1920)
    inst = batch.maxcut_instances(6, rng, n_lo=19, n_hi=20)
    diags = [batch.maxcut_diag(i["n"], i["edges"], i["w"]) for i in inst]
    b = batch.QaoaBatch(lib, [i["n"] for i in inst], diags, max_evals=8, max_p=8)
    got = b.eval(list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst],
                 [i["beta"] for i in inst])
    for k, i in enumerate(inst):
        print(f"n {i['n']} p {i['p']}", flush=True)
        want = oracle_eval(orc, i["n"], diags[k], i["p"], i["gamma"], i["beta"])
        assert abs(got[k] - want) <= 1e-12, (k, i["n"], i["p"], got[k], want)
    b.close()


def test_full_4096_batch_sample():
    """The 4096-instance workload itself (cost diagonals built on the device), autotuned, one
    launch; a seeded sample of its evaluations, always including the largest states, against
    per-instance oracle runs at 1e-12."""
    from benchmarks import qaoa as bq
    lib = orkan_amd.load(5)
    orc = engines.oracle(5)
    inst = bq.make(4096)
    dptrs = bq.device_diags(lib, inst)
    b = batch.QaoaBatch(lib, [i["n"] for i in inst], dptrs, max_evals=len(inst), max_p=8)
    for p in dptrs:
        lib.orkan_dev_free(p)
    args = (list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst], [i["beta"] for i in inst])
    _, cfg = b.autotune(*args, reps=1)
    print(f"autotuned {cfg}", flush=True)
    got = b.eval(*args)
    b.close()
    order = np.argsort([-(i["n"] * 8 + i["p"]) for i in inst], kind="stable")
    sample = list(order[:4]) + list(np.random.default_rng(4096).choice(len(inst), 20, replace=False))
    for k in sample:
        i = inst[k]
        want = oracle_eval(orc, i["n"], batch.maxcut_diag(i["n"], i["edges"], i["w"]), i["p"], i["gamma"], i["beta"])
        assert abs(got[k] - want) <= 1e-12, (k, i["n"], i["p"], got[k], want)
