"""Batched QAOA workload: 4096 MaxCut instances on G(n, 0.5) graphs, n ~ U{10..20},
weights ~ U[0, 1], p ~ U{1..8}, gamma, beta ~ U[0, pi]; one evaluation is
state_plus, p x (exp_diag(C, gamma), rx(q, 2 beta) on every qubit), mean(C).
All 4096 evaluations run in one persistent, graph-replayed launch whose tile
shape the autotuner picks.  Reported: evaluations per second (cost diagonals
already resident), the autotuned configuration, and a CPU sample of the same
evaluations through the reference library with the largest deviation from the
device results.
"""
import ctypes as C
import os
import sys
import time

import numpy as np

from orkan_amd import batch
from . import workloads as wl


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make(count=4096, seed_cfg=4):
    g = wl.rng(seed_cfg)
    return batch.maxcut_instances(count, g)


def device_diags(lib, inst):
    ptrs = []
    for i in inst:
        n = i["n"]
        p = lib.orkan_dev_alloc(1 << n)
        e = np.ascontiguousarray(i["edges"])
        ei, ej = np.ascontiguousarray(e[:, 0]), np.ascontiguousarray(e[:, 1])
        w = np.ascontiguousarray(i["w"])
        rc = lib.orkan_maxcut_diag_dev(n, len(w), ei.ctypes.data_as(batch.IP), ej.ctypes.data_as(batch.IP),
                                       w.ctypes.data_as(batch.DP), p)
        assert rc == 0
        ptrs.append(p)
    return ptrs


def run(lib, root, reps=5, cpu_budget_s=15.0, count=4096):
    inst = make(count)
    log(f"qaoa: {len(inst)} instances generated")
    dptrs = device_diags(lib, inst)
    nq = [i["n"] for i in inst]
    b = batch.QaoaBatch(lib, nq, dptrs, max_evals=len(inst), max_p=8)
    for p in dptrs:
        lib.orkan_dev_free(p)
    args = (list(range(len(inst))), [i["p"] for i in inst], [i["gamma"] for i in inst], [i["beta"] for i in inst])
    tune_ms, cfg = b.autotune(*args, reps=2)
    log(f"qaoa: autotuned {cfg} ({tune_ms:.2f} ms)")
    times, out = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        out = b.eval(*args)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    evals_s = len(inst) / t
    # bytes the p + 1 sweeps of an HBM-resident state move, per amplitude: the first sweep
    # reads the cost diagonal and writes the state (8 + 16), the p - 1 middle ones read and
    # write the state and read the diagonal (40), the last reads both (24): 40 p + 8.
    # LDS-resident states move nothing.
    moved = 0
    for i in inst:
        if i["n"] > cfg["lt"]:
            moved += (40 * i["p"] + 8) * (1 << i["n"])
    res = {"instances": len(inst), "evals_per_s": round(evals_s, 1), "batch_ms": round(t * 1e3, 3),
           "autotuned": cfg, "hbm_GBs_sweeps": round(moved / t / 1e9, 1)}
    # CPU sample through the reference library: same evaluations, per-instance call sequence
    ref_path = os.path.join(root, "oracle", "_ref", "liborkan_ref_t5.so")
    kind = "reference"
    if not os.path.exists(ref_path):
        ref_path, kind = os.path.join(root, "oracle", "lib", "liborkan_oracle_t5.so"), "port"
    if os.path.exists(ref_path) and cpu_budget_s > 0:
        from orkan_amd import _ffi
        ref = _ffi.bind(ref_path)
        t0 = time.perf_counter()
        done, worst = 0, 0.0
        order = np.random.default_rng(0).permutation(len(inst))
        for k in order:
            i = inst[k]
            diag = batch.maxcut_diag(i["n"], i["edges"], i["w"])
            st = _ffi.StateT(0, None, 0)
            ref.state_plus(C.byref(st), i["n"])
            for l in range(i["p"]):
                ref.exp_diag(C.byref(st), _ffi.dptr(diag), float(i["gamma"][l]))
                for q in range(i["n"]):
                    ref.rx(C.byref(st), q, float(2 * i["beta"][l]))
            m = ref.mean(C.byref(st), _ffi.dptr(diag))
            ref.state_free(C.byref(st))
            worst = max(worst, abs(m - out[k]))
            done += 1
            if time.perf_counter() - t0 > cpu_budget_s:
                break
        dt = time.perf_counter() - t0
        cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        res["cpu_baseline"] = {"value": round(done / dt, 2), "unit": "evals/s", "cores": cores, "kind": kind,
                               "sample": f"{done} random instances of the 4096, per-instance reference call "
                                         f"sequence, cost diagonal built on the host per instance, {dt:.1f} s"}
        res["max_abs_err_vs_cpu"] = worst
    b.close()
    return res
