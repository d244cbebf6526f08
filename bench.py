"""Benchmark of the state-update hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload sv28|dm14] [--no-cpu]

Default workload (sv28): the 28-qubit complex128 statevector brick circuit --
from |+>^28, each layer applies rx, ry, rz (seeded angles) on every qubit then
cx on alternating neighbour pairs; a step is one layer (84 rotations + 13 or 14
cx), the 40 steps of the default run are the whole 40-layer circuit, followed
by mean() of a seeded diagonal observable (reported separately).

`value` is whole-job HBM throughput: algorithmic bytes of all gates of the K
timed steps (each amplitude a gate can change read once and written once,
16 B per amplitude) divided by the time of those steps, with the state and
the observable resident in HBM.  `roofline` describes the dominant kernel (the
single-qubit pair update) from per-launch HIP events recorded on the library's
stream around every rotation of the timed steps.  `cpu_baseline` times the
reference library compiled from its own sources (oracle/_ref) on this host's
cores on a bounded sample: one full brick layer at the same 28-qubit size.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "HBM GB/s (% peak) per gate at 28q SV / 14q DM; batched circuit evals/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="sv28", choices=["sv28", "dm14", "qaoa", "surrogate"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp32", action="store_true", help="complex64 storage variant")
    ap.add_argument("--quick", action="store_true", help="statevector line only (no dm14 / qaoa blocks)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain copy of the state (2^n x 16 B read + written) for counter calibration")
    return ap.parse_args()


class Events:
    """HIP events on the library's stream (orkan_event_record / orkan_event_ms)."""

    def __init__(self, lib):
        self.lib = lib
        self.n = 0

    def rec(self):
        i = self.n
        self.n += 1
        rc = self.lib.orkan_event_record(i)
        assert rc == 0, rc
        return i

    def ms(self, a, b):
        return self.lib.orkan_event_ms(a, b)


def traffic_from_profiles(kernel_key):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE in separate counter runs
    (not in this timed run): profiles/traffic.json, written from the committed counter summaries.
    Returns (bytes, source) with the summary file the figure comes from."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None, None
    try:
        e = json.load(open(p)).get(kernel_key, {})
    except (OSError, ValueError):
        return None, None
    return e.get("bytes_per_launch"), e.get("source", "profiles/traffic.json")


# -------------------------------------------------------------------------------------------- sv28
def run_sv28(args, lib):
    from orkan_amd import ORKAN_FP32, StateT
    from benchmarks import workloads as wl

    n = wl.SV_QUBITS
    steps = args.steps if args.steps is not None else wl.SV_LAYERS
    layers, g = wl.brick_circuit(n, max(steps + args.warmup, 1))
    obs = wl.sv_observable(n, g)
    elem = 8 if args.fp32 else 16
    kind = 0 | (ORKAN_FP32 if args.fp32 else 0)

    st = StateT(kind, None, 0)
    lib.state_plus(C.byref(st), n)
    assert st.data, "state allocation failed"
    obs_dev = lib.orkan_dev_alloc(1 << n)
    assert obs_dev
    assert lib.orkan_dev_upload(obs_dev, obs.ctypes.data, 1 << n) == 0
    ps = C.byref(st)
    fns = {name: getattr(lib, name) for name in ("rx", "ry", "rz", "cx")}

    def apply(op):
        if op[0] == "cx":
            fns["cx"](ps, op[1], op[2])
        else:
            fns[op[0]](ps, op[1], op[2])

    log(f"state and observable resident ({n} qubits); warmup {args.warmup} layers")
    for l in range(args.warmup):
        for op in layers[l]:
            apply(op)
    lib.orkan_sync()
    log("timed layers start")

    ev = Events(lib)
    rot_pairs, cx_pairs, step_marks = [], [], []
    total_bytes = 0
    step_marks.append(ev.rec())
    t0 = time.perf_counter()
    for l in range(args.warmup, args.warmup + steps):
        for op in layers[l]:
            a = ev.rec()
            apply(op)
            b = ev.rec()
            (cx_pairs if op[0] == "cx" else rot_pairs).append((a, b, op[0]))
            total_bytes += wl.sv_gate_bytes(op[0], n, elem)
        step_marks.append(ev.rec())
    lib.orkan_sync()
    wall = time.perf_counter() - t0
    log(f"timed layers done in {wall:.2f} s")
    total_ms = ev.ms(step_marks[0], step_marks[-1])
    step_ms = [ev.ms(step_marks[i], step_marks[i + 1]) for i in range(len(step_marks) - 1)]
    rot_ms = np.array([ev.ms(a, b) for a, b, _ in rot_pairs])
    rot_names = np.array([nm for _, _, nm in rot_pairs])
    cx_ms = np.array([ev.ms(a, b) for a, b, _ in cx_pairs])

    # mean() timed on its own, three calls (reduction reads psi and obs: 24 B per amplitude; each call
    # includes its second stage and the read of the result to the host)
    m0 = ev.rec()
    mval = lib.mean(ps, C.cast(obs_dev, C.POINTER(C.c_double)))
    for _ in range(2):
        lib.mean(ps, C.cast(obs_dev, C.POINTER(C.c_double)))
    m1 = ev.rec()
    mean_ms = ev.ms(m0, m1) / 3
    # after the circuit: matel_diag against a copy (reads both states and obs, 40 B per amplitude) and
    # exp_diag with obs as the diagonal (reads psi and the diagonal, writes psi: 40 B per amplitude)
    obs_p = C.cast(obs_dev, C.POINTER(C.c_double))
    other = lib.state_cp(ps)
    lib.orkan_sync()
    x0 = ev.rec()
    for _ in range(3):
        lib.matel_diag(ps, C.byref(other), obs_p)
    x1 = ev.rec()
    lib.state_free(C.byref(other))
    lib.exp_diag(ps, obs_p, 0.3)
    e0 = ev.rec()
    lib.exp_diag(ps, obs_p, 0.3)
    e1 = ev.rec()
    lib.orkan_sync()
    matel_ms, expd_ms = ev.ms(x0, x1) / 3, ev.ms(e0, e1)

    if args.copy_calibration:
        cp = lib.orkan_state_convert(ps, 1 if args.fp32 else 0)
        lib.orkan_sync()
        lib.state_free(C.byref(cp))
        log("copy calibration kernel done")

    rot_bytes = wl.sv_gate_bytes("rx", n, elem)
    rot_avg = float(rot_ms.mean())
    achieved = rot_bytes / (rot_avg * 1e-3) / 1e9
    per_gate = {}
    for name in ("rx", "ry", "rz"):
        per_gate[name] = round(rot_bytes / (float(rot_ms[rot_names == name].mean()) * 1e-3) / 1e9, 1)
    cx_bytes = wl.sv_gate_bytes("cx", n, elem)
    per_gate["cx"] = round(cx_bytes / (float(cx_ms.mean()) * 1e-3) / 1e9, 1)
    per_gate["mean"] = round((elem + 8) * (1 << n) / (mean_ms * 1e-3) / 1e9, 1) if mean_ms > 0 else None
    per_gate["matel_diag"] = round((2 * elem + 8) * (1 << n) / (matel_ms * 1e-3) / 1e9, 1) if matel_ms > 0 else None
    per_gate["exp_diag"] = round((2 * elem + 8) * (1 << n) / (expd_ms * 1e-3) / 1e9, 1) if expd_ms > 0 else None

    fused = fused_layer(lib, ps, n, elem, layers[args.warmup] if steps > 0 else layers[0], rot_bytes, rot_avg)

    value = total_bytes / (total_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles("sv_pair_kernel_rot_28q" + ("_fp32" if args.fp32 else ""))
    lib.orkan_dev_free(obs_dev)
    lib.state_free(ps)
    result = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms / steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if args.fp32 else "c128",
        "data": "synthetic",
        "config": {"workload": f"statevector {n} qubits, brick layers of rx/ry/rz on every qubit + cx on "
                               f"alternating pairs, from |+>^{n}",
                   "qubits": n, "layers_timed": steps, "gates_per_step": len(layers[0]),
                   "storage": "complex64" if args.fp32 else "complex128", "seed": hex(wl.SEED_BASE + 2)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": traffic[0], "traffic_source": traffic[1],
                     "kernel": "sv_pair_kernel (rx/ry/rz)", "bytes_per_launch": rot_bytes,
                     "avg_launch_ms": round(rot_avg, 4), "launches": int(rot_ms.size)},
        "per_gate_GBs": per_gate,
        "pct_peak_per_gate": {k: (round(v / PEAK_HBM_GBS * 100, 1) if v else None) for k, v in per_gate.items()},
        "mean_value": mval,
        "fused_layer": fused,
        "wall_s": round(wall, 3),
        "step_ms_minmax": [round(min(step_ms), 3), round(max(step_ms), 3)],
    }
    return result, (layers[args.warmup] if steps > 0 else layers[0]), n


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _rot_matrix(name, th):
    c, s = np.cos(th / 2), np.sin(th / 2)
    if name == "rx":
        return np.array([[c, -1j * s], [-1j * s, c]])
    if name == "ry":
        return np.array([[c, -s], [s, c]])
    return np.array([[np.exp(-0.5j * th), 0], [0, np.exp(0.5j * th)]])


def fused_layer(lib, ps, n, elem, layer, rot_bytes, rot_avg_ms):
    """The layer's rx, ry, rz on every qubit as one orkan_layer_1q call (U = RZ RY RX per qubit),
    timed over three calls; reported apart from the per-gate figures, which time one gate per call."""
    ang = {}
    for op in layer:
        if op[0] != "cx":
            ang.setdefault(op[1], {})[op[0]] = op[2]
    mats = []
    for q in range(n):
        a = ang.get(q, {})
        mats.append(_rot_matrix("rz", a.get("rz", 0.0)) @ _rot_matrix("ry", a.get("ry", 0.0))
                    @ _rot_matrix("rx", a.get("rx", 0.0)))
    m = np.ascontiguousarray(np.concatenate([M.T.reshape(4) for M in mats]), np.complex128)
    ev = Events(lib)
    lib.orkan_sync()
    f0 = ev.rec()
    for _ in range(3):
        assert lib.orkan_layer_1q(ps, m.ctypes.data) == 0
    f1 = ev.rec()
    lib.orkan_sync()
    ms = ev.ms(f0, f1) / 3
    rots = sum(1 for op in layer if op[0] != "cx")
    sweeps = 1 + max(0, -(-(n - 12) // 8))
    return {"rotations": rots, "ms": round(ms, 4), "sweeps": sweeps,
            "hbm_GBs": round(sweeps * 2 * elem * (1 << n) / (ms * 1e-3) / 1e9, 1),
            "speedup_vs_gate_calls": round(rots * rot_avg_ms / ms, 2),
            "note": "fused rotations of one layer, one call: HBM GB/s on the bytes the sweeps move; not a per-gate figure"}


def sv_fp32_block(lib, layers, n, nlayers=2):
    """FP32-storage statevector: the same brick layers on complex64 and on
    complex128 states from |+>; rotation-kernel GB/s (8 B per amplitude) and the
    max amplitude error against complex128."""
    from orkan_amd import ORKAN_FP32, StateT
    from benchmarks import workloads as wl
    s64, s32 = StateT(0, None, 0), StateT(0 | ORKAN_FP32, None, 0)
    lib.state_plus(C.byref(s64), n)
    lib.state_plus(C.byref(s32), n)
    ev = Events(lib)
    rot = []
    for l in range(nlayers):
        for op in layers[l]:
            for st, timed in ((s64, False), (s32, True)):
                a = ev.rec() if timed else None
                if op[0] == "cx":
                    lib.cx(C.byref(st), op[1], op[2])
                else:
                    getattr(lib, op[0])(C.byref(st), op[1], op[2])
                if timed and op[0] != "cx":
                    rot.append((a, ev.rec()))
    lib.orkan_sync()
    ms = float(np.mean([ev.ms(a, b) for a, b in rot]))
    back = lib.orkan_state_convert(C.byref(s32), 0)
    N = 1 << n
    err = 0.0
    chunk = 1 << 24
    a64 = np.empty(N, np.complex128)
    b64 = np.empty(N, np.complex128)
    assert lib.orkan_state_download(C.byref(s64), a64.ctypes.data) == 0
    assert lib.orkan_state_download(C.byref(back), b64.ctypes.data) == 0
    for i in range(0, N, chunk):
        err = max(err, float(np.abs(a64[i:i + chunk] - b64[i:i + chunk]).max()))
    norm = float(np.vdot(b64, b64).real)
    del a64, b64
    for st in (s64, s32, back):
        lib.state_free(C.byref(st))
    gbs = wl.sv_gate_bytes("rx", n, elem=8) / (ms * 1e-3) / 1e9
    return {"rotation_GB/s": round(gbs, 1), "pct_peak": round(gbs / PEAK_HBM_GBS * 100, 1),
            "rotation_avg_ms": round(ms, 4), "layers": nlayers,
            "max_abs_amplitude_err_vs_fp64": err, "norm_drift": abs(norm - 1.0)}


def cpu_baseline_sv(layer, n, elem=16):
    """The reference library (oracle/_ref, compiled from its sources; the restatement where it is
    absent) on this host at the same n, OpenMP over the cores this process may use.  Per gate type
    the reference's two-level method (benchmarks/cpu.py): one sweep of the layer's positions per
    timed block (rx, ry, rz: every qubit; cx: the layer's pairs), 3 warm-up sweeps, 5 samples,
    median per application; GB/s weighs the medians by one layer's gate counts.  One whole layer in
    order is timed beside it."""
    from benchmarks import cpu as cpu_mod
    from benchmarks import workloads as wl
    threads = cpu_mod.setup_threads()
    kind, path = cpu_mod.library(5)
    if path is None:
        return None
    from orkan_amd import _ffi
    ref = _ffi.bind(path)
    st = _ffi.StateT(0, None, 0)
    ref.state_plus(C.byref(st), n)
    ps = C.byref(st)
    ang = {}
    for op in layer:
        if op[0] != "cx":
            ang.setdefault(op[0], {})[op[1]] = op[2]
    pairs = [(op[1], op[2]) for op in layer if op[0] == "cx"]
    sweeps = {g: (lambda g=g: [getattr(ref, g)(ps, q, ang[g][q]) for q in range(n)]) for g in ("rx", "ry", "rz")}
    sweeps["cx"] = lambda: [ref.cx(ps, a, b) for a, b in pairs]
    positions = {"rx": n, "ry": n, "rz": n, "cx": len(pairs)}
    per = {}
    for g, fn in sweeps.items():
        r = cpu_mod.two_level(fn, warmup=3, samples=5)
        for k in ("median_s", "min_s", "max_s"):
            r[k] /= positions[g]
        per[g] = r
        log(f"cpu baseline: {g} {r['median_s'] * 1e3:.2f} ms per gate (median of 5 sweeps)")
    gbs, t_layer = cpu_mod.weighted(per, positions, lambda g: wl.sv_gate_bytes(g, n, elem))
    t0 = time.perf_counter()
    for op in layer:
        if op[0] == "cx":
            ref.cx(ps, op[1], op[2])
        else:
            getattr(ref, op[0])(ps, op[1], op[2])
    wall = time.perf_counter() - t0
    ref.state_free(ps)
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": kind,
            "sample": f"reference two-level method per gate type (rx, ry, rz over all {n} qubits, cx over the "
                      f"layer's {len(pairs)} pairs), 5 samples of one sweep after 3 warm-up sweeps, medians weighted "
                      f"by one brick layer's gate counts; {n}-qubit complex128 state",
            "median_ms_per_gate": {g: round(v["median_s"] * 1e3, 3) for g, v in per.items()},
            "GB/s_per_gate": {g: round(wl.sv_gate_bytes(g, n, elem) / v["median_s"] / 1e9, 2) for g, v in per.items()},
            "layer_ms_weighted": round(t_layer * 1e3, 2), "layer_wall_ms": round(wall * 1e3, 2),
            "circuit_s_projected": round(40 * wall, 1), "host": cpu_mod.host_info()}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the multi-GPU sharding is a design "
              "note, not a measured configuration", file=sys.stderr)
        sys.exit(2)
    import orkan_amd
    lib = orkan_amd.load(5)
    if args.workload == "sv28":
        result, sample_layer, n = run_sv28(args, lib)
        log(f"gpu: {result['value']} GB/s; cpu baseline next")
        if not args.no_cpu:
            result["cpu_baseline"] = cpu_baseline_sv(sample_layer, n)
        if not args.quick:
            from benchmarks import workloads as wl
            from benchmarks import dm as bench_dm, qaoa as bench_qaoa, screening
            log("fp32 statevector block")
            lay, _ = wl.brick_circuit(n, 2)
            result["fp32_storage"] = sv_fp32_block(lib, lay, n)
            dm_args = argparse.Namespace(steps=2, warmup=1)
            result["dm14"] = bench_dm.run(dm_args, lib, Events, cpu=not args.no_cpu)
            result["qaoa"] = bench_qaoa.run(lib, ROOT, cpu_budget_s=0.0 if args.no_cpu else 12.0)
            log("surrogate block (config 5: 10^6 candidates per held-out graph, 4 graphs)")
            result["surrogate"] = screening.run(lib, test_graphs=4)
    elif args.workload == "dm14":
        from benchmarks import dm as bench_dm
        result = {"metric": METRIC, "dm14": bench_dm.run(args, lib, Events, cpu=not args.no_cpu)}
    elif args.workload == "qaoa":
        from benchmarks import qaoa as bench_qaoa
        result = {"metric": METRIC, "qaoa": bench_qaoa.run(lib, ROOT, cpu_budget_s=0.0 if args.no_cpu else 12.0)}
    else:
        from benchmarks import screening as bench_surrogate
        result = {"metric": METRIC, "surrogate": bench_surrogate.run(lib)}
    print(json.dumps(result))


if __name__ == "__main__":
    main()
