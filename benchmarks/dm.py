"""Density-matrix benchmark workload (14 qubits, both layouts): one step is one
layer of h and ry on every qubit, cz on alternating neighbours, amplitude
damping (gamma 0.05) and depolarizing (p 0.02) on every qubit.

Used by bench.py (--workload dm14, and the dm14 block of the default run).
Per-kernel HBM GB/s use algorithmic bytes on the L stored entries: 32 L for
mixing gates, permutations that move every entry (x, y) and channels, 16 L for
phase gates, 24 L for cx / cy / swap, 12 L for cz, 14 L for ccx.  After the
layers, the permutation gates outside the layer (x on every qubit, cx on the
layer's neighbour pairs in both directions, swap, ccx) are timed one sweep each.
The CPU baseline runs the reference library on the same starting state with the
two-level method of benchmarks/cpu.py, per gate type.
"""
import ctypes as C
import sys
import time

import numpy as np

from . import cpu as cpu_mod
from . import workloads as wl
from orkan_amd._ffi import StateT, SuperopT

PEAK = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_rho_layout(n, kind, g, rank=16, log_tile=5):
    """rho = A A^dagger / tr with A a D x rank complex Gaussian, written straight
    into the storage layout block by block (no dense D x D matrix)."""
    D = 1 << n
    A = g.standard_normal((D, rank)) + 1j * g.standard_normal((D, rank))
    A /= np.sqrt(np.sum(np.abs(A) ** 2))
    out = np.empty(wl.dm_stored_len(n, kind, log_tile), dtype=np.complex128)
    if kind == 1:   # packed, column-major lower triangle
        pos = 0
        for c in range(D):
            col = A[c:] @ A[c].conj()
            out[pos:pos + D - c] = col
            pos += D - c
    else:           # tiled, lower tiles row by row
        td = 1 << log_tile
        nt = D // td
        pos = 0
        for tr in range(nt):
            blk = A[tr * td:(tr + 1) * td] @ A[:(tr + 1) * td].conj().T      # td x (tr+1) td
            tiles = blk.reshape(td, tr + 1, td).transpose(1, 0, 2)            # (tc, lr, lc)
            out[pos:pos + (tr + 1) * td * td] = tiles.reshape(-1)
            pos += (tr + 1) * td * td
    return out


def fp32_drift(lib, kind, st64, st32, n, L):
    """max |rho64 - rho32|, trace drift |sum Re rho_ii - 1| and Hermiticity drift of the FP32
    state (max |Im rho_ii|, and for tiled the diagonal-tile mirror error)"""
    a = np.empty(L, np.complex128)
    b = np.empty(L, np.complex64)
    assert lib.orkan_state_download(C.byref(st64), a.ctypes.data) == 0
    assert lib.orkan_state_download(C.byref(st32), b.ctypes.data) == 0
    D = 1 << n
    x = np.arange(D)
    if kind == 1:
        diag = x * (2 * D - x + 1) // 2
    else:
        td = 32
        t = x >> 5
        diag = (t * (t + 1) // 2 + t) * td * td + (x & 31) * td + (x & 31)
    out = {"max_abs_err": float(np.abs(a - b.astype(np.complex128)).max()),
           "trace_drift": float(abs(np.sum(b[diag].real, dtype=np.float64) - 1.0)),
           "trace_drift_fp64": float(abs(np.sum(a[diag].real) - 1.0)),
           "herm_diag_imag_max": float(np.abs(b[diag].imag).max())}
    if kind == 2:
        worst = 0.0
        for k in range(D >> 5):
            off = (k * (k + 1) // 2 + k) * 1024
            tile = b[off:off + 1024].reshape(32, 32)
            worst = max(worst, float(np.abs(tile - tile.conj().T).max()))
        out["herm_tile_mirror_max"] = worst
    return out


def run(args, lib, Events, layouts=(2, 1), layers=None, cpu=True, fp32=True):
    n = wl.DM_QUBITS
    steps = args.steps if args.steps is not None else wl.DM_LAYERS
    g = wl.rng(3)
    thetas = g.uniform(0, 2 * np.pi, size=(steps + args.warmup, n))
    Sad = np.ascontiguousarray(wl.superop(wl.kraus_amp_damp(0.05)).reshape(-1))
    Sdp = np.ascontiguousarray(wl.superop(wl.kraus_depolarizing(0.02)).reshape(-1))
    sad = SuperopT(1, Sad.ctypes.data)
    sdp = SuperopT(1, Sdp.ctypes.data)
    res = {}
    for kind in layouts:
        name = {1: "packed", 2: "tiled"}[kind]
        L = wl.dm_stored_len(n, kind)
        data = build_rho_layout(n, kind, wl.rng(3))
        st = StateT(kind, None, 0)
        lib.state_init(C.byref(st), n, None)
        assert st.data
        assert lib.orkan_state_upload(C.byref(st), data.ctypes.data) == 0
        ps = C.byref(st)

        def apply(op):
            if op[0] == "ampdamp":
                lib.channel_1q(ps, C.byref(sad), op[1])
            elif op[0] == "depol":
                lib.channel_1q(ps, C.byref(sdp), op[1])
            elif op[0] == "cz":
                lib.cz(ps, op[1], op[2])
            elif op[0] == "ry":
                lib.ry(ps, op[1], op[2])
            else:
                getattr(lib, op[0])(ps, op[1])

        for l in range(args.warmup):
            for op in wl.dm_layer(n, l, thetas[l]):
                apply(op)
        lib.orkan_sync()
        log(f"dm14 {name}: timed layers start")
        ev = Events(lib)
        per = {}
        total_bytes = 0
        m0 = ev.rec()
        for l in range(args.warmup, args.warmup + steps):
            for op in wl.dm_layer(n, l, thetas[l]):
                a = ev.rec()
                apply(op)
                b = ev.rec()
                key = "channel" if op[0] in ("ampdamp", "depol") else op[0]
                per.setdefault(key, []).append((a, b))
                total_bytes += wl.dm_gate_bytes(key, L)
        m1 = ev.rec()
        lib.orkan_sync()
        tot = ev.ms(m0, m1)
        kern = {}
        for key, pairs in per.items():
            ms = float(np.mean([ev.ms(a, b) for a, b in pairs]))
            gbs = wl.dm_gate_bytes(key, L) / (ms * 1e-3) / 1e9
            kern[key] = {"avg_ms": round(ms, 4), "GB/s": round(gbs, 1), "pct_peak": round(gbs / PEAK * 100, 1)}
        res[name] = {"stored_entries": L, "GB/s": round(total_bytes / (tot * 1e-3) / 1e9, 1),
                     "ms_per_layer": round(tot / steps, 3), "kernels": kern}
        log(f"dm14 {name}: {res[name]['GB/s']} GB/s")
        res[name]["permutation_kernels"] = permutation_sweeps(lib, ps, n, L, Events)
        log(f"dm14 {name}: permutations {res[name]['permutation_kernels']}")
        if cpu:
            res[name]["cpu_baseline"] = cpu_baseline(kind, data, n, thetas[args.warmup], Sad, Sdp)
            log(f"dm14 {name}: cpu baseline {res[name]['cpu_baseline']['value']} GB/s")
        del data
        if fp32:
            # FP32 storage: same start state, same layers; time them and compare with FP64
            data = build_rho_layout(n, kind, wl.rng(3))
            s64 = StateT(kind, None, 0)
            lib.state_init(C.byref(s64), n, None)
            assert lib.orkan_state_upload(C.byref(s64), data.ctypes.data) == 0
            del data
            s32 = lib.orkan_state_convert(C.byref(s64), 1)
            p64, p32 = C.byref(s64), C.byref(s32)
            ev = Events(lib)
            t_ops = []
            for l in range(args.warmup, args.warmup + steps):
                for op in wl.dm_layer(n, l, thetas[l]):
                    ps = p64
                    apply(op)
                    ps = p32
                    a = ev.rec()
                    apply(op)
                    b = ev.rec()
                    t_ops.append((op[0], a, b))
            lib.orkan_sync()
            tb = sum(wl.dm_gate_bytes("channel" if o in ("ampdamp", "depol") else o, L, elem=8) for o, _, _ in t_ops)
            tt = sum(ev.ms(a, b) for _, a, b in t_ops)
            res[name]["fp32_storage"] = {"GB/s": round(tb / (tt * 1e-3) / 1e9, 1), "layers": steps,
                                         "drift_vs_fp64": fp32_drift(lib, kind, s64, s32, n, L)}
            log(f"dm14 {name} fp32: {res[name]['fp32_storage']}")
            lib.state_free(p64)
            lib.state_free(p32)
        ps = C.byref(st)
        lib.state_free(ps)
    return res


PERM_SWEEPS = (("x", 1), ("cx", 2), ("swap_gate", 2), ("ccx", 3))


def perm_positions(name, n):
    if name == "x":
        return [(q,) for q in range(n)]
    if name == "cx":
        return [(q, q + 1) for q in range(n - 1)] + [(q + 1, q) for q in range(n - 1)]
    if name == "swap_gate":
        return [(q, q + 1) for q in range(n - 1)]
    return [(q, q + 1, q + 2) for q in range(n - 2)]


def permutation_sweeps(lib, ps, n, L, Events):
    """One sweep of each permutation gate over its positions (x on every qubit, cx on the
    neighbour pairs both ways, swap on the neighbour pairs, ccx on consecutive triples), per-launch
    HIP events; GB/s on algorithmic bytes, with the slowest position."""
    ev = Events(lib)
    marks = []
    for name, _ in PERM_SWEEPS:
        fn = getattr(lib, name)
        for pos in perm_positions(name, n):
            a = ev.rec()
            fn(ps, *pos)
            marks.append((name, pos, a, ev.rec()))
    lib.orkan_sync()
    out = {}
    for name, _ in PERM_SWEEPS:
        ms = [(ev.ms(a, b), pos) for nm, pos, a, b in marks if nm == name]
        avg = float(np.mean([m for m, _ in ms]))
        worst = max(ms)
        by = wl.dm_gate_bytes(name, L)
        out[name] = {"avg_ms": round(avg, 4), "GB/s": round(by / (avg * 1e-3) / 1e9, 1),
                     "pct_peak": round(by / (avg * 1e-3) / 1e9 / PEAK * 100, 1), "positions": len(ms),
                     "slowest": {"position": list(worst[1]), "GB/s": round(by / (worst[0] * 1e-3) / 1e9, 1)}}
    return out


def cpu_baseline(kind, data, n, thetas, Sad, Sdp, warmup=3, samples=5):
    """The reference library (oracle/_ref; the restatement where it is absent) on the same starting
    state in host memory: per gate type of the layer (h, ry, cz, one channel), one sweep over the
    positions the layer uses per timed block, warm-up and samples as given; median seconds per
    application.  GB/s weighs each type's median by its count in one layer, on the same algorithmic
    bytes as the GPU figures; the whole-layer wall time of one layer is reported beside it."""
    threads = cpu_mod.setup_threads()
    lkind, path = cpu_mod.library(5)
    if path is None:
        return None
    from orkan_amd import _ffi
    ref = _ffi.bind(path)
    st = StateT(kind, None, 0)
    ref.state_init(C.byref(st), n, None)
    L = data.size
    assert st.data and ref.state_len(C.byref(st)) == L
    C.memmove(st.data, data.ctypes.data, data.nbytes)
    ps = C.byref(st)
    sad, sdp = SuperopT(1, Sad.ctypes.data), SuperopT(1, Sdp.ctypes.data)
    pairs = [(q, q + 1) for q in range(0, n - 1, 2)]
    sweeps = {
        "h": lambda: [ref.h(ps, q) for q in range(n)],
        "ry": lambda: [ref.ry(ps, q, float(thetas[q])) for q in range(n)],
        "cz": lambda: [ref.cz(ps, a, b) for a, b in pairs],
        "channel": lambda: [ref.channel_1q(ps, C.byref(sdp), q) for q in range(n)],
    }
    positions = {"h": n, "ry": n, "cz": len(pairs), "channel": n}
    per = {}
    for key, fn in sweeps.items():
        r = cpu_mod.two_level(fn, warmup=warmup, samples=samples)
        for k in ("median_s", "min_s", "max_s"):
            r[k] /= positions[key]
        per[key] = r
        log(f"dm14 cpu: {key} {r['median_s'] * 1e3:.2f} ms per application")
    layer = wl.dm_layer(n, 0, thetas)
    counts = {"h": n, "ry": n, "cz": sum(1 for op in layer if op[0] == "cz"), "channel": 2 * n}
    gbs, t_layer = cpu_mod.weighted(per, counts, lambda k: wl.dm_gate_bytes(k, L))
    # one whole layer in order (amplitude damping and depolarizing both), wall time
    t0 = time.perf_counter()
    for op in layer:
        if op[0] == "ampdamp":
            ref.channel_1q(ps, C.byref(sad), op[1])
        elif op[0] == "depol":
            ref.channel_1q(ps, C.byref(sdp), op[1])
        elif op[0] == "cz":
            ref.cz(ps, op[1], op[2])
        elif op[0] == "ry":
            ref.ry(ps, op[1], op[2])
        else:
            ref.h(ps, op[1])
    wall = time.perf_counter() - t0
    ref.state_free(ps)
    return {"value": round(gbs, 2), "unit": "GB/s", "cores": threads, "kind": lkind,
            "sample": f"reference two-level method per gate type (h, ry, cz, channel), {samples} samples of one "
                      f"sweep after {warmup} warm-up sweeps, weighted by one layer's gate counts; plus one whole "
                      f"layer of {len(layer)} operations in order",
            "median_ms_per_gate": {k: round(v["median_s"] * 1e3, 3) for k, v in per.items()},
            "GB/s_per_gate": {k: round(wl.dm_gate_bytes(k, L) / v["median_s"] / 1e9, 2) for k, v in per.items()},
            "layer_ms_weighted": round(t_layer * 1e3, 2), "layer_wall_ms": round(wall * 1e3, 2),
            "host": cpu_mod.host_info()}
