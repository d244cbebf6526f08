"""CPU baseline: the reference library compiled from its own sources (oracle/_ref), or the oracle's
restatement where that is absent, timed on this host's cores.

Method (reference docs/BENCHMARK.md:84-93, :139-154): per gate type a two-level loop -- `warmup`
untimed calls, then `samples` timed samples of `iters` calls each on CLOCK_MONOTONIC
(time.perf_counter), reporting the median seconds per call.  The whole-layer figure then weighs
each gate type's median by its count in one layer.  Threads: OMP_NUM_THREADS (all cores this
process may run on unless the caller set it), OMP_PLACES=cores, OMP_PROC_BIND=spread; these must
be in the environment before the OpenMP runtime starts, i.e. before the first CPU library loads.
"""
import os
import statistics
import subprocess
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def setup_threads():
    """Put the OpenMP settings into the environment (before any OpenMP library is loaded) and
    return the thread count."""
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    os.environ.setdefault("OMP_PLACES", "cores")
    if "OMP_NUM_THREADS" not in os.environ:
        try:
            n = len(os.sched_getaffinity(0))
        except AttributeError:
            n = os.cpu_count() or 1
        os.environ["OMP_NUM_THREADS"] = str(n)
    return int(os.environ["OMP_NUM_THREADS"])


def host_info():
    """Sockets, physical cores, NUMA nodes and model from lscpu, plus the CPUs this process may use
    and the OpenMP placement settings."""
    info = {}
    try:
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=20).stdout
        kv = {}
        for line in out.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        model = kv.get("Model name")
        sockets = int(kv.get("Socket(s)", "0") or 0)
        cps = int(kv.get("Core(s) per socket", "0") or 0)
        tpc = int(kv.get("Thread(s) per core", "0") or 0)
        info.update({"model": model, "sockets": sockets, "cores_per_socket": cps, "threads_per_core": tpc,
                     "physical_cores": sockets * cps, "logical_cpus": int(kv.get("CPU(s)", "0") or 0),
                     "numa_nodes": int(kv.get("NUMA node(s)", "0") or 0)})
    except (OSError, ValueError, subprocess.SubprocessError):
        info["lscpu"] = "unavailable"
    try:
        info["cpus_available_to_process"] = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    for k in ("OMP_NUM_THREADS", "OMP_PLACES", "OMP_PROC_BIND"):
        info[k] = os.environ.get(k)
    return info


def library(log_tile=5):
    """(kind, path) of the CPU library to time: the compiled reference, else the restatement."""
    ref = os.path.join(ROOT, "oracle", "_ref", f"liborkan_ref_t{log_tile}.so")
    if os.path.exists(ref):
        return "reference", ref
    port = os.path.join(ROOT, "oracle", "lib", f"liborkan_oracle_t{log_tile}.so")
    if os.path.exists(port):
        return "port", port
    return None, None


def two_level(fn, warmup=3, samples=5, iters=1):
    """median, min and max seconds per call of fn() over `samples` samples of `iters` calls,
    after `warmup` untimed calls"""
    for _ in range(warmup):
        fn()
    per = []
    for _ in range(samples):
        t0 = time.perf_counter()
        for _ in range(iters):
            fn()
        per.append((time.perf_counter() - t0) / iters)
    return {"median_s": statistics.median(per), "min_s": min(per), "max_s": max(per), "samples": samples,
            "warmup": warmup, "iters": iters}


def weighted(per_type, counts, bytes_of):
    """GB/s of a layer whose gate types occur counts[t] times, each at its median per call"""
    t = sum(counts[k] * per_type[k]["median_s"] for k in counts)
    b = sum(counts[k] * bytes_of(k) for k in counts)
    return b / t / 1e9, t
