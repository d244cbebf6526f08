"""Batched QAOA evaluation on the device (include/orkan_amd.h, orkan_qaoa_*).

One evaluation = the reference call sequence state_plus(PURE, n);
p x {exp_diag(C, gamma), rx(q, 2 beta) for every q}; mean(C), for the MaxCut
cost diagonal C of a weighted graph.  `QaoaBatch` binds the C entry points;
`maxcut_instances` generates the synthetic Erdos-Renyi workload.
"""
import ctypes as C

import numpy as np

VP = C.c_void_p
IP = C.POINTER(C.c_int)
DP = C.POINTER(C.c_double)

EXTRA = {
    "orkan_qaoa_create": (VP, [C.c_int, C.POINTER(C.c_ubyte), C.POINTER(VP), C.c_int, C.c_int]),
    "orkan_qaoa_destroy": (None, [VP]),
    "orkan_qaoa_eval": (C.c_int, [VP, C.c_int, IP, IP, DP, DP, DP]),
    "orkan_qaoa_set_config": (C.c_int, [VP, C.c_int, C.c_int, C.c_int]),
    "orkan_qaoa_get_config": (C.c_int, [VP, IP, IP, IP]),
    "orkan_qaoa_autotune": (C.c_int, [VP, C.c_int, IP, IP, DP, DP, C.c_int, DP]),
    "orkan_maxcut_diag_dev": (C.c_int, [C.c_int, C.c_int, IP, IP, DP, VP]),
}


def maxcut_diag(n, edges, weights):
    """C(x) = sum_(i,j) w_ij [x_i != x_j] on the host (numpy)."""
    x = np.arange(1 << n, dtype=np.int64)
    out = np.zeros(1 << n)
    for (i, j), w in zip(edges, weights):
        out += w * (((x >> i) ^ (x >> j)) & 1)
    return out


def maxcut_instances(count, rng, n_lo=10, n_hi=20, p_lo=1, p_hi=8):
    """Erdos-Renyi G(n, 0.5) graphs with n ~ U{n_lo..n_hi}, weights ~ U[0, 1];
    p ~ U{p_lo..p_hi}; gamma, beta ~ U[0, pi]."""
    inst = []
    for _ in range(count):
        n = int(rng.integers(n_lo, n_hi + 1))
        iu, ju = np.triu_indices(n, 1)
        keep = rng.random(iu.size) < 0.5
        edges = np.stack([iu[keep], ju[keep]], 1).astype(np.int32)
        w = rng.uniform(0.0, 1.0, edges.shape[0])
        p = int(rng.integers(p_lo, p_hi + 1))
        gamma = rng.uniform(0.0, np.pi, p)
        beta = rng.uniform(0.0, np.pi, p)
        inst.append({"n": n, "edges": edges, "w": w, "p": p, "gamma": gamma, "beta": beta})
    return inst


class QaoaBatch:
    """Graphs (cost diagonals) resident on the device; evaluate lists of
    (graph, p, gammas, betas)."""

    def __init__(self, lib, nqubits, diags, max_evals, max_p):
        self.lib = lib
        self.max_p = max_p
        self.max_evals = max_evals
        nq = (C.c_ubyte * len(nqubits))(*nqubits)
        self._keep = [np.ascontiguousarray(d, dtype=np.float64) if isinstance(d, np.ndarray) else d for d in diags]
        ptrs = (VP * len(diags))(*[d.ctypes.data if isinstance(d, np.ndarray) else int(d) for d in self._keep])
        self.h = lib.orkan_qaoa_create(len(nqubits), nq, ptrs, max_evals, max_p)
        if not self.h:
            raise RuntimeError("orkan_qaoa_create failed")

    def _rows(self, graph, p, gammas, betas):
        m = len(graph)
        G = np.zeros((m, self.max_p))
        B = np.zeros((m, self.max_p))
        for e in range(m):
            G[e, :p[e]] = gammas[e][:p[e]]
            B[e, :p[e]] = betas[e][:p[e]]
        return (np.ascontiguousarray(graph, dtype=np.int32), np.ascontiguousarray(p, dtype=np.int32), G, B)

    def eval(self, graph, p, gammas, betas):
        g, pp, G, B = self._rows(graph, p, gammas, betas)
        out = np.empty(len(g))
        rc = self.lib.orkan_qaoa_eval(self.h, len(g), g.ctypes.data_as(IP), pp.ctypes.data_as(IP),
                                      G.ctypes.data_as(DP), B.ctypes.data_as(DP), out.ctypes.data_as(DP))
        if rc != 0:
            raise RuntimeError(f"orkan_qaoa_eval failed ({rc})")
        return out

    def autotune(self, graph, p, gammas, betas, reps=3):
        g, pp, G, B = self._rows(graph, p, gammas, betas)
        best = C.c_double(0)
        rc = self.lib.orkan_qaoa_autotune(self.h, len(g), g.ctypes.data_as(IP), pp.ctypes.data_as(IP),
                                          G.ctypes.data_as(DP), B.ctypes.data_as(DP), reps, C.byref(best))
        if rc != 0:
            raise RuntimeError(f"orkan_qaoa_autotune failed ({rc})")
        return best.value, self.config()

    def set_config(self, lt, threads, blocks_per_cu=0):
        rc = self.lib.orkan_qaoa_set_config(self.h, lt, threads, blocks_per_cu)
        if rc != 0:
            raise RuntimeError(f"orkan_qaoa_set_config failed ({rc})")

    def config(self):
        lt, th, gr = C.c_int(), C.c_int(), C.c_int()
        self.lib.orkan_qaoa_get_config(self.h, C.byref(lt), C.byref(th), C.byref(gr))
        return {"lt": lt.value, "threads": th.value, "grid": gr.value}

    def close(self):
        if self.h:
            self.lib.orkan_qaoa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
