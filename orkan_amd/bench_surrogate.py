"""Learned screening for QAOA angle search (n = 16, p = 4).

1. Training data: `train_graphs` G(16, 0.5) MaxCut instances (weights U[0, 1]),
   `per_graph` random angle sets each (gamma, beta ~ U[0, pi]); energies from
   the batched simulator, target <C> / sum(w).
2. Training on the GPU (BF16 MFMA forward/backward, Adam), then `rounds` of
   refinement on the training graphs only: the surrogate ranks a candidate pool,
   the simulator labels its top candidates, and those labels join the training
   set -- the model learns the region it will be asked to rank.
3. Held-out graphs: MAE against the simulator on uniform angles and on the
   surrogate's own top picks.
4. Search per held-out graph over one pool of `ncand` random angle sets:
     screened -- surrogate on the whole pool, exact simulation of its top 64;
     plain    -- exact simulation of the pool in a random order, in batches,
                 until its best ratio reaches the screened one.
   The plain time counts only the evaluations needed (index of the first
   candidate reaching the ratio times the steady-state time per evaluation),
   which favours the plain search.  Approximation ratio = <C> / max_x C(x).
"""
import sys
import time

import numpy as np

from . import batch, surrogate, workloads as wl

N, P = 16, 4


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def graphs(rng, count):
    inst = batch.maxcut_instances(count, rng, n_lo=N, n_hi=N, p_lo=P, p_hi=P)
    for i in inst:
        i["diag"] = batch.maxcut_diag(N, i["edges"], i["w"])
        i["wsum"] = float(np.sum(i["w"]))
        i["cmax"] = float(i["diag"].max())
        i["gf"] = surrogate.graph_features(i["w"])
    return inst


def simulate(b, gidx, G, B, chunk=65536):
    """exact <C> for rows (graph gidx[k], gammas G[k], betas B[k])"""
    out = np.empty(len(gidx))
    for s in range(0, len(gidx), chunk):
        e = slice(s, s + chunk)
        m = len(gidx[e])
        out[e] = b.eval(list(gidx[e]), [P] * m, G[e], B[e])
    return out


class Pool:
    """candidate angle sets resident on the device for screening"""

    def __init__(self, lib, G, B):
        self.lib, self.G, self.B = lib, G, B
        ang = np.ascontiguousarray(np.concatenate([G, B], 1) / np.pi, dtype=np.float32)
        self.n = ang.shape[0]
        self.dang = lib.orkan_dev_alloc((ang.nbytes + 7) // 8)
        assert lib.orkan_copy_to_device(self.dang, ang.ctypes.data, ang.nbytes) == 0
        self.dy = lib.orkan_dev_alloc((self.n * 4 + 7) // 8)

    def rank(self, mlp, gf, top):
        """indices of the `top` best predictions (largest first); self.vals holds them"""
        import ctypes as C
        mlp.forward_candidates(gf, self.dang, self.n, self.dy)
        idx = np.empty(top, np.int32)
        self.vals = np.empty(top, np.float32)
        rc = self.lib.orkan_topk_f32(self.dy, self.n, top, idx.ctypes.data_as(C.POINTER(C.c_int)),
                                     self.vals.ctypes.data_as(C.POINTER(C.c_float)))
        assert rc == 0
        return idx.astype(np.int64)

    def close(self):
        self.lib.orkan_dev_free(self.dang)
        self.lib.orkan_dev_free(self.dy)


def run(lib, train_graphs=512, per_graph=2048, test_graphs=8, ncand=1_000_000, topk=64, epochs=4,
        batch_size=4096, plain_batch=16384, rounds=3, ref_cand=65536, ref_top=256, ref_steps=400,
        plain_cap=1 << 20):
    rng = wl.rng(5)
    tr = graphs(rng, train_graphs)
    te = graphs(rng, test_graphs)
    allg = tr + te
    sim = batch.QaoaBatch(lib, [N] * len(allg), [g["diag"] for g in allg], max_evals=65536, max_p=P)
    wsum = np.array([g["wsum"] for g in allg])
    gfs = np.stack([g["gf"] for g in allg])

    # ---- 1. uniform training data from the simulator
    t0 = time.perf_counter()
    gidx = np.repeat(np.arange(train_graphs), per_graph)
    G = rng.uniform(0, np.pi, (gidx.size, P))
    B = rng.uniform(0, np.pi, (gidx.size, P))
    E = simulate(sim, gidx, G, B)
    t_data = time.perf_counter() - t0
    log(f"surrogate: {gidx.size} training energies in {t_data:.1f} s")

    mlp = surrogate.Mlp(lib, max_batch=65536)
    mlp.init(rng)

    def train(gi, GG, BB, EE, steps, lr0):
        T = (EE / wsum[gi]).astype(np.float32)
        losses = []
        for s in range(steps):
            idx = rng.integers(0, gi.size, batch_size)
            X = surrogate.features(gfs[gi[idx]], GG[idx], BB[idx])
            losses.append(mlp.train_step(X, T[idx], lr0 * (0.1 ** (s / max(steps - 1, 1)))))
        return losses

    # ---- 2. training and refinement rounds
    t0 = time.perf_counter()
    steps0 = epochs * (gidx.size // batch_size)
    losses = train(gidx, G, B, E, steps0, 1e-3)
    log(f"surrogate: uniform phase {steps0} steps, loss {np.mean(losses[-32:]):.3e}")
    RG = rng.uniform(0, np.pi, (ref_cand, P))
    RB = rng.uniform(0, np.pi, (ref_cand, P))
    rpool = Pool(lib, RG, RB)
    xg, xG, xB, xE = [gidx], [G], [B], [E]
    for r in range(rounds):
        picks_g, picks_i = [], []
        for g in range(train_graphs):
            top = rpool.rank(mlp, gfs[g], ref_top)
            picks_g.append(np.full(top.size, g))
            picks_i.append(top)
        pg = np.concatenate(picks_g)
        pi = np.concatenate(picks_i)
        pE = simulate(sim, pg, RG[pi], RB[pi])
        xg.append(pg); xG.append(RG[pi]); xB.append(RB[pi]); xE.append(pE)
        allg_i, allG, allB, allE = (np.concatenate(v) for v in (xg, xG, xB, xE))
        # half of every batch from the refined labels, half uniform
        nref = allg_i.size - gidx.size
        mix = np.concatenate([rng.integers(0, gidx.size, nref), np.arange(gidx.size, allg_i.size)])
        losses = train(allg_i[mix], allG[mix], allB[mix], allE[mix], ref_steps, 3e-4)
        log(f"surrogate: refinement round {r + 1}: {pg.size} new labels, loss {np.mean(losses[-32:]):.3e}")
        # fresh pool for the next round so the labels keep probing new candidates
        rpool.close()
        RG = rng.uniform(0, np.pi, (ref_cand, P))
        RB = rng.uniform(0, np.pi, (ref_cand, P))
        rpool = Pool(lib, RG, RB)
    rpool.close()
    t_train = time.perf_counter() - t0

    # ---- 3. held-out accuracy (uniform angles)
    hg = np.repeat(np.arange(test_graphs), 4096) + train_graphs
    HG = rng.uniform(0, np.pi, (hg.size, P))
    HB = rng.uniform(0, np.pi, (hg.size, P))
    Eh = simulate(sim, hg, HG, HB)
    pred = mlp.forward(surrogate.features(gfs[hg], HG, HB))
    mae = float(np.mean(np.abs(pred - Eh / wsum[hg])))
    mae_energy = float(np.mean(np.abs(pred * wsum[hg] - Eh)))

    # ---- 4. search on the held-out graphs over one shared candidate pool
    CG = rng.uniform(0, np.pi, (ncand, P))
    CB = rng.uniform(0, np.pi, (ncand, P))
    pool = Pool(lib, CG, CB)
    rows, top_err = [], []
    for k, g in enumerate(te):
        gi = train_graphs + k
        t0 = time.perf_counter()
        top = pool.rank(mlp, g["gf"], topk)
        ex = simulate(sim, np.full(topk, gi), CG[top], CB[top])
        best_s = float(ex.max() / g["cmax"])
        t_s = time.perf_counter() - t0
        top_err.append(float(np.mean(np.abs(pool.vals - ex / g["wsum"]))))
        # plain batched search over the same pool, random order
        order = np.random.default_rng(k).permutation(ncand)[:plain_cap]
        t0 = time.perf_counter()
        best_p, first, done = 0.0, None, 0
        while done < order.size:
            sel = order[done:done + plain_batch]
            e = simulate(sim, np.full(sel.size, gi), CG[sel], CB[sel]) / g["cmax"]
            if first is None and e.max() >= best_s - 1e-12:
                first = done + int(np.argmax(e >= best_s - 1e-12))
            best_p = max(best_p, float(e.max()))
            done += sel.size
            if first is not None:
                break
        t_p_meas = time.perf_counter() - t0
        per_eval = t_p_meas / done
        reached = first is not None
        t_p = (first + 1) * per_eval if reached else t_p_meas
        rows.append({"graph": k, "best_ratio_screened": round(best_s, 6), "best_ratio_plain": round(best_p, 6),
                     "screened_s": round(t_s, 5), "plain_s": round(t_p, 5), "plain_s_measured": round(t_p_meas, 5),
                     "plain_evals_to_reach": (first + 1) if reached else None, "plain_evals_run": done,
                     "reached": reached, "speedup": round(t_p / t_s, 2) if reached else None})
        log(f"surrogate: graph {k}: screened {best_s:.4f} in {t_s * 1e3:.1f} ms; plain reaches it after "
            f"{(first + 1) if reached else 'no'} evals ({t_p * 1e3:.1f} ms), plain best {best_p:.4f}")
    pool.close()
    mlp.close()
    sim.close()
    sp = np.array([r["speedup"] for r in rows if r["speedup"] is not None])
    return {"train_samples_uniform": int(gidx.size), "refined_labels": int(sum(v.size for v in xg[1:])),
            "train_data_s": round(t_data, 2), "train_s": round(t_train, 2),
            "heldout_mae_normalised": round(mae, 5), "heldout_mae_energy": round(mae_energy, 5),
            "heldout_mae_top_picks_normalised": round(float(np.mean(top_err)), 5),
            "candidates_per_graph": ncand, "topk": topk, "search": rows,
            "speedup_median": float(np.median(sp)) if sp.size else None,
            "speedup_min": float(sp.min()) if sp.size else None,
            "all_reached_equal_ratio": all(r["reached"] for r in rows)}
