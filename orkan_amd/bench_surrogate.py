"""Learned screening for QAOA angle search (n = 16, p = 4).

1. Training data: `train_graphs` G(16, 0.5) MaxCut instances (weights U[0, 1]),
   `per_graph` random angle sets each (gamma, beta ~ U[0, pi]); energies from
   the batched simulator, target <C> / sum(w).
2. Training: the MLP surrogate (BF16 MFMA forward/backward, Adam) on the GPU.
3. Held-out graphs: surrogate MAE against the simulator on fresh angles.
4. Search per held-out graph over the same pool of `ncand` random angle sets:
     screened -- surrogate on all candidates, exact simulation of the top 64;
     plain    -- exact simulation of the candidates in batches, in a random
                 order, until its best approximation ratio reaches the screened one.
   Reported: wall time of each to the same best ratio and the speedup.
Approximation ratio = <C> / max_x C(x).
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


def run(lib, train_graphs=512, per_graph=2048, test_graphs=8, ncand=1_000_000, topk=64, epochs=4,
        batch_size=4096, plain_batch=8192):
    rng = wl.rng(5)
    tr = graphs(rng, train_graphs)
    te = graphs(rng, test_graphs)
    allg = tr + te
    sim = batch.QaoaBatch(lib, [N] * len(allg), [g["diag"] for g in allg], max_evals=65536, max_p=P)

    # ---- 1. training data from the simulator
    t0 = time.perf_counter()
    gidx = np.repeat(np.arange(train_graphs), per_graph)
    G = rng.uniform(0, np.pi, (gidx.size, P))
    B = rng.uniform(0, np.pi, (gidx.size, P))
    E = simulate(sim, gidx, G, B)
    t_data = time.perf_counter() - t0
    wsum = np.array([g["wsum"] for g in tr])
    T = (E / wsum[gidx]).astype(np.float32)
    gfs = np.stack([g["gf"] for g in tr])
    log(f"surrogate: {gidx.size} training energies in {t_data:.1f} s")

    # ---- 2. training
    mlp = surrogate.Mlp(lib, max_batch=max(batch_size, 65536))
    mlp.init(rng)
    t0 = time.perf_counter()
    steps = epochs * (gidx.size // batch_size)
    losses = []
    for s in range(steps):
        idx = rng.integers(0, gidx.size, batch_size)
        X = surrogate.features(gfs[gidx[idx]], G[idx], B[idx])
        lr = 1e-3 * (0.1 ** (s / max(steps - 1, 1)))
        losses.append(mlp.train_step(X, T[idx], lr))
        if s % 256 == 0:
            log(f"surrogate: step {s}/{steps} loss {losses[-1]:.3e}")
    t_train = time.perf_counter() - t0

    # ---- 3. held-out accuracy
    hg = np.repeat(np.arange(test_graphs), 4096)
    HG = rng.uniform(0, np.pi, (hg.size, P))
    HB = rng.uniform(0, np.pi, (hg.size, P))
    Eh = simulate(sim, hg + train_graphs, HG, HB)
    ws_te = np.array([g["wsum"] for g in te])
    pred = mlp.forward(surrogate.features(np.stack([g["gf"] for g in te])[hg], HG, HB))
    mae = float(np.mean(np.abs(pred - Eh / ws_te[hg])))
    mae_energy = float(np.mean(np.abs(pred * ws_te[hg] - Eh)))

    # ---- 4. search on the held-out graphs over one shared candidate pool
    CG = rng.uniform(0, np.pi, (ncand, P))
    CB = rng.uniform(0, np.pi, (ncand, P))
    ang = np.ascontiguousarray(np.concatenate([CG, CB], 1) / np.pi, dtype=np.float32)
    nbytes = ang.nbytes
    dang = lib.orkan_dev_alloc((nbytes + 7) // 8)
    assert lib.orkan_copy_to_device(dang, ang.ctypes.data, nbytes) == 0
    dy = lib.orkan_dev_alloc((ncand * 4 + 7) // 8)
    ypred = np.empty(ncand, np.float32)
    rows = []
    for k, g in enumerate(te):
        gi = train_graphs + k
        # screened search
        t0 = time.perf_counter()
        mlp.forward_candidates(g["gf"], dang, ncand, dy)
        assert lib.orkan_copy_to_host(ypred.ctypes.data, dy, ypred.nbytes) == 0
        top = np.argpartition(-ypred, topk)[:topk]
        ex = simulate(sim, np.full(topk, gi), CG[top], CB[top])
        best_s = float(ex.max() / g["cmax"])
        t_s = time.perf_counter() - t0
        # plain batched search: exact evaluation of the pool in random order until best >= best_s
        t0 = time.perf_counter()
        order = np.random.default_rng(k).permutation(ncand)
        best_p, done = 0.0, 0
        while done < ncand and best_p < best_s - 1e-12:
            sel = order[done:done + plain_batch]
            e = simulate(sim, np.full(sel.size, gi), CG[sel], CB[sel])
            best_p = max(best_p, float(e.max() / g["cmax"]))
            done += sel.size
        t_p = time.perf_counter() - t0
        rows.append({"graph": k, "best_ratio_screened": round(best_s, 6), "best_ratio_plain": round(best_p, 6),
                     "screened_s": round(t_s, 4), "plain_s": round(t_p, 4), "plain_evals": done,
                     "reached": bool(best_p >= best_s - 1e-12), "speedup": round(t_p / t_s, 2)})
        log(f"surrogate: graph {k}: screened {best_s:.4f} in {t_s:.3f} s, plain {best_p:.4f} after {done} "
            f"evals in {t_p:.3f} s")
    lib.orkan_dev_free(dang)
    lib.orkan_dev_free(dy)
    mlp.close()
    sim.close()
    sp = np.array([r["speedup"] for r in rows])
    return {"train_samples": int(gidx.size), "train_data_s": round(t_data, 2), "train_s": round(t_train, 2),
            "train_steps": steps, "final_loss": float(np.mean(losses[-32:])),
            "heldout_mae_normalised": round(mae, 5), "heldout_mae_energy": round(mae_energy, 5),
            "candidates_per_graph": ncand, "topk": topk, "search": rows,
            "speedup_median": float(np.median(sp)), "speedup_min": float(sp.min()),
            "all_reached_equal_ratio": all(r["reached"] for r in rows)}
