"""Learned screening for QAOA angle search.

An MLP predicts the normalised QAOA energy <C>/sum(w) of a MaxCut instance from
its sorted edge weights (padded to 120 = the edge count of K_16) and the 2p
angles (gamma/pi, beta/pi, p = 4): 128 inputs, 3 hidden layers of 512, GELU,
one output.  Training data come from the batched simulator (orkan_amd.batch);
forward, backward and Adam run on the library's BF16 MFMA kernels.
"""
import ctypes as C

import numpy as np

VP = C.c_void_p
FP = C.POINTER(C.c_float)

EXTRA = {
    "orkan_mlp_create": (VP, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "orkan_mlp_destroy": (None, [VP]),
    "orkan_mlp_set_weights": (C.c_int, [VP, C.POINTER(FP), C.POINTER(FP), FP, C.c_float]),
    "orkan_mlp_get_weights": (C.c_int, [VP, C.POINTER(FP), C.POINTER(FP), FP, FP]),
    "orkan_mlp_forward": (C.c_int, [VP, C.c_int, FP, FP]),
    "orkan_mlp_train_step": (C.c_int, [VP, C.c_int, FP, FP, C.c_float, FP]),
    "orkan_mlp_grads": (C.c_int, [VP, C.c_int, FP, FP, C.POINTER(FP), C.POINTER(FP), FP, FP, FP]),
    "orkan_mlp_gemm_test": (C.c_int, [C.c_int, C.c_int, C.c_int, VP, VP, VP]),
    "orkan_mlp_forward_candidates": (C.c_int, [VP, FP, C.c_int, VP, C.c_long, VP]),
    "orkan_topk_f32": (C.c_int, [VP, C.c_long, C.c_int, C.POINTER(C.c_int), FP]),
}

IN_DIM, HIDDEN, LAYERS, EDGE_SLOTS, P = 128, 512, 3, 120, 4


def _fp(a):
    return a.ctypes.data_as(FP)


class Mlp:
    def __init__(self, lib, in_dim=IN_DIM, hidden=HIDDEN, layers=LAYERS, max_batch=8192):
        self.lib = lib
        self.in_dim, self.hidden, self.layers = in_dim, hidden, layers
        self.h = lib.orkan_mlp_create(in_dim, hidden, layers, max_batch)
        if not self.h:
            raise RuntimeError("orkan_mlp_create failed")
        self.max_batch = max_batch

    def init(self, rng):
        W, b = [], []
        k = self.in_dim
        for _ in range(self.layers):
            W.append((rng.standard_normal((k, self.hidden)) * np.sqrt(2.0 / k)).astype(np.float32))
            b.append(np.zeros(self.hidden, np.float32))
            k = self.hidden
        wo = (rng.standard_normal(self.hidden) * np.sqrt(1.0 / self.hidden)).astype(np.float32)
        self.set_weights(W, b, wo, 0.0)

    def set_weights(self, W, b, w_out, b_out):
        W = [np.ascontiguousarray(w, np.float32) for w in W]
        b = [np.ascontiguousarray(v, np.float32) for v in b]
        w_out = np.ascontiguousarray(w_out, np.float32)
        Wp = (FP * len(W))(*[_fp(w) for w in W])
        bp = (FP * len(b))(*[_fp(v) for v in b])
        rc = self.lib.orkan_mlp_set_weights(self.h, Wp, bp, _fp(w_out), float(b_out))
        if rc:
            raise RuntimeError(f"orkan_mlp_set_weights failed ({rc})")
        self._shapes = [w.shape for w in W]

    def get_weights(self):
        W = [np.empty(s, np.float32) for s in self._shapes]
        b = [np.empty(self.hidden, np.float32) for _ in W]
        wo = np.empty(self.hidden, np.float32)
        bo = C.c_float()
        rc = self.lib.orkan_mlp_get_weights(self.h, (FP * len(W))(*[_fp(w) for w in W]),
                                            (FP * len(b))(*[_fp(v) for v in b]), _fp(wo), C.byref(bo))
        if rc:
            raise RuntimeError(f"orkan_mlp_get_weights failed ({rc})")
        return W, b, wo, bo.value

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float32)
        y = np.empty(x.shape[0], np.float32)
        rc = self.lib.orkan_mlp_forward(self.h, x.shape[0], _fp(x), _fp(y))
        if rc:
            raise RuntimeError(f"orkan_mlp_forward failed ({rc})")
        return y

    def forward_candidates(self, gf, angles_dev, ncand, y_dev):
        gf = np.ascontiguousarray(gf, np.float32)
        rc = self.lib.orkan_mlp_forward_candidates(self.h, _fp(gf), gf.size, angles_dev, int(ncand), y_dev)
        if rc:
            raise RuntimeError(f"orkan_mlp_forward_candidates failed ({rc})")

    def train_step(self, x, t, lr):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_float()
        rc = self.lib.orkan_mlp_train_step(self.h, x.shape[0], _fp(x), _fp(t), float(lr), C.byref(loss))
        if rc:
            raise RuntimeError(f"orkan_mlp_train_step failed ({rc})")
        return loss.value

    def grads(self, x, t):
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        dW = [np.empty(s, np.float32) for s in self._shapes]
        db = [np.empty(self.hidden, np.float32) for _ in dW]
        dwo = np.empty(self.hidden, np.float32)
        dbo, loss = C.c_float(), C.c_float()
        rc = self.lib.orkan_mlp_grads(self.h, x.shape[0], _fp(x), _fp(t), (FP * len(dW))(*[_fp(w) for w in dW]),
                                      (FP * len(db))(*[_fp(v) for v in db]), _fp(dwo), C.byref(dbo), C.byref(loss))
        if rc:
            raise RuntimeError(f"orkan_mlp_grads failed ({rc})")
        return loss.value, (dW, db, dwo, dbo.value)

    def close(self):
        if self.h:
            self.lib.orkan_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---- features ---------------------------------------------------------------------------------
def graph_features(w):
    """edge weights sorted in descending order, zero-padded to 120 slots"""
    f = np.zeros(EDGE_SLOTS, np.float32)
    s = np.sort(np.asarray(w))[::-1][:EDGE_SLOTS]
    f[:s.size] = s
    return f


def features(gf, gammas, betas):
    """rows of [sorted weights (120), gamma_1..p / pi, beta_1..p / pi]"""
    gammas = np.atleast_2d(gammas)
    betas = np.atleast_2d(betas)
    m = gammas.shape[0]
    out = np.empty((m, IN_DIM), np.float32)
    out[:, :EDGE_SLOTS] = gf
    out[:, EDGE_SLOTS:EDGE_SLOTS + P] = gammas / np.pi
    out[:, EDGE_SLOTS + P:] = betas / np.pi
    return out
