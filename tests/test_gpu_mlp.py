"""The BF16 MFMA surrogate kernels: GEMM exactness on bf16-exact operands (pins
the MFMA operand and accumulator lane layout), forward and backward against the
CPU oracle (oracle/mlp_np.py, bf16 variant), and training that lowers the loss."""
import os
import sys

import numpy as np
import pytest

import orkan_amd
from orkan_amd import surrogate

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_np  # noqa: E402

pytestmark = pytest.mark.gpu


def to_bf16_bits(a):
    u = np.ascontiguousarray(a, np.float32).view(np.uint32)
    return ((u + ((u >> 16) & 1) + np.uint32(0x7FFF)) >> 16).astype(np.uint16)


def upload(lib, arr):
    raw = np.ascontiguousarray(arr)
    p = lib.orkan_dev_alloc((raw.nbytes + 7) // 8)
    assert p and lib.orkan_copy_to_device(p, raw.ctypes.data, raw.nbytes) == 0
    return p


GEMM_TILE_HOOK = 11


@pytest.mark.parametrize("variant", [0, 1, 2])
@pytest.mark.parametrize("M,N,K", [(128, 128, 32), (256, 384, 96), (512, 128, 512), (256, 256, 64),
                                   (512, 512, 512), (256, 512, 128), (768, 256, 1024)])
def test_gemm_exact_small_integers(M, N, K, variant):
    """Every tile variant the shape admits (128 x 128; 256 x 256 register-staged when M, N % 256
    and K % 64; 256 x 256 DMA-staged when M, N % 256 and K % 128; otherwise the next smaller)."""
    lib = orkan_amd.load(5)
    assert lib.orkan_set_launch_shape(GEMM_TILE_HOOK, variant, 0) == 0
    rng = np.random.default_rng(M + N + K)
    A = rng.integers(-3, 4, (M, K)).astype(np.float32)
    B = rng.integers(-3, 4, (N, K)).astype(np.float32)
    pa, pb = upload(lib, to_bf16_bits(A)), upload(lib, to_bf16_bits(B))
    pc = lib.orkan_dev_alloc(M * N // 2 + 1)
    assert lib.orkan_mlp_gemm_test(M, N, K, pa, pb, pc) == 0
    got = np.empty(M * N, np.float32)
    assert lib.orkan_copy_to_host(got.ctypes.data, pc, got.nbytes) == 0
    assert lib.orkan_set_launch_shape(GEMM_TILE_HOOK, 2, 0) == 0
    assert np.array_equal(got.reshape(M, N), (A @ B.T).astype(np.float32))
    for p in (pa, pb, pc):
        lib.orkan_dev_free(p)


def _params(rng):
    W, b, wo, bo = mlp_np.init_params(rng, 128, 512, 3)
    b = [rng.standard_normal(512) * 0.1 for _ in b]
    return [w.astype(np.float32) for w in W], [v.astype(np.float32) for v in b], wo.astype(np.float32), 0.25


def test_forward_backward_match_oracle():
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(7)
    W, b, wo, bo = _params(rng)
    m = surrogate.Mlp(lib, max_batch=512)
    m.set_weights(W, b, wo, bo)
    x = rng.standard_normal((300, 128)).astype(np.float32)
    t = rng.standard_normal(300).astype(np.float32)
    params = ([w.astype(np.float64) for w in W], [v.astype(np.float64) for v in b], wo.astype(np.float64), bo)
    y_ref, _, _ = mlp_np.forward(params, x.astype(np.float64), exact=False)
    y = m.forward(x)
    assert np.abs(y - y_ref).max() <= 2e-3 * max(1.0, np.abs(y_ref).max())
    loss, (dW, db, dwo, dbo) = m.grads(x, t)
    lref, (rW, rb, rwo, rbo) = mlp_np.loss_and_grads(params, x.astype(np.float64), t.astype(np.float64), exact=False)
    assert abs(loss - lref) <= 1e-3 * lref

    def rel(a, r):
        return float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-30))

    for l in range(3):
        assert rel(dW[l], rW[l]) < 2e-2, (l, rel(dW[l], rW[l]))
        assert rel(db[l], rb[l]) < 2e-2, (l, rel(db[l], rb[l]))
    assert rel(dwo, rwo) < 1e-2
    assert abs(dbo - rbo) <= 1e-3 * max(1.0, abs(rbo))
    m.close()


def test_training_lowers_loss():
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(11)
    m = surrogate.Mlp(lib, max_batch=1024)
    m.init(rng)
    X = rng.uniform(0, 1, (8192, 128)).astype(np.float32)
    T = (np.sin(3 * X[:, 0]) * X[:, 1] + 0.5 * X[:, 2:6].sum(1)).astype(np.float32)
    first = None
    for step in range(300):
        idx = rng.integers(0, X.shape[0], 1024)
        loss = m.train_step(X[idx], T[idx], 1e-3)
        first = loss if first is None else first
    final = float(np.mean((m.forward(X[:2048]) - T[:2048]) ** 2))
    assert final < 0.05 * first, (first, final)
    m.close()


@pytest.mark.parametrize("n,k", [(1000, 1), (1000, 64), (1 << 20, 64), (5000, 5000)])
def test_topk_matches_numpy(n, k):
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(n + k)
    y = rng.standard_normal(n).astype(np.float32)
    y[rng.integers(0, n, n // 10)] = 0.5          # ties
    y[: min(n, 7)] = -0.0                          # signed zeros
    p = upload(lib, y)
    idx = np.empty(k, np.int32)
    val = np.empty(k, np.float32)
    import ctypes as C
    assert lib.orkan_topk_f32(p, n, k, idx.ctypes.data_as(C.POINTER(C.c_int)),
                              val.ctypes.data_as(C.POINTER(C.c_float))) == 0
    order = np.lexsort((np.arange(n), -y))          # value descending, index ascending
    want_vals = y[order[:k]]
    assert np.array_equal(val, want_vals)
    assert np.array_equal(y[idx], val)
    assert len(set(idx.tolist())) == k
    lib.orkan_dev_free(p)


@pytest.mark.parametrize("ncand,max_batch", [(300, 512), (1000, 256), (129, 128)])
def test_forward_candidates_matches_forward(ncand, max_batch):
    """Screening forward (graph share of layer 1 computed once, angles per row,
    candidates in chunks of max_batch with a ragged last chunk) against the CPU
    oracle and against the plain forward on the assembled rows."""
    lib = orkan_amd.load(5)
    rng = np.random.default_rng(ncand)
    W, b, wo, bo = _params(rng)
    m = surrogate.Mlp(lib, max_batch=max_batch)
    m.set_weights(W, b, wo, bo)
    gf = surrogate.graph_features(rng.uniform(0, 1, 60))
    ang = rng.uniform(0, 1, (ncand, 8)).astype(np.float32)
    x = np.concatenate([np.broadcast_to(gf, (ncand, 120)), ang], 1).astype(np.float32)
    params = ([w.astype(np.float64) for w in W], [v.astype(np.float64) for v in b], wo.astype(np.float64), bo)
    y_ref, _, _ = mlp_np.forward(params, x.astype(np.float64), exact=False)
    dang = upload(lib, ang)
    # one guard float past the outputs: the forward must not write beyond ncand
    ydev = lib.orkan_dev_alloc((ncand * 4 + 4 + 7) // 8)
    guard = np.full(ncand + 1, 12345.0, np.float32)
    assert lib.orkan_copy_to_device(ydev, guard.ctypes.data, guard.nbytes) == 0
    m.forward_candidates(gf, dang, ncand, ydev)
    y = np.empty(ncand + 1, np.float32)
    assert lib.orkan_copy_to_host(y.ctypes.data, ydev, y.nbytes) == 0
    assert y[ncand] == 12345.0
    y = y[:ncand]
    scale = max(1.0, np.abs(y_ref).max())
    assert np.abs(y - y_ref).max() <= 2e-3 * scale
    assert np.abs(y - m.forward(x)).max() <= 2e-3 * scale
    lib.orkan_dev_free(dang)
    lib.orkan_dev_free(ydev)
    m.close()
