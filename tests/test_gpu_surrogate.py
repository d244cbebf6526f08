"""The screening pipeline end to end at small scale: simulator-labelled data,
GPU training, held-out error, and screened search reaching the plain search's
ratio with fewer exact evaluations."""
import pytest

import orkan_amd
from orkan_amd import bench_surrogate

pytestmark = pytest.mark.gpu


def test_surrogate_pipeline_small():
    lib = orkan_amd.load(5)
    r = bench_surrogate.run(lib, train_graphs=64, per_graph=512, test_graphs=2, ncand=20000, topk=64, epochs=3,
                            batch_size=1024, plain_batch=2048, rounds=1, ref_cand=4096, ref_top=32, ref_steps=50,
                            plain_cap=20000)
    assert r["train_samples_uniform"] == 64 * 512 and r["refined_labels"] == 64 * 32
    assert r["heldout_mae_normalised"] < 0.05
    for row in r["search"]:
        assert 0.0 < row["best_ratio_screened"] <= 1.0 + 1e-12
        assert row["plain_evals_run"] >= 1
