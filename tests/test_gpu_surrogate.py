"""The screening pipeline end to end: simulator-labelled data, GPU training,
held-out error, and what the trained network gains over plain batched search
on the same candidate pool (exact expected number of plain evaluations needed
to reach the ratio the screened search finds with 64 exact evaluations)."""
import numpy as np
import pytest

import orkan_amd
from benchmarks import screening

pytestmark = pytest.mark.gpu


def test_surrogate_gain_over_plain_search():
    lib = orkan_amd.load(5)
    topk, ncand = 64, 65536
    r = screening.run(lib, train_graphs=512, per_graph=2048, test_graphs=4, ncand=ncand, topk=topk,
                      epochs=16, batch_size=4096, plain_batch=65536)
    print({k: v for k, v in r.items() if k != "search"})
    for row in r["search"]:
        print(row)
    assert r["train_samples"] == 512 * 2048
    assert r["heldout_mae_over_target_std"] < 0.6
    assert r["heldout_spearman"] > 0.3
    evals = []
    for row in r["search"]:
        assert 0.0 < row["best_ratio_screened"] <= row["pool_best_ratio"] + 1e-12 <= 1.0 + 2e-12
        assert 1 <= row["pool_at_or_above_screened"] <= ncand
        evals.append(row["plain_expected_evals"])
    # the 64 exact evaluations spent on the network's picks are worth more than
    # ten times as many evaluations of the plain search
    assert np.median(evals) > 10 * topk
