"""The optimizer-vs-single-scheme experiment (thesis tex:1276-1304) on
regime-shifting columns: the stateful optimizer's tree and blob equal the
oracle's CompressData restatement part after part, every blob decodes, and
the experiment's report is consistent."""
import numpy as np
import pytest

import gputimeseriescompressionoptimizer_amd as g
from gputimeseriescompressionoptimizer_amd import experiments as E
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", ["A+B", "Time+A", "A+B+MaxPrec"])
def test_regime_shifting_parts_match_oracle(ctx, name):
    dtype, gens = E.data_sets(3000)[name]
    parts = E.make_parts(gens, 22, 5, seed=3)
    opt = g.CompressionOptimizer(ctx)
    ref = O.Optimizer()
    for i, p in enumerate(parts):
        print(f"[experiment] {name} part {i}", flush=True)
        blob = opt.CompressData(ctx.upload(p), dtype).to_bytes()
        want = ref.compress_data(p, dtype)
        assert opt.GetOptimalTree() == ref.tree.preorder(), i
        assert blob == want, i
        assert g.CompressionTree(ctx, None, dtype).Decompress(ctx.upload(blob)).to_bytes() == p.tobytes(), i


def test_experiment_report(ctx):
    res = E.optimizer_vs_single_scheme(ctx, size=2000, parts=20, per_regime=5)
    for name, r in res.items():
        if name == "mean_gain":
            continue
        assert r["bit_exact"], name
        assert r["gain"] == pytest.approx(r["optimizer_ratio"] / r["single_scheme_ratio"], rel=1e-2)
    assert res["mean_gain"] == pytest.approx(np.mean([r["gain"] for k, r in res.items() if k != "mean_gain"]),
                                             rel=1e-3)
