"""The shipped cost model file (q1tsim_amd/data/costmodel_gfx950.npz, written
by q1tsim_amd/train_costmodel.py on an MI355X): its layout matches the
library's feature and parameter counts, it loads without unpickling, and its
recorded held-out errors are those of a usable ranking model."""
import json

import numpy as np

from oracle import mlp_ref as M
from q1tsim_amd import costmodel


def test_shipped_model_layout():
    with np.load(costmodel.DEFAULT_MODEL, allow_pickle=False) as z:
        assert int(z["nfeat"]) == costmodel.NFEAT == M.F
        assert z["params"].shape == (costmodel.NPARAMS,) and z["params"].dtype == np.float32
        assert z["mu"].shape == z["sigma"].shape == (M.F,)
        assert np.all(z["sigma"] > 0) and np.all(np.isfinite(z["params"]))
        meta = json.loads(str(z["meta"]))
    assert costmodel.NPARAMS == M.NPARAMS
    # trained on both kernel kinds, with thousands of recorded blocks
    assert meta["features"] == M.F and meta["training_blocks"] >= 5000
    for kind in ("interpreter", "specialised"):
        ho = meta["held_out"][kind]
        assert ho["samples"] >= 50
        assert ho["mean_rel_err"] < 0.25 and ho["mse_log_time"] < 0.05


def test_shipped_model_predicts_finite_times():
    with np.load(costmodel.DEFAULT_MODEL, allow_pickle=False) as z:
        p, mu, sg = (z[k].astype(np.float64) for k in ("params", "mu", "sigma"))
    x = mu + sg * np.random.default_rng(1).uniform(-1, 1, size=(256, M.F))
    y, _ = M.forward(p, x, mu, sg)
    # log(microseconds per 1024 tiles): finite and within a plausible range
    assert np.all(np.isfinite(y)) and np.all(np.abs(y) < 20)
