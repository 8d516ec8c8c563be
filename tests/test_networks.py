"""The shipped warm-start networks (akuaengine_amd/networks) are well-formed: the parameter blob
has the documented size and finite values, positive scales, a recipe file that matches, and the
FP64 restatement of the network gives finite outputs on features of the training distribution."""
import glob
import json
import os

import numpy as np
import pytest

from oracle import mlp_oracle as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NETS = sorted(glob.glob(os.path.join(ROOT, "akuaengine_amd", "networks", "*.npy")))


def test_networks_are_shipped():
    assert NETS, "no trained network files"


@pytest.mark.parametrize("path", NETS, ids=os.path.basename)
def test_network_blob_and_recipe(path):
    from akuaengine_amd import learned
    p = learned.load_network(path)
    assert p.shape == (M.NPARAMS,) and p.dtype == np.float32
    assert np.isfinite(p).all()
    q = M.unpack(p)
    assert (q["inv"] > 0).all() and (q["ymul"] > 0).all()
    with open(os.path.splitext(path)[0] + ".json") as f:
        meta = json.load(f)
    J = int(path.rsplit("_J", 1)[1].split(".")[0])
    assert meta["iterations"] == J
    assert 0.0 < meta["gate_ratio_p90"] < 2.0 and meta["loss_last50"] < meta["loss_first50"] * 1.5
    # features near the training distribution: the normalisation's own mean +- 2 sigma
    rng = np.random.default_rng(0)
    F = q["mean"] + rng.uniform(-2, 2, size=(256, 16)) / q["inv"]
    y = M.forward(p, F, exact=True)
    assert np.isfinite(y).all()


def test_load_rejects_a_wrong_blob(tmp_path):
    from akuaengine_amd import learned
    bad = tmp_path / "bad.npy"
    np.save(bad, np.zeros(10, np.float32))
    with pytest.raises(ValueError):
        learned.load_network(str(bad))
