"""Dataset.X (pyLOM/dataset.py:252-292) on the device against the oracle:
bit for bit (a gather)."""
import numpy as np
import pytest
import torch

import oracle

pytestmark = pytest.mark.gpu
pl = pytest.importorskip("pyloworder_amd")


@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_dataset_x_matches_oracle(dt):
    rng = np.random.default_rng(1)
    npts, nt = 5003, 37
    f = {"PRESS": (1, rng.standard_normal((npts, nt)).astype(dt)),
         "VELOC": (3, rng.standard_normal((3 * npts, nt)).astype(dt)),
         "TEMP": (1, rng.standard_normal((npts, nt)).astype(dt))}
    d = pl.Dataset(npts, {k: {"ndim": v[0], "value": v[1]} for k, v in f.items()})
    for names in ([], ["VELOC"], ["TEMP", "PRESS"], ["VELOC", "PRESS", "TEMP"]):
        X = d.X(*names)
        Xo = oracle.dataset_x(f, npts, names or None)
        assert X.dtype == dt and np.array_equal(X.astype(np.float64), Xo)


def test_dataset_x_feeds_pod():
    rng = np.random.default_rng(2)
    V = torch.from_numpy(rng.standard_normal((2 * 4000, 50))).cuda()
    d = pl.Dataset(4000, {"U": {"ndim": 2, "value": V}})
    X = d.X("U")
    assert isinstance(X, torch.Tensor) and torch.equal(X, V)
    U, S, Vt = pl.POD.run(X, remove_mean=True)
    assert S.shape[0] == 50
