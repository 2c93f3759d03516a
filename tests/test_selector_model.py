"""The persisted tile-selection model (numrs_amd/selector_model.npz): loads without pickles and has
the layout the library's nrk_selector_set_model expects (no GPU needed)."""
import numpy as np
import pytest

from numrs_amd import selector as S


def test_persisted_model_layout():
    if not S.MODEL_PATH.exists():
        pytest.skip("no persisted selector model in this tree")
    z = np.load(S.MODEL_PATH, allow_pickle=False)
    K = int(z["K"])
    assert K == 5
    assert z["params"].shape == (S.H * S.IN + S.H + S.H * S.H + S.H + K * S.H + K,)
    assert z["params"].dtype == np.float32 and np.all(np.isfinite(z["params"]))
    assert z["x_mean"].shape == (S.IN,) and np.all(z["x_scale"] > 0)
    # trained on the ragged-batch GEMM shapes: m, n, k in [16, 512]
    assert np.all(z["domain_lo"] >= 4.0) and np.all(z["domain_hi"] <= 9.0 + 1e-6)
    assert float(z["margin"]) >= 0.0
