"""Host-side argument checks of the autograd wrapper (no GPU needed): every
vector handed to the C-ABI is float64, on the engine's GPU and holds exactly
the real particles; the kernels index it without bounds checks."""
import types

import numpy as np
import pytest
import torch

from sph_playground_amd import autograd as AG


def _eng(n=10, device=0):
    return types.SimpleNamespace(n_real=n, device=device)


@pytest.mark.parametrize("t,msg", [
    (np.zeros(10), "torch tensor"),
    (torch.zeros(10, dtype=torch.float32), "float64"),
    (torch.zeros(10, dtype=torch.float64), "cuda:0"),
])
def test_vector_rejects_bad_inputs(t, msg):
    with pytest.raises(ValueError, match=msg):
        AG._vector(_eng(), t, "T")


def test_tensor6_shape():
    assert AG._tensor6([1, 2, 3, 4, 5, 6]).shape == (6,)
    with pytest.raises(ValueError):
        AG._tensor6([1, 2, 3])
