"""The torch-CPU apply of the benchmark's CPU baseline equals the numpy oracle (CPU)."""
import numpy as np
import torch

from oracle import hex8 as O
from oracle import torch_cpu as TC
from tests.test_oracle_hex8 import _op


def test_torch_cpu_apply_equals_the_oracle():
    op = _op(shape=(5, 4, 6), h=(1.0, 1.2, 0.8))
    rng = np.random.default_rng(7)
    x = rng.uniform(-1e-3, 1e-3, op.n_mech)
    d = rng.uniform(0, 1, op.n_red)
    ref = op.mech_action(x, op.degradation(d))
    for batch in (7, 65536):  # several voxel batches and one
        got = TC.TorchCpuVoxelApply(op, batch=batch).apply(torch.from_numpy(x), torch.from_numpy(d)).numpy()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
