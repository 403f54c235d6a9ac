"""Torch-CPU restatement of the voxel element apply: the CPU baseline of the benchmark.

Test and benchmark infrastructure only; the product never imports it.

The reference's CPU path for ``PeriodicVoxelPhaseFieldOperator.mech_action``
(solid3d/operator.py:1108-1194, batches of ``batch_size`` voxels, model.py:775) is a
chain of torch ops per voxel batch.  It gathers the 24 element dofs, forms the strains of
the 8 Gauss points with one (B,24) x (24,48) product, adds the macro strain, multiplies by
each voxel's stiffness and the degradation, applies the transposed product, and
``scatter_add_``s the result onto the reduced dofs.  This class runs the same sequence on
torch CPU tensors over the arrays of an ``oracle.hex8.VoxelOperator``.  The benchmark times
it on the host cores; ``tests/test_oracle_torch_cpu.py`` pins it to the numpy oracle.
"""
from __future__ import annotations

import numpy as np
import torch


class TorchCpuVoxelApply:
    """``y = A(g(d)) x`` for x = [u~ (3n), E (6)] with g formed from nodal damage d."""

    def __init__(self, op, batch: int = 65536):
        self.n = op.n_red
        self.E = op.E
        self.batch = int(batch)
        self.kappa = float(op.kappa)
        self.dofs = torch.from_numpy(np.ascontiguousarray(op.dofs, dtype=np.int64))
        self.nodes = torch.from_numpy(np.ascontiguousarray(op.nodes, dtype=np.int64))
        self.Bf = torch.from_numpy(np.ascontiguousarray(op.B_flat))           # (24, 48)
        self.BfT = self.Bf.t().contiguous()
        self.Nt = torch.from_numpy(np.ascontiguousarray(op.N.T))              # (8 nodes, 8 qp)
        self.w = torch.from_numpy(np.ascontiguousarray(op.w))                 # (8,)
        self.bank = torch.from_numpy(np.ascontiguousarray(op.bank))           # (K, 6, 6), symmetric
        self.index = torch.from_numpy(np.ascontiguousarray(op.index, dtype=np.int64))

    def apply(self, x: torch.Tensor, d: torch.Tensor) -> torch.Tensor:
        n3 = 3 * self.n
        u, Em = x[:n3], x[n3:]
        r_u = torch.zeros(n3, dtype=x.dtype)
        r_E = torch.zeros(6, dtype=x.dtype)
        k = self.kappa
        for s in range(0, self.E, self.batch):
            sl = slice(s, min(s + self.batch, self.E))
            eps = (u[self.dofs[sl]] @ self.Bf).view(-1, 8, 6) + Em           # strain at 8 Gauss points
            d_q = d[self.nodes[sl]] @ self.Nt                                 # damage at the Gauss points
            gw = ((1.0 - k) * (1.0 - d_q) ** 2 + k) * self.w                 # g(d_q) times the weight
            sig = torch.bmm(eps, self.bank[self.index[sl]]) * gw.unsqueeze(-1)
            r_E += sig.sum(dim=(0, 1))
            r_u.index_add_(0, self.dofs[sl].reshape(-1), (sig.reshape(-1, 48) @ self.BfT).reshape(-1))
        return torch.cat((r_u, r_E))
