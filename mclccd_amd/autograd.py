"""PyTorch autograd over the FP64 ACCD times of impact.

``toi_with_grad(tree, V0, V1, F)`` runs update + traverse on the GPU and returns
the TOIs of all CCD hits as a differentiable tensor (plus the hit stencils);
backward calls the library's hand-written adjoint kernel (``mclg_toi_grad``)
to produce dL/dV0 and dL/dV1.  The hit set is fixed by the forward pass;
the gradient is that of each hit's computed TOI wherever it is smooth.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from ._lib import lib


class TOIFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, V0, V1, tree, F):
        torch.cuda.current_stream().synchronize()
        tree.update(V0, V1, F)
        tree.run(V0, V1, F)
        sten, is_vf, toi = (x.clone() for x in tree.device_hits())
        ctx.tree = tree
        ctx.serial = tree.serial
        ctx.shape = V0.shape
        ctx.mark_non_differentiable(sten, is_vf)
        torch.cuda.current_stream().synchronize()
        return toi, sten, is_vf

    @staticmethod
    def backward(ctx, gtoi, _gs, _gv):
        tree = ctx.tree
        if tree.serial != ctx.serial:
            raise RuntimeError("the tree ran another update or traverse since this forward pass")
        g = gtoi.contiguous().to(torch.float64)
        dV0 = torch.empty(ctx.shape, dtype=torch.float64, device=g.device)
        dV1 = torch.empty(ctx.shape, dtype=torch.float64, device=g.device)
        torch.cuda.current_stream().synchronize()
        tree._check(lib.mclg_toi_grad(tree._h, C.c_void_p(g.data_ptr()), C.c_void_p(dV0.data_ptr()),
                                      C.c_void_p(dV1.data_ptr()), 1))
        return dV0, dV1, None, None


def toi_with_grad(tree, V0, V1, F):
    """(toi, stencils, is_vf) of every CCD hit; toi is differentiable w.r.t. V0, V1."""
    return TOIFunction.apply(V0, V1, tree, F)


def toi_grad(tree, w):
    """dL/dV0, dL/dV1 (numpy) of L = sum_i w_i toi_i for the tree's last traverse."""
    w = np.ascontiguousarray(w, np.float64)
    shape = (tree.nv, tree.dim)
    g0 = np.empty(shape)
    g1 = np.empty(shape)
    tree._check(lib.mclg_toi_grad(tree._h, w.ctypes.data_as(C.c_void_p), g0.ctypes.data_as(C.c_void_p),
                                  g1.ctypes.data_as(C.c_void_p), 0))
    return g0, g1
