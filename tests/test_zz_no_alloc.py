"""mclg_toi_grad_noalloc (the adjoint with its scratch in the library's static
device array) against mclg_toi_grad on the same traverse, through the C ABI:
per entry within 1e-12 relative (plus 1e-15 of the largest entry: the two
differ only in the order of the floating-point sums), on two calls in a row,
the second on a smaller scene, and its refusal of a hit list larger than the
static scratch."""
import ctypes as C

import numpy as np
import pytest

from mclccd_amd import scenes
from tests.parity import grad_close

pytestmark = pytest.mark.gpu


def _tree(V0, V1, F, th):
    import torch

    from mclccd_amd.tree import BVHTree

    t = BVHTree(V0.shape[1])
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    dev = torch.device("cuda:0")
    d = [torch.from_numpy(x).to(dev) for x in (V0, V1, F)]
    t.update(*d)
    t.run(*d)
    return t


def _grads(t, w, noalloc):
    import torch

    from mclccd_amd._lib import lib

    g0 = torch.full((t.nv, t.dim), 7.0, dtype=torch.float64, device=w.device)  # stale values must be cleared
    g1 = torch.full_like(g0, -3.0)
    torch.cuda.synchronize()
    p = [C.c_void_p(x.data_ptr()) for x in (w, g0, g1)]
    if noalloc:
        t._check(lib.mclg_toi_grad_noalloc(t._h, *p))
    else:
        t._check(lib.mclg_toi_grad(t._h, *p, 1))
    return np.concatenate([g0.cpu().numpy().ravel(), g1.cpu().numpy().ravel()])


@pytest.mark.parametrize("first,second", [("layers", "cloth"), ("cloth", "disk")])
def test_noalloc_adjoint_matches_allocating_one(gpu, first, second):
    import torch

    cases = {"layers": lambda: scenes.stacked_layers(n=24),
             "cloth": lambda: scenes.cloth_fold(n=48),
             "disk": lambda: scenes.disk2d(npts=4000)}
    prev = None
    for name in (first, second):
        V0, V1, F, th = cases[name]()
        t = _tree(V0, V1, F, th)
        n = t.counts()["hits"]
        assert 100 < n <= (1 << 20), n
        if prev is not None:
            assert n < prev  # the second call sees a smaller hit list than the array's last contents
        prev = n
        gen = torch.Generator(device="cuda:0").manual_seed(3)
        w = torch.rand(n, dtype=torch.float64, device="cuda:0", generator=gen)
        ref = _grads(t, w, noalloc=False)
        got = _grads(t, w, noalloc=True)
        assert np.abs(ref).max() > 0
        assert grad_close(got, ref, rtol=1e-12, floor=1e-15) <= 1.0
        again = _grads(t, w, noalloc=True)
        assert grad_close(again, ref, rtol=1e-12, floor=1e-15) <= 1.0


def test_noalloc_adjoint_rejects_more_hits_than_its_scratch(gpu):
    import torch

    from mclccd_amd import _lib
    from mclccd_amd._lib import lib

    V0, V1, F, th = scenes.stacked_layers(n=60)
    t = _tree(V0, V1, F, th)
    n = t.counts()["hits"]
    assert n > _lib.NOALLOC_MAX_HITS, n
    w = torch.ones(n, dtype=torch.float64, device="cuda:0")
    g = torch.zeros((2, t.nv, 3), dtype=torch.float64, device="cuda:0")
    rc = lib.mclg_toi_grad_noalloc(t._h, C.c_void_p(w.data_ptr()), C.c_void_p(g[0].data_ptr()),
                                   C.c_void_p(g[1].data_ptr()))
    assert rc == _lib.ERR_CAPACITY
    assert b"exceed" in lib.mclg_last_error(t._h)
    assert float(g.abs().max()) == 0.0  # nothing written
