"""TOI adjoint at full size: 16 stacked 180x180 cloth layers (1,036,800
triangles, about 8.3e8 CCD hits), w_i ~ U(0,1).

(a) The GPU gradient at 1,000 seeded random coordinates of V0 and V1 equals
    the oracle's dual-number gradient summed over every hit touching that
    vertex (per entry: relative 1e-6, plus 1e-12 of the largest sampled entry).
(b) Central differences (h = 1e-7) of the oracle's forward ACCD agree with
    the GPU gradient (relative 1e-6) on 1,000 seeded hits, one coordinate
    each; w is non-zero on a vertex-disjoint subset of them so each checked
    coordinate's gradient comes from one hit.  A sample is skipped when the
    +-h evaluations take a different branch (l_p argmax, closest-point region
    or clamp, step count): the rest pose is 16 identical flat grids, so many
    edge pairs start exactly parallel, where the TOI has a kink.
"""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as orc
from tests.parity import grad_close

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def full_scene(gpu):
    import torch

    from mclccd_amd.tree import BVHTree

    V0, V1, F, th = scenes.stacked_layers()
    dev = torch.device("cuda:0")
    t = BVHTree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    tV0, tV1, tF = (torch.from_numpy(x).to(dev) for x in (V0, V1, F))
    t.update(tV0, tV1, tF)
    t.run(tV0, tV1, tF)
    yield t, V0, V1, F, th
    del t
    torch.cuda.empty_cache()


def _grad(t, w):
    import ctypes as C

    import torch

    from mclccd_amd._lib import lib

    dV0 = torch.empty((t.nv, 3), dtype=torch.float64, device=w.device)
    dV1 = torch.empty_like(dV0)
    torch.cuda.synchronize()
    t._check(lib.mclg_toi_grad(t._h, C.c_void_p(w.data_ptr()), C.c_void_p(dV0.data_ptr()),
                               C.c_void_p(dV1.data_ptr()), 1))
    return dV0.cpu().numpy(), dV1.cpu().numpy()


def test_full_size_adjoint_vs_dual_numbers(full_scene):
    import torch

    t, V0, V1, F, th = full_scene
    sten, is_vf, toi = t.device_hits()
    n = len(toi)
    assert n > 5e8
    gen = torch.Generator(device="cuda:0").manual_seed(4)
    w = torch.rand(n, dtype=torch.float64, device="cuda:0", generator=gen)
    g0, g1 = _grad(t, w)
    rng = np.random.default_rng(4)
    which = rng.integers(0, 2, 1000)
    vi = rng.integers(0, len(V0), 1000)
    k = rng.integers(0, 3, 1000)
    mark = torch.zeros(len(V0) + 1, dtype=torch.bool, device="cuda:0")
    mark[torch.from_numpy(np.unique(vi)).to("cuda:0")] = True
    parts = []
    for lo in range(0, n, 1 << 26):  # hits touching a sampled vertex, in chunks
        blk = sten[lo:lo + (1 << 26)].long()
        blk[blk < 0] = len(V0)
        parts.append(mark[blk].any(dim=1).nonzero().squeeze(1) + lo)
        del blk
    touch = torch.cat(parts)
    s_h = sten[touch].cpu().numpy()
    v_h = is_vf[touch].cpu().numpy()
    w_h = w[touch].cpu().numpy()
    del w, touch
    r0, r1 = orc.toi_grad_list(V0, V1, s_h, v_h, w_h, th, th)
    got = np.where(which == 0, g0[vi, k], g1[vi, k])
    ref = np.where(which == 0, r0[vi, k], r1[vi, k])
    assert np.count_nonzero(ref) > 900
    # per entry: 1e-6 relative, plus 1e-12 of the largest sampled entry (summation order near zero)
    worst = grad_close(got, ref)
    assert worst <= 1.0, worst


def test_full_size_adjoint_vs_central_differences(full_scene):
    import torch

    t, V0, V1, F, th = full_scene
    sten, is_vf, toi = t.device_hits()
    rng = np.random.default_rng(7)
    pick = torch.from_numpy(rng.choice(len(toi), 1000, replace=False)).to("cuda:0")
    s_p, v_p, t_p = sten[pick].cpu().numpy(), is_vf[pick].cpu().numpy(), toi[pick].cpu().numpy()
    used = set()
    keep = []
    for j, st in enumerate(s_p):
        vs = [int(x) for x in st if x >= 0]
        if used.isdisjoint(vs):
            used.update(vs)
            keep.append(j)
    keep = np.array(keep)
    wv = rng.uniform(0.5, 1.5, len(keep))
    w = torch.zeros(len(toi), dtype=torch.float64, device="cuda:0")
    w[pick[torch.from_numpy(keep).to("cuda:0")]] = torch.from_numpy(wv).to("cuda:0")
    g0, g1 = _grad(t, w)
    del w
    h = 1e-7
    checked = skipped = 0
    worst = 0.0
    for j, wj in zip(keep, wv):
        st = s_p[j][s_p[j] >= 0]
        nv = len(st)
        arr, m, k = int(rng.integers(0, 2)), int(rng.integers(0, nv)), int(rng.integers(0, 3))
        g = (g0 if arr == 0 else g1)[st[m], k]
        A = np.stack([V0[st]] * 3)
        B = np.stack([V1[st]] * 3)
        (A if arr == 0 else B)[1, m, k] += h
        (A if arr == 0 else B)[2, m, k] -= h
        ss = np.array([[c * nv + x for x in range(nv)] + [-1] * (4 - nv) for c in range(3)], np.int32)
        tt, sig = orc.toi_list(A.reshape(-1, 3), B.reshape(-1, 3), ss, np.repeat(v_p[j], 3), th, th,
                               signature=True)
        assert tt[0] == t_p[j]  # the oracle's forward TOI equals the GPU's bit for bit
        if not (sig[0] == sig[1] == sig[2]) or tt[1] < 0 or tt[2] < 0:
            skipped += 1
            continue
        fd = wj * (tt[1] - tt[2]) / (2 * h)
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
        checked += 1
    assert checked >= 500, (checked, skipped)
    assert worst <= 1e-6, worst
