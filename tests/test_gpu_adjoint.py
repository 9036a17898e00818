"""TOI adjoint on the GPU vs (a) the oracle's gradient by forward-mode dual
numbers over every hit (per entry: within 1e-6 relative, plus 1e-12 of the
largest entry for entries near zero), (b) central
finite differences (h = 1e-7) of the oracle's forward ACCD on 1,000 random
coordinates (relative 1e-6), (c) torch autograd."""
import numpy as np
import pytest

from mclccd_amd import scenes
from oracle import oracle as orc
from tests.parity import grad_close, quirk_stencils

pytestmark = pytest.mark.gpu


def gpu_run(V0, V1, F, th, strict_rank=None):
    from mclccd_amd.tree import BVHTree

    t = BVHTree(V0.shape[1])
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    t.update(V0, V1, F)
    if strict_rank is not None:
        t.set_rank(strict_rank)
    t.run(V0, V1, F)
    return t


def quirk_scene(count=64):
    """Pairs (static-ish triangle, small moving triangle) whose moving vertex
    starts in the reference's point_on_triangle quirk region; pairs are
    spaced 20 apart."""
    V0, V1, F = [], [], []
    for i, (v0, v1) in enumerate(quirk_stencils(count)):
        off = np.array([20.0 * i, 0.0, 0.0])
        d = v1[0] - v0[0]
        q0 = np.array([v0[0], v0[0] + [0.0, 0.0, 3.0], v0[0] + [0.0, 3.0, 0.0]])
        base = 6 * i
        V0 += list(v0[1:] + off) + list(q0 + off)
        V1 += list(v1[1:] + off) + list(q0 + d + off)
        F += [[base, base + 1, base + 2], [base + 3, base + 4, base + 5]]
    return np.array(V0), np.array(V1), np.array(F, np.int32), 1e-2


@pytest.mark.parametrize("case", ["cloth", "layers", "disk", "soup", "quirk"])
def test_adjoint_matches_dual_numbers(gpu, case):
    from mclccd_amd.autograd import toi_grad

    if case == "cloth":
        V0, V1, F, th = scenes.cloth_fold()
    elif case == "layers":
        V0, V1, F, th = scenes.stacked_layers(n=20)
    elif case == "disk":
        V0, V1, F, th = scenes.disk2d(npts=3000)
    elif case == "soup":
        V0, V1, F, th = scenes.random_soup(300, seed=1, share=True, motion=0.2)
        th = 1e-2
    else:
        V0, V1, F, th = quirk_scene()
    o = orc.run_scene(V0, V1, F, th)
    t = gpu_run(V0, V1, F, th, strict_rank=o.rank())
    s, v, toi = t.hits()
    assert len(s) > (50 if case == "quirk" else 100)
    w = np.random.default_rng(4).uniform(0, 1, len(s))
    g0, g1 = toi_grad(t, w)
    r0, r1 = orc.toi_grad_list(V0, V1, s, v, w, th, th)
    assert max(np.abs(r0).max(), np.abs(r1).max()) > 0
    ref, got = np.concatenate([r0.ravel(), r1.ravel()]), np.concatenate([g0.ravel(), g1.ravel()])
    assert grad_close(got, ref) <= 1.0


def test_adjoint_first_and_later_frames(gpu):
    # 66 % of this soup's candidates hit, more than a fresh context's first hit
    # list holds (half the candidates): the first traverse builds its list in
    # the compaction pass (hits in candidate order), the second appends it
    # directly from the ACCD waves (another order).  Both gradients match the
    # oracle's.
    from mclccd_amd.autograd import toi_grad

    V0, V1, F, _ = scenes.random_soup(300, seed=1, share=True, motion=0.2)
    th = 1e-2
    o = orc.run_scene(V0, V1, F, th)
    t = gpu_run(V0, V1, F, th, strict_rank=o.rank())
    c = t.counts()
    assert c["hits"] > c["candidates"] // 2
    for frame in range(2):
        if frame:
            t.run(V0, V1, F)
        s, v, _ = t.hits()
        w = np.random.default_rng(6).uniform(0, 1, len(s))
        g0, g1 = toi_grad(t, w)
        r0, r1 = orc.toi_grad_list(V0, V1, s, v, w, th, th)
        ref, got = np.concatenate([r0.ravel(), r1.ravel()]), np.concatenate([g0.ravel(), g1.ravel()])
        assert grad_close(got, ref) <= 1.0, frame


def test_adjoint_central_differences(gpu):
    # w is non-zero on a vertex-disjoint set of hits, so each coordinate's
    # gradient comes from one hit; FD re-evaluates that hit's TOI at +-h
    from mclccd_amd.autograd import toi_grad

    # general position: the grid rest pose of cloth_fold has exact parallel
    # edges and tied displacement norms, i.e. kinks of the TOI
    V0, V1, F, th = scenes.cloth_fold()
    rg = np.random.default_rng(5)
    V0 = V0 + rg.normal(0, 1e-4, V0.shape)
    V1 = V1 + rg.normal(0, 1e-3, V1.shape)
    t = gpu_run(V0, V1, F, th)
    s, v, toi = t.hits()
    # the hit list's order varies from run to run (leaf pairs are flushed
    # through a counter): sample from a canonical order so the test is the same
    order = np.lexsort((s[:, 3], s[:, 2], s[:, 1], s[:, 0], v))
    rng = np.random.default_rng(7)
    used = np.zeros(len(V0), bool)
    owner = np.full(len(V0), -1, np.int64)
    w = np.zeros(len(s))
    for i in order[rng.permutation(len(s))]:
        st = s[i][s[i] >= 0]
        if used[st].any():
            continue
        used[st] = True
        owner[st] = i
        w[i] = rng.uniform(0.5, 1.5)
    assert (w > 0).sum() > 500
    g0, g1 = toi_grad(t, w)
    assert np.all(g0[~used] == 0) and np.all(g1[~used] == 0)
    h = 1e-7
    verts = np.nonzero(used)[0]
    checked = kinks = 0
    worst = 0.0
    for _ in range(1000):
        which = int(rng.integers(0, 2))
        vi = int(rng.choice(verts))
        k = int(rng.integers(0, 3))
        i = owner[vi]
        g = (g0 if which == 0 else g1)[vi, k]
        A0, A1 = V0.copy(), V1.copy()
        (A0 if which == 0 else A1)[vi, k] += h
        tp, sp = orc.toi_list(A0, A1, s[i:i + 1], v[i:i + 1], th, th, signature=True)
        A0, A1 = V0.copy(), V1.copy()
        (A0 if which == 0 else A1)[vi, k] -= h
        tm, sm = orc.toi_list(A0, A1, s[i:i + 1], v[i:i + 1], th, th, signature=True)
        _, s0 = orc.toi_list(V0, V1, s[i:i + 1], v[i:i + 1], th, th, signature=True)
        tp, tm = tp[0], tm[0]
        # skip when the hit is lost, or when +-h takes another branch (closest-point
        # region or clamp, l_p argmax, step count): a kink of the TOI within +-h
        if tp < 0 or tm < 0 or not (sp[0] == sm[0] == s0[0]):
            kinks += 1
            continue
        fd = w[i] * (tp - tm) / (2 * h)
        worst = max(worst, abs(fd - g) / max(1.0, abs(g)))
        checked += 1
    assert checked >= 950, (checked, kinks)
    # central differences in FP64 carry ~1e-9 absolute noise at h = 1e-7
    assert worst <= 1e-6, worst


def test_autograd_function(gpu):
    import torch

    from mclccd_amd.autograd import toi_with_grad
    from mclccd_amd.tree import BVHTree

    V0, V1, F, th = scenes.cloth_fold(n=24)
    dev = torch.device("cuda:0")
    tV0 = torch.from_numpy(V0).to(dev).requires_grad_(True)
    tV1 = torch.from_numpy(V1).to(dev).requires_grad_(True)
    tF = torch.from_numpy(F).to(dev)
    t = BVHTree(3)
    t.options.box_eta = t.options.vf_ccd_eta = t.options.ee_ccd_eta = th
    toi, sten, is_vf = toi_with_grad(t, tV0, tV1, tF)
    assert toi.numel() > 0
    w = torch.rand(toi.numel(), dtype=torch.float64, device=dev)
    (w * toi).sum().backward()
    r0, r1 = orc.toi_grad_list(V0, V1, sten.cpu().numpy(), is_vf.cpu().numpy(), w.cpu().numpy(), th, th)
    ref = np.concatenate([r0.ravel(), r1.ravel()])
    got = np.concatenate([tV0.grad.cpu().numpy().ravel(), tV1.grad.cpu().numpy().ravel()])
    assert grad_close(got, ref) <= 1.0


def test_adjoint_refuses_hits_of_an_older_update(gpu):
    """An update or refit after the traverse replaces the vertex records the
    hits would be evaluated at (and may change nv): the adjoint refuses
    instead of writing a gradient of mismatched data, through the C ABI and
    through the autograd wrapper."""
    import torch

    from mclccd_amd import _lib
    from mclccd_amd.autograd import toi_grad, toi_with_grad
    from mclccd_amd.tree import BVHTree

    V0, V1, F, th = scenes.cloth_fold(n=24)
    t = gpu_run(V0, V1, F, th)
    n = t.counts()["hits"]
    assert n > 0
    w = np.ones(n)
    toi_grad(t, w)  # fine right after the traverse
    t.update(V0, V1, F, refit=True)
    with pytest.raises(_lib.MclgError) as e:
        toi_grad(t, w)
    assert e.value.code == _lib.ERR_STATE
    # a larger mesh after the forward pass: backward must not write dim*new_nv values
    dev = torch.device("cuda:0")
    tV0 = torch.from_numpy(V0).to(dev).requires_grad_(True)
    tV1 = torch.from_numpy(V1).to(dev).requires_grad_(True)
    tF = torch.from_numpy(F).to(dev)
    t2 = BVHTree(3)
    t2.options.box_eta = t2.options.vf_ccd_eta = t2.options.ee_ccd_eta = th
    toi, _, _ = toi_with_grad(t2, tV0, tV1, tF)
    W0, W1, G, _ = scenes.cloth_fold(n=32)
    t2.update(W0, W1, G)
    with pytest.raises(RuntimeError):
        toi.sum().backward()
