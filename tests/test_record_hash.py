"""The multiset check the full-size GPU test uses to compare the hit list with
the candidates (tests/test_gpu_full_size.py): the count plus a wrapping int64
sum of a 64-bit mix of each (stencil, toi) record.  It must not depend on the
record order and must change for any plausible mistake in one record."""
import numpy as np
import torch

from tests.test_gpu_full_size import _candidate_hashes, _hit_hashes, _record_hash


def _records(n, seed):
    rng = np.random.default_rng(seed)
    sten = torch.as_tensor(rng.integers(0, 600_000, (n, 4), dtype=np.int32))
    toi = torch.as_tensor(rng.random(n))
    return sten, toi


def test_order_invariant():
    sten, toi = _records(50_000, 1)
    h = int(_record_hash(sten, toi).sum())
    p = torch.as_tensor(np.random.default_rng(2).permutation(len(sten)))
    assert int(_record_hash(sten[p], toi[p]).sum()) == h


def test_detects_single_record_changes():
    sten, toi = _records(50_000, 3)
    h = int(_record_hash(sten, toi).sum())
    t = toi.clone()
    t[17] = np.nextafter(t[17].item(), 2.0)  # one ulp
    assert int(_record_hash(sten, t).sum()) != h
    for a, b in ((0, 1), (1, 2), (2, 3), (0, 3)):  # two indices of one stencil swapped
        s = sten.clone()
        s[5, [a, b]] = s[5, [b, a]]
        assert int(_record_hash(s, toi).sum()) != h
    assert int(_record_hash(sten[1:], toi[1:]).sum()) != h  # a dropped record
    dup = torch.cat([sten, sten[:1]]), torch.cat([toi, toi[:1]])
    assert int(_record_hash(*dup).sum()) != h  # a duplicated record


def test_hits_against_candidates():
    vs, vt = _records(20_000, 4)
    es, et = _records(30_000, 5)
    rng = np.random.default_rng(6)
    vt[torch.as_tensor(rng.random(len(vt)) < 0.3)] = -1.0
    et[torch.as_tensor(rng.random(len(et)) < 0.3)] = -1.0
    mv, me = vt >= 0, et >= 0
    hs = torch.cat([vs[mv], es[me]])
    ht = torch.cat([vt[mv], et[me]])
    hv = torch.cat([torch.ones(int(mv.sum()), dtype=torch.uint8), torch.zeros(int(me.sum()), dtype=torch.uint8)])
    nvf = int(mv.sum())
    p = torch.cat([torch.randperm(nvf), nvf + torch.randperm(len(hs) - nvf)])  # any order within each kind
    want = _candidate_hashes(vs, vt, es, et)
    assert _hit_hashes(hs[p], hv[p], ht[p]) == want
    hv2 = hv.clone()
    hv2[0] = 0  # a VF hit labelled EE
    assert _hit_hashes(hs, hv2, ht) != want
