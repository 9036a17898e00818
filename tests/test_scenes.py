"""Sizes of the synthetic workloads (the benchmark configurations)."""
import numpy as np

from mclccd_amd import scenes


def test_cloth_fold_sizes():
    V0, V1, F, th = scenes.cloth_fold()
    assert V0.shape == (4225, 3) and F.shape == (8192, 3) and th == 1e-3
    assert np.all(V0[:, 2] == 0) and F.dtype == np.int32
    # deterministic
    assert np.array_equal(V1, scenes.cloth_fold()[1])


def test_stacked_layers_sizes():
    V0, V1, F, th = scenes.stacked_layers()
    assert V0.shape == (524176, 3) and F.shape == (1036800, 3) and th == 1e-4
    assert F.max() == len(V0) - 1


def test_disk2d_sizes():
    V0, V1, E, th = scenes.disk2d(npts=3000)
    assert V0.shape == (3000, 2) and E.shape[1] == 2
    assert np.all(np.linalg.norm(V0, axis=1) <= 1)
    Vb, _, Eb, _ = scenes.disk2d(npts=3000, edges="boundary")
    assert len(Eb) < len(E)


def test_batch_sheets():
    b = scenes.batch_sheets(count=512)
    assert len(b) == 512
    ntri = np.array([len(s[2]) for s in b])
    assert ntri.min() >= 2 * 16 * 16 and ntri.max() <= 2 * 180 * 180
