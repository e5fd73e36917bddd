import numpy as np

from ufednl_amd import datagen


def test_configs_shapes_and_labels():
    p1 = datagen.config1()
    assert p1.features.shape == (50, 221, 68)
    assert set(np.unique(p1.labels)) <= {-1.0, 1.0}
    assert np.all(p1.features[:, :, -1] == 1.0)
    p2 = datagen.config2()
    assert p2.features.shape == (142, 350, 300)
    pos = (p2.labels == 1).mean()
    assert 0.015 < pos < 0.06
    dens = p2.features[:, :, :-1].mean()
    assert 0.035 < dens < 0.045
    p3 = datagen.config3()
    assert p3.features.shape == (80, 407, 123)
    assert 0.4 < (p3.labels == 1).mean() < 0.6


def test_deterministic():
    a = datagen.config1()
    b = datagen.config1()
    assert np.array_equal(a.features, b.features) and np.array_equal(a.labels, b.labels)


def test_batched_and_serving_shapes():
    sh = datagen.batched_shapes()
    assert len(sh) == 64 and all(32 <= d <= 512 and 16 <= n <= 256 and 100 <= m <= 600 for d, n, m in sh)
    sv = datagen.serving_shapes(50)
    assert all(64 <= d <= 2048 and 32 <= n <= 512 and 200 <= m <= 1000 for d, n, m in sv)
