"""The optimizer restatements of the oracle (Adam with its L2 term, the Ema
weight average), pinned by closed forms: Adam's first step moves every
weight by lr against the sign of its gradient (plus L2 term) whatever the
gradient's size; a zero gradient with L2 regularisation shrinks weights
towards 0; the debiased average of constant weights is those weights, and of
the weights 1, 2, ..., t it is the decay-weighted mean."""
import numpy as np

from oracle.oracle import adam_update_d, ema_update_d


def test_adam_first_step_is_a_signed_step():
    rng = np.random.default_rng(0)
    w = rng.standard_normal(1000)
    g = rng.standard_normal(1000) * 10.0 ** rng.uniform(-4, 3, 1000)
    w1, m, v = adam_update_d(w, g, 0 * w, 0 * w, 1, 1e-2, 0.9, 0.99, 1e-15)
    assert np.allclose(w1, w - 1e-2 * np.sign(g), rtol=0, atol=1e-10)
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.01 * g * g)


def test_adam_l2_term_joins_the_gradient():
    w = np.array([2.0, -3.0, 0.5])
    g = np.zeros(3)
    w1, _, _ = adam_update_d(w, g, 0 * w, 0 * w, 1, 1e-3, 0.9, 0.99, 1e-15, l2_reg=1e-6)
    # zero loss gradient: the L2 term alone moves each weight towards 0 by lr
    assert np.allclose(w1, w - 1e-3 * np.sign(w), atol=1e-12)
    # a mistake that subtracted the term (or left it out) fails here
    w2, _, _ = adam_update_d(w, np.array([1e-7, 1e-7, 1e-7]), 0 * w, 0 * w, 1, 1e-3, 0.9, 0.99, 1e-15, l2_reg=1.0)
    assert np.allclose(w2, w - 1e-3 * np.sign(w + 1e-7), atol=1e-12)


def test_adam_second_step_closed_form():
    w, g1, g2 = np.array([1.0]), np.array([0.3]), np.array([-0.1])
    w1, m, v = adam_update_d(w, g1, 0 * w, 0 * w, 1, 0.01, 0.9, 0.99, 0.0)
    w2, m, v = adam_update_d(w1, g2, m, v, 2, 0.01, 0.9, 0.99, 0.0)
    mh = (0.9 * 0.1 * 0.3 + 0.1 * -0.1) / (1 - 0.81)
    vh = (0.99 * 0.01 * 0.09 + 0.01 * 0.01) / (1 - 0.9801)
    assert np.allclose(w2, 1.0 - 0.01 - 0.01 * mh / np.sqrt(vh), atol=1e-14)


def test_ema_debiased_average():
    a = np.zeros(4)
    w = np.array([1.0, -2.0, 3.5, 0.0])
    for t in range(1, 6):
        a, deb = ema_update_d(a, w, 0.95, t)
        assert np.allclose(deb, w, atol=1e-12)
    a = np.zeros(1)
    d = 0.9
    for t in range(1, 8):
        a, deb = ema_update_d(a, np.array([float(t)]), d, t)
        ws = np.arange(1, t + 1, dtype=np.float64)
        wt = (1 - d) * d ** (t - ws)
        assert np.allclose(deb, (wt * ws).sum() / wt.sum(), atol=1e-12)
