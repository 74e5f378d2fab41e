"""CPU: the oracle of the learned selection / AQ network (oracle/jxl_mlp.c).

Its analytic gradient is pinned by central finite differences of its own
loss in float64, its loss by a direct numpy restatement of the definition,
and the per-block DCT8 costs it exposes by the selection they come from."""

from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 96, 64, 7


def random_problem(jo, n, seed):
    rng = np.random.default_rng(seed)
    params = (rng.standard_normal(jo.MLP_PARAMS) * 0.2).astype(np.float32)
    x = rng.standard_normal((n, N_IN)).astype(np.float32)
    lab = np.zeros((n, jo.MLP_LABELS), np.float32)
    lab[:, 0] = rng.integers(0, 2, n)
    lab[:, 1:5] = rng.integers(0, 2, (n, 4))
    lab[:, 5] = rng.standard_normal(n)
    lab[:, 6] = rng.integers(0, 2, n)
    lab[:, 7:11] = rng.integers(0, 2, (n, 4))
    return params, x, lab


def numpy_loss(params, x, lab, lam):
    p = params.astype(np.float64)
    w1 = p[: N_HID * N_IN].reshape(N_HID, N_IN)
    b1 = p[N_HID * N_IN : N_HID * N_IN + N_HID]
    o = N_HID * N_IN + N_HID
    w2 = p[o : o + N_OUT * N_HID].reshape(N_OUT, N_HID)
    b2 = p[o + N_OUT * N_HID :]
    h = np.maximum(x.astype(np.float64) @ w1.T + b1, 0.0)
    z = h @ w2.T + b2
    lse = np.logaddexp(z[:, 0], z[:, 1])
    ce = lse - np.where(lab[:, 0] > 0.5, z[:, 1], z[:, 0])
    bce = np.logaddexp(0.0, z[:, 2:6]) - lab[:, 1:5] * z[:, 2:6]
    mse = (z[:, 6] - lab[:, 5]) ** 2
    return float(np.mean(lab[:, 6] * ce + (lab[:, 7:11] * bce).sum(1) + lam * mse)), z


def test_loss_and_outputs_match_numpy(jo):
    params, x, lab = random_problem(jo, 50, 1)
    loss, _ = jo.mlp_loss_grad(params, x, lab, 0.3, want_grad=False)
    want, z = numpy_loss(params, x, lab, float(np.float32(0.3)))
    assert abs(loss - want) <= 1e-10 * max(1.0, abs(want))
    np.testing.assert_allclose(jo.mlp_forward(params, x), z, rtol=1e-12, atol=1e-12)


def test_gradient_matches_finite_differences(jo):
    """Central differences in float64 on a float32 parameter grid: the
    perturbation is applied to a float64 copy through the numpy restatement
    (whose loss equals the oracle's), so only the oracle's backward pass is
    under test."""
    params, x, lab = random_problem(jo, 20, 2)
    lam = 0.5  # exact in float32, as the oracle takes it
    _, grad = jo.mlp_loss_grad(params, x, lab, lam)
    rng = np.random.default_rng(3)
    idx = np.concatenate([rng.choice(N_HID * N_IN, 40, replace=False),
                          N_HID * N_IN + rng.choice(N_HID, 8, replace=False),
                          N_HID * N_IN + N_HID + rng.choice(N_OUT * N_HID, 20, replace=False),
                          N_HID * N_IN + N_HID + N_OUT * N_HID + np.arange(N_OUT)])
    p64 = params.astype(np.float64)
    for k in idx:
        eps = 1e-6
        a = p64.copy()
        a[k] += eps
        b = p64.copy()
        b[k] -= eps
        fd = (numpy_loss(a, x, lab, lam)[0] - numpy_loss(b, x, lab, lam)[0]) / (2 * eps)
        assert abs(fd - grad[k]) <= 1e-6 + 1e-5 * abs(fd), (int(k), fd, grad[k])


def test_masks_remove_terms(jo):
    params, x, lab = random_problem(jo, 10, 4)
    lab[:, 6:11] = 0.0
    loss, grad = jo.mlp_loss_grad(params, x, lab, 0.0)
    assert loss == 0.0 and not grad.any()


def test_block_costs8_are_the_selection_costs(jo):
    """A flat plane costs next to nothing (every AC coefficient quantizes to
    zero; only the basis' rounding leaves a squared error near 1e-15); a plane
    whose only content is one block's texture costs only there, and its cost
    is the selection's own DCT8 cost of that block (the same value the
    selection compares against the larger transforms)."""
    w, h = 64, 32
    y = np.full((h, w), 0.4, np.float32)
    assert jo.block_costs8(y, 1.0).max() < 1e-12
    rng = np.random.default_rng(5)
    y[8:16, 16:24] += rng.standard_normal((8, 8)).astype(np.float32) * 0.05
    c = jo.block_costs8(y, 1.0).reshape(h // 8, w // 8)
    assert c[1, 2] > 1e-3 and np.count_nonzero(c > 1e-12) == 1
    cost, dist, rate = jo.block_costs8(y, 1.0, parts=True)
    np.testing.assert_array_equal(cost, (64.0 * dist.astype(np.float32)).astype(np.float32) +
                                  np.float32(6.0e-4) * rate)
    # The textured block is too costly for a DCT16 or DCT32 to cover it:
    # the selection keeps the region around it at DCT8 only where it must.
    acs = jo.select_transforms(y, 1.0).reshape(h // 8, w // 8)
    assert acs[0, 0] in (16, 32)
