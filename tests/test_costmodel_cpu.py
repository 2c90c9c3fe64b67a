"""CPU checks of the schedule cost model's host side (no GPU calls).

* the float64 network restatement in oracle/mlp_ref.py (the checker of the HIP kernels) is
  pinned against central finite differences, a closed-form linear case and a hand-computed
  Adam step;
* the feature encoding, the candidate sets and the analytic picks exported by the library are
  checked against facts of the signed-digit bucket method.
"""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "oracle"))

import mlp_ref  # noqa: E402


def _net(dims, seed):
    rng = np.random.default_rng(seed)
    n = sum(i * o + o for i, o in zip(dims[:-1], dims[1:]))
    return rng.normal(0, 0.5, size=n)


@pytest.mark.parametrize("dims", [[3, 1], [5, 7, 1], [11, 8, 6, 1], [4, 5, 3]])
def test_oracle_gradient_matches_finite_differences(dims):
    rng = np.random.default_rng(len(dims))
    p = _net(dims, 3)
    X = rng.normal(size=(9, dims[0]))
    t = rng.normal(size=(9, dims[-1]))
    _, g = mlp_ref.loss_and_grad(dims, p, X, t)
    fd = mlp_ref.finite_diff_grad(dims, p, X, t)
    assert np.allclose(g, fd, rtol=1e-5, atol=1e-7)


def test_oracle_linear_closed_form():
    # one linear layer: y = W x + b, loss = mean (y - t)^2, dL/dW = 2/B sum (y - t) x^T
    dims = [2, 1]
    p = np.array([0.5, -1.0, 0.25])
    X = np.array([[1.0, 2.0], [3.0, -1.0]])
    t = np.array([[0.0], [1.0]])
    y = X @ np.array([0.5, -1.0]) + 0.25            # [-1.25, 2.75]
    e = y - t[:, 0]                                  # [-1.25, 1.75]
    loss, g = mlp_ref.loss_and_grad(dims, p, X, t)
    assert math.isclose(loss, (1.25 ** 2 + 1.75 ** 2) / 2)
    assert np.allclose(g, [(e * X[:, 0]).sum(), (e * X[:, 1]).sum(), e.sum()])


def test_oracle_adam_first_step_is_sign_step():
    # with zero moments the bias-corrected first Adam step is lr * g / (|g| + eps) ~ lr * sign(g)
    dims = [3, 1]
    p = np.array([0.3, -0.2, 0.1, 0.05])
    X = np.array([[1.0, 0.0, 2.0], [0.7, 1.0, -1.0]])
    t = np.array([[1.0], [-1.0]])
    _, g = mlp_ref.loss_and_grad(dims, p, X, t)
    p1, _ = mlp_ref.adam(dims, p, X, t, epochs=1, lr=0.01)
    assert np.allclose(p1, p - 0.01 * np.sign(g), atol=1e-6)


def test_oracle_relu_masks_dead_units():
    # a hidden unit with a large negative bias is dead: its incoming weights get no gradient
    dims = [2, 2, 1]
    p = np.array([1.0, 1.0, 0.5, -0.5, 0.0, -100.0, 1.0, 1.0, 0.0])
    X = np.array([[1.0, 2.0], [-1.0, 0.5]])
    t = np.array([[0.0], [1.0]])
    _, g = mlp_ref.loss_and_grad(dims, p, X, t)
    assert g[2] == 0 and g[3] == 0 and g[5] == 0


# ---------------------------------------------------------------------------------------
# host-side schedule code in the library (CPU-only entry points)

@pytest.fixture(scope="module")
def zkh():
    sys.path.insert(0, str(REPO / "zikkurat-algebra_amd"))
    import zkgpu
    zkgpu.lib()
    return zkgpu


def layout(bits, c):
    """Balanced windows: W = ceil((bits + 1) / c) windows of width ceil((bits + 1) / W) or one less."""
    W = -(-(bits + 1) // c)
    cm = -(-(bits + 1) // W)
    wide = bits + 1 - W * (cm - 1)
    assert 1 <= wide <= W and wide * cm + (W - wide) * (cm - 1) == bits + 1
    return W, cm


def test_window_layout_covers_exactly():
    for bits in range(0, 300):
        for c in range(1, 21):
            W, cm = layout(bits, c)
            assert cm <= c and W == -(-(bits + 1) // c)
            assert layout(bits, cm) == (W, cm)  # the effective width is its own layout


def shape(curve, bits, n):
    """The bucket problem a job runs as: BN254 scalars wider than 128 bits take the GLV split
    (2n points, 127-bit halves, two digits per window from each full scalar)."""
    if curve == 0 and bits > 128:
        return 127, 2 * n, 2
    return bits, n, 1


@pytest.mark.parametrize("curve", [0, 1])
def test_features_encode_window_counts(zkh, curve):
    # W = ceil((bits + 1) / c): 255-bit BLS12-381 scalars at c = 16 need 16 windows, 64-bit
    # ones 5; BN254 scalars over 128 bits run as 2n points of 127 bits (8 windows at c = 16)
    for bits, c in [(254, 16), (64, 16), (63, 16), (0, 8), (255, 13), (254, 14), (254, 17), (128, 16), (129, 16)]:
        n = 1 << 16
        b, pts, per = shape(curve, bits, n)
        W, cm = layout(b, c)
        st = zkh.stats_from(n, 0, 0, bits)
        x = zkh.msm_features(curve, st, c, 5)
        assert math.isclose(x[6] * 64, W, abs_tol=1e-4), (bits, c)
        assert math.isclose(x[4] * 20, cm, abs_tol=1e-5) and math.isclose(x[5] * 8, 5, abs_tol=1e-5)
        assert math.isclose(x[1] * 24, math.log2(pts), abs_tol=1e-4)
        assert math.isclose(x[2] * 256, b, abs_tol=1e-3)
        assert math.isclose(2 ** (x[7] * 32) - 1, W * per * n, rel_tol=1e-4)
    assert layout(254, 16) == (16, 16) and layout(127, 16) == (8, 16) and layout(127, 17) == (8, 16)


def test_features_sparse_counts_entries(zkh):
    # zeros cost nothing, ones one entry: entries = W * (n - zeros - ones) + ones
    n, z, o = 1 << 14, 6000, 7000
    st = zkh.stats_from(n, z, o, 254)
    x = zkh.msm_features(1, st, 10, 4)
    W = math.ceil(255 / 10)
    assert math.isclose(2 ** (x[7] * 32) - 1, W * (n - z - o) + o, rel_tol=1e-4)
    assert math.isclose(x[3], (z + o) / n, rel_tol=1e-6)
    assert x[0] == 1.0


@pytest.mark.parametrize("curve", [0, 1])
@pytest.mark.parametrize("logn", [10, 14, 18, 22])
def test_candidates_contain_both_analytic_picks(zkh, curve, logn):
    for bits, zeros, ones in [(254, 0, 0), (64, 0, 0), (254, 0.4, 0.4)]:
        n = 1 << logn
        st = zkh.stats_from(n, int(zeros * n), int(ones * n), bits)
        cands = zkh.msm_candidates(curve, st)
        assert len(cands) == len(set(cands)) > 0
        assert all(2 <= c <= 18 and 2 <= s <= 7 for c, s in cands)
        b = shape(curve, bits, n)[0]
        assert all(layout(b, c)[1] == c for c, _ in cands)  # listed by effective width
        assert tuple(zkh.msm_pick(curve, st, zkh.SCHED_STATS)) in cands
        cp, sp = zkh.msm_pick(curve, st, zkh.SCHED_PLAIN)
        assert min(c for c, _ in cands) <= cp <= max(c for c, _ in cands) and sp == 5


def test_plain_pick_matches_op_count_model(zkh):
    # the plain pick minimises W * (10 n' + 28 * 2^(c-1)) over c <= 16 (windows whose bucket
    # bases fit LDS): BLS12-381 with W = ceil(256 / c) over n points, uniform c-bit windows;
    # BN254 through the GLV split, W = ceil(128 / c) over 2n points, reported as the balanced
    # width the job runs with
    for logn in range(8, 24, 3):
        n = 1 << logn
        best = min(range(2, 17), key=lambda c: (math.ceil(256 / c) * (10 * n + 28 * 2 ** (c - 1)), c))
        assert zkh.msm_pick(1, zkh.stats_from(n, 0, 0, 255), zkh.SCHED_PLAIN)[0] == best
        best = min(range(2, 17), key=lambda c: (math.ceil(128 / c) * (20 * n + 28 * 2 ** (c - 1)), c))
        assert zkh.msm_pick(0, zkh.stats_from(n, 0, 0, 254), zkh.SCHED_PLAIN)[0] == layout(127, best)[1]


def test_stats_pick_narrow_scalars_use_fewer_windows(zkh):
    n = 1 << 20
    cw, _ = zkh.msm_pick(0, zkh.stats_from(n, 0, 0, 254), zkh.SCHED_STATS)
    cn, _ = zkh.msm_pick(0, zkh.stats_from(n, 0, 0, 64), zkh.SCHED_STATS)
    assert math.ceil(65 / cn) < math.ceil(255 / cw)
    # mostly zero/one scalars: fewer entries per window, so no wider window than for uniform ones
    cs, _ = zkh.msm_pick(0, zkh.stats_from(n, int(0.4 * n), int(0.4 * n), 254), zkh.SCHED_STATS)
    assert cs <= cw
