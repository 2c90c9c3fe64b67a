"""Schedule selection on the GPU: the cost-model network's HIP kernels against the float64
restatement in oracle/mlp_ref.py, the scalar-statistics pass against numpy, and bit-exactness
of every policy (statistics-trimmed windows, learned picks, batched uneven MSMs) against the
oracle's MSM (the reference's bucket method, lib/cbits/curves/g1/proj/bn128_G1_proj.c:506-673)
and the plain path.  The last test tunes a small job set exhaustively, trains the network on
part of it and reports what its picks gain on the rest."""
import json
import math
import sys
from pathlib import Path

import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, ints_to_array, random_scalars

REPO = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(REPO / "oracle"))
import mlp_ref  # noqa: E402

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


def _params(dims, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    n = sum(i * o + o for i, o in zip(dims[:-1], dims[1:]))
    return rng.normal(0, scale, size=n).astype(np.float32)


@pytest.mark.parametrize("dims", [[11, 32, 32, 1], [11, 16, 8, 1], [4, 6, 3], [3, 1]])
def test_mlp_forward_and_gradients_match_oracle(zk, dims):
    rng = np.random.default_rng(sum(dims))
    net = zk.Mlp(dims, seed=5)
    p = _params(dims, 9)
    net.set_params(p)
    assert np.array_equal(net.get_params(), p)
    B = 300
    X = rng.normal(size=(B, dims[0])).astype(np.float32)
    t = rng.normal(size=(B, dims[-1])).astype(np.float32)
    y_ref, _ = mlp_ref.forward(dims, p.astype(np.float64), X.astype(np.float64))
    assert np.allclose(net.forward(X), y_ref, rtol=1e-4, atol=1e-4)
    g, loss = net.gradients(X, t)
    loss_ref, g_ref = mlp_ref.loss_and_grad(dims, p.astype(np.float64), X.astype(np.float64), t)
    assert math.isclose(loss, loss_ref, rel_tol=1e-4)
    # float32 sums over 300 rows against float64: relative to the gradient's scale
    assert np.max(np.abs(g - g_ref)) <= 1e-4 * max(1.0, np.max(np.abs(g_ref)))
    net.close()


def test_mlp_adam_steps_match_oracle(zk):
    dims = [11, 16, 8, 1]
    rng = np.random.default_rng(3)
    X = rng.normal(size=(128, 11)).astype(np.float32)
    t = (X[:, :1] * 0.5 - X[:, 1:2] ** 2).astype(np.float32)
    p0 = _params(dims, 4)
    net = zk.Mlp(dims)
    net.set_params(p0)
    loss = net.train(X, t, epochs=20, lr=1e-2)
    p_ref, loss_ref = mlp_ref.adam(dims, p0.astype(np.float64), X.astype(np.float64), t, epochs=20, lr=1e-2)
    assert math.isclose(loss, loss_ref, rel_tol=2e-3)
    assert np.max(np.abs(net.get_params() - p_ref)) < 2e-3
    net.close()


def test_mlp_training_fits_a_smooth_target(zk):
    rng = np.random.default_rng(8)
    X = rng.uniform(-1, 1, size=(512, 11)).astype(np.float32)
    t = (np.sin(2 * X[:, :1]) + X[:, 1:2] * X[:, 2:3]).astype(np.float32)
    net = zk.Mlp([11, 32, 32, 1], seed=2)
    _, l0 = net.gradients(X, t)
    l1 = net.train(X, t, epochs=1500, lr=3e-3)
    assert l1 < 0.05 * l0
    net.close()


# ---------------------------------------------------------------------------------------
# statistics pass

def _stats_numpy(ints):
    return dict(n=len(ints), n_zero=sum(1 for x in ints if x == 0), n_one=sum(1 for x in ints if x == 1),
                max_bits=max((x.bit_length() for x in ints), default=0))


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_scalar_stats_exact(zk, C):
    rng = np.random.default_rng(21)
    cases = {
        "uniform": random_scalars(1, 5000, C.r),
        "u64": [int(x) for x in rng.integers(0, 2**63, size=3000)],
        "sparse": [int(x) if rng.random() < 0.2 else int(rng.integers(0, 2)) for x in random_scalars(2, 4000, C.r)],
        "zeros": [0] * 700,
        "ones_and_max": [1] * 300 + [C.r - 1],
        "bit_boundaries": [(1 << k) - 1 for k in range(1, 200)] + [1 << 199],
    }
    for name, ints in cases.items():
        S = ints_to_array(ints, 4)
        d = zk.DeviceArray.from_host(S)
        assert zk.scalar_stats(C.id, len(ints), d).as_dict() == _stats_numpy(ints), name
        Sm = ints_to_array([C.fr_to_mont(x % C.r) for x in ints], 4)
        dm = zk.DeviceArray.from_host(Sm)
        assert zk.scalar_stats(C.id, len(ints), dm, flags=zk.MSM_MONT).as_dict() == _stats_numpy(ints), name


# ---------------------------------------------------------------------------------------
# every policy gives the oracle's point

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_trimmed_windows_bit_exact(zk, C):
    """Windows cover only max_bits + 1 bits; the top window's carry must still land."""
    o = Oracle(C)
    n = 300
    rng = np.random.default_rng(31)
    P = o.points_from_scalars(random_scalars(40, n, C.r))
    dP = zk.DeviceArray.from_host(P)
    for c in (3, 5, 8, 13, 16):
        for bits in sorted({1, 2, c - 1, c, c + 1, 2 * c - 1, 2 * c, 63, 64, 65, 128, 253, C.r.bit_length()}):
            if bits < 1:
                continue
            top = min(1 << bits, C.r)
            ss = [int(x) % top for x in random_scalars(50 + bits, n, C.r)]
            ss[0] = top - 1                      # largest value of that width: carries up to the top
            ss[1] = (1 << (bits - 1)) if bits > 1 else 1
            S = ints_to_array(ss, 4)
            dS = zk.DeviceArray.from_host(S)
            exp = o.msm_affine(S, P)
            for seg in (2, 5):
                got = zk.msm_device(C.id, n, dS, dP, window=c, seg_log=seg, policy=zk.SCHED_STATS)
                assert (got == exp).all(), (c, bits, seg)
    # default choice under the statistics policy, all-zero and all-one vectors
    for ss in ([0] * n, [1] * n, [0] * (n - 1) + [C.r - 1]):
        S = ints_to_array(ss, 4)
        dS = zk.DeviceArray.from_host(S)
        got = zk.msm_device(C.id, n, dS, dP, policy=zk.SCHED_STATS)
        assert (got == o.msm_affine(S, P)).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_heavy_bucket_trees_bit_exact(zk, C):
    """Buckets spanning many accumulation segments are summed by pairwise trees; a top window
    with only a few distinct digits gives a handful of buckets with thousands of partials beside
    many buckets with a few (the per-segment round path), small seg_log and fix_limit force it."""
    n = 1 << 15
    P = Oracle(C).points_from_scalars(random_scalars(80, 4096, C.r))
    P = np.ascontiguousarray(np.tile(P, (n // 4096, 1)))
    dP = zk.DeviceArray.from_host(P)
    S = ints_to_array(random_scalars(81, n, C.r), 4)
    dS = zk.DeviceArray.from_host(S)
    ref = zk.msm_device(C.id, n, dS, dP)
    for c, seg, lim in [(14, 1, 1), (14, 2, 1), (14, 2, 0), (12, 1, 1), (16, 1, 1), (9, 1, 2), (13, 3, 4)]:
        for pol in (zk.SCHED_PLAIN, zk.SCHED_STATS):
            got = zk.msm_device(C.id, n, dS, dP, window=c, seg_log=seg, policy=pol, fix_limit=lim)
            assert (got == ref).all(), (c, seg, lim, pol)
    # closed form of the tiled input: sum_i s_i k_(i mod 4096)
    ks = random_scalars(80, 4096, C.r)
    ss = random_scalars(81, n, C.r)
    e = sum(s * ks[i % 4096] for i, s in enumerate(ss)) % C.r
    assert C.words_to_affine(ref) == C.mul(e, C.g)


def _dist_scalars(C, n, dist, seed):
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        return random_scalars(seed, n, C.r)
    if dist == "u64":
        return [int(x) for x in rng.integers(0, 2**64, size=n, dtype=np.uint64)]
    return [int(x) if rng.random() < 0.2 else int(rng.integers(0, 2)) for x in random_scalars(seed, n, C.r)]


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_learned_policy_and_batch_bit_exact(zk, C):
    # any installed network only changes the schedule: train one on arbitrary targets
    rng = np.random.default_rng(7)
    net = zk.Mlp([zk.MSM_NFEAT, 32, 32, 1], seed=3)
    X = rng.uniform(0, 1, size=(256, zk.MSM_NFEAT)).astype(np.float32)
    net.train(X, rng.normal(size=(256, 1)).astype(np.float32), epochs=50, lr=1e-2)
    net.install()
    o = Oracle(C)
    sizes = [1 << 8, 777, 1 << 10, 3000, 1 << 12, 1 << 13, 10000, 1 << 14]
    nmax = max(sizes)
    P = o.points_from_scalars(random_scalars(60, nmax, C.r))
    dP = zk.DeviceArray.from_host(P)
    jobs = []
    for k, n in enumerate(sizes):
        dist = ("uniform", "u64", "sparse")[k % 3]
        S = ints_to_array(_dist_scalars(C, n, dist, 70 + k), 4)
        jobs.append((n, S, zk.DeviceArray.from_host(S)))
    plain = [zk.msm_device(C.id, n, dS, dP) for n, _, dS in jobs]
    for (n, S, _), r in zip(jobs[:3], plain[:3]):
        assert (r == o.msm_affine(S, P[:n])).all()
    for pol in (zk.SCHED_STATS, zk.SCHED_LEARNED):
        for (n, _, dS), r in zip(jobs, plain):
            assert (zk.msm_device(C.id, n, dS, dP, policy=pol) == r).all(), (pol, n)
        for streams in (1, 4):
            outs, _ = zk.msm_batch(C.id, [n for n, _, _ in jobs], [d for _, _, d in jobs], [dP] * len(jobs),
                                   policy=pol, nstreams=streams)
            assert all((a == b).all() for a, b in zip(outs, plain)), (pol, streams)
    zk.Mlp.uninstall()
    net.close()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_sparse_compaction_bit_exact(zk, C):
    """Measured sparse scalar sets (half or more in {0, 1}) run as a compacted dense MSM plus a
    direct sum of the points with scalar 1: every mix of zeros, ones and others, Montgomery and
    one-limb scalars, external and prepared points, equals the plain path and the closed form."""
    n = (1 << 14) + 37
    ks = random_scalars(90 + C.id, n, C.r)
    o = Oracle(C)
    P = o.points_from_scalars(ks[:4096])
    P = np.ascontiguousarray(np.tile(P, (n // 4096 + 1, 1))[:n])
    kk = [ks[i % 4096] for i in range(n)]
    dP = zk.DeviceArray.from_host(P)
    prep = zk.PreparedPoints(C.id, dP, n)
    rng = np.random.default_rng(91)
    others = random_scalars(92, n, C.r)
    cases = {
        "sparse": [int(others[i]) if rng.random() < 0.2 else int(rng.integers(0, 2)) for i in range(n)],
        "ones_only": [1] * n,
        "zeros_and_ones": [int(rng.integers(0, 2)) for _ in range(n)],
        "zeros_only": [0] * n,
        "few_others": [int(others[i]) if i % 997 == 0 else 0 for i in range(n)],
        "one_other": [5] + [1] * (n - 1),
    }
    for name, ss in cases.items():
        S = ints_to_array(ss, 4)
        dS = zk.DeviceArray.from_host(S)
        ref = zk.msm_device(C.id, n, dS, dP)
        e = sum(s * k for s, k in zip(ss, kk)) % C.r
        if e:
            assert C.words_to_affine(ref) == C.mul(e, C.g), name
        for pol in (zk.SCHED_STATS, zk.SCHED_LEARNED):
            assert (zk.msm_device(C.id, n, dS, dP, policy=pol) == ref).all(), (name, pol)
            assert (zk.msm_device(C.id, n, dS, prep, policy=pol) == ref).all(), (name, pol, "prepared")
        Sm = ints_to_array([C.fr_to_mont(x) for x in ss], 4)
        dSm = zk.DeviceArray.from_host(Sm)
        assert (zk.msm_device(C.id, n, dSm, dP, flags=zk.MSM_MONT, policy=zk.SCHED_STATS) == ref).all(), name
        if all(x < 2**64 for x in ss):
            S1 = np.ascontiguousarray(S[:, :1])
            got = zk.msm_device(C.id, n, zk.DeviceArray.from_host(S1), dP, expo_nlimbs=1, policy=zk.SCHED_STATS)
            assert (got == ref).all(), (name, "one limb")
    prep.free()


def test_cost_model_gain_over_heuristics(zk):
    """Exhaustive tuning of a small BN254 job set; the network trained on the grid picks schedules
    for a held-out mix.  Reports (and bounds) its time against the plain path, the analytic pick
    on the measured scalars, and the exhaustive best."""
    import costmodel
    pts = costmodel.PointSets(1 << 17)
    train = [costmodel.tune_job(pts, c, n, d, s, reps=3) for c, n, d, s in costmodel.grid_jobs([0], 13, 17, 101)]
    held = [costmodel.tune_job(pts, c, n, d, s, reps=3) for c, n, d, s in costmodel.mix_jobs([0], 12, 202, 13.0, 17.0)]
    assert all(r["identical"] for r in train + held)
    net, loss, _ = costmodel.train(train, epochs=3000)
    rep = costmodel.evaluate(held, net)
    summary = {k: v for k, v in rep.items() if k != "rows"}
    print(json.dumps(summary))
    # the network's picks beat the plain path and the analytic pick on the measured scalars (the
    # method without the model) and stay near the exhaustive optimum
    assert rep["mlp"]["total_ms"] < rep["plain"]["total_ms"]
    assert rep["mlp"]["total_ms"] < rep["stats_heuristic"]["total_ms"]
    assert rep["mlp"]["geomean_vs_best"] < 1.3
    net.close()


def test_batch_mixed_curves_bit_exact(zk):
    """One batch holding BN254 and BLS12-381 jobs of uneven sizes equals one plain call per job."""
    specs = [(BN, 3000, "uniform"), (BLS, 2500, "sparse"), (BN, 1 << 12, "u64"), (BLS, 700, "uniform"),
             (BN, 1, "uniform"), (BLS, 5000, "u64"), (BN, 0, "uniform")]
    pts = {C.id: zk.DeviceArray.from_host(Oracle(C).points_from_scalars(random_scalars(90 + C.id, 5000, C.r)))
           for C in CURVES}
    jobs = []
    for k, (C, n, dist) in enumerate(specs):
        S = ints_to_array(_dist_scalars(C, max(n, 1), dist, 95 + k), 4)
        jobs.append((C.id, n, zk.DeviceArray.from_host(S)))
    ref = [zk.msm_device(c, n, d, pts[c]) for c, n, d in jobs]
    for pol in (zk.SCHED_PLAIN, zk.SCHED_STATS):
        outs, wall = zk.msm_batch([c for c, _, _ in jobs], [n for _, n, _ in jobs], [d for _, _, d in jobs],
                                  [pts[c] for c, _, _ in jobs], policy=pol, nstreams=3)
        assert wall > 0
        assert all(o.shape == r.shape and (o == r).all() for o, r in zip(outs, ref)), pol
