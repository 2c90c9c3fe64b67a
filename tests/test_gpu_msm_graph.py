"""The captured MSM batch (zkg_msm_batch_create / _run / _destroy): one hipGraph holding every
job's pipeline on concurrent branches.  Every replay equals one plain zkg_msm per job bit for bit
(both curves, uneven sizes, uniform / 64-bit / sparse / all-equal scalars, so heavy buckets go
through the device-driven tree instead of the host-driven rounds), replays reread the buffers'
contents, and a replay whose scalars outgrow the windows fitted at create is refused rather than
answered.  The reference computes such jobs one MSM call at a time
(lib/cbits/curves/g1/proj/bn128_G1_proj.c:506-586)."""
import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, ints_to_array, random_scalars

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]


def _scalars(C, n, dist, seed):
    rng = np.random.default_rng(seed)
    if dist == "uniform":
        return random_scalars(seed, n, C.r)
    if dist == "u64":
        return [int(x) for x in rng.integers(0, 2**64, size=n, dtype=np.uint64)]
    if dist == "equal":  # one bucket per window holds every entry: heavy buckets
        v = random_scalars(seed, 1, C.r)[0]
        return [v] * n
    return [int(x) if rng.random() < 0.2 else int(rng.integers(0, 2)) for x in random_scalars(seed, n, C.r)]


@pytest.fixture(scope="module")
def points(zk):
    return {C.id: zk.DeviceArray.from_host(Oracle(C).points_from_scalars(random_scalars(190 + C.id, 6000, C.r)))
            for C in CURVES}


def _jobs(zk, specs, seed):
    jobs = []
    for k, (C, n, dist) in enumerate(specs):
        S = ints_to_array(_scalars(C, max(n, 1), dist, seed + k), 4)
        jobs.append((C.id, n, zk.DeviceArray.from_host(S)))
    return jobs


SPECS = [(BN, 6000, "uniform"), (BLS, 2500, "sparse"), (BN, 1 << 12, "u64"), (BLS, 700, "uniform"),
         (BN, 1, "uniform"), (BLS, 5000, "u64"), (BN, 0, "uniform"), (BN, 5000, "equal"), (BLS, 3000, "equal"),
         (BN, 4000, "sparse")]


@pytest.mark.parametrize("policy", ["plain", "stats"])
@pytest.mark.parametrize("nstreams", [1, 4])
def test_graph_batch_equals_plain_calls(zk, points, policy, nstreams):
    pol = {"plain": zk.SCHED_PLAIN, "stats": zk.SCHED_STATS}[policy]
    jobs = _jobs(zk, SPECS, 300)
    ref = [zk.msm_device(c, n, d, points[c]) for c, n, d in jobs]
    print(f"graph batch {policy} streams={nstreams}", flush=True)
    plan = zk.MsmBatch([c for c, _, _ in jobs], [n for _, n, _ in jobs], [d for _, _, d in jobs],
                       [points[c] for c, _, _ in jobs], policy=pol, nstreams=nstreams)
    for rep in range(3):
        outs, wall = plan.run()
        assert wall > 0
        for k, (o, r) in enumerate(zip(outs, ref)):
            assert o.shape == r.shape and (o == r).all(), (rep, k, SPECS[k])
    plan.close()


def test_graph_batch_learned_with_committed_model(zk, points):
    """The committed cost model picks the schedules; results as the plain calls."""
    import costmodel
    net, _ = costmodel.load()
    net.install()
    try:
        jobs = _jobs(zk, SPECS, 400)
        ref = [zk.msm_device(c, n, d, points[c]) for c, n, d in jobs]
        plan = zk.MsmBatch([c for c, _, _ in jobs], [n for _, n, _ in jobs], [d for _, _, d in jobs],
                           [points[c] for c, _, _ in jobs], policy=zk.SCHED_LEARNED, nstreams=4)
        outs, _ = plan.run()
        assert all((o == r).all() for o, r in zip(outs, ref))
        plan.close()
    finally:
        zk.Mlp.uninstall()
        net.close()


def test_graph_batch_rereads_inputs_and_refuses_wider_scalars(zk, points):
    """A replay reads the buffers as they are then; new scalars of the measured width give the new
    results; full-width scalars in a plan fitted to 64-bit ones are refused (statistics policy),
    and the plan works again once the buffers hold scalars it covers."""
    specs = [(BN, 3000, "u64"), (BLS, 2000, "u64"), (BN, 1500, "uniform")]
    jobs = _jobs(zk, specs, 500)
    plan = zk.MsmBatch([c for c, _, _ in jobs], [n for _, n, _ in jobs], [d for _, _, d in jobs],
                       [points[c] for c, _, _ in jobs], policy=zk.SCHED_STATS, nstreams=2)
    outs, _ = plan.run()
    assert all((o == zk.msm_device(c, n, d, points[c])).all() for o, (c, n, d) in zip(outs, jobs))
    # same widths, new values
    for k, (C, n, dist) in enumerate(specs):
        jobs[k][2].copy_from_host(ints_to_array(_scalars(C, n, dist, 600 + k), 4))
    outs, _ = plan.run()
    assert all((o == zk.msm_device(c, n, d, points[c])).all() for o, (c, n, d) in zip(outs, jobs))
    # a full-width scalar where 64-bit ones were measured
    wide = ints_to_array(_scalars(BN, 3000, "uniform", 700), 4)
    jobs[0][2].copy_from_host(wide)
    with pytest.raises(zk.ZkgError):
        plan.run()
    jobs[0][2].copy_from_host(ints_to_array(_scalars(BN, 3000, "u64", 701), 4))
    outs, _ = plan.run()
    assert all((o == zk.msm_device(c, n, d, points[c])).all() for o, (c, n, d) in zip(outs, jobs))
    plan.close()


@pytest.mark.parametrize("C", CURVES, ids=["bn254", "bls12-381"])
def test_graph_batch_heavy_buckets_closed_form(zk, C):
    """All-equal scalars at 2^15 points: every window's entries land in one bucket with ~1000
    continuation partials (the device tree's case); the result is (s sum k_i) G."""
    n = 1 << 15
    K = random_scalars(800 + C.id, n, C.r)
    dP = zk.scalar_mul_gen(C.id, zk.DeviceArray.from_host(ints_to_array(K, 4)), n)
    s = random_scalars(801 + C.id, 1, C.r)[0]
    dS = zk.DeviceArray.from_host(ints_to_array([s] * n, 4))
    plan = zk.MsmBatch([C.id], [n], [dS], [dP], policy=zk.SCHED_PLAIN, nstreams=1)
    outs, _ = plan.run()
    assert C.words_to_affine(outs[0]) == C.mul(s * sum(K) % C.r, C.g)
    assert (outs[0] == zk.msm_device(C.id, n, dS, dP)).all()
    plan.close()
