"""Batched uneven NTTs (one hipGraph replay per run) against the oracle for small jobs and
against the plain one-call-per-job path for large ones; repeated replays are bit-identical.
Semantics per job follow lib/cbits/curves/poly/mont/bn128_poly_mont.c:452-638 (forward,
inverse, forward_shifted: x_i eta^i before the transform, inverse_shifted: eta^i after)."""
import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, ints_to_array, random_scalars, to_limbs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("C", [BN, BLS], ids=["bn254", "bls12-381"])
def test_ntt_batch_small_vs_oracle(zk, C):
    o = Oracle(C)
    specs = []
    for m in (1, 2, 3, 5, 8, 9, 11, 12):
        for inv in (0, 1):
            for coset in (False, True):
                specs.append((m, inv, coset))
    jobs, host = [], []
    for k, (m, inv, coset) in enumerate(specs):
        X = ints_to_array([C.fr_to_mont(x) for x in random_scalars(40 + k, 1 << m, C.r)], 4)
        eta = np.array(to_limbs(C.fr_to_mont(random_scalars(900 + k, 1, C.r)[0]), 4), np.uint64) if coset else None
        dX, dY = zk.DeviceArray.from_host(X), zk.DeviceArray(X.size)
        jobs.append((m, inv, eta, dX, dY))
        host.append((X, eta))
    plan = zk.NttBatch(C.id, jobs)
    # the outputs are cleared before each replay, so that every replay has to write them
    for rep in range(3):
        for (_, _, _, _, dY), (X, _) in zip(jobs, host):
            dY.copy_from_host(np.zeros_like(X))
        plan.run()
        for (m, inv, coset), (X, eta), (_, _, _, _, dY) in zip(specs, host, jobs):
            g = o.fft_generator(m)
            if coset:
                exp = o.ntt_shifted(eta, m, g, X, inverse=bool(inv))
            else:
                exp = o.ntt(m, g, X, inverse=bool(inv))
            assert (dY.to_host(X.shape) == exp).all(), (rep, m, inv, coset)
    plan.close()


def test_ntt_batch_large_vs_plain(zk):
    rng = np.random.default_rng(5)
    specs = [(int(m), int(rng.integers(0, 2)), bool(rng.integers(0, 2))) for m in rng.integers(10, 21, size=24)]
    jobs, plain = [], []
    for k, (m, inv, coset) in enumerate(specs):
        N = 1 << m
        dX = zk.fill_random(0, 1, N, 77 + k)
        dY, dP = zk.DeviceArray(4 * N), zk.DeviceArray(4 * N)
        eta = np.array(to_limbs(BN.fr_to_mont(random_scalars(5000 + k, 1, BN.r)[0]), 4), np.uint64) if coset else None
        jobs.append((m, inv, eta, dX, dY))
        zk.ntt_device(0, m, zk.fft_generator(0, m), dX, dP, inverse=bool(inv), eta=eta)
        plain.append(dP)
    plan = zk.NttBatch(0, jobs)
    for rep in range(3):  # three replays of the plan's graph, the outputs cleared before each
        for (m, _, _), (_, _, _, _, dY) in zip(specs, jobs):
            dY.copy_from_host(np.zeros(4 << m, np.uint64))
        plan.run()
        for (m, _, _), (_, _, _, _, dY), dP in zip(specs, jobs, plain):
            assert (dY.to_host() == dP.to_host()).all(), (rep, m)
    plan.close()


def test_ntt_tuning_table_bit_exact(zk):
    """Per-size NTT schedules from the autotuner (and deliberately small ones) give the oracle's
    output; the batch plan picks them up per job."""
    o = Oracle(BN)
    try:
        for m in (10, 13):
            smax, tl, ms, same = zk.ntt_tune(BN.id, m, reps=1, install=True)
            assert same and ms > 0 and 5 <= smax <= 11 and 7 <= tl <= 12
        zk.ntt_set_tuning(5, 7, curve=BN.id, m=12)
        for m in (10, 12, 13):
            X = ints_to_array([BN.fr_to_mont(x) for x in random_scalars(700 + m, 1 << m, BN.r)], 4)
            g = o.fft_generator(m)
            dX, dY = zk.DeviceArray.from_host(X), zk.DeviceArray(X.size)
            zk.ntt_device(BN.id, m, g, dX, dY)
            assert (dY.to_host(X.shape) == o.ntt(m, g, X)).all(), m
            plan = zk.NttBatch(BN.id, [(m, 1, None, dX, dY)])
            plan.run()
            assert (dY.to_host(X.shape) == o.ntt(m, g, X, inverse=True)).all(), m
            plan.close()
    finally:
        zk.ntt_set_tuning(8, 9)
