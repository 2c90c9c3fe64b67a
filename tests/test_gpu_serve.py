"""KZG-commit serving loop: every served commitment equals the same job run alone, equals the
closed form f(tau) G (f interpolated from the evaluations, examples/KZG.hs:91-97), and the
admission control never exceeds the HBM budget it was given."""
import numpy as np
import pytest

from zkmath import BLS, BN, array_to_ints, ints_to_array, random_scalars

pytestmark = pytest.mark.gpu


def barycentric(values, logn, tau, C):
    """f(tau) for the f with f(w^k) = values[k] (standard-form ints), w the size-n root."""
    n = 1 << logn
    w = C.fft_root(logn)
    r = C.r
    acc, wk = 0, 1
    for v in values:
        acc = (acc + v * wk * pow(tau - wk, -1, r)) % r
        wk = wk * w % r
    return (pow(tau, n, r) - 1) * pow(n, -1, r) % r * acc % r


@pytest.mark.parametrize("C", [BLS, BN], ids=["bls12-381", "bn254"])
def test_served_commitments(zk, C):
    tau = random_scalars(2024, 1, C.r)[0]
    srs_n = 1 << 12
    d_srs = zk.srs_generate(C.id, srs_n, tau)
    srs = zk.PreparedPoints(C.id, d_srs, srs_n)
    rng = np.random.default_rng(3)
    lognes = [int(x) for x in rng.integers(4, 13, size=12)]
    vals_std = [random_scalars(70 + k, 1 << m, C.r) for k, m in enumerate(lognes)]
    vals = [ints_to_array([C.fr_to_mont(v) for v in vs], 4) for vs in vals_std]
    # alone: one worker, one job at a time
    alone = zk.Server(C.id, srs, 1 << 34, 1)
    ref = []
    for m, v in zip(lognes, vals):
        _, out = alone.submit(m, v)
        alone.drain()
        ref.append(out.copy())
    alone.close()
    # in the mix: four workers, everything submitted at once
    mix = zk.Server(C.id, srs, 1 << 34, 4)
    outs = [mix.submit(m, v)[1] for m, v in zip(lognes, vals)]
    mix.drain()
    for a, b in zip(ref, outs):
        assert (a == b).all()
    mix.close()
    # closed form for two of them
    for k in (0, 5):
        ft = barycentric(vals_std[k], lognes[k], tau, C)
        assert C.words_to_affine(outs[k]) == C.mul(ft, C.g)


def test_large_jobs_run_alone(zk):
    """With the size rule at 2^10 points, no job starts beside a job of 2^10 or 2^11 points (so
    with only such jobs at most one runs at a time); every result equals the job run alone."""
    C = BN
    srs_n = 1 << 12
    d_srs = zk.srs_generate(C.id, srs_n, 777)
    srs = zk.PreparedPoints(C.id, d_srs, srs_n)
    lognes = [10, 11, 10, 11, 10]
    vals = [ints_to_array([C.fr_to_mont(v) for v in random_scalars(600 + k, 1 << m, C.r)], 4)
            for k, m in enumerate(lognes)]
    alone = zk.Server(C.id, srs, 1 << 34, 1)
    ref = []
    for m, v in zip(lognes, vals):
        ref.append(alone.submit(m, v)[1])
        alone.drain()
    alone.close()
    big = zk.Server(C.id, srs, 1 << 34, 4)
    big.set_exclusive_log(10)
    outs = [big.submit(m, v)[1] for m, v in zip(lognes, vals)]
    big.drain()
    assert big.stats()[1] == 1
    big.close()
    assert all((a == b).all() for a, b in zip(ref, outs))
    small = zk.Server(C.id, srs, 1 << 34, 4)
    small.set_exclusive_log(12)
    for _ in range(3):
        outs = [small.submit(m, v)[1] for m, v in zip(lognes, vals)]
        small.drain()
        assert all((a == b).all() for a, b in zip(ref, outs))
    small.close()
    srs.free()


def test_admission_respects_budget(zk):
    C = BLS
    srs_n = 1 << 12
    d_srs = zk.srs_generate(C.id, srs_n, 12345)
    srs = zk.PreparedPoints(C.id, d_srs, srs_n)
    vals = [ints_to_array([C.fr_to_mont(v) for v in random_scalars(500 + k, 1 << 12, C.r)], 4) for k in range(6)]
    tight = zk.Server(C.id, srs, 1, 4)  # smaller than any job: one job at a time
    for v in vals:
        tight.submit(12, v)
    tight.drain()
    _, conc = tight.stats()
    assert conc == 1
    tight.close()
