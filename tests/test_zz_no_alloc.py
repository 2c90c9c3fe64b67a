"""zkg_ntt_noalloc (the NTT with its scratch in a fixed device array, no allocation) against
zkg_ntt, through the C ABI: two calls in a row, the second smaller, for forward, inverse and
coset transforms of both curves, in place and out of place; sizes past the scratch and
generators of another order are refused."""
import numpy as np
import pytest

from zkmath import BLS, BN, random_scalars, to_limbs

pytestmark = pytest.mark.gpu


def _eta(C, seed):
    return np.array(to_limbs(C.fr_to_mont(random_scalars(seed, 1, C.r)[0]), 4), np.uint64)


@pytest.mark.parametrize("C", [BN, BLS], ids=["bn254", "bls12-381"])
def test_noalloc_matches_ntt(zk, C):
    for case, (m1, m2) in enumerate([(18, 11), (16, 1), (13, 2)]):
        for inverse in (False, True):
            for coset in (False, True):
                for m in (m1, m2):  # two calls in a row, the second smaller
                    N = 1 << m
                    g = zk.fft_generator(C.id, m)
                    eta = _eta(C, 100 * case + m) if coset else None
                    dX = zk.fill_random(C.id, 1, N, 1000 * case + 10 * m + inverse)
                    dA, dB = zk.DeviceArray(4 * N), zk.DeviceArray(4 * N)
                    print(f"noalloc {C.name} m={m} inverse={inverse} coset={coset}", flush=True)
                    zk.ntt_device(C.id, m, g, dX, dA, inverse=inverse, eta=eta)
                    zk.ntt_device_noalloc(C.id, m, g, dX, dB, inverse=inverse, eta=eta)
                    exp = dA.to_host()
                    assert (dB.to_host() == exp).all(), (m, inverse, coset)
                    # in place
                    zk.ntt_device_noalloc(C.id, m, g, dX, dX, inverse=inverse, eta=eta)
                    assert (dX.to_host() == exp).all(), (m, inverse, coset, "in place")


def test_noalloc_size_zero(zk):
    C = BN
    d = zk.fill_random(C.id, 1, 1, 7)
    out = zk.DeviceArray(4)
    zk.ntt_device_noalloc(C.id, 0, zk.fft_generator(C.id, 0), d, out)
    assert (out.to_host() == d.to_host()).all()


def test_noalloc_refuses_what_it_cannot_hold(zk):
    C = BN
    d = zk.fill_random(C.id, 1, 1 << 19, 5)
    out = zk.DeviceArray(4 << 19)
    with pytest.raises(zk.ZkgError, match="0..18"):
        zk.ntt_device_noalloc(C.id, 19, zk.fft_generator(C.id, 19), d, out)
    other = np.array(to_limbs(C.fr_to_mont(987654321), 4), np.uint64)
    with pytest.raises(zk.ZkgError, match="order 2\\^m"):
        zk.ntt_device_noalloc(C.id, 10, other, d, out)
    # the library still works afterwards
    small = zk.fill_random(C.id, 1, 1 << 10, 6)
    a, b = zk.DeviceArray(4 << 10), zk.DeviceArray(4 << 10)
    g = zk.fft_generator(C.id, 10)
    zk.ntt_device(C.id, 10, g, small, a)
    zk.ntt_device_noalloc(C.id, 10, g, small, b)
    assert (a.to_host() == b.to_host()).all()
