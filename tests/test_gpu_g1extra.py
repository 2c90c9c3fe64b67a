"""GPU binary MSMs, batch conversions and the Jacobian G1 FFT, called through the reference's
entry points (lib/cbits/curves/g1/proj/bn128_G1_proj.c:146-166, :684-782,
lib/cbits/curves/g1/jac/bn128_G1_jac.c:139-158, :726-950) and the device API, against the
golden fixtures of the reference's C, the oracle on fresh inputs, and at large sizes the
closed forms: a binary MSM over k_i G is (sum of the selected k_i) G, and projective or
Jacobian forms (x z, y z, z) / (x z^2, y z^3, z) of a point convert back to its affine words."""
import json
from pathlib import Path

import numpy as np
import pytest

from cref import Oracle
from zkmath import BN, BLS, array_to_ints, ints_to_array, random_scalars, to_limbs

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]
GOLDEN = Path(__file__).resolve().parent / "golden"
ALL_FF = np.uint64(0xFFFFFFFFFFFFFFFF)


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN / "golden_small.npz", allow_pickle=False)


@pytest.fixture(scope="module")
def manifest():
    return json.loads((GOLDEN / "manifest.json").read_text())


def jac_from_proj_normalised(C, p):
    """normalised projective words -> the normalised Jacobian words the library writes:
    the same (x, y, 1), infinity [1 : 1 : 0]."""
    L = C.lp
    if not p[2 * L:].any():
        one = np.array(to_limbs(C.fp_to_mont(1), L), np.uint64)
        return np.concatenate([one, one, np.zeros(L, np.uint64)])
    return p


# ---------------------------------------------------------------- binary MSM

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_binary_msm_vs_golden(zk, golden, manifest, C):
    o = Oracle(C)
    keys = [k for k in manifest["binary"] if k.startswith(C.name + "_")]
    assert keys
    for key in keys:
        P, bits, rows = golden[key + "_points"], golden[key + "_bits"], golden[key + "_rows"]
        exp = golden[key + "_affine"]
        for coord in ("proj", "jac"):
            assert (zk.msm_binary(C.id, bits, P, coord=coord) == exp).all(), (key, coord)
        proj = zk.msm_binary(C.id, bits, P, coord="proj", out="proj")
        assert (proj == o.msm_binary(bits, P)).all(), key
        jac = zk.msm_binary(C.id, bits, P, coord="jac", out="jac")
        assert (jac == jac_from_proj_normalised(C, o.msm_binary(bits, P))).all(), key
        multi = zk.multi_msm_binary(C.id, rows, P)
        assert (multi == golden[key + "_multi_affine"]).all(), key
        mj = zk.multi_msm_binary(C.id, rows, P, coord="jac")
        assert (mj == golden[key + "_multi_jac_affine"]).all(), key


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n,density", [(0, 0.5), (1, 1.0), (2, 0.5), (64, 0.0), (64, 1.0), (1000, 0.3),
                                       (4097, 0.5), (20000, 0.05)])
def test_binary_msm_vs_oracle(zk, C, n, density):
    o = Oracle(C)
    ks = random_scalars(3000 + n, n, C.r)
    if n >= 4:
        ks[1] = 0
        ks[2] = ks[3]
    P = o.points_from_scalars(ks) if n else np.zeros((0, 2 * C.lp), np.uint64)
    rng = np.random.default_rng(n)
    bits = (rng.random(n) < density).astype(np.uint8) * rng.integers(1, 256, n).astype(np.uint8)
    exp = o.msm_binary(bits, P)
    assert (zk.msm_binary(C.id, bits, P, out="proj") == exp).all()
    aff = zk.msm_binary(C.id, bits, P)
    if exp[2 * C.lp:].any():
        assert (aff == o.proj_to_affine(exp)).all()
    else:
        assert (aff == ALL_FF).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("nrows,n", [(0, 10), (1, 50), (5, 0), (33, 257), (100, 64)])
def test_multi_binary_msm_vs_oracle(zk, C, nrows, n):
    o = Oracle(C)
    ks = random_scalars(3500 + n, n, C.r)
    P = o.points_from_scalars(ks) if n else np.zeros((0, 2 * C.lp), np.uint64)
    rng = np.random.default_rng(nrows * 1000 + n)
    rows = (rng.random((nrows, n)) < 0.5).astype(np.uint8)
    if nrows > 2:
        rows[0] = 0
        rows[1] = 1
    exp = o.multi_msm_binary(rows, P)
    got = zk.multi_msm_binary(C.id, rows, P, out="proj")
    assert got.shape == (nrows, 3 * C.lp)
    assert (got == exp).all()


@pytest.mark.parametrize("C,logn", [(BN, 18), (BLS, 16)], ids=["bn254-2^18", "bls12-381-2^16"])
def test_binary_msm_device_closed_form(zk, C, logn):
    n = 1 << logn
    d_k = zk.fill_random(C.id, 1, n, 4242 + logn)
    ks = array_to_ints(d_k.to_host((n, 4)))
    d_pts = zk.scalar_mul_gen(C.id, d_k, n)
    o = Oracle(C)
    pts = d_pts.to_host((n, 2 * C.lp))
    for i in (0, 1, n // 2, n - 1):  # the device inputs are k_i G
        assert (pts[i] == o.points_from_scalars([ks[i]])[0]).all()
    rng = np.random.default_rng(logn)
    for density in (0.5, 0.01):
        bits = (rng.random(n) < density).astype(np.uint8)
        e = sum(k for k, b in zip(ks, bits) if b) % C.r
        exp = o.points_from_scalars([e])[0]
        d_bits = zk.DeviceArray.from_bytes(bits)
        assert (zk.msm_binary_device(C.id, n, d_bits, d_pts) == exp).all(), density
        d_int = zk.PreparedPoints(C.id, d_pts, n)
        got = zk.msm_binary_device(C.id, n, d_bits, d_int, flags=zk.MSM_POINTS_INTERNAL)
        assert (got == exp).all(), density
        d_int.free()


# ---------------------------------------------------------------- batch conversions

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_batch_conversions_vs_golden(zk, golden, C):
    key = f"{C.name}_conv"
    assert (zk.batch_to_affine(C.id, golden[key + "_proj"]) == golden[key + "_proj_to_affine"]).all()
    assert (zk.batch_to_affine(C.id, golden[key + "_jac"], coord="jac") == golden[key + "_jac_to_affine"]).all()
    assert (zk.batch_from_affine(C.id, golden[key + "_affine"]) == golden[key + "_proj_from_affine"]).all()
    assert (zk.batch_from_affine(C.id, golden[key + "_affine"], coord="jac") ==
            golden[key + "_jac_from_affine"]).all()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("n", [0, 1, 15, 16, 17, 33, 300])
def test_batch_conversions_vs_oracle(zk, C, n):
    """Ragged chunks of the 16-point batch inversion, infinities at chunk edges and a chunk of
    infinities only."""
    o = Oracle(C)
    ks = random_scalars(5000 + n, n, C.r)
    for i in (0, 15, 16, 31):
        if i < n:
            ks[i] = 0
    if n >= 64:
        for i in range(32, 48):
            ks[i] = 0
    for jac in (False, True):
        coord = "jac" if jac else "proj"
        P = o.jac_points(ks) if jac else o.proj_points(ks)
        P = P.reshape(n, 3 * C.lp)
        assert (zk.batch_to_affine(C.id, P, coord=coord) == o.batch_to_affine(P, jac=jac)).all(), coord
        A = o.points_from_scalars(ks).reshape(n, 2 * C.lp)
        assert (zk.batch_from_affine(C.id, A, coord=coord) == o.batch_from_affine(A, jac=jac)).all(), coord


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_batch_to_affine_device_large(zk, C):
    """(x z, y z, z) and (x z^2, y z^3, z) of 2^20 device points convert back to (x, y); the
    scaling uses the device field product, itself checked against the oracle elsewhere."""
    n = 1 << 20
    L = C.lp
    d_k = zk.fill_random(C.id, 1, n, 777)
    A = zk.scalar_mul_gen(C.id, d_k, n).to_host((n, 2 * L))
    Z = zk.fill_random(C.id, 0, n, 778).to_host((n, L))
    Z[5] = 0  # one infinity
    d_z = zk.DeviceArray.from_host(Z)
    d_x, d_y = zk.DeviceArray.from_host(np.ascontiguousarray(A[:, :L])), zk.DeviceArray.from_host(
        np.ascontiguousarray(A[:, L:]))
    d_t, d_u = zk.DeviceArray(n * L), zk.DeviceArray(n * L)
    FM = zk.FOP["mul"]
    for jac in (False, True):
        if jac:
            d_z2 = zk.DeviceArray(n * L)
            zk.field_op(C.id, 0, FM, d_z, d_z, d_z2, n)
            zk.field_op(C.id, 0, FM, d_x, d_z2, d_t, n)
            zk.field_op(C.id, 0, FM, d_z2, d_z, d_z2, n)
            zk.field_op(C.id, 0, FM, d_y, d_z2, d_u, n)
        else:
            zk.field_op(C.id, 0, FM, d_x, d_z, d_t, n)
            zk.field_op(C.id, 0, FM, d_y, d_z, d_u, n)
        P = np.concatenate([d_t.to_host((n, L)), d_u.to_host((n, L)), Z], axis=1)
        d_p, d_a = zk.DeviceArray.from_host(P), zk.DeviceArray(n * 2 * L)
        zk.batch_to_affine_device(C.id, n, d_p, d_a, jac=jac)
        got = d_a.to_host((n, 2 * L))
        exp = A.copy()
        exp[5] = ALL_FF
        bad = np.nonzero((got != exp).any(axis=1))[0]
        assert bad.size == 0, (jac, bad[:8])


# ---------------------------------------------------------------- Jacobian G1 FFT

@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_jac_g1_fft_vs_golden(zk, golden, manifest, C):
    keys = [k for k in manifest["g1fft_jac"] if k.startswith(C.name + "_")]
    assert keys
    for key in keys:
        m = int(key.rsplit("_", 1)[1])
        P, g = golden[key + "_src"], golden[key + "_gen"]
        assert (zk.g1_fft(C.id, m, g, P, coord="jac") == golden[key + "_fwd"]).all(), key
        assert (zk.g1_fft(C.id, m, g, P, inverse=True, coord="jac") == golden[key + "_inv"]).all(), key


@pytest.mark.parametrize("C", CURVES, ids=IDS)
@pytest.mark.parametrize("m", [0, 2, 5])
def test_jac_g1_fft_vs_oracle(zk, C, m):
    o = Oracle(C)
    ks = random_scalars(5500 + m, 1 << m, C.r)
    if m >= 2:
        ks[2] = 0
    P = o.jac_points(ks)
    for g in (o.fft_generator(m), np.array(to_limbs(C.fr_to_mont(99991 + m), 4), np.uint64)):
        for inv in (False, True):
            exp = o.g1_fft(m, g, P, inverse=inv, jac=True)
            assert (zk.g1_fft(C.id, m, g, P, inverse=inv, coord="jac") == exp).all(), (m, inv)


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_multi_binary_msm_device(zk, C):
    """The device-buffer form over prepared points equals the reference entry point and the
    oracle, row by row."""
    o = Oracle(C)
    n, nrows = 3000, 9
    ks = random_scalars(8800 + C.id, n, C.r)
    P = o.points_from_scalars(ks)
    rng = np.random.default_rng(8)
    rows = (rng.random((nrows, n)) < 0.3).astype(np.uint8)
    rows[2] = 0
    d_p = zk.DeviceArray.from_host(P)
    prep = zk.PreparedPoints(C.id, d_p, n)
    d_bits = zk.DeviceArray.from_bytes(rows)
    got = zk.multi_msm_binary_device(C.id, nrows, n, d_bits, prep, flags=zk.MSM_PROJ_OUT)
    assert (got == o.multi_msm_binary(rows, P)).all()
    aff = zk.multi_msm_binary_device(C.id, nrows, n, d_bits, d_p)
    assert (aff == zk.multi_msm_binary(C.id, rows, P)).all()
    prep.free()
