"""Device-resident SRS in the reference's flat-array format (lib/src/ZK/Algebra/Class/Flat.hs:
211-233: raw little-endian limbs, no header, length from the file size): files written from
device memory hold exactly the words of the array, read back bit for bit, a loaded and
prepared SRS commits like one prepared from memory, and a size that is not a whole number of
elements is refused as readFlatArray refuses it."""
import ctypes

import numpy as np
import pytest

from zkmath import BN, BLS

pytestmark = pytest.mark.gpu
CURVES = [BN, BLS]
IDS = ["bn254", "bls12-381"]


@pytest.mark.parametrize("C,logn", [(BN, 20), (BLS, 16)], ids=["bn254-2^20", "bls12-381-2^16"])
def test_srs_roundtrip_and_commit(zk, tmp_path, C, logn):
    n = (1 << logn) + 3
    w = 2 * C.lp
    d_srs = zk.srs_generate(C.id, n, 0xC0FFEE + C.id)
    host = d_srs.to_host((n, w))
    path = tmp_path / f"srs_{C.name}.bin"
    zk.flat_write(path, d_srs, w, n)
    assert path.stat().st_size == n * w * 8
    assert (np.fromfile(path, dtype="<u8").reshape(n, w) == host).all()
    assert zk.flat_file_count(path, w) == n
    back = zk.flat_read(path, w)
    assert (back.to_host((n, w)) == host).all()
    d_aff, n2 = zk.srs_load(C.id, path, prepare=False)
    assert n2 == n and (d_aff.to_host((n, w)) == host).all()
    prep, n3 = zk.srs_load(C.id, path, prepare=True)
    ref_prep = zk.PreparedPoints(C.id, d_srs, n)
    m = 1 << (logn - 1)
    d_v = zk.fill_random(C.id, 1, m, 99)
    a = zk.msm_device(C.id, m, d_v, prep, flags=zk.MSM_MONT | zk.MSM_POINTS_INTERNAL)
    b = zk.msm_device(C.id, m, d_v, ref_prep, flags=zk.MSM_MONT | zk.MSM_POINTS_INTERNAL)
    c = zk.msm_device(C.id, m, d_v, d_srs, flags=zk.MSM_MONT)
    assert (a == b).all() and (a == c).all()
    prep.free()
    ref_prep.free()


@pytest.mark.parametrize("C", CURVES, ids=IDS)
def test_flat_file_size_checks(zk, tmp_path, C):
    w = 2 * C.lp
    path = tmp_path / "bad.bin"
    np.zeros(w * 3 + 1, np.uint64).tofile(path)
    with pytest.raises(zk.ZkgError, match="invalid size"):
        zk.flat_file_count(path, w)
    with pytest.raises(zk.ZkgError):
        zk.srs_load(C.id, path)
    empty = tmp_path / "empty.bin"
    empty.write_bytes(b"")
    assert zk.flat_file_count(empty, w) == 0
    small = zk.DeviceArray(w)
    np.zeros(w * 2, np.uint64).tofile(path)
    got = ctypes.c_size_t()
    st = zk.lib().zkg_flat_read(str(path).encode(), 8 * w, small.ptr, 1, ctypes.byref(got))
    assert st != 0 and b"too small" in zk.lib().zkg_last_error()
    with pytest.raises(zk.ZkgError, match="cannot open"):
        zk.flat_file_count(tmp_path / "missing.bin", w)
