"""Writes the golden fixtures under tests/golden/ from the reference's own C.

The reference path (lib/cbits/...) is compiled from its sources by `make -C oracle ref` into
oracle/_ref/libzkref.so; this script calls its entry points on seeded inputs and stores inputs
and outputs (small cases) or output SHA-256 digests (large cases, inputs regenerated from the
seed by tests/zkmath.py:seeded_fr_vector).  Run from the repository root:

    make -C oracle ref && python tests/golden/gen_golden.py            # everything
    make -C oracle ref && python tests/golden/gen_golden.py --large    # only missing large entries
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))

from cref import Oracle, Ref, ref  # noqa: E402
from zkmath import BN, BLS, ints_to_array, random_scalars, seeded_fr_vector, to_limbs  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.uint64).tobytes()).hexdigest()


def msm_inputs(C, n, seed):
    ks = random_scalars(seed, n, C.r)
    ss = random_scalars(seed + 1, n, C.r)
    if n >= 8:
        ss[0], ss[1], ss[2], ss[3] = 0, C.r - 1, (1 << 64) - 1, (1 << 253) + 12345
        ks[4] = ks[5]
        ss[4] = ss[5]
        ks[6] = (C.r - ks[7]) % C.r
        ss[6] = ss[7]
        ks[3] = 0
    return ks, ss


def main():
    assert ref() is not None, "build the reference first: make -C oracle ref"
    out = {}
    manifest = {"generator": "tests/golden/gen_golden.py", "reference": "oracle/_ref/libzkref.so (lib/cbits sources, clang -O2)",
                "msm": [], "ntt": [], "large": [], "g1fft": [], "binary": [],
                "conv": [], "g1fft_jac": [], "poly": [], "kzg": []}
    for C in (BN, BLS):
        R, O = Ref(C), Oracle(C)
        for n in (1, 3, 17, 256, 1000):
            ks, ss = msm_inputs(C, n, 4000 + n + 100 * C.id)
            P = O.points_from_scalars(ks)
            S = ints_to_array(ss, 4)
            Sm = ints_to_array([C.fr_to_mont(s % C.r) for s in ss], 4)
            key = f"{C.name}_msm_{n}"
            out[key + "_points"] = P
            out[key + "_scalars"] = S
            out[key + "_scalars_mont"] = Sm
            for coord in ("proj", "jac"):
                for kind, sc in (("std", S), ("mont", Sm)):
                    out[f"{key}_{coord}_{kind}_affine"] = R.msm(sc, P, coord=coord, mont=kind == "mont")
            manifest["msm"].append(key)
        for m in range(2, 13):
            X = seeded_fr_vector(7000 + m + 100 * C.id, 1 << m, C.r)
            X = ints_to_array([C.fr_to_mont(int(sum(int(w) << (64 * i) for i, w in enumerate(row)))) for row in X], 4)
            g = np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)
            key = f"{C.name}_ntt_{m}"
            out[key + "_src"] = X
            out[key + "_gen"] = g
            out[key + "_fwd"] = R.ntt(m, g, X)
            out[key + "_inv"] = R.ntt(m, g, X, inverse=True)
            if m in (2, 5, 9):
                eta = np.array(to_limbs(C.fr_to_mont(random_scalars(8000 + m, 1, C.r)[0]), 4), np.uint64)
                out[key + "_eta"] = eta
                out[key + "_fwd_shifted"] = R.ntt_shifted(eta, m, g, X)
                out[key + "_inv_shifted"] = R.ntt_shifted(eta, m, g, X, inverse=True)
            manifest["ntt"].append(key)
        for ms, mt in ((2, 4), (3, 7), (5, 6)):
            X = out[f"{C.name}_ntt_{ms}_src"]
            gs, gt = out[f"{C.name}_ntt_{ms}_gen"], np.array(to_limbs(C.fr_to_mont(C.fft_root(mt)), 4), np.uint64)
            out[f"{C.name}_asym_{ms}_{mt}"] = R.ntt_asym(ms, mt, gs, gt, X)
        # G1 FFT (bn128_G1_proj.c:791-902): projective points k_i G with Z != 1 and an infinity
        for m in (0, 1, 2, 3, 5):
            ks = random_scalars(8000 + m + 100 * C.id, 1 << m, C.r)
            if m >= 2:
                ks[2] = 0
            P = O.proj_points(ks)
            key = f"{C.name}_g1fft_{m}"
            out[key + "_src"] = P
            gens = {"root": np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)}
            if m == 3:
                gens["other"] = np.array(to_limbs(C.fr_to_mont(987654321), 4), np.uint64)
            for gname, g in gens.items():
                out[f"{key}_{gname}_gen"] = g
                out[f"{key}_{gname}_fwd"] = R.g1_fft(m, g, P)
                out[f"{key}_{gname}_inv"] = R.g1_fft(m, g, P, inverse=True)
                manifest["g1fft"].append(f"{key}_{gname}")
        # Jacobian G1 FFT (bn128_G1_jac.c:840-950): Jacobian points with Z != 1 and an infinity
        for m in (1, 3, 4):
            ks = random_scalars(8500 + m + 100 * C.id, 1 << m, C.r)
            ks[1] = 0
            P = O.jac_points(ks)
            key = f"{C.name}_g1fftjac_{m}"
            g = np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)
            out[key + "_src"], out[key + "_gen"] = P, g
            out[key + "_fwd"] = R.g1_fft(m, g, P, coord="jac")
            out[key + "_inv"] = R.g1_fft(m, g, P, inverse=True, coord="jac")
            manifest["g1fft_jac"].append(key)
        # binary and multi-binary MSMs (bn128_G1_proj.c:684-782): byte coefficients, any non-zero
        # byte selects the point; points include an infinity, a repeated point and a negation
        rng = np.random.default_rng(13000 + C.id)
        for n in (1, 40, 300):
            ks, _ = msm_inputs(C, n, 13100 + n + 100 * C.id)
            P = O.points_from_scalars(ks)
            bits = (rng.random(n) < 0.5).astype(np.uint8) * rng.integers(1, 256, n).astype(np.uint8)
            if n >= 8:
                bits[3:8] = 1  # infinity, the repeated pair and the cancelling pair selected
            rows = (rng.random((21, n)) < 0.4).astype(np.uint8)
            rows[0] = 0
            rows[1] = 1
            key = f"{C.name}_binary_{n}"
            out[key + "_points"], out[key + "_bits"], out[key + "_rows"] = P, bits, rows
            out[key + "_affine"] = R.msm_binary(bits, P)
            out[key + "_jac_affine"] = R.msm_binary(bits, P, coord="jac")
            out[key + "_multi_affine"] = R.multi_msm_binary(rows, P)
            out[key + "_multi_jac_affine"] = R.multi_msm_binary(rows, P, coord="jac")
            manifest["binary"].append(key)
        # batch conversions (bn128_G1_proj.c:146-166, bn128_G1_jac.c:139-158)
        ks = random_scalars(13500 + 100 * C.id, 37, C.r)
        ks[0] = ks[20] = 0
        key = f"{C.name}_conv"
        out[key + "_proj"], out[key + "_jac"] = O.proj_points(ks), O.jac_points(ks)
        out[key + "_affine"] = O.points_from_scalars(ks)
        out[key + "_proj_to_affine"] = R.batch_to_affine(out[key + "_proj"])
        out[key + "_jac_to_affine"] = R.batch_to_affine(out[key + "_jac"], coord="jac")
        out[key + "_proj_from_affine"] = R.batch_from_affine(out[key + "_affine"])
        out[key + "_jac_from_affine"] = R.batch_from_affine(out[key + "_affine"], coord="jac")
        manifest["conv"].append(key)
        # polynomial evaluation and division by x^n - eta (bn128_poly_mont.c:225-243, :317-413)
        one = np.array(to_limbs(C.fr_to_mont(1), 4), np.uint64)
        for n1 in (1, 5, 64, 300):
            cs = random_scalars(14000 + n1 + 100 * C.id, n1, C.r)
            a = ints_to_array([C.fr_to_mont(c) for c in cs], 4)
            if n1 >= 5:
                a[-1] = 0  # a trailing zero coefficient: the degree is below the length
            x = np.array(to_limbs(C.fr_to_mont(random_scalars(14500 + n1, 1, C.r)[0]), 4), np.uint64)
            key = f"{C.name}_poly_{n1}"
            out[key + "_coeffs"], out[key + "_x"] = a, x
            out[key + "_eval"] = R.poly_eval(a, x)
            for n in (1, 3, 16, 64):
                for ename, eta in (("one", one), ("x", x)):
                    nq, nr = n1 + 2, n + 1
                    q, r = R.div_by_vanishing(a, n, eta, nq, nr)
                    out[f"{key}_div_{n}_{ename}_quot"], out[f"{key}_div_{n}_{ename}_rem"] = q, r
            manifest["poly"].append(key)
        # KZG opening (examples/KZG.hs:108-114) over an SRS tau^i G of 16 points
        tau = random_scalars(15000 + C.id, 1, C.r)[0]
        srs = O.points_from_scalars([pow(tau, i, C.r) for i in range(16)])
        cs = random_scalars(15100 + C.id, 16, C.r)
        a = ints_to_array([C.fr_to_mont(c) for c in cs], 4)
        z = np.array(to_limbs(C.fr_to_mont(random_scalars(15200 + C.id, 1, C.r)[0]), 4), np.uint64)
        v = R.poly_eval(a, z)
        pm = a.copy()
        pm[0] = O.f_sub(1, a[0], v)
        ok, q = R.quot_by_vanishing(pm, 1, z, 15)
        assert ok == 1
        key = f"{C.name}_kzg_16"
        out[key + "_srs"], out[key + "_coeffs"], out[key + "_z"], out[key + "_tau"] = srs, a, z, np.array(
            to_limbs(tau, 4), np.uint64)
        out[key + "_value"] = v
        out[key + "_proof"] = R.msm(q, srs[:15], mont=True)
        manifest["kzg"].append(key)
        # large NTTs: output digests only (inputs from seeded_fr_vector, taken as Montgomery words)
        for m in LARGE_NTT_LOGS:
            manifest["large"].append(large_ntt_entry(C, m))
        print(C.name, "done", flush=True)
    # large MSM: reference output for points k_i G and scalars s_i from seeded vectors
    for C, logn in LARGE_MSMS:
        manifest["large"].append(large_msm_entry(C, logn))
        print(C.name, "msm", logn, flush=True)
    np.savez_compressed(HERE / "golden_small.npz", **out)
    (HERE / "manifest.json").write_text(json.dumps(manifest, indent=1))


# digests of the reference NTT at these sizes (2^22: the BN254 forward + inverse workload)
LARGE_NTT_LOGS = (16, 20, 22)
# reference MSM results at these sizes (BN254 2^16: the plain single-MSM workload; 2^20: the
# BN254 and BLS12-381 MSMs of the headline and of the BLS12-381 timing workload)
LARGE_MSMS = ((BN, 14), (BN, 16), (BLS, 14), (BN, 20), (BLS, 20))


def large_ntt_entry(C, m):
    R = Ref(C)
    X = seeded_fr_vector(9000 + m + 100 * C.id, 1 << m, C.r)
    g = np.array(to_limbs(C.fr_to_mont(C.fft_root(m)), 4), np.uint64)
    return {"curve": C.name, "kind": "ntt_fwd", "m": m, "seed": 9000 + m + 100 * C.id,
            "sha256": sha(R.ntt(m, g, X)), "inv_sha256": sha(R.ntt(m, g, X, inverse=True))}


def _points_chunk(args):
    cid, rows = args
    C = BN if cid == BN.id else BLS
    return Oracle(C).points_from_scalars([int.from_bytes(r.tobytes(), "little") for r in rows])


def large_msm_entry(C, logn, workers=8):
    """Points k_i G (the oracle's scalar multiplication, in worker processes) and the reference's
    MSM over them; the points are stored as a digest (the tests rebuild them on the device)."""
    import multiprocessing as mp
    n = 1 << logn
    kseed, sseed = 11000 + logn + 100 * C.id, 12000 + logn + 100 * C.id
    K = seeded_fr_vector(kseed, n, C.r)
    S = seeded_fr_vector(sseed, n, C.r)
    chunks = [(C.id, K[i:i + 4096]) for i in range(0, n, 4096)]
    if len(chunks) > 1:
        with mp.get_context("fork").Pool(workers) as pool:
            P = np.concatenate(pool.map(_points_chunk, chunks))
    else:
        P = _points_chunk(chunks[0])
    res = Ref(C).msm(S, P)
    return {"curve": C.name, "kind": "msm_std_affine", "logn": logn, "kseed": kseed, "sseed": sseed,
            "result": [int(x) for x in res], "points_sha256": sha(P)}


def add_missing_large():
    """Adds the large entries the manifest lacks (same seeds and rules as a full run), leaving the
    small fixtures untouched."""
    assert ref() is not None, "build the reference first: make -C oracle ref"
    path = HERE / "manifest.json"
    manifest = json.loads(path.read_text())
    have = {(e["curve"], e["kind"], e.get("m", e.get("logn"))) for e in manifest["large"]}
    for C in (BN, BLS):
        for m in LARGE_NTT_LOGS:
            if (C.name, "ntt_fwd", m) not in have:
                manifest["large"].append(large_ntt_entry(C, m))
                print(C.name, "ntt", m, flush=True)
    for C, logn in LARGE_MSMS:
        if (C.name, "msm_std_affine", logn) not in have:
            manifest["large"].append(large_msm_entry(C, logn))
            print(C.name, "msm", logn, flush=True)
    path.write_text(json.dumps(manifest, indent=1))


if __name__ == "__main__":
    if "--large" in sys.argv[1:]:
        add_missing_large()
    else:
        main()
