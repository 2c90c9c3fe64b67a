"""MSM schedule autotuning and the learned cost model.

The reference fixes the MSM window from n alone (lib/cbits/curves/g1/proj/bn128_G1_proj.c:600,
after the op-count model of docs/algebraic_primitives.tex:659-683).  Here a schedule is a
window width c and a bucket partition seg_log (sorted entries per accumulation thread), and it
is chosen per job from (curve, n, scalar statistics):

* ``collect``  -- exhaustive autotuning: every candidate schedule of a job is run and timed
  (zkg_msm_tune, which also checks that all candidates give the same point), beside the plain
  path (analytic window on the declared scalar width);
* ``train``    -- fits the cost model, an MLP from the job/candidate features to ln(ms),
  trained by the library's own HIP forward/backward/Adam kernels (zkg_mlp_*);
* ``evaluate`` -- the network's pick against the analytic heuristics and the exhaustive best,
  on timings of jobs the network was not trained on.

A trained model is a flat little-endian float32 parameter file plus a JSON manifest; it is
installed into the library with ``Mlp.install()`` and used by the SCHED_LEARNED policy.

    python costmodel.py collect --curves 0 1 --logn 14 22 --seed 7 --out train.json
    python costmodel.py collect --curves 0 1 --mix 24 --seed 8 --out eval.json
    python costmodel.py train --data train.json --model tuned/msm_cost_gfx950
    python costmodel.py evaluate --data eval.json --model tuned/msm_cost_gfx950
    python costmodel.py ntt --out tuned/ntt_tuning_gfx950.json

The NTT has its own, smaller autotuner: per (curve, log size) the pass split bound smax and
the LDS tile size are timed exhaustively and the table is saved and reinstalled the same way.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
sys.path.insert(0, str(PKG))

import zkgpu  # noqa: E402

R = {0: 21888242871839275222246405745257275088548364400416034343698204186575808495617,
     1: 0x73eda753299d7d483339d80809a1d80553bda402fffe5bfeffffffff00000001}
DISTS = ("uniform", "u64", "sparse")
HIDDEN = (32, 32)
DEFAULT_MODEL = PKG / "tuned" / "msm_cost_gfx950"
NTT_TUNING = PKG / "tuned" / "ntt_tuning_gfx950.json"


def uniform_below(rng, n, r):
    """n integers uniform in [0, r) as (n, 4) little-endian u64 words (rejection sampling)."""
    bits = r.bit_length()
    out = np.empty((n, 4), np.uint64)
    filled = 0
    top_mask = np.uint64((1 << (bits - 192)) - 1)
    r_top = np.uint64(r >> 192)
    while filled < n:
        k = int((n - filled) * 1.4) + 16
        cand = rng.integers(0, 2**64, size=(k, 4), dtype=np.uint64)
        cand[:, 3] &= top_mask
        ok = cand[:, 3] < r_top
        for i in np.nonzero(cand[:, 3] == r_top)[0]:
            ok[i] = sum(int(cand[i, j]) << (64 * j) for j in range(4)) < r
        acc = cand[ok][: n - filled]
        out[filled:filled + len(acc)] = acc
        filled += len(acc)
    return out


def scalars(curve, n, dist, rng):
    """Standard-form scalars: uniform in [0, r); uniform 64-bit; or sparse (80% drawn from
    {0, 1} with equal odds, the rest uniform in [0, r))."""
    if dist == "uniform":
        return uniform_below(rng, n, R[curve])
    if dist == "u64":
        out = np.zeros((n, 4), np.uint64)
        out[:, 0] = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        return out
    if dist == "sparse":
        out = uniform_below(rng, n, R[curve])
        small = rng.random(n) < 0.8
        out[small] = 0
        out[small, 0] = rng.integers(0, 2, size=int(small.sum()), dtype=np.uint64)
        return out
    raise ValueError(f"unknown scalar distribution {dist}")


def _median_ms(fn, reps):
    fn()
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append((time.perf_counter() - t0) * 1e3)
    return float(np.median(t))


class PointSets:
    """One resident point set per curve (tau^i G, affine Montgomery), prefixes shared by jobs."""

    def __init__(self, max_n, seed=11):
        rng = np.random.default_rng(seed)
        self.max_n = max_n
        self.d = {}
        for c in (0, 1):
            tau = int(uniform_below(rng, 1, R[c])[0][0])
            self.d[c] = zkgpu.srs_generate(c, max_n, tau)


def tune_job(points, curve, n, dist, seed, reps=3, log=None):
    """Exhaustive autotuning of one job; returns its record."""
    rng = np.random.default_rng(seed)
    S = scalars(curve, n, dist, rng)
    dS = zkgpu.DeviceArray.from_host(S)
    dP = points.d[curve]
    st = zkgpu.scalar_stats(curve, n, dS)
    cands = zkgpu.msm_candidates(curve, st)
    ms, same = zkgpu.msm_tune(curve, n, dS, dP, cands, reps=reps)
    ref = zkgpu.msm_device(curve, n, dS, dP)
    plain_ms = _median_ms(lambda: zkgpu.msm_device(curve, n, dS, dP), reps)
    cand_res = zkgpu.msm_device(curve, n, dS, dP, window=cands[0][0], seg_log=cands[0][1],
                                policy=zkgpu.SCHED_STATS)
    rec = dict(curve=curve, n=n, dist=dist, seed=seed, stats=st.as_dict(), cands=cands,
               ms=[round(x, 5) for x in ms], identical=bool(same and (ref == cand_res).all()),
               plain_pick=list(zkgpu.msm_pick(curve, st, zkgpu.SCHED_PLAIN)),
               stats_pick=list(zkgpu.msm_pick(curve, st, zkgpu.SCHED_STATS)), plain_ms=round(plain_ms, 5))
    dS.free()
    if log:
        b = int(np.argmin(ms))
        log(f"tuned curve={curve} n={n} dist={dist}: plain {plain_ms:.3f} ms, best {ms[b]:.3f} ms at "
            f"c={cands[b][0]} seg={cands[b][1]} ({len(cands)} candidates, identical={rec['identical']})")
    return rec


def grid_jobs(curves, lo, hi, seed):
    jobs = []
    for c in curves:
        for lg in range(lo, hi + 1):
            for d in DISTS:
                jobs.append((c, 1 << lg, d, seed + 1000 * c + 10 * lg + DISTS.index(d)))
    return jobs


def mix_jobs(curves, count, seed, lo=14.0, hi=22.0):
    """count jobs: curve and distribution uniform, log2 n uniform in [lo, hi] (real-valued)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for k in range(count):
        c = int(curves[rng.integers(0, len(curves))])
        d = DISTS[int(rng.integers(0, 3))]
        n = int(round(2 ** rng.uniform(lo, hi)))
        jobs.append((c, n, d, seed * 7919 + k))
    return jobs


def features(rec, cands=None):
    st = zkgpu.stats_from(**rec["stats"])
    cands = cands if cands is not None else rec["cands"]
    return np.stack([zkgpu.msm_features(rec["curve"], st, c, s) for c, s in cands])


def dataset(records):
    X = np.concatenate([features(r) for r in records])
    y = np.concatenate([np.log(np.asarray(r["ms"], np.float64)) for r in records]).astype(np.float32)
    return X, y.reshape(-1, 1)


def train(records, epochs=4000, lr=3e-3, seed=1, hidden=HIDDEN):
    X, y = dataset(records)
    net = zkgpu.Mlp([zkgpu.MSM_NFEAT, *hidden, 1], seed=seed)
    loss = net.train(X, y, epochs, lr)
    return net, float(loss), X.shape[0]


def save(net, stem, meta):
    stem = Path(stem)
    stem.parent.mkdir(parents=True, exist_ok=True)
    p = net.get_params()
    p.astype("<f4").tofile(str(stem) + ".f32")
    man = dict(format="float32 little-endian, W_0 (out x in, row-major), b_0, W_1, b_1, ...",
               dims=net.dims, nparams=int(p.size), features=zkgpu.MSM_NFEAT, target="ln(ms) of one MSM", **meta)
    Path(str(stem) + ".json").write_text(json.dumps(man, indent=1))


def load(stem=DEFAULT_MODEL):
    stem = Path(stem)
    man = json.loads(Path(str(stem) + ".json").read_text())
    net = zkgpu.Mlp(man["dims"])
    p = np.fromfile(str(stem) + ".f32", dtype="<f4")
    net.set_params(p)
    return net, man


def evaluate(records, net):
    """Per job: measured ms of the plain path, of the analytic pick on the measured scalars, of
    the installed network's pick and of the exhaustive best (all from the job's timing table)."""
    net.install()
    rows = []
    for r in records:
        st = zkgpu.stats_from(**r["stats"])
        table = {tuple(c): m for c, m in zip(r["cands"], r["ms"])}
        pick = zkgpu.msm_pick(r["curve"], st, zkgpu.SCHED_LEARNED)
        sp = tuple(r["stats_pick"])
        best = min(r["ms"])
        rows.append(dict(curve=r["curve"], n=r["n"], dist=r["dist"], plain_ms=r["plain_ms"],
                         stats_ms=table[sp], mlp_ms=table[tuple(pick)], best_ms=best, mlp_pick=list(pick)))
    zkgpu.Mlp.uninstall()

    def agg(key):
        return dict(total_ms=round(sum(x[key] for x in rows), 3),
                    geomean_vs_best=round(math.exp(np.mean([math.log(x[key] / x["best_ms"]) for x in rows])), 4))
    return dict(jobs=len(rows), plain=agg("plain_ms"), stats_heuristic=agg("stats_ms"), mlp=agg("mlp_ms"),
                exhaustive=agg("best_ms"), rows=rows)


def tune_ntt(curves=(0, 1), lo=10, hi=22, reps=3, log=None):
    """NTT autotuning per (curve, log size): forward transform, every (smax, tile_log); the
    winner is installed in the library and returned."""
    table = []
    for c in curves:
        for m in range(lo, hi + 1):
            smax, tl, ms, same = zkgpu.ntt_tune(c, m, reps=reps, install=True)
            table.append(dict(curve=c, m=m, smax=smax, tile_log=tl, ms=round(ms, 5), identical=same))
            if log:
                log(f"ntt curve={c} m={m}: smax={smax} tile_log={tl} {ms:.4f} ms identical={same}")
    return table


def install_ntt_tuning(path=NTT_TUNING):
    """Installs a saved NTT tuning table; returns the number of entries (0 if none saved)."""
    path = Path(path)
    if not path.exists():
        return 0
    rows = json.loads(path.read_text())["table"]
    for r in rows:
        zkgpu.ntt_set_tuning(r["smax"], r["tile_log"], curve=r["curve"], m=r["m"])
    return len(rows)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    c = sub.add_parser("collect")
    c.add_argument("--curves", type=int, nargs="+", default=[0, 1])
    c.add_argument("--logn", type=int, nargs=2, default=[14, 22])
    c.add_argument("--mix", type=int, default=0, help="random mix of this many jobs instead of the grid")
    c.add_argument("--seed", type=int, default=7)
    c.add_argument("--reps", type=int, default=3)
    c.add_argument("--out", required=True)
    t = sub.add_parser("train")
    t.add_argument("--data", nargs="+", required=True)
    t.add_argument("--model", default=str(DEFAULT_MODEL))
    t.add_argument("--epochs", type=int, default=4000)
    t.add_argument("--lr", type=float, default=3e-3)
    nt = sub.add_parser("ntt")
    nt.add_argument("--curves", type=int, nargs="+", default=[0, 1])
    nt.add_argument("--logn", type=int, nargs=2, default=[10, 22])
    nt.add_argument("--reps", type=int, default=5)
    nt.add_argument("--out", default=str(NTT_TUNING))
    e = sub.add_parser("evaluate")
    e.add_argument("--data", nargs="+", required=True)
    e.add_argument("--model", default=str(DEFAULT_MODEL))
    e.add_argument("--out", default=None)
    a = ap.parse_args()

    def log(*x):
        print(*x, flush=True)

    if a.cmd == "ntt":
        table = tune_ntt(a.curves, a.logn[0], a.logn[1], a.reps, log)
        Path(a.out).write_text(json.dumps(dict(device=zkgpu.device_info(), table=table), indent=1))
        log(f"all schedules identical: {all(r['identical'] for r in table)}")
        return
    if a.cmd == "collect":
        jobs = (mix_jobs(a.curves, a.mix, a.seed, a.logn[0], a.logn[1]) if a.mix
                else grid_jobs(a.curves, a.logn[0], a.logn[1], a.seed))
        pts = PointSets(max(n for _, n, _, _ in jobs))
        log(f"{len(jobs)} jobs, points ready")
        recs = [tune_job(pts, cv, n, d, s, a.reps, log) for cv, n, d, s in jobs]
        Path(a.out).write_text(json.dumps(dict(device=zkgpu.device_info(), records=recs)))
        log(f"all candidates identical: {all(r['identical'] for r in recs)}")
    else:
        recs = [r for f in a.data for r in json.loads(Path(f).read_text())["records"]]
        if a.cmd == "train":
            net, loss, rows = train(recs, a.epochs, a.lr)
            save(net, a.model, dict(trained_on=[str(Path(f).name) for f in a.data], rows=rows, jobs=len(recs),
                                    epochs=a.epochs, lr=a.lr, final_mse=loss))
            log(f"trained on {rows} rows ({len(recs)} jobs): mse {loss:.5f}")
        else:
            net, _ = load(a.model)
            rep = evaluate(recs, net)
            log(json.dumps({k: v for k, v in rep.items() if k != "rows"}))
            if a.out:
                Path(a.out).write_text(json.dumps(rep, indent=1))


if __name__ == "__main__":
    main()
