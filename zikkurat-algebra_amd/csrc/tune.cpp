// Schedule selection for the MSM: candidate enumeration, the cost-model input encoding, the
// learned pick, the exhaustive autotuner, and batched MSMs of uneven sizes.
//
// A schedule is (window width c, seg_log): c sets the number of signed-digit windows
// W = ceil((bits + 1) / c) and buckets per window 2^(c-1); seg_log sets how many sorted bucket
// entries one accumulation thread sums (the bucket partition of the load-balanced
// accumulation).  The reference hard-codes its window (lib/cbits/curves/g1/proj/
// bn128_G1_proj.c:600, round(log2 n - 3.5) per docs/algebraic_primitives.tex:659-683); every
// schedule here gives the same point, only the time differs.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <map>
#include <string>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "../../include/zkgpu.h"
#include "mlp.hpp"
#include "msm.hpp"
#include "runtime.hpp"

namespace zk {

static_assert(ZKG_MSM_NFEAT == 11, "feature count");

// The bucket problem a job runs as (msm_resolve): with the GLV split (BN254, 4-limb scalars
// wider than 128 bits, points given as affine words) 2n points P_i, phi(P_i) with scalar halves
// below 2^127, each full scalar giving a digit per window in both halves; otherwise n points
// with windows over the measured bits.
struct MsmShape {
  int bits;
  double points, per_full;  // bucket-problem points; nonzero digits per window of a full scalar
};
static MsmShape msm_shape(const zkg_scalar_stats &st, bool glv) {
  const double n = st.n > 0 ? (double)st.n : 1.0;
  if (glv) return MsmShape{127, 2.0 * n, 2.0};
  return MsmShape{st.max_bits, n, 1.0};
}
static bool msm_glv_shape(int curve, const zkg_scalar_stats &st, bool glv_ok) {
  return glv_ok && curve == 0 && st.max_bits > 128;
}

// Cost-model input row.  Every feature is a plain function of the job and the candidate as it
// runs (the GLV shape included); the last one is the analytic cost of msm_stats_window, so the
// network learns a correction to it.
void msm_features(int curve, const zkg_scalar_stats &st, int c_req, int seg_log, bool glv_ok, float *x) {
  const bool glv = msm_glv_shape(curve, st, glv_ok);
  const MsmShape sh = msm_shape(st, glv);
  int w, c, wide;
  msm_window_layout(sh.bits, c_req, true, &w, &c, &wide);
  const double n = st.n > 0 ? (double)st.n : 1.0;
  const double nfull = (double)(st.n - st.n_zero - st.n_one), ones = (double)st.n_one;
  const double W = (double)w;
  const double entries = W * sh.per_full * nfull + ones;
  const double buckets = W * (double)(1ll << (c - 1));
  const double threads = entries / (double)(1ll << seg_log) + 1.0;
  const double cost = 10.0 * entries + 3.0 * W * sh.points + 28.0 * buckets;
  x[0] = (float)curve;
  x[1] = (float)(log2(sh.points) / 24.0);
  x[2] = (float)(sh.bits / 256.0);
  x[3] = (float)((st.n_zero + st.n_one) / n);
  x[4] = (float)(c / 20.0);
  x[5] = (float)(seg_log / 8.0);
  x[6] = (float)(W / 64.0);
  x[7] = (float)(log2(entries + 1.0) / 32.0);
  x[8] = (float)(log2(buckets) / 32.0);
  x[9] = (float)(log2(threads) / 32.0);
  x[10] = (float)(log2(cost + 1.0) / 48.0);
}

// Windows around both analytic choices (declared width and measured scalars; with the GLV
// split, the choice for 2n points of 127 bits) up to 18 bits, seg_log 2..7; requested widths
// with the same balanced layout on the bits the job runs with are one candidate (listed by
// their effective width).
int msm_candidates(int curve, const zkg_scalar_stats &st, bool glv_ok, int max, int *cs, int *segs) {
  const bool glv = msm_glv_shape(curve, st, glv_ok);
  const MsmShape sh = msm_shape(st, glv);
  const int fr_bits = curve == 0 ? 254 : 255;
  const int hp = glv ? msm_heuristic_window(curve, 2ll * st.n, 127) : msm_heuristic_window(curve, st.n, fr_bits);
  const int hs = glv ? hp : msm_stats_window(curve, st);
  const int lo = std::max(2, std::min(hp, hs) - 5), hi = std::min(18, std::max(hp, hs) + 3);
  int k = 0, last = -1;
  for (int c = lo; c <= hi; c++) {
    int w, cm, wide;
    msm_window_layout(sh.bits, c, true, &w, &cm, &wide);
    if (cm == last) continue;
    last = cm;
    for (int s = 2; s <= 7; s++) {
      if (k >= max) return k;
      cs[k] = cm;
      segs[k] = s;
      k++;
    }
  }
  return k;
}

// ---------------------------------------------------------------------------------------
// installed cost model

static std::mutex g_model_mu;
static Mlp *g_model = nullptr;
static std::map<std::tuple<int, long long, int, int, int, bool>, std::pair<int, int>> g_pick_cache;

int cost_model_set(Mlp *m) {
  std::lock_guard<std::mutex> lk(g_model_mu);
  if (m && (mlp_dim(m, 0) != ZKG_MSM_NFEAT || mlp_dim(m, mlp_layers(m)) != 1)) {
    set_error("cost model: the network must map ZKG_MSM_NFEAT features to one output");
    return ZK_EINVAL;
  }
  g_model = m;
  g_pick_cache.clear();
  return ZK_OK;
}

void cost_model_forget(Mlp *m) {
  std::lock_guard<std::mutex> lk(g_model_mu);
  if (g_model == m) {
    g_model = nullptr;
    g_pick_cache.clear();
  }
}

bool cost_model_pick(int curve, const zkg_scalar_stats &st, bool glv_ok, int *c, int *seg_log) {
  std::lock_guard<std::mutex> lk(g_model_mu);
  if (!g_model) return false;
  // jobs with the same size, width, shape and (to 1/64) share of zeros and ones share a decision
  const auto key = std::make_tuple(curve, st.n, st.max_bits, (int)(64.0 * st.n_zero / std::max(1ll, st.n)),
                                   (int)(64.0 * st.n_one / std::max(1ll, st.n)), glv_ok);
  auto it = g_pick_cache.find(key);
  if (it != g_pick_cache.end()) {
    *c = it->second.first;
    *seg_log = it->second.second;
    return true;
  }
  int cs[256], ss[256];
  const int k = msm_candidates(curve, st, glv_ok, 256, cs, ss);
  if (k == 0) return false;
  std::vector<float> X((size_t)k * ZKG_MSM_NFEAT), y(k);
  for (int i = 0; i < k; i++)
    msm_features(curve, st, cs[i], ss[i], glv_ok, X.data() + (size_t)i * ZKG_MSM_NFEAT);
  if (mlp_forward(g_model, k, X.data(), y.data()) != ZK_OK) return false;
  const int best = (int)(std::min_element(y.begin(), y.end()) - y.begin());
  *c = cs[best];
  *seg_log = ss[best];
  g_pick_cache[key] = {*c, *seg_log};
  return true;
}

// ---------------------------------------------------------------------------------------
// exhaustive autotuner

int msm_tune(int curve, int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
             bool points_internal, int ncand, const int *cs, const int *segs, int reps, double *ms_out,
             int *identical, hipStream_t st) {
  uint64_t first[12] = {0}, cur[12] = {0};
  bool same = true;
  reps = std::max(1, reps);
  std::vector<double> t(reps);
  for (int k = 0; k < ncand; k++) {
    zkg_msm_schedule s;
    memset(&s, 0, sizeof s);
    s.window = cs[k];
    s.seg_log = segs[k];
    s.policy = ZKG_SCHED_STATS;
    MsmResult r;
    for (int rep = -1; rep < reps; rep++) {
      const auto t0 = std::chrono::steady_clock::now();
      ZK_TRY(msm_device(curve, n, d_scalars, enl, mont, d_points, points_internal, &s, st, &r));
      const auto t1 = std::chrono::steady_clock::now();
      if (rep >= 0) t[rep] = std::chrono::duration<double, std::milli>(t1 - t0).count();
    }
    std::sort(t.begin(), t.end());
    ms_out[k] = t[reps / 2];
    ZK_TRY(msm_write_result(curve, r, k == 0 ? first : cur, 0));
    if (k > 0 && memcmp(first, cur, sizeof first) != 0) same = false;
  }
  if (identical) *identical = same ? 1 : 0;
  return ZK_OK;
}

// ---------------------------------------------------------------------------------------
// batched MSMs of uneven sizes: worker threads, one stream each, largest job first

int msm_batch(int njobs, const int *curves, const int *ns, const uint64_t *const *d_scalars, int enl, int mont,
              const uint64_t *const *d_points, bool points_internal, uint64_t *const *outs, int out_flags,
              const zkg_msm_schedule *sched, int nstreams, double *wall_ms) {
  std::vector<int> order(njobs);
  for (int k = 0; k < njobs; k++) order[k] = k;
  // largest work first (a BLS12-381 point op costs about twice a BN254 one)
  auto work = [&](int k) { return (double)ns[k] * (curves[k] == 1 ? 2.0 : 1.0); };
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return work(a) > work(b); });
  zkg_msm_schedule s;
  memset(&s, 0, sizeof s);
  s.policy = sched ? sched->policy : ZKG_SCHED_PLAIN;
  std::atomic<int> next{0}, status{ZK_OK};
  std::mutex err_mu;
  std::string err;
  nstreams = std::max(1, std::min(nstreams, 16));
  const auto t0 = std::chrono::steady_clock::now();
  auto worker = [&]() {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      status = ZK_EHIP;
      return;
    }
    for (;;) {
      const int i = next.fetch_add(1);
      if (i >= njobs || status.load() != ZK_OK) break;
      const int k = order[i];
      MsmResult r;
      int e = msm_device(curves[k], ns[k], d_scalars[k], enl, mont, d_points[k], points_internal, &s, st, &r);
      if (e == ZK_OK) e = msm_write_result(curves[k], r, outs[k], out_flags);
      if (e != ZK_OK) {
        std::lock_guard<std::mutex> lk(err_mu);
        err = last_error();
        status = e;
      }
    }
    (void)hipStreamDestroy(st);
  };
  std::vector<std::thread> th;
  for (int w = 0; w < nstreams; w++) th.emplace_back(worker);
  for (auto &x : th) x.join();
  if (status.load() != ZK_OK) set_error(err.empty() ? "msm batch: worker stream creation failed" : err);
  if (wall_ms) *wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return status.load();
}

}  // namespace zk
