#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace zk {

enum NttFlags : int {
  NTT_FORWARD = 0,
  NTT_INVERSE = 1,
};

// Schedule knobs of the fast path: passes of at most 2^smax points, LDS tiles of 2^tile_log
// field elements per workgroup.
struct NttTuning {
  int smax;
  int tile_log;
};

struct NttPlan {
  int m = 0;
  int npass = 0;
  int s[8] = {0};
  int tile_log = 11;
};

NttPlan ntt_plan(int m, int smax, int tile_log);
void ntt_set_tuning(const NttTuning &t);  // default for every size; clears the per-size table
NttTuning ntt_get_tuning();
void ntt_set_tuning_for(int curve, int m, const NttTuning &t);
NttTuning ntt_tuning_for(int curve, int m);
// times every (smax, tile_log) for one size; best is not installed
int ntt_tune(int curve, int m, int inverse, int reps, NttTuning *best, double *best_ms, int *identical,
             hipStream_t st);

// true when gen has multiplicative order exactly 2^m (the fast path's precondition)
bool ntt_generator_ok(int curve, int m, const uint64_t *gen);

// Device-buffer NTT.  gen/eta are host pointers to Montgomery Fr (4 x u64); eta == nullptr
// means no coset shift.  d_in may equal d_out.  Enqueued on `st`; returns after the work is
// complete (temporary buffers are released).
int ntt_device(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const void *d_in,
               void *d_out, hipStream_t st, const NttTuning *tune);

// The same transform with no device allocation: scratch in a fixed file-scope device array
// (m <= 18, generators of order 2^m); calls that overlap on one GPU race.
int ntt_device_noalloc(int curve, int m, const uint64_t *gen, const uint64_t *eta, int flags, const void *d_in,
                       void *d_out, hipStream_t st);

struct NttBatchJob {
  int m;
  int inverse;
  const uint64_t *eta;  // host Montgomery Fr coset shift, or nullptr
  const void *d_src;
  void *d_tgt;
};
struct NttBatchPlan;
int ntt_batch_create(int curve, int njobs, const NttBatchJob *jobs, NttBatchPlan **out);
int ntt_batch_run(NttBatchPlan *pl, double *ms);
void ntt_batch_destroy(NttBatchPlan *pl);
uint64_t ntt_batch_bytes(const NttBatchPlan *pl);

}  // namespace zk
