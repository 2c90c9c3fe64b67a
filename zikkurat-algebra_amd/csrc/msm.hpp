#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/zkgpu.h"

namespace zk {

// Final MSM value in XYZZ Montgomery coordinates on the host.
struct MsmResult {
  int curve = 0;
  uint64_t x[6] = {0}, y[6] = {0}, zz[6] = {0}, zzz[6] = {0};
};

// Resolved schedule of one MSM.
// Windows 0..wide-1 are c bits wide, the others c-1 bits.  With measured scalars (largest bit
// length known) the widths are balanced: W = ceil((bits + 1) / c_requested), c = ceil((bits +
// 1) / W), together exactly bits + 1, so no window is left with only a few bits (a few
// buckets holding most of the entries).  With only the declared width known, every window is
// c_requested wide and the top one holds the few high bits the scalars may have (usually only
// the signed-digit carry).
struct MsmSchedule {
  int c = 0;          // widest window; 2^(c-1) buckets per window
  int windows = 0;    // number of signed-digit windows
  int wide = 0;       // windows of width c (the rest have c - 1)
  int seg_log = 5;    // log2 sorted entries per accumulation thread
  int bits = 0;       // scalar bits the windows cover
  int reduce = 0;     // scalars are reduced mod r before the digits (BN254: r P = 0 on the curve)
  int glv = 0;        // BN254 GLV split: 2n points (P_i, phi(P_i)) with scalars below 2^127
  int start(int w) const { return w < wide ? w * c : wide * c + (w - wide) * (c - 1); }
};
// the window layout for a requested width (balanced or uniform, see above)
void msm_window_layout(int bits, int c_req, bool balanced, int *W, int *c, int *wide);

// analytic window choice: minimises W * (10 n + 28 * 2^(c-1)) group-op cost (madd ~ 10
// mulmods, bucket pyramid ~ 2 adds of 14 mulmods per bucket) over c in [1, 20]
int msm_heuristic_window(int curve, long long n, int scalar_bits);
// the same cost model on measured scalars: windows cover max_bits, zeros cost nothing, ones
// land in window 0 only, and each digit costs a little sort work
int msm_stats_window(int curve, const zkg_scalar_stats &st);
int msm_scalar_stats(int curve, int n, const uint64_t *d_scalars, int enl, int mont, hipStream_t st,
                     zkg_scalar_stats *out);
// cost-model choice for a job (false when no model is installed).  glv_ok as for msm_resolve:
// the candidates and features then describe the GLV shape the job runs with.
bool cost_model_pick(int curve, const zkg_scalar_stats &st, bool glv_ok, int *c, int *seg_log);
void msm_features(int curve, const zkg_scalar_stats &st, int c, int seg_log, bool glv_ok, float *x);
int msm_candidates(int curve, const zkg_scalar_stats &st, bool glv_ok, int max, int *cs, int *segs);
// glv_ok: a BN254 job may take the GLV split (2n points P_i and phi(P_i), scalars below 2^127;
// affine words are converted with phi in the digit pass, prepared sets carry phi): under the
// plain schedule always, under the statistics policies when the measured scalars are wider
// than 128 bits
int msm_resolve(int curve, int n, int enl, const zkg_msm_schedule *user, const zkg_scalar_stats *st,
                MsmSchedule *out, bool glv_ok = false);

void msm_infinity(int curve, MsmResult *r);
int msm_write_result(int curve, const MsmResult &r, uint64_t *tgt, int flags);

// Device MSM; d_scalars: n * enl u64 (standard form, or Montgomery Fr when mont != 0),
// d_points: n affine Montgomery points.  Synchronous; the result is left in *out.
// points_internal: d_points already holds the reduced-radix form written by msm_prepare_points
int msm_device(int curve, int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
               bool points_internal, const zkg_msm_schedule *sched, hipStream_t st, MsmResult *out);
// external affine points -> internal records (msm_internal_point_bytes each); BN254 records
// come in (P_i, phi(P_i)) pairs, the phi half written only when phi is set (MSMs take the GLV
// split on a prepared set with it; a binary sum never reads it)
int msm_prepare_points(int curve, int n, const uint64_t *d_ext, void *d_int, hipStream_t st, bool phi = true);
size_t msm_internal_point_bytes(int curve);
size_t msm_footprint(int curve, int n, int enl, bool points_internal, const zkg_msm_schedule *sched);

int g1_scalar_mul_gen(int curve, int n, const uint64_t *d_scalars, uint64_t *d_out, hipStream_t st);
// batch conversions (external Montgomery words) and binary-coefficient MSMs
int g1_batch_to_affine(int curve, bool jac, int n, const uint64_t *d_src, uint64_t *d_dst, hipStream_t st);
int g1_batch_from_affine(int curve, bool jac, int n, const uint64_t *d_src, uint64_t *d_dst, hipStream_t st);
// sums of the points selected by the non-zero bytes of each of nrows rows (row r at
// d_bits + r * n); out holds nrows results
int binary_sum_device(int curve, int nrows, int n, const uint8_t *d_bits, const uint64_t *d_points,
                      bool points_internal, hipStream_t st, MsmResult *out);
// polynomial evaluation (result to host memory), degree, and division by x^n - eta
int poly_eval_device(int curve, int n, const uint64_t *d_coeffs, const uint64_t *loc, uint64_t *out_host,
                     hipStream_t st);
int poly_degree_device(int n, const uint64_t *d_coeffs, int *deg_host, hipStream_t st);
int vanish_div_device(int curve, int D, const uint64_t *d_a, int n, const uint64_t *eta, uint64_t *d_quot,
                      uint64_t *d_rem, hipStream_t st);
// FFT of 2^m projective G1 points (external Montgomery words), normalised outputs
int g1_fft_device(int curve, int m, const uint64_t *gen, bool inverse, bool jac, const uint64_t *d_src,
                  uint64_t *d_dst, hipStream_t st);
int srs_generate(int curve, int n, const uint64_t *tau_std, uint64_t *d_out, hipStream_t st);

struct Server;
int serve_create(int curve, const void *d_srs_internal, int srs_n, size_t budget, int nworkers, Server **out);
int serve_submit(Server *s, int logn, const uint64_t *values, uint64_t *out, long long *id);
int serve_set_exclusive_log(Server *s, int logn);
int serve_drain(Server *s);
int serve_job_times(Server *s, long long id, double *queue_ms, double *run_ms, double *latency_ms);
int serve_stats(Server *s, size_t *peak_bytes, int *max_concurrent);
int serve_run_trace(Server *s, int njobs, const int *logn, const double *arrival_s, const uint64_t *const *values,
                    uint64_t *const *outs, double *latency_ms, double *wall_s);
void serve_destroy(Server *s);

int msm_tune(int curve, int n, const uint64_t *d_scalars, int enl, int mont, const uint64_t *d_points,
             bool points_internal, int ncand, const int *cs, const int *segs, int reps, double *ms_out,
             int *identical, hipStream_t st);
int msm_batch(int njobs, const int *curves, const int *ns, const uint64_t *const *d_scalars, int enl, int mont,
              const uint64_t *const *d_points, bool points_internal, uint64_t *const *outs, int out_flags,
              const zkg_msm_schedule *sched, int nstreams, double *wall_ms);

// captured batch of MSMs (one hipGraph over nstreams concurrent branches); run() replays it
// and writes every job's result to outs[k] (host words, msm_write_result flags)
struct MsmGraph;
int msm_graph_create(int njobs, const int *curves, const int *ns, const uint64_t *const *d_scalars, int enl, int mont,
                     const uint64_t *const *d_points, bool points_internal, const zkg_msm_schedule *sched,
                     int nstreams, MsmGraph **out);
int msm_graph_run(MsmGraph *g, uint64_t *const *outs, int out_flags, double *wall_ms);
void msm_graph_destroy(MsmGraph *g);

int msm_last_timing(double *ms_total, double *ms_phases, int max_phases, int *nphases);
int msm_set_phase_timing(int on);
int msm_last_stats(long long *out4);

}  // namespace zk
