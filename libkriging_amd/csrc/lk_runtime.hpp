// Host runtime of the batched likelihood pipeline: device memory for designs
// and batch slots, the launch sequence, hipGraph capture, and the
// jitter-retry ladder of LinearAlgebra::safe_chol_lower_retry
// (src/lib/LinearAlgebra.cpp:68-98) driven from the per-slot device status.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "lk_kernels.h"

namespace lk {

struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void hip_check(hipError_t e, const char* what);

// LinearAlgebra statics (LinearAlgebra.cpp:33,53,63,100)
struct CholConfig {
  double num_nugget = 1e-10;
  bool rcond_check = true;
  int max_inc = 10;
  double min_rcond = 1e-18;
};

// The reference's message for an evaluation whose factorisation failed
// (LinearAlgebra.cpp:78-86): "Exceed max numerical nugget (inc x 1e<log10
// num_nugget>) ..." or "Cannot add numerical nugget which is not strictly
// positive: <num_nugget>".  status is an EvalStatus.
std::string chol_failure_message(int status, int inc, double num_nugget);

// Estimation flags of a model (Kriging::_logLikelihood, Kriging.cpp:214-245)
struct ModelFlags {
  bool est_sigma2 = true;
  bool est_nugget = true;
  double sigma2 = 1.0;
  double nugget = 0.0;
  double alpha = 1.0;  // member alpha used when gamma has d entries
  // factorisation overrides of one request (NestedKriging's chol(R + 1e-10 I)
  // is inc0 = 1, max_inc = 0, rcond_check = 0: one fixed jitter, no retry)
  int inc0 = 0;          // jitter ladder step of the first attempt
  int max_inc = -1;      // last step retried (-1: CholConfig::max_inc)
  int rcond_check = -1;  // -1: CholConfig::rcond_check
};

struct EvalRequest {
  int problem = 0;
  const double* gamma = nullptr;  // d or d+1 entries
  int gamma_n = 0;
  ModelFlags flags;
  // an inactive request keeps its slot (same layout, same captured graph)
  // but is not evaluated; its result is left with ll = NaN
  bool active = true;
  // MODE_KERNEL (the likelihood), or, for a problem made by
  // add_matrix_problem, MODE_MATRIX (factor M with the retry ladder) /
  // MODE_FACTOR (M is the factor)
  int mode = MODE_KERNEL;
};

enum EvalStatus { EVAL_OK = 0, EVAL_CHOL_FAILED = 1, EVAL_NUGGET_NOT_POSITIVE = 2 };

struct EvalResult {
  double ll = 0.0;
  double grad[DMAX + 1] = {};
  double beta[PMAX] = {};
  double sse = 0.0, logdet = 0.0, sigma2 = 0.0, rcond2 = 0.0;
  int status = EVAL_OK;
  int inc = 0;
};

struct DeviceProblem {
  int n = 0, d = 0, p = 0, kernel = 0, noise = 0;
  double* X = nullptr;
  double* y = nullptr;
  double* F = nullptr;
  double* M = nullptr;  // n x n (ld = n) matrix of the matrix / factor modes
  bool owns_M = false;  // M allocated by add_matrix_problem (else the caller's)
  std::vector<double> range;  // max - min of each input column (prior of the LMP objective)
};

// Launch parameters of one size bucket: the slots whose matrices have at most
// max_nblk 64-row blocks.  Chosen per bucket by the autotuner
// (libkriging_amd/autotune.py), which writes them to the table the runtime
// loads at start (tuning_<arch>.json beside the library).
struct BucketTuning {
  int max_nblk = 0;       // upper bound of the bucket, in 64-row blocks
  int panel_blocks = 4;   // block columns per Cholesky panel
  int streams = 4;        // parts (1, 2, 4) a batch of this bucket alone is split into
  int wide = 0;           // trailing Cholesky updates on 128 x 64 tiles (else 64 x 64)
};

// Per-evaluation timings filled when a bench map is requested (Bench::toc
// semantics, src/lib/Bench.cpp:16-35), milliseconds.
using BenchMap = std::map<std::string, double>;

class Runtime {
 public:
  explicit Runtime(int device);
  ~Runtime();
  Runtime(const Runtime&) = delete;
  Runtime& operator=(const Runtime&) = delete;

  int add_problem(const double* X, int n, int d, const double* y, const double* F, int p, int kernel, int noise);
  // a problem whose matrix is given (n x n column-major, host or device
  // memory): evaluated with EvalRequest::mode MODE_MATRIX or MODE_FACTOR
  int add_matrix_problem(const double* M, int n, bool device_memory = false);

  // LinearAlgebra::safe_chol_lower (LinearAlgebra.cpp:43-98) on the
  // device: L (n x n, lower) of M + the ladder's jitter and the jitter
  // steps; throws the reference's messages when the ladder fails
  void safe_chol(const double* M, int n, double* L, int* inc);
  // LinearAlgebra::rcond_chol (LinearAlgebra.cpp:106-115): rcond(L)^2 by
  // the 1-norm estimator, for a given lower factor L
  double rcond_chol(const double* L, int n);
  // LinearAlgebra::chol_block (LinearAlgebra.cpp:257-299): the Cholesky
  // factor of C (n x n) from the factor Loo of its leading no x no block:
  // L_uo^T = Loo^-1 C_ou, L_uu = safe_chol(C_uu - L_uo L_uo^T), and, when
  // that fails, safe_chol(C) from scratch
  void chol_block(const double* C, int n, const double* Loo, int no, double* L);
  // the same on device buffers: C (n x n, ld = n), Loo (ld = ldo) and its
  // inverse LooInv (lower, ld = ldi, a multiple of 64 with identity padding,
  // diagonal blocks zero above the diagonal, as a slot's W) -> L (n x n,
  // ld = n).  Returns false when it fell back to the full factor of C.
  bool chol_block_device(const double* C, int n, const double* Loo, int ldo, const double* LooInv, int ldi, int no,
                         double* L);
  // the factor a MODE_FACTOR evaluation of problem id takes (n x n, ld = n,
  // device memory the caller keeps alive until the evaluation returns);
  // nullptr clears it
  void set_problem_factor(int id, const double* L);
  void remove_problem(int id);
  const DeviceProblem& problem(int id) const;

  // Evaluate LL (objective OBJ_LL), the leave-one-out objective (OBJ_LOO)
  // or the log marginal posterior (OBJ_LMP) -- EvalResult::ll holds the
  // objective's value -- and its gradient over gamma for every request in
  // one batched pipeline.  Per-request failures (chol still
  // failing after the retry ladder) are reported in EvalResult::status.
  void evaluate(const std::vector<EvalRequest>& reqs, bool want_grad, std::vector<EvalResult>& res,
                BenchMap* bench = nullptr, int objective = OBJ_LL);

  // The same LL evaluation with no device or pinned-host allocation: slot
  // tables and workspaces in the device's static scratch array, parameters
  // and results through host arrays of this object, launches issued
  // directly (no graph), one stream.  Throws when the batch needs more than
  // NOALLOC_SCRATCH_BYTES or has more than NOALLOC_MAX_ITEMS items.  The
  // allocating path's workspace is kept and laid out again by its next call.
  void evaluate_noalloc(const std::vector<EvalRequest>& reqs, bool want_grad, std::vector<EvalResult>& res);

  // Device pointers of slot s after the last evaluate(): the factor L (A,
  // lower, ld = npad), L^-1 (W lower) and the vector block V.
  struct SlotView {
    const double* A;
    const double* W;
    const double* V;
    int n, npad, p;
  };
  SlotView slot_view(int s) const;
  // host copy of a slot buffer after the last evaluate(); `what` as
  // LKGPU_READ_* in include/lkgpu.h
  void read_slot(int s, int what, double* out);
  // R^-1 (n x n, full, column-major) of slot s after the last evaluate() of
  // the L^-1 path; overwrites the slot's L
  void read_rinv(int s, double* out);

  CholConfig chol;
  bool use_graph = true;
  // most parts (1, 2 or 4) an untimed evaluation splits its batch into, one
  // stream each; a batch is split into the largest S <= this with slots a
  // multiple of 8 S.  Also set by LKGPU_STREAMS.
  void set_streams(int S);
  int streams() const { return streams_; }

  // Size buckets.  A batch whose slots fall in more than one bucket runs
  // as one part per bucket (at most 4, the smallest buckets merged), each on
  // its own stream with launches sized to its own largest matrix: the small
  // matrices of an uneven batch do not wait through the chain of the
  // largest one's diagonal blocks.  Slots are ordered by bucket; results are
  // indexed by request as always.
  void set_tuning(const std::vector<BucketTuning>& t);
  const std::vector<BucketTuning>& tuning() const { return tuning_; }
  // load a table ({"buckets": [{"max_nblk", "panel_blocks", "streams"}, ...]})
  void load_tuning_json(const std::string& text);
  int bucket_of(int nblk) const;
  // look-ahead Cholesky for one-slot batches of at least kLookaheadMinBlocks
  // blocks (on by default; LKGPU_LOOKAHEAD=0 or this call turns it off):
  // results are the same either way
  void set_lookahead(bool on) { lookahead_ = on; ++layout_version_; }
  bool lookahead_active() const { return lookahead_ && nslots_ == 1 && nblk_max_ >= kLookaheadMinBlocks; }
  // parts of the last run (slot ranges in bucket order) and their bucket
  int last_parts() const { return last_parts_; }

  // Live timing of one pipeline stage with a HIP event pair recorded on the
  // pipeline stream (inside the captured graph when graphs are used).
  // Stage ids as LKGPU_STAGE_* in include/lkgpu.h; -1 disables.
  void set_timed_stage(int stage);
  void take_stage_time(double* avg_ms, int* count);
  hipStream_t stream() const { return stream_; }
  int device() const { return device_; }

 private:
  void ensure_slots(const std::vector<EvalRequest>& reqs);
  int slot_of(int req) const;  // slot of request `req` of the last evaluate()
  void add_lmp_prior(const std::vector<EvalRequest>& reqs, std::vector<EvalResult>& res, bool want_grad) const;
  void run_pipeline(bool want_grad, BenchMap* bench, int objective = OBJ_LL);

 public:
  struct Step {
    unsigned stages;  // bit s set when the launch belongs to stage s
    const char* name;
    std::function<hipError_t()> fn;
  };

 private:
  std::vector<Step> pipeline_steps(const Batch& B, bool want_grad, int objective, int panel_blocks, int wide,
                                   bool lookahead = false) const;
  // a contiguous slot range run as one unit of launches
  struct Part {
    int first = 0, count = 0, nblk_max = 0, dmax = 0, panel_blocks = 4, wide = 0;
  };
  void apply_tuning(Part& q) const;
  std::vector<Part> plan_parts(bool timed, BenchMap* bench) const;
  Batch part_batch(const Part& q, hipStream_t s) const;
  hipGraph_t capture(const std::vector<Step>& v, size_t a, size_t b);

  int device_ = 0;
  hipStream_t stream_ = nullptr;
  std::map<int, DeviceProblem> problems_;
  int next_problem_ = 0;

  // slot workspace
  std::vector<int> slot_problem_;
  std::vector<int> slot_of_req_;      // request index -> slot (slots ordered by bucket)
  std::vector<Part> buckets_;         // slot ranges of the buckets present (<= kMaxStreams)
  std::vector<BucketTuning> tuning_;  // ascending max_nblk, the last one unbounded
  int last_parts_ = 1;
  // look-ahead Cholesky of a one-slot batch: the next panel's chain on the
  // main stream beside the bulk trailing update on la_stream_, ordered by the
  // two events.  la_stream_ is restricted to all CUs but la_reserve_ (a CU
  // mask), which stay free for the chain's launches; the mask is a property
  // of the stream's hardware queue and does not carry into graph replays,
  // so the look-ahead pipeline is issued without a graph.
  // la_workgroups_ > 0 makes the bulk update a persistent launch of that
  // many workgroups (slower: kept as an option).
  hipStream_t la_stream_ = nullptr;
  hipEvent_t la_ev_[2] = {nullptr, nullptr};
  int la_workgroups_ = 0;
  int la_reserve_ = 32;
  bool lookahead_ = true;
  // measured at one slot of n = 1024 .. 8192: a win from 128 blocks on
  static constexpr int kLookaheadMinBlocks = 128;
  // allocation-free calls: the device scratch array and the host arrays
  // that stand in for the pinned parameter / result tables
  char* scratch_ = nullptr;
  bool noalloc_ = false;
  SlotParams np_params_[NOALLOC_MAX_ITEMS] = {};
  SlotOut np_out_[NOALLOC_MAX_ITEMS] = {};
  std::vector<size_t> slot_bytes_;
  std::vector<Slot> h_slots_;
  Slot* d_slots_ = nullptr;
  SlotParams* h_params_ = nullptr;  // pinned
  SlotParams* d_params_ = nullptr;
  SlotOut* h_out_ = nullptr;        // pinned
  SlotOut* d_out_ = nullptr;
  size_t slot_cap_ = 0;
  double* pool_ = nullptr;
  size_t pool_bytes_ = 0;
  int nslots_ = 0, nblk_max_ = 0, dmax_ = 0;
  uint64_t layout_version_ = 0;

  // graph cache: (layout_version, objective, want_grad) -> exec
  struct GraphEntry {
    uint64_t version = ~0ull;
    int timed_stage = -1;
    int split = 1;
    hipGraphExec_t exec = nullptr;
  };
  GraphEntry graphs_[2 * OBJ_COUNT];
  int last_objective_ = OBJ_LL;
  int timed_stage_ = -1;
  // the batch split into S equal parts on S streams: the latency-bound
  // launches of one part (diagonal blocks, panels, finalisation) run beside
  // the MFMA launches of the others
  int streams_ = 4;
  static constexpr int kMaxStreams = 4;
  hipStream_t xstream_[kMaxStreams - 1] = {};  // streams of parts 1 .. S-1
  hipEvent_t fork_ = nullptr;
  hipEvent_t join_[kMaxStreams - 1] = {};
  void issue_split(const std::vector<Step>* parts, int S);
  hipEvent_t tev_[2] = {nullptr, nullptr};
  double timed_sum_ms_ = 0.0;
  int timed_count_ = 0;
};

}  // namespace lk
