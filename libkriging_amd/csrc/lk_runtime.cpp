#include "lk_runtime.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>

namespace lk {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

Runtime::Runtime(int device) : device_(device) {
  int count = 0;
  hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) throw std::runtime_error("no HIP device " + std::to_string(device));
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
}

Runtime::~Runtime() {
  hipSetDevice(device_);
  for (auto& g : graphs_)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (auto& kv : problems_) {
    hipFree(kv.second.X);
    hipFree(kv.second.y);
    hipFree(kv.second.F);
  }
  if (pool_) hipFree(pool_);
  if (d_slots_) hipFree(d_slots_);
  if (d_params_) hipFree(d_params_);
  if (d_out_) hipFree(d_out_);
  if (h_params_) hipHostFree(h_params_);
  if (h_out_) hipHostFree(h_out_);
  if (stream_) hipStreamDestroy(stream_);
}

int Runtime::add_problem(const double* X, int n, int d, const double* y, const double* F, int p, int kernel,
                         int noise) {
  if (n <= 0) throw std::invalid_argument("design must have at least one point");
  if (d <= 0 || d > DMAX) throw std::invalid_argument("input dimension must be in [1, " + std::to_string(DMAX) + "]");
  if (p < 0 || p > PMAX) throw std::invalid_argument("trend width must be in [0, " + std::to_string(PMAX) + "]");
  if (kernel < K_GAUSS || kernel > K_MATERN52) throw std::invalid_argument("unknown kernel id");
  if (noise != NOISE_NONE && noise != NOISE_NUGGET) throw std::invalid_argument("unknown noise model id");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  DeviceProblem P;
  P.n = n;
  P.d = d;
  P.p = p;
  P.kernel = kernel;
  P.noise = noise;
  hip_check(hipMalloc(&P.X, sizeof(double) * (size_t)n * d), "hipMalloc X");
  hip_check(hipMalloc(&P.y, sizeof(double) * (size_t)n), "hipMalloc y");
  hip_check(hipMalloc(&P.F, sizeof(double) * (size_t)n * std::max(p, 1)), "hipMalloc F");
  hip_check(hipMemcpy(P.X, X, sizeof(double) * (size_t)n * d, hipMemcpyHostToDevice), "upload X");
  hip_check(hipMemcpy(P.y, y, sizeof(double) * (size_t)n, hipMemcpyHostToDevice), "upload y");
  if (p > 0) hip_check(hipMemcpy(P.F, F, sizeof(double) * (size_t)n * p, hipMemcpyHostToDevice), "upload F");
  const int id = next_problem_++;
  problems_[id] = P;
  return id;
}

void Runtime::remove_problem(int id) {
  auto it = problems_.find(id);
  if (it == problems_.end()) throw std::invalid_argument("unknown problem id");
  hipSetDevice(device_);
  hipStreamSynchronize(stream_);
  hipFree(it->second.X);
  hipFree(it->second.y);
  hipFree(it->second.F);
  problems_.erase(it);
  // slots that pointed at it must be rebuilt
  slot_problem_.clear();
}

const DeviceProblem& Runtime::problem(int id) const {
  auto it = problems_.find(id);
  if (it == problems_.end()) throw std::invalid_argument("unknown problem id");
  return it->second;
}

void Runtime::ensure_slots(const std::vector<EvalRequest>& reqs) {
  const int ns = (int)reqs.size();
  std::vector<int> mapping(ns);
  for (int s = 0; s < ns; ++s) mapping[s] = reqs[s].problem;
  if (mapping == slot_problem_ && ns == nslots_) return;

  std::vector<size_t> off(ns);
  size_t total = 0;
  int nblk_max = 0, dmax = 0;
  h_slots_.assign(ns, Slot{});
  for (int s = 0; s < ns; ++s) {
    const DeviceProblem& P = problem(mapping[s]);
    const int npad = (P.n + NB - 1) / NB * NB;
    const int nblk = npad / NB;
    off[s] = total;
    total += align256(sizeof(double) * (size_t)npad * npad) * 2;       // A, W
    total += align256(sizeof(double) * (size_t)npad * (P.p + 7));       // V
    total += align256(sizeof(double) * red_size(nblk));                 // red
    nblk_max = std::max(nblk_max, nblk);
    dmax = std::max(dmax, P.d);
  }
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamSynchronize(stream_), "sync");
  if (total > pool_bytes_) {
    if (pool_) hip_check(hipFree(pool_), "hipFree pool");
    pool_ = nullptr;
    hip_check(hipMalloc(&pool_, total), "hipMalloc slot pool");
    pool_bytes_ = total;
  }
  char* base = reinterpret_cast<char*>(pool_);
  for (int s = 0; s < ns; ++s) {
    const DeviceProblem& P = problem(mapping[s]);
    const int npad = (P.n + NB - 1) / NB * NB;
    const int nblk = npad / NB;
    char* q = base + off[s];
    Slot& S = h_slots_[s];
    S.X = P.X;
    S.y = P.y;
    S.F = P.F;
    S.A = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * npad);
    S.W = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * npad);
    S.V = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * (P.p + 7));
    S.red = reinterpret_cast<double*>(q);
    S.n = P.n;
    S.npad = npad;
    S.nblk = nblk;
    S.d = P.d;
    S.p = P.p;
    S.kernel = P.kernel;
    S.noise = P.noise;
  }
  if ((size_t)ns > slot_cap_) {
    if (d_slots_) hipFree(d_slots_);
    if (d_params_) hipFree(d_params_);
    if (d_out_) hipFree(d_out_);
    if (h_params_) hipHostFree(h_params_);
    if (h_out_) hipHostFree(h_out_);
    const size_t cap = std::max<size_t>(ns, 8);
    hip_check(hipMalloc(&d_slots_, sizeof(Slot) * cap), "hipMalloc slots");
    hip_check(hipMalloc(&d_params_, sizeof(SlotParams) * cap), "hipMalloc params");
    hip_check(hipMalloc(&d_out_, sizeof(SlotOut) * cap), "hipMalloc out");
    hip_check(hipHostMalloc(&h_params_, sizeof(SlotParams) * cap, hipHostMallocDefault), "hipHostMalloc params");
    hip_check(hipHostMalloc(&h_out_, sizeof(SlotOut) * cap), "hipHostMalloc out");
    slot_cap_ = cap;
  }
  hip_check(hipMemcpy(d_slots_, h_slots_.data(), sizeof(Slot) * ns, hipMemcpyHostToDevice), "upload slots");
  slot_problem_ = mapping;
  nslots_ = ns;
  nblk_max_ = nblk_max;
  dmax_ = dmax;
  ++layout_version_;
}

void Runtime::launch_all(const Batch& B, bool want_grad) {
  hip_check(launch_build(B), "build");
  for (int k = 0; k < B.nblk_max; ++k) {
    hip_check(launch_potrf_diag(B, k), "potrf_diag");
    hip_check(launch_potrf_panel(B, k), "potrf_panel");
    hip_check(launch_potrf_update(B, k), "potrf_update");
  }
  const int lv = trtri_levels(B.nblk_max);
  for (int l = 0; l < lv; ++l) {
    hip_check(launch_trtri_level(B, 1 << l, 0), "trtri step 0");
    hip_check(launch_trtri_level(B, 1 << l, 1), "trtri step 1");
  }
  hip_check(launch_solve_fy(B), "solve_fy");
  hip_check(launch_gls(B), "gls");
  hip_check(launch_solve_e(B), "solve_e");
  hip_check(launch_solve_x(B), "solve_x");
  hip_check(launch_finalize_ll(B), "finalize_ll");
  hip_check(launch_rcond_estimate(B), "rcond_estimate");
  if (want_grad) {
    hip_check(launch_lauum_adjoint(B), "lauum_adjoint");
    hip_check(launch_finalize_grad(B), "finalize_grad");
  }
}

void Runtime::run_pipeline(bool want_grad, BenchMap* bench) {
  Batch B;
  B.slots = d_slots_;
  B.params = d_params_;
  B.out = d_out_;
  B.nslots = nslots_;
  B.nblk_max = nblk_max_;
  B.dmax = dmax_;
  B.stream = stream_;
  hip_check(hipMemcpyAsync(d_params_, h_params_, sizeof(SlotParams) * nslots_, hipMemcpyHostToDevice, stream_),
            "upload params");
  if (bench != nullptr) {
    // timed, uncaptured run: one event pair around the whole pipeline
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream_), "event record");
    launch_all(B, want_grad);
    hip_check(hipEventRecord(e1, stream_), "event record");
    hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    hipEventElapsedTime(&ms, e0, e1);
    (*bench)["gpu pipeline"] += ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  } else if (use_graph) {
    GraphEntry& G = graphs_[want_grad ? 1 : 0];
    if (G.version != layout_version_ || G.exec == nullptr) {
      if (G.exec) hip_check(hipGraphExecDestroy(G.exec), "graph exec destroy");
      G.exec = nullptr;
      hipGraph_t graph;
      hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
      try {
        launch_all(B, want_grad);
      } catch (...) {
        hipStreamEndCapture(stream_, &graph);
        throw;
      }
      hip_check(hipStreamEndCapture(stream_, &graph), "end capture");
      hip_check(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0), "graph instantiate");
      hip_check(hipGraphDestroy(graph), "graph destroy");
      G.version = layout_version_;
    }
    hip_check(hipGraphLaunch(G.exec, stream_), "graph launch");
  } else {
    launch_all(B, want_grad);
  }
  hip_check(hipMemcpyAsync(h_out_, d_out_, sizeof(SlotOut) * nslots_, hipMemcpyDeviceToHost, stream_),
            "download results");
  hip_check(hipStreamSynchronize(stream_), "stream sync");
}

void Runtime::evaluate(const std::vector<EvalRequest>& reqs, bool want_grad, std::vector<EvalResult>& res,
                       BenchMap* bench) {
  const int ns = (int)reqs.size();
  res.assign(ns, EvalResult{});
  if (ns == 0) return;
  for (int s = 0; s < ns; ++s) {
    const DeviceProblem& P = problem(reqs[s].problem);
    if (reqs[s].gamma_n != P.d && reqs[s].gamma_n != P.d + 1)
      throw std::invalid_argument("gamma must have d or d+1 entries (d = " + std::to_string(P.d) + ")");
  }
  hip_check(hipSetDevice(device_), "hipSetDevice");
  ensure_slots(reqs);
  for (int s = 0; s < ns; ++s) {
    const EvalRequest& R = reqs[s];
    const DeviceProblem& P = problem(R.problem);
    SlotParams& Q = h_params_[s];
    std::memset(&Q, 0, sizeof(Q));
    for (int k = 0; k < P.d; ++k) Q.theta[k] = R.gamma[k];
    double alpha = (R.gamma_n > P.d) ? R.gamma[P.d] : R.flags.alpha;
    // Kriging.cpp:241-242: nugget model with sigma2 and nugget both fixed
    if (P.noise == NOISE_NUGGET && !R.flags.est_sigma2 && !R.flags.est_nugget)
      alpha = R.flags.sigma2 / (R.flags.sigma2 + R.flags.nugget);
    Q.alpha = alpha;
    Q.sigma2 = R.flags.sigma2;
    Q.nugget = R.flags.nugget;
    Q.num_nugget = chol.num_nugget;
    Q.min_rcond = chol.min_rcond;
    Q.inc = 0;
    Q.active = 1;
    Q.est_sigma2 = R.flags.est_sigma2;
    Q.est_nugget = R.flags.est_nugget;
    Q.rcond_check = chol.rcond_check;
    Q.want_grad = want_grad;
    Q.gamma_has_alpha = R.gamma_n > P.d;
  }
  for (;;) {
    run_pipeline(want_grad, bench);
    bool retry = false;
    for (int s = 0; s < ns; ++s) {
      SlotParams& Q = h_params_[s];
      if (!Q.active) continue;
      const SlotOut& O = h_out_[s];
      const DeviceProblem& P = problem(reqs[s].problem);
      EvalResult& E = res[s];
      if (O.status == ST_OK) {
        E.ll = O.ll;
        const int gn = reqs[s].gamma_n;
        if (want_grad) {
          for (int k = 0; k < P.d; ++k) E.grad[k] = O.grad[k];
          if (gn > P.d) E.grad[P.d] = (P.noise == NOISE_NUGGET) ? O.grad[P.d] : 0.0;
        }
        for (int a = 0; a < P.p; ++a) E.beta[a] = O.beta[a];
        E.sse = O.sse;
        E.logdet = O.logdet;
        E.sigma2 = O.sigma2;
        E.rcond2 = O.rcond2;
        E.inc = Q.inc;
        E.status = EVAL_OK;
        Q.active = 0;
        continue;
      }
      // failed factorization or rcond below threshold: next step of the
      // ladder (LinearAlgebra.cpp:78-90)
      if (Q.inc > chol.max_inc) {
        E.status = EVAL_CHOL_FAILED;
        E.inc = Q.inc;
        Q.active = 0;
      } else if (chol.num_nugget <= 0.0) {
        E.status = EVAL_NUGGET_NOT_POSITIVE;
        E.inc = Q.inc;
        Q.active = 0;
      } else {
        Q.inc += 1;
        retry = true;
      }
    }
    if (!retry) break;
  }
  // leave every slot active for the next call's graph (params are rewritten)
}

Runtime::SlotView Runtime::slot_view(int s) const {
  if (s < 0 || s >= nslots_) throw std::out_of_range("slot index");
  const Slot& S = h_slots_[s];
  return SlotView{S.A, S.W, S.V, S.n, S.npad, S.p};
}

void Runtime::read_slot(int s, int what, double* out) {
  const SlotView v = slot_view(s);
  const int n = v.n, ld = v.npad;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamSynchronize(stream_), "sync");
  auto copy_cols = [&](const double* src, int ncols) {
    hip_check(hipMemcpy2D(out, sizeof(double) * n, src, sizeof(double) * ld, sizeof(double) * n, ncols,
                          hipMemcpyDeviceToHost),
              "read slot");
  };
  switch (what) {
    case 0:
    case 1: {
      copy_cols(what == 0 ? v.A : v.W, n);
      for (int j = 1; j < n; ++j)
        for (int i = 0; i < j; ++i) out[i + (size_t)j * n] = 0.0;
      break;
    }
    case 2:
      if (v.p > 0) copy_cols(v.V, v.p);
      break;
    case 3:
      copy_cols(v.V + (size_t)v.p * ld, 1);
      break;
    case 4:
      copy_cols(v.V + (size_t)(v.p + 2) * ld, 1);
      break;
    case 5:
      copy_cols(v.V + (size_t)(v.p + 3) * ld, 1);
      break;
    default:
      throw std::invalid_argument("unknown buffer id");
  }
}

}  // namespace lk
