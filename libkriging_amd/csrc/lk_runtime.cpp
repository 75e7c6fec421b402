#include "lk_runtime.hpp"

#include <dlfcn.h>

#include <algorithm>
#include <climits>
#include <fstream>
#include <sstream>
#include <cmath>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <limits>

#include "model_io.hpp"

namespace lk {

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

std::string chol_failure_message(int status, int inc, double num_nugget) {
  if (status == EVAL_NUGGET_NOT_POSITIVE)
    return "[ERROR] Cannot add numerical nugget which is not strictly positive: " + std::to_string(num_nugget);
  return "[ERROR] Exceed max numerical nugget (" + std::to_string(inc) + " x 1e" +
         std::to_string(std::log10(num_nugget)) + ") added to force chol matrix";
}

static size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

Runtime::Runtime(int device) : device_(device) {
  int count = 0;
  hip_check(hipGetDeviceCount(&count), "hipGetDeviceCount");
  if (device < 0 || device >= count) throw std::runtime_error("no HIP device " + std::to_string(device));
  hip_check(hipSetDevice(device), "hipSetDevice");
  hip_check(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking), "hipStreamCreate");
  for (auto& e : la_ev_) hip_check(hipEventCreateWithFlags(&e, hipEventDisableTiming), "hipEventCreate");
  {
    int cus = 0;
    hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    if (const char* e = std::getenv("LKGPU_LA_RESERVE")) la_reserve_ = std::max(0, std::atoi(e));
    if (la_reserve_ > 0 && la_reserve_ < cus) {
      // the bulk update's stream on the first cus - la_reserve_ CUs
      // (n = 8192 Cholesky: 6.44 / 6.44 / 6.60 ms with 24 / 32 / 48 CUs kept
      // free, 7.00 ms without a mask, 7.05 ms without the look-ahead)
      std::vector<uint32_t> mask((cus + 31) / 32, 0u);
      for (int c = 0; c < cus - la_reserve_; ++c) mask[c / 32] |= 1u << (c % 32);
      hip_check(hipExtStreamCreateWithCUMask(&la_stream_, (uint32_t)mask.size(), mask.data()),
                "hipExtStreamCreateWithCUMask");
    } else {
      hip_check(hipStreamCreateWithFlags(&la_stream_, hipStreamNonBlocking), "hipStreamCreate");
    }
  }
  if (const char* e = std::getenv("LKGPU_LOOKAHEAD")) lookahead_ = std::atoi(e) != 0;
  if (const char* e = std::getenv("LKGPU_LOOKAHEAD_WG")) la_workgroups_ = std::max(0, std::atoi(e));
  // the allocation-free entry point's scratch: its device address, looked
  // up once per device with this device current
  hip_check(noalloc_scratch(device, &scratch_), "hipGetSymbolAddress(scratch)");
  int S = 4;
  if (const char* e = std::getenv("LKGPU_STREAMS")) {
    const int v = std::atoi(e);
    S = (v == 2 || v == 4) ? v : 1;
  }
  set_streams(S);
  // size buckets of 4, 8, ..., 128 blocks and the rest; panel width and
  // stream parts from the tuned table when one is found, else 4 and 4
  std::vector<BucketTuning> t;
  for (int b : {4, 8, 16, 32, 64, 128, INT_MAX}) t.push_back(BucketTuning{b, 4, 4});
  set_tuning(t);
  std::string path;
  if (const char* e = std::getenv("LKGPU_TUNING")) {
    path = e;
  } else {
    Dl_info info;
    if (dladdr(reinterpret_cast<void*>(&hip_check), &info) != 0 && info.dli_fname != nullptr) {
      std::string so = info.dli_fname;
      const size_t slash = so.rfind('/');
      path = (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/tuning_gfx950.json";
    }
  }
  if (!path.empty() && path != "none") {
    std::ifstream f(path);
    if (f) {
      std::stringstream ss;
      ss << f.rdbuf();
      load_tuning_json(ss.str());
    }
  }
  // one panel width for every bucket (experiments)
  if (const char* e = std::getenv("LKGPU_PANEL_BLOCKS")) {
    const int pb = std::max(1, std::atoi(e));
    for (BucketTuning& b : tuning_) b.panel_blocks = pb;
  }
}

void Runtime::set_tuning(const std::vector<BucketTuning>& t) {
  if (t.empty()) throw std::invalid_argument("tuning table must have at least one bucket");
  std::vector<BucketTuning> v(t);
  std::sort(v.begin(), v.end(), [](const BucketTuning& a, const BucketTuning& b) { return a.max_nblk < b.max_nblk; });
  for (const BucketTuning& b : v) {
    if (b.max_nblk < 1) throw std::invalid_argument("bucket bound must be positive");
    if (b.panel_blocks < 1 || b.panel_blocks > 64) throw std::invalid_argument("panel_blocks must be in [1, 64]");
    if (b.streams != 1 && b.streams != 2 && b.streams != 4) throw std::invalid_argument("streams must be 1, 2 or 4");
  }
  v.back().max_nblk = INT_MAX;
  tuning_ = v;
  slot_problem_.clear();  // the next batch recomputes its bucket layout
}

void Runtime::load_tuning_json(const std::string& text) {
  const io::Json j = io::parse(text);
  std::vector<BucketTuning> t;
  for (const io::Json& b : j.at("buckets").arr) {
    BucketTuning q;
    const double mx = b.at("max_nblk").type == io::Json::Null ? (double)INT_MAX : b.at("max_nblk").as_number();
    q.max_nblk = mx >= (double)INT_MAX ? INT_MAX : (int)mx;
    q.panel_blocks = (int)b.at("panel_blocks").as_number();
    q.streams = (int)b.at("streams").as_number();
    if (const io::Json* w = b.find("wide")) q.wide = w->as_number() != 0.0 ? 1 : 0;
    t.push_back(q);
  }
  set_tuning(t);
}

int Runtime::bucket_of(int nblk) const {
  for (size_t q = 0; q < tuning_.size(); ++q)
    if (nblk <= tuning_[q].max_nblk) return (int)q;
  return (int)tuning_.size() - 1;
}

int Runtime::slot_of(int req) const {
  if (req < 0 || req >= (int)slot_of_req_.size()) throw std::invalid_argument("item index out of range");
  return slot_of_req_[req];
}

void Runtime::set_streams(int S) {
  if (S != 1 && S != 2 && S != 4) throw std::invalid_argument("streams must be 1, 2 or 4");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  if (fork_ == nullptr) hip_check(hipEventCreateWithFlags(&fork_, hipEventDisableTiming), "hipEventCreate");
  // all three extra streams whatever S: with the look-ahead's CU-masked
  // stream created before them, a context holding only one extra stream ran
  // its two-part batches 1.5x slower (3.68 vs 2.49 ms per 64 x n = 1024
  // batch); with three, as by default, it does not
  for (int q = 0; q + 1 < kMaxStreams; ++q) {
    if (xstream_[q] == nullptr)
      hip_check(hipStreamCreateWithFlags(&xstream_[q], hipStreamNonBlocking), "hipStreamCreate");
    if (join_[q] == nullptr) hip_check(hipEventCreateWithFlags(&join_[q], hipEventDisableTiming), "hipEventCreate");
  }
  streams_ = S;
}

Runtime::~Runtime() {
  hipSetDevice(device_);
  for (auto& e : tev_)
    if (e) hipEventDestroy(e);
  for (auto& g : graphs_)
    if (g.exec) hipGraphExecDestroy(g.exec);
  for (auto& kv : problems_) {
    hipFree(kv.second.X);
    hipFree(kv.second.y);
    hipFree(kv.second.F);
    if (kv.second.owns_M && kv.second.M) hipFree(kv.second.M);
  }
  if (pool_) hipFree(pool_);
  if (d_slots_) hipFree(d_slots_);
  if (d_params_) hipFree(d_params_);
  if (d_out_) hipFree(d_out_);
  if (h_params_) hipHostFree(h_params_);
  if (h_out_) hipHostFree(h_out_);
  if (fork_) hipEventDestroy(fork_);
  for (auto& e : la_ev_)
    if (e) hipEventDestroy(e);
  if (la_stream_) hipStreamDestroy(la_stream_);
  for (int q = 0; q < kMaxStreams - 1; ++q) {
    if (join_[q]) hipEventDestroy(join_[q]);
    if (xstream_[q]) hipStreamDestroy(xstream_[q]);
  }
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
  P.range.assign(d, 0.0);
  for (int k = 0; k < d; ++k) {
    double mn = X[(size_t)k * n], mx = mn;
    for (int i = 1; i < n; ++i) {
      mn = std::min(mn, X[i + (size_t)k * n]);
      mx = std::max(mx, X[i + (size_t)k * n]);
    }
    P.range[k] = mx - mn;
  }
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

int Runtime::add_matrix_problem(const double* M, int n, bool device_memory) {
  if (n <= 0) throw std::invalid_argument("matrix must have at least one row");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  DeviceProblem P;
  P.n = n;
  P.d = 1;
  P.p = 0;
  P.range.assign(1, 0.0);
  // the likelihood stages that run after the factorisation read X, y: zeros
  std::vector<double> zeros(n, 0.0);
  hip_check(hipMalloc(&P.X, sizeof(double) * (size_t)n), "hipMalloc X");
  hip_check(hipMalloc(&P.y, sizeof(double) * (size_t)n), "hipMalloc y");
  hip_check(hipMalloc(&P.F, sizeof(double)), "hipMalloc F");
  hip_check(hipMalloc(&P.M, sizeof(double) * (size_t)n * n), "hipMalloc M");
  P.owns_M = true;
  hip_check(hipMemcpy(P.X, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice), "upload X");
  hip_check(hipMemcpy(P.y, zeros.data(), sizeof(double) * n, hipMemcpyHostToDevice), "upload y");
  hip_check(hipMemcpy(P.M, M, sizeof(double) * (size_t)n * n,
                      device_memory ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice),
            "upload M");
  const int id = next_problem_++;
  problems_[id] = P;
  return id;
}

namespace {
// one evaluation of a matrix problem in the given mode
EvalResult eval_matrix(Runtime& rt, int pid, int mode) {
  const double one = 1.0;
  EvalRequest R;
  R.problem = pid;
  R.gamma = &one;
  R.gamma_n = 1;
  R.mode = mode;
  std::vector<EvalRequest> reqs{R};
  std::vector<EvalResult> res;
  rt.evaluate(reqs, false, res);
  return res[0];
}

// problem ids of matrix problems, removed on scope exit
struct MatrixProblems {
  Runtime& rt;
  std::vector<int> ids;
  ~MatrixProblems() {
    for (int id : ids) {
      try {
        rt.remove_problem(id);
      } catch (...) {
      }
    }
  }
  int add(const double* M, int n, bool device_memory = false) {
    ids.push_back(rt.add_matrix_problem(M, n, device_memory));
    return ids.back();
  }
};
}  // namespace

void Runtime::safe_chol(const double* M, int n, double* L, int* inc) {
  MatrixProblems mp{*this, {}};
  const int pid = mp.add(M, n);
  const EvalResult r = eval_matrix(*this, pid, MODE_MATRIX);
  if (r.status != EVAL_OK) throw std::runtime_error(chol_failure_message(r.status, r.inc, chol.num_nugget));
  read_slot(0, 0, L);
  if (inc) *inc = r.inc;
}

double Runtime::rcond_chol(const double* L, int n) {
  MatrixProblems mp{*this, {}};
  const int pid = mp.add(L, n);
  const EvalResult r = eval_matrix(*this, pid, MODE_FACTOR);
  return r.status == EVAL_OK ? r.rcond2 : 0.0;
}

void Runtime::chol_block(const double* C, int n, const double* Loo, int no, double* L) {
  if (no <= 0 || no > n) throw std::invalid_argument("chol_block: the old block must have 1..n rows");
  for (int i = 0; i < no; ++i)
    if (Loo[i + (size_t)i * no] < 0.0)
      throw std::invalid_argument("chol_block: the old factor must have a non-negative diagonal");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  MatrixProblems mp{*this, {}};
  // Loo^-1 from the factor mode (the slot's W, lower, ld = npad)
  const int pid = mp.add(Loo, no);
  const EvalResult r = eval_matrix(*this, pid, MODE_FACTOR);
  if (r.status != EVAL_OK) throw std::runtime_error("chol_block: the old factor has a zero diagonal entry");
  const SlotView v = slot_view(0);
  double *dC = nullptr, *dL = nullptr;
  struct Free {
    double** p[2];
    ~Free() {
      for (double** q : p)
        if (*q) hipFree(*q);
    }
  } fr{{&dC, &dL}};
  const size_t bn = sizeof(double) * (size_t)n * n;
  hip_check(hipMalloc(&dC, bn), "hipMalloc");
  hip_check(hipMalloc(&dL, bn), "hipMalloc");
  hip_check(hipMemcpy(dC, C, bn, hipMemcpyHostToDevice), "upload C");
  chol_block_device(dC, n, problem(pid).M, no, v.W, v.npad, no, dL);
  hip_check(hipMemcpy(L, dL, bn, hipMemcpyDeviceToHost), "read L");
}

bool Runtime::chol_block_device(const double* dC, int n, const double* dLoo, int ldo, const double* dLooInv, int ldi,
                                int no, double* dL) {
  if (no <= 0 || no > n) throw std::invalid_argument("chol_block: the old block must have 1..n rows");
  if (ldi % NB != 0 || ldi < no) throw std::invalid_argument("chol_block: bad leading dimension of Loo^-1");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  const int nu = n - no, mpad = (nu + NB - 1) / NB * NB;
  BlockAssembleArgs ba{};
  ba.Loo = dLoo;
  ba.L = dL;
  ba.n = n;
  ba.no = no;
  ba.ldo = ldo;
  if (nu == 0) {
    ba.U = dLoo;
    ba.Luu = dLoo;
    ba.ldu = ba.ldl = 1;
    hip_check(launch_block_assemble(ba, stream_), "assemble L");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return true;
  }
  double *dB = nullptr, *dU = nullptr, *dCuu = nullptr, *dS = nullptr;
  struct Free {
    double** p[4];
    ~Free() {
      for (double** q : p)
        if (*q) hipFree(*q);
    }
  } fr{{&dB, &dU, &dCuu, &dS}};
  const size_t bu = sizeof(double) * (size_t)ldi * mpad, bs = sizeof(double) * (size_t)nu * nu;
  hip_check(hipMalloc(&dB, bu), "hipMalloc");
  hip_check(hipMalloc(&dU, bu), "hipMalloc");
  hip_check(hipMalloc(&dCuu, bs), "hipMalloc");
  hip_check(hipMalloc(&dS, bs), "hipMalloc");
  // C_ou (no x nu) into a zero-padded ldi x mpad block, C_uu (nu x nu)
  hip_check(hipMemsetAsync(dB, 0, bu, stream_), "clear");
  hip_check(hipMemcpy2DAsync(dB, sizeof(double) * ldi, dC + (size_t)no * n, sizeof(double) * n, sizeof(double) * no,
                             nu, hipMemcpyDeviceToDevice, stream_),
            "copy C_ou");
  hip_check(hipMemcpy2DAsync(dCuu, sizeof(double) * nu, dC + no + (size_t)no * n, sizeof(double) * n,
                             sizeof(double) * nu, nu, hipMemcpyDeviceToDevice, stream_),
            "copy C_uu");
  // U = Loo^-1 C_ou = L_ou' (LinearAlgebra.cpp:271), S = C_uu - U' U
  PredictArgs t{};
  t.Linv = dLooInv;
  t.Ron = dB;
  t.C = dU;
  t.npad = ldi;
  t.mpad = mpad;
  hip_check(launch_trmm_lower(t, stream_), "trmm");
  SchurArgs sa{};
  sa.U = dU;
  sa.Cuu = dCuu;
  sa.S = dS;
  sa.kdim = ldi;
  sa.m = nu;
  sa.mpad = mpad;
  hip_check(launch_schur(sa, stream_), "Schur complement");
  hip_check(hipStreamSynchronize(stream_), "sync");
  // L_uu = safe_chol(S); when its ladder fails, safe_chol(C) from scratch
  // (LinearAlgebra.cpp:286-296)
  MatrixProblems mp{*this, {}};
  const int ps = mp.add(dS, nu, true);
  const EvalResult rs = eval_matrix(*this, ps, MODE_MATRIX);
  if (rs.status == EVAL_OK) {
    const SlotView v = slot_view(0);
    ba.U = dU;
    ba.ldu = ldi;
    ba.Luu = v.A;
    ba.ldl = v.npad;
    hip_check(launch_block_assemble(ba, stream_), "assemble L");
    hip_check(hipStreamSynchronize(stream_), "sync");
    return true;
  }
  const int pc = mp.add(dC, n, true);
  const EvalResult rc = eval_matrix(*this, pc, MODE_MATRIX);
  if (rc.status != EVAL_OK) throw std::runtime_error(chol_failure_message(rc.status, rc.inc, chol.num_nugget));
  const SlotView v = slot_view(0);
  BlockAssembleArgs fa{};
  fa.L = dL;
  fa.n = n;
  fa.no = 0;
  fa.Loo = v.A;
  fa.U = v.A;
  fa.Luu = v.A;
  fa.ldo = fa.ldu = fa.ldl = v.npad;
  hip_check(launch_block_assemble(fa, stream_), "assemble L");
  hip_check(hipStreamSynchronize(stream_), "sync");
  return false;
}

void Runtime::set_problem_factor(int id, const double* L) {
  auto it = problems_.find(id);
  if (it == problems_.end()) throw std::invalid_argument("unknown problem id");
  if (it->second.owns_M) throw std::invalid_argument("a matrix problem's matrix is its own");
  it->second.M = const_cast<double*>(L);
  // the slot table carries the pointer: lay it out again
  slot_problem_.clear();
}

void Runtime::remove_problem(int id) {
  auto it = problems_.find(id);
  if (it == problems_.end()) throw std::invalid_argument("unknown problem id");
  hipSetDevice(device_);
  hipStreamSynchronize(stream_);
  hipFree(it->second.X);
  hipFree(it->second.y);
  hipFree(it->second.F);
  if (it->second.owns_M && it->second.M) hipFree(it->second.M);
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

  // slots in bucket order (stable): slot_of_req_[r] is request r's slot
  std::vector<int> bucket(ns), order(ns);
  for (int r = 0; r < ns; ++r) bucket[r] = bucket_of((problem(mapping[r]).n + NB - 1) / NB);
  for (int r = 0; r < ns; ++r) order[r] = r;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return bucket[a] < bucket[b]; });
  std::vector<int> slot_of_req(ns);
  for (int s = 0; s < ns; ++s) slot_of_req[order[s]] = s;
  std::vector<int> slot_prob(ns);
  for (int s = 0; s < ns; ++s) slot_prob[s] = mapping[order[s]];

  std::vector<size_t> off(ns);
  size_t total = 0;
  int nblk_max = 0, dmax = 0;
  h_slots_.assign(ns, Slot{});
  for (int s = 0; s < ns; ++s) {
    const DeviceProblem& P = problem(slot_prob[s]);
    const int npad = (P.n + NB - 1) / NB * NB;
    const int nblk = npad / NB;
    off[s] = total;
    total += align256(sizeof(double) * (size_t)npad * npad) * 2;       // A, W
    total += align256(sizeof(double) * (size_t)npad * v_cols(P.p));     // V
    total += align256(sizeof(double) * red_size(nblk));                 // red
    nblk_max = std::max(nblk_max, nblk);
    dmax = std::max(dmax, P.d);
  }
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamSynchronize(stream_), "sync");
  char* base = nullptr;
  if (noalloc_) {
    // slot tables first, then the workspaces, all in the scratch array
    const size_t t1 = align256(sizeof(Slot) * ns), t2 = align256(sizeof(SlotParams) * ns),
                 t3 = align256(sizeof(SlotOut) * ns);
    if (ns > NOALLOC_MAX_ITEMS)
      throw std::invalid_argument("the allocation-free entry point takes at most " +
                                  std::to_string(NOALLOC_MAX_ITEMS) + " items, got " + std::to_string(ns));
    if (t1 + t2 + t3 + total > NOALLOC_SCRATCH_BYTES)
      throw std::invalid_argument("this batch needs " + std::to_string(t1 + t2 + t3 + total) +
                                  " bytes of workspace; the allocation-free entry point has " +
                                  std::to_string(NOALLOC_SCRATCH_BYTES) + " (use lkgpu_ll_grad_batch)");
    d_slots_ = reinterpret_cast<Slot*>(scratch_);
    d_params_ = reinterpret_cast<SlotParams*>(scratch_ + t1);
    d_out_ = reinterpret_cast<SlotOut*>(scratch_ + t1 + t2);
    h_params_ = np_params_;
    h_out_ = np_out_;
    base = scratch_ + t1 + t2 + t3;
  } else if (total > pool_bytes_) {
    if (pool_) hip_check(hipFree(pool_), "hipFree pool");
    pool_ = nullptr;
    hip_check(hipMalloc(&pool_, total), "hipMalloc slot pool");
    pool_bytes_ = total;
  }
  if (!noalloc_) base = reinterpret_cast<char*>(pool_);
  for (int s = 0; s < ns; ++s) {
    const DeviceProblem& P = problem(slot_prob[s]);
    const int npad = (P.n + NB - 1) / NB * NB;
    const int nblk = npad / NB;
    char* q = base + off[s];
    Slot& S = h_slots_[s];
    S.X = P.X;
    S.y = P.y;
    S.F = P.F;
    S.M = P.M;
    S.A = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * npad);
    S.W = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * npad);
    S.V = reinterpret_cast<double*>(q);
    q += align256(sizeof(double) * (size_t)npad * v_cols(P.p));
    S.red = reinterpret_cast<double*>(q);
    S.n = P.n;
    S.npad = npad;
    S.nblk = nblk;
    S.d = P.d;
    S.p = P.p;
    S.kernel = P.kernel;
    S.noise = P.noise;
  }
  if (!noalloc_ && (size_t)ns > slot_cap_) {
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
  // contiguous slot range of each bucket present; more than kMaxStreams
  // buckets: the smallest adjacent ones merged
  std::vector<Part> parts;
  for (int s = 0; s < ns; ++s) {
    const int bk = bucket[order[s]];
    if (parts.empty() || bucket[order[parts.back().first]] != bk) parts.push_back(Part{s, 0, 0, 0, 4});
    Part& q = parts.back();
    ++q.count;
    q.nblk_max = std::max(q.nblk_max, h_slots_[s].nblk);
    q.dmax = std::max(q.dmax, h_slots_[s].d);
  }
  while ((int)parts.size() > kMaxStreams) {
    parts[0].count += parts[1].count;
    parts[0].nblk_max = std::max(parts[0].nblk_max, parts[1].nblk_max);
    parts[0].dmax = std::max(parts[0].dmax, parts[1].dmax);
    parts.erase(parts.begin() + 1);
  }
  for (Part& q : parts) apply_tuning(q);
  buckets_ = parts;
  slot_of_req_ = slot_of_req;
  slot_problem_ = mapping;
  nslots_ = ns;
  nblk_max_ = nblk_max;
  dmax_ = dmax;
  ++layout_version_;
}

// The pipeline as an ordered list of launches, each tagged with the stages
// it belongs to (bit s = stage s): 0 build, 1 Cholesky, 2 first trailing
// update, 3 trtri, 4 triangular solves + LL, 5 fused R^-1 / adjoint.
void Runtime::apply_tuning(Part& q) const {
  const BucketTuning& t = tuning_[bucket_of(q.nblk_max)];
  q.panel_blocks = t.panel_blocks;
  q.wide = t.wide;
}

std::vector<Runtime::Step> Runtime::pipeline_steps(const Batch& B, bool want_grad, int objective,
                                                   int panel_blocks, int wide, bool lookahead) const {
  std::vector<Step> v;
  v.push_back({1u << 0, "build", [B] { return launch_build(B); }});
  // blocked right-looking Cholesky, two levels: panels of P block columns;
  // inside a panel each step updates only the panel's remaining columns,
  // then one trailing update with K = 64 P per panel
  const int P = panel_blocks;
  bool first_update = true;
  if (lookahead && B.nslots == 1 && B.nblk_max >= kLookaheadMinBlocks) {
    // look-ahead: after panel p, its update of panel p + 1's columns runs on
    // the main stream, then panel p + 1's chain (diagonal blocks, panels,
    // in-panel updates) beside panel p's update of all later columns on the
    // second stream.  Every tile still receives the panels' updates in panel
    // order: the main stream waits for the bulk update of panel p - 1 before
    // it applies panel p's update to panel p + 1.
    Batch BL = B;
    BL.stream = la_stream_;
    const hipStream_t sa = B.stream, sl = la_stream_;
    const hipEvent_t evA = la_ev_[0], evB = la_ev_[1];
    const int wg = la_workgroups_;
    bool pending = false;
    for (int K0 = 0; K0 < B.nblk_max; K0 += P) {
      const int K1 = std::min(K0 + P, B.nblk_max);
      for (int k = K0; k < K1; ++k) {
        v.push_back({1u << 1, "potrf_diag", [B, k] { return launch_potrf_diag(B, k); }});
        v.push_back({1u << 1, "potrf_panel", [B, k] { return launch_potrf_panel(B, k); }});
        if (k + 1 < K1)
          v.push_back({1u << 1, "potrf_update (panel)", [B, k, K1] { return launch_potrf_update(B, k, k + 1, k + 1, K1 - 1); }});
      }
      if (K1 >= B.nblk_max) break;
      const int K2 = std::min(K1 + P, B.nblk_max);
      if (pending) v.push_back({1u << 1, "wait bulk update", [sa, evB] { return hipStreamWaitEvent(sa, evB, 0); }});
      v.push_back({(1u << 1) | (first_update ? 1u << 2 : 0u), "potrf_update (next panel)",
                   [B, K0, K1, K2] { return launch_potrf_update(B, K0, K1, K1, K2 - 1); }});
      first_update = false;
      if (K2 < B.nblk_max) {
        v.push_back({1u << 1, "record chain", [sa, evA] { return hipEventRecord(evA, sa); }});
        v.push_back({1u << 1, "fork bulk update", [sl, evA] { return hipStreamWaitEvent(sl, evA, 0); }});
        if (wg > 0)
          v.push_back({1u << 1, "potrf_update (bulk, persistent)",
                       [BL, K0, K1, K2, wg] { return launch_potrf_update_persistent(BL, K0, K1, K2, -1, wg); }});
        else
          v.push_back({1u << 1, "potrf_update (bulk)",
                       [BL, K0, K1, K2] { return launch_potrf_update(BL, K0, K1, K2, -1); }});
        v.push_back({1u << 1, "record bulk update", [sl, evB] { return hipEventRecord(evB, sl); }});
        pending = true;
      }
    }
    if (pending) v.push_back({1u << 1, "join bulk update", [sa, evB] { return hipStreamWaitEvent(sa, evB, 0); }});
  } else
  for (int K0 = 0; K0 < B.nblk_max; K0 += P) {
    const int K1 = std::min(K0 + P, B.nblk_max);
    for (int k = K0; k < K1; ++k) {
      v.push_back({1u << 1, "potrf_diag", [B, k] { return launch_potrf_diag(B, k); }});
      v.push_back({1u << 1, "potrf_panel", [B, k] { return launch_potrf_panel(B, k); }});
      if (k + 1 < K1) {
        v.push_back({(1u << 1) | (first_update ? 1u << 2 : 0u), "potrf_update (panel)",
                     [B, k, K1] { return launch_potrf_update(B, k, k + 1, k + 1, K1 - 1); }});
        first_update = false;
      }
    }
    if (K1 < B.nblk_max) {
      v.push_back({(1u << 1) | (first_update ? 1u << 2 : 0u), "potrf_update (trailing)",
                   [B, K0, K1, wide] {
                     return wide ? launch_potrf_update_trailing(B, K0, K1, K1) : launch_potrf_update(B, K0, K1, K1, -1);
                   }});
      first_update = false;
    }
  }
  const int lv = trtri_levels(B.nblk_max);
  for (int l = 0; l < lv; ++l) {
    v.push_back({1u << 3, "trtri step 0", [B, l] { return launch_trtri_level(B, 1 << l, 0); }});
    v.push_back({1u << 3, "trtri step 1", [B, l] { return launch_trtri_level(B, 1 << l, 1); }});
  }
  v.push_back({1u << 4, "solve_fy", [B] { return launch_solve_fy(B); }});
  v.push_back({1u << 4, "gls", [B] { return launch_gls(B); }});
  v.push_back({1u << 4, "solve_e", [B] { return launch_solve_e(B); }});
  v.push_back({1u << 4, "solve_x", [B] { return launch_solve_x(B); }});
  v.push_back({1u << 4, "finalize_ll", [B] { return launch_finalize_ll(B); }});
  v.push_back({1u << 4, "rcond_estimate", [B] { return launch_rcond_estimate(B); }});
  if (objective == OBJ_LOO) {
    // leave-one-out value (diag B from the columns of L^-1), then its
    // gradient: B w, Bs = B diag(sqrt a) into A, Bs Bs' contracted with dR
    v.push_back({1u << 4, "loo_g", [B] { return launch_loo_g(B); }});
    v.push_back({1u << 4, "loo_value", [B] { return launch_loo_value(B); }});
    if (want_grad) {
      v.push_back({1u << 5, "loo_wv", [B] { return launch_loo_wv(B); }});
      v.push_back({1u << 5, "loo_bw", [B] { return launch_loo_bw(B); }});
      v.push_back({1u << 5, "loo_bs", [B] { return launch_loo_bs(B); }});
      v.push_back({1u << 5, "loo_adjoint", [B] { return launch_loo_adjoint(B); }});
      v.push_back({0u, "loo_finalize_grad", [B] { return launch_loo_finalize_grad(B); }});
    }
    return v;
  }
  if (objective == OBJ_LMP) {
    // log marginal posterior: the value comes out of finalize_ll; its
    // gradient is the likelihood's contraction with B = R^-1 - G'G
    if (want_grad) {
      v.push_back({1u << 5, "loo_g", [B] { return launch_loo_g(B); }});
      v.push_back({1u << 5, "lmp_adjoint", [B] { return launch_lmp_adjoint(B); }});
      v.push_back({0u, "finalize_grad", [B] { return launch_finalize_grad(B); }});
    }
    return v;
  }
  if (want_grad) {
    v.push_back({1u << 5, "lauum_adjoint", [B] { return launch_lauum_adjoint(B); }});
    v.push_back({0u, "finalize_grad", [B] { return launch_finalize_grad(B); }});
  }
  return v;
}

static void run_steps(const std::vector<Runtime::Step>& v, size_t a, size_t b) {
  for (size_t i = a; i < b; ++i) hip_check(v[i].fn(), v[i].name);
}

// [first, last) range of the steps of `stage` (empty when it has none)
static std::pair<size_t, size_t> stage_range(const std::vector<Runtime::Step>& v, int stage) {
  size_t a = v.size(), b = v.size();
  if (stage < 0) return {a, b};
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i].stages & (1u << stage)) {
      if (a == v.size()) a = i;
      b = i + 1;
    }
  if (a == v.size()) b = a;
  return {a, b};
}

hipGraph_t Runtime::capture(const std::vector<Step>& v, size_t a, size_t b) {
  hipGraph_t g = nullptr;
  hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    run_steps(v, a, b);
  } catch (...) {
    hipStreamEndCapture(stream_, &g);
    if (g) hipGraphDestroy(g);
    throw;
  }
  hip_check(hipStreamEndCapture(stream_, &g), "end capture");
  return g;
}

void Runtime::set_timed_stage(int stage) {
  if (stage < -1 || stage > 5) throw std::invalid_argument("unknown stage id");
  hip_check(hipSetDevice(device_), "hipSetDevice");
  for (auto& e : tev_)
    if (e == nullptr) hip_check(hipEventCreate(&e), "hipEventCreate");
  timed_stage_ = stage;
  timed_sum_ms_ = 0.0;
  timed_count_ = 0;
}

void Runtime::take_stage_time(double* avg_ms, int* count) {
  *avg_ms = timed_count_ > 0 ? timed_sum_ms_ / timed_count_ : 0.0;
  *count = timed_count_;
  timed_sum_ms_ = 0.0;
  timed_count_ = 0;
}

// Parts of one run.  Timed runs (stage events on the main stream) and bench
// runs: one part.  A batch spanning several size buckets: one part per
// bucket.  A batch of one bucket: the largest S <= min(streams_, the
// bucket's tuned streams) equal parts whose slot counts are multiples of 8
// (each part keeps the XCD-aware placement).
std::vector<Runtime::Part> Runtime::plan_parts(bool timed, BenchMap* bench) const {
  Part all{0, nslots_, nblk_max_, dmax_};
  apply_tuning(all);
  if (timed || bench != nullptr || noalloc_) return {all};
  if (buckets_.size() > 1 && streams_ > 1) {
    std::vector<Part> v = buckets_;
    while ((int)v.size() > streams_) {
      v[0].count += v[1].count;
      v[0].nblk_max = std::max(v[0].nblk_max, v[1].nblk_max);
      v[0].dmax = std::max(v[0].dmax, v[1].dmax);
      apply_tuning(v[0]);
      v.erase(v.begin() + 1);
    }
    return v;
  }
  const int smax = std::min(streams_, tuning_[bucket_of(nblk_max_)].streams);
  for (int S = smax; S > 1; S /= 2)
    if (nslots_ % (8 * S) == 0) {
      std::vector<Part> v(S, all);
      for (int q = 0; q < S; ++q) {
        v[q].first = q * (nslots_ / S);
        v[q].count = nslots_ / S;
      }
      return v;
    }
  return {all};
}

Batch Runtime::part_batch(const Part& q, hipStream_t s) const {
  Batch B;
  B.slots = d_slots_ + q.first;
  B.params = d_params_ + q.first;
  B.out = d_out_ + q.first;
  B.nslots = q.count;
  B.nblk_max = q.nblk_max;
  B.dmax = q.dmax;
  B.stream = s;
  return B;
}

// launches of the S parts, interleaved, between a fork from and a join back
// to the main stream (also valid inside a stream capture, which the other
// streams then join)
void Runtime::issue_split(const std::vector<Step>* parts, int S) {
  hip_check(hipEventRecord(fork_, stream_), "fork record");
  for (int q = 1; q < S; ++q) hip_check(hipStreamWaitEvent(xstream_[q - 1], fork_, 0), "fork wait");
  size_t len = 0;
  for (int q = 0; q < S; ++q) len = std::max(len, parts[q].size());
  for (size_t i = 0; i < len; ++i)
    for (int q = 0; q < S; ++q)
      if (i < parts[q].size()) hip_check(parts[q][i].fn(), parts[q][i].name);
  for (int q = 1; q < S; ++q) {
    hip_check(hipEventRecord(join_[q - 1], xstream_[q - 1]), "join record");
    hip_check(hipStreamWaitEvent(stream_, join_[q - 1], 0), "join wait");
  }
}

void Runtime::run_pipeline(bool want_grad, BenchMap* bench, int objective) {
  hip_check(hipMemcpyAsync(d_params_, h_params_, sizeof(SlotParams) * nslots_, hipMemcpyHostToDevice, stream_),
            "upload params");
  Part all{0, nslots_, nblk_max_, dmax_};
  apply_tuning(all);
  const bool la = lookahead_active();
  const std::vector<Step> steps = pipeline_steps(part_batch(all, stream_), want_grad, objective, all.panel_blocks,
                                                 all.wide, la);
  const auto rng = stage_range(steps, timed_stage_);
  const bool timed = rng.first < rng.second;
  const std::vector<Part> plan = plan_parts(timed, bench);
  const int S = (int)plan.size();
  const bool split = S > 1;
  last_parts_ = S;
  std::vector<Step> part[kMaxStreams];
  if (split) {
    for (int q = 0; q < S; ++q)
      part[q] = pipeline_steps(part_batch(plan[q], q == 0 ? stream_ : xstream_[q - 1]), want_grad, objective,
                               plan[q].panel_blocks, plan[q].wide);
  }
  if (bench != nullptr) {
    // uncaptured run with one event pair around the whole pipeline
    hipEvent_t e0, e1;
    hip_check(hipEventCreate(&e0), "event");
    hip_check(hipEventCreate(&e1), "event");
    hip_check(hipEventRecord(e0, stream_), "event record");
    run_steps(steps, 0, steps.size());
    hip_check(hipEventRecord(e1, stream_), "event record");
    hip_check(hipEventSynchronize(e1), "event sync");
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, e0, e1), "event elapsed");
    (*bench)["gpu pipeline"] += ms;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  } else if (use_graph && !noalloc_ && !la) {
    if (objective < 0 || objective >= OBJ_COUNT) throw std::invalid_argument("unknown objective");
    GraphEntry& G = graphs_[2 * objective + (want_grad ? 1 : 0)];
    if (G.version != layout_version_ || G.exec == nullptr || G.timed_stage != timed_stage_ || G.split != S) {
      if (G.exec) hip_check(hipGraphExecDestroy(G.exec), "graph exec destroy");
      G.exec = nullptr;
      hipGraph_t graph = nullptr;
      if (split) {
        hip_check(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal), "begin capture");
        try {
          issue_split(part, S);
        } catch (...) {
          hipStreamEndCapture(stream_, &graph);
          if (graph) hipGraphDestroy(graph);
          throw;
        }
        hip_check(hipStreamEndCapture(stream_, &graph), "end capture");
      } else if (!timed) {
        graph = capture(steps, 0, steps.size());
      } else {
        // events captured from a stream are not timeable here, so the timed
        // stage sits between explicit event-record nodes of a parent graph
        // whose other nodes are the captured segments before and after it
        hip_check(hipGraphCreate(&graph, 0), "graph create");
        hipGraphNode_t prev = nullptr, node = nullptr;
        auto add_child = [&](size_t a, size_t b) {
          if (a >= b) return;
          hipGraph_t child = capture(steps, a, b);
          hip_check(hipGraphAddChildGraphNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, child),
                    "add child graph");
          hip_check(hipGraphDestroy(child), "graph destroy");
          prev = node;
        };
        auto add_event = [&](hipEvent_t e) {
          hip_check(hipGraphAddEventRecordNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, e),
                    "add event node");
          prev = node;
        };
        add_child(0, rng.first);
        add_event(tev_[0]);
        add_child(rng.first, rng.second);
        add_event(tev_[1]);
        add_child(rng.second, steps.size());
      }
      hip_check(hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0), "graph instantiate");
      hip_check(hipGraphDestroy(graph), "graph destroy");
      G.version = layout_version_;
      G.timed_stage = timed_stage_;
      G.split = S;
    }
    hip_check(hipGraphLaunch(G.exec, stream_), "graph launch");
  } else if (split) {
    issue_split(part, S);
  } else {
    run_steps(steps, 0, rng.first);
    if (timed) hip_check(hipEventRecord(tev_[0], stream_), "timing event");
    run_steps(steps, rng.first, rng.second);
    if (timed) hip_check(hipEventRecord(tev_[1], stream_), "timing event");
    run_steps(steps, rng.second, steps.size());
  }
  hip_check(hipMemcpyAsync(h_out_, d_out_, sizeof(SlotOut) * nslots_, hipMemcpyDeviceToHost, stream_),
            "download results");
  hip_check(hipStreamSynchronize(stream_), "stream sync");
  if (timed && bench == nullptr) {
    float ms = 0.f;
    hip_check(hipEventElapsedTime(&ms, tev_[0], tev_[1]), "stage timing");
    timed_sum_ms_ += ms;
    ++timed_count_;
  }
}

void Runtime::evaluate(const std::vector<EvalRequest>& reqs, bool want_grad, std::vector<EvalResult>& res,
                       BenchMap* bench, int objective) {
  if (objective < 0 || objective >= OBJ_COUNT) throw std::invalid_argument("unknown objective");
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
    SlotParams& Q = h_params_[slot_of_req_[s]];
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
    Q.inc = R.flags.inc0;
    Q.active = R.active ? 1 : 0;
    if (!R.active) res[s].ll = std::numeric_limits<double>::quiet_NaN();
    Q.est_sigma2 = R.flags.est_sigma2;
    Q.est_nugget = R.flags.est_nugget;
    Q.rcond_check = R.flags.rcond_check >= 0 ? R.flags.rcond_check : chol.rcond_check;
    Q.want_grad = want_grad;
    Q.gamma_has_alpha = R.gamma_n > P.d;
    Q.objective = objective;
    Q.mode = R.mode;
    if (R.mode != MODE_KERNEL && P.M == nullptr) throw std::invalid_argument("matrix mode needs a matrix problem");
  }
  for (;;) {
    run_pipeline(want_grad, bench, objective);
    bool retry = false;
    for (int s = 0; s < ns; ++s) {
      SlotParams& Q = h_params_[slot_of_req_[s]];
      if (!Q.active) continue;
      const SlotOut& O = h_out_[slot_of_req_[s]];
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
      if (Q.mode == MODE_FACTOR) {  // a given factor with a zero (or NaN) diagonal
        E.status = EVAL_CHOL_FAILED;
        E.inc = 0;
        E.rcond2 = 0.0;
        Q.active = 0;
        continue;
      }
      // failed factorization or rcond below threshold: next step of the
      // ladder (LinearAlgebra.cpp:78-90)
      const int max_inc = reqs[s].flags.max_inc >= 0 ? reqs[s].flags.max_inc : chol.max_inc;
      if (Q.inc > max_inc) {
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
  last_objective_ = objective;
  if (objective == OBJ_LMP) add_lmp_prior(reqs, res, want_grad);
  // leave every slot active for the next call's graph (params are rewritten)
}

void Runtime::evaluate_noalloc(const std::vector<EvalRequest>& reqs, bool want_grad,
                               std::vector<EvalResult>& res) {
  if (scratch_ == nullptr) throw std::runtime_error("no scratch array on this device");
  // the allocating path's workspace, kept aside and restored
  Slot* s_slots = d_slots_;
  SlotParams* s_dp = d_params_;
  SlotOut* s_do = d_out_;
  SlotParams* s_hp = h_params_;
  SlotOut* s_ho = h_out_;
  auto restore = [&] {
    d_slots_ = s_slots;
    d_params_ = s_dp;
    d_out_ = s_do;
    h_params_ = s_hp;
    h_out_ = s_ho;
    noalloc_ = false;
    // the next allocating call lays its slots out again (in its own pool)
    slot_problem_.clear();
    nslots_ = 0;
  };
  noalloc_ = true;
  slot_problem_.clear();
  nslots_ = 0;
  try {
    evaluate(reqs, want_grad, res, nullptr, OBJ_LL);
  } catch (...) {
    restore();
    throw;
  }
  restore();
}

// The approximate reference prior of the LMP objective (Kriging.cpp:585-598,
// 600-606, 618): a = 0.2, b = (a + d) / n^(1/d), CL = range(X) / n^(1/d),
// t = sum CL / theta (+ (1 - alpha) / alpha for the nugget model), prior
// -b t + a log t.  With sigma2 fixed the reference's gradient is a forward
// difference (step 1e-6 in theta) of the whole objective, reproduced here by
// d more batched evaluations; the alpha component is then 0.
void Runtime::add_lmp_prior(const std::vector<EvalRequest>& reqs, std::vector<EvalResult>& res,
                            bool want_grad) const {
  const int ns = (int)reqs.size();
  for (int s = 0; s < ns; ++s) {
    EvalResult& E = res[s];
    if (!reqs[s].active || E.status != EVAL_OK) continue;
    const DeviceProblem& P = problem(reqs[s].problem);
    const int d = P.d;
    const double n = P.n;
    const double a = 0.2, b = 1.0 / std::pow(n, 1.0 / d) * (a + d);
    const bool nug = P.noise == NOISE_NUGGET;
    const double alpha = h_params_[slot_of_req_[s]].alpha;
    double t = nug ? (1.0 - alpha) / alpha : 0.0;
    for (int k = 0; k < d; ++k) t += P.range[k] / std::pow(n, 1.0 / d) / reqs[s].gamma[k];
    E.ll += -b * t + a * std::log(t);
    if (!want_grad) continue;
    for (int k = 0; k < d; ++k) {
      const double cl = P.range[k] / std::pow(n, 1.0 / d), th = reqs[s].gamma[k];
      E.grad[k] -= (a * cl / t - b * cl) / (th * th);
    }
    if (nug && reqs[s].gamma_n > d) E.grad[d] -= (a / t - b) / (alpha * alpha);
  }
  if (!want_grad) return;
  // fixed sigma2: forward differences of the objective (Kriging.cpp:634-646)
  std::vector<int> fd;
  for (int s = 0; s < ns; ++s)
    if (reqs[s].active && res[s].status == EVAL_OK && !reqs[s].flags.est_sigma2) fd.push_back(s);
  if (fd.empty()) return;
  Runtime* self = const_cast<Runtime*>(this);
  int dmax = 0;
  for (int s : fd) dmax = std::max(dmax, problem(reqs[s].problem).d);
  std::vector<std::vector<double>> gam(ns);
  std::vector<EvalRequest> r2(reqs);
  std::vector<EvalResult> res2;
  const double eps = 1e-6;
  for (int k = 0; k < dmax; ++k) {
    for (int s = 0; s < ns; ++s) {
      r2[s].active = false;
      if (std::find(fd.begin(), fd.end(), s) == fd.end() || k >= problem(reqs[s].problem).d) continue;
      gam[s].assign(reqs[s].gamma, reqs[s].gamma + reqs[s].gamma_n);
      gam[s][k] += eps;
      r2[s].gamma = gam[s].data();
      r2[s].active = true;
    }
    self->evaluate(r2, false, res2, nullptr, OBJ_LMP);
    for (int s : fd)
      if (k < problem(reqs[s].problem).d) res[s].grad[k] = (res2[s].ll - res[s].ll) / eps;
  }
  for (int s : fd) {
    const DeviceProblem& P = problem(reqs[s].problem);
    if (P.noise == NOISE_NUGGET && reqs[s].gamma_n > P.d) res[s].grad[P.d] = 0.0;
  }
}

Runtime::SlotView Runtime::slot_view(int req) const {
  if (req < 0 || req >= nslots_) throw std::out_of_range("slot index");
  const Slot& S = h_slots_[slot_of(req)];
  return SlotView{S.A, S.W, S.V, S.n, S.npad, S.p};
}

void Runtime::read_rinv(int s, double* out) {
  if (s < 0 || s >= nslots_) throw std::invalid_argument("slot out of range");
  // s is a request index; slot_view translates it
  hip_check(hipSetDevice(device_), "hipSetDevice");
  Batch B;
  B.slots = d_slots_;
  B.params = d_params_;
  B.out = d_out_;
  B.nslots = nslots_;
  B.nblk_max = nblk_max_;
  B.dmax = dmax_;
  B.stream = stream_;
  hip_check(launch_rinv_store(B), "R^-1");
  const SlotView v = slot_view(s);
  hip_check(hipStreamSynchronize(stream_), "sync");
  hip_check(hipMemcpy2D(out, sizeof(double) * v.n, v.A, sizeof(double) * v.npad, sizeof(double) * v.n, v.n,
                        hipMemcpyDeviceToHost),
            "read R^-1");
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
    case 7:  // leave-one-out errors
    case 8:  // sigma2LOO
      if (last_objective_ != OBJ_LOO) throw std::runtime_error("the last evaluation was not a leave-one-out one");
      copy_cols(v.V + (size_t)(what == 7 ? v_loo_err(v.p) : v_loo_s2(v.p)) * ld, 1);
      break;
    case 6: {  // R* (p x p, upper), stored by the GLS kernel
      const Slot& S = h_slots_[slot_of(s)];
      if (v.p > 0)
        hip_check(hipMemcpy(out, S.red + red_rstar_off(S.nblk), sizeof(double) * v.p * v.p, hipMemcpyDeviceToHost),
                  "read R*");
      break;
    }
    default:
      throw std::invalid_argument("unknown buffer id");
  }
}

}  // namespace lk
