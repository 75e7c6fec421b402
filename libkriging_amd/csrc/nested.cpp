#include "nested.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <thread>
#include <limits>
#include <random>
#include <sstream>
#include <stdexcept>

#include "lk_kernels.h"

namespace lk {

namespace {
constexpr double NK_JITTER = 1e-10;  // NestedKriging::jitter (NestedKriging.hpp:165)

template <class T>
void dfree(T*& p) {
  if (p != nullptr) hipFree((void*)p);
  p = nullptr;
}

// pseudo-inverse of a symmetric p x p matrix (column-major) by cyclic Jacobi
// eigen-decomposition, singular values below max(p) * max|lambda| * eps
// dropped (arma::pinv's default tolerance); returns pinv(A) b
std::vector<double> sym_pinv_solve(std::vector<double> A, const std::vector<double>& b, int p) {
  std::vector<double> V((size_t)p * p, 0.0);
  for (int i = 0; i < p; ++i) V[(size_t)i * p + i] = 1.0;
  auto at = [&](int i, int j) -> double& { return A[(size_t)j * p + i]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int j = 0; j < p; ++j)
      for (int i = 0; i < j; ++i) off += at(i, j) * at(i, j);
    if (off < 1e-300) break;
    for (int q = 1; q < p; ++q)
      for (int r = 0; r < q; ++r) {
        const double apq = at(r, q);
        if (apq == 0.0) continue;
        const double theta = (at(q, q) - at(r, r)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < p; ++k) {  // columns r, q
          const double akr = at(k, r), akq = at(k, q);
          at(k, r) = c * akr - s * akq;
          at(k, q) = s * akr + c * akq;
        }
        for (int k = 0; k < p; ++k) {  // rows r, q
          const double ark = at(r, k), aqk = at(q, k);
          at(r, k) = c * ark - s * aqk;
          at(q, k) = s * ark + c * aqk;
        }
        for (int k = 0; k < p; ++k) {
          const double vkr = V[(size_t)r * p + k], vkq = V[(size_t)q * p + k];
          V[(size_t)r * p + k] = c * vkr - s * vkq;
          V[(size_t)q * p + k] = s * vkr + c * vkq;
        }
      }
  }
  double smax = 0.0;
  for (int i = 0; i < p; ++i) smax = std::max(smax, std::fabs(at(i, i)));
  const double tol = p * smax * std::numeric_limits<double>::epsilon();
  std::vector<double> x(p, 0.0);
  for (int e = 0; e < p; ++e) {
    const double lam = at(e, e);
    if (std::fabs(lam) <= tol) continue;
    double vb = 0.0;
    for (int k = 0; k < p; ++k) vb += V[(size_t)e * p + k] * b[k];
    vb /= lam;
    for (int k = 0; k < p; ++k) x[k] += V[(size_t)e * p + k] * vb;
  }
  return x;
}
}  // namespace

NestedKriging::Aggregation NestedKriging::aggregation_from_string(const std::string& s) {
  if (s == "PoE") return Aggregation::PoE;
  if (s == "gPoE") return Aggregation::gPoE;
  if (s == "BCM") return Aggregation::BCM;
  if (s == "rBCM") return Aggregation::rBCM;
  if (s == "NK") return Aggregation::NK;
  throw std::invalid_argument("Unknown aggregation: " + s + " (expected PoE, gPoE, BCM, rBCM or NK)");
}

std::string NestedKriging::aggregation_to_string(Aggregation a) {
  switch (a) {
    case Aggregation::PoE: return "PoE";
    case Aggregation::gPoE: return "gPoE";
    case Aggregation::BCM: return "BCM";
    case Aggregation::rBCM: return "rBCM";
    case Aggregation::NK: return "NK";
  }
  return "?";
}

NestedKriging::Partition NestedKriging::partition_from_string(const std::string& s) {
  if (s == "random") return Partition::Random;
  if (s == "kmeans" || s.empty()) return Partition::KMeans;
  throw std::invalid_argument("Unknown partition: '" + s + "'. Expected 'kmeans' or 'random'.");
}

NestedKriging::NestedKriging(const std::string& kernel) : kernel_(kernel) {
  kernel_id(kernel);  // Covariance::resolve validates the name
  device_ = default_device();
}

NestedKriging::~NestedKriging() {
  release_nk();
  for (auto& e : pev_)
    if (e) hipEventDestroy(e);
}

void NestedKriging::set_pair_timing(bool on) {
  if (on)
    for (auto& e : pev_)
      if (e == nullptr) hip_check(hipEventCreate(&e), "hipEventCreate");
  time_pairs_ = on;
  pair_ms_ = 0.0;
  pair_launches_ = pair_points_ = 0;
}

void NestedKriging::take_pair_time(double* ms, int* launches, int* points) {
  if (ms) *ms = pair_ms_;
  if (launches) *launches = pair_launches_;
  if (points) *points = pair_points_;
  pair_ms_ = 0.0;
  pair_launches_ = pair_points_ = 0;
}

void NestedKriging::release_nk() {
  if (rt_ != nullptr) hipSetDevice(device_);
  dfree(d_Xg_);
  dfree(d_Linv_);
  dfree(d_alpha_);
  dfree(d_groups_);
  dfree(d_tile_group_);
  dfree(d_pairs_);
  npairs_ = 0;
  Npad_ = 0;
}

const Kriging& NestedKriging::submodel(int i) const {
  if (i < 0 || i >= (int)sub_.size()) throw std::out_of_range("submodel index");
  return *sub_[i];
}
Kriging& NestedKriging::submodel(int i) {
  if (i < 0 || i >= (int)sub_.size()) throw std::out_of_range("submodel index");
  return *sub_[i];
}

// argument checks of NestedKriging::fit (NestedKriging.cpp:177-187)
void NestedKriging::check_args(int n, int d, int nb_groups, RegModel regmodel) const {
  if (n <= 0 || d <= 0) throw std::invalid_argument("y and X should have the same number of rows");
  if (nb_groups < 1 || nb_groups > n / (d + 2)) throw std::invalid_argument("nb_groups should be in [1, n/(d+2)]");
  if (aggregation_ == Aggregation::NK && regmodel != RegModel::Constant)
    throw std::invalid_argument("NK aggregation requires a Constant trend; use PoE/gPoE/BCM/rBCM otherwise");
}

// make_partition (NestedKriging.cpp:124-166): k-means (random-subset seeds,
// 10 Lloyd iterations) with a fallback to a balanced random split when a
// cluster has fewer than d + 2 points; the random split assigns perm(i) to
// group i mod nb_groups.  The generator is std::mt19937(seed), so partitions
// are reproducible for a seed but are not Armadillo's.
void NestedKriging::make_partition(int nb_groups) {
  const int n = n_, d = d_;
  std::mt19937 eng((unsigned)seed_);
  std::vector<int> assign(n, 0);
  bool kmeans_ok = false;
  if (partition_ == Partition::KMeans && nb_groups > 1) {
    // seeds: a random subset of nb_groups distinct points
    std::vector<int> idx(n);
    for (int i = 0; i < n; ++i) idx[i] = i;
    for (int i = 0; i < nb_groups; ++i) {
      std::uniform_int_distribution<int> u(i, n - 1);
      std::swap(idx[i], idx[u(eng)]);
    }
    std::vector<double> cen((size_t)nb_groups * d);
    for (int g = 0; g < nb_groups; ++g)
      for (int k = 0; k < d; ++k) cen[(size_t)g * d + k] = X_[idx[g] + (size_t)k * n];
    auto nearest = [&](int i) {
      int best = 0;
      double bd = std::numeric_limits<double>::infinity();
      for (int g = 0; g < nb_groups; ++g) {
        double s = 0.0;
        for (int k = 0; k < d; ++k) {
          const double v = X_[i + (size_t)k * n] - cen[(size_t)g * d + k];
          s += v * v;
        }
        if (s < bd) {
          bd = s;
          best = g;
        }
      }
      return best;
    };
    kmeans_ok = true;
    for (int iter = 0; iter < 10; ++iter) {
      for (int i = 0; i < n; ++i) assign[i] = nearest(i);
      std::vector<double> sum((size_t)nb_groups * d, 0.0);
      std::vector<int> cnt(nb_groups, 0);
      for (int i = 0; i < n; ++i) {
        cnt[assign[i]]++;
        for (int k = 0; k < d; ++k) sum[(size_t)assign[i] * d + k] += X_[i + (size_t)k * n];
      }
      for (int g = 0; g < nb_groups; ++g) {
        if (cnt[g] == 0) {
          kmeans_ok = false;
          break;
        }
        for (int k = 0; k < d; ++k) cen[(size_t)g * d + k] = sum[(size_t)g * d + k] / cnt[g];
      }
      if (!kmeans_ok) break;
    }
    if (kmeans_ok) {
      std::vector<int> cnt(nb_groups, 0);
      for (int i = 0; i < n; ++i) cnt[assign[i] = nearest(i)]++;
      for (int g = 0; g < nb_groups; ++g)
        if (cnt[g] < d + 2) kmeans_ok = false;
    }
  }
  if (!kmeans_ok) {
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = n - 1; i > 0; --i) {
      std::uniform_int_distribution<int> u(0, i);
      std::swap(perm[i], perm[u(eng)]);
    }
    for (int i = 0; i < n; ++i) assign[perm[i]] = i % nb_groups;
  }
  groups_.assign(nb_groups, {});
  for (int i = 0; i < n; ++i) groups_[assign[i]].push_back(i);
}

void NestedKriging::fit(const double* y, const double* X, int n, int d, int nb_groups, RegModel regmodel,
                        const std::string& optim, const std::string& objective, const Parameters& parameters) {
  check_args(n, d, nb_groups, regmodel);
  n_ = n;
  d_ = d;
  X_.assign(X, X + (size_t)n * d);
  y_.assign(y, y + n);
  make_partition(nb_groups);
  fit_impl(regmodel, optim, objective, parameters);
}

void NestedKriging::fit_groups(const double* y, const double* X, int n, int d, const int* group_of, int nb_groups,
                               RegModel regmodel, const std::string& optim, const std::string& objective,
                               const Parameters& parameters) {
  if (n <= 0 || d <= 0) throw std::invalid_argument("y and X should have the same number of rows");
  if (nb_groups < 1 || nb_groups > n) throw std::invalid_argument("nb_groups should be in [1, n]");
  if (aggregation_ == Aggregation::NK && regmodel != RegModel::Constant)
    throw std::invalid_argument("NK aggregation requires a Constant trend; use PoE/gPoE/BCM/rBCM otherwise");
  n_ = n;
  d_ = d;
  X_.assign(X, X + (size_t)n * d);
  y_.assign(y, y + n);
  groups_.assign(nb_groups, {});
  for (int i = 0; i < n; ++i) {
    if (group_of[i] < 0 || group_of[i] >= nb_groups) throw std::invalid_argument("group index out of range");
    groups_[group_of[i]].push_back(i);
  }
  for (int g = 0; g < nb_groups; ++g)
    if (groups_[g].empty()) throw std::invalid_argument("group " + std::to_string(g) + " is empty");
  fit_impl(regmodel, optim, objective, parameters);
}

// NestedKriging::fit, plain path (NestedKriging.cpp:257-275): one Kriging fit
// per group, all groups' optimiser steps sharing batched pipelines
void NestedKriging::fit_impl(RegModel regmodel, const std::string& optim, const std::string& objective,
                             const Parameters& parameters) {
  fitted_ = false;
  release_nk();
  regmodel_ = regmodel;
  const int p = (int)groups_.size(), n = n_, d = d_;
  sub_.clear();
  std::vector<std::unique_ptr<FitJob>> jobs;
  std::vector<FitJob*> jp;
  std::vector<double> Xg, yg;
  auto gather = [&](int g, std::vector<double>& Xo, std::vector<double>& yo) {
    const std::vector<int>& idx = groups_[g];
    const int ng = (int)idx.size();
    Xo.resize((size_t)ng * d);
    yo.resize(ng);
    for (int a = 0; a < ng; ++a) {
      yo[a] = y_[idx[a]];
      for (int k = 0; k < d; ++k) Xo[a + (size_t)k * ng] = X_[idx[a] + (size_t)k * n];
    }
  };
  // the groups' theta bounds (each an O(n_g^2 d) host pair loop, in the
  // reference's order) computed for all groups at once on the host cores
  std::vector<std::vector<double>> blo(p), bup(p);
  if (optim != "none") {
    const int nt = std::max(1, std::min<int>(p, (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()))));
    std::atomic<int> next{0};
    auto work = [&] {
      std::vector<double> Xl, yl;
      for (int g = next++; g < p; g = next++) {
        gather(g, Xl, yl);
        Kriging::theta_bounds_of(Xl.data(), yl.data(), (int)groups_[g].size(), d, blo[g], bup[g]);
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
  }
  for (int g = 0; g < p; ++g) {
    const int ng = (int)groups_[g].size();
    gather(g, Xg, yg);
    sub_.push_back(std::make_unique<Kriging>(kernel_));
    if (optim != "none") sub_.back()->use_theta_bounds(blo[g], bup[g]);
    jobs.push_back(sub_.back()->begin_fit(yg.data(), Xg.data(), ng, d, regmodel, false, optim, objective, parameters));
    jp.push_back(jobs.back().get());
  }
  Kriging::run_fit_jobs(jp);
  fit_evals_ = 0;
  fit_batches_ = 0;
  for (const auto& k : sub_) {
    fit_evals_ += k->fit_evaluations();
    fit_batches_ = std::max(fit_batches_, k->fit_batches());
  }
  rt_ = sub_[0]->runtime();
  unify(objective);
  if (aggregation_ == Aggregation::NK) precompute_nk();
  fitted_ = true;
}

// unify_hyperparameters, plain path (NestedKriging.cpp:308-338)
void NestedKriging::unify(const std::string& objective) {
  const int p = (int)groups_.size(), d = d_, n = n_;
  std::vector<double> logt(d, 0.0);
  double s2 = 0.0, b0 = 0.0;
  gtheta_.assign((size_t)p * d, 0.0);
  gsigma2_.assign(p, 0.0);
  gbeta0_.assign(p, 0.0);
  for (int g = 0; g < p; ++g) {
    const double w = (double)groups_[g].size() / (double)n;
    const std::vector<double>& th = sub_[g]->theta();
    for (int k = 0; k < d; ++k) gtheta_[(size_t)g * d + k] = th[k];
    gsigma2_[g] = sub_[g]->sigma2();
    if (regmodel_ == RegModel::Constant) gbeta0_[g] = sub_[g]->beta()[0];
    for (int k = 0; k < d; ++k) logt[k] += w * std::log(th[k]);
    s2 += w * sub_[g]->sigma2();
    if (regmodel_ == RegModel::Constant) b0 += w * sub_[g]->beta()[0];
  }
  theta_.resize(d);
  for (int k = 0; k < d; ++k) theta_[k] = std::exp(logt[k]);
  sigma2_ = s2;
  beta0_ = b0;

  Parameters fixed;
  fixed.theta = theta_;
  fixed.theta_rows = 1;
  fixed.is_theta_estim = false;
  fixed.sigma2 = sigma2_;
  fixed.is_sigma2_estim = false;
  if (regmodel_ == RegModel::Constant) {
    fixed.beta = std::vector<double>{beta0_};
    fixed.is_beta_estim = false;
  }
  std::vector<std::unique_ptr<FitJob>> jobs;
  std::vector<FitJob*> jp;
  std::vector<double> Xg, yg;
  for (int g = 0; g < p; ++g) {
    const std::vector<int>& idx = groups_[g];
    const int ng = (int)idx.size();
    Xg.resize((size_t)ng * d);
    yg.resize(ng);
    for (int a = 0; a < ng; ++a) {
      yg[a] = y_[idx[a]];
      for (int k = 0; k < d; ++k) Xg[a + (size_t)k * ng] = X_[idx[a] + (size_t)k * n];
    }
    jobs.push_back(sub_[g]->begin_fit(yg.data(), Xg.data(), ng, d, regmodel_, false, "none", objective, fixed));
    jp.push_back(jobs.back().get());
  }
  Kriging::run_fit_jobs(jp);
}

// precompute_nk (NestedKriging.cpp:340-351): L_g = chol(R_g + 1e-10 I) and
// alpha_g = R_g^-1 (y_g - beta0), from one batched pipeline over the groups
// (fixed jitter, no retry ladder, no rcond check, as arma::chol)
void NestedKriging::precompute_nk() {
  const int p = (int)groups_.size(), d = d_, n = n_;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  goff_.assign(p, 0);
  gpad_.assign(p, 0);
  int off = 0;
  size_t lsize = 0;
  std::vector<NkGroup> G(p);
  for (int g = 0; g < p; ++g) {
    const int ng = (int)groups_[g].size();
    gpad_[g] = (ng + NB - 1) / NB * NB;
    goff_[g] = off;
    G[g].off = off;
    G[g].n = ng;
    G[g].npad = gpad_[g];
    G[g].pad_ = 0;
    G[g].loff = (long long)lsize;
    off += gpad_[g];
    lsize += (size_t)gpad_[g] * gpad_[g];
  }
  Npad_ = off;
  std::vector<double> Xg((size_t)Npad_ * d, 0.0);
  for (int g = 0; g < p; ++g)
    for (int a = 0; a < (int)groups_[g].size(); ++a)
      for (int k = 0; k < d; ++k) Xg[goff_[g] + a + (size_t)k * Npad_] = X_[groups_[g][a] + (size_t)k * n];
  std::vector<int> tg(Npad_ / NB);
  for (int g = 0; g < p; ++g)
    for (int t = 0; t < gpad_[g] / NB; ++t) tg[goff_[g] / NB + t] = g;
  // group pairs i < j, largest work first (load balance of the pair kernel)
  std::vector<int2> pairs;
  pairs.reserve((size_t)p * (p - 1) / 2);
  for (int i = 0; i < p; ++i)
    for (int j = i + 1; j < p; ++j) pairs.push_back(make_int2(i, j));
  std::stable_sort(pairs.begin(), pairs.end(), [&](const int2& a, const int2& b) {
    return (long long)gpad_[a.x] * gpad_[a.y] > (long long)gpad_[b.x] * gpad_[b.y];
  });
  npairs_ = (int)pairs.size();

  hip_check(hipMalloc(&d_Xg_, sizeof(double) * Xg.size()), "hipMalloc");
  hip_check(hipMalloc(&d_Linv_, sizeof(double) * lsize), "hipMalloc");
  hip_check(hipMalloc(&d_alpha_, sizeof(double) * Npad_), "hipMalloc");
  hip_check(hipMalloc(&d_groups_, sizeof(NkGroup) * p), "hipMalloc");
  hip_check(hipMalloc(&d_tile_group_, sizeof(int) * tg.size()), "hipMalloc");
  if (npairs_ > 0) hip_check(hipMalloc(&d_pairs_, sizeof(int2) * pairs.size()), "hipMalloc");
  hip_check(hipMemcpy(d_Xg_, Xg.data(), sizeof(double) * Xg.size(), hipMemcpyHostToDevice), "upload");
  hip_check(hipMemcpy(d_groups_, G.data(), sizeof(NkGroup) * p, hipMemcpyHostToDevice), "upload");
  hip_check(hipMemcpy(d_tile_group_, tg.data(), sizeof(int) * tg.size(), hipMemcpyHostToDevice), "upload");
  if (npairs_ > 0)
    hip_check(hipMemcpy(d_pairs_, pairs.data(), sizeof(int2) * pairs.size(), hipMemcpyHostToDevice), "upload");

  // factorisations of all groups in one batch
  std::vector<EvalRequest> reqs(p);
  for (int g = 0; g < p; ++g) {
    reqs[g].problem = sub_[g]->problem_id();
    reqs[g].gamma = theta_.data();
    reqs[g].gamma_n = d;
    reqs[g].flags.inc0 = 1;
    reqs[g].flags.max_inc = 0;
    reqs[g].flags.rcond_check = 0;
  }
  std::vector<EvalResult> res;
  rt_->evaluate(reqs, false, res);
  std::vector<double> r((size_t)Npad_, 0.0), ys, Fs;
  for (int g = 0; g < p; ++g) {
    if (res[g].status != EVAL_OK) throw std::runtime_error("chol(): decomposition failed (NestedKriging group " +
                                                           std::to_string(g) + ")");
    const Runtime::SlotView v = rt_->slot_view(g);
    if (v.npad != gpad_[g]) throw std::logic_error("NestedKriging: slot padding mismatch");
    hip_check(hipMemcpyAsync(d_Linv_ + G[g].loff, v.W, sizeof(double) * (size_t)v.npad * v.npad,
                             hipMemcpyDeviceToDevice, rt_->stream()),
              "copy factor");
    const int ng = G[g].n;
    ys.resize(ng);
    Fs.resize(ng);
    rt_->read_slot(g, 3, ys.data());  // L^-1 y
    rt_->read_slot(g, 2, Fs.data());  // L^-1 1
    for (int a = 0; a < ng; ++a) r[goff_[g] + a] = ys[a] - beta0_ * Fs[a];
  }
  double* d_r = nullptr;
  hip_check(hipMalloc(&d_r, sizeof(double) * Npad_), "hipMalloc");
  try {
    hip_check(hipMemcpyAsync(d_r, r.data(), sizeof(double) * Npad_, hipMemcpyHostToDevice, rt_->stream()), "upload");
    NkArgs a{};
    a.Linv = d_Linv_;
    a.alpha = d_alpha_;
    a.groups = (const NkGroup*)d_groups_;
    a.tile_group = d_tile_group_;
    a.Npad = Npad_;
    hip_check(launch_nk_alpha(a, d_r, rt_->stream()), "nk alpha");
    hip_check(hipStreamSynchronize(rt_->stream()), "sync");
  } catch (...) {
    hipFree(d_r);
    throw;
  }
  hipFree(d_r);
}

void NestedKriging::predict(const double* Xn, int m, bool want_stdev, double* mean, double* stdev) {
  if (!fitted_) throw std::runtime_error("NestedKriging is not fitted");
  if (m <= 0) return;
  if (aggregation_ == Aggregation::NK)
    predict_nk(Xn, m, want_stdev, mean, stdev);
  else
    predict_poe(Xn, m, want_stdev, mean, stdev);
}

// predict_poe_family (NestedKriging.cpp:365-421): submodel means/variances
// from the GPU predict of each submodel, combined per point
void NestedKriging::predict_poe(const double* Xn, int m, bool want_stdev, double* mean, double* stdev) {
  const int p = (int)groups_.size();
  constexpr double tiny = 1e-12;
  const double prior_var = sigma2_;
  std::vector<double> mus((size_t)m * p), vars((size_t)m * p), mu(m), sd(m);
  for (int g = 0; g < p; ++g) {
    sub_[g]->predict(Xn, m, true, mu.data(), sd.data());
    for (int t = 0; t < m; ++t) {
      mus[(size_t)g * m + t] = mu[t];
      vars[(size_t)g * m + t] = sd[t] * sd[t] + tiny;
    }
  }
  for (int t = 0; t < m; ++t) {
    double prec = 0.0, num = 0.0, bsum = 0.0;
    switch (aggregation_) {
      case Aggregation::PoE:
        for (int g = 0; g < p; ++g) prec += 1.0 / vars[(size_t)g * m + t];
        for (int g = 0; g < p; ++g) num += mus[(size_t)g * m + t] / vars[(size_t)g * m + t];
        mean[t] = num / prec;
        break;
      case Aggregation::gPoE:
        for (int g = 0; g < p; ++g) prec += 1.0 / vars[(size_t)g * m + t];
        prec /= (double)p;
        for (int g = 0; g < p; ++g) num += mus[(size_t)g * m + t] / vars[(size_t)g * m + t];
        mean[t] = (num / (double)p) / prec;
        break;
      case Aggregation::BCM:
        for (int g = 0; g < p; ++g) prec += 1.0 / vars[(size_t)g * m + t];
        prec -= ((double)p - 1.0) / prior_var;
        prec = std::max(prec, tiny);
        for (int g = 0; g < p; ++g) num += mus[(size_t)g * m + t] / vars[(size_t)g * m + t];
        mean[t] = num / prec;
        break;
      case Aggregation::rBCM:
        for (int g = 0; g < p; ++g) {
          const double v = vars[(size_t)g * m + t];
          const double beta = 0.5 * (std::log(prior_var) - std::log(v));
          prec += beta / v;
          bsum += beta;
          num += beta * mus[(size_t)g * m + t] / v;
        }
        prec += (1.0 - bsum) / prior_var;
        prec = std::max(prec, tiny);
        mean[t] = num / prec;
        break;
      default:
        throw std::logic_error("unreachable");
    }
    if (want_stdev && stdev != nullptr) stdev[t] = std::sqrt(std::max(1.0 / prec, 0.0));
  }
}

// predict_nk (NestedKriging.cpp:353-465) on the device, chunk by chunk
void NestedKriging::predict_nk(const double* Xn, int m, bool want_stdev, double* mean, double* stdev) {
  const int p = (int)groups_.size(), d = d_;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  // chunk: explicit, else as many points as fit in ~8 GiB of workspace
  int qc = chunk_;
  if (qc <= 0) {
    const double per_point = 8.0 * (2.0 * Npad_ + 2.0 * (double)p * p + 2.0 * p + 4.0 + d);
    qc = (int)std::min<double>(8192.0, std::max(64.0, (8.0 * (1ull << 30)) / per_point));
    qc = std::max(64, qc / 64 * 64);
  }
  qc = std::min(qc, m);
  const int qpad_max = (qc + NK_QT - 1) / NK_QT * NK_QT;
  double *dXc = nullptr, *dC = nullptr, *dU = nullptr, *dM = nullptr, *dKd = nullptr, *dKM = nullptr, *dKF = nullptr,
         *dout = nullptr;
  int* dinfo = nullptr;
  auto cleanup = [&] {
    dfree(dXc);
    dfree(dC);
    dfree(dU);
    dfree(dM);
    dfree(dKd);
    dfree(dKM);
    dfree(dKF);
    dfree(dout);
    dfree(dinfo);
  };
  fallbacks_ = 0;
  try {
    const size_t pp = (size_t)p * p;
    hip_check(hipMalloc(&dXc, sizeof(double) * (size_t)qc * d), "hipMalloc");
    hip_check(hipMalloc(&dC, sizeof(double) * (size_t)Npad_ * qpad_max), "hipMalloc");
    hip_check(hipMalloc(&dU, sizeof(double) * (size_t)Npad_ * qpad_max), "hipMalloc");
    hip_check(hipMalloc(&dM, sizeof(double) * (size_t)qc * p), "hipMalloc");
    hip_check(hipMalloc(&dKd, sizeof(double) * (size_t)qc * p), "hipMalloc");
    hip_check(hipMalloc(&dKM, sizeof(double) * (size_t)qc * pp), "hipMalloc");
    hip_check(hipMalloc(&dKF, sizeof(double) * (size_t)qc * pp), "hipMalloc");
    hip_check(hipMalloc(&dout, sizeof(double) * 2 * (size_t)qc), "hipMalloc");
    hip_check(hipMalloc(&dinfo, sizeof(int) * (size_t)qc), "hipMalloc");
    std::vector<double> xc((size_t)qc * d), hout(2 * (size_t)qc);
    std::vector<int> hinfo(qc);
    hipStream_t st = rt_->stream();
    for (int s0 = 0; s0 < m; s0 += qc) {
      const int mc = std::min(qc, m - s0);
      for (int k = 0; k < d; ++k)
        for (int t = 0; t < mc; ++t) xc[t + (size_t)k * mc] = Xn[s0 + t + (size_t)k * m];
      hip_check(hipMemcpyAsync(dXc, xc.data(), sizeof(double) * (size_t)mc * d, hipMemcpyHostToDevice, st), "upload");
      NkArgs a{};
      a.Xg = d_Xg_;
      a.Linv = d_Linv_;
      a.alpha = d_alpha_;
      a.groups = (const NkGroup*)d_groups_;
      a.tile_group = d_tile_group_;
      a.pairs = (const int2*)d_pairs_;
      a.Xc = dXc;
      a.C = dC;
      a.U = dU;
      a.M = dM;
      a.Kd = dKd;
      a.KM = dKM;
      a.KF = dKF;
      a.out = dout;
      a.info = dinfo;
      for (int k = 0; k < d; ++k) a.theta[k] = theta_[k];
      for (int k = 0; k < DMAX; ++k) a.itheta[k] = (k < d) ? 1.0 / theta_[k] : 1.0;
      a.beta0 = beta0_;
      a.sigma2 = sigma2_;
      a.jitter = NK_JITTER;
      a.Npad = Npad_;
      a.d = d;
      a.p = p;
      a.qc = mc;
      a.qpad = (mc + NK_QT - 1) / NK_QT * NK_QT;
      a.kernel = kernel_id(kernel_);
      a.npairs = npairs_;
      hip_check(launch_nk_prepare(a, st), "nk prepare");
      if (time_pairs_) hip_check(hipEventRecord(pev_[0], st), "event record");
      hip_check(launch_nk_pairs(a, st), "nk pairs");
      if (time_pairs_) hip_check(hipEventRecord(pev_[1], st), "event record");
      hip_check(launch_nk_solve(a, st), "nk solve");
      hip_check(hipMemcpyAsync(hout.data(), dout, sizeof(double) * 2 * mc, hipMemcpyDeviceToHost, st), "download");
      hip_check(hipMemcpyAsync(hinfo.data(), dinfo, sizeof(int) * mc, hipMemcpyDeviceToHost, st), "download");
      hip_check(hipStreamSynchronize(st), "sync");
      if (time_pairs_) {
        float ms = 0.f;
        hip_check(hipEventElapsedTime(&ms, pev_[0], pev_[1]), "event elapsed");
        pair_ms_ += ms;
        pair_points_ += mc;
        ++pair_launches_;
      }
      for (int t = 0; t < mc; ++t) {
        double mu = hout[t], var = hout[mc + t];
        if (hinfo[t] != 0) ++fallbacks_;
        if (hinfo[t] == 2) {
          // singular KM: w = pinv(KM) kM (NestedKriging.cpp:457-458)
          std::vector<double> K(pp), kM(p), Mv(p);
          hip_check(hipMemcpy(K.data(), dKM + (size_t)t * pp, sizeof(double) * pp, hipMemcpyDeviceToHost), "download");
          hip_check(hipMemcpy(Mv.data(), dM + (size_t)t * p, sizeof(double) * p, hipMemcpyDeviceToHost), "download");
          for (int i = 0; i < p; ++i) {
            for (int j = 0; j < i; ++j) K[(size_t)j * p + i] = K[(size_t)i * p + j];  // symmatu
            kM[i] = K[(size_t)i * p + i];
          }
          const std::vector<double> w = sym_pinv_solve(K, kM, p);
          double sm = 0.0, sv = 0.0;
          for (int i = 0; i < p; ++i) {
            sm += w[i] * (Mv[i] - beta0_);
            sv += w[i] * kM[i];
          }
          mu = beta0_ + sm;
          var = sigma2_ * std::max(0.0, 1.0 - sv);
        }
        mean[s0 + t] = mu;
        if (want_stdev && stdev != nullptr) stdev[s0 + t] = std::sqrt(var);
      }
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

std::string NestedKriging::summary() const {
  std::ostringstream oss;
  if (!fitted_) {
    oss << "* covariance: " << kernel_ << " (not fitted)";
    return oss.str();
  }
  oss << "* data: " << n_ << " x " << d_ << " -> " << n_ << "\n";
  oss << "* groups: " << groups_.size() << " (";
  for (size_t g = 0; g < groups_.size(); ++g) oss << (g ? "," : "") << groups_[g].size();
  oss << ")\n";
  oss << "* aggregation: " << aggregation_to_string(aggregation_) << "\n";
  oss << "* trend (constant): " << beta0_ << "\n";
  oss << "* variance: " << sigma2_ << "\n";
  oss << "* covariance: " << kernel_ << ", range:";
  for (double t : theta_) oss << " " << t;
  return oss.str();
}

}  // namespace lk
