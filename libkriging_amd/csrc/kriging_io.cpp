// Kriging model files (Kriging::save / Kriging::load, src/lib/Kriging.cpp:
// 2691-2750; the common fields of KrigingImpl::dump_common_to_json,
// src/lib/KrigingImpl.cpp:644-711).  save writes every field the reference's
// loader reads, the factors taken from one device evaluation at the committed
// parameters; load restores data, options and parameters and recomputes the
// factors on the device.
#include <algorithm>
#include <cmath>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>

#include "kriging.hpp"
#include "model_io.hpp"

namespace lk {

const char* regmodel_name(RegModel r) {
  switch (r) {
    case RegModel::None:
      return "none";
    case RegModel::Constant:
      return "constant";
    default:
      return "linear";
  }
}

namespace {

// correlation of the covariance family (Covariance.cpp:24-161) at dx
double corr(const std::string& kernel, const double* dx, const std::vector<double>& theta) {
  const int d = (int)theta.size();
  double s = 0.0, pr = 1.0;
  if (kernel == "gauss") {
    for (int k = 0; k < d; ++k) s += (dx[k] / theta[k]) * (dx[k] / theta[k]);
    return std::exp(-0.5 * s);
  }
  if (kernel == "exp") {
    for (int k = 0; k < d; ++k) s += std::fabs(dx[k] / theta[k]);
    return std::exp(-s);
  }
  const bool m52 = kernel == "matern5_2";
  const double c = m52 ? std::sqrt(5.0) : std::sqrt(3.0);
  for (int k = 0; k < d; ++k) {
    const double u = c * std::fabs(dx[k] / theta[k]);
    s += u;
    pr *= m52 ? 1.0 + u + u * u / 3.0 : 1.0 + u;
  }
  return std::exp(-s) * pr;
}

}  // namespace

// T = L (n x n), M = F* (n x p) and z = E* (or y* - M beta when beta is
// fixed, Kriging.cpp:2170-2172) of the committed model, as the commit (or a
// block update) left them on the device (the reference's m_T, m_M, m_z);
// R^-1 when asked from one device evaluation that takes that L as given
void Kriging::factors(std::vector<double>* T, std::vector<double>* M, std::vector<double>* z,
                      std::vector<double>* Rinv) const {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  if (d_L_ == nullptr || d_vec_ == nullptr) throw std::runtime_error("model has no committed factor");
  const int n = n_, p = p_;
  hip_check(hipSetDevice(device_), "hipSetDevice");
  hip_check(hipStreamSynchronize(rt_->stream()), "sync");
  if (T) {
    T->assign((size_t)n * n, 0.0);
    hip_check(hipMemcpy(T->data(), d_L_, sizeof(double) * (size_t)n * n, hipMemcpyDeviceToHost), "read T");
    // the device copy's strict upper triangle is workspace: T is lower
    for (int j = 1; j < n; ++j)
      for (int i = 0; i < j; ++i) (*T)[i + (size_t)j * n] = 0.0;
  }
  if (M) {
    M->assign((size_t)n * p, 0.0);
    if (p > 0)
      hip_check(hipMemcpy2D(M->data(), sizeof(double) * n, d_vec_, sizeof(double) * npad_, sizeof(double) * n, p,
                            hipMemcpyDeviceToHost),
                "read M");
  }
  if (z) {
    z->assign(n, 0.0);
    hip_check(hipMemcpy(z->data(), d_vec_ + (size_t)npad_ * p, sizeof(double) * n, hipMemcpyDeviceToHost), "read z");
  }
  if (Rinv) {
    std::vector<double> g(theta_);
    if (noise_ == NoiseModel::Nugget) g.push_back(alpha_);
    std::vector<EvalRequest> reqs{commit_request(g)};
    reqs[0].flags.alpha = alpha_;
    reqs[0].mode = MODE_FACTOR;
    std::vector<EvalResult> res;
    rt_->set_problem_factor(problem_, d_L_);
    try {
      rt_->evaluate(reqs, false, res);
    } catch (...) {
      rt_->set_problem_factor(problem_, nullptr);
      throw;
    }
    rt_->set_problem_factor(problem_, nullptr);
    if (res[0].status != EVAL_OK) throw std::runtime_error("the committed factor has a zero diagonal entry");
    Rinv->assign((size_t)n * n, 0.0);
    rt_->read_rinv(0, Rinv->data());
  }
}

void Kriging::save(const std::string& filename) {
  const std::string text = to_json();
  std::ofstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open '" + filename + "' for writing");
  f << text;
  if (!f) throw std::runtime_error("cannot write '" + filename + "'");
}

std::string Kriging::to_json() const {
  if (is_empty_) throw std::runtime_error("model is not fitted");
  const int n = n_, d = d_, p = p_;
  const bool nug = noise_ == NoiseModel::Nugget;
  // factors at the committed parameters: L, F*, E* / y*, then R^-1
  std::vector<double> T, M, z, Rinv;
  factors(&T, &M, &z, &Rinv);
  // R (un-jittered: factor * k off the diagonal, 1 on it), dX and maxdX
  // (LinearAlgebra::compute_dX: column i n + j and j n + i hold x_i - x_j)
  const double factor = nug ? alpha_ : 1.0;
  std::vector<double> R((size_t)n * n), dX((size_t)d * n * n, 0.0), maxdX(d, 0.0), dx(d);
  for (int i = 0; i < n; ++i) {
    R[i + (size_t)i * n] = 1.0;
    for (int j = i + 1; j < n; ++j) {
      for (int k = 0; k < d; ++k) {
        dx[k] = X_[i + (size_t)k * n] - X_[j + (size_t)k * n];
        dX[k + ((size_t)i * n + j) * d] = dx[k];
        dX[k + ((size_t)j * n + i) * d] = dx[k];
        maxdX[k] = std::max(maxdX[k], std::fabs(dx[k]));
      }
      const double r = factor * corr(kernel_, dx.data(), theta_);
      R[i + (size_t)j * n] = r;
      R[j + (size_t)i * n] = r;
    }
  }
  std::ostringstream o;
  auto field = [&](const char* key, const std::string& value, bool last = false) {
    o << "    " << io::quote(key) << ": " << value << (last ? "\n" : ",\n");
  };
  auto boolean = [](bool b) { return std::string(b ? "true" : "false"); };
  o << "{\n";
  field("version", "2");
  field("content", io::quote("Kriging"));
  field("covType", io::quote(kernel_));
  field("X", io::mat_node(X_.data(), n, d));
  field("centerX", io::vec_node("rowvec", centerX_.data(), d));
  field("scaleX", io::vec_node("rowvec", scaleX_.data(), d));
  field("y", io::vec_node("colvec", y_.data(), n));
  field("centerY", io::number(centerY_));
  field("scaleY", io::number(scaleY_));
  field("normalize", boolean(normalize_));
  field("regmodel", io::quote(regmodel_name(regmodel_)));
  field("optim", io::quote(optim_));
  field("objective", io::quote(objective_));
  field("dX", io::mat_node(dX.data(), d, (size_t)n * n));
  field("maxdX", io::vec_node("colvec", maxdX.data(), d));
  field("F", io::mat_node(F_.data(), n, p));
  field("T", io::mat_node(T.data(), n, n));
  field("R", io::mat_node(R.data(), n, n));
  field("M", io::mat_node(M.data(), n, p));
  field("star", io::mat_node(nullptr, 0, 0));
  field("circ", io::mat_node(Rstar_.data(), p, p));
  field("z", io::vec_node("colvec", z.data(), n));
  field("Rinv", io::mat_node(Rinv.data(), n, n));
  field("beta", io::vec_node("colvec", beta_.data(), p));
  field("est_beta", boolean(est_beta_));
  field("theta", io::vec_node("colvec", theta_.data(), d));
  field("est_theta", boolean(est_theta_));
  field("sigma2", io::number(sigma2_));
  field("est_sigma2", boolean(est_sigma2_));
  field("noise_model", io::quote(nug ? "Nugget" : "None"));
  if (nug) {
    field("nugget", io::number(nugget_));
    field("est_nugget", boolean(est_nugget_));
    field("alpha", io::number(alpha_));
  }
  field("nystrom_light", "false", true);
  o << "}\n";
  return o.str();
}

Kriging* Kriging::load(const std::string& filename) {
  std::ifstream f(filename, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open '" + filename + "'");
  std::stringstream buf;
  buf << f.rdbuf();
  return from_json(buf.str(), filename);
}

// Kriging(const Kriging&, ExplicitCopySpecifier) (Kriging.cpp:164): the same
// state through the model file's fields, the factors recomputed on the device
Kriging* Kriging::copy() const { return from_json(to_json(), "(copy)"); }

Kriging* Kriging::from_json(const std::string& text, const std::string& filename) {
  const io::Json j = io::parse(text);
  const int version = (int)j.at("version").as_number();
  if (version != 2)
    throw std::runtime_error("Bad version to load from '" + filename + "'; found " + std::to_string(version) +
                             ", requires 2");
  const std::string content = j.at("content").as_string();
  if (content != "Kriging")
    throw std::runtime_error("Bad content to load from '" + filename + "'; found '" + content +
                             "', requires 'Kriging'");
  const io::Json* nmj = j.find("noise_model");
  const std::string nm = nmj ? nmj->as_string() : "None";
  if (nm == "Heterogeneous") throw std::runtime_error("heterogeneous noise models are not supported on this path");
  const io::Json* nyj = j.find("nystrom_light");
  if (nyj != nullptr && nyj->as_bool()) throw std::runtime_error("Nystrom models are not supported on this path");
  const NoiseModel noise = nm == "Nugget" ? NoiseModel::Nugget : NoiseModel::None;
  std::unique_ptr<Kriging> k(new Kriging(j.at("covType").as_string(), noise));

  size_t n = 0, d = 0, r = 0, c = 0;
  k->X_ = io::array_from(j.at("X"), &n, &d);
  k->y_ = io::array_from(j.at("y"), &r, &c);
  if (r * c != n) throw std::runtime_error("model file: X and y sizes differ");
  k->centerX_ = io::array_from(j.at("centerX"), &r, &c);
  k->scaleX_ = io::array_from(j.at("scaleX"), &r, &c);
  if (k->centerX_.size() != d || k->scaleX_.size() != d) throw std::runtime_error("model file: bad centerX / scaleX");
  k->centerY_ = j.at("centerY").as_number();
  k->scaleY_ = j.at("scaleY").as_number();
  k->normalize_ = j.at("normalize").as_bool();
  k->regmodel_ = regmodel_from_string(j.at("regmodel").as_string());
  k->optim_ = j.at("optim").as_string();
  k->objective_ = j.at("objective").as_string();
  k->n_ = (int)n;
  k->d_ = (int)d;
  k->upload_data();

  k->theta_ = io::array_from(j.at("theta"), &r, &c);
  if (k->theta_.size() != d) throw std::runtime_error("model file: theta must have d entries");
  k->est_theta_ = j.at("est_theta").as_bool();
  k->beta_ = io::array_from(j.at("beta"), &r, &c);
  if ((int)k->beta_.size() != k->p_) throw std::runtime_error("model file: beta must have p entries");
  k->est_beta_ = j.at("est_beta").as_bool();
  k->sigma2_ = j.at("sigma2").as_number();
  k->est_sigma2_ = j.at("est_sigma2").as_bool();
  k->alpha_ = 1.0;
  if (noise == NoiseModel::Nugget) {
    k->nugget_ = j.at("nugget").as_number();
    k->est_nugget_ = j.at("est_nugget").as_bool();
    k->alpha_ = j.at("alpha").as_number();
  }
  // factors at the saved parameters; the saved scalars are kept as saved
  const std::vector<double> beta = k->beta_;
  const double sigma2 = k->sigma2_, nugget = k->nugget_;
  k->commit(k->theta_, k->alpha_);
  k->beta_ = beta;
  k->sigma2_ = sigma2;
  k->nugget_ = nugget;
  return k.release();
}

}  // namespace lk
