// extern "C" boundary of liblkgpu.so (declarations in include/lkgpu.h).
#include <exception>
#include <string>
#include <vector>

#include "../../include/lkgpu.h"
#include "lk_runtime.hpp"

struct lkgpu_context {
  lk::Runtime rt;
  explicit lkgpu_context(int dev) : rt(dev) {}
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return -1;
}
}  // namespace

namespace lk {
void set_last_error(const std::string& s) { g_last_error = s; }
}  // namespace lk

#ifndef LKGPU_BUILD_ID
#define LKGPU_BUILD_ID "unknown"
#endif

extern "C" {

const char* lk_get_last_error(void) { return g_last_error.c_str(); }

const char* lkgpu_build_id(void) { return LKGPU_BUILD_ID; }

int lkgpu_context_create(int device, lkgpu_context** out) {
  return guarded([&] {
    if (out == nullptr) throw std::invalid_argument("out is NULL");
    *out = new lkgpu_context(device);
  });
}

void lkgpu_context_destroy(lkgpu_context* ctx) { delete ctx; }

int lkgpu_set_chol_config(lkgpu_context* ctx, double num_nugget, int rcond_check, int max_inc, double min_rcond) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.chol.num_nugget = num_nugget;
    ctx->rt.chol.rcond_check = rcond_check != 0;
    ctx->rt.chol.max_inc = max_inc;
    ctx->rt.chol.min_rcond = min_rcond;
  });
}

int lkgpu_set_graph(lkgpu_context* ctx, int enable) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.use_graph = enable != 0;
  });
}

int lkgpu_problem_create(lkgpu_context* ctx, const double* X, int n, int d, const double* y, const double* F, int p,
                         int kernel, int noise, int* id_out) {
  return guarded([&] {
    if (ctx == nullptr || X == nullptr || y == nullptr || id_out == nullptr)
      throw std::invalid_argument("NULL argument");
    if (p > 0 && F == nullptr) throw std::invalid_argument("F is NULL with p > 0");
    *id_out = ctx->rt.add_problem(X, n, d, y, F, p, kernel, noise);
  });
}

int lkgpu_problem_destroy(lkgpu_context* ctx, int id) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.remove_problem(id);
  });
}

namespace {
int ll_grad_batch(lkgpu_context* ctx, int nb, const int* problem_ids, const double* gamma, int gamma_n,
                  const lkgpu_model_flags* flags, int flags_n, int return_grad, double* ll_out, double* grad_out,
                  int* status_out, int* inc_out, bool noalloc) {
  return guarded([&] {
    if (ctx == nullptr || problem_ids == nullptr || gamma == nullptr || ll_out == nullptr)
      throw std::invalid_argument("NULL argument");
    if (nb < 0) throw std::invalid_argument("nb < 0");
    if (return_grad && grad_out == nullptr) throw std::invalid_argument("grad_out is NULL");
    if (flags_n != 0 && flags_n != 1 && flags_n != nb) throw std::invalid_argument("flags_n must be 0, 1 or nb");
    std::vector<lk::EvalRequest> reqs(nb);
    for (int i = 0; i < nb; ++i) {
      lk::EvalRequest& R = reqs[i];
      R.problem = problem_ids[i];
      R.gamma = gamma + (size_t)i * gamma_n;
      R.gamma_n = gamma_n;
      if (flags != nullptr && flags_n > 0) {
        const lkgpu_model_flags& f = flags[flags_n == 1 ? 0 : i];
        R.flags.est_sigma2 = f.est_sigma2 != 0;
        R.flags.est_nugget = f.est_nugget != 0;
        R.flags.sigma2 = f.sigma2;
        R.flags.nugget = f.nugget;
        R.flags.alpha = f.alpha;
      }
    }
    std::vector<lk::EvalResult> res;
    if (noalloc)
      ctx->rt.evaluate_noalloc(reqs, return_grad != 0, res);
    else
      ctx->rt.evaluate(reqs, return_grad != 0, res);
    for (int i = 0; i < nb; ++i) {
      ll_out[i] = res[i].ll;
      if (return_grad)
        for (int k = 0; k < gamma_n; ++k) grad_out[(size_t)i * gamma_n + k] = res[i].grad[k];
      if (status_out) status_out[i] = res[i].status;
      if (inc_out) inc_out[i] = res[i].inc;
    }
  });
}
}  // namespace

int lkgpu_ll_grad_batch(lkgpu_context* ctx, int nb, const int* problem_ids, const double* gamma, int gamma_n,
                        const lkgpu_model_flags* flags, int flags_n, int return_grad, double* ll_out,
                        double* grad_out, int* status_out, int* inc_out) {
  return ll_grad_batch(ctx, nb, problem_ids, gamma, gamma_n, flags, flags_n, return_grad, ll_out, grad_out,
                       status_out, inc_out, false);
}

int lkgpu_ll_grad_batch_noalloc(lkgpu_context* ctx, int nb, const int* problem_ids, const double* gamma,
                                int gamma_n, const lkgpu_model_flags* flags, int flags_n, int return_grad,
                                double* ll_out, double* grad_out, int* status_out, int* inc_out) {
  return ll_grad_batch(ctx, nb, problem_ids, gamma, gamma_n, flags, flags_n, return_grad, ll_out, grad_out,
                       status_out, inc_out, true);
}

int lkgpu_synchronize(lkgpu_context* ctx) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    lk::hip_check(hipSetDevice(ctx->rt.device()), "hipSetDevice");
    lk::hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}


int lkgpu_set_streams(lkgpu_context* ctx, int streams) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.set_streams(streams);
  });
}

int lkgpu_set_lookahead(lkgpu_context* ctx, int enable) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.set_lookahead(enable != 0);
  });
}

int lkgpu_set_tuning(lkgpu_context* ctx, const char* json_text) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    if (json_text == nullptr) {
      std::vector<lk::BucketTuning> t;
      for (int b : {4, 8, 16, 32, 64, 128, 1 << 30}) t.push_back(lk::BucketTuning{b, 4, 4});
      ctx->rt.set_tuning(t);
    } else {
      ctx->rt.load_tuning_json(json_text);
    }
  });
}

int lkgpu_last_parts(lkgpu_context* ctx) { return ctx ? ctx->rt.last_parts() : -1; }


int lkgpu_set_timed_stage(lkgpu_context* ctx, int stage) {
  return guarded([&] {
    if (ctx == nullptr) throw std::invalid_argument("context is NULL");
    ctx->rt.set_timed_stage(stage);
  });
}

int lkgpu_take_stage_time(lkgpu_context* ctx, double* avg_ms, int* count) {
  return guarded([&] {
    if (ctx == nullptr || avg_ms == nullptr || count == nullptr) throw std::invalid_argument("NULL argument");
    ctx->rt.take_stage_time(avg_ms, count);
  });
}

int lkgpu_read_item(lkgpu_context* ctx, int item, int what, double* out) {
  return guarded([&] {
    if (ctx == nullptr || out == nullptr) throw std::invalid_argument("NULL argument");
    ctx->rt.read_slot(item, what, out);
  });
}

}  // extern "C"
