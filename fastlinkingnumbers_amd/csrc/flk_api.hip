// C ABI of the linking-number engine (include/flk.h).  Translates C++
// exceptions into status codes and keeps the last error message per context.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstring>

#include "flk_internal.h"

namespace flk {
void stage_vertices_launch(flk_ctx *ctx);
}

using flk::Error;

static thread_local std::string g_create_err;

#define FLK_GUARD(ctx, ...)                                  \
  try {                                                      \
    __VA_ARGS__                                              \
  } catch (const flk::Error &e) {                            \
    if (ctx) (ctx)->err = e.msg;                             \
    return e.code;                                           \
  } catch (const std::bad_alloc &) {                         \
    if (ctx) (ctx)->err = "host allocation failed";          \
    return FLK_ERR_NOMEM;                                    \
  } catch (const std::exception &e) {                        \
    if (ctx) (ctx)->err = e.what();                          \
    return FLK_ERR_STATE;                                    \
  }

namespace flk {

int64_t curves_layout(flk_ctx *ctx, const int64_t *offsets, int32_t ncurves) {
  Curves &cv = ctx->curves;
  if (ncurves > 0 && offsets[0] != 0) throw Error(FLK_ERR_ARG, "offsets[0] must be 0");
  for (int32_t c = 0; c < ncurves; ++c)
    if (offsets[c + 1] < offsets[c]) throw Error(FLK_ERR_ARG, "offsets must be non-decreasing");
  for (int32_t c = 0; c < ncurves; ++c)
    if (offsets[c + 1] - offsets[c] > INT32_MAX - 1)
      throw Error(FLK_ERR_ARG, "a curve has more than 2^31-2 vertices");
  const int64_t nv = ncurves ? offsets[ncurves] : 0;
  cv.ncurves = ncurves;
  if (ncurves == 0) cv.in_off.assign(1, 0);
  else cv.in_off.assign(offsets, offsets + ncurves + 1);
  cv.n.resize(ncurves);
  cv.off.resize(ncurves + 1);
  for (int32_t c = 0; c < ncurves; ++c) {
    cv.n[c] = (int32_t)(offsets[c + 1] - offsets[c]);
    cv.off[c] = offsets[c] + c;  // one closing vertex per curve
  }
  cv.off[ncurves] = nv + ncurves;
  cv.nvert_dev = nv + ncurves;
  const int64_t nd = std::max<int64_t>(cv.nvert_dev, 1);
  cv.raw.reserve(sizeof(double) * 3 * std::max<int64_t>(nv, 1));
  cv.d_off.reserve(sizeof(int64_t) * (ncurves + 1));
  cv.dpts.reserve(sizeof(double4) * nd);
  cv.hi.reserve(sizeof(float4) * nd);  // also the FP32 landing area of the upload
  cv.lo.reserve(sizeof(float4) * nd);
  cv.ed.reserve(sizeof(float4) * nd);
  cv.inexact = false;
  cv.bh_budget = 0;
  ++ctx->curves_version;
  for (auto &g : ctx->ds_graph) g.reset();
  ctx->cert_graph.reset();
  return nv;
}

void curves_clear(flk_ctx *ctx) {
  Curves &cv = ctx->curves;
  cv.ncurves = 0;
  cv.in_off.assign(1, 0);
  cv.off.assign(1, 0);
  cv.n.clear();
  cv.box.clear();
  cv.nvert_dev = 0;
  cv.inexact = false;
  cv.bh_budget = 0;
  ++ctx->curves_version;
  for (auto &g : ctx->ds_graph) g.reset();
  ctx->cert_graph.reset();
}

void curves_finish(flk_ctx *ctx) {
  Curves &cv = ctx->curves;
  const int32_t ncurves = cv.ncurves;
  hipStream_t s = ctx->stream;
  curve_boxes_run(ctx, cv.box);
  // Origin and power-of-two scale: stored = (x - origin) * scale with the
  // model extent near 1.  The origin stays 0 unless the model is far from it
  // (then it is snapped to a coarse grid so fp32 inputs remain exact).
  double lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int32_t c = 0; c < ncurves; ++c)
    if (cv.box[6 * c] <= cv.box[6 * c + 3])
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::min(lo[d], cv.box[6 * c + d]);
        hi[d] = std::max(hi[d], cv.box[6 * c + 3 + d]);
      }
  double extent = 0.0, far = 0.0;
  for (int d = 0; d < 3; ++d)
    if (lo[d] <= hi[d]) {
      extent = std::max(extent, hi[d] - lo[d]);
      far = std::max(far, std::fabs(0.5 * (lo[d] + hi[d])));
    }
  cv.origin[0] = cv.origin[1] = cv.origin[2] = 0.0;
  if (extent > 0 && far > 4 * extent) {
    const double grid = std::ldexp(1.0, (int)std::floor(std::log2(extent)));
    for (int d = 0; d < 3; ++d) cv.origin[d] = grid * std::round(0.5 * (lo[d] + hi[d]) / grid);
  }
  // The scale is set by the largest curve, not by the model: two curves whose
  // boxes overlap lie within twice the largest box diagonal of each other,
  // so every relative vector of a candidate pair stays within length 1 (the
  // FP32 direct sum's products of six of them, and of twelve in its paired
  // short path, stay far from overflow), while segments keep their size
  // relative to that of a curve.  Scaling by the model's span instead shrank a
  // 160-unit-tall lattice of radius-1 rings by 256, and the sixth powers of
  // its 1e-7 relative vectors underflowed (a pair's FP32 sum landed 0.33 off
  // its integer).
  double dmax = 0.0, span = 0.0;
  for (int32_t c = 0; c < ncurves; ++c)
    if (cv.box[6 * c] <= cv.box[6 * c + 3]) {
      double dsq = 0.0;
      for (int d = 0; d < 3; ++d) dsq += (cv.box[6 * c + 3 + d] - cv.box[6 * c + d]) * (cv.box[6 * c + 3 + d] - cv.box[6 * c + d]);
      dmax = std::max(dmax, std::sqrt(dsq));
    }
  for (int d = 0; d < 3; ++d)
    if (lo[d] <= hi[d])
      span = std::max(span, std::max(std::fabs(hi[d] - cv.origin[d]), std::fabs(lo[d] - cv.origin[d])));
  const double ref = dmax > 0 ? 2.0 * dmax : span;
  cv.scale = ref > 0 ? std::ldexp(1.0, -(int)std::ceil(std::log2(ref))) : 1.0;
  if (!std::isfinite(cv.scale) || cv.scale == 0) cv.scale = 1.0;
  if (cv.nvert_dev > 0) {
    stage_vertices_launch(ctx);
    ctx->d_count.reserve(sizeof(int32_t));
    FLK_HIP(hipMemsetAsync(ctx->d_count.ptr, 0, sizeof(int32_t), s));
    convert_vertices_launch(s, cv.dpts.as<double4>(), cv.nvert_dev, cv.hi.as<float4>(), cv.lo.as<float4>(),
                            cv.ed.as<float4>(), ctx->d_count.as<int32_t>());
    int32_t flag = 0;
    FLK_HIP(hipMemcpyAsync(&flag, ctx->d_count.ptr, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    FLK_HIP(hipStreamSynchronize(s));
    cv.inexact = flag != 0;
  }
}

}  // namespace flk

namespace {

// Fallback and judge: the FP64 direct sum replaces every pair that lands
// farther than fallback_tol from an integer (or is not finite); then the NaN
// flags and the caller's outputs.  `noalloc` keeps the recomputation in the
// fixed direct-sum array.
void judge_and_finish(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params &prm,
                      std::vector<double> &v, const std::vector<double> &e, std::vector<uint32_t> &f, double *lk,
                      double *err_est, uint32_t *flags, bool noalloc) {
  if (prm.fallback_tol > 0 && !(prm.method == FLK_DIRECTSUM && prm.fp64)) {
    std::vector<int32_t> redo;
    std::vector<int64_t> idx;
    for (int64_t k = 0; k < npairs; ++k) {
      const double d = std::fabs(v[k] - std::round(v[k]));
      if (!(d <= prm.fallback_tol)) {
        redo.push_back(pairs[2 * k]);
        redo.push_back(pairs[2 * k + 1]);
        idx.push_back(k);
      }
    }
    if (!idx.empty()) {
      std::vector<double> w;
      float ms = 0.f;
      if (noalloc) flk::ds_run_noalloc(ctx, redo.data(), (int64_t)idx.size(), true, w, &ms);
      else flk::ds_run(ctx, redo.data(), (int64_t)idx.size(), true, w, &ms, prm.graph != 0);
      ctx->stats.ms_fallback = ms;
      ctx->stats.fallbacks = (int64_t)idx.size();
      for (size_t t = 0; t < idx.size(); ++t) {
        v[idx[t]] = w[t];
        f[idx[t]] |= FLK_PAIR_FALLBACK;
      }
    }
  }
  for (int64_t k = 0; k < npairs; ++k) {
    if (std::isnan(v[k])) f[k] |= FLK_PAIR_NAN;
    lk[k] = v[k];
    if (err_est) err_est[k] = e[k];
    if (flags) flags[k] = f[k];
  }
}

}  // namespace

extern "C" {

int32_t flk_default_params(flk_params *p) {
  if (!p) return FLK_ERR_ARG;
  std::memset(p, 0, sizeof *p);
  p->method = FLK_BARNESHUT;
  p->criterion = FLK_CRIT_BETA;
  p->beta_init = 2.0;
  p->beta_limit = 10.0;
  p->eps_sq = 1e-24;
  p->target_error = 0.1;
  p->err_scale = 0.05;
  p->fallback_tol = 0.05;
  p->bh_direct_threshold = 0;
  p->leaf_size = 1;
  p->fp64 = 0;
  p->mlp_log_tol = -6.0;
  p->graph = 1;
  return FLK_OK;
}

int32_t flk_create(int32_t device, flk_ctx **out) {
  if (!out) return FLK_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    g_create_err = "no HIP device available";
    return FLK_ERR_DEVICE;
  }
  if (device < 0 || device >= n) return FLK_ERR_ARG;
  flk_ctx *ctx = new flk_ctx();
  ctx->device = device;
  try {
    FLK_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    FLK_HIP(hipGetDeviceProperties(&prop, device));
    ctx->num_cus = prop.multiProcessorCount;
    ctx->name = prop.gcnArchName;
    FLK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    FLK_HIP(hipEventCreate(&ctx->ev0));
    FLK_HIP(hipEventCreate(&ctx->ev1));
    FLK_HIP(hipEventCreate(&ctx->ev2));
    FLK_HIP(hipEventCreate(&ctx->ev3));
    FLK_HIP(hipStreamCreateWithFlags(&ctx->aux, hipStreamNonBlocking));
    FLK_HIP(hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming));
    FLK_HIP(hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming));
    ctx->noalloc_scratch = flk::noalloc_scratch_init(device);  // once per device, this device current
  } catch (const flk::Error &e) {
    g_create_err = e.msg;
    delete ctx;
    return e.code;
  }
  *out = ctx;
  return FLK_OK;
}

void flk_destroy(flk_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  flk::DevBuf *bufs[] = {&ctx->curves.raw, &ctx->curves.dpts, &ctx->curves.hi, &ctx->curves.lo,
                         &ctx->curves.ed, &ctx->curves.d_off, &ctx->d_pairs, &ctx->d_tasks,
                         &ctx->d_cost, &ctx->d_task_theta, &ctx->d_task_corr,
                         &ctx->d_pair_task0, &ctx->d_pair_out, &ctx->d_pair_corr,
                         &ctx->d_boxes, &ctx->d_pls, &ctx->d_count, &ctx->d_misc,
                         &ctx->mlp.d_params, &ctx->mlp.d_norm, &ctx->mlp.d_bf16,
                         &ctx->mlp.d_image};
  for (auto *b : bufs) b->release();
  ctx->cert_graph.reset();
  ctx->bh.release();
  ctx->stage.release();
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
  if (ctx->ev3) (void)hipEventDestroy(ctx->ev3);
  if (ctx->ev_fork) (void)hipEventDestroy(ctx->ev_fork);
  if (ctx->ev_join) (void)hipEventDestroy(ctx->ev_join);
  if (ctx->aux) (void)hipStreamDestroy(ctx->aux);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

const char *flk_last_error(const flk_ctx *ctx) {
  return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

int32_t flk_device_name(const flk_ctx *ctx, char *buf, int32_t len) {
  if (!ctx || !buf || len <= 0) return FLK_ERR_ARG;
  std::snprintf(buf, (size_t)len, "%s", ctx->name.c_str());
  return FLK_OK;
}

int32_t flk_set_curves(flk_ctx *ctx, const double *xyz, const int64_t *offsets, int32_t ncurves) {
  if (!ctx || ncurves < 0 || (ncurves > 0 && !offsets)) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    if (ncurves > 0 && offsets[ncurves] > 0 && !xyz) throw Error(FLK_ERR_ARG, "xyz is NULL");
    try {
      const int64_t nv = flk::curves_layout(ctx, offsets, ncurves);
      flk::Curves &cv = ctx->curves;
      flk::ingest_points(ctx, xyz, false, nv, cv.raw.as<double>(), cv.hi.as<float>(), cv.hi.bytes);
      flk::curves_finish(ctx);
    } catch (...) {
      flk::curves_clear(ctx);  // never leave a half-built curve set behind
      throw;
    }
    return FLK_OK;
  })
}

int32_t flk_set_curves_f32(flk_ctx *ctx, const float *xyz, const int64_t *offsets, int32_t ncurves) {
  if (!ctx || ncurves < 0 || (ncurves > 0 && !offsets)) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    if (ncurves > 0 && offsets[ncurves] > 0 && !xyz) throw Error(FLK_ERR_ARG, "xyz is NULL");
    try {
      const int64_t nv = flk::curves_layout(ctx, offsets, ncurves);
      flk::Curves &cv = ctx->curves;
      flk::ingest_points(ctx, xyz, true, nv, cv.raw.as<double>(), cv.hi.as<float>(), cv.hi.bytes);
      flk::curves_finish(ctx);
    } catch (...) {
      flk::curves_clear(ctx);  // never leave a half-built curve set behind
      throw;
    }
    return FLK_OK;
  })
}

int32_t flk_curve_boxes(flk_ctx *ctx, double *boxes_out) {
  if (!ctx || (!boxes_out && ctx->curves.ncurves)) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    std::copy(ctx->curves.box.begin(), ctx->curves.box.end(), boxes_out);
    return FLK_OK;
  })
}

int32_t flk_potential_links(flk_ctx *ctx, const double *boxes, int32_t ncurves,
                            const int32_t **pairs, int64_t *npairs) {
  if (!ctx || !pairs || !npairs || ncurves < 0) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    const double *b = boxes;
    if (!b) {
      if (ncurves != ctx->curves.ncurves)
        throw Error(FLK_ERR_ARG, "ncurves does not match the uploaded curves");
      b = ctx->curves.box.data();
    }
    flk::pls_run(ctx, b, ncurves, ctx->pls_pairs);
    *pairs = ctx->pls_pairs.data();
    *npairs = (int64_t)ctx->pls_pairs.size() / 2;
    return FLK_OK;
  })
}

int32_t flk_linking_numbers(flk_ctx *ctx, const int32_t *pairs, int64_t npairs,
                            const flk_params *params, double *lk, double *err_est,
                            uint32_t *flags) {
  if (!ctx || npairs < 0 || (npairs > 0 && (!pairs || !lk))) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    flk_params prm;
    if (params) prm = *params;
    else flk_default_params(&prm);
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    const flk::Curves &cv = ctx->curves;
    for (int64_t k = 0; k < 2 * npairs; ++k)
      if (pairs[k] < 0 || pairs[k] >= cv.ncurves) throw Error(FLK_ERR_ARG, "pair index out of range");
    for (int64_t k = 0; k < npairs; ++k)
      ctx->stats.seg_pairs += (int64_t)cv.n[pairs[2 * k]] * cv.n[pairs[2 * k + 1]];
    std::vector<double> v, e(npairs, 0.0);
    std::vector<uint32_t> f(npairs, 0u);
    if (prm.method == FLK_DIRECTSUM) {
      float ms = 0.f;
      flk::ds_run(ctx, pairs, npairs, prm.fp64 != 0, v, &ms, prm.graph != 0);
      ctx->stats.ms_kernel = ms;
      ctx->stats.ms_ds_kernel = ms;
    } else if (prm.method == FLK_BARNESHUT) {
      if (prm.bh_batch_ksegs < 0) throw Error(FLK_ERR_ARG, "bh_batch_ksegs must be >= 0");
      flk::bh_run(ctx, pairs, npairs, prm, v, e, f);
    } else {
      throw Error(FLK_ERR_ARG, "invalid method");
    }
    judge_and_finish(ctx, pairs, npairs, prm, v, e, f, lk, err_est, flags, false);
    ctx->stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FLK_OK;
  })
}

int32_t flk_linking_numbers_noalloc(flk_ctx *ctx, const int32_t *pairs, int64_t npairs,
                                    const flk_params *params, double *lk, double *err_est,
                                    uint32_t *flags) {
  if (!ctx || npairs < 0 || (npairs > 0 && (!pairs || !lk))) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    flk_params prm;
    if (params) prm = *params;
    else flk_default_params(&prm);
    if (prm.method != FLK_DIRECTSUM)
      throw Error(FLK_ERR_ARG, "flk_linking_numbers_noalloc runs the direct sum only (method FLK_DIRECTSUM)");
    const auto t0 = std::chrono::steady_clock::now();
    std::memset(&ctx->stats, 0, sizeof ctx->stats);
    const flk::Curves &cv = ctx->curves;
    for (int64_t k = 0; k < 2 * npairs; ++k)
      if (pairs[k] < 0 || pairs[k] >= cv.ncurves) throw Error(FLK_ERR_ARG, "pair index out of range");
    for (int64_t k = 0; k < npairs; ++k)
      ctx->stats.seg_pairs += (int64_t)cv.n[pairs[2 * k]] * cv.n[pairs[2 * k + 1]];
    std::vector<double> v, e(npairs, 0.0);
    std::vector<uint32_t> f(npairs, 0u);
    float ms = 0.f;
    flk::ds_run_noalloc(ctx, pairs, npairs, prm.fp64 != 0, v, &ms);
    ctx->stats.ms_kernel = ms;
    ctx->stats.ms_ds_kernel = ms;
    judge_and_finish(ctx, pairs, npairs, prm, v, e, f, lk, err_est, flags, true);
    ctx->stats.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return FLK_OK;
  })
}

int32_t flk_get_stats(const flk_ctx *ctx, flk_stats *out) {
  if (!ctx || !out) return FLK_ERR_ARG;
  *out = ctx->stats;
  return FLK_OK;
}

int32_t flk_tree_moments(flk_ctx *ctx, int32_t curve, int32_t leaf_size, flk_tree_node *nodes,
                         int32_t cap, int32_t *nnodes, int32_t *seg_order) {
  if (!ctx || !nnodes) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    if (curve < 0 || curve >= ctx->curves.ncurves) throw Error(FLK_ERR_ARG, "curve index out of range");
    if (leaf_size < 1) throw Error(FLK_ERR_ARG, "leaf_size must be >= 1");
    std::vector<flk_tree_node> t;
    std::vector<int32_t> order;
    flk::tree_export(ctx, curve, leaf_size, t, order);
    *nnodes = (int32_t)t.size();
    if (nodes) {
      if (cap < (int32_t)t.size()) throw Error(FLK_ERR_ARG, "cap is smaller than the node count");
      std::copy(t.begin(), t.end(), nodes);
    }
    if (seg_order) std::copy(order.begin(), order.end(), seg_order);
    return FLK_OK;
  })
}

int32_t flk_discretize(flk_ctx *ctx, const double *ctrl_xyz, const int64_t *offsets,
                       const int32_t *segment_types, const uint8_t *looped, int32_t ncurves,
                       const int32_t *pairs, int64_t npairs, const double **out_xyz,
                       const int64_t **out_offsets, flk_disc_info *info) {
  if (!ctx || ncurves < 0 || npairs < 0) return FLK_ERR_ARG;
  if (ncurves > 0 && (!offsets || !segment_types || !looped)) return FLK_ERR_ARG;
  if (npairs > 0 && !pairs) return FLK_ERR_ARG;
  auto fill = [&]() {
    if (!info) return;
    std::memset(info, 0, sizeof *info);
    const flk::DiscResult &r = ctx->disc;
    info->segments = r.segments;
    info->vertices = r.out_offsets.empty() ? 0 : r.out_offsets.back();
    info->rounds = r.rounds;
    info->pieces = r.retired;
    info->small_curve = r.small_curve;
    info->small_segment = r.small_segment;
    info->small_size = r.small_size;
    for (size_t k = 0; k < r.overlaps.size() && k < 55; ++k) info->overlaps[k] = r.overlaps[k];
  };
  try {
    FLK_HIP(hipSetDevice(ctx->device));
    if (ncurves > 0 && offsets[0] != 0) throw Error(FLK_ERR_ARG, "offsets[0] must be 0");
    std::vector<double> current;
    if (!ctrl_xyz && ncurves > 0 && offsets[ncurves] > 0) {
      // the control points are the current curve set's input array; the
      // reference's guard and extent loops read them on the host
      const flk::Curves &cv = ctx->curves;
      if (cv.ncurves != ncurves || !std::equal(offsets, offsets + ncurves + 1, cv.in_off.begin()))
        throw Error(FLK_ERR_ARG, "offsets do not match the current curve set");
      current.resize(3 * (size_t)offsets[ncurves]);
      FLK_HIP(hipMemcpyAsync(current.data(), cv.raw.ptr, sizeof(double) * current.size(), hipMemcpyDeviceToHost,
                             ctx->stream));
      FLK_HIP(hipStreamSynchronize(ctx->stream));
      ctrl_xyz = current.data();
    }
    flk::discretize_run(ctx, ctrl_xyz, offsets, segment_types, looped, ncurves, pairs, npairs, ctx->disc,
                        out_xyz != nullptr);
  } catch (const flk::Error &e) {
    ctx->disc.rounds = (int32_t)ctx->disc.overlaps.size();
    fill();
    ctx->err = e.msg;
    return e.code;
  } catch (const std::bad_alloc &) {
    ctx->err = "host allocation failed";
    return FLK_ERR_NOMEM;
  } catch (const std::exception &e) {
    ctx->err = e.what();
    return FLK_ERR_STATE;
  }
  fill();
  if (out_xyz) *out_xyz = ctx->disc.xyz.data();
  if (out_offsets) *out_offsets = ctx->disc.out_offsets.data();
  return FLK_OK;
}

int32_t flk_load_bcc(flk_ctx *ctx, const char *path, flk_bcc_info *info) {
  if (!ctx || !path) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    flk::load_bcc(ctx, path, ctx->bcc);
    if (info) {
      std::memset(info, 0, sizeof *info);
      info->ncurves = ctx->curves.ncurves;
      info->segment_type = ctx->bcc.type;
      info->points = ctx->bcc.points;
      info->offsets = ctx->bcc.offsets.data();
      info->looped = ctx->bcc.looped.data();
    }
    return FLK_OK;
  })
}

int32_t flk_get_tune(const flk_ctx *ctx, flk_tune *out) {
  if (!ctx || !out) return FLK_ERR_ARG;
  *out = ctx->tune;
  return FLK_OK;
}

int32_t flk_set_tune(flk_ctx *ctx, const flk_tune *tune) {
  if (!ctx || !tune) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    if (tune->ds_rows_per_lane != 2 && tune->ds_rows_per_lane != 4 && tune->ds_rows_per_lane != 8)
      throw Error(FLK_ERR_ARG, "ds_rows_per_lane must be 2, 4 or 8");
    if (tune->ds_blocks_per_cu < 0 || tune->ds_blocks_per_cu > 64)
      throw Error(FLK_ERR_ARG, "ds_blocks_per_cu must be in [0, 64]");
    if (tune->ds_tasks_per_wave < 1 || tune->ds_tasks_per_wave > 1024)
      throw Error(FLK_ERR_ARG, "ds_tasks_per_wave must be in [1, 1024]");
    if (tune->bh_items_per_thread != 0 && tune->bh_items_per_thread != 2 && tune->bh_items_per_thread != 4 &&
        tune->bh_items_per_thread != 8)
      throw Error(FLK_ERR_ARG, "bh_items_per_thread must be 0 (by criterion), 2, 4 or 8");
    if (tune->ingest_threads < 0 || tune->ingest_threads > 64)
      throw Error(FLK_ERR_ARG, "ingest_threads must be in [0, 64]");
    if (tune->bh_level_capacity != 0 && tune->bh_level_capacity < 64)
      throw Error(FLK_ERR_ARG, "bh_level_capacity must be 0 (automatic) or at least 64");
    ctx->tune = *tune;
  })
  return FLK_OK;
}

// Direct-sum candidates run on every k-th 64R-row block of the pairs' direct-
// sum work, k chosen so one run covers about a tenth of it (at least 2e10
// segment pairs, or all of it: smaller samples over-weight the end-of-launch
// tail); Barnes–Hut candidates run the Barnes–Hut pairs whole.  Each
// candidate runs twice and keeps its faster time.
int32_t flk_autotune(flk_ctx *ctx, const int32_t *pairs, int64_t npairs, const flk_params *params,
                     flk_tune *out) {
  if (!ctx || (npairs > 0 && !pairs) || npairs < 0) return FLK_ERR_ARG;
  FLK_GUARD(ctx, {
    FLK_HIP(hipSetDevice(ctx->device));
    flk_params prm;
    if (params) prm = *params;
    else flk_default_params(&prm);
    const flk::Curves &cv = ctx->curves;
    for (int64_t k = 0; k < 2 * npairs; ++k)
      if (pairs[k] < 0 || pairs[k] >= cv.ncurves) throw Error(FLK_ERR_ARG, "pair index out of range");
    // split as flk_linking_numbers would
    std::vector<int32_t> ds_pairs, bh_pairs;
    for (int64_t k = 0; k < npairs; ++k) {
      const int32_t a = pairs[2 * k], b = pairs[2 * k + 1];
      const bool direct = prm.method == FLK_DIRECTSUM || cv.n[a] < 2 || cv.n[b] < 2 ||
                          (int64_t)cv.n[a] * cv.n[b] < prm.bh_direct_threshold;
      (direct ? ds_pairs : bh_pairs).insert((direct ? ds_pairs : bh_pairs).end(), {a, b});
    }
    const flk_stats saved = ctx->stats;
    flk_tune best = ctx->tune;
    const int64_t nds = (int64_t)ds_pairs.size() / 2;
    if (nds > 0 && !(prm.method == FLK_DIRECTSUM && prm.fp64)) {
      const int64_t total = flk::ds_work(ctx, ds_pairs.data(), nds);
      const int64_t target = std::max<int64_t>((int64_t)2e10, total / 10);
      double best_rate = -1.0;
      // every stride-th row block: about target / total of the work
      const int64_t stride = std::max<int64_t>(1, total / target);
      for (int R : {2, 4, 8}) {
        for (int bpc : {0, 4}) {  // all blocks resident, or 4 per CU
          for (int tpw : {8, 16}) {
            flk_tune t = ctx->tune;
            t.ds_rows_per_lane = R;
            t.ds_tasks_per_wave = tpw;
            t.ds_blocks_per_cu = bpc;
            float ms = 1e30f;
            int64_t work = 0;
            for (int rep = 0; rep < 2; ++rep)
              ms = std::min(ms, flk::ds_time_sample(ctx, ds_pairs.data(), nds, t, stride, &work));
            const double rate = ms > 0.f ? (double)work / ms : 0.0;
            if (rate > best_rate) {
              best_rate = rate;
              best.ds_rows_per_lane = t.ds_rows_per_lane;
              best.ds_blocks_per_cu = t.ds_blocks_per_cu;
              best.ds_tasks_per_wave = t.ds_tasks_per_wave;
            }
          }
        }
      }
    }
    const int64_t nbh = (int64_t)bh_pairs.size() / 2;
    if (nbh > 0) {
      double best_ms = 1e300;
      flk_params bp = prm;
      bp.method = FLK_BARNESHUT;
      bp.bh_direct_threshold = 0;
      bp.timing = 0;
      for (int ipt : {2, 4, 8}) {
        ctx->tune.bh_items_per_thread = ipt;
        double ms = 1e300;
        for (int rep = 0; rep < 2; ++rep) {
          std::vector<double> v, e;
          std::vector<uint32_t> f;
          const auto t0 = std::chrono::steady_clock::now();
          flk::bh_run(ctx, bh_pairs.data(), nbh, bp, v, e, f);
          ms = std::min(ms, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
        if (ms < best_ms) {
          best_ms = ms;
          best.bh_items_per_thread = ipt;
        }
      }
    }
    ctx->tune = best;
    ctx->stats = saved;
    if (out) *out = best;
  })
  return FLK_OK;
}

}  // extern "C"
