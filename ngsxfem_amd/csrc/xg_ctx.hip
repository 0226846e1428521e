// xg_ctx.hip -- context, error reporting and scratch management of the C ABI.
//
// The reference runs every hot call inside an NGSolve LocalHeap on
// TaskManager threads (utils/ngsxstd.cpp:3-31) and reports failures with C++
// exceptions.  Here a context binds a device + stream + scratch; failures
// are status codes with a thread-local message, and data-dependent failures
// detected inside kernels (the conditions on which the reference throws,
// e.g. straightcutrule.cpp:148,193) raise a sticky device flag that
// xg_sync() reports and clears.
#include <stdio.h>

#include <string>

#include "xg_internal.hpp"

namespace {
thread_local std::string g_last_error;
}

namespace xg {

int set_error(int code, const std::string &msg) {
  g_last_error = msg;
  return code;
}

int check_hip(hipError_t e, const char *what) {
  g_last_error = std::string("HIP error ") + hipGetErrorString(e) + " in " + what;
  return XG_ERR_HIP;
}

int ensure_buf(xg_ctx *ctx, void **buf, size_t *size, size_t bytes) {
  if (*size >= bytes) return XG_OK;
  if (*buf) {
    XG_CHECK(hipStreamSynchronize(ctx->stream));
    XG_CHECK(hipFree(*buf));
    *buf = nullptr;
    *size = 0;
  }
  size_t sz = 1 << 20;
  while (sz < bytes) sz <<= 1;
  XG_CHECK(hipMalloc(buf, sz));
  *size = sz;
  return XG_OK;
}

int ensure_ws(xg_ctx *ctx, size_t bytes) { return ensure_buf(ctx, &ctx->ws, &ctx->ws_size, bytes); }

KernelTimer::KernelTimer(xg_ctx *c, const char *name) : ctx(c) {
  if (!ctx->timing) return;
  for (size_t k = 0; k < ctx->tnames.size(); ++k)
    if (ctx->tnames[k] == name) slot = (int)k;
  if (slot < 0) {
    slot = (int)ctx->tnames.size();
    ctx->tnames.push_back(name);
    ctx->tevents.emplace_back();
  }
  if (hipEventCreate(&a) != hipSuccess || hipEventRecord(a, ctx->stream) != hipSuccess) {
    if (a) hipEventDestroy(a);
    a = nullptr;
  }
}

KernelTimer::~KernelTimer() {
  if (!a) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) != hipSuccess || hipEventRecord(b, ctx->stream) != hipSuccess) {
    if (b) hipEventDestroy(b);
    hipEventDestroy(a);
    return;
  }
  ctx->tevents[slot].emplace_back(a, b);
}

void timing_clear(xg_ctx *ctx) {
  if (ctx->tnames.empty()) return;
  hipStreamSynchronize(ctx->stream);
  for (auto &v : ctx->tevents)
    for (auto &e : v) {
      hipEventDestroy(e.first);
      hipEventDestroy(e.second);
    }
  ctx->tnames.clear();
  ctx->tevents.clear();
}

}  // namespace xg

extern "C" const char *xg_last_error(void) { return g_last_error.c_str(); }

extern "C" int xg_timing_enable(xg_ctx *ctx, int on) {
  if (!ctx) return xg::set_error(XG_ERR_ARG, "xg_timing_enable: null ctx");
  if (on) xg::timing_clear(ctx);
  ctx->timing = on != 0;
  return XG_OK;
}

extern "C" int xg_timing_read(xg_ctx *ctx, const char *name, double *total_ms, int *launches) {
  if (!ctx || !name || !total_ms || !launches) return xg::set_error(XG_ERR_ARG, "xg_timing_read: null argument");
  *total_ms = 0.0;
  *launches = 0;
  XG_CHECK(hipStreamSynchronize(ctx->stream));
  for (size_t k = 0; k < ctx->tnames.size(); ++k) {
    if (ctx->tnames[k] != name) continue;
    for (auto &e : ctx->tevents[k]) {
      float ms = 0.0f;
      XG_CHECK(hipEventElapsedTime(&ms, e.first, e.second));
      *total_ms += ms;
      *launches += 1;
    }
  }
  return XG_OK;
}

extern "C" int xg_ctx_create(int device, xg_ctx **out) {
  if (!out) return xg::set_error(XG_ERR_ARG, "xg_ctx_create: null out");
  int ndev = 0;
  XG_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return xg::set_error(XG_ERR_ARG, "xg_ctx_create: no such device");
  XG_CHECK(hipSetDevice(device));
  xg_ctx *c = new xg_ctx();
  c->device = device;
  c->stream = nullptr;
  hipError_t e = hipMalloc(&c->d_err, sizeof(int32_t));
  if (e != hipSuccess) {
    delete c;
    return xg::check_hip(e, "hipMalloc(err)");
  }
  e = hipHostMalloc(&c->h_err, sizeof(int32_t), hipHostMallocDefault);
  if (e != hipSuccess) {
    hipFree(c->d_err);
    delete c;
    return xg::check_hip(e, "hipHostMalloc(err)");
  }
  *c->h_err = 0;
  e = hipMemset(c->d_err, 0, sizeof(int32_t));
  if (e != hipSuccess) {
    hipFree(c->d_err);
    hipHostFree(c->h_err);
    delete c;
    return xg::check_hip(e, "hipMemset(err)");
  }
  // per-device set-up of the no-alloc entry point's scratch address
  const int r = xg::noalloc_scratch(device, &c->noalloc_scratch);
  if (r != XG_OK) {
    hipFree(c->d_err);
    hipHostFree(c->h_err);
    delete c;
    return r;
  }
  *out = c;
  return XG_OK;
}

extern "C" int xg_ctx_destroy(xg_ctx *ctx) {
  if (!ctx) return XG_OK;
  hipSetDevice(ctx->device);
  if (ctx->stream) hipStreamSynchronize(ctx->stream);
  else hipDeviceSynchronize();
  xg::timing_clear(ctx);
  if (ctx->ws) hipFree(ctx->ws);
  if (ctx->sdx_ws) hipFree(ctx->sdx_ws);
  if (ctx->sdx_verts) hipFree(ctx->sdx_verts);
  for (int k = 0; k < 4; ++k)
    if (ctx->sdx_mom[k]) hipFree(ctx->sdx_mom[k]);
  if (ctx->d_err) hipFree(ctx->d_err);
  if (ctx->h_err) hipHostFree(ctx->h_err);
  delete ctx;
  return XG_OK;
}

extern "C" int xg_ctx_set_stream(xg_ctx *ctx, void *stream) {
  if (!ctx) return xg::set_error(XG_ERR_ARG, "xg_ctx_set_stream: null ctx");
  ctx->stream = (hipStream_t)stream;
  return XG_OK;
}

extern "C" int xg_sync(xg_ctx *ctx) {
  if (!ctx) return xg::set_error(XG_ERR_ARG, "xg_sync: null ctx");
  XG_CHECK(hipMemcpyAsync(ctx->h_err, ctx->d_err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  XG_CHECK(hipStreamSynchronize(ctx->stream));
  const int32_t flag = *ctx->h_err;
  if (flag) {
    XG_CHECK(hipMemsetAsync(ctx->d_err, 0, sizeof(int32_t), ctx->stream));
    XG_CHECK(hipStreamSynchronize(ctx->stream));
    if (flag & 2) return xg::set_error(XG_ERR_SUBDIV, "subdivision rule: did not expect this.. -3-");
    if (flag & 4) return xg::set_error(XG_ERR_CAPACITY, "output capacity exceeded");
    if (flag & 8) return xg::set_error(XG_ERR_ARG, "node or dof number out of range");
    if (flag & 16)
      return xg::set_error(XG_ERR_DEGENERATE,
                           "quadrilateral cut rule: if_scale_factor larger than bound, or a NaN direction bound "
                           "(straightcutrule.cpp:283-287, :438)");
    return xg::set_error(XG_ERR_DEGENERATE,
                         "You tried to cut a simplex with a plain geometry lset function "
                         "(a vertex level-set value is exactly 0 next to a negative one)");
  }
  return XG_OK;
}
