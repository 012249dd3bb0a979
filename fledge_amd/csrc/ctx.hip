// Context, error reporting, device memory helpers and kernel timers of the
// fledge HIP library.  The ownership convention mirrors fledge's Resource
// arena (src/resource.hpp:45-120): handles own what they allocate; the
// failure convention replaces SErr/assert aborts (src/debug.hpp:60-69) with a
// status code plus flg_last_error().
#include <cstdio>

#include "bvh.hpp"
#include "halton.hpp"

namespace flg {

int set_error(flg_ctx* ctx, int status, const std::string& msg) {
  if (ctx) {
    try {
      ctx->last_error = msg;
    } catch (...) {  // the status still reports the error
      ctx->last_error.clear();
    }
  }
  return status;
}

int check_hip(flg_ctx* ctx, hipError_t e, const char* what) {
  char buf[512];
  snprintf(buf, sizeof buf, "%s failed: %s", what, hipGetErrorString(e));
  int st = (e == hipErrorOutOfMemory) ? FLG_ERR_OOM
           : (e == hipErrorNoDevice || e == hipErrorInvalidDevice) ? FLG_ERR_NO_DEVICE
                                                                   : FLG_ERR_HIP;
  return set_error(ctx, st, buf);
}

void* scratch(flg_ctx* ctx, const char* key, size_t bytes, int* status) {
  auto& slot = ctx->scratch[key];
  if (slot.second >= bytes && slot.first) return slot.first;
  if (slot.first) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(slot.first);
    slot.first = nullptr;
    slot.second = 0;
  }
  size_t cap = bytes < 256 ? 256 : bytes + bytes / 8;
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, cap);
  if (e != hipSuccess) {
    *status = check_hip(ctx, e, key);
    return nullptr;
  }
  slot.first = p;
  slot.second = cap;
  *status = FLG_OK;
  return p;
}

int graph_launch(flg_ctx* ctx, const std::string& key, const std::function<void(hipStream_t)>& enqueue,
                 const char* what) {
  hipGraphExec_t exec = nullptr;
  for (auto& kv : ctx->graphs)
    if (kv.first == key) exec = kv.second;
  if (!exec) {
    if (!ctx->capture_stream)
      FLG_CHECK_HIP(ctx, hipStreamCreateWithFlags(&ctx->capture_stream, hipStreamNonBlocking));
    FLG_CHECK_HIP(ctx, hipStreamBeginCapture(ctx->capture_stream, hipStreamCaptureModeThreadLocal));
    enqueue(ctx->capture_stream);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->capture_stream, &graph);
    if (e != hipSuccess) return check_hip(ctx, e, (std::string("hipStreamEndCapture (") + what + ")").c_str());
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return check_hip(ctx, e, (std::string("hipGraphInstantiate (") + what + ")").c_str());
    if (ctx->graphs.size() >= 8) {
      (void)hipStreamSynchronize(ctx->stream);
      (void)hipGraphExecDestroy(ctx->graphs.front().second);
      ctx->graphs.erase(ctx->graphs.begin());
    }
    ctx->graphs.emplace_back(key, exec);
  }
  FLG_CHECK_HIP(ctx, hipGraphLaunch(exec, ctx->stream));
  return FLG_OK;
}

// Events are recorded around the launch without waiting; they are resolved
// (synchronised and accumulated) when the timers are read, or when too many
// are pending, so timing does not serialise the stream.
void resolve_timers(flg_ctx* ctx) {
  for (auto& p : ctx->pending) {
    if (hipEventSynchronize(p.b) == hipSuccess) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, p.a, p.b);
      auto& t = ctx->timers[p.name];
      t.ms += ms;
      t.launches += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  ctx->pending.clear();
}

ScopedTimer::ScopedTimer(flg_ctx* c, const char* n) : ctx(c), name(n) {
  if (!ctx || !ctx->profiling) return;  // a null context times nothing
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    a = b = nullptr;
    return;
  }
  (void)hipEventRecord(a, ctx->stream);
}

ScopedTimer::~ScopedTimer() {
  if (!a) return;
  (void)hipEventRecord(b, ctx->stream);
  ctx->pending.push_back(PendingTimer{name, a, b});
  if (ctx->pending.size() > 4096) resolve_timers(ctx);
}

// ------------------------------------------------------------ debug eval ---
// Test hook (flg_debug_eval): left out of builds with FLG_NO_TEST_HOOKS.
#ifndef FLG_NO_TEST_HOOKS
__global__ void debug_eval_kernel(int kind, const void* in, void* out, int64_t n,
                                  HaltonTables tab) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (kind == 0) {
    ((float*)out)[i] = flg_logf(((const float*)in)[i]);
  } else if (kind == 1 || kind == 2) {
    float s, c;
    flg_sincosf(((const float*)in)[i], &s, &c);
    ((float*)out)[i] = kind == 1 ? s : c;
  } else if (kind == 3) {
    const uint32_t* p = (const uint32_t*)in;
    ((float*)out)[i] = halton_get1d(tab, p[2 * i], p[2 * i + 1], 128, 243);
  } else if (kind == 4) {
    const uint32_t* p = (const uint32_t*)in;
    ((float*)out)[i] = walk_draw(walk_key(p[3 * i], p[3 * i + 1]), p[3 * i + 2]);
  }
}
#endif

}  // namespace flg

using namespace flg;

extern "C" {

const char* flg_status_string(int status) {
  switch (status) {
    case FLG_OK: return "ok";
    case FLG_ERR_INVALID: return "invalid argument";
    case FLG_ERR_HIP: return "HIP runtime error";
    case FLG_ERR_NO_DEVICE: return "no usable GPU";
    case FLG_ERR_OOM: return "device out of memory";
    case FLG_ERR_RANGE: return "input beyond a documented limit";
    default: return "unknown status";
  }
}

int flg_ctx_create(int device, void* stream, flg_ctx** out) {
  return host_guard(nullptr, "flg_ctx_create", [&]() -> int {
    if (!out) return FLG_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) return FLG_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return FLG_ERR_INVALID;
    if (hipSetDevice(device) != hipSuccess) return FLG_ERR_NO_DEVICE;
    flg_ctx* ctx = new flg_ctx;
    ctx->device = device;
    ctx->stream = (hipStream_t)stream;  // NULL: the device's default stream
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    // set-up that later calls must not repeat: the Halton tables, and this
    // device's address of the fixed scratch array (no allocation later in
    // flg_volpath_render_noalloc)
    HaltonTables tab;
    int st = get_halton_tables(ctx, &tab);
    if (st == FLG_OK) st = noalloc_scratch_address(device, &ctx->noalloc_scratch);
    if (st != FLG_OK) {
      (void)hipStreamSynchronize(ctx->stream);
      release_halton_tables(ctx);
      delete ctx;
      return st;
    }
    *out = ctx;
    return FLG_OK;
  });
}

int flg_ctx_destroy(flg_ctx* ctx) {
  return host_guard(ctx, "flg_ctx_destroy", [&]() -> int {
    if (!ctx) return FLG_ERR_INVALID;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    resolve_timers(ctx);
    for (auto& kv : ctx->graphs) (void)hipGraphExecDestroy(kv.second);
    if (ctx->capture_stream) (void)hipStreamDestroy(ctx->capture_stream);
    for (auto& kv : ctx->scratch)
      if (kv.second.first) (void)hipFree(kv.second.first);
    bvh_pool_release(ctx);
    release_halton_tables(ctx);
    delete ctx;
    return FLG_OK;
  });
}

int flg_ctx_synchronize(flg_ctx* ctx) {
  return host_guard(ctx, "flg_ctx_synchronize", [&]() -> int {
    if (!ctx) return FLG_ERR_INVALID;
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FLG_OK;
  });
}

void* flg_ctx_stream(flg_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

const char* flg_last_error(flg_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "null context"; }

int flg_ctx_set_profiling(flg_ctx* ctx, int enable) {
  return host_guard(ctx, "flg_ctx_set_profiling", [&]() -> int {
    if (!ctx) return FLG_ERR_INVALID;
    ctx->profiling = enable != 0;
    return FLG_OK;
  });
}

int flg_ctx_kernel_time(flg_ctx* ctx, const char* name, double* ms, int64_t* launches) {
  return host_guard(ctx, "flg_ctx_kernel_time", [&]() -> int {
    if (!ctx || !name) return FLG_ERR_INVALID;
    resolve_timers(ctx);
    auto it = ctx->timers.find(name);
    if (it == ctx->timers.end()) return set_error(ctx, FLG_ERR_INVALID, std::string("no timer ") + name);
    if (ms) *ms = it->second.ms;
    if (launches) *launches = it->second.launches;
    return FLG_OK;
  });
}

int flg_ctx_reset_profiling(flg_ctx* ctx) {
  return host_guard(ctx, "flg_ctx_reset_profiling", [&]() -> int {
    if (!ctx) return FLG_ERR_INVALID;
    resolve_timers(ctx);
    ctx->timers.clear();
    return FLG_OK;
  });
}

int flg_malloc(flg_ctx* ctx, size_t bytes, void** d_out) {
  return host_guard(ctx, "flg_malloc", [&]() -> int {
    if (!ctx || !d_out) return FLG_ERR_INVALID;
    *d_out = nullptr;
    FLG_CHECK_HIP(ctx, hipMalloc(d_out, bytes ? bytes : 1));
    return FLG_OK;
  });
}

int flg_free(flg_ctx* ctx, void* d_ptr) {
  return host_guard(ctx, "flg_free", [&]() -> int {
    if (!ctx) return FLG_ERR_INVALID;
    if (!d_ptr) return FLG_OK;
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FLG_CHECK_HIP(ctx, hipFree(d_ptr));
    return FLG_OK;
  });
}

int flg_memcpy_htod(flg_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  return host_guard(ctx, "flg_memcpy_htod", [&]() -> int {
    if (!ctx || (!d_dst && bytes) || (!h_src && bytes)) return FLG_ERR_INVALID;
    if (!bytes) return FLG_OK;
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FLG_OK;
  });
}

int flg_memcpy_dtoh(flg_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  return host_guard(ctx, "flg_memcpy_dtoh", [&]() -> int {
    if (!ctx || (!d_src && bytes) || (!h_dst && bytes)) return FLG_ERR_INVALID;
    if (!bytes) return FLG_OK;
    FLG_CHECK_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    FLG_CHECK_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return FLG_OK;
  });
}

#ifndef FLG_NO_TEST_HOOKS
int flg_debug_eval(flg_ctx* ctx, int32_t kind, const void* d_in, void* d_out, int64_t n) {
  return host_guard(ctx, "flg_debug_eval", [&]() -> int {
    if (!ctx || kind < 0 || kind > 4 || n < 0 || (n && (!d_in || !d_out)))
      return set_error(ctx, FLG_ERR_INVALID, "flg_debug_eval: bad arguments");
    if (n == 0) return FLG_OK;
    HaltonTables tab;
    int st = get_halton_tables(ctx, &tab);
    if (st) return st;
    debug_eval_kernel<<<div_up(n, 256), 256, 0, ctx->stream>>>(kind, d_in, d_out, n, tab);
    FLG_CHECK_LAUNCH(ctx);
    return FLG_OK;
  });
}

#endif

}  // extern "C"
