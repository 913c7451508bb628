// capi_util.hpp -- helpers shared by the extern "C" translation units: exception to
// status conversion, device selection and host/device staging.
#pragma once

#include <hip/hip_runtime.h>

#include <new>

#include "nwg_internal.hpp"

namespace nwg {

template <class F>
int guarded(nwg_ctx* ctx, F&& f) {
  try {
    f();
    if (ctx) ctx->last_error.clear();
    return NWG_OK;
  } catch (const Error& e) {
    if (ctx) ctx->last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    if (ctx) ctx->last_error = "host allocation failed";
    return NWG_ENOMEM;
  } catch (const std::exception& e) {
    if (ctx) ctx->last_error = e.what();
    return NWG_EHIP;
  } catch (...) {
    if (ctx) ctx->last_error = "unknown failure";
    return NWG_EHIP;
  }
}

struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) NWG_HIP(hipSetDevice(dev));
  }
  ~DeviceGuard() {
    int cur = -1;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

inline void check_graph(nwg_ctx* ctx, const nwg_graph* g) {
  NWG_REQUIRE(g != nullptr, "graph is null");
  NWG_REQUIRE(g->device == ctx->device, "graph belongs to another device");
}

inline void check_mem_kind(int k) {
  NWG_REQUIRE(k == NWG_MEM_HOST || k == NWG_MEM_DEVICE, "mem_kind must be 0 or 1");
}

// copy `count` elements in (host or device source) to a device buffer
template <class T>
void stage_in(hipStream_t s, DevBuf<T>& dst, const T* src, size_t count, int mem_kind) {
  dst.alloc(count ? count : 1);
  if (count)
    NWG_HIP(hipMemcpyAsync(dst.p, src, count * sizeof(T),
                           mem_kind == NWG_MEM_DEVICE ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, s));
}

// A device view of a caller buffer: the buffer itself for device memory, else a staging
// copy (copied in when `in`, copied out by finish()).
template <class T>
struct Staged {
  T* user;
  size_t count;
  int kind;
  DevBuf<T> tmp;
  T* dev = nullptr;
  Staged(hipStream_t s, T* p, size_t c, int k, bool in) : user(p), count(c), kind(k) {
    if (!p) return;
    if (k == NWG_MEM_DEVICE) {
      dev = p;
    } else {
      tmp.alloc(c ? c : 1);
      dev = tmp.p;
      if (in && c) NWG_HIP(hipMemcpyAsync(tmp.p, p, c * sizeof(T), hipMemcpyHostToDevice, s));
    }
  }
  void finish(hipStream_t s) {
    if (user && kind == NWG_MEM_HOST && count) {
      NWG_HIP(hipMemcpyAsync(user, tmp.p, count * sizeof(T), hipMemcpyDeviceToHost, s));
      NWG_HIP(hipStreamSynchronize(s));
    }
  }
};

}  // namespace nwg
