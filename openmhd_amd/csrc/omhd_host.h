// omhd_host.h -- host-side helpers shared by the libomhd translation units:
// error recording, HIP call checking, host/device pointer staging.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/omhd.h"

namespace omhd {

omhd_error_info &tls_error();

inline int set_error(int code, const char *msg, int i = 0, int j = 0, int k = 0, double value = 0.0) {
  omhd_error_info &e = tls_error();
  e.code = code;
  e.i = i;
  e.j = j;
  e.k = k;
  e.value = value;
  snprintf(e.msg, sizeof(e.msg), "%s", msg ? msg : "");
  return code;
}

inline int ok() {
  tls_error().code = OMHD_OK;
  tls_error().msg[0] = 0;
  return OMHD_OK;
}

#define OMHD_HIP(call)                                                                 \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      char b_[160];                                                                    \
      snprintf(b_, sizeof b_, "%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return ::omhd::set_error(OMHD_ERR_HIP, b_);                                      \
    }                                                                                  \
  } while (0)

inline bool is_device_ptr(const void *p) {
  if (!p) return false;
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

// Device mirrors of host arrays (omhd_mirror_mode).  When mirroring is on, a
// host array handed to a drop-in entry point keeps its device copy from call
// to call instead of crossing the bus each time, and its host copy is brought
// up to date only by omhd_mirror_sync (which the fileio shims call).
bool mirror_on();
// device address of host [p, p + count) in its mirror, made or grown as
// needed (host contents are uploaded where no mirror holds the range);
// `writes`: the call writes into the range, so the host copy goes stale
int mirror_acquire(const double *p, size_t count, bool writes, double **dev);
// bring host [p, p + count) up to date from the mirrors that hold it
int mirror_sync(const double *p, size_t count);

// A device view of a caller array: used in place when it is device memory;
// for host memory, the array's mirror when mirroring is on, otherwise a
// device copy (uploaded if `in`, downloaded at finish() if `out`).  A copy is
// freed, stream-ordered, by release() or when the view goes out of scope, so
// an entry point that returns early on an error frees it too.
struct Staged {
  double *dev = nullptr;
  double *host = nullptr;
  size_t n = 0;
  bool owned = false, out = false;
  hipStream_t stream = nullptr;
  Staged() = default;
  Staged(const Staged &) = delete;
  Staged &operator=(const Staged &) = delete;
  ~Staged() { release(stream); }
  int init(const double *p, size_t count, bool in, bool out_, hipStream_t s) {
    n = count;
    out = out_;
    stream = s;
    if (is_device_ptr(p)) {
      dev = const_cast<double *>(p);
      return OMHD_OK;
    }
    host = const_cast<double *>(p);
    if (mirror_on()) return mirror_acquire(p, count, out_, &dev);  // not owned: the mirror outlives the call
    OMHD_HIP(hipMallocAsync((void **)&dev, n * sizeof(double), s));
    owned = true;
    if (in) OMHD_HIP(hipMemcpyAsync(dev, host, n * sizeof(double), hipMemcpyHostToDevice, s));
    return OMHD_OK;
  }
  int finish(hipStream_t s) {
    if (owned && out) OMHD_HIP(hipMemcpyAsync(host, dev, n * sizeof(double), hipMemcpyDeviceToHost, s));
    return OMHD_OK;
  }
  void release(hipStream_t s) {
    if (owned && dev) (void)hipFreeAsync(dev, s);
    dev = nullptr;
    owned = false;
  }
};

hipStream_t dropin_stream();

}  // namespace omhd
