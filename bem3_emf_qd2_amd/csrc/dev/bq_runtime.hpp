// bq_runtime.hpp -- device-side runtime helpers shared by the kernel translation units:
// error reporting that fails loudly, one library stream per device, scratch buffers.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>

namespace bq {

void set_error(const char* fmt, ...);
hipStream_t stream();     // library stream on the current device (created on first use)
int ensure_device();      // 0 if a gfx950-capable device is usable, else -1 (error recorded)

// Per-kernel timing with HIP events recorded on the launch stream around each launch;
// off unless bq_kernel_timing(1) was called.  ktime_begin returns a slot (or -1 when off)
// that ktime_end closes; durations are resolved when bq_kernel_times() is read.
int ktime_begin(const char* name, hipStream_t s);
void ktime_end(int slot, hipStream_t s);

}  // namespace bq

// time one launch (or a library call issuing launches) on stream s under a name
#define BQ_KTIME(name, s, ...)                       \
  do {                                               \
    int kt_slot_ = bq::ktime_begin((name), (s));     \
    __VA_ARGS__;                                     \
    bq::ktime_end(kt_slot_, (s));                    \
  } while (0)

#define BQ_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      bq::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)
