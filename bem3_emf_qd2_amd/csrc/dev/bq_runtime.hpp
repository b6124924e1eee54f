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

}  // namespace bq

#define BQ_HIP(call)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (call);                                                            \
    if (e_ != hipSuccess) {                                                            \
      bq::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
      return -1;                                                                       \
    }                                                                                  \
  } while (0)
