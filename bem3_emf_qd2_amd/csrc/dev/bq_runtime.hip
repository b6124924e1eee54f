// bq_runtime.hip -- library runtime (errors, device, stream) and the batch entry points for
// the special functions: Faddeeva w(z) and the quasi-periodic Green's function.
#include <cstdarg>
#include <cstring>
#include <mutex>
#include <algorithm>
#include "bq_runtime.hpp"
#include "bq_qpgf.hpp"
#include "../../../include/bq_amd.h"

namespace bq {

static thread_local char g_err[1024];
static std::mutex g_mu;
static hipStream_t g_stream[64];
static int g_ready[64];

void set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  if (getenv("BQ_VERBOSE")) fprintf(stderr, "bem3qd2: %s\n", g_err);
}

int ensure_device()
{
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error("no HIP device available (%s)", hipGetErrorString(e));
    return -1;
  }
  if (dev < 0 || dev >= 64) {
    set_error("device index %d out of range", dev);
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ready[dev]) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
      set_error("hipGetDeviceProperties failed");
      return -1;
    }
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
      set_error("device %d is %s; this library is built for gfx950 only", dev, p.gcnArchName);
      return -1;
    }
    if (hipStreamCreateWithFlags(&g_stream[dev], hipStreamNonBlocking) != hipSuccess) {
      set_error("stream creation failed");
      return -1;
    }
    g_ready[dev] = 1;
  }
  return 0;
}

hipStream_t stream()
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  return g_stream[dev];
}

// ---- kernel timing ----------------------------------------------------------------
// Each timed launch gets a ticket; its start/end events are resolved into per-name totals
// once both are recorded.  Failed event calls are cleared from HIP's last-error state so
// that they never surface in a later hipGetLastError() check of the library.
namespace {
constexpr int kMaxNames = 64, kMaxPending = 4096;
struct KStat {
  char name[48];
  double ms;
  long long n;
};
struct KPending {
  int stat;
  long long ticket;
  bool ended;
  hipEvent_t a, b;
};
std::mutex t_mu;
int t_on = 0, t_nstat = 0, t_npend = 0;
long long t_ticket = 0;
KStat t_stat[kMaxNames];
KPending t_pend[kMaxPending];

// resolve every pending pair whose end was recorded (t_mu held)
void t_flush()
{
  int keep = 0;
  for (int i = 0; i < t_npend; ++i) {
    KPending& p = t_pend[i];
    if (!p.ended) {
      t_pend[keep++] = p;
      continue;
    }
    float ms = 0.f;
    if (p.stat >= 0 && hipEventSynchronize(p.b) == hipSuccess && hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
      t_stat[p.stat].ms += ms;
      t_stat[p.stat].n += 1;
    }
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
  }
  t_npend = keep;
  (void)hipGetLastError();
}
}  // namespace

int ktime_begin(const char* name, hipStream_t s)
{
  std::lock_guard<std::mutex> lk(t_mu);
  if (!t_on) return -1;
  if (t_npend == kMaxPending) t_flush();
  if (t_npend == kMaxPending) return -1;
  int k = 0;
  while (k < t_nstat && strncmp(t_stat[k].name, name, sizeof t_stat[k].name - 1) != 0) ++k;
  if (k == t_nstat) {
    if (t_nstat == kMaxNames) return -1;
    snprintf(t_stat[k].name, sizeof t_stat[k].name, "%s", name);
    t_stat[k].ms = 0.0;
    t_stat[k].n = 0;
    ++t_nstat;
  }
  KPending& p = t_pend[t_npend];
  if (hipEventCreate(&p.a) != hipSuccess) {
    (void)hipGetLastError();
    return -1;
  }
  if (hipEventCreate(&p.b) != hipSuccess) {
    (void)hipEventDestroy(p.a);
    (void)hipGetLastError();
    return -1;
  }
  p.stat = k;
  p.ended = false;
  p.ticket = t_ticket++;
  if (hipEventRecord(p.a, s) != hipSuccess) {
    (void)hipEventDestroy(p.a);
    (void)hipEventDestroy(p.b);
    (void)hipGetLastError();
    return -1;
  }
  ++t_npend;
  return (int)(p.ticket & 0x7fffffff);
}

void ktime_end(int ticket, hipStream_t s)
{
  if (ticket < 0) return;
  std::lock_guard<std::mutex> lk(t_mu);
  for (int i = t_npend - 1; i >= 0; --i) {
    KPending& p = t_pend[i];
    if ((int)(p.ticket & 0x7fffffff) == ticket && !p.ended) {
      if (hipEventRecord(p.b, s) == hipSuccess) p.ended = true;
      else (void)hipGetLastError();
      return;
    }
  }
}

}  // namespace bq

// ------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_faddeeva(int n, const double* __restrict__ z, double* __restrict__ w)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx v = bq::faddeeva_w(z[2 * i], z[2 * i + 1]);
  w[2 * i] = v.re;
  w[2 * i + 1] = v.im;
}

__global__ void __launch_bounds__(256) k_erfcx(int n, const double* __restrict__ x, double* __restrict__ out)
{
  __shared__ double tab[bq::kErfcxTab];
  bq::erfcx_stage(tab);
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = bq::erfcx_any(x[i], tab);
}

__global__ void __launch_bounds__(256) k_qpgf_v1(bq_qplan P, int z0_ref, int n, const double* __restrict__ r,
                                                 double* __restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bq::QG q;
  bq::qpgf_v1(P, r[3 * i], r[3 * i + 1], r[3 * i + 2], q, z0_ref != 0);
  double* o = out + 8 * i;
  o[0] = q.g.re; o[1] = q.g.im;
  for (int c = 0; c < 3; ++c) { o[2 + 2 * c] = q.d[c].re; o[3 + 2 * c] = q.d[c].im; }
}

__global__ void __launch_bounds__(256) k_qpgf_v2(bq_qplan P, int n, const double* __restrict__ r,
                                                 double* __restrict__ out)
{
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bq::QG2 q;
  bq::qpgf_v2(P, r[3 * i], r[3 * i + 1], r[3 * i + 2], q);
  double* o = out + 20 * i;
  o[0] = q.g.re; o[1] = q.g.im;
  for (int c = 0; c < 3; ++c) { o[2 + 2 * c] = q.d[c].re; o[3 + 2 * c] = q.d[c].im; }
  for (int c = 0; c < 6; ++c) { o[8 + 2 * c] = q.h[c].re; o[9 + 2 * c] = q.h[c].im; }
}

// ---- allocation-free QPGF batch ------------------------------------------------------
// One error word per chunk of kNoAllocChunk points in a file-scope device array (each device
// holds its own copy); its device address is looked up once per device and used for every
// launch and copy.  Calls that overlap on one device share the array and race.
constexpr int kNoAllocChunk = 1 << 20, kNoAllocWords = 64;
static __device__ __align__(256) int scratch[kNoAllocWords];
static int* g_scratch_ptr[64];

__global__ void k_fill_int(int* __restrict__ p, int n, int v)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

extern "C" {

const char* bq_last_error(void) { return bq::g_err; }

int bq_device_ok(void) { return bq::ensure_device(); }

int bq_faddeeva_w_batch(int n, const double* z, double* w)
{
  if (n < 0 || (n > 0 && (!z || !w))) { bq::set_error("bq_faddeeva_w_batch: bad arguments"); return -2; }
  if (n == 0) return 0;
  if (bq::ensure_device()) return -1;
  hipStream_t s = bq::stream();
  double *dz = nullptr, *dw = nullptr;
  size_t bytes = sizeof(double) * 2 * (size_t)n;
  BQ_HIP(hipMallocAsync((void**)&dz, bytes, s));
  BQ_HIP(hipMallocAsync((void**)&dw, bytes, s));
  BQ_HIP(hipMemcpyAsync(dz, z, bytes, hipMemcpyHostToDevice, s));
  k_faddeeva<<<(n + 255) / 256, 256, 0, s>>>(n, dz, dw);
  BQ_HIP(hipGetLastError());
  BQ_HIP(hipMemcpyAsync(w, dw, bytes, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dz, s));
  BQ_HIP(hipFreeAsync(dw, s));
  BQ_HIP(hipStreamSynchronize(s));
  return 0;
}

int bq_erfcx_batch(int n, const double* x, double* out)
{
  if (n < 0 || (n > 0 && (!x || !out))) { bq::set_error("bq_erfcx_batch: bad arguments"); return -2; }
  if (n == 0) return 0;
  if (bq::ensure_device()) return -1;
  hipStream_t s = bq::stream();
  double *dx = nullptr, *dy = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n;
  BQ_HIP(hipMallocAsync((void**)&dx, bytes, s));
  BQ_HIP(hipMallocAsync((void**)&dy, bytes, s));
  BQ_HIP(hipMemcpyAsync(dx, x, bytes, hipMemcpyHostToDevice, s));
  k_erfcx<<<(n + 255) / 256, 256, 0, s>>>(n, dx, dy);
  BQ_HIP(hipGetLastError());
  BQ_HIP(hipMemcpyAsync(out, dy, bytes, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dx, s));
  BQ_HIP(hipFreeAsync(dy, s));
  BQ_HIP(hipStreamSynchronize(s));
  return 0;
}

int bq_qpgf_plan(const double* qd8, double eps, bq_qplan* plan)
{
  if (!qd8 || !plan) return -2;
  return bq_qplan_init(plan, qd8[0], qd8 + 1, qd8 + 4, qd8 + 6, eps);
}

int bq_qpgf_v1_dev(const bq_qplan* plan, int variant, int n, const double* d_r, double* d_out, hipStream_t s)
{
  if (n <= 0) return 0;
  if (variant == 2)
    k_qpgf_v2<<<(n + 255) / 256, 256, 0, s>>>(*plan, n, d_r, d_out);
  else
    k_qpgf_v1<<<(n + 255) / 256, 256, 0, s>>>(*plan, variant == 0, n, d_r, d_out);
  BQ_HIP(hipGetLastError());
  return 0;
}

int bq_qpgf_batch(const double* qd8, int variant, int n, const double* r, double* out)
{
  if (n < 0 || !qd8 || variant < 0 || variant > 2 || (n > 0 && (!r || !out))) {
    bq::set_error("bq_qpgf_batch: bad arguments");
    return -2;
  }
  bq_qplan P;
  if (bq_qpgf_plan(qd8, 1e-15, &P)) { bq::set_error("bq_qpgf_batch: degenerate lattice"); return -3; }
  if (n == 0) return 0;
  if (bq::ensure_device()) return -1;
  hipStream_t s = bq::stream();
  double *dr = nullptr, *dout = nullptr;
  int* derr = nullptr;
  BQ_HIP(hipMallocAsync((void**)&dr, sizeof(double) * 3 * (size_t)n, s));
  const size_t nv = variant == 2 ? 20 : 8;
  BQ_HIP(hipMallocAsync((void**)&dout, sizeof(double) * nv * (size_t)n, s));
  BQ_HIP(hipMallocAsync((void**)&derr, sizeof(int), s));
  BQ_HIP(hipMemsetAsync(derr, 0, sizeof(int), s));
  P.err = derr;
  BQ_HIP(hipMemcpyAsync(dr, r, sizeof(double) * 3 * (size_t)n, hipMemcpyHostToDevice, s));
  if (bq_qpgf_v1_dev(&P, variant, n, dr, dout, s)) return -1;
  int err = 0;
  BQ_HIP(hipMemcpyAsync(out, dout, sizeof(double) * nv * (size_t)n, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipMemcpyAsync(&err, derr, sizeof(int), hipMemcpyDeviceToHost, s));
  BQ_HIP(hipFreeAsync(dr, s));
  BQ_HIP(hipFreeAsync(dout, s));
  BQ_HIP(hipFreeAsync(derr, s));
  BQ_HIP(hipStreamSynchronize(s));
  if (err) {
    bq::set_error("quasi-periodic Green's function: a reciprocal order has beta_n == 0 (code %d)", err);
    return err;
  }
  return 0;
}

int bq_qpgf_batch_noalloc(const double* qd8, int variant, int n, const double* d_r, double* d_out, void* stream)
{
  if (n < 0 || !qd8 || variant < 0 || variant > 2 || (n > 0 && (!d_r || !d_out))) {
    bq::set_error("bq_qpgf_batch_noalloc: bad arguments");
    return -2;
  }
  if ((long)n > (long)kNoAllocChunk * kNoAllocWords) {
    bq::set_error("bq_qpgf_batch_noalloc: %d points exceed the maximum of %ld (%d chunks of %d)", n,
                  (long)kNoAllocChunk * kNoAllocWords, kNoAllocWords, kNoAllocChunk);
    return -2;
  }
  bq_qplan P;
  if (bq_qpgf_plan(qd8, 1e-15, &P)) {
    bq::set_error("bq_qpgf_batch_noalloc: degenerate lattice");
    return -3;
  }
  if (n == 0) return 0;
  if (bq::ensure_device()) return -1;
  int dev = 0;
  BQ_HIP(hipGetDevice(&dev));
  {
    std::lock_guard<std::mutex> lk(bq::g_mu);
    if (!g_scratch_ptr[dev]) {
      void* p = nullptr;
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) {
        bq::set_error("bq_qpgf_batch_noalloc: hipGetSymbolAddress failed (%s)", hipGetErrorString(e));
        return -1;
      }
      g_scratch_ptr[dev] = (int*)p;
    }
  }
  int* ptr = g_scratch_ptr[dev];
  hipStream_t s = (hipStream_t)stream;
  const int nch = (int)(((long)n + kNoAllocChunk - 1) / kNoAllocChunk);
  const size_t nv = variant == 2 ? 20 : 8;
  k_fill_int<<<1, kNoAllocWords, 0, s>>>(ptr, nch, 0);
  BQ_HIP(hipGetLastError());
  for (int c = 0; c < nch; ++c) {
    const long p0 = (long)c * kNoAllocChunk;
    const int m = (int)std::min((long)kNoAllocChunk, (long)n - p0);
    P.err = ptr + c;
    if (bq_qpgf_v1_dev(&P, variant, m, d_r + 3 * p0, d_out + nv * p0, s)) return -1;
  }
  int h[kNoAllocWords];
  BQ_HIP(hipMemcpyAsync(h, ptr, sizeof(int) * nch, hipMemcpyDeviceToHost, s));
  BQ_HIP(hipStreamSynchronize(s));
  int err = 0;
  for (int c = 0; c < nch; ++c) err = std::min(err, h[c]);
  if (err) {
    bq::set_error("quasi-periodic Green's function: a reciprocal order has beta_n == 0 (code %d)", err);
    return err;
  }
  return 0;
}

int bq_qpgf_v1_batch(const double* qd8, int n, const double* r, double* out)
{
  return bq_qpgf_batch(qd8, 0, n, r, out);
}

int bq_kernel_timing(int on)
{
  std::lock_guard<std::mutex> lk(bq::t_mu);
  bq::t_flush();
  bq::t_on = on ? 1 : 0;
  bq::t_nstat = 0;
  // pairs still open stay pending but no longer count
  for (int i = 0; i < bq::t_npend; ++i) bq::t_pend[i].stat = -1;
  return 0;
}

int bq_kernel_times(int max, char* names, double* total_ms, long long* count)
{
  std::lock_guard<std::mutex> lk(bq::t_mu);
  bq::t_flush();
  int n = bq::t_nstat < max ? bq::t_nstat : max;
  for (int i = 0; i < n; ++i) {
    if (names) snprintf(names + 48 * i, 48, "%s", bq::t_stat[i].name);
    if (total_ms) total_ms[i] = bq::t_stat[i].ms;
    if (count) count[i] = bq::t_stat[i].n;
  }
  return n;
}

}  // extern "C"
