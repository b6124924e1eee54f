// bq_mfma.hip -- complex single-precision GEMM update C -= A B on the bf16 matrix cores, for
// the trailing updates and many-right-hand-side solves of the FP32 LU (bq_lu.hip).
//
// Each FP32 operand x is split into two bf16 numbers, x = hi + lo + r with hi = bf16(x),
// lo = bf16(x - hi) and |r| <= 2^-17 |x|.  A real product is taken as
//   x y ~ hi_x hi_y + hi_x lo_y + lo_x hi_y       (three bf16 MFMA products, FP32 accumulation)
// so each term carries a relative error of about 2^-16; the complex product takes four real
// products.  This is Henry, Tang and Heinecke's bf16x3 scheme ("Leveraging the bfloat16
// artificial intelligence datatype for higher-precision computations", ARITH 2019) applied
// to complex operands.  The FP32 factors it produces are a preconditioner of the FP64 FGMRES
// solve, which restores the FP64 backward error; MI355X's dense bf16 rate (2.5 PFLOP/s) is
// 16x its FP32 matrix rate, so three products per term still leave ~5x.
//
// Data flow per call:
//   k_pack_a / k_pack_b  write A and B once as operand tiles: for every 32 x 16 (complex)
//                        block, four planes (re hi, re lo, im hi, im lo) of 64 lanes x 8 bf16,
//                        in exactly the order the MFMA reads them (one 16-byte load per lane);
//                        the block counts are padded to multiples of 4 with zero tiles.
//   k_cgemm_split_d      one workgroup (4 waves) per 128 x 128 tile of C; per 16-wide k step
//                        the workgroup's A and B tiles (16 KB each) are loaded straight into LDS
//                        (global_load_lds_dwordx4), double-buffered with one barrier per step; each wave
//                        owns 64 x 64 of C as 2 x 2 blocks of v_mfma_f32_32x32x16_bf16 with
//                        separate real and imaginary accumulators (128 floats per lane), and
//                        subtracts them from C at the end (B's tile goes into the MFMA's
//                        first operand, so that each block comes out with C's rows along
//                        the lanes and the read-modify-write of C is coalesced).
// Workgroups are dispatched round-robin over the 8 XCDs; the tile index is remapped so that
// each XCD takes a contiguous run of tiles, walked in groups of 8 tile rows, and the A and B
// tiles in flight on one XCD are shared through its L2.
#include <hip/hip_runtime.h>
#include <mutex>
#include <vector>
#include "bq_runtime.hpp"
#include "../bq_linalg.h"

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kTile = 128;           // rows and columns of C per workgroup
constexpr int kBlk = 32;             // MFMA block edge
constexpr int kKStep = 16;           // complex k per stage (the MFMA's k)
constexpr int kTileU4 = 64;          // one 32 x 16 plane = 64 lanes x 16 bytes
constexpr int kStageU4 = 4 * 4 * kTileU4;  // 4 blocks x 4 planes = 16 KB
constexpr int kGroupRows = 8;        // tile rows per swizzle group

__device__ __forceinline__ unsigned bf16_rne(float x)
{
  const unsigned u = __float_as_uint(x);
  return (u + 0x7fffu + ((u >> 16) & 1u)) >> 16;
}

// hi and lo halves of x (x = hi + lo + r, |r| <= 2^-17 |x|)
__device__ __forceinline__ void split(float x, unsigned& hi, unsigned& lo)
{
  hi = bf16_rne(x);
  lo = bf16_rne(x - __uint_as_float(hi << 16));
}

// 8 values (4 planes x 8 bf16) of one lane -> 4 uint4 at dst[p * kTileU4]
__device__ __forceinline__ void store_planes(const float2 (&v)[8], uint4* dst)
{
  unsigned w[4][4];
#pragma unroll
  for (int e = 0; e < 8; e += 2) {
    unsigned rh0, rl0, ih0, il0, rh1, rl1, ih1, il1;
    split(v[e].x, rh0, rl0);
    split(v[e].y, ih0, il0);
    split(v[e + 1].x, rh1, rl1);
    split(v[e + 1].y, ih1, il1);
    w[0][e / 2] = rh0 | (rh1 << 16);
    w[1][e / 2] = rl0 | (rl1 << 16);
    w[2][e / 2] = ih0 | (ih1 << 16);
    w[3][e / 2] = il0 | (il1 << 16);
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) dst[p * kTileU4] = make_uint4(w[p][0], w[p][1], w[p][2], w[p][3]);
}

// A: m x k (column-major, lda).  Tile (kb, rb) lane l holds row rb*32 + l%32 and k = kb*16 +
// 8*(l/32) + 0..7; it sits at Ap[((kb * Mb + rb) * 4 + plane) * 64 + l].
__global__ void k_pack_a(int m, int k, const float2* __restrict__ A, long lda, int Mb, int Kb, uint4* __restrict__ Ap)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)Kb * Mb * 64) return;
  const int l = (int)(t & 63);
  const long q = t >> 6;
  const int rb = (int)(q % Mb), kb = (int)(q / Mb);
  const int row = rb * kBlk + (l & 31), k0 = kb * kKStep + 8 * (l >> 5);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e)
    v[e] = (row < m && k0 + e < k) ? A[(long)(k0 + e) * lda + row] : make_float2(0.f, 0.f);
  store_planes(v, Ap + ((long)(kb * Mb + rb) * 4) * kTileU4 + l);
}

// B: k x n (column-major, ldb).  Tile (kb, cb) lane l holds column cb*32 + l%32 and the same
// k as A's lanes; it sits at Bp[((kb * Nb + cb) * 4 + plane) * 64 + l].
__global__ void k_pack_b(int k, int n, const float2* __restrict__ B, long ldb, int Nb, int Kb, uint4* __restrict__ Bp)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)Kb * Nb * 64) return;
  const int l = (int)(t & 63);
  const long q = t >> 6;
  const int cb = (int)(q % Nb), kb = (int)(q / Nb);
  const int col = cb * kBlk + (l & 31), k0 = kb * kKStep + 8 * (l >> 5);
  float2 v[8];
#pragma unroll
  for (int e = 0; e < 8; ++e)
    v[e] = (col < n && k0 + e < k) ? B[(long)col * ldb + k0 + e] : make_float2(0.f, 0.f);
  store_planes(v, Bp + ((long)(kb * Nb + cb) * 4) * kTileU4 + l);
}

__device__ __forceinline__ bf16x8 as_bf(uint4 u) { return __builtin_bit_cast(bf16x8, u); }
__device__ __forceinline__ bf16x8 neg_bf(uint4 u)
{
  const unsigned s = 0x80008000u;
  return __builtin_bit_cast(bf16x8, make_uint4(u.x ^ s, u.y ^ s, u.z ^ s, u.w ^ s));
}

__device__ __forceinline__ f32x16 mma(bf16x8 a, bf16x8 b, f32x16 c)
{
  return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// ---- C (m x n, ldc) -= A B from the packed operands (STORE: C = -A B, C not read) ---------
// grid = tiles_r * tiles_c workgroups of 256.  global_load_lds_dwordx4 puts each lane's 16 bytes
// at the wavefront's LDS base + 16 x lane, the layout the MFMA reads, so a stage needs no
// staging registers and no LDS stores; the two stage buffers are separate LDS objects and the
// loop is unrolled by two, so that the loads into one buffer do not make the reads of the other
// wait for them.
__device__ __forceinline__ void stage_direct(uint4* la, uint4* lb, const uint4* pa, const uint4* pb, int tid, int w)
{
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    __builtin_amdgcn_global_load_lds((const void*)(pa + tid + 256 * q),
                                     (__attribute__((address_space(3))) void*)(la + 64 * w + 256 * q), 16, 0, 0);
    __builtin_amdgcn_global_load_lds((const void*)(pb + tid + 256 * q),
                                     (__attribute__((address_space(3))) void*)(lb + 64 * w + 256 * q), 16, 0, 0);
  }
}

// the 48 MFMAs of one 16-wide k step on a wave's 64 x 64 part of the tile
__device__ __forceinline__ void stage_mma(const uint4* la, const uint4* lb, int lane, int wr, int wc,
                                          f32x16 (&cr)[2][2], f32x16 (&ci)[2][2])
{
  uint4 b_op[2][4];
#pragma unroll
  for (int j = 0; j < 2; ++j)
#pragma unroll
    for (int p = 0; p < 4; ++p) b_op[j][p] = lb[((2 * wc + j) * 4 + p) * kTileU4 + lane];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    uint4 a_op[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) a_op[p] = la[((2 * wr + i) * 4 + p) * kTileU4 + lane];
    const bf16x8 arh = as_bf(a_op[0]), arl = as_bf(a_op[1]), aih = as_bf(a_op[2]), ail = as_bf(a_op[3]);
    const bf16x8 naih = neg_bf(a_op[2]), nail = neg_bf(a_op[3]);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const bf16x8 brh = as_bf(b_op[j][0]), brl = as_bf(b_op[j][1]), bih = as_bf(b_op[j][2]), bil = as_bf(b_op[j][3]);
      f32x16 r = cr[i][j], q = ci[i][j];
      r = mma(brh, arh, r);
      q = mma(bih, arh, q);
      r = mma(brl, arh, r);
      q = mma(bil, arh, q);
      r = mma(brh, arl, r);
      q = mma(bih, arl, q);
      r = mma(bih, naih, r);
      q = mma(brh, aih, q);
      r = mma(bil, naih, r);
      q = mma(brl, aih, q);
      r = mma(bih, nail, r);
      q = mma(brh, ail, q);
      cr[i][j] = r;
      ci[i][j] = q;
    }
  }
}

template <bool STORE>
__global__ __launch_bounds__(256, 2) void k_cgemm_split_d(int m, int n, int tiles_r, int tiles_c, int Kb,
                                                          const uint4* __restrict__ Ap, const uint4* __restrict__ Bp,
                                                          float2* __restrict__ C, long ldc)
{
  __shared__ uint4 l0a[kStageU4], l0b[kStageU4], l1a[kStageU4], l1b[kStageU4];
  const int ntiles = tiles_r * tiles_c;
  int t = blockIdx.x;
  if ((ntiles & 7) == 0) t = (t & 7) * (ntiles >> 3) + (t >> 3);  // contiguous run per XCD
  const int span = kGroupRows * tiles_c, g = t / span, first = g * kGroupRows;
  const int gsz = min(kGroupRows, tiles_r - first), in = t - g * span;
  const int tr = first + in % gsz, tc = in / gsz;
  const int tid = threadIdx.x, lane = tid & 63, w = __builtin_amdgcn_readfirstlane(tid >> 6), wr = w >> 1, wc = w & 1;
  const int Mb = tiles_r * 4, Nb = tiles_c * 4;
  const uint4* a_src = Ap + (long)(tr * 4) * 4 * kTileU4;
  const uint4* b_src = Bp + (long)(tc * 4) * 4 * kTileU4;
  const long a_step = (long)Mb * 4 * kTileU4, b_step = (long)Nb * 4 * kTileU4;
  f32x16 cr[2][2], ci[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      cr[i][j] = f32x16{};
      ci[i][j] = f32x16{};
    }
  stage_direct(l0a, l0b, a_src, b_src, tid, w);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  for (int kb = 0; kb < Kb; kb += 2) {
    if (kb + 1 < Kb) stage_direct(l1a, l1b, a_src + (long)(kb + 1) * a_step, b_src + (long)(kb + 1) * b_step, tid, w);
    stage_mma(l0a, l0b, lane, wr, wc, cr, ci);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (kb + 1 < Kb) {
      if (kb + 2 < Kb) stage_direct(l0a, l0b, a_src + (long)(kb + 2) * a_step, b_src + (long)(kb + 2) * b_step, tid, w);
      stage_mma(l1a, l1b, lane, wr, wc, cr, ci);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int row = tr * kTile + (2 * wr + i) * kBlk + (lane & 31);
      if (row >= m) continue;
      const int c0 = tc * kTile + (2 * wc + j) * kBlk + 4 * (lane >> 5);
      float2* cr0 = C + row;
      float2 v[16];
#pragma unroll
      for (int e = 0; e < 16; ++e)
        v[e] = STORE ? make_float2(0.f, 0.f) : cr0[(long)min(c0 + 8 * (e >> 2) + (e & 3), n - 1) * ldc];
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int col = c0 + 8 * (e >> 2) + (e & 3);
        v[e].x -= cr[i][j][e];
        v[e].y -= ci[i][j][e];
        if (col < n) cr0[(long)col * ldc] = v[e];
      }
    }
}

// packed-operand workspace per (device, stream), grown on demand and kept for the process
struct Ws {
  int dev;
  hipStream_t s;
  void* p;
  size_t bytes;
};
std::mutex g_wmu;
std::vector<Ws> g_ws;

void* workspace(hipStream_t s, size_t bytes)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_wmu);
  Ws* w = nullptr;
  for (Ws& q : g_ws)
    if (q.dev == dev && q.s == s) w = &q;
  if (!w) {
    g_ws.push_back(Ws{dev, s, nullptr, 0});
    w = &g_ws.back();
  }
  if (w->bytes < bytes) {
    if (w->p) {
      // earlier calls on this stream may still read the old buffer
      if (hipStreamSynchronize(s) != hipSuccess || hipFree(w->p) != hipSuccess) return nullptr;
      w->p = nullptr;
      w->bytes = 0;
    }
    if (hipMalloc(&w->p, bytes) != hipSuccess) {
      w->p = nullptr;
      return nullptr;
    }
    w->bytes = bytes;
  }
  return w->p;
}

}  // namespace

// C -= A B (store = false) or C = -A B (store = true); the operands are packed before the
// product is written, so with store = true B may be C itself (C <- -A C in place).
int split_gemm(hipStream_t s, int m, int n, int k, const float* A, long lda, const float* B, long ldb, float* C,
               long ldc, bool store)
{
  if (m <= 0 || n <= 0) return 0;
  if (k <= 0 && !store) return 0;
  const int tiles_r = (m + kTile - 1) / kTile, tiles_c = (n + kTile - 1) / kTile;
  const int Mb = tiles_r * 4, Nb = tiles_c * 4, Kb = max(1, (k + kKStep - 1) / kKStep);
  const long na = (long)Kb * Mb * 4 * kTileU4, nb = (long)Kb * Nb * 4 * kTileU4;  // uint4 counts
  if ((long)tiles_r * tiles_c > 0x7fffffffL) {
    bq::set_error("bq_cgemm_split_dev: %d x %d tiles exceed one launch", tiles_r, tiles_c);
    return -1;
  }
  uint4* ws = (uint4*)workspace(s, sizeof(uint4) * (size_t)(na + nb));
  if (!ws) {
    bq::set_error("bq_cgemm_split_dev: workspace of %.1f MB not allocated", 16.0 * (na + nb) / 1e6);
    return -1;
  }
  uint4 *Ap = ws, *Bp = ws + na;
  const long ta = (long)Kb * Mb * 64, tb = (long)Kb * Nb * 64;
  k_pack_a<<<(unsigned)((ta + 255) / 256), 256, 0, s>>>(m, k, (const float2*)A, lda, Mb, Kb, Ap);
  k_pack_b<<<(unsigned)((tb + 255) / 256), 256, 0, s>>>(k, n, (const float2*)B, ldb, Nb, Kb, Bp);
  const unsigned grid = (unsigned)(tiles_r * tiles_c);
  if (store)
    BQ_KTIME("k_cgemm_split", s,
             k_cgemm_split_d<true><<<grid, 256, 0, s>>>(m, n, tiles_r, tiles_c, Kb, Ap, Bp, (float2*)C, ldc));
  else
    BQ_KTIME("k_cgemm_split", s,
             k_cgemm_split_d<false><<<grid, 256, 0, s>>>(m, n, tiles_r, tiles_c, Kb, Ap, Bp, (float2*)C, ldc));
  BQ_HIP(hipGetLastError());
  return 0;
}

extern "C" {

// Free the packed-operand workspace kept for this stream on the current device (call before
// the stream is destroyed).
void bq_cgemm_split_release(void* stream)
{
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_wmu);
  for (size_t i = 0; i < g_ws.size(); ++i)
    if (g_ws[i].dev == dev && g_ws[i].s == (hipStream_t)stream) {
      if (g_ws[i].p) {
        (void)hipStreamSynchronize((hipStream_t)stream);
        (void)hipFree(g_ws[i].p);
      }
      g_ws.erase(g_ws.begin() + i);
      return;
    }
}

// C -= A B (complex FP32, column-major): A m x k, B k x n, C m x n.  bf16x3 products on the
// matrix cores with FP32 accumulation (relative error ~2^-16 per product term).
int bq_cgemm_split_dev(void* stream, int m, int n, int k, const float* A, long lda, const float* B, long ldb,
                       float* C, long ldc)
{
  return split_gemm((hipStream_t)stream, m, n, k, A, lda, B, ldb, C, ldc, false);
}

// C = -A B on the same kernel; B may be C itself (C <- -A C in place: the operands are packed
// before the product is written).
int bq_cgemm_split_store_dev(void* stream, int m, int n, int k, const float* A, long lda, const float* B, long ldb,
                             float* C, long ldc)
{
  return split_gemm((hipStream_t)stream, m, n, k, A, lda, B, ldb, C, ldc, true);
}

}  // extern "C"
