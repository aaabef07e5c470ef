// VALU issue-rate probe for gfx950: how many wave-instructions per CU per
// clock each instruction the step kernels lean on can issue, with
// independent work (8 chains per lane) and as a dependent chain (1 chain),
// at 1 to 8 waves per SIMD; and the exact random stream of the draw work
// (xorshift64*, the collision column, one cell after another, as the step
// kernels run it) in cells per CU per clock at the same occupancies.
//
// Every workgroup is 4 waves (256 threads, one per SIMD of a CU) and the grid
// is CUs x waves-per-SIMD workgroups, so each CU holds that many waves per
// SIMD at once. Thread 0 of each workgroup reads the shader clock
// (s_memtime) and the 100 MHz real-time clock (s_memrealtime) around its
// loop; the rate is the grid's work over the span from the first start to
// the last end on the real-time clock (shared by every CU), converted to
// shader clocks at the median workgroup's clock.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "lgca_internal.h"

namespace {

// One instruction of `op` on chain register(s); dependent on the chain's own
// previous value only.
template <int OP>
__device__ __forceinline__ void probe_op(uint32_t& a, uint64_t& x, uint32_t b, uint32_t c) {
  if constexpr (OP == 0) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 1) asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0x96" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == 2) asm volatile("v_lshlrev_b64 %0, 13, %0" : "+v"(x));
  if constexpr (OP == 3) asm volatile("v_alignbit_b32 %0, %0, %1, 19" : "+v"(a) : "v"(b));
  if constexpr (OP == 4) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 5) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 6) asm volatile("v_mad_u64_u32 %0, vcc, %1, %2, %0" : "+v"(x) : "v"(b), "v"(c) : "vcc");
  if constexpr (OP == 7) asm volatile("v_mul_u32_u24 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 8) asm volatile("v_mul_hi_u32_u24 %0, %0, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 9) asm volatile("v_lshl_or_b32 %0, %0, 3, %1" : "+v"(a) : "v"(b));
  if constexpr (OP == 10) asm volatile("v_mov_b32_dpp %0, %0 wave_ror:1 row_mask:0xf bank_mask:0xf" : "+v"(a));
  if constexpr (OP == 11) asm volatile("v_lshrrev_b64 %0, 7, %0" : "+v"(x));
  if constexpr (OP == 12) asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == 13) asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(a) : "v"(b), "v"(c));
  if constexpr (OP == 14) asm volatile("v_lshlrev_b32 %0, 13, %0" : "+v"(a));
}

constexpr int PROBE_OPS = 15;
const char* const kOpNames[PROBE_OPS] = {
    "v_xor_b32",      "v_bitop3_b32",     "v_lshlrev_b64",  "v_alignbit_b32",
    "v_mul_lo_u32",   "v_mul_hi_u32",     "v_mad_u64_u32",  "v_mul_u32_u24",
    "v_mul_hi_u32_u24", "v_lshl_or_b32",  "v_mov_b32_dpp",  "v_lshrrev_b64",
    "v_add3_u32",     "v_mad_u32_u24",    "v_lshlrev_b32"};

// 16 instructions per chain per iteration, chains interleaved.
template <int OP, int CH>
__global__ void __launch_bounds__(256) lg_probe_op_kernel(uint32_t iters, uint32_t seed,
                                                          unsigned long long* __restrict clk,
                                                          uint32_t* __restrict sink) {
  uint32_t a[CH];
  uint64_t x[CH];
  const uint32_t b = seed ^ threadIdx.x, c = seed * 3u + blockIdx.x;
#pragma unroll
  for (int i = 0; i < CH; ++i) {
    a[i] = seed + i * 7919u + threadIdx.x;
    x[i] = ((uint64_t)a[i] << 32) | (a[i] ^ 0x9E3779B9u);
  }
  __syncthreads();
  const unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (uint32_t it = 0; it < iters; ++it) {
#pragma unroll
    for (int u = 0; u < 16; ++u) {
#pragma unroll
      for (int i = 0; i < CH; ++i) probe_op<OP>(a[i], x[i], b, c);
    }
  }
  __syncthreads();
  const unsigned long long t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  uint32_t s = 0;
#pragma unroll
  for (int i = 0; i < CH; ++i) s ^= a[i] ^ (uint32_t)x[i] ^ (uint32_t)(x[i] >> 32);
  sink[blockIdx.x * 256 + threadIdx.x] = s;
  if (threadIdx.x == 0) {
    clk[4 * blockIdx.x] = t1 - t0;
    clk[4 * blockIdx.x + 1] = r1 - r0;
    clk[4 * blockIdx.x + 2] = r0;
    clk[4 * blockIdx.x + 3] = r1;
  }
}

// The draw work's exact stream, one cell after another on one chain per lane
// (lg_draw_half's loop: xorshift64, the high half of x * M, the column k =
// (r * 30) >> 32, one shift-or deposit per cell), iters x 32 cells per lane.
__global__ void __launch_bounds__(256) lg_probe_stream_kernel(uint32_t iters, uint32_t seed,
                                                              unsigned long long* __restrict clk,
                                                              uint32_t* __restrict sink) {
  uint64_t st = lg_rng_init(seed ^ (blockIdx.x * 256u + threadIdx.x));
  uint32_t acc = 0;
  __syncthreads();
  const unsigned long long t0 = __builtin_amdgcn_s_memtime(), r0 = __builtin_amdgcn_s_memrealtime();
  for (uint32_t it = 0; it < iters; ++it) {
    uint32_t K[LG_COLLIDE_REST_KBITS];
    st = lg_draw_half<LG_COLLIDE_REST_KBITS>(st, LG_COLLIDE_REST_STRIDE, K);
#pragma unroll
    for (int bq = 0; bq < LG_COLLIDE_REST_KBITS; ++bq) acc ^= K[bq];
  }
  __syncthreads();
  const unsigned long long t1 = __builtin_amdgcn_s_memtime(), r1 = __builtin_amdgcn_s_memrealtime();
  sink[blockIdx.x * 256 + threadIdx.x] = acc ^ (uint32_t)st;
  if (threadIdx.x == 0) {
    clk[4 * blockIdx.x] = t1 - t0;
    clk[4 * blockIdx.x + 1] = r1 - r0;
    clk[4 * blockIdx.x + 2] = r0;
    clk[4 * blockIdx.x + 3] = r1;
  }
}

template <int OP>
void launch_op(uint32_t chains, dim3 grid, uint32_t iters, uint32_t seed, unsigned long long* clk,
               uint32_t* sink) {
  if (chains == 1)
    hipLaunchKernelGGL((lg_probe_op_kernel<OP, 1>), grid, dim3(256), 0, 0, iters, seed, clk, sink);
  else
    hipLaunchKernelGGL((lg_probe_op_kernel<OP, 8>), grid, dim3(256), 0, 0, iters, seed, clk, sink);
}

template <int... OPS>
bool dispatch_op(int op, uint32_t chains, dim3 grid, uint32_t iters, uint32_t seed,
                 unsigned long long* clk, uint32_t* sink, std::integer_sequence<int, OPS...>) {
  bool hit = false;
  ((op == OPS ? (launch_op<OPS>(chains, grid, iters, seed, clk, sink), hit = true) : false), ...);
  return hit;
}

}  // namespace

using namespace lgca_internal;

extern "C" {

const char* lg_probe_op_name(int op) {
  if (op >= 0 && op < PROBE_OPS) return kOpNames[op];
  if (op == LG_PROBE_STREAM) return "stream (xorshift64* + column, per cell)";
  return nullptr;
}

int lg_probe_valu(int op, uint32_t waves_per_simd, uint32_t chains, uint32_t iters,
                  double* per_cu_cycle, double* clock_ghz) {
  if (!per_cu_cycle || !clock_ghz) return fail(LG_EINVAL, "null argument");
  if (!lg_probe_op_name(op)) return fail(LG_EINVAL, "unknown probe op");
  if (waves_per_simd < 1 || waves_per_simd > 8) return fail(LG_EINVAL, "waves per SIMD must be 1..8");
  if (chains != 1 && chains != 8) return fail(LG_EINVAL, "chains must be 1 or 8");
  if (iters == 0 || iters > (1u << 20)) return fail(LG_EINVAL, "iters must be 1..2^20");
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  hipDeviceProp_t p;
  HIP_TRY(hipGetDeviceProperties(&p, dev));
  const uint32_t ncu = (uint32_t)p.multiProcessorCount;
  const dim3 grid(ncu * waves_per_simd);
  DevBuf<unsigned long long> clk;
  DevBuf<uint32_t> sink;
  HIP_TRY(clk.alloc(4 * (size_t)grid.x));
  HIP_TRY(sink.alloc(256 * (size_t)grid.x));
  for (int rep = 0; rep < 2; ++rep) {  // the first launch warms the clocks and code
    if (op == LG_PROBE_STREAM)
      hipLaunchKernelGGL(lg_probe_stream_kernel, grid, dim3(256), 0, 0, iters, 12345u, clk.p, sink.p);
    else
      dispatch_op(op, chains, grid, iters, 12345u, clk.p, sink.p,
                  std::make_integer_sequence<int, PROBE_OPS>{});
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
  }
  std::vector<unsigned long long> h(4 * (size_t)grid.x);
  HIP_TRY(hipMemcpy(h.data(), clk.p, h.size() * 8, hipMemcpyDeviceToHost));
  std::vector<double> ghz(grid.x);
  unsigned long long rmin = ~0ull, rmax = 0;
  for (uint32_t i = 0; i < grid.x; ++i) {
    ghz[i] = (double)h[4 * i] / ((double)h[4 * i + 1] * 10.0);  // real-time clock: 10 ns ticks
    rmin = std::min(rmin, h[4 * i + 2]);
    rmax = std::max(rmax, h[4 * i + 3]);
  }
  std::sort(ghz.begin(), ghz.end());
  *clock_ghz = ghz[ghz.size() / 2];
  // the whole grid's span on the real-time clock (common to every CU), in
  // shader clocks at the median workgroup's rate
  const double span_cycles = (double)(rmax - rmin) * 10.0 * *clock_ghz;
  const double units = op == LG_PROBE_STREAM ? 64.0 * 32.0 * iters  // cells per wave
                                             : 16.0 * chains * iters;  // instructions per wave
  *per_cu_cycle = units * 4.0 * grid.x / ((double)ncu * span_cycles);
  return LG_OK;
}

}  // extern "C"
