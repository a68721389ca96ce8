// Shared helpers of the gfx950 DMRG kernels: error reporting, dtype access,
// XCD-aware block remapping and wavefront reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>

#include "../../include/tnalg_hip.h"

#define TNX_VERSION 10000

// thread-local message of the last failure (tnx_last_error)
void tnx_set_error(const char* fmt, ...);

// sampled kernel timing (prof.hip): a Lanczos call asks whether it is
// sampled, switches timing on for its first Krylov step; launch sites bracket
// their kernel with mark/done when timing is on.
enum {
  TNX_PC_GEMM_F64_1 = 0, TNX_PC_GEMM_F64_2, TNX_PC_GEMM_F32_1, TNX_PC_GEMM_F32_2,
  TNX_PC_GEMM_BF16_1, TNX_PC_GEMM_BF16_2, TNX_PC_MIX, TNX_PC_DOTS, TNX_PC_UPDATE_DOTS,
  TNX_PC_UPDATE_NORM
};
bool tnx_prof_call_begin();      // this call is sampled (timing stays off)
void tnx_prof_set(bool on);      // timing on / off for the calling thread
bool tnx_prof_on();
int tnx_prof_mark(int cls, hipStream_t s);                // -1: not timed
unsigned long long* tnx_prof_ts(int mark);                // the launch's stamp slot (or null)
void tnx_prof_done(int mark, hipStream_t s, double work);
void tnx_prof_count(int cls, double work);   // one launch of the class and its work (while armed)

// Launch-span stamps of a timed launch: the slot holds TNX_TS_PAIRS pairs
// (earliest block start, latest block end) of the device's constant wall
// clock; block b updates pair b % TNX_TS_PAIRS with vector atomics (spread
// over pairs, so no address takes more than a few dozen), and the host
// takes the minimum start and maximum end.  A null slot: not timed.
constexpr int TNX_TS_PAIRS = 64;
__device__ __forceinline__ void tnx_stamp_begin(unsigned long long* ts) {
  if (ts != nullptr && threadIdx.x == 0)
    atomicMin(ts + 2 * (blockIdx.x % TNX_TS_PAIRS), (unsigned long long)wall_clock64());
}
__device__ __forceinline__ void tnx_stamp_end(unsigned long long* ts) {
  if (ts != nullptr && threadIdx.x == 0)
    atomicMax(ts + 2 * (blockIdx.x % TNX_TS_PAIRS) + 1, (unsigned long long)wall_clock64());
}
// Grouped GEMM launch with a stamp slot (gemm.hip); tnx_gemm_grouped is this
// with a null slot.
int tnx_gemm_grouped_ts(int mode, const tnx_gemm_problem* prob, int nprob,
                        const tnx_gemm_seg* seg, int ntiles, const int32_t* active,
                        tnx_stream_t stream, unsigned long long* ts);

#define TNX_CHECK_HIP(call)                                                      \
  do {                                                                           \
    hipError_t e_ = (call);                                                      \
    if (e_ != hipSuccess) {                                                      \
      tnx_set_error("%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),       \
                    __FILE__, __LINE__);                                         \
      return TNX_EHIP;                                                           \
    }                                                                            \
  } while (0)

#define TNX_CHECK_LAUNCH()                                                       \
  do {                                                                           \
    hipError_t e_ = hipGetLastError();                                           \
    if (e_ != hipSuccess) {                                                      \
      tnx_set_error("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_),   \
                    __FILE__, __LINE__);                                         \
      return TNX_EHIP;                                                           \
    }                                                                            \
  } while (0)

typedef double tnx_d4 __attribute__((ext_vector_type(4)));
typedef float tnx_f4 __attribute__((ext_vector_type(4)));
typedef __bf16 tnx_bf8 __attribute__((ext_vector_type(8)));

// Load element idx of a buffer of runtime dtype as double.
__device__ __forceinline__ double tnx_load(const void* p, int dtype, int64_t idx) {
  if (dtype == TNX_F64) return ((const double*)p)[idx];
  if (dtype == TNX_F32) return (double)((const float*)p)[idx];
  return (double)(float)((const __bf16*)p)[idx];
}

__device__ __forceinline__ void tnx_store(void* p, int dtype, int64_t idx, double v) {
  if (dtype == TNX_F64) ((double*)p)[idx] = v;
  else if (dtype == TNX_F32) ((float*)p)[idx] = (float)v;
  else ((__bf16*)p)[idx] = (__bf16)(float)v;
}

// Workgroups are dispatched round-robin over the 8 XCDs of MI355X, each with
// its own L2.  Remap the hardware block id so that consecutive logical tiles
// (which share operands) land on the same XCD.
__device__ __forceinline__ int tnx_xcd_remap(int bid, int nblocks) {
  const int nx = 8;
  if (nblocks < nx) return bid;
  int x = bid % nx, i = bid / nx;
  int q = nblocks / nx, r = nblocks % nx;
  return x * q + (x < r ? x : r) + i;
}

__device__ __forceinline__ double tnx_wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Largest index i in [0, n) whose 32-bit field at byte offset `off` of the
// i-th record (records of `rec` bytes) is <= key.  The fields are prefix
// sums (non-decreasing, first one 0, key >= 0), so i is the record owning
// block `key`.  Called by every lane of a full wavefront with uniform
// arguments: each round the 64 lanes probe 64 evenly spaced records of the
// remaining range [lo, hi] (lane 0 probes lo, known to qualify) and the range
// shrinks to the gap after the last qualifying probe, so up to 4096 records
// cost two dependent loads instead of twelve.
__device__ __forceinline__ int tnx_find_owner(const void* base, int rec, int off, int n,
                                              int key) {
  const char* b = (const char*)base + off;
  const int lane = threadIdx.x & 63;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int step = (hi - lo + 64) >> 6;  // ceil((hi - lo + 1) / 64)
    const int idx = lo + lane * step;
    bool ok = false;
    if (idx <= hi) ok = *(const int32_t*)(b + (int64_t)idx * rec) <= key;
    const unsigned long long m = __ballot(ok);
    lo += (63 - __clzll((long long)m)) * step;
    hi = min(hi, lo + step - 1);
  }
  return lo;
}
