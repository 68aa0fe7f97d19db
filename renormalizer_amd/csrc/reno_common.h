// Shared definitions for the libreno_hip.so translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include <vector>

#include "../../include/reno_hip.h"

namespace reno {

// Per-thread text of the last runtime error, read by reno_last_error().
void set_error(const char* fmt, ...);

#define RENO_HIP_CHECK(expr)                                                   \
  do {                                                                         \
    hipError_t _e = (expr);                                                    \
    if (_e != hipSuccess) {                                                    \
      ::reno::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,           \
                        hipGetErrorString(_e));                                \
      return RENO_EHIP;                                                        \
    }                                                                          \
  } while (0)

#define RENO_CHECK_LAUNCH()                                                    \
  do {                                                                         \
    hipError_t _e = hipGetLastError();                                         \
    if (_e != hipSuccess) {                                                    \
      ::reno::set_error("%s:%d launch -> %s", __FILE__, __LINE__,              \
                        hipGetErrorString(_e));                                \
      return RENO_EHIP;                                                        \
    }                                                                          \
  } while (0)

#define RENO_TRY(expr)                                                         \
  do {                                                                         \
    int _s = (expr);                                                           \
    if (_s != RENO_OK) return _s;                                              \
  } while (0)

// Interleaved complex storage types (match torch complex128 / complex64).
struct __align__(16) z128 {
  double re, im;
};
struct __align__(8) c64 {
  float re, im;
};

template <typename S> struct real_of;
template <> struct real_of<z128> { typedef double type; };
template <> struct real_of<c64> { typedef float type; };

// Generic batched strided complex GEMM:
//   C[b](m,n) (+)= sum_k opA(A[b])(m,k) * opB(B[b])(k,n)
// Element (m,k) of A lives at A + b*sAb + m*sAm + k*sAk (units: elements of
// the storage type).  Either sAk == 1 or sAm == 1 (same for B with sBk/sBn);
// C must have sCn == 1.  a_real: A is a real matrix (storage type float or
// double, matching the compute precision) and the imaginary MFMA products are
// skipped.  conj flags conjugate A / B on load.  beta 0 overwrites, 1 adds.
enum StorageKind { ST_Z128 = 0, ST_C64 = 1, ST_F64 = 2, ST_F32 = 3 };

struct GemmCall {
  int M, N, K, batch;
  int prec;  // RENO_PREC_FP64 -> FP64 MFMA, RENO_PREC_MIXED -> FP32 MFMA
  const void* A;
  int a_kind;
  int64_t sAm, sAk, sAb;
  int conjA;
  const void* B;
  int b_kind;
  int64_t sBk, sBn, sBb;
  int conjB;
  void* C;
  int c_kind;
  int64_t sCm, sCn, sCb;
  int beta;
  // optional device flag: when *skip != 0 the launch does nothing (used by
  // the Lanczos loop to cancel speculatively queued iterations)
  const int* skip;
  // optional scratch for split-K partial sums (see splitk_bytes); without it
  // the GEMM runs unsplit
  void* part;
  size_t part_bytes;
  // set by launch_gemm: K range per blockIdx.z; launch-executed marker used
  // by the timing hook (so launches cancelled through `skip` are not counted)
  int kchunk;
  int* exec;
  // optional fused dot product with the result (complex128, batch 1,
  // contiguous rows): dot_part[b] = partial sums of Re(sum conj(C) dot_v)
  // over dot_np blocks (a fixed-order split; the caller folds them)
  const double2* dot_v;
  double* dot_part;
  int dot_np;
  // set by launch_gemm: tiles numbered column-major (each XCD's run of
  // tiles then covers a block of columns and reads all of A and its own
  // part of B) rather than row-major
  int tile_colmajor;
};

int launch_gemm(const GemmCall& g, hipStream_t stream);
// While an object of this type lives on a thread, the GEMM launches of that
// thread bypass the timing hook (which allocates its marker array and
// synchronises the device when it harvests markers).
struct TimingHookSuspend {
  TimingHookSuspend();
  ~TimingHookSuspend();
};
// bytes of split-K scratch launch_gemm would use for this shape (0: no split)
size_t splitk_bytes(int prec, int M, int N, int K, int batch);
// tile autotuner (gemm.hip): record this thread's GEMM calls instead of
// launching them (between begin and end); whether a call's signature is in
// the tuned table; time the candidates of one call and store the fastest
int gemm_record_begin(std::vector<GemmCall>* calls);
void gemm_record_end();
bool gemm_is_tuned(const GemmCall& g);
int gemm_tune(const GemmCall& g, hipStream_t stream, int reps, reno_tile_cfg* out);
int gemm_tune_reset();
int gemm_tune_entries();

inline size_t align_up(size_t x, size_t a = 256) { return (x + a - 1) / a * a; }

// Sum of v over the 256 threads of a block in a fixed order (every call
// with the same inputs gives the same bits, and every thread returns the
// same value): a butterfly over the 64 lanes of each wave (partners
// lane ^ 32, ..., lane ^ 1; both partners add the same two operands), then
// the four wave sums through sh[0..4).  Two barriers: the second lets the
// caller reuse sh at once.
__device__ __forceinline__ double block_sum256(double v, double* sh) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  const double s = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  __syncthreads();
  return s;
}

inline size_t elem_bytes(int prec) { return prec == RENO_PREC_FP64 ? 16 : 8; }

}  // namespace reno
