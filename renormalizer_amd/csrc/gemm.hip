// Batched strided complex GEMM on CDNA4 matrix cores.
//
// One kernel family serves every contraction of the TDVP/DMRG local update:
// the three stages of the effective-Hamiltonian matvec, the three stages of
// the environment update and the bond absorption (see contract.hip).
//
// Tiling: a 256-thread workgroup owns a 64x64 output tile and walks K in
// steps of 16.  Operands are staged through LDS with real and imaginary parts
// split into separate planes (conjugation applied on the way in), so that
// each wave feeds v_mfma_{f64,f32}_16x16x4 directly from LDS: a wave owns a
// 32x32 sub-tile = 2x2 MFMA blocks.  A complex multiply-accumulate is three
// real MFMAs (Gauss: Ar Br, Ai Bi, (Ar + Ai)(Br + Bi)), two when A is real
// (the MPO stage).  The next K-tile is fetched
// into registers while the current one is multiplied.
//
// MFMA 16x16x4 lane layout (checked on the device by reno_probe_mfma16):
//   A operand: lane l holds A[l % 16][l / 16]
//   B operand: lane l holds B[l / 16][l % 16]
//   C/D f32  : lane l, register i holds C[4 * (l / 16) + i][l % 16]
//   C/D f64  : lane l, register i holds C[4 * i + l / 16][l % 16]
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include "reno_common.h"

namespace reno {

typedef double d4_t __attribute__((ext_vector_type(4)));
typedef float f4_t __attribute__((ext_vector_type(4)));

template <typename R> struct Mfma;
template <> struct Mfma<double> {
  typedef d4_t acc;
  __device__ static __forceinline__ acc mac(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // output row of register r in a lane whose lane/16 is q
  __device__ static __forceinline__ int out_row(int q, int r) { return 4 * r + q; }
};
template <> struct Mfma<float> {
  typedef f4_t acc;
  __device__ static __forceinline__ acc mac(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static __forceinline__ int out_row(int q, int r) { return 4 * q + r; }
};

// Element access with the storage kind fixed at compile time (a runtime
// switch per element would split every load into branches and serialise the
// memory latency).
template <typename R, int KIND>
__device__ __forceinline__ void load_k(const void* base, int64_t off, R& re, R& im) {
  if constexpr (KIND == ST_Z128) {
    const double2 v = reinterpret_cast<const double2*>(base)[off];
    re = (R)v.x;
    im = (R)v.y;
  } else if constexpr (KIND == ST_C64) {
    const float2 v = reinterpret_cast<const float2*>(base)[off];
    re = (R)v.x;
    im = (R)v.y;
  } else if constexpr (KIND == ST_F64) {
    re = (R)reinterpret_cast<const double*>(base)[off];
    im = (R)0;
  } else {
    re = (R)reinterpret_cast<const float*>(base)[off];
    im = (R)0;
  }
}

template <typename R, int KIND>
__device__ __forceinline__ void store_k(void* base, int64_t off, R re, R im, int beta) {
  if constexpr (KIND == ST_Z128) {
    double2* p = reinterpret_cast<double2*>(base) + off;
    double2 v = make_double2((double)re, (double)im);
    if (beta) {
      const double2 o = *p;
      v.x += o.x;
      v.y += o.y;
    }
    *p = v;
  } else {
    float2* p = reinterpret_cast<float2*>(base) + off;
    float2 v = make_float2((float)re, (float)im);
    if (beta) {
      const float2 o = *p;
      v.x += o.x;
      v.y += o.y;
    }
    *p = v;
  }
}

// runtime-kind store (split-K reduction epilogue)
template <typename R>
__device__ __forceinline__ void store_elem(void* base, int kind, int64_t off, R re, R im, int beta) {
  if (kind == ST_Z128) store_k<R, ST_Z128>(base, off, re, im, beta);
  else store_k<R, ST_C64>(base, off, re, im, beta);
}

constexpr int kBM = 64;  // tile of k_cgemm; k_cgemm_s uses 32 x 32
constexpr int kBN = 64;
constexpr int kBK = 16;
constexpr int kThreads = 256;
constexpr int kPad = 1;  // odd row stride: conflict-free LDS writes along k
constexpr int kSmallTile = 32;
constexpr int64_t kSmallGridTiles = 64;  // 64 x 64 tiles (over the batch) up to which 32 x 32 is used

// AK: A is contiguous along k (sAk == 1), else along m.
// BKC: B is contiguous along k (sBk == 1), else along n.
// AREAL: A has no imaginary part.
// TM: square output tile (64: four waves of 32 x 32 = 2 x 2 MFMA blocks;
// 32: four waves of 16 x 16 = one block, for products whose 64 x 64 tile
// grid would leave most of the chip idle).
template <typename R, bool AK, bool BKC, int AKIND, int BKIND, int CKIND, int TM>
__device__ __forceinline__ void cgemm_body(const GemmCall& g) {
  constexpr int kBM = TM, kBN = TM;
  constexpr int MB = TM / 32;  // MFMA blocks per wave along m and n
  constexpr int kPerThread = (kBM * kBK) / kThreads;  // A and B elements per thread
  constexpr bool AREAL = (AKIND == ST_F64 || AKIND == ST_F32);
  // double-buffered: [buffer][re / im plane][k][m or n]
  __shared__ R sA[2][2][kBK][kBM + kPad];
  __shared__ R sB[2][2][kBK][kBN + kPad];
  typedef typename Mfma<R>::acc acc_t;
  if (g.skip && *g.skip) return;
  if (g.exec && blockIdx.x == 0 && blockIdx.y == 0 && blockIdx.z == 0 && threadIdx.x == 0)
    *g.exec = 1;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;

  // XCD-aware tile order: workgroups are dealt round-robin to the 8 XCDs, so
  // give each XCD a contiguous run of tiles.  Row-major runs share A row
  // panels and read all of B on every XCD; column-major runs share B column
  // panels and read all of A: the launcher picks the order that replicates
  // the smaller operand across the XCDs' L2 caches.
  const int tilesN = (g.N + kBN - 1) / kBN;
  const int tilesM = (g.M + kBM - 1) / kBM;
  const int ntiles = gridDim.x;
  const int t8 = (ntiles / 8) * 8;
  const int bid = blockIdx.x;
  const int tile = bid < t8 ? (bid % 8) * (ntiles / 8) + bid / 8 : bid;
  const int tm = g.tile_colmajor ? tile % tilesM : tile / tilesN;
  const int tn = g.tile_colmajor ? tile / tilesM : tile - (tile / tilesN) * tilesN;
  const int m0 = tm * kBM;
  const int n0 = tn * kBN;
  const int64_t bat = blockIdx.y;
  const int64_t aoff = bat * g.sAb;
  const int64_t boff = bat * g.sBb;
  // split-K: blockIdx.z owns K range [kb, ke) and writes its own partial
  const int kb = blockIdx.z * g.kchunk;
  const int ke = min(g.K, kb + g.kchunk);
  const int64_t coff = bat * g.sCb + (int64_t)blockIdx.z * gridDim.y * g.sCb;

  // Complex x complex products use three real products per block (Gauss):
  // P1 = Ar Br, P2 = Ai Bi, P3 = (Ar + Ai)(Br + Bi); then Re = P1 - P2,
  // Im = P3 - P1 - P2.  cr holds P1, ci holds P3 and cq P2.  In FP32 this
  // runs the tail iterations of the mixed mode, whose rounding is weighted
  // down by small Krylov coefficients (117.9 -> 103.7 us per C4 matvec).
  constexpr bool K3M = !AREAL;
  acc_t cr[MB][MB], ci[MB][MB], cq[MB][MB];
#pragma unroll
  for (int i = 0; i < MB; ++i)
#pragma unroll
    for (int j = 0; j < MB; ++j) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        cr[i][j][r] = (R)0;
        ci[i][j][r] = (R)0;
        cq[i][j][r] = (R)0;
      }
    }

  R pa_re[kPerThread], pa_im[kPerThread], pb_re[kPerThread], pb_im[kPerThread];

  auto a_coord = [&](int e, int& mm, int& kk) {
    if (AK) {
      kk = e % kBK;
      mm = e / kBK;
    } else {
      mm = e % kBM;
      kk = e / kBM;
    }
  };
  auto b_coord = [&](int e, int& kk, int& nn) {
    if (BKC) {
      kk = e % kBK;
      nn = e / kBK;
    } else {
      nn = e % kBN;
      kk = e / kBN;
    }
  };

  // Per-thread element offsets, set for the first K step and advanced by
  // kBK times the k stride after every load; a workgroup whose tile and K
  // range lie wholly inside the matrices loads without bounds tests.
  int64_t a_off[kPerThread], b_off[kPerThread];
  int a_kk[kPerThread], b_kk[kPerThread];
  bool a_in[kPerThread], b_in[kPerThread];
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    const int e = tid + i * kThreads;
    int mm, kk, nn;
    a_coord(e, mm, kk);
    a_in[i] = m0 + mm < g.M;
    a_kk[i] = kk;
    a_off[i] = a_in[i] ? aoff + (int64_t)(m0 + mm) * g.sAm + (int64_t)(kb + kk) * g.sAk : 0;
    b_coord(e, kk, nn);
    b_in[i] = n0 + nn < g.N;
    b_kk[i] = kk;
    b_off[i] = b_in[i] ? boff + (int64_t)(kb + kk) * g.sBk + (int64_t)(n0 + nn) * g.sBn : 0;
  }
  const int64_t a_step = (int64_t)kBK * g.sAk, b_step = (int64_t)kBK * g.sBk;
  const bool full = m0 + kBM <= g.M && n0 + kBN <= g.N && (ke - kb) % kBK == 0;

  // loads the K tile starting at k0 (absolute) and advances the offsets
  auto gload = [&](int k0, auto checked) {
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      R re = (R)0, im = (R)0;
      if (!decltype(checked)::value || (a_in[i] && k0 + a_kk[i] < ke))
        load_k<R, AKIND>(g.A, a_off[i], re, im);
      pa_re[i] = re;
      pa_im[i] = im;
      re = (R)0;
      im = (R)0;
      if (!decltype(checked)::value || (b_in[i] && k0 + b_kk[i] < ke))
        load_k<R, BKIND>(g.B, b_off[i], re, im);
      pb_re[i] = re;
      pb_im[i] = im;
      a_off[i] += a_step;
      b_off[i] += b_step;
    }
  };

  // conjugation is applied here, after the multiply of the current tile:
  // touching the prefetched registers any earlier would make the wave wait
  // for the global loads before its matrix-core work
  const R sgnA = g.conjA ? (R)-1 : (R)1;
  const R sgnB = g.conjB ? (R)-1 : (R)1;
  auto sstore = [&](int buf) {
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const int e = tid + i * kThreads;
      int mm, kk;
      a_coord(e, mm, kk);
      sA[buf][0][kk][mm] = pa_re[i];
      if (!AREAL) sA[buf][1][kk][mm] = sgnA * pa_im[i];
      int nn;
      b_coord(e, kk, nn);
      sB[buf][0][kk][nn] = pb_re[i];
      sB[buf][1][kk][nn] = sgnB * pb_im[i];
    }
  };

  const int r16 = lane & 15;
  const int kq = lane >> 4;

  // one barrier per K step: the next tile goes to the other LDS buffer while
  // this one is multiplied (all waves left the other buffer at the previous
  // barrier)
  auto run = [&](auto checked) {
    gload(kb, checked);
    sstore(0);
    __syncthreads();
    int cur = 0;
    for (int k0 = kb; k0 < ke; k0 += kBK) {
      const bool more = k0 + kBK < ke;
      if (more) gload(k0 + kBK, checked);
#pragma unroll
      for (int ks = 0; ks < kBK / 4; ++ks) {
        const int kr = ks * 4 + kq;
        R a_re[MB], a_im[MB], b_re[MB], b_im[MB];
#pragma unroll
        for (int mb = 0; mb < MB; ++mb) {
          a_re[mb] = sA[cur][0][kr][wm * (TM / 2) + mb * 16 + r16];
          a_im[mb] = AREAL ? (R)0 : sA[cur][1][kr][wm * (TM / 2) + mb * 16 + r16];
        }
#pragma unroll
        for (int nb = 0; nb < MB; ++nb) {
          b_re[nb] = sB[cur][0][kr][wn * (TM / 2) + nb * 16 + r16];
          b_im[nb] = sB[cur][1][kr][wn * (TM / 2) + nb * 16 + r16];
        }
        if constexpr (K3M) {
          R a_s[MB], b_s[MB];
#pragma unroll
          for (int mb = 0; mb < MB; ++mb) a_s[mb] = a_re[mb] + a_im[mb];
#pragma unroll
          for (int nb = 0; nb < MB; ++nb) b_s[nb] = b_re[nb] + b_im[nb];
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < MB; ++nb) {
              cr[mb][nb] = Mfma<R>::mac(a_re[mb], b_re[nb], cr[mb][nb]);
              cq[mb][nb] = Mfma<R>::mac(a_im[mb], b_im[nb], cq[mb][nb]);
              ci[mb][nb] = Mfma<R>::mac(a_s[mb], b_s[nb], ci[mb][nb]);
            }
        } else {
#pragma unroll
          for (int mb = 0; mb < MB; ++mb)
#pragma unroll
            for (int nb = 0; nb < MB; ++nb) {
              cr[mb][nb] = Mfma<R>::mac(a_re[mb], b_re[nb], cr[mb][nb]);
              ci[mb][nb] = Mfma<R>::mac(a_re[mb], b_im[nb], ci[mb][nb]);
              if (!AREAL) {
                cr[mb][nb] = Mfma<R>::mac(-a_im[mb], b_im[nb], cr[mb][nb]);
                ci[mb][nb] = Mfma<R>::mac(a_im[mb], b_re[nb], ci[mb][nb]);
              }
            }
        }
      }
      if (more) sstore(cur ^ 1);
      __syncthreads();
      cur ^= 1;
    }
  };
  if (full)
    run(std::integral_constant<bool, false>());
  else
    run(std::integral_constant<bool, true>());

#pragma unroll
  for (int mb = 0; mb < MB; ++mb)
#pragma unroll
    for (int nb = 0; nb < MB; ++nb)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * (TM / 2) + mb * 16 + Mfma<R>::out_row(kq, r);
        const int col = n0 + wn * (TM / 2) + nb * 16 + r16;
        R re = cr[mb][nb][r], im = ci[mb][nb][r];
        if constexpr (K3M) {
          re = cr[mb][nb][r] - cq[mb][nb][r];
          im = ci[mb][nb][r] - cr[mb][nb][r] - cq[mb][nb][r];
        }
        if (row < g.M && col < g.N)
          store_k<R, CKIND>(g.C, coff + (int64_t)row * g.sCm + (int64_t)col * g.sCn, re, im, g.beta);
      }
}

template <typename R, bool AK, bool BKC, int AKIND, int BKIND, int CKIND>
__global__ __launch_bounds__(kThreads) void k_cgemm(GemmCall g) {
  cgemm_body<R, AK, BKC, AKIND, BKIND, CKIND, kBM>(g);
}
template <typename R, bool AK, bool BKC, int AKIND, int BKIND, int CKIND>
__global__ __launch_bounds__(kThreads) void k_cgemm_s(GemmCall g) {
  cgemm_body<R, AK, BKC, AKIND, BKIND, CKIND, kSmallTile>(g);
}

// ---- instantiations -----------------------------------------------------
// The (precision, A k-contiguous, B k-contiguous, A/B/C storage) combinations
// the contractions use; launch_gemm rejects any other.  Their order is the
// "kind" index of the timing hook (reno_prof_kind_info).
typedef void (*GemmKernel)(GemmCall);
struct GemmKind {
  int prec, ak, bk, akind, bkind, ckind;
  GemmKernel fn;    // 64 x 64 tiles
  GemmKernel fn_s;  // 32 x 32 tiles
};
#define RENO_KIND(R, P, AK, BK, A, B, C) \
  {P, AK, BK, A, B, C, k_cgemm<R, (bool)AK, (bool)BK, A, B, C>, k_cgemm_s<R, (bool)AK, (bool)BK, A, B, C>}
static const GemmKind kGemmKinds[] = {
    // FP64: complex128 operands, real MPO stage
    RENO_KIND(double, RENO_PREC_FP64, 1, 0, ST_Z128, ST_Z128, ST_Z128),
    RENO_KIND(double, RENO_PREC_FP64, 1, 1, ST_Z128, ST_Z128, ST_Z128),
    RENO_KIND(double, RENO_PREC_FP64, 0, 0, ST_Z128, ST_Z128, ST_Z128),
    RENO_KIND(double, RENO_PREC_FP64, 0, 1, ST_Z128, ST_Z128, ST_Z128),
    RENO_KIND(double, RENO_PREC_FP64, 1, 0, ST_F64, ST_Z128, ST_Z128),
    // mixed: complex64 environments / intermediates, complex128 site vectors
    RENO_KIND(float, RENO_PREC_MIXED, 1, 0, ST_C64, ST_Z128, ST_C64),   // S1
    RENO_KIND(float, RENO_PREC_MIXED, 1, 0, ST_F32, ST_C64, ST_C64),    // S2
    RENO_KIND(float, RENO_PREC_MIXED, 1, 1, ST_C64, ST_C64, ST_Z128),   // S3 -> vector
    RENO_KIND(float, RENO_PREC_MIXED, 1, 1, ST_C64, ST_C64, ST_C64),    // S3 split-K partials
    RENO_KIND(float, RENO_PREC_MIXED, 0, 0, ST_Z128, ST_C64, ST_C64),   // env L S3
    RENO_KIND(float, RENO_PREC_MIXED, 1, 1, ST_Z128, ST_C64, ST_C64),   // env R S1 / S3
};
#undef RENO_KIND
static constexpr int kNumGemmKinds = (int)(sizeof(kGemmKinds) / sizeof(kGemmKinds[0]));

// The kernels address A and B through both strides, so the k-contiguity
// flags only choose the coalescing pattern: without an exact match (a
// degenerate dimension of size 1 makes both patterns valid) any
// instantiation with the same precision and storage kinds is correct.
static int find_kind(const GemmCall& g, bool ak, bool bk) {
  int any = -1;
  for (int i = 0; i < kNumGemmKinds; ++i) {
    const GemmKind& k = kGemmKinds[i];
    if (k.prec != g.prec || k.akind != g.a_kind || k.bkind != g.b_kind || k.ckind != g.c_kind)
      continue;
    if (k.ak == (int)ak && k.bk == (int)bk) return i;
    if (any < 0) any = i;
  }
  return any;
}

// Tile edge of the planner: 32 x 32 when the 64 x 64 grid is small (the
// bond-matrix products of the zero-site step: four times the workgroups,
// less split-K) or when M or N is at most 32, where a 64-wide tile would
// leave at least half of it idle (the batched MPO stage, M = d D_r = 30 at
// the spin-boson workload).
static int planned_tile(int M, int N, int batch) {
  const int64_t tiles64 = (int64_t)((M + kBM - 1) / kBM) * ((N + kBN - 1) / kBN) * batch;
  return (tiles64 <= kSmallGridTiles || M <= kSmallTile || N <= kSmallTile) ? kSmallTile : kBM;
}

// ---- split-K -------------------------------------------------------------
// Split K into chunks (multiples of the K tile) when the output tile grid
// would leave the chip idle; each chunk writes its own partial and
// k_split_reduce sums them in a fixed order (deterministic, no atomics).
// The split minimises an estimate: rounds of resident workgroups (two per
// CU: 512 on the 256 CUs) times the K extent of one workgroup plus a fixed
// per-workgroup cost (prologue, tile store) of about 48 K steps; chunks stay
// at least 64 deep.
static void split_plan(int M, int N, int K, int batch, int* nsplit, int* kchunk, int tile = kBM) {
  const int64_t tiles = (int64_t)((M + tile - 1) / tile) * ((N + tile - 1) / tile) * batch;
  *nsplit = 1;
  *kchunk = K > 0 ? K : 1;
  constexpr int64_t kSlots = 512;
  constexpr int kMinChunk = 64;
  constexpr int64_t kFixed = 48;
  if (tiles >= kSlots * 7 / 8 || K < 2 * kMinChunk) return;
  auto est = [&](int64_t ns, int64_t chunk) { return ((tiles * ns + kSlots - 1) / kSlots) * (chunk + kFixed); };
  int64_t best = est(1, K);
  for (int s = 2; s <= K / kMinChunk; ++s) {
    int chunk = (K + s - 1) / s;
    chunk = (chunk + kBK - 1) / kBK * kBK;
    const int ns = (K + chunk - 1) / chunk;
    const int64_t e = est(ns, chunk);
    if (e < best) {
      best = e;
      *nsplit = ns;
      *kchunk = chunk;
    }
  }
}

// C[b](m,n) (+)= sum_s P[s][b][m][n]
template <typename R>
__global__ void k_split_reduce(const typename std::conditional<sizeof(R) == 8, double2, float2>::type* __restrict__ P,
                               int nsplit, int batch, int M, int N, void* C, int c_kind,
                               int64_t sCm, int64_t sCb, int beta, const int* skip) {
  if (skip && *skip) return;
  const int64_t per = (int64_t)M * N;
  const int64_t total = per * batch;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    R re = (R)0, im = (R)0;
    for (int s = 0; s < nsplit; ++s) {
      const auto v = P[(int64_t)s * total + t];
      re += v.x;
      im += v.y;
    }
    const int64_t b = t / per, r = t - b * per;
    const int64_t m = r / N, n = r - m * N;
    store_elem<R>(C, c_kind, b * sCb + m * sCm + n, re, im, beta);
  }
}

// 256-thread fixed-order sum (deterministic)
// Split-K reduction of a contiguous complex128 result fused with the dot
// product the Lanczos recurrence needs next: C = sum_s P[s] (+ C), and
// dot_part[block] = sum over the block's share of Re(conj(C) v).
template <typename R>
__global__ __launch_bounds__(256) void k_split_reduce_dot(
    const typename std::conditional<sizeof(R) == 8, double2, float2>::type* __restrict__ P,
    int nsplit, int64_t total, double2* __restrict__ C, int beta, const double2* __restrict__ v,
    double* __restrict__ dot_part, const int* skip) {
  if (skip && *skip) return;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t t = blockIdx.x * (int64_t)256 + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * 256) {
    R re = (R)0, im = (R)0;
    for (int s = 0; s < nsplit; ++s) {
      const auto x = P[(int64_t)s * total + t];
      re += x.x;
      im += x.y;
    }
    double2 c = make_double2((double)re, (double)im);
    if (beta) {
      const double2 o = C[t];
      c.x += o.x;
      c.y += o.y;
    }
    C[t] = c;
    const double2 w = v[t];
    acc += c.x * w.x + c.y * w.y;
  }
  const double sum = block_sum256(acc, sh);
  if (threadIdx.x == 0) dot_part[blockIdx.x] = sum;
}

// dot_part[block] = partial sums of Re(conj(c) v) (unsplit GEMM results)
__global__ __launch_bounds__(256) void k_dot_part(const double2* __restrict__ c,
                                                  const double2* __restrict__ v, int64_t n,
                                                  double* __restrict__ dot_part, const int* skip) {
  if (skip && *skip) return;
  __shared__ double sh[256];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double2 a = c[i], b = v[i];
    acc += a.x * b.x + a.y * b.y;
  }
  const double sum = block_sum256(acc, sh);
  if (threadIdx.x == 0) dot_part[blockIdx.x] = sum;
}

// ---- optional launch timing (used by the benchmark's roofline figure) ----
// When enabled, every GEMM launch is recorded with its template
// instantiation and algorithmic FLOP count (8 per complex MAC, 4 per real x
// complex MAC) and given a marker slot in device memory that the kernel sets
// when it runs (launches cancelled on the device by the Krylov stop flag
// leave it 0).  A pseudo-random 1-in-period sample of the launches is also
// timed by two HIP events bound to the dispatch (hipExtLaunchKernel).
struct ProfRec {
  hipEvent_t a, b;  // null unless the launch is in the timed sample
  double flop;
  int kind;
  int ran;  // marker value once harvested, -1 before
};
static std::mutex g_prof_mu;
static int g_prof_period = 0;  // 0: off
static uint64_t g_prof_seq = 0;
static std::vector<ProfRec> g_prof;
static size_t g_prof_harvested = 0;  // records [0, this) have `ran` set
static std::vector<hipEvent_t> g_event_pool;
static constexpr size_t kExecSlots = 1 << 20;
static int* g_exec = nullptr;  // kExecSlots markers, allocated on first enable

static hipEvent_t pooled_event() {
  if (!g_event_pool.empty()) {
    hipEvent_t e = g_event_pool.back();
    g_event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// splitmix64 finaliser: the sample must not follow the periodic launch
// pattern of the Lanczos iteration
static uint64_t mix64(uint64_t x) {
  x += 0x9e3779b97f4a7c15ull;
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}

// Copy the markers of the records not yet harvested (at most kExecSlots of
// them, in slot order) into their `ran` fields and clear the slots.  Waits
// for the device.  Caller holds g_prof_mu.
static int harvest_markers() {
  const size_t n = g_prof.size() - g_prof_harvested;
  if (n == 0 || !g_exec) return RENO_OK;
  std::vector<int> ran(n);
  RENO_HIP_CHECK(hipDeviceSynchronize());
  RENO_HIP_CHECK(hipMemcpy(ran.data(), g_exec, n * sizeof(int), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < n; ++i) g_prof[g_prof_harvested + i].ran = ran[i];
  g_prof_harvested += n;
  RENO_HIP_CHECK(hipMemset(g_exec, 0, kExecSlots * sizeof(int)));
  RENO_HIP_CHECK(hipDeviceSynchronize());
  return RENO_OK;
}

static thread_local int t_hook_suspended = 0;
TimingHookSuspend::TimingHookSuspend() { ++t_hook_suspended; }
TimingHookSuspend::~TimingHookSuspend() { --t_hook_suspended; }

// ---- tile autotuner state --------------------------------------------------
// Tuned (tile, split) choices keyed by everything that selects the kernel and
// its grid: precision, shape, batch, storage kinds, contiguity, conjugation.
// The planner above is the default; reno_autotune measures the candidates
// of one operator's GEMMs on the device and stores the fastest here.
struct TuneKey {
  int prec, M, N, K, batch, a_kind, b_kind, c_kind, ak, bk, conjA, conjB;
  bool operator==(const TuneKey& o) const { return std::memcmp(this, &o, sizeof o) == 0; }
};
struct TuneKeyHash {
  size_t operator()(const TuneKey& k) const {
    const int* w = &k.prec;
    uint64_t h = 1469598103934665603ull;
    for (size_t i = 0; i < sizeof k / sizeof(int); ++i) h = (h ^ (uint32_t)w[i]) * 1099511628211ull;
    return (size_t)h;
  }
};
struct TileChoice {
  int tile, nsplit;
};
static std::mutex g_tune_mu;
static std::unordered_map<TuneKey, TileChoice, TuneKeyHash> g_tuned;
static std::atomic<int> g_tuned_count{0};
// tuner only: a forced choice for this thread's launches, or a list that
// receives this thread's GEMM calls instead of launching them
static thread_local const TileChoice* t_force = nullptr;
static thread_local std::vector<GemmCall>* t_record = nullptr;

static TuneKey tune_key(const GemmCall& g) {
  TuneKey k;
  std::memset(&k, 0, sizeof k);
  k.prec = g.prec; k.M = g.M; k.N = g.N; k.K = g.K; k.batch = g.batch;
  k.a_kind = g.a_kind; k.b_kind = g.b_kind; k.c_kind = g.c_kind;
  k.ak = g.sAk == 1; k.bk = g.sBk == 1; k.conjA = g.conjA; k.conjB = g.conjB;
  return k;
}

// K split into (about) s chunks that are multiples of the K tile
static void split_fixed(int K, int s, int* nsplit, int* kchunk) {
  if (K <= 0 || s <= 1) {
    *nsplit = 1;
    *kchunk = K > 0 ? K : 1;
    return;
  }
  int chunk = (K + s - 1) / s;
  chunk = (chunk + kBK - 1) / kBK * kBK;
  *kchunk = chunk;
  *nsplit = (K + chunk - 1) / chunk;
}

int launch_gemm(const GemmCall& g, hipStream_t stream) {
  if (g.M < 0 || g.N < 0 || g.K < 0 || g.batch < 0) {
    set_error("gemm: negative size M=%d N=%d K=%d batch=%d", g.M, g.N, g.K, g.batch);
    return RENO_EINVAL;
  }
  if (g.M == 0 || g.N == 0 || g.batch == 0) return RENO_OK;
  if (g.sCn != 1 || (g.sAk != 1 && g.sAm != 1) || (g.sBk != 1 && g.sBn != 1)) {
    set_error("gemm: unsupported strides sAm=%lld sAk=%lld sBk=%lld sBn=%lld sCn=%lld",
              (long long)g.sAm, (long long)g.sAk, (long long)g.sBk, (long long)g.sBn,
              (long long)g.sCn);
    return RENO_EINVAL;
  }
  if (g.c_kind != ST_Z128 && g.c_kind != ST_C64) {
    set_error("gemm: output must be complex");
    return RENO_EINVAL;
  }
  if (g.batch > 65535) {
    set_error("gemm: batch %d exceeds grid limit", g.batch);
    return RENO_EINVAL;
  }
  const bool areal = (g.a_kind == ST_F64 || g.a_kind == ST_F32);
  if (g.b_kind == ST_F64 || g.b_kind == ST_F32) {
    set_error("gemm: B must be complex");
    return RENO_EINVAL;
  }
  if (g.dot_part && (g.batch != 1 || g.sCm != g.N || g.c_kind != ST_Z128 || !g.dot_v ||
                     g.dot_np < 1)) {
    set_error("gemm: fused dot needs one contiguous complex128 result");
    return RENO_EINVAL;
  }
  if (t_record) {
    t_record->push_back(g);
    return RENO_OK;
  }
  int tile = planned_tile(g.M, g.N, g.batch);
  // split K when the tile grid cannot fill the chip and scratch is given
  int nsplit = 1, kchunk = g.K > 0 ? g.K : 1;
  split_plan(g.M, g.N, g.K, g.batch, &nsplit, &kchunk, tile);
  const size_t cell = (size_t)g.batch * g.M * g.N * elem_bytes(g.prec);
  if (nsplit > 1 && (!g.part || g.part_bytes < nsplit * cell)) {
    nsplit = 1;
    kchunk = g.K > 0 ? g.K : 1;
  }
  // a tuned choice replaces the planner's when its partials fit the scratch
  {
    TileChoice ch{0, 0};
    if (t_force) {
      ch = *t_force;
    } else if (g_tuned_count.load(std::memory_order_acquire) > 0) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tuned.find(tune_key(g));
      if (it != g_tuned.end()) ch = it->second;
    }
    if (ch.tile) {
      int ns, chunk;
      split_fixed(g.K, ch.nsplit, &ns, &chunk);
      if (ns == 1 || (g.part && g.part_bytes >= ns * cell)) {
        tile = ch.tile;
        nsplit = ns;
        kchunk = chunk;
      }
    }
  }
  const int64_t tiles = (int64_t)((g.M + tile - 1) / tile) * ((g.N + tile - 1) / tile);
  if (tiles > 0x7fffffff) {
    set_error("gemm: too many tiles");
    return RENO_EINVAL;
  }
  const bool ak = (g.sAk == 1);
  const bool bk = (g.sBk == 1);
  GemmCall k = g;
  k.kchunk = kchunk;
  // same K for both operands: M < N means A is the smaller one
  k.tile_colmajor = g.M < g.N;
  int kid = -1;
  if (nsplit > 1) {
    // partial sums [split][batch][M][N] in the compute precision, overwritten
    k.C = g.part;
    k.c_kind = g.prec == RENO_PREC_FP64 ? ST_Z128 : ST_C64;
    k.sCm = g.N;
    k.sCn = 1;
    k.sCb = (int64_t)g.M * g.N;
    k.beta = 0;
    kid = find_kind(k, ak, bk);
    if (kid < 0) {  // no instantiation for the partials: run unsplit
      nsplit = 1;
      k = g;
      k.kchunk = g.K > 0 ? g.K : 1;
      k.tile_colmajor = g.M < g.N;
    }
  }
  if (kid < 0) kid = find_kind(k, ak, bk);
  if (kid < 0) {
    set_error("gemm: no kernel for prec=%d A k-contiguous=%d B k-contiguous=%d kinds A=%d B=%d C=%d",
              g.prec, (int)ak, (int)bk, g.a_kind, g.b_kind, k.c_kind);
    return RENO_EINVAL;
  }
  dim3 grid((unsigned)tiles, (unsigned)g.batch, (unsigned)nsplit);
  const GemmKernel kfn = tile == kSmallTile ? kGemmKinds[kid].fn_s : kGemmKinds[kid].fn;
  // K == 0 needs no special case: the kernel stores its zero accumulators.
  ProfRec rec{nullptr, nullptr, 0.0, 0, -1};
  bool prof = false, timed = false;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (g_prof_period > 0 && g_exec && !t_hook_suspended) {
      // every slot in use: collect the markers before reusing them
      if (g_prof.size() - g_prof_harvested == kExecSlots) {
        const int st = harvest_markers();
        if (st != RENO_OK) return st;
      }
      prof = true;
      k.exec = g_exec + (g_prof.size() - g_prof_harvested);
      if (mix64(g_prof_seq++) % (uint64_t)g_prof_period == 0) {
        rec.a = pooled_event();
        rec.b = pooled_event();
        timed = rec.a && rec.b;
      }
    }
  }
  {
    void* args[] = {(void*)&k};
    if (timed) {
      // the events are bound to this dispatch: they time the matrix-core
      // kernel alone, without the queue latency in front of it (the split-K
      // reduction is a separate, bandwidth-bound kernel)
      RENO_HIP_CHECK(hipExtLaunchKernel((const void*)kfn, grid, dim3(kThreads), args,
                                        0, stream, rec.a, rec.b, 0));
    } else {
      RENO_HIP_CHECK(hipLaunchKernel((const void*)kfn, grid, dim3(kThreads), args, 0,
                                     stream));
    }
  }
  RENO_CHECK_LAUNCH();
  const bool fuse_dot = g.dot_part != nullptr;
  if (fuse_dot && nsplit > 1) {
    const int64_t total = (int64_t)g.M * g.N;
    if (g.prec == RENO_PREC_FP64)
      hipLaunchKernelGGL(k_split_reduce_dot<double>, dim3(g.dot_np), dim3(256), 0, stream,
                         (const double2*)g.part, nsplit, total, (double2*)g.C, g.beta, g.dot_v,
                         g.dot_part, g.skip);
    else
      hipLaunchKernelGGL(k_split_reduce_dot<float>, dim3(g.dot_np), dim3(256), 0, stream,
                         (const float2*)g.part, nsplit, total, (double2*)g.C, g.beta, g.dot_v,
                         g.dot_part, g.skip);
    RENO_CHECK_LAUNCH();
  } else if (nsplit > 1) {
    const int64_t total = (int64_t)g.batch * g.M * g.N;
    const int blocks = (int)((total + 255) / 256 < 8192 ? (total + 255) / 256 : 8192);
    if (g.prec == RENO_PREC_FP64)
      hipLaunchKernelGGL(k_split_reduce<double>, dim3(blocks), dim3(256), 0, stream,
                         (const double2*)g.part, nsplit, g.batch, g.M, g.N, g.C, g.c_kind, g.sCm,
                         g.sCb, g.beta, g.skip);
    else
      hipLaunchKernelGGL(k_split_reduce<float>, dim3(blocks), dim3(256), 0, stream,
                         (const float2*)g.part, nsplit, g.batch, g.M, g.N, g.C, g.c_kind, g.sCm,
                         g.sCb, g.beta, g.skip);
    RENO_CHECK_LAUNCH();
  }
  if (fuse_dot && nsplit == 1) {
    hipLaunchKernelGGL(k_dot_part, dim3(g.dot_np), dim3(256), 0, stream, (const double2*)g.C,
                       g.dot_v, (int64_t)g.M * g.N, g.dot_part, g.skip);
    RENO_CHECK_LAUNCH();
  }
  if (prof) {
    rec.flop = (areal ? 4.0 : 8.0) * (double)g.M * (double)g.N * (double)g.K * (double)g.batch;
    // timing-hook kinds: [0, n) 64 x 64 tiles, [n, 2n) 32 x 32 tiles
    rec.kind = tile == kSmallTile ? kid + kNumGemmKinds : kid;
    if (!timed) rec.a = rec.b = nullptr;
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(rec);
  }
  return RENO_OK;
}

size_t splitk_bytes(int prec, int M, int N, int K, int batch) {
  int nsplit = 1, kchunk = 1;
  split_plan(M, N, K, batch, &nsplit, &kchunk, planned_tile(M, N, batch));
  return nsplit > 1 ? (size_t)nsplit * batch * M * N * elem_bytes(prec) : 0;
}

// ---- tile autotuner ---------------------------------------------------------
int gemm_record_begin(std::vector<GemmCall>* calls) {
  t_record = calls;
  return RENO_OK;
}
void gemm_record_end() { t_record = nullptr; }

bool gemm_is_tuned(const GemmCall& g) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return g_tuned.count(tune_key(g)) != 0;
}

// Time every candidate (tile 64 or 32) x (split of K into 1, 2, 3, 4, 6, 8,
// 12 or 16 chunks of at least 32, partials within the call's scratch) of
// one GEMM call on `stream` with HIP events, `reps` launches each after one
// untimed launch, and store the fastest.  The planner's own choice is kept
// unless a candidate beats it by more than 2 %.  Synchronises the stream.
int gemm_tune(const GemmCall& g, hipStream_t stream, int reps, reno_tile_cfg* out) {
  TimingHookSuspend quiet;
  TileChoice planned{planned_tile(g.M, g.N, g.batch), 1};
  {
    int ns = 1, chunk = 1;
    split_plan(g.M, g.N, g.K, g.batch, &ns, &chunk, planned.tile);
    const size_t cell = (size_t)g.batch * g.M * g.N * elem_bytes(g.prec);
    planned.nsplit = (ns > 1 && g.part && g.part_bytes >= ns * cell) ? ns : 1;
  }
  std::vector<TileChoice> cand{planned};
  const size_t cell = (size_t)g.batch * g.M * g.N * elem_bytes(g.prec);
  for (int tile : {kBM, kSmallTile}) {
    std::vector<int> seen;
    for (int s : {1, 2, 3, 4, 6, 8, 12, 16}) {
      int ns, chunk;
      split_fixed(g.K, s, &ns, &chunk);
      if (ns > 1 && (chunk < 32 || !g.part || g.part_bytes < ns * cell)) continue;
      if (std::find(seen.begin(), seen.end(), ns) != seen.end()) continue;
      seen.push_back(ns);
      if (tile == planned.tile && ns == planned.nsplit) continue;
      cand.push_back(TileChoice{tile, ns});
    }
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  RENO_HIP_CHECK(hipEventCreate(&e0));
  RENO_HIP_CHECK(hipEventCreate(&e1));
  std::vector<float> us(cand.size(), 0.0f);
  int st = RENO_OK;
  for (size_t i = 0; i < cand.size() && st == RENO_OK; ++i) {
    t_force = &cand[i];
    st = launch_gemm(g, stream);
    if (st == RENO_OK && hipEventRecord(e0, stream) != hipSuccess) st = RENO_EHIP;
    for (int r = 0; r < reps && st == RENO_OK; ++r) st = launch_gemm(g, stream);
    if (st == RENO_OK && hipEventRecord(e1, stream) != hipSuccess) st = RENO_EHIP;
    if (st == RENO_OK && hipEventSynchronize(e1) != hipSuccess) st = RENO_EHIP;
    float ms = 0.0f;
    if (st == RENO_OK && hipEventElapsedTime(&ms, e0, e1) != hipSuccess) st = RENO_EHIP;
    us[i] = 1e3f * ms / (float)(reps > 0 ? reps : 1);
    t_force = nullptr;
  }
  t_force = nullptr;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (st != RENO_OK) {
    if (st == RENO_EHIP) set_error("autotune: HIP event failure");
    return st;
  }
  size_t best = 0;
  for (size_t i = 1; i < cand.size(); ++i)
    if (us[i] < us[best]) best = i;
  if (best != 0 && us[best] > 0.98f * us[0]) best = 0;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[tune_key(g)] = cand[best];
    g_tuned_count.store((int)g_tuned.size(), std::memory_order_release);
  }
  if (out) {
    out->M = g.M; out->N = g.N; out->K = g.K; out->batch = g.batch; out->prec = g.prec;
    out->tile = cand[best].tile; out->nsplit = cand[best].nsplit;
    out->planned_tile = planned.tile; out->planned_nsplit = planned.nsplit;
    out->us_tuned = us[best]; out->us_planned = us[0];
  }
  return RENO_OK;
}

int gemm_tune_reset() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tuned.clear();
  g_tuned_count.store(0, std::memory_order_release);
  return RENO_OK;
}

int gemm_tune_entries() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (int)g_tuned.size();
}

int prof_enable(int period) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  if (period < 0) {  // resume the current record
    if (!g_exec) {
      set_error("prof_enable: nothing to resume");
      return RENO_EINVAL;
    }
    g_prof_period = -period;
    return RENO_OK;
  }
  g_prof_period = period;
  if (period > 0) {
    for (auto& r : g_prof) {
      if (r.a) g_event_pool.push_back(r.a);
      if (r.b) g_event_pool.push_back(r.b);
    }
    g_prof.clear();
    g_prof_harvested = 0;
    g_prof_seq = 0;
    if (!g_exec) RENO_HIP_CHECK(hipMalloc((void**)&g_exec, kExecSlots * sizeof(int)));
    RENO_HIP_CHECK(hipDeviceSynchronize());
    RENO_HIP_CHECK(hipMemset(g_exec, 0, kExecSlots * sizeof(int)));
    RENO_HIP_CHECK(hipDeviceSynchronize());
  }
  return RENO_OK;
}

// Per instantiation: over the timed sample, executed launches, their GPU
// milliseconds and FLOP; over all recorded launches, the executed count and
// FLOP and the count of all.  Waits for the device.
int prof_collect(int nkinds, double* ms, double* flop_timed, int* timed, double* flop,
                 int* executed, int* queued) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int k = 0; k < nkinds; ++k) {
    if (ms) ms[k] = 0.0;
    if (flop_timed) flop_timed[k] = 0.0;
    if (timed) timed[k] = 0;
    if (flop) flop[k] = 0.0;
    if (executed) executed[k] = 0;
    if (queued) queued[k] = 0;
  }
  const int st = harvest_markers();
  if (st != RENO_OK) return st;
  for (auto& r : g_prof) {
    if (r.kind >= nkinds) continue;
    if (queued) queued[r.kind] += 1;
    if (!r.ran) continue;
    if (executed) executed[r.kind] += 1;
    if (flop) flop[r.kind] += r.flop;
    if (!r.a) continue;
    RENO_HIP_CHECK(hipEventSynchronize(r.b));
    float t = 0.f;
    RENO_HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
    if (ms) ms[r.kind] += t;
    if (flop_timed) flop_timed[r.kind] += r.flop;
    if (timed) timed[r.kind] += 1;
  }
  return RENO_OK;
}

// Per instantiation, the event-timed launches that were cancelled on the
// device (they return at once): their count and summed event milliseconds,
// i.e. what the event bracket itself adds to a launch.  Waits for the device.
int prof_cancelled(int nkinds, double* ms, int* count) {
  std::lock_guard<std::mutex> lk(g_prof_mu);
  for (int k = 0; k < nkinds; ++k) {
    ms[k] = 0.0;
    count[k] = 0;
  }
  const int st = harvest_markers();
  if (st != RENO_OK) return st;
  for (auto& r : g_prof) {
    if (r.kind >= nkinds || r.ran || !r.a) continue;
    RENO_HIP_CHECK(hipEventSynchronize(r.b));
    float t = 0.f;
    RENO_HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
    ms[r.kind] += t;
    count[r.kind] += 1;
  }
  return RENO_OK;
}

// ---- lane-layout probe: one MFMA 16x16x4 on row-major a (16x4), b (4x16);
// writes the raw accumulator registers c[lane*4 + r] -------------------------
template <typename R>
__global__ void k_probe_mfma16(const R* a, const R* b, R* c) {
  const int lane = threadIdx.x;
  typename Mfma<R>::acc acc;
  for (int r = 0; r < 4; ++r) acc[r] = (R)0;
  const R av = a[(lane % 16) * 4 + lane / 16];
  const R bv = b[(lane / 16) * 16 + lane % 16];
  acc = Mfma<R>::mac(av, bv, acc);
  for (int r = 0; r < 4; ++r) c[lane * 4 + r] = acc[r];
}

}  // namespace reno

extern "C" int reno_probe_mfma16(int prec, const void* a, const void* b, void* c,
                                 void* stream) {
  hipStream_t s = (hipStream_t)stream;
  if (prec == RENO_PREC_FP64)
    hipLaunchKernelGGL(reno::k_probe_mfma16<double>, dim3(1), dim3(64), 0, s,
                       (const double*)a, (const double*)b, (double*)c);
  else
    hipLaunchKernelGGL(reno::k_probe_mfma16<float>, dim3(1), dim3(64), 0, s,
                       (const float*)a, (const float*)b, (float*)c);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_prof_enable(int period) { return reno::prof_enable(period); }

extern "C" int reno_prof_kind_info(int kind, int* info) {
  if (kind >= 0 && kind < 2 * reno::kNumGemmKinds && info) {
    const reno::GemmKind& k = reno::kGemmKinds[kind % reno::kNumGemmKinds];
    info[0] = k.prec;
    info[1] = k.ak;
    info[2] = k.bk;
    info[3] = k.akind;
    info[4] = k.bkind;
    info[5] = k.ckind;
    info[6] = kind < reno::kNumGemmKinds ? reno::kBM : reno::kSmallTile;
  }
  return 2 * reno::kNumGemmKinds;
}

extern "C" int reno_prof_cancelled(int nkinds, double* ms, int* count) {
  if (nkinds < 0 || (nkinds > 0 && (!ms || !count))) {
    reno::set_error("prof_cancelled: bad arguments");
    return RENO_EINVAL;
  }
  return reno::prof_cancelled(nkinds, ms, count);
}

extern "C" int reno_prof_collect(int nkinds, double* ms, double* flop_timed, int* timed,
                                 double* flop, int* executed, int* queued) {
  return reno::prof_collect(nkinds, ms, flop_timed, timed, flop, executed, queued);
}

extern "C" int reno_zgemm(int M, int N, int K, const void* A, const void* B, void* C,
                          void* stream) {
  reno::GemmCall g{};
  g.M = M;
  g.N = N;
  g.K = K;
  g.batch = 1;
  g.prec = RENO_PREC_FP64;
  g.A = A;
  g.a_kind = reno::ST_Z128;
  g.sAm = K;
  g.sAk = 1;
  g.B = B;
  g.b_kind = reno::ST_Z128;
  g.sBk = N;
  g.sBn = 1;
  g.C = C;
  g.c_kind = reno::ST_Z128;
  g.sCm = N;
  g.sCn = 1;
  return reno::launch_gemm(g, (hipStream_t)stream);
}
