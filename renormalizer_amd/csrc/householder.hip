// Blocked Householder QR of a column-major complex128 matrix, on the device.
//
// LAPACK conventions (zgeqrf/zlarfg/zlarft/zungqr): reflector
// H_j = I - tau_j v_j v_j^H with v_j[j] = 1, beta_j = -sign(Re alpha) |(alpha,x)|
// real, so R has a real diagonal; Q = H_0 H_1 ... H_{k-1}.
//
// R phase, per panel of NB = 8 columns: (a) a small launch applies the
// previous panel's block reflector to the new panel's columns (one workgroup
// per column, column and reflector rows held in registers); (b) k_qr_step:
// workgroup 0 factors the panel with its columns held in registers (rows
// r0 + tid + 512 i), one reduction per column (norm of x, the dot products
// with the later panel columns; the pivot row goes through LDS), and forms
// the panel's compact-WY factor T (zlarft, forward columnwise); the other
// workgroups meanwhile apply the previous panel's reflector to one trailing
// column each.  512 threads per workgroup keep every register array in
// VGPRs (two waves per SIMD).  A reduction adds quad partners by DPP, passes
// the quad sums through LDS and finishes each value in one 16-lane row;
// barriers order LDS only.
//
// Q phase (zungqr): with the explicit V (m x k, unit lower trapezoidal) and
// the full k x k factor T of H_0...H_{k-1} = I - V T V^H,
//   Q = (I - V T V^H) [I; 0] = E + V (N V1^H),  N = -T, V1 = V[0:k, :],
// i.e. two matrix-core GEMMs.  N is assembled from the per-panel factors by
// recursive doubling, N12 = N11 (V1^H V2) N22 for adjacent blocks, with the
// Gram matrix V^H V from one more GEMM.
#include <algorithm>

#include "householder.h"

namespace reno {

constexpr int kQrNB = 8;
constexpr int kPanelThreads = 512;  // two waves per SIMD: up to 256 registers per lane
constexpr int kQuads = kPanelThreads / 4;
constexpr int kMaxRpt = 5;          // rows per thread: m <= 2560

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// The row loops below are FP64 VALU-bound: their complex multiply-adds are
// written as four fused multiply-adds each (a separate product and sum
// would take six instructions).
// (re, im) += conj(a) b
__device__ __forceinline__ void acc_conj_mul(double& re, double& im, double2 a, double2 b) {
  re = fma(a.x, b.x, fma(a.y, b.y, re));
  im = fma(a.x, b.y, fma(-a.y, b.x, im));
}
// y + a b
__device__ __forceinline__ double2 cmul_add(double2 y, double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, y.x)), fma(a.x, b.y, fma(a.y, b.x, y.y)));
}
// y - a b
__device__ __forceinline__ double2 cmul_sub(double2 y, double2 a, double2 b) {
  return make_double2(fma(-a.x, b.x, fma(a.y, b.y, y.x)), fma(-a.x, b.y, fma(-a.y, b.x, y.y)));
}

// Workgroup barrier ordering LDS only: every exchange between the threads
// of these kernels goes through LDS, so a barrier need not wait for the
// global stores of the factored columns (a plain __syncthreads would).
__device__ __forceinline__ void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A double moved between lanes by a DPP pattern that reads a valid lane for
// every active lane (quad permutations, row mirrors).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Sum NV doubles over the workgroup in a fixed order (deterministic);
// results in out[0..NV), visible to every thread on return.  Stage 1: quad
// sums by two DPP exchanges (quad_perm [1,0,3,2], [2,3,0,1]), stored by the
// first lane of each quad (sh holds NV * kQuads doubles).  Stage 2: one
// 16-lane row per value; lane p adds quad sums p, p + 16, ... (consecutive
// lanes read consecutive words), then the row is summed by the two quad
// exchanges and the half-row and row mirrors.  Every lane of a row ends
// with the same bits (each exchange adds the same two operands).
template <int NV>
__device__ __forceinline__ void lds_reduce(double (&vals)[NV], double* sh, double* out) {
  static_assert(16 * NV <= kPanelThreads, "one 16-lane row per value");
  const int tid = threadIdx.x;
  const bool lead = (tid & 3) == 0;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    double v = vals[i];
    v += dpp_f64<0xB1>(v);
    v += dpp_f64<0x4E>(v);
    if (lead) sh[i * kQuads + (tid >> 2)] = v;
  }
  lds_barrier();
  if (tid < 16 * NV) {
    const int v = tid >> 4, p = tid & 15;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kQuads / 16; ++k) s += sh[v * kQuads + k * 16 + p];
    s += dpp_f64<0xB1>(s);
    s += dpp_f64<0x4E>(s);
    s += dpp_f64<0x141>(s);
    s += dpp_f64<0x140>(s);
    if (p == 0) out[v] = s;
  }
  lds_barrier();
}

// Column c (rows jp..m-1) <- (I + V F V^H) column, V = Vall[:, jp:jp+jbp],
// F the row-major 8x8 factor (-T^H of panel jp) in LDS.  The column and the
// V rows jp + tid + 512 i are loaded once (all loads in flight together) and
// both passes run from registers.
template <int RPT>
__device__ __forceinline__ void apply_prev_column(double2* __restrict__ X, int m, int ld, int c, int jp, int jbp,
                                  const double2* __restrict__ Vall, const double2* F,
                                  double* sh, double* red, double2* sG) {
  const int tid = threadIdx.x;
  const double2 zero = make_double2(0.0, 0.0);
  double2 x[RPT];
  double2 v[RPT][kQrNB];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = jp + tid + kPanelThreads * i;
    const bool in = r < m;
    x[i] = in ? X[r + (int64_t)c * ld] : zero;
#pragma unroll
    for (int q = 0; q < kQrNB; ++q) v[i][q] = (in && q < jbp) ? Vall[r + (int64_t)(jp + q) * m] : zero;
  }
  double w[2 * kQrNB];
#pragma unroll
  for (int q = 0; q < 2 * kQrNB; ++q) w[q] = 0.0;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
#pragma unroll
    for (int q = 0; q < kQrNB; ++q) {
      acc_conj_mul(w[2 * q], w[2 * q + 1], v[i][q], x[i]);
    }
  }
  lds_reduce<2 * kQrNB>(w, sh, red);
  if (tid < kQrNB) {
    const int q = tid;
    double2 acc = zero;
    for (int p = 0; p < kQrNB; ++p) acc = cadd(acc, cmul(F[q * kQrNB + p], make_double2(red[2 * p], red[2 * p + 1])));
    sG[q] = acc;
  }
  lds_barrier();
  double2 g[kQrNB];
#pragma unroll
  for (int q = 0; q < kQrNB; ++q) g[q] = sG[q];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = jp + tid + kPanelThreads * i;
    double2 y = x[i];
#pragma unroll
    for (int q = 0; q < kQrNB; ++q) y = cmul_add(y, v[i][q], g[q]);
    if (r < m) X[r + (int64_t)c * ld] = y;
  }
}

// One column jj of the panel.  The registers hold rows r0 + tid + 512 i
// (r0 <= j0, rows from r0 to j0 - 1 lie above the panel and pass through
// unchanged) of the panel in a rotating slot order: column jj + c sits in
// slot (OFF + c) mod 8 for c < nc, and reflector v_q (q < jj) in slot
// (OFF + 8 - jj + q) mod 8, so that every register index is a compile-time
// constant for the caller's OFF.  The column's reflector replaces it in its
// slot, which is where the next column (OFF + 1) expects v_jj.  One
// reduction per column gives |x|^2, the dot products
// g_c = sum_{r>j} conj(a_rj) a_rc with the later columns, and the same sums
// with the earlier reflectors, from which column jj of V^H V and hence
// column jj of the compact-WY factor T (zlarft, forward columnwise) follow
// without another pass over V.  The pivot row (alpha, a_jc and the
// v_q(j)) goes through LDS (double-buffered by column parity).  Each column
// is written to X / Vall as soon as it is final.  Rows of chunk i >= 1 lie
// below every panel column (r >= r0 + 512 > j) and, for i < RPT - 1, inside
// the matrix; rows past m hold zeros throughout.
template <int RPT, int OFF>
__device__ __forceinline__ void factor_column(double2 (&a)[RPT][kQrNB], int jj, int r0, int m, int ld,
                                              int j0, int jb, double2* __restrict__ X,
                                              double2* __restrict__ Vall, double* sh, double* red,
                                              double2 (*piv)[kQrNB], double2* sTau,
                                              double2 (*sT)[kQrNB]) {
  const int tid = threadIdx.x;
  const int j = j0 + jj;
  const int nc = jb - jj;  // live columns
  double2* pv = piv[jj & 1];
  double vals[1 + 2 * kQrNB];
#pragma unroll
  for (int t = 0; t < 1 + 2 * kQrNB; ++t) vals[t] = 0.0;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = r0 + tid + kPanelThreads * i;
    if (i > 0 || r > j) {
      const double2 x = a[i][OFF];
      vals[0] = fma(x.x, x.x, fma(x.y, x.y, vals[0]));
#pragma unroll
      for (int c = 1; c < kQrNB; ++c)
        acc_conj_mul(vals[1 + 2 * c], vals[2 + 2 * c], x, a[i][(OFF + c) & (kQrNB - 1)]);
    } else if (r == j) {
#pragma unroll
      for (int c = 0; c < kQrNB; ++c) pv[c] = a[i][(OFF + c) & (kQrNB - 1)];
    }
  }
  lds_reduce<1 + 2 * kQrNB>(vals, sh, red);
  const double xnorm2 = red[0];
  const double2 alpha = pv[0];
  double2 tj = make_double2(0.0, 0.0);
  double2 scale = make_double2(1.0, 0.0);
  double beta = alpha.x;
  if (!(xnorm2 == 0.0 && alpha.y == 0.0)) {
    const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
    beta = alpha.x >= 0.0 ? -nrm : nrm;
    const double ib = 1.0 / beta;
    tj = make_double2((beta - alpha.x) * ib, -alpha.y * ib);
    // scale = 1 / (alpha - beta)
    const double dr = alpha.x - beta, di = alpha.y;
    const double iden = 1.0 / (dr * dr + di * di);
    scale = make_double2(dr * iden, -di * iden);
  }
  if (tid == 0) sTau[jj] = tj;
  // T[q][jj] = -tau_jj sum_{p=q}^{jj-1} T[q][p] S[p][jj] for q < jj, with
  // S[p][jj] = v_p^H v_jj = conj(v_p(j)) + scale conj(G_p), G_p the sum of
  // conj(x) v_p over the rows below j (relative slot 8 - jj + p of the
  // reduction).  Wave 0 forms it while the other waves update their rows.
  if (tid < jj) {
    const int q = tid;
    double2 acc = make_double2(0.0, 0.0);
    for (int p = q; p < jj; ++p) {
      const int sl = kQrNB - jj + p;
      const double2 gp = make_double2(red[1 + 2 * sl], -red[2 + 2 * sl]);  // conj(G_p)
      const double2 vpj = pv[sl];
      const double2 spj = cadd(make_double2(vpj.x, -vpj.y), cmul(scale, gp));
      acc = cadd(acc, cmul(sT[q][p], spj));
    }
    const double2 v = cmul(tj, acc);
    sT[q][jj] = make_double2(-v.x, -v.y);
  } else if (tid == jj) {
    sT[jj][jj] = tj;
  }
  const bool refl = !(tj.x == 0.0 && tj.y == 0.0);
  // f_c = conj(tau) (a_jc + conj(scale) g_c): coefficient of v in H^H a_c.
  // Without a reflection scale is 1 and every f_c is zero.
  double2 f[kQrNB];
  const double2 ct = make_double2(tj.x, -tj.y);
  const double2 cs = make_double2(scale.x, -scale.y);
#pragma unroll
  for (int c = 0; c < kQrNB; ++c)
    f[c] = (refl && c > 0 && c < nc)
               ? cmul(ct, cadd(pv[c], cmul(cs, make_double2(red[1 + 2 * c], red[2 + 2 * c]))))
               : make_double2(0.0, 0.0);
  // the reflector's rows and the final column first ...
  double2 vr[RPT];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = r0 + tid + kPanelThreads * i;
    const double2 x = a[i][OFF];
    if (i == 0) {
      vr[0] = r > j ? cmul(x, scale) : x;  // v_r below the diagonal
      double2 out = vr[0];                 // final column value
      if (r == j) {
        out = make_double2(beta, 0.0);
        vr[0] = make_double2(1.0, 0.0);
      }
      if (r < j) out = x;
      if (r < m) {
        X[r + (int64_t)j * ld] = out;
        Vall[r + (int64_t)j * m] = r < j ? make_double2(0.0, 0.0) : vr[0];
      }
    } else {
      vr[i] = cmul(x, scale);  // zero for rows past m
      if (i < RPT - 1 || r < m) {
        X[r + (int64_t)j * ld] = vr[i];
        Vall[r + (int64_t)j * m] = vr[i];
      }
    }
  }
  // ... then the later columns, a_c -= v f_c on rows >= j, one slot at a
  // time behind a wave-uniform test: the slots c >= nc hold reflectors
  // (f_c = 0) and are skipped, about half of the update work
  const bool upd0 = r0 + tid >= j;
#pragma unroll
  for (int c = 1; c < kQrNB; ++c) {
    if (c < nc) {
#pragma unroll
      for (int i = 0; i < RPT; ++i)
        if (i > 0 || upd0) a[i][(OFF + c) & (kQrNB - 1)] = cmul_sub(a[i][(OFF + c) & (kQrNB - 1)], vr[i], f[c]);
    }
  }
  // v_jj takes the column's slot
  {
    const int r = r0 + tid;
    a[0][OFF] = (r < j || r >= m) ? make_double2(0.0, 0.0) : vr[0];
#pragma unroll
    for (int i = 1; i < RPT; ++i) a[i][OFF] = vr[i];
  }
}

// The panel, two columns per loop iteration with their slot offsets fixed
// at compile time, then the slots rotate by two (one register rotation per
// two columns instead of one per column; a fully unrolled panel overflows
// the instruction cache, and four columns per iteration exceed the register
// budget).
template <int RPT>
__device__ __forceinline__ void factor_panel(double2 (&a)[RPT][kQrNB], int r0, int m, int ld, int j0, int jb,
                             double2* __restrict__ X, double2* __restrict__ Vall, double* sh,
                             double* red, double2 (*piv)[kQrNB], double2* sTau,
                             double2 (*sT)[kQrNB]) {
#pragma unroll 1
  for (int jj = 0; jj < jb; jj += 2) {
    factor_column<RPT, 0>(a, jj, r0, m, ld, j0, jb, X, Vall, sh, red, piv, sTau, sT);
    if (jj + 1 < jb) factor_column<RPT, 1>(a, jj + 1, r0, m, ld, j0, jb, X, Vall, sh, red, piv, sTau, sT);
    if (jj + 2 < jb) {
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        double2 t[kQrNB];
#pragma unroll
        for (int c = 0; c < kQrNB; ++c) t[c] = a[i][(c + 2) & (kQrNB - 1)];
#pragma unroll
        for (int c = 0; c < kQrNB; ++c) a[i][c] = t[c];
      }
    }
  }
}

// The panel columns into registers (rows r0 + tid + 512 i, column j0 + c in
// slot c; zeros past m and past the panel).
template <int RPT>
__device__ __forceinline__ void load_panel(double2 (&a)[RPT][kQrNB], const double2* __restrict__ X,
                                           int m, int ld, int r0, int j0, int jb) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = r0 + tid + kPanelThreads * i;
#pragma unroll
    for (int c = 0; c < kQrNB; ++c)
      a[i][c] = (r < m && c < jb) ? X[r + (int64_t)(j0 + c) * ld] : make_double2(0.0, 0.0);
  }
}

// Apply the block reflector of panel jp (-T^H in Tall) to columns
// c_begin + blockIdx.x (one workgroup per column, rows jp + tid + 512 i).
template <int RPT>
__global__ __launch_bounds__(kPanelThreads) void k_qr_apply(double2* __restrict__ X, int m, int ld,
                                                            int jp, int jbp, int c_begin, int c_end,
                                                            const double2* __restrict__ Vall,
                                                            const double2* __restrict__ Tall) {
  __shared__ double sh[2 * kQrNB * kQuads];
  __shared__ double red[2 * kQrNB];
  __shared__ double2 sF[kQrNB * kQrNB];
  __shared__ double2 sG[kQrNB];
  const int c = c_begin + (int)blockIdx.x;
  if (c >= c_end) return;
  if (threadIdx.x < kQrNB * kQrNB) sF[threadIdx.x] = Tall[(size_t)(jp / kQrNB) * 2 * kQrNB * kQrNB + threadIdx.x];
  lds_barrier();
  apply_prev_column<RPT>(X, m, ld, c, jp, jbp, Vall, sF, sh, red, sG);
}

// One step of the R phase: workgroup 0 factors panel [j0, j0+jb) (already
// updated by all previous reflectors); workgroups 1.. apply the reflector of
// panel jp (jp < 0: none) to columns j0+jb, j0+jb+1, ..., n-1.  Both hold
// rows from r0 = jp (j0 for the first panel), RPT per thread.
// Outputs: R / v in X, tau[j0..), explicit V columns in Vall (zeros above
// the unit diagonal) and Tall[p*128 ..] = -T^H (row-major 8x8) followed by -T.
template <int RPT>
__global__ __launch_bounds__(kPanelThreads) void k_qr_step(double2* __restrict__ X, int m, int ld,
                                                           int n, int jp, int jbp, int j0, int jb,
                                                           double2* __restrict__ tau,
                                                           double2* __restrict__ Vall,
                                                           double2* __restrict__ Tall) {
  __shared__ double sh[(1 + 2 * kQrNB) * kQuads];
  __shared__ double red[1 + 2 * kQrNB];
  __shared__ double2 sT[kQrNB][kQrNB];
  __shared__ double2 sTau[kQrNB];
  __shared__ double2 sPiv[2][kQrNB];
  __shared__ double2 sF[kQrNB * kQrNB];
  __shared__ double2 sG[kQrNB];
  const int tid = threadIdx.x;
  if (blockIdx.x > 0) {
    const int c = j0 + jb + (int)blockIdx.x - 1;
    if (c >= n || jp < 0) return;
    if (tid < kQrNB * kQrNB) sF[tid] = Tall[(size_t)(jp / kQrNB) * 2 * kQrNB * kQrNB + tid];
    lds_barrier();
    apply_prev_column<RPT>(X, m, ld, c, jp, jbp, Vall, sF, sh, red, sG);
    return;
  }
  if (jb == 0) return;
  const int r0 = jp >= 0 ? jp : j0;
  for (int t = tid; t < kQrNB * kQrNB; t += kPanelThreads) sT[t / kQrNB][t % kQrNB] = make_double2(0.0, 0.0);
  double2 a[RPT][kQrNB];
  load_panel<RPT>(a, X, m, ld, r0, j0, jb);
  factor_panel<RPT>(a, r0, m, ld, j0, jb, X, Vall, sh, red, sPiv, sTau, sT);
  for (int r = tid; r < r0; r += kPanelThreads)
    for (int c = 0; c < jb; ++c) Vall[r + (int64_t)(j0 + c) * m] = make_double2(0.0, 0.0);
  lds_barrier();
  double2* Tb = Tall + (size_t)(j0 / kQrNB) * 2 * kQrNB * kQrNB;
  if (tid < kQrNB * kQrNB) {
    const int q = tid / kQrNB, c = tid % kQrNB;
    const double2 th = sT[c][q];  // (-T^H)[q][c] = -conj(T[c][q])
    Tb[q * kQrNB + c] = make_double2(-th.x, th.y);
    const double2 t = sT[q][c];
    Tb[kQrNB * kQrNB + q * kQrNB + c] = make_double2(-t.x, -t.y);
  }
  if (tid < jb) tau[j0 + tid] = sTau[tid];
}

// N (kp x kp, row-major) <- 0 except the per-panel diagonal blocks -T_p;
// S (kp x kp) <- 0 (the Gram GEMM fills its k x k corner)
__global__ void k_qr_tinit(double2* __restrict__ N, double2* __restrict__ S, int kp, int np,
                           const double2* __restrict__ Tall) {
  const int64_t total = (int64_t)kp * kp;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / kp), j = (int)(t - (int64_t)i * kp);
    const int p = i / kQrNB;
    double2 v = make_double2(0.0, 0.0);
    if (p < np && j / kQrNB == p)
      v = Tall[(size_t)p * 2 * kQrNB * kQrNB + kQrNB * kQrNB + (i % kQrNB) * kQrNB + (j % kQrNB)];
    N[t] = v;
    S[t] = make_double2(0.0, 0.0);
  }
}

__global__ void k_eye(double2* Q, int m, int k) {
  const int64_t total = (int64_t)m * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = t / m, r = t - c * m;
    Q[t] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
}

struct QrLayout {
  int k, np, kp;
  size_t off_V, off_tau, off_T, off_N, off_S, off_Y, off_X, off_part, part_bytes, total;
};

static QrLayout qr_layout(int m, int n) {
  QrLayout L{};
  L.k = m < n ? m : n;
  L.np = (L.k + kQrNB - 1) / kQrNB;
  L.kp = kQrNB;
  while (L.kp < L.np * kQrNB) L.kp *= 2;
  const int k = L.k, kp = L.kp;
  size_t part = 0;
  part = std::max(part, splitk_bytes(RENO_PREC_FP64, k, k, m, 1));  // Gram
  part = std::max(part, splitk_bytes(RENO_PREC_FP64, k, k, k, 1));  // X
  part = std::max(part, splitk_bytes(RENO_PREC_FP64, k, m, k, 1));  // Q
  for (int b = kQrNB; b < kp; b *= 2) part = std::max(part, splitk_bytes(RENO_PREC_FP64, b, b, b, kp / (2 * b)));
  size_t o = 0;
  L.off_V = o; o += align_up((size_t)m * k * 16);
  L.off_tau = o; o += align_up((size_t)(k + 1) * 16);
  L.off_T = o; o += align_up((size_t)L.np * 2 * kQrNB * kQrNB * 16);
  L.off_N = o; o += align_up((size_t)kp * kp * 16);
  L.off_S = o; o += align_up((size_t)kp * kp * 16);
  L.off_Y = o; o += align_up((size_t)kp * kp / 4 * 16 + 16);
  L.off_X = o; o += align_up((size_t)k * k * 16);
  L.off_part = o; o += align_up(part);
  L.part_bytes = align_up(part);
  L.total = o;
  return L;
}

size_t qr_ws_bytes(int m, int n) { return qr_layout(m, n).total; }

bool qr_supported(int m, int n) {
  (void)n;
  return m <= kMaxRpt * kPanelThreads;
}

template <int RPT>
static void launch_step(hipStream_t s, int grid, double2* X, int m, int n, int jp, int jbp, int j0,
                        int jb, double2* tau, double2* Vall, double2* Tall) {
  if (jp >= 0 && jb > 0)
    hipLaunchKernelGGL(k_qr_apply<RPT>, dim3(jb), dim3(kPanelThreads), 0, s, X, m, m, jp, jbp, j0, j0 + jb,
                       (const double2*)Vall, (const double2*)Tall);
  hipLaunchKernelGGL(k_qr_step<RPT>, dim3(grid), dim3(kPanelThreads), 0, s, X, m, m, n, jp, jbp, j0,
                     jb, tau, Vall, Tall);
}

static GemmCall zgemm_call(int M, int N, int K, int batch, const void* A, int64_t sAm, int64_t sAk,
                           int64_t sAb, int conjA, const void* B, int64_t sBk, int64_t sBn,
                           int64_t sBb, int conjB, void* C, int64_t sCm, int64_t sCb, int beta,
                           void* part, size_t part_bytes) {
  GemmCall g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch; g.prec = RENO_PREC_FP64;
  g.A = A; g.a_kind = ST_Z128; g.sAm = sAm; g.sAk = sAk; g.sAb = sAb; g.conjA = conjA;
  g.B = B; g.b_kind = ST_Z128; g.sBk = sBk; g.sBn = sBn; g.sBb = sBb; g.conjB = conjB;
  g.C = C; g.c_kind = ST_Z128; g.sCm = sCm; g.sCn = 1; g.sCb = sCb; g.beta = beta;
  g.part = part; g.part_bytes = part_bytes;
  return g;
}

int qr_householder(hipStream_t s, int m, int n, double2* X, double2* Q, void* ws) {
  const QrLayout L = qr_layout(m, n);
  const int k = L.k;
  if (k <= 0) return RENO_OK;
  if (!qr_supported(m, n)) {
    set_error("qr_householder: %d rows exceed the register-resident panel", m);
    return RENO_EINVAL;
  }
  char* base = (char*)ws;
  double2* Vall = (double2*)(base + L.off_V);
  double2* tau = (double2*)(base + L.off_tau);
  double2* Tall = (double2*)(base + L.off_T);
  double2* Nm = (double2*)(base + L.off_N);
  double2* Sm = (double2*)(base + L.off_S);
  double2* Y = (double2*)(base + L.off_Y);
  double2* Xm = (double2*)(base + L.off_X);
  void* part = base + L.off_part;
  // ---- R phase: per panel, the previous reflector on the panel columns,
  // then the panel factorisation alongside the rest of the trailing update;
  // a final launch brings the last reflector to columns k..n-1
  for (int ip = 0; ip <= L.np; ++ip) {
    const int j0 = ip < L.np ? ip * kQrNB : k;
    const int jb = ip < L.np ? std::min(kQrNB, k - j0) : 0;
    const int jp = ip > 0 ? (ip - 1) * kQrNB : -1;
    const int jbp = ip > 0 ? std::min(kQrNB, k - jp) : 0;
    const int trail = ip > 0 ? std::max(0, n - (j0 + jb)) : 0;
    if (jb == 0 && trail == 0) break;
    // rows from r0 = jp (j0 for the first panel): those the previous
    // reflector and the panel touch
    const int r0 = jp >= 0 ? jp : j0;
    const int rpt = std::max(1, (m - r0 + kPanelThreads - 1) / kPanelThreads);
    const int grid = 1 + trail;
    switch (rpt) {
      case 1: launch_step<1>(s, grid, X, m, n, jp, jbp, j0, jb, tau, Vall, Tall); break;
      case 2: launch_step<2>(s, grid, X, m, n, jp, jbp, j0, jb, tau, Vall, Tall); break;
      case 3: launch_step<3>(s, grid, X, m, n, jp, jbp, j0, jb, tau, Vall, Tall); break;
      case 4: launch_step<4>(s, grid, X, m, n, jp, jbp, j0, jb, tau, Vall, Tall); break;
      default: launch_step<5>(s, grid, X, m, n, jp, jbp, j0, jb, tau, Vall, Tall); break;
    }
    RENO_CHECK_LAUNCH();
  }
  // ---- Q phase
  const int kp = L.kp;
  {
    const int64_t tot = (int64_t)kp * kp;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 4096);
    hipLaunchKernelGGL(k_qr_tinit, dim3(blocks), dim3(256), 0, s, Nm, Sm, kp, L.np,
                       (const double2*)Tall);
    RENO_CHECK_LAUNCH();
  }
  if (L.np > 1) {
    // Gram S = V^H V (k x k corner of the kp x kp buffer)
    RENO_TRY(launch_gemm(zgemm_call(k, k, m, 1, Vall, m, 1, 0, 1, Vall, 1, m, 0, 0, Sm, kp, 0, 0,
                                    part, L.part_bytes), s));
    // recursive doubling: N12 = N11 S12 N22 for adjacent b x b blocks
    for (int b = kQrNB; b < kp; b *= 2) {
      const int pairs = kp / (2 * b);
      const int64_t bs = (int64_t)2 * b * (kp + 1);
      // Y_i = S[o1:, o2:] N[o2:, o2:]
      RENO_TRY(launch_gemm(zgemm_call(b, b, b, pairs, Sm + b, kp, 1, bs, 0, Nm + (int64_t)b * kp + b, kp, 1,
                                      bs, 0, Y, b, (int64_t)b * b, 0, part, L.part_bytes), s));
      // N[o1:, o2:] = N[o1:, o1:] Y_i
      RENO_TRY(launch_gemm(zgemm_call(b, b, b, pairs, Nm, kp, 1, bs, 0, Y, b, 1, (int64_t)b * b, 0,
                                      Nm + b, kp, bs, 0, part, L.part_bytes), s));
    }
  }
  // X = N V1^H  (k x k)
  RENO_TRY(launch_gemm(zgemm_call(k, k, k, 1, Nm, kp, 1, 0, 0, Vall, m, 1, 0, 1, Xm, k, 0, 0, part,
                                  L.part_bytes), s));
  // Q = E + V X, computed as Q^T = E^T + X^T V^T into the column-major Q
  {
    const int64_t tot = (int64_t)m * k;
    const int blocks = (int)std::min<int64_t>((tot + 255) / 256, 4096);
    hipLaunchKernelGGL(k_eye, dim3(blocks), dim3(256), 0, s, Q, m, k);
    RENO_CHECK_LAUNCH();
  }
  RENO_TRY(launch_gemm(zgemm_call(k, m, k, 1, Xm, 1, k, 0, 0, Vall, m, 1, 0, 0, Q, m, 0, 1, part,
                                  L.part_bytes), s));
  return RENO_OK;
}

}  // namespace reno
