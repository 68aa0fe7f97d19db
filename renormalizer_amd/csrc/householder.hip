// Blocked Householder QR of a column-major complex128 matrix, on the device.
//
// LAPACK conventions (zgeqrf/zlarfg/zlarft/zungqr): reflector
// H_j = I - tau_j v_j v_j^H with v_j[j] = 1, beta_j = -sign(Re alpha) |(alpha,x)|
// real, so R has a real diagonal; Q = H_0 H_1 ... H_{k-1}.
//
// Panels of NB = 8 columns.  One 512-thread workgroup factors a panel with the
// panel held in registers (rows j0 + tid + 512 i), so each column step is a
// fused norm reduction, reflector, dot-product reduction and update with no
// trips to memory; it also forms the compact-WY factor T (zlarft, forward
// columnwise).  The trailing matrix is updated as A += V (-T^H) (V^H A) by
// one fused kernel per panel (each workgroup owns 4 columns: reduce V^H A
// over the rows, multiply by the 8x8 factor, apply the rank-8 update).  Q is
// formed by backward accumulation Q <- (I - V T V^H) Q with the same kernel.
#include "householder.h"

namespace reno {

constexpr int kQrNB = 8;
constexpr int kPanelThreads = 512;
constexpr int kApplyThreads = 256;
constexpr int kApplyCols = 4;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {  // conj(a) * b
  return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// Sum `nv` doubles across the workgroup; result broadcast in out[0..nv).
// sh must hold (blockDim/64) * nv doubles.
template <int NT, int NV>
__device__ __forceinline__ void block_reduce(double (&vals)[NV], int nv, double* sh, double* out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  constexpr int NW = NT / 64;
#pragma unroll
  for (int i = 0; i < NV; ++i) {
    if (i >= nv) break;
    double v = vals[i];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if (lane == 0) sh[wave * nv + i] = v;
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nv; i += NT) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) s += sh[w * nv + i];
    out[i] = s;
  }
  __syncthreads();
}

// Factor columns [j0, j0+jb) of X (m x ld, column-major), rows j0..m-1.
// Writes R / v in place, tau[j0..j0+jb), the explicit V into Vall
// (column j0+c, all rows, zeros above the unit diagonal), and into Tb the
// row-major 8x8 factors -T^H (Tb[0..63], trailing update) and -T
// (Tb[64..127], forming Q).
template <int RPT>
__global__ __launch_bounds__(kPanelThreads) void k_panel_qr(double2* __restrict__ X, int m, int ld,
                                                            int j0, int jb,
                                                            double2* __restrict__ tau,
                                                            double2* __restrict__ Vall,
                                                            double2* __restrict__ Tb) {
  __shared__ double sh[(kPanelThreads / 64) * 64];
  __shared__ double red[64];
  __shared__ double2 sT[kQrNB][kQrNB];
  __shared__ double2 sS[kQrNB][kQrNB];
  __shared__ double2 s_alpha;
  const int tid = threadIdx.x;
  double2 a[RPT][kQrNB];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = j0 + tid + kPanelThreads * i;
#pragma unroll
    for (int c = 0; c < kQrNB; ++c)
      a[i][c] = (r < m && c < jb) ? X[r + (int64_t)(j0 + c) * ld] : make_double2(0.0, 0.0);
  }
  double2 taus[kQrNB];
#pragma unroll
  for (int jj = 0; jj < kQrNB; ++jj) {
    taus[jj] = make_double2(0.0, 0.0);
    if (jj >= jb) continue;
    const int j = j0 + jj;
    // every thread has read the previous column's s_alpha
    __syncthreads();
    // ---- reflector for column jj: alpha = X[j][j], x = X[j+1:, j]
    double part[1] = {0.0};
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = j0 + tid + kPanelThreads * i;
      if (r > j && r < m) part[0] += a[i][jj].x * a[i][jj].x + a[i][jj].y * a[i][jj].y;
    }
    if (tid == jj) s_alpha = a[0][jj];
    block_reduce<kPanelThreads>(part, 1, sh, red);
    const double xnorm2 = red[0];
    const double2 alpha = s_alpha;
    double2 tj = make_double2(0.0, 0.0);
    if (!(xnorm2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xnorm2);
      const double beta = alpha.x >= 0.0 ? -nrm : nrm;
      tj = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      // scale = 1 / (alpha - beta)
      const double dr = alpha.x - beta, di = alpha.y;
      const double den = dr * dr + di * di;
      const double2 scale = make_double2(dr / den, -di / den);
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = j0 + tid + kPanelThreads * i;
        if (r > j && r < m) a[i][jj] = cmul(a[i][jj], scale);
      }
      if (tid == jj) a[0][jj] = make_double2(beta, 0.0);
    }
    taus[jj] = tj;
    // ---- apply H^H = I - conj(tau) v v^H to columns jj+1 .. jb-1
    if (jj + 1 < jb && !(tj.x == 0.0 && tj.y == 0.0)) {
      double w[2 * kQrNB];
#pragma unroll
      for (int c = 0; c < 2 * kQrNB; ++c) w[c] = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = j0 + tid + kPanelThreads * i;
        if (r < j || r >= m) continue;
        const double2 v = (r == j) ? make_double2(1.0, 0.0) : a[i][jj];
#pragma unroll
        for (int c = 0; c < kQrNB; ++c) {
          if (c <= jj || c >= jb) continue;
          const double2 p = cmulc(v, a[i][c]);
          w[2 * c] += p.x;
          w[2 * c + 1] += p.y;
        }
      }
      block_reduce<kPanelThreads>(w, 2 * kQrNB, sh, red);
      const double2 ct = make_double2(tj.x, -tj.y);  // conj(tau)
#pragma unroll
      for (int c = 0; c < kQrNB; ++c) {
        if (c <= jj || c >= jb) continue;
        const double2 f = cmul(ct, make_double2(red[2 * c], red[2 * c + 1]));
#pragma unroll
        for (int i = 0; i < RPT; ++i) {
          const int r = j0 + tid + kPanelThreads * i;
          if (r < j || r >= m) continue;
          const double2 v = (r == j) ? make_double2(1.0, 0.0) : a[i][jj];
          const double2 d = cmul(v, f);
          a[i][c].x -= d.x;
          a[i][c].y -= d.y;
        }
      }
    }
  }
  // ---- write back the panel and the explicit V
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = j0 + tid + kPanelThreads * i;
    if (r >= m) continue;
#pragma unroll
    for (int c = 0; c < kQrNB; ++c) {
      if (c >= jb) continue;
      X[r + (int64_t)(j0 + c) * ld] = a[i][c];
      const int jr = j0 + c;
      Vall[r + (int64_t)(j0 + c) * m] =
          r < jr ? make_double2(0.0, 0.0) : (r == jr ? make_double2(1.0, 0.0) : a[i][c]);
    }
  }
  // rows j0 .. j0+jb-1 above the panel rows of other threads are covered above;
  // rows < j0 of Vall column are zero
  for (int r = tid; r < j0; r += kPanelThreads)
    for (int c = 0; c < jb; ++c) Vall[r + (int64_t)(j0 + c) * m] = make_double2(0.0, 0.0);
  // ---- S[q][c] = v_q^H v_c for q < c (one reduction per column c, to keep
  // the register footprint at the panel itself), then T (zlarft, forward,
  // columnwise)
  {
#pragma unroll
    for (int c = 1; c < kQrNB; ++c) {
      if (c >= jb) continue;
      double s[2 * kQrNB];
#pragma unroll
      for (int t = 0; t < 2 * kQrNB; ++t) s[t] = 0.0;
#pragma unroll
      for (int i = 0; i < RPT; ++i) {
        const int r = j0 + tid + kPanelThreads * i;
        if (r >= m || r < j0 + c) continue;  // v_c is zero above row j0+c
        const double2 vc = (r == j0 + c) ? make_double2(1.0, 0.0) : a[i][c];
#pragma unroll
        for (int q = 0; q < kQrNB; ++q) {
          if (q >= c) continue;
          const double2 p = cmulc(a[i][q], vc);  // r > j0+q, so v_q[r] = a[i][q]
          s[2 * q] += p.x;
          s[2 * q + 1] += p.y;
        }
      }
      block_reduce<kPanelThreads>(s, 2 * c, sh, red);
      if (tid < c) sS[tid][c] = make_double2(red[2 * tid], red[2 * tid + 1]);
    }
    __syncthreads();
    if (tid == 0) {
      for (int q = 0; q < kQrNB; ++q)
        for (int c = 0; c < kQrNB; ++c) sT[q][c] = make_double2(0.0, 0.0);
      for (int c = 0; c < jb; ++c) {
        const double2 tc = taus[c];
        sT[c][c] = tc;
        // T[0:c, c] = -tau_c T[0:c, 0:c] S[0:c, c]
        for (int q = 0; q < c; ++q) {
          double2 acc = make_double2(0.0, 0.0);
          for (int p = q; p < c; ++p) acc = cadd(acc, cmul(sT[q][p], sS[p][c]));
          const double2 v = cmul(tc, acc);
          sT[q][c] = make_double2(-v.x, -v.y);
        }
      }
    }
    __syncthreads();
    if (tid < kQrNB * kQrNB) {
      const int q = tid / kQrNB, c = tid % kQrNB;
      const double2 th = sT[c][q];  // (-T^H)[q][c] = -conj(T[c][q])
      Tb[q * kQrNB + c] = make_double2(-th.x, th.y);
      const double2 t = sT[q][c];
      Tb[kQrNB * kQrNB + q * kQrNB + c] = make_double2(-t.x, -t.y);
    }
#pragma unroll
    for (int c = 0; c < kQrNB; ++c)
      if (tid == c && c < jb) tau[j0 + c] = taus[c];
  }
}

// A[j0:m, c] += V (F (V^H A[j0:m, c])) for columns c in [c_begin, c_end), with
// V = Vall[:, j0:j0+jb] and F the row-major 8x8 factor (-T^H or -T).
__global__ __launch_bounds__(kApplyThreads) void k_apply_wy(double2* __restrict__ A, int lda, int m,
                                                            int j0, int jb, int c_begin, int c_end,
                                                            const double2* __restrict__ Vall,
                                                            const double2* __restrict__ F) {
  __shared__ double sh[(kApplyThreads / 64) * 2 * kQrNB * kApplyCols];
  __shared__ double red[2 * kQrNB * kApplyCols];
  __shared__ double2 sW2[kQrNB][kApplyCols];
  const int tid = threadIdx.x;
  const int c0 = c_begin + blockIdx.x * kApplyCols;
  const int nc = min(kApplyCols, c_end - c0);
  double w[2 * kQrNB * kApplyCols];
#pragma unroll
  for (int t = 0; t < 2 * kQrNB * kApplyCols; ++t) w[t] = 0.0;
  for (int r = j0 + tid; r < m; r += kApplyThreads) {
    double2 v[kQrNB];
#pragma unroll
    for (int q = 0; q < kQrNB; ++q)
      v[q] = q < jb ? Vall[r + (int64_t)(j0 + q) * m] : make_double2(0.0, 0.0);
#pragma unroll
    for (int cc = 0; cc < kApplyCols; ++cc) {
      if (cc >= nc) continue;
      const double2 x = A[r + (int64_t)(c0 + cc) * lda];
#pragma unroll
      for (int q = 0; q < kQrNB; ++q) {
        const double2 p = cmulc(v[q], x);
        w[2 * (q * kApplyCols + cc)] += p.x;
        w[2 * (q * kApplyCols + cc) + 1] += p.y;
      }
    }
  }
  block_reduce<kApplyThreads>(w, 2 * kQrNB * kApplyCols, sh, red);
  if (tid < kQrNB * kApplyCols) {
    const int q = tid / kApplyCols, cc = tid % kApplyCols;
    double2 acc = make_double2(0.0, 0.0);
    for (int p = 0; p < jb; ++p)
      acc = cadd(acc, cmul(F[q * kQrNB + p], make_double2(red[2 * (p * kApplyCols + cc)],
                                                         red[2 * (p * kApplyCols + cc) + 1])));
    sW2[q][cc] = acc;
  }
  __syncthreads();
  for (int r = j0 + tid; r < m; r += kApplyThreads) {
    double2 v[kQrNB];
#pragma unroll
    for (int q = 0; q < kQrNB; ++q)
      v[q] = q < jb ? Vall[r + (int64_t)(j0 + q) * m] : make_double2(0.0, 0.0);
#pragma unroll
    for (int cc = 0; cc < kApplyCols; ++cc) {
      if (cc >= nc) continue;
      double2 x = A[r + (int64_t)(c0 + cc) * lda];
#pragma unroll
      for (int q = 0; q < kQrNB; ++q) x = cadd(x, cmul(v[q], sW2[q][cc]));
      A[r + (int64_t)(c0 + cc) * lda] = x;
    }
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

size_t qr_ws_bytes(int m, int n) {
  const int k = m < n ? m : n;
  const int np = (k + kQrNB - 1) / kQrNB;
  return align_up((size_t)m * k * 16) /* Vall */ + align_up((size_t)(k + 1) * 16) /* tau */ +
         align_up((size_t)np * 2 * kQrNB * kQrNB * 16) /* T factors */;
}

bool qr_supported(int m, int n) {
  (void)n;
  return m <= 5 * kPanelThreads;
}

template <int RPT>
static void launch_panel(hipStream_t s, double2* X, int m, int j0, int jb, double2* tau,
                         double2* Vall, double2* Tb) {
  hipLaunchKernelGGL(k_panel_qr<RPT>, dim3(1), dim3(kPanelThreads), 0, s, X, m, m, j0, jb, tau,
                     Vall, Tb);
}

int qr_householder(hipStream_t s, int m, int n, double2* X, double2* Q, void* ws) {
  const int k = m < n ? m : n;
  if (k <= 0) return RENO_OK;
  if (!qr_supported(m, n)) {
    set_error("qr_householder: %d rows exceed the register-resident panel", m);
    return RENO_EINVAL;
  }
  const int np = (k + kQrNB - 1) / kQrNB;
  char* p = (char*)ws;
  double2* Vall = (double2*)p;
  p += align_up((size_t)m * k * 16);
  double2* tau = (double2*)p;
  p += align_up((size_t)(k + 1) * 16);
  double2* Tall = (double2*)p;
  for (int ip = 0; ip < np; ++ip) {
    const int j0 = ip * kQrNB;
    const int jb = (k - j0) < kQrNB ? (k - j0) : kQrNB;
    const int rpt = (m - j0 + kPanelThreads - 1) / kPanelThreads;
    double2* Tb = Tall + (size_t)ip * 2 * kQrNB * kQrNB;
    switch (rpt) {
      case 1: launch_panel<1>(s, X, m, j0, jb, tau, Vall, Tb); break;
      case 2: launch_panel<2>(s, X, m, j0, jb, tau, Vall, Tb); break;
      case 3: launch_panel<3>(s, X, m, j0, jb, tau, Vall, Tb); break;
      case 4: launch_panel<4>(s, X, m, j0, jb, tau, Vall, Tb); break;
      default: launch_panel<5>(s, X, m, j0, jb, tau, Vall, Tb); break;
    }
    RENO_CHECK_LAUNCH();
    const int cb = j0 + jb;
    if (cb < n) {
      const int blocks = (n - cb + kApplyCols - 1) / kApplyCols;
      hipLaunchKernelGGL(k_apply_wy, dim3(blocks), dim3(kApplyThreads), 0, s, X, m, m, j0, jb, cb,
                         n, (const double2*)Vall, (const double2*)Tb);
      RENO_CHECK_LAUNCH();
    }
  }
  // Q = H_0 ... H_{k-1} applied to the first k columns of I (zungqr)
  {
    const int64_t tot = (int64_t)m * k;
    const int blocks = (int)((tot + 255) / 256 < 4096 ? (tot + 255) / 256 : 4096);
    hipLaunchKernelGGL(k_eye, dim3(blocks), dim3(256), 0, s, Q, m, k);
    RENO_CHECK_LAUNCH();
  }
  for (int ip = np - 1; ip >= 0; --ip) {
    const int j0 = ip * kQrNB;
    const int jb = (k - j0) < kQrNB ? (k - j0) : kQrNB;
    double2* Tb = Tall + (size_t)ip * 2 * kQrNB * kQrNB;
    // Q[j0:, j0:k] <- (I - V T V^H) Q[j0:, j0:k]  (factor -T in the second slot)
    const int blocks = (k - j0 + kApplyCols - 1) / kApplyCols;
    hipLaunchKernelGGL(k_apply_wy, dim3(blocks), dim3(kApplyThreads), 0, s, Q, m, m, j0, jb, j0, k,
                       (const double2*)Vall, (const double2*)(Tb + kQrNB * kQrNB));
    RENO_CHECK_LAUNCH();
  }
  return RENO_OK;
}

}  // namespace reno
