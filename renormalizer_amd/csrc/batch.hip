// Batched one-site TDVP-PS sweep for many small independent states (static
// disorder samples, parameter scans): one launch per sweep position, one
// workgroup per state.  Each workgroup runs its state's whole site update
// (reference mps/mps.py:1296-1397) without leaving the kernel:
//   1. expm(-i dt/2 H_eff) on the site tensor by Lanczos with the
//      reference's stopping rule (lib/krylov/krylov.py:27-84);
//   2. the quantum-number blocked QR / RQ of the evolved tensor
//      (svd_qn.py:99-240, QR branch; LAPACK zgeqr2 / zung2r conventions);
//   3. the environment update of the new isometry (lib.py:169-250);
//   4. expm(+i dt/2 H_bond) on the bond matrix;
//   5. the bond matrix absorbed into the neighbouring site.
// Small states leave a single-state launch sequence latency-bound (every
// stage is a few microseconds of work behind several microseconds of launch
// and dependency); here the states of one launch keep every CU busy and no
// stage waits for the host.  All contractions are the same strided complex
// GEMMs the single-state path launches (contract.hip), executed by the
// workgroup tile by tile on the FP64 matrix cores.
#include <math.h>

#include "contract.h"
#include "tricoef.h"

namespace reno {

constexpr int kBT = 512;      // threads per workgroup: eight waves, two per SIMD
constexpr int kWaves = kBT / 64;
constexpr int kBTile = 64;    // GEMM output tile (four waves of 32 x 32)
constexpr int kEngT = 256;    // threads of one GEMM tile engine (two per workgroup)
constexpr int kBK = 16;       // GEMM K step

typedef double d4b_t __attribute__((ext_vector_type(4)));

struct BatchLds {
  double ta[2][2][kBK][kBTile + 1];  // [engine] A tile, re / im planes, [k][m]
  double tb[2][2][kBK][kBTile + 1];  // [engine] B tile, re / im planes, [k][n]
  double red[kBT];
  double J[kWaves][kChebMax];         // Bessel terms, one set per wave
  int K[kWaves];
};

// fixed-order workgroup sum, visible to every thread on return
__device__ double wg_sum(double v, double* red) {
  const int tid = threadIdx.x;
  red[tid] = v;
  __syncthreads();
  for (int off = kBT / 2; off > 0; off >>= 1) {
    if (tid < off) red[tid] += red[tid + off];
    __syncthreads();
  }
  const double r = red[0];
  __syncthreads();
  return r;
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// C[b] (+)= op(A[b]) op(B[b]) for one GemmCall (the strided batched product
// of reno_common.h; A complex128 or float64, B and C complex128), computed by
// the calling workgroup: two 64 x 64 output tiles at a time (two engines of
// four waves, in step), K in steps of 16 staged through LDS with the next
// step's elements already loaded into registers while the current one is
// multiplied; each wave of an engine owns a 32 x 32
// quadrant of 2 x 2 v_mfma_f64_16x16x4f64 blocks (three real products per
// complex product, two when A is real).  MFMA 16x16x4 f64 lane layout:
// A lane l = A[l % 16][l / 16], B lane l = B[l / 16][l % 16], D lane l
// register r = D[4 r + l / 16][l % 16].
__device__ void wg_gemm(const GemmCall& g, BatchLds& s) {
  // two tile engines of four waves each work on consecutive tiles in step
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int eng = wave >> 2, tid = threadIdx.x & (kEngT - 1);
  const int wm = (wave & 3) >> 1, wn = wave & 1;
  const bool areal = g.a_kind == ST_F64;
  const int tm = (g.M + kBTile - 1) / kBTile, tn = (g.N + kBTile - 1) / kBTile;
  const double2* A2 = (const double2*)g.A;
  const double* A1 = (const double*)g.A;
  const double2* B = (const double2*)g.B;
  double2* C = (double2*)g.C;
  const double sa = g.conjA ? -1.0 : 1.0, sb = g.conjB ? -1.0 : 1.0;
  constexpr int kPer = kBTile * kBK / kEngT;  // elements of each operand per thread
  __syncthreads();  // the operands may have been written by this workgroup
  const int64_t ntile = (int64_t)g.batch * tm * tn;
  for (int64_t bt0 = 0; bt0 < ntile; bt0 += 2) {
    // the second engine idles (zero operands, no store) past the last tile
    // but keeps to the workgroup's barriers
    const bool valid = bt0 + eng < ntile;
    const int64_t bt = valid ? bt0 + eng : bt0;
    const int64_t b = bt / (tm * tn);
    const int t = (int)(bt - b * tm * tn);
    const int m0 = (t / tn) * kBTile, n0 = (t % tn) * kBTile;
    double pa_r[kPer], pa_i[kPer], pb_r[kPer], pb_i[kPer];
    auto fetch = [&](int k0) {
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const int e = tid + kEngT * i;
        const int x = e % kBTile, kk = e / kBTile;
        double re = 0.0, im = 0.0;
        if (valid && m0 + x < g.M && k0 + kk < g.K) {
          const int64_t off = b * g.sAb + (int64_t)(m0 + x) * g.sAm + (int64_t)(k0 + kk) * g.sAk;
          if (areal) {
            re = A1[off];
          } else {
            const double2 v = A2[off];
            re = v.x;
            im = sa * v.y;
          }
        }
        pa_r[i] = re;
        pa_i[i] = im;
        re = 0.0;
        im = 0.0;
        if (valid && n0 + x < g.N && k0 + kk < g.K) {
          const double2 v = B[b * g.sBb + (int64_t)(k0 + kk) * g.sBk + (int64_t)(n0 + x) * g.sBn];
          re = v.x;
          im = sb * v.y;
        }
        pb_r[i] = re;
        pb_i[i] = im;
      }
    };
    auto stash = [&]() {
#pragma unroll
      for (int i = 0; i < kPer; ++i) {
        const int e = tid + kEngT * i;
        const int x = e % kBTile, kk = e / kBTile;
        s.ta[eng][0][kk][x] = pa_r[i];
        s.ta[eng][1][kk][x] = pa_i[i];
        s.tb[eng][0][kk][x] = pb_r[i];
        s.tb[eng][1][kk][x] = pb_i[i];
      }
    };
    d4b_t cr[2][2], ci[2][2], cq[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) cr[i][j] = ci[i][j] = cq[i][j] = (d4b_t){0.0, 0.0, 0.0, 0.0};
    fetch(0);
    stash();
    __syncthreads();
    for (int k0 = 0; k0 < g.K; k0 += kBK) {
      const bool more = k0 + kBK < g.K;
      if (more) fetch(k0 + kBK);
#pragma unroll
      for (int ks = 0; ks < kBK / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          ar[q] = s.ta[eng][0][kr][wm * 32 + q * 16 + (lane & 15)];
          ai[q] = s.ta[eng][1][kr][wm * 32 + q * 16 + (lane & 15)];
          br[q] = s.tb[eng][0][kr][wn * 32 + q * 16 + (lane & 15)];
          bi[q] = s.tb[eng][1][kr][wn * 32 + q * 16 + (lane & 15)];
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j) {
            cr[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cr[i][j], 0, 0, 0);
            if (areal) {
              ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], ci[i][j], 0, 0, 0);
            } else {
              cq[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], bi[j], cq[i][j], 0, 0, 0);
              ci[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i] + ai[i], br[j] + bi[j], ci[i][j], 0,
                                                                0, 0);
            }
          }
      }
      __syncthreads();
      if (more) {
        stash();
        __syncthreads();
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = m0 + wm * 32 + i * 16 + 4 * r + (lane >> 4);
          const int col = n0 + wn * 32 + j * 16 + (lane & 15);
          if (valid && row < g.M && col < g.N) {
            const double re = areal ? cr[i][j][r] : cr[i][j][r] - cq[i][j][r];
            const double im = areal ? ci[i][j][r] : ci[i][j][r] - cr[i][j][r] - cq[i][j][r];
            double2* p = C + b * g.sCb + (int64_t)row * g.sCm + col;
            double2 v = make_double2(re, im);
            if (g.beta) {
              const double2 o = *p;
              v.x += o.x;
              v.y += o.y;
            }
            *p = v;
          }
        }
  }
  __syncthreads();
}

__device__ GemmCall gcall(int M, int N, int K, int batch, const void* A, int a_kind, int64_t sAm,
                          int64_t sAk, int64_t sAb, int conjA, const void* B, int64_t sBk,
                          int64_t sBn, int64_t sBb, void* C, int64_t sCm, int64_t sCb) {
  GemmCall g{};
  g.M = M; g.N = N; g.K = K; g.batch = batch; g.prec = RENO_PREC_FP64;
  g.A = A; g.a_kind = a_kind; g.sAm = sAm; g.sAk = sAk; g.sAb = sAb; g.conjA = conjA;
  g.B = B; g.b_kind = ST_Z128; g.sBk = sBk; g.sBn = sBn; g.sBb = sBb;
  g.C = C; g.c_kind = ST_Z128; g.sCm = sCm; g.sCn = 1; g.sCb = sCb;
  return g;
}

// The effective operator of one Krylov call (the stages of hop_apply_core).
struct WgHop {
  int nsite;  // 1: one-site (L, Wl, R), 0: bond (L, R)
  int Ml, d, Mr, Dl, Dr;
  const double2* L;
  const double* Wl;
  const double2* R;
  double2* t1;
  double2* t2;
};

__device__ void wg_hop(const WgHop& h, const double2* cin, double2* cout, BatchLds& s) {
  const int64_t Ml = h.Ml, d = h.d, Mr = h.Mr, Dl = h.Dl, Dr = h.Dr;
  if (h.nsite == 1) {
    wg_gemm(gcall(Ml * Dl, d * Mr, Ml, 1, h.L, ST_Z128, Ml, 1, 0, 0, cin, d * Mr, 1, 0, h.t1, d * Mr, 0), s);
    wg_gemm(gcall(d * Dr, Mr, Dl * d, Ml, h.Wl, ST_F64, Dl * d, 1, 0, 0, h.t1, Mr, 1, Dl * d * Mr, h.t2,
                  Mr, d * Dr * Mr), s);
    wg_gemm(gcall(Ml * d, Mr, Dr * Mr, 1, h.t2, ST_Z128, Dr * Mr, 1, 0, 0, h.R, 1, Dr * Mr, 0, cout, Mr, 0),
            s);
  } else {
    wg_gemm(gcall(Ml * Dl, Mr, Ml, 1, h.L, ST_Z128, Ml, 1, 0, 0, cin, Mr, 1, 0, h.t1, Mr, 0), s);
    wg_gemm(gcall(Ml, Mr, Dl * Mr, 1, h.t1, ST_Z128, Dl * Mr, 1, 0, 0, h.R, 1, Dl * Mr, 0, cout, Mr, 0), s);
  }
}

// Krylov scratch of one call
struct WgKry {
  double2* V;      // kcap vectors of n
  double2* w;      // n
  double* alpha;   // kcap
  double* beta;    // kcap
  double2* cnew;   // kcap
  double2* cold;   // kcap
  int kcap;
};

// coef[0..m) = |v0| expm(dt T_m) e_0 from alpha / beta (all waves compute it,
// wave 0 stores); visible to every thread on return
__device__ void wg_tri(const WgKry& k, double v0n, double dt_re, double dt_im, int m, double2* coef,
                       BatchLds& s) {
  const int wave = threadIdx.x >> 6;
  double2* out = wave == 0 ? coef : nullptr;
  if (m <= 64)
    tri_coef_wave<1>(k.alpha, k.beta, v0n, dt_re, dt_im, m, s.J[wave], &s.K[wave], out);
  else if (m <= 128)
    tri_coef_wave<2>(k.alpha, k.beta, v0n, dt_re, dt_im, m, s.J[wave], &s.K[wave], out);
  else
    tri_coef_wave<4>(k.alpha, k.beta, v0n, dt_re, dt_im, m, s.J[wave], &s.K[wave], out);
  __syncthreads();
}

// out[i] = sum_{j<m} coef[j] V[j][i]
__device__ void wg_combine(const WgKry& k, int64_t n, int m, const double2* coef, double2* out) {
  for (int64_t i = threadIdx.x; i < n; i += kBT) {
    double xr = 0.0, xi = 0.0;
    for (int j = 0; j < m; ++j) {
      const double2 c = coef[j];
      const double2 v = k.V[(int64_t)j * n + i];
      xr += v.x * c.x - v.y * c.y;
      xi += v.x * c.y + v.y * c.x;
    }
    out[i] = make_double2(xr, xi);
  }
  __syncthreads();
}

// expm(dt H) v0 -> out by Lanczos without re-orthogonalisation with the
// reference's rule: breakdown beta_j < 100 n eps, the full-space exit at
// j = n - 1, and at every even j > 3 numpy.allclose(res, new_res) (rtol
// 1e-5, atol 1e-8, elementwise).  Returns the matvec count (reference j + 1),
// -1 when kcap vectors did not suffice (out not written), -2 for a zero v0.
__device__ int wg_expm(const WgHop& h, const double2* v0, double2* out, double dt_re, double dt_im,
                       const WgKry& k, BatchLds& s) {
  const int tid = threadIdx.x;
  const int64_t n = (int64_t)h.Ml * h.d * h.Mr;
  double acc = 0.0;
  for (int64_t i = tid; i < n; i += kBT) acc += v0[i].x * v0[i].x + v0[i].y * v0[i].y;
  const double nrm = sqrt(wg_sum(acc, s.red));
  if (!(nrm > 0.0)) return -2;
  const double inv = 1.0 / nrm;
  for (int64_t i = tid; i < n; i += kBT) k.V[i] = make_double2(v0[i].x * inv, v0[i].y * inv);
  const double thr = 100.0 * (double)n * 2.220446049250313e-16;
  bool have_old = false;
  double beta_prev = 0.0;
  for (int j = 0;; ++j) {
    if (j >= k.kcap) return -1;
    double2* vj = k.V + (int64_t)j * n;
    wg_hop(h, vj, k.w, s);
    double a = 0.0;
    for (int64_t i = tid; i < n; i += kBT) a += k.w[i].x * vj[i].x + k.w[i].y * vj[i].y;
    a = wg_sum(a, s.red);
    if (tid == 0) k.alpha[j] = a;
    __syncthreads();
    if (j == n - 1) {
      wg_tri(k, nrm, dt_re, dt_im, j + 1, k.cnew, s);
      wg_combine(k, n, j + 1, k.cnew, out);
      return j + 1;
    }
    const double2* vp = j > 0 ? vj - n : nullptr;
    double b2 = 0.0;
    for (int64_t i = tid; i < n; i += kBT) {
      double2 x = k.w[i];
      x.x -= a * vj[i].x;
      x.y -= a * vj[i].y;
      if (vp) {
        x.x -= beta_prev * vp[i].x;
        x.y -= beta_prev * vp[i].y;
      }
      k.w[i] = x;
      b2 += x.x * x.x + x.y * x.y;
    }
    const double bj = sqrt(wg_sum(b2, s.red));
    if (tid == 0) k.beta[j] = bj;
    __syncthreads();
    if (bj < thr) {
      wg_tri(k, nrm, dt_re, dt_im, j + 1, k.cnew, s);
      wg_combine(k, n, j + 1, k.cnew, out);
      return j + 1;
    }
    if (j > 3 && j % 2 == 0) {
      wg_tri(k, nrm, dt_re, dt_im, j + 1, k.cnew, s);
      if (have_old) {
        // res_new (written to out) against res_old, element by element
        double viol = 0.0;
        for (int64_t i = tid; i < n; i += kBT) {
          double xr = 0.0, xi = 0.0, orr = 0.0, oi = 0.0;
          for (int q = 0; q <= j; ++q) {
            const double2 v = k.V[(int64_t)q * n + i];
            const double2 c = k.cnew[q];
            xr += v.x * c.x - v.y * c.y;
            xi += v.x * c.y + v.y * c.x;
            if (q <= j - 2) {
              const double2 o = k.cold[q];
              orr += v.x * o.x - v.y * o.y;
              oi += v.x * o.y + v.y * o.x;
            }
          }
          out[i] = make_double2(xr, xi);
          const double dr = orr - xr, di = oi - xi;
          if (!(sqrt(dr * dr + di * di) <= 1e-8 + 1e-5 * sqrt(xr * xr + xi * xi))) viol += 1.0;
        }
        if (wg_sum(viol, s.red) == 0.0) return j + 1;
      }
      for (int q = tid; q <= j; q += kBT) k.cold[q] = k.cnew[q];
      __syncthreads();
      have_old = true;
    }
    if (j + 1 >= k.kcap) return -1;
    const double ib = 1.0 / bj;
    double2* vn = vj + n;
    for (int64_t i = tid; i < n; i += kBT) vn[i] = make_double2(k.w[i].x * ib, k.w[i].y * ib);
    __syncthreads();
    beta_prev = bj;
  }
}

__device__ __forceinline__ double2 zmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// Householder QR of the column-major m x n matrix X (ld m) in place, LAPACK
// zgeqr2 conventions (zlarfg: beta = -sign(Re alpha) |(alpha, x)|, real R
// diagonal), then Q (m x k, k = min(m, n)) formed as zung2r does.  Column
// steps are workgroup-wide; within a step every wave takes whole columns,
// its lanes the rows (wave reductions by shuffles: fixed order).
__device__ void wg_householder(double2* X, int m, int n, double2* Q, double2* tau, BatchLds& s) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k = m < n ? m : n;
  __syncthreads();
  for (int j = 0; j < k; ++j) {
    double2* xj = X + (int64_t)j * m;
    double acc = 0.0;
    for (int r = j + 1 + tid; r < m; r += kBT) acc += xj[r].x * xj[r].x + xj[r].y * xj[r].y;
    const double xn2 = wg_sum(acc, s.red);
    const double2 alpha = xj[j];
    double2 t = make_double2(0.0, 0.0);
    double2 scale = make_double2(1.0, 0.0);
    double beta = alpha.x;
    if (!(xn2 == 0.0 && alpha.y == 0.0)) {
      const double nrm = sqrt(alpha.x * alpha.x + alpha.y * alpha.y + xn2);
      beta = alpha.x >= 0.0 ? -nrm : nrm;
      t = make_double2((beta - alpha.x) / beta, -alpha.y / beta);
      const double dr = alpha.x - beta, di = alpha.y;
      const double iden = 1.0 / (dr * dr + di * di);
      scale = make_double2(dr * iden, -di * iden);
    }
    __syncthreads();  // every thread has read alpha before it changes
    for (int r = j + 1 + tid; r < m; r += kBT) xj[r] = zmul(xj[r], scale);
    if (tid == 0) {
      xj[j] = make_double2(beta, 0.0);
      tau[j] = t;
    }
    __syncthreads();
    if (!(t.x == 0.0 && t.y == 0.0)) {
      const double2 ct = make_double2(t.x, -t.y);
      for (int c = j + 1 + wave; c < n; c += kWaves) {
        double2* xc = X + (int64_t)c * m;
        double dr = 0.0, di = 0.0;
        for (int r = j + 1 + lane; r < m; r += 64) {
          const double2 v = xj[r], a = xc[r];
          dr += v.x * a.x + v.y * a.y;  // conj(v) a
          di += v.x * a.y - v.y * a.x;
        }
        dr = wave_sum(dr);
        di = wave_sum(di);
        const double2 dot = make_double2(xc[j].x + dr, xc[j].y + di);
        const double2 f = zmul(ct, dot);
        for (int r = j + 1 + lane; r < m; r += 64) {
          const double2 p = zmul(xj[r], f);
          xc[r] = make_double2(xc[r].x - p.x, xc[r].y - p.y);
        }
        if (lane == 0) xc[j] = make_double2(xc[j].x - f.x, xc[j].y - f.y);
      }
    }
    __syncthreads();
  }
  // Q = H_0 ... H_{k-1} [I; 0] (zung2r)
  for (int64_t e = tid; e < (int64_t)m * k; e += kBT) {
    const int64_t c = e / m, r = e - c * m;
    Q[e] = make_double2(r == c ? 1.0 : 0.0, 0.0);
  }
  __syncthreads();
  for (int i = k - 1; i >= 0; --i) {
    const double2 t = tau[i];
    const double2* xi = X + (int64_t)i * m;
    // H_i on columns c > i of Q: q_c -= tau v (v^H q_c), v = [1; X[i+1:, i]]
    for (int c = i + 1 + wave; c < k; c += kWaves) {
      double2* qc = Q + (int64_t)c * m;
      double dr = 0.0, di = 0.0;
      for (int r = i + 1 + lane; r < m; r += 64) {
        const double2 v = xi[r], a = qc[r];
        dr += v.x * a.x + v.y * a.y;
        di += v.x * a.y - v.y * a.x;
      }
      dr = wave_sum(dr);
      di = wave_sum(di);
      const double2 f = zmul(t, make_double2(qc[i].x + dr, qc[i].y + di));
      for (int r = i + 1 + lane; r < m; r += 64) {
        const double2 p = zmul(xi[r], f);
        qc[r] = make_double2(qc[r].x - p.x, qc[r].y - p.y);
      }
      if (lane == 0) qc[i] = make_double2(qc[i].x - f.x, qc[i].y - f.y);
    }
    __syncthreads();
    double2* qi = Q + (int64_t)i * m;
    for (int r = tid; r < m; r += kBT) {
      if (r < i) qi[r] = make_double2(0.0, 0.0);
      else if (r == i) qi[r] = make_double2(1.0 - t.x, -t.y);
      else qi[r] = make_double2(-(t.x * xi[r].x - t.y * xi[r].y), -(t.x * xi[r].y + t.y * xi[r].x));
    }
    __syncthreads();
  }
}

}  // namespace reno

using namespace reno;

// one state's site update at one sweep position (declared in reno_hip.h)
__global__ __launch_bounds__(kBT) void k_batch_site(const reno_batch_site* __restrict__ sites,
                                                     double dt) {
  __shared__ BatchLds s;
  const reno_batch_site& st = sites[blockIdx.x];
  if (!st.active) return;
  const int tid = threadIdx.x;
  const int64_t Ml = st.Ml, d = st.d, Mr = st.Mr, Dl = st.Dl, Dr = st.Dr, kt = st.kt;
  const int64_t nsite = Ml * d * Mr;
  // scratch: V | w | t1 | t2 | y | z | X | Q | tau | alpha | beta | cnew | cold
  const int64_t tsz = Ml * d * Mr * (Dl > Dr ? Dl : Dr);
  double2* base = (double2*)st.scratch;
  WgKry kr;
  kr.kcap = st.kcap;
  kr.V = base;
  kr.w = kr.V + (int64_t)st.kcap * nsite;
  double2* t1 = kr.w + nsite;
  double2* t2 = t1 + tsz;
  double2* X = t2 + tsz;
  double2* Q = X + nsite;
  double2* tau = Q + nsite;
  kr.alpha = (double*)(tau + nsite);
  kr.beta = kr.alpha + st.kcap;
  kr.cnew = (double2*)(kr.beta + st.kcap);
  kr.cold = kr.cnew + st.kcap;
  int* counts = st.counts;
  // 1. forward one-site step
  WgHop h1{1, st.Ml, st.d, st.Mr, st.Dl, st.Dr, (const double2*)st.L, (const double*)st.Wl,
           (const double2*)st.R, t1, t2};
  const int n1 = wg_expm(h1, (const double2*)st.v0, (double2*)st.kout, 0.0, -0.5 * dt, kr, s);
  if (tid == 0) counts[0] = n1;
  if (n1 < 0 || st.last) return;
  // 2. blocked QR (dir 1, left to right: (Ml d) x Mr) or RQ (dir 0: Ml x (d Mr))
  const double2* coef = (const double2*)st.kout;
  const int64_t rows = st.dir ? Ml * d : Ml;
  const int64_t cols = st.dir ? Mr : d * Mr;
  double2* u = st.dir ? (double2*)st.site_out : (double2*)st.bond;        // rows x kt
  double2* vt = st.dir ? (double2*)st.bond : (double2*)st.site_out;       // kt x cols
  for (int64_t e = tid; e < rows * kt; e += kBT) u[e] = make_double2(0.0, 0.0);
  for (int64_t e = tid; e < kt * cols; e += kBT) vt[e] = make_double2(0.0, 0.0);
  int64_t off = 0;
  for (int b = 0; b < st.nblocks; ++b) {
    const int nr = st.bnr[b], nc = st.bnc[b];
    const int64_t* ri = st.brows[b];
    const int64_t* ci = st.bcols[b];
    const int kb = nr < nc ? nr : nc;
    __syncthreads();
    if (st.dir) {  // QR of the block (nr x nc), column-major X
      for (int64_t e = tid; e < (int64_t)nr * nc; e += kBT) {
        const int64_t j = e / nr, i = e - j * nr;
        X[e] = coef[ri[i] * cols + ci[j]];
      }
      wg_householder(X, nr, nc, Q, tau, s);
      for (int64_t e = tid; e < (int64_t)nr * kb; e += kBT) {
        const int64_t j = e / nr, i = e - j * nr;
        u[ri[i] * kt + off + j] = Q[i + j * nr];
      }
      for (int64_t e = tid; e < (int64_t)nc * kb; e += kBT) {
        const int64_t j = e / nc, c = e - j * nc;
        if (c >= j) vt[(off + j) * cols + ci[c]] = X[j + c * nr];
      }
    } else {  // RQ via the QR of the reversed conjugate transpose (nc x nr)
      const int m = nc;
      for (int64_t e = tid; e < (int64_t)nr * nc; e += kBT) {
        const int64_t j = e / nc, i = e - j * nc;
        const double2 v = coef[ri[nr - 1 - j] * cols + ci[i]];
        X[e] = make_double2(v.x, -v.y);
      }
      wg_householder(X, m, nr, Q, tau, s);
      // u[rows[i], off + a] = conj(R'[k-1-a, nr-1-i]) where i - a <= nr - k
      for (int64_t e = tid; e < (int64_t)nr * kb; e += kBT) {
        const int64_t i = e / kb, a = e - i * kb;
        if (i - a <= nr - kb) {
          const double2 v = X[(kb - 1 - a) + (int64_t)(nr - 1 - i) * m];
          u[ri[i] * kt + off + a] = make_double2(v.x, -v.y);
        }
      }
      // vt[off + a, cols[c]] = conj(Q'[c, k-1-a])
      for (int64_t e = tid; e < (int64_t)nc * kb; e += kBT) {
        const int64_t c = e / kb, a = e - c * kb;
        const double2 v = Q[c + (int64_t)(kb - 1 - a) * m];
        vt[(off + a) * cols + ci[c]] = make_double2(v.x, -v.y);
      }
    }
    off += kb;
  }
  __syncthreads();
  // 3. environment update of the new isometry (the stages of env_update_core)
  const double2* ms = (const double2*)st.site_out;
  if (st.dir) {  // L[i] (kt, Dr, kt) from L[i-1] (Ml, Dl, Ml); site (Ml, d, kt)
    const int64_t Mo = kt;
    wg_gemm(gcall(Ml * Dl, d * Mo, Ml, 1, st.env_in, ST_Z128, Ml, 1, 0, 0, ms, d * Mo, 1, 0, t1, d * Mo, 0), s);
    wg_gemm(gcall(d * Dr, Mo, Dl * d, Ml, st.Wl, ST_F64, Dl * d, 1, 0, 0, t1, Mo, 1, Dl * d * Mo, t2, Mo,
                  d * Dr * Mo), s);
    wg_gemm(gcall(Mo, Dr * Mo, Ml * d, 1, ms, ST_Z128, 1, Mo, 0, 1, t2, Dr * Mo, 1, 0, st.env_out, Dr * Mo, 0),
            s);
  } else {  // R[i] (kt, Dl, kt) from R[i+1] (Mr, Dr, Mr); site (kt, d, Mr)
    const int64_t Mo = kt;
    wg_gemm(gcall(Mo * d, Mr, Mr, Dr, ms, ST_Z128, Mr, 1, 0, 0, st.env_in, 1, Dr * Mr, Mr, t1, Dr * Mr, Mr), s);
    wg_gemm(gcall(Dl * d, Mr, d * Dr, Mo, st.W, ST_F64, d * Dr, 1, 0, 0, t1, Mr, 1, d * Dr * Mr, t2, Mr,
                  Dl * d * Mr), s);
    wg_gemm(gcall(Mo, Mo, d * Mr, Dl, ms, ST_Z128, d * Mr, 1, 0, 1, t2, 1, Dl * d * Mr, d * Mr, st.env_out,
                  Dl * Mo, Mo), s);
  }
  // 4. backward bond step, expm(+i dt/2 H_bond)
  WgHop h0{0, st.dir ? (int)kt : st.Ml, 1, st.dir ? st.Mr : (int)kt, st.D0, st.D0,
           (const double2*)st.L0, nullptr, (const double2*)st.R0, t1, t2};
  const int n0 = wg_expm(h0, (const double2*)st.bond, (double2*)st.bond_out, 0.0, 0.5 * dt, kr, s);
  if (tid == 0) counts[1] = n0;
  if (n0 < 0) return;
  // 5. absorb into the neighbour
  if (st.dir)  // carry (kt x nbr_b) = bond_out (kt x Mr) nbr (Mr x nbr_b)
    wg_gemm(gcall(kt, st.nbr_b, Mr, 1, st.bond_out, ST_Z128, Mr, 1, 0, 0, st.nbr, st.nbr_b, 1, 0, st.carry,
                  st.nbr_b, 0), s);
  else  // carry (nbr_a x kt) = nbr (nbr_a x Ml) bond_out (Ml x kt)
    wg_gemm(gcall(st.nbr_a, kt, Ml, 1, st.nbr, ST_Z128, Ml, 1, 0, 0, st.bond_out, kt, 1, 0, st.carry, kt, 0),
            s);
}

// Frobenius norm of the centre site of every state and its rescaling to 1
// (normalize("mps_only") of a canonical state, mps.py:2025-2059).
__global__ __launch_bounds__(kBT) void k_batch_normalize(double2* const* __restrict__ centre,
                                                         const int64_t* __restrict__ len,
                                                         double* __restrict__ norms) {
  __shared__ double red[kBT];
  double2* c = centre[blockIdx.x];
  const int64_t n = len[blockIdx.x];
  double acc = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += kBT) acc += c[i].x * c[i].x + c[i].y * c[i].y;
  const double nrm = sqrt(wg_sum(acc, red));
  if (threadIdx.x == 0) norms[blockIdx.x] = nrm;
  const double inv = nrm > 0.0 ? 1.0 / nrm : 1.0;
  for (int64_t i = threadIdx.x; i < n; i += kBT) c[i] = make_double2(c[i].x * inv, c[i].y * inv);
}

extern "C" size_t reno_batch_scratch_bytes(int Ml, int d, int Mr, int Dl, int Dr, int kcap) {
  if (Ml <= 0 || d <= 0 || Mr <= 0 || Dl <= 0 || Dr <= 0 || kcap < 2) return 0;
  const int64_t n = (int64_t)Ml * d * Mr;
  const int64_t tsz = n * (Dl > Dr ? Dl : Dr);
  return (size_t)(((int64_t)kcap * n + n + 2 * tsz + 3 * n) * 16 + (int64_t)kcap * (2 * 8 + 2 * 16));
}

extern "C" int reno_batch_site_step(int nstates, const reno_batch_site* sites, double dt, void* stream) {
  if (nstates < 0 || (nstates > 0 && !sites)) {
    set_error("batch_site_step: bad arguments");
    return RENO_EINVAL;
  }
  if (nstates == 0) return RENO_OK;
  hipLaunchKernelGGL(k_batch_site, dim3(nstates), dim3(kBT), 0, (hipStream_t)stream, sites, dt);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_batch_normalize(int nstates, void* const* centre, const int64_t* len, double* norms,
                                    void* stream) {
  if (nstates < 0 || (nstates > 0 && (!centre || !len || !norms))) {
    set_error("batch_normalize: bad arguments");
    return RENO_EINVAL;
  }
  if (nstates == 0) return RENO_OK;
  hipLaunchKernelGGL(k_batch_normalize, dim3(nstates), dim3(kBT), 0, (hipStream_t)stream,
                     (double2* const*)centre, len, norms);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}
