// The small exponential of the Lanczos tridiagonal matrix, shared by the
// Krylov kernels (krylov.hip) and the per-instance persistent kernels of the
// batched sweep (batch.hip).
#pragma once

#include "reno_common.h"

namespace reno {

__device__ __forceinline__ double wave_min(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// Neighbour exchange across the whole wavefront by DPP wave shifts (one VALU
// move per 32-bit half, no LDS): lane l receives lane l - 1 (shr) or
// lane l + 1 (shl); the lane without a source receives 0.
__device__ __forceinline__ double wave_shr1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_shl1(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

constexpr int kChebMax = 512;  // Bessel terms of the Chebyshev branch

// J_k(x), k = 0 .. nmax - 1, for x > 0 by Miller's backward recurrence
// J_{k-1} = (2k / x) J_k - J_{k+1} from a start well above nmax.  Every lane
// of the calling wave runs the same recurrence, so there is no divergent
// branch in the loop: lane l keeps J_{64 q + l} while block q of indices
// goes by, and the wave stores each finished block of 64 at once.  Writes
// the unnormalised values to J[0..nmax) and the normalisation sum
// J_0 + 2 sum_{k>=1} J_{2k} to J[nmax] (J holds nmax + 1 doubles).  The
// overflow test runs every 8 terms: 8 terms grow by less than
// (2 * 512 / 0.05)^8 ~ 3e34, far inside the double range above its 1e200
// threshold.
__device__ inline void bessel_j_miller(double x, int nmax, double* J) {
  const int lane = threadIdx.x & 63;
  const int start = nmax + 24 + (nmax & 1);  // even
  const double ix = 1.0 / x;
  double jp1 = 0.0, jk = 1e-30, norm = 0.0, cur = 0.0;
  // one term: the value of index i (= k - 1), indices descending
  auto step = [&](int i) {
    const double c = (double)(2 * (i + 1)) * ix;
    const double jm1 = c * jk - jp1;
    cur = ((i & 63) == lane && i < nmax) ? jm1 : cur;
    if ((i & 1) == 0) norm = i > 0 ? norm + 2.0 * jm1 : norm + jm1;
    jp1 = jk;
    jk = jm1;
  };
  // after the term of index i, i a multiple of 8
  auto block_end = [&](int i) {
    if (fabs(jk) > 1e200) {  // rescale everything computed so far
      jp1 *= 1e-200;
      jk *= 1e-200;
      norm *= 1e-200;
      cur *= 1e-200;
      for (int t = ((i >> 6) + 1) * 64 + lane; t < nmax; t += 64) J[t] *= 1e-200;
    }
    if ((i & 63) == 0) {
      if (i + lane < nmax) J[i + lane] = cur;
      cur = 0.0;
    }
  };
  int i = start - 1;
  const int top = i & 7;
  for (int t = 0; t <= top; ++t, --i) step(i);
  block_end(i + 1);
  while (i >= 0) {
#pragma unroll
    for (int t = 0; t < 8; ++t) step(i - t);
    i -= 8;
    block_end(i + 1);
  }
  if (lane == 0) J[nmax] = norm;
}

__device__ __forceinline__ int wave_max_int(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}

// coef = |v0| expm(dt T) e_0 for the m x m tridiagonal Lanczos matrix T
// (krylov.py:15-24 forms the same vector from eigh_tridiagonal).  With c the
// centre and r the half-width of T's Gershgorin interval and Y = (T - c)/r:
// * real-time steps (dt = i tau, the TDVP path): the Chebyshev expansion
//   exp(i w y) = J_0(w) + 2 sum_{k>=1} i^k J_k(w) T_k(y), w = tau r, on
//   Y e_0, whose Chebyshev vectors T_k(Y) e_0 are real (three-term
//   recurrence); J_k from bessel_j_miller (normalised by the wave), about
//   w + 25 terms;
// * other dt: exp(dt c) [Taylor_K(h (T - c I))]^q e_0 with q = ceil(|dt| r / 4),
//   h = dt / q: every factor has ||h (T - c I)|| <= 4, and K = 34 terms leave
//   a truncation error below 4^35 / 35! ~ 1e-19.
// One wavefront (lanes threadIdx.x & 63); lane l holds entries
// l*R .. l*R+R-1; neighbours cross lanes by DPP wave shifts.  Fixed
// operation order.  Every thread of the workgroup must call it (it holds
// workgroup barriers); sJ (kChebMax doubles) and sK are the calling wave's
// LDS, coef (m entries) is written unless null.
// tri_coef_core: the same with this lane's entries of T already in
// registers (a = alpha, bl / br = the off-diagonal to the left / right,
// zero past m), for callers that load them early.
template <int R>
__device__ __forceinline__ void tri_coef_core(double (&a)[R], double (&bl)[R], double (&br)[R],
                                              double v0n, double dt_re, double dt_im, int m,
                                              double* sJ, int* sK, double2* __restrict__ coef);

template <int R>
__device__ __forceinline__ void tri_coef_wave(const double* __restrict__ alpha,
                                              const double* __restrict__ beta, double v0n,
                                              double dt_re, double dt_im, int m, double* sJ,
                                              int* sK, double2* __restrict__ coef) {
  const int lane = threadIdx.x & 63;
  double a[R], bl[R], br[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane * R + r;
    const bool in = i < m;
    a[r] = in ? alpha[i] : 0.0;
    bl[r] = (in && i > 0) ? beta[i - 1] : 0.0;
    br[r] = (in && i + 1 < m) ? beta[i] : 0.0;
  }
  tri_coef_core<R>(a, bl, br, v0n, dt_re, dt_im, m, sJ, sK, coef);
}

template <int R>
__device__ __forceinline__ void tri_coef_core(double (&a)[R], double (&bl)[R], double (&br)[R],
                                              double v0n, double dt_re, double dt_im, int m,
                                              double* sJ, int* sK, double2* __restrict__ coef) {
  const int lane = threadIdx.x & 63;
  double lo = 1e300, hi = -1e300;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane * R + r;
    const bool in = i < m;
    if (in) {
      lo = fmin(lo, a[r] - fabs(bl[r]) - fabs(br[r]));
      hi = fmax(hi, a[r] + fabs(bl[r]) + fabs(br[r]));
    }
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  const double c = 0.5 * (lo + hi);
  const double rad = 0.5 * (hi - lo);
#pragma unroll
  for (int r = 0; r < R; ++r) a[r] -= c;
  const double w = dt_im * rad;
  const double x = fabs(w);
  const int nterm = (int)(x + 12.0 * cbrt(x) + 32.0);
  if (dt_re == 0.0 && x >= 0.05 && nterm < kChebMax - 24) {
    bessel_j_miller(x, nterm, sJ);
    __syncthreads();
    // normalise across the wave; K = number of leading terms to keep (past
    // the last one every |J_k| < 1e-18)
    {
      const double in = 1.0 / sJ[nterm];
      int keep = 1;
      for (int k = lane; k < nterm; k += 64) {
        const double v = sJ[k] * in;
        sJ[k] = v;
        if (fabs(v) >= 1e-18) keep = k + 1;
      }
      keep = wave_max_int(keep);
      if (lane == 0) *sK = keep;
    }
    __syncthreads();
    const int K = *sK;
    const double ir = 1.0 / rad;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      a[r] *= ir;
      bl[r] *= ir;
      br[r] *= ir;
    }
    // i^k for w > 0, (-i)^k for w < 0: phase index k mod 4 and the sign
    const double s = w < 0.0 ? -1.0 : 1.0;
    double yp[R], yc[R], accr[R], acci[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      yp[r] = (lane == 0 && r == 0) ? 1.0 : 0.0;  // T_0(Y) e_0
      yc[r] = 0.0;
      accr[r] = sJ[0] * yp[r];
      acci[r] = 0.0;
    }
    // one Chebyshev term k with phase index ph = k mod 4 and J_k = jk
    auto term = [&](int k, int ph, double jk) {
      const double lr = wave_shr1(yp[R - 1]);
      const double rr = wave_shl1(yp[0]);
      double pv = lr;
      double t[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double nv = r + 1 < R ? yp[r + 1] : rr;
        t[r] = a[r] * yp[r] + bl[r] * pv + br[r] * nv;  // Y T_{k-1}(Y) e_0
        pv = yp[r];
      }
      const double ck = 2.0 * jk;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        // T_k = 2 Y T_{k-1} - T_{k-2}  (T_1 = Y T_0)
        const double yn = k == 1 ? t[r] : 2.0 * t[r] - yc[r];
        yc[r] = yp[r];
        yp[r] = yn;
        const double v = ck * yn;
        if (ph == 0) accr[r] += v;
        else if (ph == 1) acci[r] += s * v;
        else if (ph == 2) accr[r] -= v;
        else acci[r] -= s * v;
      }
    };
    // k = 1, then blocks of four from k = 2 (phases 2, 3, 0, 1 fixed at
    // compile time, the four J_k loaded together), then the rest
    int k = 1;
    if (K > 1) term(1, 1, sJ[1]);
    for (k = 2; k + 3 < K; k += 4) {
      const double j0 = sJ[k], j1 = sJ[k + 1], j2 = sJ[k + 2], j3 = sJ[k + 3];
      term(k, 2, j0);
      term(k + 1, 3, j1);
      term(k + 2, 0, j2);
      term(k + 3, 1, j3);
    }
    for (; k < K; ++k) term(k, k & 3, sJ[k]);
    const double pr = v0n * cos(dt_im * c), pi = v0n * sin(dt_im * c);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int i = lane * R + r;
      if (coef && i < m) coef[i] = make_double2(accr[r] * pr - acci[r] * pi, accr[r] * pi + acci[r] * pr);
    }
    return;
  }
  const double adt = sqrt(dt_re * dt_re + dt_im * dt_im);
  int q = (int)ceil(adt * rad / 4.0);
  if (q < 1) q = 1;
  const double hr = dt_re / q, hi_ = dt_im / q;
  double yr[R], yi[R], xr[R], xi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    yr[r] = (lane == 0 && r == 0) ? v0n : 0.0;
    yi[r] = 0.0;
  }
  for (int step = 0; step < q; ++step) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      xr[r] = yr[r];
      xi[r] = yi[r];
    }
    for (int k = 1; k <= 34; ++k) {
      // neighbours across lanes: the entry before this lane's first and the
      // one after its last (zero-weighted at the ends via bl / br)
      const double lr = wave_shr1(xr[R - 1]), li = wave_shr1(xi[R - 1]);
      const double rr = wave_shl1(xr[0]), ri = wave_shl1(xi[0]);
      const double fr = hr / (double)k, fi = hi_ / (double)k;
      double pr = lr, pi = li;  // old x[r-1]
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double nr = r + 1 < R ? xr[r + 1] : rr;
        const double ni = r + 1 < R ? xi[r + 1] : ri;
        const double tr = a[r] * xr[r] + bl[r] * pr + br[r] * nr;
        const double ti = a[r] * xi[r] + bl[r] * pi + br[r] * ni;
        pr = xr[r];
        pi = xi[r];
        xr[r] = tr * fr - ti * fi;
        xi[r] = tr * fi + ti * fr;
        yr[r] += xr[r];
        yi[r] += xi[r];
      }
    }
  }
  const double e = exp(dt_re * c);
  const double pr = e * cos(dt_im * c), pi = e * sin(dt_im * c);
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane * R + r;
    if (coef && i < m) coef[i] = make_double2(yr[r] * pr - yi[r] * pi, yr[r] * pi + yi[r] * pr);
  }
}

}  // namespace reno
