// DMRG local-eigenproblem helpers (reference mps/gs.py:302-477).
//
// The iterative eigensolver applies H_eff through reno_hop_apply; these two
// kernels supply what it needs besides the matvec:
//   * the diagonal of the one-site H_eff, the Davidson preconditioner
//     (gs.py:424-436):  hdiag[a,d,l] = sum_{b,f} L[a,b,a] W[b,d,d,f] R[l,f,l];
//   * the dense H_eff restricted to the allowed quantum-number entries, for
//     the direct diagonalisation of small sites (gs.py:302-330, used when
//     fewer than 1000 coefficients are allowed, gs.py:245):
//       H[i][j] = sum_{b,f} L[a,b,c] W[b,d,e,f] R[l,f,k],
//       (a,d,l) = idx[i], (c,e,k) = idx[j] (flat indices of the site tensor).
// Two-site problems use the same kernels with the two MPO sites merged into
// one of physical dimension d0*d1.  FP64 (complex128 L/R, float64 W).
#include "reno_common.h"

namespace reno {

__global__ void k_heff_diag(int Ml, int d, int Mr, int Dl, int Dr, const double2* __restrict__ L,
                            const double* __restrict__ W, const double2* __restrict__ R,
                            double2* __restrict__ out) {
  const int64_t n = (int64_t)Ml * d * Mr;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int a = (int)(t / ((int64_t)d * Mr));
    const int rem = (int)(t - (int64_t)a * d * Mr);
    const int dd = rem / Mr, l = rem - dd * Mr;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < Dl; ++b) {
      const double2 lv = L[((int64_t)a * Dl + b) * Ml + a];
      for (int f = 0; f < Dr; ++f) {
        const double w = W[(((int64_t)b * d + dd) * d + dd) * Dr + f];
        if (w == 0.0) continue;
        const double2 rv = R[((int64_t)l * Dr + f) * Mr + l];
        // (lv * rv) * w
        sr += w * (lv.x * rv.x - lv.y * rv.y);
        si += w * (lv.x * rv.y + lv.y * rv.x);
      }
    }
    out[t] = make_double2(sr, si);
  }
}

__global__ void k_heff_dense(int Ml, int d, int Mr, int Dl, int Dr, const double2* __restrict__ L,
                             const double* __restrict__ W, const double2* __restrict__ R,
                             const int64_t* __restrict__ idx, int nidx,
                             double2* __restrict__ out) {
  const int64_t total = (int64_t)nidx * nidx;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / nidx), j = (int)(t - (int64_t)i * nidx);
    const int64_t ri = idx[i], cj = idx[j];
    const int a = (int)(ri / ((int64_t)d * Mr));
    const int ra = (int)(ri - (int64_t)a * d * Mr);
    const int dd = ra / Mr, l = ra - dd * Mr;
    const int c = (int)(cj / ((int64_t)d * Mr));
    const int rc = (int)(cj - (int64_t)c * d * Mr);
    const int e = rc / Mr, k = rc - e * Mr;
    double sr = 0.0, si = 0.0;
    for (int b = 0; b < Dl; ++b) {
      const double2 lv = L[((int64_t)a * Dl + b) * Ml + c];
      if (lv.x == 0.0 && lv.y == 0.0) continue;
      for (int f = 0; f < Dr; ++f) {
        const double w = W[(((int64_t)b * d + dd) * d + e) * Dr + f];
        if (w == 0.0) continue;
        const double2 rv = R[((int64_t)l * Dr + f) * Mr + k];
        sr += w * (lv.x * rv.x - lv.y * rv.y);
        si += w * (lv.x * rv.y + lv.y * rv.x);
      }
    }
    out[t] = make_double2(sr, si);
  }
}

static int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 8192) b = 8192;
  return (int)b;
}

}  // namespace reno

using namespace reno;

extern "C" int reno_heff_diag(int Ml, int d, int Mr, int Dl, int Dr, const void* L, const void* W,
                              const void* R, void* out, void* stream) {
  if (Ml < 1 || d < 1 || Mr < 1 || Dl < 1 || Dr < 1 || !L || !W || !R || !out) {
    set_error("heff_diag: bad arguments Ml=%d d=%d Mr=%d Dl=%d Dr=%d", Ml, d, Mr, Dl, Dr);
    return RENO_EINVAL;
  }
  const int64_t n = (int64_t)Ml * d * Mr;
  hipLaunchKernelGGL(k_heff_diag, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, Ml, d, Mr,
                     Dl, Dr, (const double2*)L, (const double*)W, (const double2*)R, (double2*)out);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_heff_dense(int Ml, int d, int Mr, int Dl, int Dr, const void* L, const void* W,
                               const void* R, const int64_t* idx, int nidx, void* out,
                               void* stream) {
  if (Ml < 1 || d < 1 || Mr < 1 || Dl < 1 || Dr < 1 || !L || !W || !R || !out || nidx < 0 ||
      (nidx > 0 && !idx)) {
    set_error("heff_dense: bad arguments Ml=%d d=%d Mr=%d Dl=%d Dr=%d nidx=%d", Ml, d, Mr, Dl, Dr,
              nidx);
    return RENO_EINVAL;
  }
  if (nidx == 0) return RENO_OK;
  const int64_t total = (int64_t)nidx * nidx;
  hipLaunchKernelGGL(k_heff_dense, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream, Ml, d,
                     Mr, Dl, Dr, (const double2*)L, (const double*)W, (const double2*)R, idx, nidx,
                     (double2*)out);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}
