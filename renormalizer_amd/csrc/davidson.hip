// Vector algebra of the Davidson-Liu eigensolver of the DMRG local problem
// (reference lib/davidson/davidson.py:154-441 on gs.py:380-407): the basis
// V and its images AV stay on the device as rows of length n (row i at
// V + i * ldv, complex128); the host keeps only the small subspace matrix
// and its eigen-decomposition.
//
//   reno_dav_orth     x <- x - V^T (conj(V) x), twice (classical Gram-Schmidt
//                     with one re-orthogonalisation pass), |x|^2 and, above a
//                     threshold, x / |x|;
//   reno_dav_dots     conj(V) y: a column of the subspace matrix V^H A V;
//   reno_dav_ritz     Ritz vectors X = U^T V, their images AX = U^T AV, the
//                     residuals R = AX - e X and |R_j|^2 in one pass over V, AV;
//   reno_dav_precond  the diagonal preconditioner r / (hdiag - e + 1e-4) on
//                     the allowed entries (reference gs.py:424-436).
//
// Every sum over n runs as per-block partial sums in a fixed order followed
// by one fold kernel, so results are reproducible bit for bit.  All of it is
// bandwidth-bound (one to four passes over k vectors of n complex numbers).
#include "reno_common.h"

namespace reno {

constexpr int kDavThreads = 256;
constexpr int kDavMaxBlocks = 256;  // partial sums per reduced value
constexpr int kDavMaxK = 64;        // basis vectors / roots per call

static int dav_blocks(int64_t n) {
  int64_t b = (n + 4 * kDavThreads - 1) / (4 * kDavThreads);
  if (b < 1) b = 1;
  if (b > kDavMaxBlocks) b = kDavMaxBlocks;
  return (int)b;
}

// part[(i * nb + bx) * 2 + {0,1}] = partial of sum_r conj(V[i][r]) y[r] over
// the rows r = bx * chunk ... of block bx (grid nb x k)
__global__ __launch_bounds__(kDavThreads) void k_dav_dot_part(int64_t n, const double2* __restrict__ V,
                                                              int64_t ldv, const double2* __restrict__ y,
                                                              double* __restrict__ part) {
  __shared__ double sh[4];
  const int i = blockIdx.y, nb = gridDim.x;
  const int64_t chunk = (n + nb - 1) / nb;
  const int64_t r0 = blockIdx.x * chunk;
  const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
  const double2* v = V + (int64_t)i * ldv;
  double sr = 0.0, si = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kDavThreads) {
    const double2 a = v[r], b = y[r];
    sr += a.x * b.x + a.y * b.y;
    si += a.x * b.y - a.y * b.x;
  }
  sr = block_sum256(sr, sh);
  si = block_sum256(si, sh);
  if (threadIdx.x == 0) {
    part[((int64_t)i * nb + blockIdx.x) * 2] = sr;
    part[((int64_t)i * nb + blockIdx.x) * 2 + 1] = si;
  }
}

// out[i] = sum_b part[i][b] (block i), in the order b = tid, tid + 256, ...
// then the block sum
__global__ __launch_bounds__(kDavThreads) void k_dav_fold(const double* __restrict__ part, int nb,
                                                          double2* __restrict__ out) {
  __shared__ double sh[4];
  const int i = blockIdx.x;
  double sr = 0.0, si = 0.0;
  for (int b = threadIdx.x; b < nb; b += kDavThreads) {
    sr += part[((int64_t)i * nb + b) * 2];
    si += part[((int64_t)i * nb + b) * 2 + 1];
  }
  sr = block_sum256(sr, sh);
  si = block_sum256(si, sh);
  if (threadIdx.x == 0) out[i] = make_double2(sr, si);
}

// x[r] -= sum_i V[i][r] c[i]
__global__ __launch_bounds__(kDavThreads) void k_dav_sub(int64_t n, int k, const double2* __restrict__ V,
                                                         int64_t ldv, const double2* __restrict__ c,
                                                         double2* __restrict__ x) {
  __shared__ double2 sc[kDavMaxK];
  for (int i = threadIdx.x; i < k; i += kDavThreads) sc[i] = c[i];
  __syncthreads();
  for (int64_t r = blockIdx.x * (int64_t)kDavThreads + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * kDavThreads) {
    double2 a = x[r];
    for (int i = 0; i < k; ++i) {
      const double2 v = V[(int64_t)i * ldv + r], s = sc[i];
      a.x -= v.x * s.x - v.y * s.y;
      a.y -= v.x * s.y + v.y * s.x;
    }
    x[r] = a;
  }
}

// nrm2[0] = Re dot[0] (= |x|^2); x <- x / |x| when nrm2 >= thresh
__global__ __launch_bounds__(kDavThreads) void k_dav_scale(int64_t n, const double2* __restrict__ dot,
                                                           double thresh, double* __restrict__ nrm2,
                                                           double2* __restrict__ x) {
  const double s2 = dot[0].x;
  if (blockIdx.x == 0 && threadIdx.x == 0) nrm2[0] = s2;
  if (!(s2 >= thresh)) return;
  const double inv = 1.0 / sqrt(s2);
  for (int64_t r = blockIdx.x * (int64_t)kDavThreads + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * kDavThreads) {
    const double2 a = x[r];
    x[r] = make_double2(a.x * inv, a.y * inv);
  }
}

// X[j] = sum_i U[i][j] V[i], AX[j] = sum_i U[i][j] AV[i], R[j] = AX[j] - e[j] X[j]
// and the partial sums of |R[j]|^2 (grid nb; rows r = bx * chunk ...)
__global__ __launch_bounds__(kDavThreads) void k_dav_ritz(int64_t n, int m, int k, const double2* __restrict__ V,
                                                          const double2* __restrict__ AV, int64_t ldv,
                                                          const double2* __restrict__ U,
                                                          const double* __restrict__ e,
                                                          double2* __restrict__ X, double2* __restrict__ AX,
                                                          double2* __restrict__ R, double* __restrict__ part) {
  __shared__ double2 sU[kDavMaxK * 16];
  __shared__ double sh[4];
  for (int t = threadIdx.x; t < m * k; t += kDavThreads) sU[t] = U[t];
  __syncthreads();
  const int nb = gridDim.x;
  const int64_t chunk = (n + nb - 1) / nb;
  const int64_t r0 = blockIdx.x * chunk;
  const int64_t r1 = r0 + chunk < n ? r0 + chunk : n;
  for (int j = 0; j < k; ++j) {
    const double ej = e[j];
    double acc = 0.0;
    for (int64_t r = r0 + threadIdx.x; r < r1; r += kDavThreads) {
      double xr = 0.0, xi = 0.0, ar = 0.0, ai = 0.0;
      for (int i = 0; i < m; ++i) {
        const double2 u = sU[i * k + j];
        const double2 v = V[(int64_t)i * ldv + r], w = AV[(int64_t)i * ldv + r];
        xr += u.x * v.x - u.y * v.y;
        xi += u.x * v.y + u.y * v.x;
        ar += u.x * w.x - u.y * w.y;
        ai += u.x * w.y + u.y * w.x;
      }
      const double rr = ar - ej * xr, ri = ai - ej * xi;
      X[(int64_t)j * n + r] = make_double2(xr, xi);
      AX[(int64_t)j * n + r] = make_double2(ar, ai);
      R[(int64_t)j * n + r] = make_double2(rr, ri);
      acc += rr * rr + ri * ri;
    }
    acc = block_sum256(acc, sh);
    if (threadIdx.x == 0) {
      part[((int64_t)j * nb + blockIdx.x) * 2] = acc;
      part[((int64_t)j * nb + blockIdx.x) * 2 + 1] = 0.0;
    }
  }
}

__global__ void k_dav_real(const double2* __restrict__ z, int k, double* __restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < k) out[j] = z[j].x;
}

__global__ __launch_bounds__(kDavThreads) void k_dav_precond(int64_t n, const double2* __restrict__ r,
                                                             const double* __restrict__ hdiag,
                                                             const unsigned char* __restrict__ mask, double e,
                                                             double2* __restrict__ out) {
  for (int64_t t = blockIdx.x * (int64_t)kDavThreads + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * kDavThreads) {
    double2 o = make_double2(0.0, 0.0);
    if (mask[t]) {
      const double2 a = r[t];
      const double inv = 1.0 / (hdiag[t] - e + 1e-4);
      o = make_double2(a.x * inv, a.y * inv);
    }
    out[t] = o;
  }
}

static int grid_rows(int64_t n) {
  int64_t b = (n + kDavThreads - 1) / kDavThreads;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

struct DavWs {
  double* part;
  double2* c;
};

static size_t dav_ws_need(int k) {
  const int kk = k < 1 ? 1 : k;
  return align_up((size_t)kk * kDavMaxBlocks * 2 * sizeof(double)) + align_up((size_t)kDavMaxK * 16);
}

static int dav_ws(void* ws, size_t ws_bytes, int k, DavWs* out) {
  if (!ws || ws_bytes < dav_ws_need(k)) {
    set_error("dav: workspace %zu < %zu", ws_bytes, dav_ws_need(k));
    return RENO_ENOSPACE;
  }
  char* p = (char*)ws;
  out->part = (double*)p;
  p += align_up((size_t)(k < 1 ? 1 : k) * kDavMaxBlocks * 2 * sizeof(double));
  out->c = (double2*)p;
  return RENO_OK;
}

// conj(V) y into out (k values) through the workspace partials
static int dav_dots(hipStream_t s, int64_t n, int k, const double2* V, int64_t ldv, const double2* y,
                    double2* out, const DavWs& w) {
  const int nb = dav_blocks(n);
  hipLaunchKernelGGL(k_dav_dot_part, dim3(nb, k), dim3(kDavThreads), 0, s, n, V, ldv, y, w.part);
  hipLaunchKernelGGL(k_dav_fold, dim3(k), dim3(kDavThreads), 0, s, (const double*)w.part, nb, out);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

}  // namespace reno

using namespace reno;

extern "C" size_t reno_dav_workspace_bytes(int64_t n, int k) {
  (void)n;
  return dav_ws_need(k > kDavMaxK ? kDavMaxK : k);
}

extern "C" int reno_dav_orth(int64_t n, int k, const void* V, int64_t ldv, void* x, double thresh, void* nrm2,
                             void* ws, size_t ws_bytes, void* stream) {
  if (n <= 0 || k < 0 || k > kDavMaxK || (k > 0 && (!V || ldv < n)) || !x || !nrm2) {
    set_error("dav_orth: bad arguments (n %lld, k %d, ldv %lld)", (long long)n, k, (long long)ldv);
    return RENO_EINVAL;
  }
  DavWs w;
  RENO_TRY(dav_ws(ws, ws_bytes, k, &w));
  hipStream_t s = (hipStream_t)stream;
  double2* xz = (double2*)x;
  const double2* Vz = (const double2*)V;
  for (int pass = 0; pass < 2 && k > 0; ++pass) {
    RENO_TRY(dav_dots(s, n, k, Vz, ldv, xz, w.c, w));
    hipLaunchKernelGGL(k_dav_sub, dim3(grid_rows(n)), dim3(kDavThreads), 0, s, n, k, Vz, ldv,
                       (const double2*)w.c, xz);
    RENO_CHECK_LAUNCH();
  }
  RENO_TRY(dav_dots(s, n, 1, xz, n, xz, w.c, w));
  hipLaunchKernelGGL(k_dav_scale, dim3(grid_rows(n)), dim3(kDavThreads), 0, s, n, (const double2*)w.c, thresh,
                     (double*)nrm2, xz);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_dav_dots(int64_t n, int k, const void* V, int64_t ldv, const void* y, void* out, void* ws,
                             size_t ws_bytes, void* stream) {
  if (n <= 0 || k <= 0 || k > kDavMaxK || !V || ldv < n || !y || !out) {
    set_error("dav_dots: bad arguments (n %lld, k %d, ldv %lld)", (long long)n, k, (long long)ldv);
    return RENO_EINVAL;
  }
  DavWs w;
  RENO_TRY(dav_ws(ws, ws_bytes, k, &w));
  return dav_dots((hipStream_t)stream, n, k, (const double2*)V, ldv, (const double2*)y, (double2*)out, w);
}

extern "C" int reno_dav_ritz(int64_t n, int m, int k, const void* V, const void* AV, int64_t ldv, const void* U,
                             const void* e, void* X, void* AX, void* R, void* rnorm2, void* ws, size_t ws_bytes,
                             void* stream) {
  if (n <= 0 || m <= 0 || m > kDavMaxK || k <= 0 || k > m || m * k > kDavMaxK * 16 || !V || !AV || ldv < n ||
      !U || !e || !X || !AX || !R || !rnorm2) {
    set_error("dav_ritz: bad arguments (n %lld, m %d, k %d, ldv %lld)", (long long)n, m, k, (long long)ldv);
    return RENO_EINVAL;
  }
  DavWs w;
  RENO_TRY(dav_ws(ws, ws_bytes, k, &w));
  hipStream_t s = (hipStream_t)stream;
  const int nb = dav_blocks(n);
  hipLaunchKernelGGL(k_dav_ritz, dim3(nb), dim3(kDavThreads), 0, s, n, m, k, (const double2*)V,
                     (const double2*)AV, ldv, (const double2*)U, (const double*)e, (double2*)X, (double2*)AX,
                     (double2*)R, w.part);
  hipLaunchKernelGGL(k_dav_fold, dim3(k), dim3(kDavThreads), 0, s, (const double*)w.part, nb, w.c);
  hipLaunchKernelGGL(k_dav_real, dim3(1), dim3(kDavMaxK), 0, s, (const double2*)w.c, k, (double*)rnorm2);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_dav_precond(int64_t n, const void* r, const void* hdiag, const void* mask, double e, void* out,
                                void* stream) {
  if (n <= 0 || !r || !hdiag || !mask || !out) {
    set_error("dav_precond: bad arguments");
    return RENO_EINVAL;
  }
  hipLaunchKernelGGL(k_dav_precond, dim3(grid_rows(n)), dim3(kDavThreads), 0, (hipStream_t)stream, n,
                     (const double2*)r, (const double*)hdiag, (const unsigned char*)mask, e, (double2*)out);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}
