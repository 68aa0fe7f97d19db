// Quantum-number-blocked QR / RQ of the site coefficient matrix
// (svd_qn QR branch, reference svd_qn.py:99-240 with QR=True, economic).
//
// A row-major (rows x cols) block is, bit for bit, the column-major
// (cols x rows) matrix X = block^T, so no transposes are needed:
//   QR of block  <=>  LQ of X   (rocsolver_zgelqf + zunglq)
//   RQ of block  <=>  QL of X   (rocsolver_zgeqlf + zungql)
// Householder factorizations as in LAPACK: the triangular factor has a real
// diagonal, so the result matches scipy.linalg.qr / rq up to the sign of each
// column/row (and to rounding).  The block gather from the QN index sets and
// the scatter into the zero-filled u / v (blockrecover, svd_qn.py:88-96) are
// small device kernels.
#include <rocsolver/rocsolver.h>

#include <map>
#include <mutex>

#include "reno_common.h"

namespace reno {

static std::mutex g_handle_mu;
static std::map<int, rocblas_handle> g_handles;

static int get_handle(rocblas_handle* out) {
  int dev = 0;
  RENO_HIP_CHECK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_handle_mu);
  auto it = g_handles.find(dev);
  if (it != g_handles.end()) {
    *out = it->second;
    return RENO_OK;
  }
  rocblas_handle h;
  if (rocblas_create_handle(&h) != rocblas_status_success) {
    set_error("rocblas_create_handle failed on device %d", dev);
    return RENO_ESOLVER;
  }
  g_handles[dev] = h;
  *out = h;
  return RENO_OK;
}

__global__ void k_zero_z(double2* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    p[i] = make_double2(0.0, 0.0);
  }
}

// X[i*nc + j] = coef[rows[i]*cols_total + cols[j]]
__global__ void k_gather(const double2* __restrict__ coef, int64_t cols_total,
                         const int64_t* __restrict__ rows, const int64_t* __restrict__ cols,
                         int nr, int nc, double2* __restrict__ X) {
  const int64_t total = (int64_t)nr * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = t / nc, j = t - (t / nc) * nc;
    X[t] = coef[rows[i] * cols_total + cols[j]];
  }
}

// dst[idx[i]*di + (off + j)*dj] = X[i*si + j*sj + x0] if cond(i,j) else 0
// for i < ni, j < k.  cond: mode 0 -> i >= j + shift ; mode 1 -> i <= j + shift;
// mode 2 -> always.
__global__ void k_scatter(const double2* __restrict__ X, int64_t x0, int64_t si, int64_t sj,
                          const int64_t* __restrict__ idx, int ni, int k, int64_t di,
                          int64_t dj, int64_t off, int mode, int shift,
                          double2* __restrict__ dst) {
  const int64_t total = (int64_t)ni * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / k), j = (int)(t - (t / k) * k);
    bool ok = true;
    if (mode == 0) ok = (i >= j + shift);
    else if (mode == 1) ok = (i <= j + shift);
    const double2 v = ok ? X[x0 + (int64_t)i * si + (int64_t)j * sj] : make_double2(0.0, 0.0);
    dst[idx[i] * di + (off + j) * dj] = v;
  }
}

static inline int grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 4096) b = 4096;
  return (int)b;
}

#define RENO_SOLVER_CHECK(expr)                                                \
  do {                                                                         \
    rocblas_status _st = (expr);                                               \
    if (_st != rocblas_status_success) {                                       \
      ::reno::set_error("%s:%d %s -> rocblas_status %d", __FILE__, __LINE__,   \
                        #expr, (int)_st);                                      \
      return RENO_ESOLVER;                                                     \
    }                                                                          \
  } while (0)

}  // namespace reno

using namespace reno;

extern "C" size_t reno_qr_workspace_bytes(int rows, int cols, int nblocks,
                                          const reno_qn_block* blocks) {
  (void)rows;
  (void)cols;
  size_t xmax = 0, kmax = 0;
  for (int b = 0; b < nblocks; ++b) {
    const size_t x = (size_t)blocks[b].nrows * blocks[b].ncols;
    const size_t k = (size_t)(blocks[b].nrows < blocks[b].ncols ? blocks[b].nrows : blocks[b].ncols);
    if (x > xmax) xmax = x;
    if (k > kmax) kmax = k;
  }
  return align_up(xmax * 16) + align_up((kmax + 1) * 16);
}

extern "C" int reno_block_qr(int system, int rows, int cols, const void* coef, int nblocks,
                             const reno_qn_block* blocks, void* u, void* v, int v_transposed,
                             void* ws, size_t ws_bytes, void* stream) {
  if (system != 0 && system != 1) {
    set_error("block_qr: system %d (0 = L/QR, 1 = R/RQ)", system);
    return RENO_EINVAL;
  }
  if (rows <= 0 || cols <= 0 || !coef || !u || !v || nblocks < 0 || (nblocks && !blocks)) {
    set_error("block_qr: bad arguments");
    return RENO_EINVAL;
  }
  if (nblocks == 0) {
    set_error("block_qr: Invalid quantum number (no block)");
    return RENO_EQN;
  }
  int64_t kt = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (blocks[b].nrows <= 0 || blocks[b].ncols <= 0 || blocks[b].nrows > rows ||
        blocks[b].ncols > cols || !blocks[b].rows || !blocks[b].cols) {
      set_error("block_qr: bad block %d (%d x %d)", b, blocks[b].nrows, blocks[b].ncols);
      return RENO_EINVAL;
    }
    kt += blocks[b].nrows < blocks[b].ncols ? blocks[b].nrows : blocks[b].ncols;
  }
  const size_t need = reno_qr_workspace_bytes(rows, cols, nblocks, blocks);
  if (!ws || ws_bytes < need) {
    set_error("block_qr: workspace %zu < %zu", ws_bytes, need);
    return RENO_ENOSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  rocblas_handle h;
  RENO_TRY(get_handle(&h));
  RENO_SOLVER_CHECK(rocblas_set_stream(h, s));

  size_t xmax = 0;
  for (int b = 0; b < nblocks; ++b) {
    const size_t x = (size_t)blocks[b].nrows * blocks[b].ncols;
    if (x > xmax) xmax = x;
  }
  double2* X = (double2*)ws;
  double2* tau = (double2*)((char*)ws + align_up(xmax * 16));
  rocblas_double_complex* Xr = (rocblas_double_complex*)X;
  rocblas_double_complex* taur = (rocblas_double_complex*)tau;

  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)rows * kt)), dim3(256), 0, s,
                     (double2*)u, (int64_t)rows * kt);
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)cols * kt)), dim3(256), 0, s,
                     (double2*)v, (int64_t)cols * kt);
  RENO_CHECK_LAUNCH();

  // u is (rows x kt); v is (cols x kt), or its transpose (kt x cols)
  const int64_t vdi = v_transposed ? 1 : kt;
  const int64_t vdj = v_transposed ? (int64_t)cols : 1;
  int64_t off = 0;
  for (int b = 0; b < nblocks; ++b) {
    const int nr = blocks[b].nrows, nc = blocks[b].ncols;
    const int k = nr < nc ? nr : nc;
    hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)nr * nc)), dim3(256), 0, s,
                       (const double2*)coef, (int64_t)cols, blocks[b].rows, blocks[b].cols, nr,
                       nc, X);
    RENO_CHECK_LAUNCH();
    if (system == 0) {
      // block = Q R  <=>  X (nc x nr, col-major) = L Qlq,  Q = Qlq^T, R = L^T
      RENO_SOLVER_CHECK(rocsolver_zgelqf(h, nc, nr, Xr, nc, taur));
      // v = R^T = L: v[cols[i], off+j] = X[i + j*nc] for i >= j
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)0, (int64_t)1, (int64_t)nc, blocks[b].cols,
                         nc, k, vdi, vdj, off, 0, 0, (double2*)v);
      RENO_CHECK_LAUNCH();
      RENO_SOLVER_CHECK(rocsolver_zunglq(h, k, nr, k, Xr, nc, taur));
      // u = Q = Qlq^T: u[rows[i], off+j] = X[j + i*nc]
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)0, (int64_t)nc, (int64_t)1, blocks[b].rows,
                         nr, k, kt, (int64_t)1, off, 2, 0, (double2*)u);
      RENO_CHECK_LAUNCH();
    } else {
      // block = R Q  <=>  X (nc x nr, col-major) = Qql L,  R = L^T, Q = Qql^T
      RENO_SOLVER_CHECK(rocsolver_zgeqlf(h, nc, nr, Xr, nc, taur));
      const int r0 = nc > nr ? nc - nr : 0;
      const int q0 = nr > nc ? nr - nc : 0;
      // u = R = L^T: u[rows[i], off+j] = X[(r0 + j) + i*nc] for i <= j + q0
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)r0, (int64_t)nc, (int64_t)1, blocks[b].rows,
                         nr, k, kt, (int64_t)1, off, 1, q0, (double2*)u);
      RENO_CHECK_LAUNCH();
      RENO_SOLVER_CHECK(rocsolver_zungql(h, nc, k, k, Xr + (int64_t)q0 * nc, nc, taur));
      // v = Q^T^T = Qql: v[cols[i], off+j] = X[i + (q0 + j)*nc]
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)q0 * nc, (int64_t)1, (int64_t)nc,
                         blocks[b].cols, nc, k, vdi, vdj, off, 2, 0, (double2*)v);
      RENO_CHECK_LAUNCH();
    }
    off += k;
  }
  return RENO_OK;
}
