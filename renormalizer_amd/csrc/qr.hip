// Quantum-number-blocked QR / RQ of the site coefficient matrix
// (svd_qn QR branch, reference svd_qn.py:99-240 with QR=True, economic).
//
// Each block B (nrows x ncols, gathered from the QN index sets) is factored
// by the device Householder QR of householder.hip:
//   system "L":  B = Q R                     -> QR of X = B (column-major)
//   system "R":  B = R Q  (scipy.linalg.rq)  -> QR of X = B^H J (columns of B^H
//                in reverse order): X = Q' R' gives R = J R'^H J, Q = J Q'^H.
// Householder factors with a real diagonal, so the result equals LAPACK's
// up to the sign of each column / row (and rounding).  Blocks with more rows
// than the register-resident panel covers go to rocSOLVER (zgelqf/zunglq,
// zgeqlf/zungql on the same memory viewed column-major).  The scatter into the
// zero-filled u / v (blockrecover, svd_qn.py:88-96) is one small kernel.
#include <rocsolver/rocsolver.h>

#include <map>
#include <mutex>

#include "householder.h"

namespace reno {

// One rocBLAS handle per host thread and device: a handle carries the stream
// it was last set to, so threads driving different streams must not share
// one.  Kept until the process ends.
static thread_local std::map<int, rocblas_handle> g_handles;

static int get_handle(rocblas_handle* out) {
  int dev = 0;
  RENO_HIP_CHECK(hipGetDevice(&dev));
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

// X[i + j*ld] = coef[rows[ri(i,j)] * cols_total + cols[ci(i,j)]] with
// mode 0: X (nr x nc) = B                 (ri = i, ci = j, ld = nr)
// mode 1: X (nc x nr) = conj(B)^T, cols reversed (ri = nr-1-j, ci = i, ld = nc)
// mode 2: X (nc x nr) = B^T row-major view (rocSOLVER path: X[j*nc + i] = B[j][i])
__global__ void k_gather(const double2* __restrict__ coef, int64_t cols_total,
                         const int64_t* __restrict__ rows, const int64_t* __restrict__ cols,
                         int nr, int nc, int mode, double2* __restrict__ X) {
  const int64_t total = (int64_t)nr * nc;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    if (mode == 0) {
      const int64_t j = t / nr, i = t - j * nr;
      X[t] = coef[rows[i] * cols_total + cols[j]];
    } else if (mode == 1) {
      const int64_t j = t / nc, i = t - j * nc;
      const double2 b = coef[rows[nr - 1 - j] * cols_total + cols[i]];
      X[t] = make_double2(b.x, -b.y);
    } else {
      const int64_t i = t / nc, j = t - i * nc;
      X[t] = coef[rows[i] * cols_total + cols[j]];
    }
  }
}

// dst[idx[i]*di + (off + j)*dj] = f(X[x0 + i*si + j*sj]) for i < ni, j < k, where
// f conjugates when cj and yields 0 unless cond(i - j): mode 0 -> i - j >= shift,
// mode 1 -> i - j <= shift, mode 2 -> always.
__global__ void k_scatter(const double2* __restrict__ X, int64_t x0, int64_t si, int64_t sj,
                          const int64_t* __restrict__ idx, int ni, int k, int64_t di,
                          int64_t dj, int64_t off, int mode, int shift, int cj,
                          double2* __restrict__ dst) {
  const int64_t total = (int64_t)ni * k;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(t / k), j = (int)(t - (t / k) * k);
    bool ok = true;
    if (mode == 0) ok = (i - j >= shift);
    else if (mode == 1) ok = (i - j <= shift);
    double2 v = make_double2(0.0, 0.0);
    if (ok) {
      v = X[x0 + (int64_t)i * si + (int64_t)j * sj];
      if (cj) v.y = -v.y;
    }
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

static size_t block_ws_bytes(int nr, int nc) {
  const size_t x = align_up((size_t)nr * nc * 16);
  const int k = nr < nc ? nr : nc;
  const int m = nr > nc ? nr : nc;  // rows of the factored matrix (either system)
  // X + Q (m x k) + Householder scratch, or X + tau for rocSOLVER
  return x + align_up((size_t)m * k * 16) + qr_ws_bytes(m, k) + align_up((size_t)(k + 1) * 16);
}

// one block through rocSOLVER (rows beyond the register-resident panel)
static int block_rocsolver(int system, hipStream_t s, const double2* coef, int cols_total,
                           const reno_qn_block& b, int64_t kt, int64_t off, int64_t vdi, int64_t vdj,
                           double2* u, double2* v, double2* X, double2* tau) {
  rocblas_handle h;
  RENO_TRY(get_handle(&h));
  RENO_SOLVER_CHECK(rocblas_set_stream(h, s));
  const int nr = b.nrows, nc = b.ncols;
  const int k = nr < nc ? nr : nc;
  rocblas_double_complex* Xr = (rocblas_double_complex*)X;
  rocblas_double_complex* taur = (rocblas_double_complex*)tau;
  hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)nr * nc)), dim3(256), 0, s, coef,
                     (int64_t)cols_total, b.rows, b.cols, nr, nc, 2, X);
  RENO_CHECK_LAUNCH();
  if (system == 0) {
    // row-major block = column-major X (nc x nr): block = Q R <=> X = L Qlq
    RENO_SOLVER_CHECK(rocsolver_zgelqf(h, nc, nr, Xr, nc, taur));
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s, (const double2*)X,
                       (int64_t)0, (int64_t)1, (int64_t)nc, b.cols, nc, k, vdi, vdj, off, 0, 0, 0, v);
    RENO_CHECK_LAUNCH();
    RENO_SOLVER_CHECK(rocsolver_zunglq(h, k, nr, k, Xr, nc, taur));
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s, (const double2*)X,
                       (int64_t)0, (int64_t)nc, (int64_t)1, b.rows, nr, k, kt, (int64_t)1, off, 2, 0,
                       0, u);
    RENO_CHECK_LAUNCH();
  } else {
    // block = R Q <=> X (nc x nr) = Qql L
    RENO_SOLVER_CHECK(rocsolver_zgeqlf(h, nc, nr, Xr, nc, taur));
    const int r0 = nc > nr ? nc - nr : 0;
    const int q0 = nr > nc ? nr - nc : 0;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s, (const double2*)X,
                       (int64_t)r0, (int64_t)nc, (int64_t)1, b.rows, nr, k, kt, (int64_t)1, off, 1, q0,
                       0, u);
    RENO_CHECK_LAUNCH();
    RENO_SOLVER_CHECK(rocsolver_zungql(h, nc, k, k, Xr + (int64_t)q0 * nc, nc, taur));
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s, (const double2*)X,
                       (int64_t)q0 * nc, (int64_t)1, (int64_t)nc, b.cols, nc, k, vdi, vdj, off, 2, 0,
                       0, v);
    RENO_CHECK_LAUNCH();
  }
  return RENO_OK;
}

}  // namespace reno

using namespace reno;

extern "C" size_t reno_qr_workspace_bytes(int rows, int cols, int nblocks,
                                          const reno_qn_block* blocks) {
  (void)rows;
  (void)cols;
  size_t need = 0;
  for (int b = 0; b < nblocks; ++b) {
    const size_t x = block_ws_bytes(blocks[b].nrows, blocks[b].ncols);
    if (x > need) need = x;
  }
  return need;
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
  const double2* cf = (const double2*)coef;
  double2* du = (double2*)u;
  double2* dv = (double2*)v;
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)rows * kt)), dim3(256), 0, s, du,
                     (int64_t)rows * kt);
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)cols * kt)), dim3(256), 0, s, dv,
                     (int64_t)cols * kt);
  RENO_CHECK_LAUNCH();
  // u is (rows x kt); v is (cols x kt), or its transpose (kt x cols)
  const int64_t vdi = v_transposed ? 1 : kt;
  const int64_t vdj = v_transposed ? (int64_t)cols : 1;
  int64_t off = 0;
  for (int ib = 0; ib < nblocks; ++ib) {
    const reno_qn_block& b = blocks[ib];
    const int nr = b.nrows, nc = b.ncols;
    const int k = nr < nc ? nr : nc;
    char* p = (char*)ws;
    double2* X = (double2*)p;
    p += align_up((size_t)nr * nc * 16);
    const int m = system == 0 ? nr : nc;  // rows of the factored matrix
    const int n = system == 0 ? nc : nr;
    double2* Q = (double2*)p;
    p += align_up((size_t)m * k * 16);
    if (!qr_supported(m, n)) {
      RENO_TRY(block_rocsolver(system, s, cf, cols, b, kt, off, vdi, vdj, du, dv, X, Q));
      off += k;
      continue;
    }
    hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)nr * nc)), dim3(256), 0, s, cf,
                       (int64_t)cols, b.rows, b.cols, nr, nc, system == 0 ? 0 : 1, X);
    RENO_CHECK_LAUNCH();
    RENO_TRY(qr_householder(s, m, n, X, Q, p));
    if (system == 0) {
      // u[rows[i], off+j] = Q[i + j*m];  v[cols[c], off+j] = R[j, c] = X[j + c*m], j <= c
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s,
                         (const double2*)Q, (int64_t)0, (int64_t)1, (int64_t)m, b.rows, nr, k, kt,
                         (int64_t)1, off, 2, 0, 0, du);
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)0, (int64_t)m, (int64_t)1, b.cols, nc, k, vdi,
                         vdj, off, 0, 0, 0, dv);
    } else {
      // u[rows[r], off+a] = conj(R'[k-1-a, nr-1-r]), valid for r - a <= nr - k
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, s,
                         (const double2*)X, (int64_t)(k - 1) + (int64_t)(nr - 1) * m, (int64_t)-m,
                         (int64_t)-1, b.rows, nr, k, kt, (int64_t)1, off, 1, nr - k, 1, du);
      // v[cols[c], off+a] = conj(Q'[c, k-1-a])
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, s,
                         (const double2*)Q, (int64_t)(k - 1) * m, (int64_t)1, (int64_t)-m, b.cols,
                         nc, k, vdi, vdj, off, 2, 0, 1, dv);
    }
    RENO_CHECK_LAUNCH();
    off += k;
  }
  return RENO_OK;
}

// ---- svd_qn SVD branch (full_matrices=True) --------------------------------
// Per block B_b (nr x nc, gathered column-major), rocSOLVER zgesvd with both
// bases complete: B = U S V^H.  u (rows x sum nr_b) and v (cols x sum nc_b)
// take, as svd_qn lays them out (svd_qn.py:205-240 with full_matrices):
// first every block's dim_b = min(nr_b, nc_b) singular columns, block by
// block, then every block's remaining nr_b - dim_b (nc_b - dim_b) completion
// columns, block by block; v holds V without conjugation (the reference's
// v = vt.T, coef = u diag(s) v^T).  s receives the dim_b singular values of
// each block in the same order (descending within a block).

namespace reno {

static size_t svd_block_ws_bytes(int nr, int nc) {
  const int k = nr < nc ? nr : nc;
  return align_up((size_t)nr * nc * 16) + align_up((size_t)nr * nr * 16) + align_up((size_t)nc * nc * 16) +
         align_up((size_t)(k + 1) * sizeof(double)) + align_up(sizeof(int));
}

}  // namespace reno

extern "C" size_t reno_svd_workspace_bytes(int rows, int cols, int nblocks, const reno_qn_block* blocks) {
  (void)rows;
  (void)cols;
  size_t need = 0;
  for (int b = 0; b < nblocks; ++b) need = std::max(need, svd_block_ws_bytes(blocks[b].nrows, blocks[b].ncols));
  return need;
}

extern "C" int reno_block_svd(int rows, int cols, const void* coef, int nblocks, const reno_qn_block* blocks,
                              void* u, void* v, void* s, void* ws, size_t ws_bytes, void* stream) {
  if (rows <= 0 || cols <= 0 || !coef || !u || !v || !s || nblocks < 0 || (nblocks && !blocks)) {
    set_error("block_svd: bad arguments");
    return RENO_EINVAL;
  }
  if (nblocks == 0) {
    set_error("block_svd: Invalid quantum number (no block)");
    return RENO_EQN;
  }
  int64_t ut = 0, vt = 0, kt = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (blocks[b].nrows <= 0 || blocks[b].ncols <= 0 || blocks[b].nrows > rows || blocks[b].ncols > cols ||
        !blocks[b].rows || !blocks[b].cols) {
      set_error("block_svd: bad block %d (%d x %d)", b, blocks[b].nrows, blocks[b].ncols);
      return RENO_EINVAL;
    }
    ut += blocks[b].nrows;
    vt += blocks[b].ncols;
    kt += std::min(blocks[b].nrows, blocks[b].ncols);
  }
  const size_t need = reno_svd_workspace_bytes(rows, cols, nblocks, blocks);
  if (!ws || ws_bytes < need) {
    set_error("block_svd: workspace %zu < %zu", ws_bytes, need);
    return RENO_ENOSPACE;
  }
  hipStream_t st = (hipStream_t)stream;
  rocblas_handle h;
  RENO_TRY(get_handle(&h));
  RENO_SOLVER_CHECK(rocblas_set_stream(h, st));
  const double2* cf = (const double2*)coef;
  double2* du = (double2*)u;
  double2* dv = (double2*)v;
  double* ds = (double*)s;
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)rows * ut)), dim3(256), 0, st, du, (int64_t)rows * ut);
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)cols * vt)), dim3(256), 0, st, dv, (int64_t)cols * vt);
  RENO_CHECK_LAUNCH();
  int64_t off_s = 0;           // singular columns (and s) of the blocks so far
  int64_t off_eu = kt, off_ev = kt;  // completion columns so far
  for (int ib = 0; ib < nblocks; ++ib) {
    const reno_qn_block& b = blocks[ib];
    const int nr = b.nrows, nc = b.ncols;
    const int k = std::min(nr, nc);
    char* p = (char*)ws;
    double2* X = (double2*)p;
    p += align_up((size_t)nr * nc * 16);
    double2* U = (double2*)p;
    p += align_up((size_t)nr * nr * 16);
    double2* VH = (double2*)p;
    p += align_up((size_t)nc * nc * 16);
    double* E = (double*)p;
    p += align_up((size_t)(k + 1) * sizeof(double));
    int* info = (int*)p;
    hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)nr * nc)), dim3(256), 0, st, cf, (int64_t)cols, b.rows,
                       b.cols, nr, nc, 0, X);
    RENO_CHECK_LAUNCH();
    RENO_SOLVER_CHECK(rocsolver_zgesvd(h, rocblas_svect_all, rocblas_svect_all, nr, nc,
                                       (rocblas_double_complex*)X, nr, ds + off_s, (rocblas_double_complex*)U,
                                       nr, (rocblas_double_complex*)VH, nc, E, rocblas_outofplace, info));
    // u[rows_b, off_s + j] = U[:, j] (j < k); u[rows_b, off_eu + j - k] = U[:, j] (j >= k)
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * k)), dim3(256), 0, st, (const double2*)U,
                       (int64_t)0, (int64_t)1, (int64_t)nr, b.rows, nr, k, (int64_t)ut, (int64_t)1, off_s, 2, 0, 0,
                       du);
    if (nr > k)
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nr * (nr - k))), dim3(256), 0, st,
                         (const double2*)U, (int64_t)k * nr, (int64_t)1, (int64_t)nr, b.rows, nr, nr - k,
                         (int64_t)ut, (int64_t)1, off_eu, 2, 0, 0, du);
    // v[cols_b, off + j] = VH[j, c] (the reference's v = vt.T)
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * k)), dim3(256), 0, st, (const double2*)VH,
                       (int64_t)0, (int64_t)nc, (int64_t)1, b.cols, nc, k, (int64_t)vt, (int64_t)1, off_s, 2, 0, 0,
                       dv);
    if (nc > k)
      hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nc * (nc - k))), dim3(256), 0, st,
                         (const double2*)VH, (int64_t)k, (int64_t)nc, (int64_t)1, b.cols, nc, nc - k, (int64_t)vt,
                         (int64_t)1, off_ev, 2, 0, 0, dv);
    RENO_CHECK_LAUNCH();
    off_s += k;
    off_eu += nr - k;
    off_ev += nc - k;
  }
  return RENO_OK;
}

// ---- eigh_qn: per quantum-number sector of a reduced density matrix --------
namespace reno {

// X (n x n, column-major) <- (X + X^H) / 2
__global__ void k_hermitize(double2* X, int n) {
  const int64_t total = (int64_t)n * n;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = t / n, i = t - j * n;
    if (i < j) continue;  // each pair once, by its lower element
    const double2 a = X[i + j * n], b = X[j + i * n];
    const double2 lo = make_double2(0.5 * (a.x + b.x), 0.5 * (a.y - b.y));
    X[i + j * n] = lo;
    X[j + i * n] = make_double2(lo.x, -lo.y);
  }
}

static size_t eigh_block_ws_bytes(int nb) {
  return align_up((size_t)nb * nb * 16) + align_up((size_t)(nb + 1) * sizeof(double)) + align_up(2 * sizeof(int));
}

// flag[0] = 1 when the eigensolver reported failure (info != 0) or left a
// non-finite eigenvalue or eigenvector entry, else 0.  One workgroup.
__global__ __launch_bounds__(256) void k_eigh_bad(const double2* __restrict__ X, int64_t nx,
                                                  const double* __restrict__ w, int nw,
                                                  const int* __restrict__ info, int* flag) {
  __shared__ int any;
  if (threadIdx.x == 0) any = *info != 0;
  __syncthreads();
  int bad = 0;
  for (int64_t i = threadIdx.x; i < nx; i += blockDim.x) {
    const double2 v = X[i];
    bad |= !(isfinite(v.x) && isfinite(v.y));
  }
  for (int i = threadIdx.x; i < nw; i += blockDim.x) bad |= !isfinite(w[i]);
  if (bad) any = 1;
  __syncthreads();
  if (threadIdx.x == 0) flag[0] = any;
}

}  // namespace reno

extern "C" size_t reno_eigh_workspace_bytes(int n, int nblocks, const reno_qn_block* blocks) {
  (void)n;
  size_t need = 0;
  for (int b = 0; b < nblocks; ++b) need = std::max(need, eigh_block_ws_bytes(blocks[b].nrows));
  return need;
}

extern "C" int reno_block_eigh(int n, const void* dm, int nblocks, const reno_qn_block* blocks, void* u, void* w,
                               void* ws, size_t ws_bytes, void* stream) {
  if (n <= 0 || !dm || !u || !w || nblocks < 0 || (nblocks && !blocks)) {
    set_error("block_eigh: bad arguments");
    return RENO_EINVAL;
  }
  if (nblocks == 0) {
    set_error("block_eigh: Invalid quantum number (no block)");
    return RENO_EQN;
  }
  int64_t ut = 0;
  for (int b = 0; b < nblocks; ++b) {
    if (blocks[b].nrows <= 0 || blocks[b].nrows != blocks[b].ncols || blocks[b].nrows > n || !blocks[b].rows ||
        !blocks[b].cols) {
      set_error("block_eigh: bad block %d (%d x %d)", b, blocks[b].nrows, blocks[b].ncols);
      return RENO_EINVAL;
    }
    ut += blocks[b].nrows;
  }
  const size_t need = reno_eigh_workspace_bytes(n, nblocks, blocks);
  if (!ws || ws_bytes < need) {
    set_error("block_eigh: workspace %zu < %zu", ws_bytes, need);
    return RENO_ENOSPACE;
  }
  hipStream_t st = (hipStream_t)stream;
  rocblas_handle h;
  RENO_TRY(get_handle(&h));
  RENO_SOLVER_CHECK(rocblas_set_stream(h, st));
  double2* du = (double2*)u;
  double* dw = (double*)w;
  hipLaunchKernelGGL(k_zero_z, dim3(grid_for((int64_t)n * ut)), dim3(256), 0, st, du, (int64_t)n * ut);
  RENO_CHECK_LAUNCH();
  int64_t off = 0;
  for (int ib = 0; ib < nblocks; ++ib) {
    const reno_qn_block& b = blocks[ib];
    const int nb = b.nrows;
    char* p = (char*)ws;
    double2* X = (double2*)p;
    p += align_up((size_t)nb * nb * 16);
    double* E = (double*)p;
    p += align_up((size_t)(nb + 1) * sizeof(double));
    int* info = (int*)p;
    int* flag = info + 1;
    auto load = [&]() {
      hipLaunchKernelGGL(k_gather, dim3(grid_for((int64_t)nb * nb)), dim3(256), 0, st, (const double2*)dm,
                         (int64_t)n, b.rows, b.cols, nb, nb, 0, X);
      hipLaunchKernelGGL(k_hermitize, dim3(grid_for((int64_t)nb * nb)), dim3(256), 0, st, X, nb);
    };
    load();
    RENO_CHECK_LAUNCH();
    // Jacobi eigenvalues, divide-and-conquer eigenvectors; rocSOLVER's
    // zheevd was measured returning NaN eigenvectors with info = 0 for
    // low-rank Gram matrices (eigenvalues 1 down to 1e-47), so every result
    // is checked on the device and a block that is not finite is solved
    // again by QR iteration (zheev).  The check waits for the stream once
    // per block.
    RENO_SOLVER_CHECK(rocsolver_zheevdj(h, rocblas_evect_original, rocblas_fill_lower, nb,
                                        (rocblas_double_complex*)X, nb, dw + off, info));
    hipLaunchKernelGGL(k_eigh_bad, dim3(1), dim3(256), 0, st, (const double2*)X, (int64_t)nb * nb,
                       (const double*)(dw + off), nb, (const int*)info, flag);
    RENO_CHECK_LAUNCH();
    int host_flag = 0;
    RENO_HIP_CHECK(hipMemcpyAsync(&host_flag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
    RENO_HIP_CHECK(hipStreamSynchronize(st));
    if (host_flag) {
      load();
      RENO_CHECK_LAUNCH();
      RENO_SOLVER_CHECK(rocsolver_zheev(h, rocblas_evect_original, rocblas_fill_lower, nb,
                                         (rocblas_double_complex*)X, nb, dw + off, E, info));
    }
    // u[rows_b, off + j] = X[:, j]
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((int64_t)nb * nb)), dim3(256), 0, st, (const double2*)X,
                       (int64_t)0, (int64_t)1, (int64_t)nb, b.rows, nb, nb, (int64_t)ut, (int64_t)1, off, 2, 0, 0,
                       du);
    RENO_CHECK_LAUNCH();
    off += nb;
  }
  return RENO_OK;
}
