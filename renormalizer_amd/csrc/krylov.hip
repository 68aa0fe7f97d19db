// Lanczos approximation of expm(dt*H) v (reference lib/krylov/krylov.py:15-84).
//
// The three-term recurrence runs on the device in FP64 with deterministic
// two-pass reductions (per-block partial sums, then one block folds them in a
// fixed order), so repeated calls give bit-identical results.  The host reads
// back alpha_j and beta_j once per iteration (the reference does the same:
// krylov.py:57,71 are device->host syncs), solves the small symmetric
// tridiagonal eigenproblem itself (cyclic Jacobi, FP64) and, at every even
// j > 3, forms the Krylov combination on the device while counting elements
// that violate allclose(res, new_res) in the same pass (krylov.py:76-81).
#include <math.h>
#include <string.h>

#include <complex>
#include <vector>

#include "contract.h"

namespace reno {

constexpr int kRedThreads = 256;
constexpr int kMaxPartials = 1024;

static int n_partials(int64_t n) {
  int64_t b = (n + kRedThreads * 4 - 1) / (kRedThreads * 4);
  if (b < 1) b = 1;
  if (b > kMaxPartials) b = kMaxPartials;
  return (int)b;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int off = kRedThreads / 2; off > 0; off >>= 1) {
    if (tid < off) sh[tid] += sh[tid + off];
    __syncthreads();
  }
  return sh[0];
}

// partial[b] = sum over this block's grid-stride share of Re(conj(w) v)
__global__ __launch_bounds__(kRedThreads) void k_dot_re(const double2* __restrict__ w,
                                                        const double2* __restrict__ v,
                                                        int64_t n, double* partial) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const double2 a = w[i], b = v[i];
    acc += a.x * b.x + a.y * b.y;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// partial[b] = sum |v|^2
__global__ __launch_bounds__(kRedThreads) void k_norm2(const double2* __restrict__ v,
                                                       int64_t n, double* partial) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const double2 a = v[i];
    acc += a.x * a.x + a.y * a.y;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// w -= alpha v_j + beta v_{j-1}  (krylov.py:70), partial[b] = sum |w|^2
__global__ __launch_bounds__(kRedThreads) void k_recur_norm2(
    double2* __restrict__ w, const double2* __restrict__ vj, const double2* __restrict__ vjm1,
    const double* __restrict__ alpha_j, const double* __restrict__ beta_jm1, int64_t n,
    double* partial) {
  __shared__ double sh[kRedThreads];
  const double al = *alpha_j;
  const double be = vjm1 ? *beta_jm1 : 0.0;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    double2 a = w[i];
    const double2 b = vj[i];
    double tr = al * b.x, ti = al * b.y;
    if (vjm1) {
      const double2 c = vjm1[i];
      tr += be * c.x;
      ti += be * c.y;
    }
    a.x -= tr;
    a.y -= ti;
    w[i] = a;
    acc += a.x * a.x + a.y * a.y;
  }
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out[0] = sum(partial) (mode 0) or sqrt(sum) (mode 1), fixed order
__global__ __launch_bounds__(kRedThreads) void k_fold(const double* __restrict__ partial,
                                                      int np, double* out, int mode) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) out[0] = mode ? sqrt(s) : s;
}

// dst = src / (*den) or src / den_host when den == nullptr
__global__ void k_div(const double2* __restrict__ src, double2* __restrict__ dst,
                      const double* __restrict__ den, double den_host, int64_t n) {
  const double q = den ? *den : den_host;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = src[i];
    double2 r;
    r.x = a.x / q;
    r.y = a.y / q;
    dst[i] = r;
  }
}

// res_new[i] = sum_j coef[j] V[j][i]; if res_old: partial count of elements
// violating |res_old - res_new| <= atol + rtol |res_new| (numpy.allclose(res_old, res_new)).
__global__ __launch_bounds__(kRedThreads) void k_combine(const double2* __restrict__ V,
                                                         int64_t n, int64_t vstride, int nv,
                                                         const double2* __restrict__ coef,
                                                         const double2* __restrict__ res_old,
                                                         double2* __restrict__ res_new,
                                                         double* partial) {
  __shared__ double sh[kRedThreads];
  __shared__ double2 sc[64];
  for (int j = threadIdx.x; j < nv && j < 64; j += kRedThreads) sc[j] = coef[j];
  __syncthreads();
  double viol = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    double xr = 0.0, xi = 0.0;
    for (int j = 0; j < nv; ++j) {
      const double2 c = j < 64 ? sc[j] : coef[j];
      const double2 v = V[(int64_t)j * vstride + i];
      xr += v.x * c.x - v.y * c.y;
      xi += v.x * c.y + v.y * c.x;
    }
    double2 r;
    r.x = xr;
    r.y = xi;
    res_new[i] = r;
    if (res_old) {
      const double2 o = res_old[i];
      const double dr = o.x - xr, di = o.y - xi;
      const double lhs = sqrt(dr * dr + di * di);
      const double rhs = 1e-8 + 1e-5 * sqrt(xr * xr + xi * xi);
      if (!(lhs <= rhs)) viol += 1.0;
    }
  }
  if (partial) {
    const double s = block_sum(viol, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
  }
}

// ---------------------------------------------------------------------------
// Host side: eigen-decomposition of the symmetric tridiagonal Lanczos matrix
// by cyclic Jacobi rotations (FP64), then coef = U (nrm exp(dt w) U[0,:]).
static void jacobi_eigh(int m, std::vector<double>& a, std::vector<double>& w,
                        std::vector<double>& u) {
  u.assign((size_t)m * m, 0.0);
  for (int i = 0; i < m; ++i) u[(size_t)i * m + i] = 1.0;
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = 0; q < m; ++q) {
        const double x = a[(size_t)p * m + q];
        tot += x * x;
        if (p != q) off += x * x;
      }
    if (off <= 1e-34 * tot || off == 0.0) break;
    for (int p = 0; p < m - 1; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = a[(size_t)p * m + q];
        if (apq == 0.0) continue;
        const double app = a[(size_t)p * m + p], aqq = a[(size_t)q * m + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0);
        const double s = t * c;
        for (int k = 0; k < m; ++k) {  // A <- A J
          const double akp = a[(size_t)k * m + p], akq = a[(size_t)k * m + q];
          a[(size_t)k * m + p] = c * akp - s * akq;
          a[(size_t)k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {  // A <- J^T A
          const double apk = a[(size_t)p * m + k], aqk = a[(size_t)q * m + k];
          a[(size_t)p * m + k] = c * apk - s * aqk;
          a[(size_t)q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {  // U <- U J
          const double ukp = u[(size_t)k * m + p], ukq = u[(size_t)k * m + q];
          u[(size_t)k * m + p] = c * ukp - s * ukq;
          u[(size_t)k * m + q] = s * ukp + c * ukq;
        }
      }
  }
  w.resize(m);
  for (int i = 0; i < m; ++i) w[i] = a[(size_t)i * m + i];
}

// coefficients of the Krylov vectors for expm(dt T) (nrm e_0), krylov.py:15-24
static void tridiag_expm_coef(int m, const double* alpha, const double* beta, double nrm,
                              double dt_re, double dt_im, std::vector<double2>& coef) {
  std::vector<double> a((size_t)m * m, 0.0), w, u;
  for (int i = 0; i < m; ++i) a[(size_t)i * m + i] = alpha[i];
  for (int i = 0; i + 1 < m; ++i) {
    a[(size_t)i * m + i + 1] = beta[i];
    a[(size_t)(i + 1) * m + i] = beta[i];
  }
  jacobi_eigh(m, a, w, u);
  std::vector<std::complex<double>> g(m);
  const std::complex<double> dt(dt_re, dt_im);
  for (int k = 0; k < m; ++k) g[k] = nrm * std::exp(dt * w[k]) * u[k];  // u[0*m+k] = U[0][k]
  coef.resize(m);
  for (int i = 0; i < m; ++i) {
    std::complex<double> acc(0.0, 0.0);
    for (int k = 0; k < m; ++k) acc += u[(size_t)i * m + k] * g[k];
    coef[i].x = acc.real();
    coef[i].y = acc.imag();
  }
}

struct KrylovLayout {
  size_t hop_wl, hop_t1, hop_t2;
  size_t off_wl, off_t1, off_t2, off_V, off_w, off_r0, off_r1, off_part, off_scal, off_coef;
  size_t total;
};

static KrylovLayout krylov_layout(const reno_hop_desc* op, int kcap) {
  KrylovLayout k{};
  const HopLayout h = hop_layout(op);
  const int64_t n = hop_vec_len(op);
  const size_t vb = align_up((size_t)n * 16);
  size_t o = 0;
  k.off_wl = o; o += h.wl_bytes;
  k.off_t1 = o; o += h.t1_bytes;
  k.off_t2 = o; o += h.t2_bytes;
  k.off_V = o; o += vb * (size_t)kcap;
  k.off_w = o; o += vb;
  k.off_r0 = o; o += vb;
  k.off_r1 = o; o += vb;
  k.off_part = o; o += align_up(kMaxPartials * sizeof(double));
  k.off_scal = o; o += align_up((size_t)(2 * kcap + 16) * sizeof(double));
  k.off_coef = o; o += align_up((size_t)kcap * 16);
  k.total = o;
  return k;
}

// Pinned host staging for scalars read back every iteration (per thread).
struct Pinned {
  double* p = nullptr;
  ~Pinned() {
    if (p) (void)hipHostFree(p);
  }
};
static thread_local Pinned t_pinned;
static constexpr int kPinnedDoubles = 8192;

static double* pinned_buf() {
  if (!t_pinned.p) {
    if (hipHostMalloc((void**)&t_pinned.p, kPinnedDoubles * sizeof(double),
                      hipHostMallocDefault) != hipSuccess) {
      t_pinned.p = nullptr;
    }
  }
  return t_pinned.p;
}

}  // namespace reno

using namespace reno;

extern "C" size_t reno_krylov_workspace_bytes(const reno_hop_desc* op, int kcap) {
  if (check_hop(op) != RENO_OK || kcap < 2) return 0;
  return krylov_layout(op, kcap).total;
}

extern "C" int reno_expm_krylov(const reno_hop_desc* op, double dt_re, double dt_im,
                                const void* v0, void* out, int* n_matvec, size_t* needed,
                                void* ws, size_t ws_bytes, void* stream) {
  RENO_TRY(check_hop(op));
  if (!v0 || !out || !n_matvec) {
    set_error("expm_krylov: null argument");
    return RENO_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = hop_vec_len(op);
  // capacity of the caller's workspace in Krylov vectors
  int kcap = 2;
  while (krylov_layout(op, kcap + 1).total <= ws_bytes && kcap < 4096) ++kcap;
  const KrylovLayout L = krylov_layout(op, kcap);
  if (!ws || L.total > ws_bytes) {
    if (needed) *needed = krylov_layout(op, 16).total;
    set_error("expm_krylov: workspace %zu too small", ws_bytes);
    return RENO_ENOSPACE;
  }
  double* pin = pinned_buf();
  if (!pin) {
    set_error("expm_krylov: pinned host buffer allocation failed");
    return RENO_EHIP;
  }
  char* base = (char*)ws;
  void* wl = base + L.off_wl;
  void* t1 = base + L.off_t1;
  void* t2 = base + L.off_t2;
  double2* V = (double2*)(base + L.off_V);
  const size_t vstride = align_up((size_t)n * 16) / 16;
  double2* w = (double2*)(base + L.off_w);
  double2* res[2] = {(double2*)(base + L.off_r0), (double2*)(base + L.off_r1)};
  double* part = (double*)(base + L.off_part);
  double* scal = (double*)(base + L.off_scal);  // [0]=nrm [1..] alpha, beta
  double* d_alpha = scal + 8;
  double* d_beta = scal + 8 + kcap;
  double2* d_coef = (double2*)(base + L.off_coef);
  const int np = n_partials(n);
  const int ew_blocks = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);

  if (op->nsite == 1) RENO_TRY(permute_w(op->prec, op->W, wl, op->Dl, op->d, op->Dr, s));

  // normalise the starting vector (krylov.py:42-45)
  hipLaunchKernelGGL(k_norm2, dim3(np), dim3(kRedThreads), 0, s, (const double2*)v0, n, part);
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, part, np, scal, 1);
  RENO_CHECK_LAUNCH();
  RENO_HIP_CHECK(hipMemcpyAsync(pin, scal, sizeof(double), hipMemcpyDeviceToHost, s));
  RENO_HIP_CHECK(hipStreamSynchronize(s));
  const double nrmv = pin[0];
  if (!(nrmv > 0.0)) {
    set_error("expm_krylov: zero starting vector (krylov.py:44 assert)");
    return RENO_EINVAL;
  }
  hipLaunchKernelGGL(k_div, dim3(ew_blocks), dim3(256), 0, s, (const double2*)v0, V,
                     (const double*)nullptr, nrmv, n);
  RENO_CHECK_LAUNCH();

  std::vector<double> alpha, beta;
  std::vector<double2> coef;
  int cur = 0;
  bool have_res = false;
  const double eps = 2.220446049250313e-16;

  auto finish = [&](int m, double2* dst) -> int {
    tridiag_expm_coef(m, alpha.data(), beta.data(), nrmv, dt_re, dt_im, coef);
    memcpy(pin + 4096, coef.data(), sizeof(double2) * m);
    RENO_HIP_CHECK(hipMemcpyAsync(d_coef, pin + 4096, sizeof(double2) * m,
                                  hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_combine, dim3(np), dim3(kRedThreads), 0, s, (const double2*)V,
                       n, (int64_t)vstride, m, (const double2*)d_coef, (const double2*)nullptr, dst,
                       (double*)nullptr);
    RENO_CHECK_LAUNCH();
    // the pinned coefficients must outlive the copy
    RENO_HIP_CHECK(hipStreamSynchronize(s));
    return RENO_OK;
  };

  for (int64_t j = 0; j < n; ++j) {
    double2* vj = V + (size_t)j * vstride;
    RENO_TRY(hop_apply_core(op, wl, vj, w, t1, t2, s));
    hipLaunchKernelGGL(k_dot_re, dim3(np), dim3(kRedThreads), 0, s, (const double2*)w,
                       (const double2*)vj, n, part);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, part, np, d_alpha + j, 0);
    RENO_CHECK_LAUNCH();
    if (j == n - 1) {
      // Krylov space equals the full space (krylov.py:59-61)
      RENO_HIP_CHECK(hipMemcpyAsync(pin, d_alpha + j, sizeof(double), hipMemcpyDeviceToHost, s));
      RENO_HIP_CHECK(hipStreamSynchronize(s));
      alpha.push_back(pin[0]);
      RENO_TRY(finish((int)j + 1, (double2*)out));
      *n_matvec = (int)j + 1;
      return RENO_OK;
    }
    hipLaunchKernelGGL(k_recur_norm2, dim3(np), dim3(kRedThreads), 0, s, w,
                       (const double2*)vj,
                       j > 0 ? (const double2*)(V + (size_t)(j - 1) * vstride) : (const double2*)nullptr,
                       (const double*)(d_alpha + j), (const double*)(j > 0 ? d_beta + j - 1 : d_beta),
                       n, part);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, part, np, d_beta + j, 1);
    RENO_CHECK_LAUNCH();
    RENO_HIP_CHECK(hipMemcpyAsync(pin, d_alpha + j, sizeof(double), hipMemcpyDeviceToHost, s));
    RENO_HIP_CHECK(hipMemcpyAsync(pin + 1, d_beta + j, sizeof(double), hipMemcpyDeviceToHost, s));
    RENO_HIP_CHECK(hipStreamSynchronize(s));
    alpha.push_back(pin[0]);
    beta.push_back(pin[1]);
    if (beta[j] < 100.0 * (double)n * eps) {
      // Lanczos breakdown (krylov.py:72-74)
      RENO_TRY(finish((int)j + 1, (double2*)out));
      *n_matvec = (int)j + 1;
      return RENO_OK;
    }
    if (j > 3 && j % 2 == 0) {
      const int m = (int)j + 1;
      tridiag_expm_coef(m, alpha.data(), beta.data(), nrmv, dt_re, dt_im, coef);
      memcpy(pin + 4096, coef.data(), sizeof(double2) * m);
      RENO_HIP_CHECK(hipMemcpyAsync(d_coef, pin + 4096, sizeof(double2) * m,
                                    hipMemcpyHostToDevice, s));
      double2* rnew = res[cur ^ 1];
      hipLaunchKernelGGL(k_combine, dim3(np), dim3(kRedThreads), 0, s, (const double2*)V, n,
                         (int64_t)vstride, m, (const double2*)d_coef,
                         have_res ? (const double2*)res[cur] : (const double2*)nullptr, rnew,
                         have_res ? part : (double*)nullptr);
      RENO_CHECK_LAUNCH();
      if (have_res) {
        hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, part, np, scal + 1, 0);
        RENO_CHECK_LAUNCH();
        RENO_HIP_CHECK(hipMemcpyAsync(pin + 2, scal + 1, sizeof(double), hipMemcpyDeviceToHost, s));
      }
      RENO_HIP_CHECK(hipStreamSynchronize(s));
      if (have_res && pin[2] == 0.0) {
        RENO_HIP_CHECK(hipMemcpyAsync(out, rnew, (size_t)n * 16, hipMemcpyDeviceToDevice, s));
        *n_matvec = (int)j + 1;
        return RENO_OK;
      }
      cur ^= 1;
      have_res = true;
    }
    if (j + 1 >= kcap) {
      if (needed) *needed = krylov_layout(op, 2 * kcap).total;
      set_error("expm_krylov: %d Krylov vectors do not fit the workspace", kcap);
      return RENO_ENOSPACE;
    }
    hipLaunchKernelGGL(k_div, dim3(ew_blocks), dim3(256), 0, s, (const double2*)w,
                       V + (size_t)(j + 1) * vstride, (const double*)(d_beta + j), 0.0, n);
    RENO_CHECK_LAUNCH();
  }
  set_error("expm_krylov: unreachable");
  return RENO_EINVAL;
}
