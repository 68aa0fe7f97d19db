// Lanczos approximation of expm(dt*H) v (reference lib/krylov/krylov.py:15-84).
//
// The three-term recurrence runs on the device in FP64 with deterministic
// two-pass reductions (per-block partial sums, then one block folds them in a
// fixed order), so repeated calls give bit-identical results.
//
// The host keeps the reference's control flow (breakdown test, full-space
// exit, allclose test at every even j > 3) but never idles the device on it:
// iteration j+1 is queued before the host waits for alpha_j / beta_j, and a
// device-side stop flag - raised by the breakdown test or by a converged
// allclose check - turns every kernel of an iteration queued past the end into
// a no-op.  Only the allclose checks need host values (the small tridiagonal
// eigenproblem is solved on the host by cyclic Jacobi in FP64), so the device
// waits for the host once per two iterations, for a few microseconds.
#include <math.h>
#include <string.h>

#include <complex>
#include <vector>

#include "contract.h"

namespace reno {

constexpr int kRedThreads = 256;
constexpr int kMaxPartials = 1024;
constexpr int kMaxKrylov = 2048;

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
                                                        int64_t n, double* partial,
                                                        const int* stop) {
  if (stop && *stop) return;
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
    double* partial, const int* stop) {
  if (stop && *stop) return;
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

// Fold the partials in a fixed order.
// mode 0: out = sum                         (alpha_j)
// mode 1: out = sqrt(sum); raise the stop flag when out < thr   (beta_j, breakdown)
// mode 2: out = sum; raise the stop flag when sum == 0           (allclose violations)
// mode 3: out = sqrt(sum); clear the stop flag                   (|v0|, start of a call)
__global__ __launch_bounds__(kRedThreads) void k_fold(const double* __restrict__ partial,
                                                      int np, double* out, int mode, double thr,
                                                      int* stop) {
  if (mode != 3 && stop && *stop) return;
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) {
    const double v = (mode == 1 || mode == 3) ? sqrt(s) : s;
    out[0] = v;
    if (mode == 1 && v < thr) *stop = 1;
    if (mode == 2 && v == 0.0) *stop = 1;
    if (mode == 3) *stop = 0;
  }
}

// dst = src / (*den)
__global__ void k_div(const double2* __restrict__ src, double2* __restrict__ dst,
                      const double* __restrict__ den, int64_t n, const int* stop) {
  if (stop && *stop) return;
  const double q = *den;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = src[i];
    dst[i] = make_double2(a.x / q, a.y / q);
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
    res_new[i] = make_double2(xr, xi);
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
  size_t off_wl, off_t1, off_t2, off_split, off_V, off_w, off_r0, off_r1, off_part, off_scal, off_coef;
  size_t split_bytes;
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
  k.off_split = o; o += h.part_bytes;
  k.split_bytes = h.part_bytes;
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

// Per-thread host staging: pinned scalars (alpha_j, beta_j, stop flag and
// allclose count per iteration; Krylov coefficients) and a pool of events.
struct HostStage {
  double* pin = nullptr;  // [2*j] alpha_j, [2*j+1] beta_j, then counts, then coefs
  std::vector<hipEvent_t> events;
  ~HostStage() {
    if (pin) (void)hipHostFree(pin);
    for (auto e : events) (void)hipEventDestroy(e);
  }
};
static thread_local HostStage t_stage;
static constexpr int kPinCoef = 2 * kMaxKrylov + 64;                 // coefficient area (doubles)
static constexpr int kPinDoubles = kPinCoef + 2 * kMaxKrylov + 64;  // total

static double* pinned_buf() {
  if (!t_stage.pin) {
    if (hipHostMalloc((void**)&t_stage.pin, kPinDoubles * sizeof(double), hipHostMallocDefault) !=
        hipSuccess)
      t_stage.pin = nullptr;
  }
  return t_stage.pin;
}

static hipEvent_t stage_event(size_t i) {
  while (t_stage.events.size() <= i) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    t_stage.events.push_back(e);
  }
  return t_stage.events[i];
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
  int kcap = 2;
  while (kcap < kMaxKrylov && krylov_layout(op, kcap + 1).total <= ws_bytes) ++kcap;
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
  void* split = base + L.off_split;
  double2* V = (double2*)(base + L.off_V);
  const size_t vstride = align_up((size_t)n * 16) / 16;
  double2* w = (double2*)(base + L.off_w);
  double2* res[2] = {(double2*)(base + L.off_r0), (double2*)(base + L.off_r1)};
  double* part = (double*)(base + L.off_part);
  double* scal = (double*)(base + L.off_scal);  // [0]=|v0| [1]=violations [4]=stop flag (int)
  int* stop = (int*)(scal + 4);
  double* d_alpha = scal + 8;
  double* d_beta = scal + 8 + kcap;
  double2* d_coef = (double2*)(base + L.off_coef);
  const int np = n_partials(n);
  const int ew_blocks = (int)((n + 255) / 256 < 2048 ? (n + 255) / 256 : 2048);
  const double thr = 100.0 * (double)n * 2.220446049250313e-16;  // krylov.py:72

  if (op->nsite == 1) RENO_TRY(permute_w(op->prec, op->W, wl, op->Dl, op->d, op->Dr, s));

  // |v0| and V[0] = v0 / |v0| (krylov.py:42-45); clears the stop flag
  hipLaunchKernelGGL(k_norm2, dim3(np), dim3(kRedThreads), 0, s, (const double2*)v0, n, part);
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, (const double*)part, np, scal, 3,
                     0.0, stop);
  hipLaunchKernelGGL(k_div, dim3(ew_blocks), dim3(256), 0, s, (const double2*)v0, V,
                     (const double*)scal, n, (const int*)nullptr);
  RENO_CHECK_LAUNCH();
  RENO_HIP_CHECK(hipMemcpyAsync(pin + kPinCoef - 8, scal, sizeof(double), hipMemcpyDeviceToHost, s));

  // queue iteration j: matvec, alpha_j, and (j < n-1) the recurrence, beta_j
  // and V[j+1]; then stage alpha_j, beta_j and the stop flag to the host.
  auto queue_iter = [&](int64_t j) -> int {
    double2* vj = V + (size_t)j * vstride;
    RENO_TRY(hop_apply_core(op, wl, vj, w, t1, t2, s, stop, split, L.split_bytes));
    hipLaunchKernelGGL(k_dot_re, dim3(np), dim3(kRedThreads), 0, s, (const double2*)w,
                       (const double2*)vj, n, part, (const int*)stop);
    hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, (const double*)part, np,
                       d_alpha + j, 0, 0.0, stop);
    if (j < n - 1) {
      hipLaunchKernelGGL(k_recur_norm2, dim3(np), dim3(kRedThreads), 0, s, w, (const double2*)vj,
                         j > 0 ? (const double2*)(vj - vstride) : (const double2*)nullptr,
                         (const double*)(d_alpha + j), (const double*)(j > 0 ? d_beta + j - 1 : d_beta),
                         n, part, (const int*)stop);
      hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, (const double*)part, np,
                         d_beta + j, 1, thr, stop);
      if (j + 1 < kcap)
        hipLaunchKernelGGL(k_div, dim3(ew_blocks), dim3(256), 0, s, (const double2*)w,
                           V + (size_t)(j + 1) * vstride, (const double*)(d_beta + j), n,
                           (const int*)stop);
    }
    RENO_CHECK_LAUNCH();
    RENO_HIP_CHECK(hipMemcpyAsync(pin + 2 * j, d_alpha + j, sizeof(double), hipMemcpyDeviceToHost, s));
    if (j < n - 1)
      RENO_HIP_CHECK(hipMemcpyAsync(pin + 2 * j + 1, d_beta + j, sizeof(double), hipMemcpyDeviceToHost, s));
    hipEvent_t e = stage_event((size_t)j);
    if (!e) {
      set_error("expm_krylov: event creation failed");
      return RENO_EHIP;
    }
    RENO_HIP_CHECK(hipEventRecord(e, s));
    return RENO_OK;
  };

  std::vector<double> alpha, beta;
  std::vector<double2> coef;
  int cur = 0;
  bool have_res = false;
  double nrmv = -1.0;

  // the combination sum_i coef_i V[i] for m vectors; with check, also count
  // allclose violations against res[cur] and raise the stop flag when none
  auto queue_combine = [&](int m, double2* dst, bool check) -> int {
    tridiag_expm_coef(m, alpha.data(), beta.data(), nrmv, dt_re, dt_im, coef);
    memcpy(pin + kPinCoef, coef.data(), sizeof(double2) * m);
    RENO_HIP_CHECK(hipMemcpyAsync(d_coef, pin + kPinCoef, sizeof(double2) * m,
                                  hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_combine, dim3(np), dim3(kRedThreads), 0, s, (const double2*)V, n,
                       (int64_t)vstride, m, (const double2*)d_coef,
                       check ? (const double2*)res[cur] : (const double2*)nullptr, dst,
                       check ? part : (double*)nullptr);
    if (check)
      hipLaunchKernelGGL(k_fold, dim3(1), dim3(kRedThreads), 0, s, (const double*)part, np,
                         scal + 1, 2, 0.0, stop);
    RENO_CHECK_LAUNCH();
    return RENO_OK;
  };

  RENO_TRY(queue_iter(0));
  for (int64_t j = 0; j < n; ++j) {
    const bool last = (j == n - 1);
    const bool check_pt = (j > 3 && j % 2 == 0 && !last);
    // keep the device busy: queue the next iteration now unless this one
    // needs a host decision first (allclose checks)
    if (!check_pt && !last && j + 1 < kcap) RENO_TRY(queue_iter(j + 1));
    RENO_HIP_CHECK(hipEventSynchronize(stage_event((size_t)j)));
    if (nrmv < 0) {
      nrmv = pin[kPinCoef - 8];
      if (!(nrmv > 0.0)) {
        set_error("expm_krylov: zero starting vector (krylov.py:44 assert)");
        return RENO_EINVAL;
      }
    }
    alpha.push_back(pin[2 * j]);
    if (last) {
      // Krylov space equals the full space (krylov.py:59-61)
      RENO_TRY(queue_combine((int)j + 1, (double2*)out, false));
      RENO_HIP_CHECK(hipStreamSynchronize(s));  // pinned coefficients in flight
      *n_matvec = (int)j + 1;
      return RENO_OK;
    }
    beta.push_back(pin[2 * j + 1]);
    if (beta[j] < thr) {
      // Lanczos breakdown (krylov.py:72-74); the queued iteration j+1 was a no-op
      RENO_TRY(queue_combine((int)j + 1, (double2*)out, false));
      RENO_HIP_CHECK(hipStreamSynchronize(s));
      *n_matvec = (int)j + 1;
      return RENO_OK;
    }
    if (check_pt) {
      const int m = (int)j + 1;
      RENO_TRY(queue_combine(m, res[cur ^ 1], have_res));
      hipEvent_t ec = stage_event((size_t)kMaxKrylov);
      if (have_res) {
        RENO_HIP_CHECK(hipMemcpyAsync(pin + kPinCoef - 4, scal + 1, sizeof(double),
                                      hipMemcpyDeviceToHost, s));
      }
      RENO_HIP_CHECK(hipEventRecord(ec, s));
      // the next iteration goes behind the check; it is cancelled on device
      // if the check converged
      if (j + 1 < kcap) RENO_TRY(queue_iter(j + 1));
      RENO_HIP_CHECK(hipEventSynchronize(ec));
      if (have_res && pin[kPinCoef - 4] == 0.0) {
        RENO_HIP_CHECK(hipMemcpyAsync(out, res[cur ^ 1], (size_t)n * 16, hipMemcpyDeviceToDevice, s));
        *n_matvec = (int)j + 1;
        return RENO_OK;
      }
      cur ^= 1;
      have_res = true;
    }
    if (j + 1 >= kcap) {
      if (needed) *needed = krylov_layout(op, 2 * kcap).total;
      RENO_HIP_CHECK(hipStreamSynchronize(s));
      set_error("expm_krylov: %d Krylov vectors do not fit the workspace", kcap);
      return RENO_ENOSPACE;
    }
  }
  set_error("expm_krylov: unreachable");
  return RENO_EINVAL;
}
