// Lanczos approximation of expm(dt*H) v (reference lib/krylov/krylov.py:15-84).
//
// The three-term recurrence runs on the device in FP64 with deterministic
// two-pass reductions (per-block partial sums folded in a fixed order, by the
// last block of the same launch or by the next kernel), so repeated calls
// give bit-identical results.  Krylov vectors are stored unnormalised: slot
// j holds V[j] = v_j / sig_j with sig_j = 1 / beta_{j-1} kept beside alpha
// and beta, and every consumer applies the factor (the matvec is linear, so
// w = H V[j] is scaled afterwards).  The recurrence thereby writes the next
// vector in the same pass that computes its norm.
//
// The reference's control flow runs on the device too: the breakdown test
// (beta_j < 100 n eps), the full-space exit, and at every even j > 3 the
// small exponential expm(dt T) e_0 of the tridiagonal Lanczos matrix
// followed by numpy.allclose(res, new_res) on the combined vector.  A
// device-side state block records why and where the iteration ended, and a
// stop flag turns every kernel queued past that point into a no-op, so the
// host queues iterations ahead without waiting for any of them:
//   * reno_expm_krylov queues chunks of iterations and reads the state once
//     per chunk (normally once per call);
//   * reno_expm_krylov_async queues up to max_matvec iterations plus the
//     finalisation and returns at once; the outcome lands in a device info
//     array that the caller reads whenever it likes.
// A convergence check first tests one block of 256 elements in every 4096:
// a violation there settles that allclose fails, which is the usual outcome
// until the last check, and the full pass over the vectors is skipped.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "contract.h"
#include "tricoef.h"

namespace reno {

constexpr int kRedThreads = 256;
constexpr int kMaxPartials = 1024;
constexpr int kMaxKrylov = 1024;  // 64 lanes x 16 entries in k_tri_coef

// device state block (ints) at scal + 2
enum {
  KS_STOP = 0,     // nonzero: every later iteration kernel returns at once
  KS_STATUS = 1,   // 0 running, 1 converged (allclose), 2 breakdown, 3 full space, -1 zero v0
  KS_MFINAL = 2,   // Krylov vectors of the final combination (status 2, 3)
  KS_RESIDX = 3,   // result buffer holding the converged vector (status 1)
  KS_NMATVEC = 4,  // matvecs used (reference j + 1)
  KS_COUNT = 5,    // blocks finished in the current launch (last-block folds)
  KS_BITER = 6,    // reference iteration j of a breakdown / full-space exit (else INT_MAX)
  KS_CONV = 7,     // 1 + iteration j of the converged allclose check (else 0)
  KS_PFAIL = 8,    // iteration j whose probe found an allclose violation
  KS_INTS = 12     // the block holds at most 12 ints (scal[2..8))
};

// How the call ended, from the state block: 1 converged (result in
// res[RESIDX], CONV matvecs), 2 breakdown / 3 full space (combination of
// MFINAL vectors, NMATVEC matvecs), 0 iteration budget exhausted, -1 zero v0.
// A check of iteration j is queued behind the recurrence of iteration j,
// which can already have found the breakdown of iteration j - 1 (it folds
// beta_{j-1}); the earlier event wins, as in the reference's loop (a
// breakdown and a converged check at the same iteration give the same
// vector, and the breakdown is taken).
__host__ __device__ __forceinline__ int final_kind(const int* st) {
  const int status = st[KS_STATUS];
  if (status < 0) return -1;
  const int conv = st[KS_CONV];
  if (conv && conv - 1 < st[KS_BITER]) return 1;
  return status;
}

// A check at iteration j is void once a check converged, v0 was zero, or the
// recurrence stopped at or before j.
__device__ __forceinline__ bool check_void(const int* st, int j) {
  return st[KS_CONV] != 0 || st[KS_STATUS] < 0 || st[KS_BITER] <= j;
}

// one block per 256 elements up to kMaxPartials blocks: the bond-matrix
// vectors (65 536 elements at M = 256) still spread over 256 CUs
static int n_partials(int64_t n) {
  int64_t b = (n + kRedThreads - 1) / kRedThreads;
  if (b < 1) b = 1;
  if (b > kMaxPartials) b = kMaxPartials;
  return (int)b;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
  static_assert(kRedThreads == 256, "block_sum256 sums 256 threads");
  return block_sum256(v, sh);
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

// beta_{j-1} = |V[j]| folded from the norm partials the previous recurrence
// left (every block folds them in the same order), sig_j = 1 / beta_{j-1};
// block 0 stores both.  Breakdown (beta_{j-1} < thr, krylov.py:72-74) ends
// the iteration with the j vectors built before it (the matvec of V[j]
// queued meanwhile is discarded).  Returns false on breakdown.
__device__ __forceinline__ bool fold_beta(const double* __restrict__ norm_part, int np, int j,
                                          double thr, double* __restrict__ beta,
                                          double* __restrict__ sig, int* st, double* sh,
                                          double* be_out) {
  double b2 = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) b2 += norm_part[i];
  const double be = sqrt(block_sum(b2, sh));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    beta[j - 1] = be;
    sig[j] = 1.0 / be;
    if (be < thr) {
      st[KS_STATUS] = 2;
      st[KS_MFINAL] = j;
      st[KS_NMATVEC] = j;
      st[KS_BITER] = j - 1;
      st[KS_STOP] = 1;
    }
  }
  *be_out = be;
  return !(be < thr);
}

// Iteration j of the recurrence (krylov.py:57, 70-74), with w = H V[j]:
// sig_j from the previous iteration's norm partials (fold_beta; sig_0 = 1),
// alpha_j = Re<H v_j, v_j> = sig_j^2 Re<w, V[j]> folded from the dot partials
// of the matvec (every block folds them in the same order; block 0 stores
// alpha_j), then
//   V[j+1] = sig_j w - alpha_j sig_j V[j] - beta_{j-1} sig_{j-1} V[j-1]
// (= beta_j v_{j+1}) into vnext, and per-block partial sums of |V[j+1]|^2
// for the next iteration.
__global__ __launch_bounds__(kRedThreads) void k_recur(
    const double2* __restrict__ w, const double2* __restrict__ vj, const double2* __restrict__ vjm1,
    double2* vnext, const double* __restrict__ dot_part, const double* __restrict__ norm_prev,
    int np, double* __restrict__ alpha, double* __restrict__ beta, double* __restrict__ sig,
    int j, int64_t n, double thr, double* __restrict__ norm_part, int* st) {
  if (st[KS_STOP]) return;
  __shared__ double sh[kRedThreads];
  double bprev = 1.0;  // beta_{j-1}
  if (j > 0 && !fold_beta(norm_prev, np, j, thr, beta, sig, st, sh, &bprev)) return;
  const double sj = j > 0 ? 1.0 / bprev : 1.0;
  double a = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) a += dot_part[i];
  const double al = block_sum(a, sh) * sj * sj;
  if (blockIdx.x == 0 && threadIdx.x == 0) alpha[j] = al;
  const double ca = al * sj;
  // beta_{j-1} was folded above (block 0 stores it concurrently, so it is
  // not read back); sig_{j-1} was stored by an earlier launch
  const double cb = vjm1 ? bprev * sig[j - 1] : 0.0;
  double acc = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const double2 x = w[i];
    const double2 b = vj[i];
    double xr = sj * x.x - ca * b.x, xi = sj * x.y - ca * b.y;
    if (vjm1) {
      const double2 c = vjm1[i];
      xr -= cb * c.x;
      xi -= cb * c.y;
    }
    vnext[i] = make_double2(xr, xi);
    acc += xr * xr + xi * xi;
  }
  const double s2 = block_sum(acc, sh);
  if (threadIdx.x == 0) norm_part[blockIdx.x] = s2;
}

// Start of a call: scal[0] = |v0|, state cleared; a zero v0 (the reference's
// assert, krylov.py:44) ends the call with status -1.
__global__ __launch_bounds__(kRedThreads) void k_kry_init(const double* __restrict__ partial,
                                                          int np, double* scal, double* sig,
                                                          int* st) {
  __shared__ double sh[kRedThreads];
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x < KS_INTS) st[threadIdx.x] = threadIdx.x == KS_BITER ? 0x7fffffff : 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double nrm = sqrt(s);
    scal[0] = nrm;
    sig[0] = 1.0;
    if (!(nrm > 0.0)) {
      st[KS_STOP] = 1;
      st[KS_STATUS] = -1;
    }
  }
}

// Last possible iteration j = n - 1 (one block): beta_{j-1} / sig_j with the
// breakdown test as in k_recur, then alpha_j = sig_j^2 (sum of the dot
// partials).
__global__ __launch_bounds__(kRedThreads) void k_fold_alpha(const double* __restrict__ partial,
                                                            const double* __restrict__ norm_prev,
                                                            int np, double* __restrict__ alpha,
                                                            double* __restrict__ beta,
                                                            double* __restrict__ sig, int j,
                                                            double thr, int* st) {
  if (st[KS_STOP]) return;
  __shared__ double sh[kRedThreads];
  double bprev = 1.0;
  if (j > 0 && !fold_beta(norm_prev, np, j, thr, beta, sig, st, sh, &bprev)) return;
  const double sj = j > 0 ? 1.0 / bprev : 1.0;
  double acc = 0.0;
  for (int i = threadIdx.x; i < np; i += kRedThreads) acc += partial[i];
  const double s = block_sum(acc, sh);
  if (threadIdx.x == 0) alpha[j] = s * sj * sj;
}

// Krylov space equals the full space after iteration n-1 (krylov.py:59-61)
__global__ void k_mark_full(int n, int* st) {
  if (st[KS_STOP]) return;
  st[KS_STATUS] = 3;
  st[KS_MFINAL] = n;
  st[KS_NMATVEC] = n;
  st[KS_BITER] = n - 1;
  st[KS_STOP] = 1;
}

// dst = src / (*den)
__global__ void k_div(const double2* __restrict__ src, double2* __restrict__ dst,
                      const double* __restrict__ den, int64_t n, const int* stop) {
  if (*stop) return;
  const double q = *den;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double2 a = src[i];
    dst[i] = make_double2(a.x / q, a.y / q);
  }
}

// k_tri_coef: tri_coef_wave for the Krylov loop.
// final == 0: the check of iteration m-1 (skipped once stopped);
// final == 1: end of the call, m = MFINAL, run only for status 2 / 3.
template <int R>
__global__ __launch_bounds__(64) void k_tri_coef(const double* __restrict__ alpha,
                                                 const double* __restrict__ beta,
                                                 const double* __restrict__ scal, double dt_re,
                                                 double dt_im, int m_in, int final,
                                                 const int* st, double2* __restrict__ coef) {
  // the operands first (alpha and beta hold at least m_in entries), so that
  // their latency overlaps the reads of the state block
  const int lane = threadIdx.x & 63;
  double a[R], bl[R], br[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane * R + r;
    a[r] = i < m_in ? alpha[i] : 0.0;
    bl[r] = (i > 0 && i < m_in) ? beta[i - 1] : 0.0;
    br[r] = i + 1 < m_in ? beta[i] : 0.0;
  }
  const double v0n = scal[0];
  if (!final && check_void(st, m_in - 1)) return;
  if (final) {
    const int kind = final_kind(st);
    if (kind != 2 && kind != 3) return;
  }
  __shared__ double sJ[kChebMax];
  __shared__ int sK;
  const int m = final ? st[KS_MFINAL] : m_in;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int i = lane * R + r;
    if (i >= m) a[r] = bl[r] = 0.0;
    if (i + 1 >= m) br[r] = 0.0;
  }
  tri_coef_core<R>(a, bl, br, v0n, dt_re, dt_im, m, sJ, &sK, coef);
}

// Convergence check of iteration j (krylov.py:76-81): with
//   res_new[i] = sum_{k<nv} coef_new[k] sig_k V[k][i]   (this check) and
//   res_old[i] = sum_{k<nv-2} coef_old[k] sig_k V[k][i] (the previous one),
// numpy.allclose(res_old, res_new) holds when no element violates
// |res_old - res_new| <= atol + rtol |res_new|.  Both combinations are formed
// from V in one pass (res_old is not kept between checks), in the same
// order as k_final_out forms a result, so the values are those the stored
// vectors would hold.
//   probe == 1: blocks of 256 consecutive elements, one per 4096; any
//     violation there proves the check fails (st[KS_PFAIL] = j), and the
//     full pass is skipped;
//   probe == 0: every element, res_new kept in res (buffer residx); the last
//     block to finish sees whether any block counted a violation (low 12
//     bits of the completion atomic: blocks done, the rest: blocks with a
//     violation): none means converged at iteration j.
__global__ __launch_bounds__(kRedThreads) void k_check(const double2* __restrict__ V, int64_t n,
                                                       int64_t vstride, int nv,
                                                       const double2* __restrict__ coef_new,
                                                       const double2* __restrict__ coef_old,
                                                       const double* __restrict__ sig,
                                                       double2* __restrict__ res, int j, int residx,
                                                       int probe, int* st) {
  if (check_void(st, j)) return;
  if (!probe && st[KS_PFAIL] == j) return;
  __shared__ double sh[kRedThreads];
  __shared__ double2 sc[2][64];
  const int nvo = nv - 2;
  for (int k = threadIdx.x; k < nv && k < 64; k += kRedThreads) {
    const double2 c = coef_new[k];
    sc[0][k] = make_double2(c.x * sig[k], c.y * sig[k]);
    const double2 o = k < nvo ? coef_old[k] : make_double2(0.0, 0.0);
    sc[1][k] = make_double2(o.x * sig[k], o.y * sig[k]);
  }
  __syncthreads();
  double viol = 0.0;
  const int64_t first = probe ? (int64_t)blockIdx.x * 4096 + threadIdx.x
                              : blockIdx.x * (int64_t)kRedThreads + threadIdx.x;
  const int64_t step = probe ? n : (int64_t)gridDim.x * kRedThreads;
  for (int64_t i = first; i < n; i += step) {
    double xr = 0.0, xi = 0.0, orr = 0.0, oi = 0.0;
    for (int k = 0; k < nv; ++k) {
      double2 c, o;
      if (k < 64) {
        c = sc[0][k];
        o = sc[1][k];
      } else {
        c = coef_new[k];
        c.x *= sig[k];
        c.y *= sig[k];
        o = k < nvo ? coef_old[k] : make_double2(0.0, 0.0);
        o.x *= sig[k];
        o.y *= sig[k];
      }
      const double2 v = V[(int64_t)k * vstride + i];
      xr += v.x * c.x - v.y * c.y;
      xi += v.x * c.y + v.y * c.x;
      if (k < nvo) {
        orr += v.x * o.x - v.y * o.y;
        oi += v.x * o.y + v.y * o.x;
      }
    }
    if (!probe) res[i] = make_double2(xr, xi);
    const double dr = orr - xr, di = oi - xi;
    const double lhs = sqrt(dr * dr + di * di);
    const double rhs = 1e-8 + 1e-5 * sqrt(xr * xr + xi * xi);
    if (!(lhs <= rhs)) viol += 1.0;
  }
  const double sv = block_sum(viol, sh);
  if (threadIdx.x != 0) return;
  if (probe) {
    if (sv > 0.0) st[KS_PFAIL] = j;
    return;
  }
  // the verdict travels in the same atomic as the completion count, so no
  // partial sums and no device-wide fence are needed
  const int inc = 1 + (sv > 0.0 ? (1 << 12) : 0);
  const int tot = atomicAdd(st + KS_COUNT, inc) + inc;
  if ((tot & ((1 << 12) - 1)) == (int)gridDim.x) {
    st[KS_COUNT] = 0;
    if ((tot >> 12) == 0) {
      st[KS_RESIDX] = residx;
      st[KS_CONV] = j + 1;
      st[KS_STOP] = 1;
    }
  }
}

// End of a call: out = the converged combination (status 1) or the
// combination of MFINAL vectors (breakdown / full space, coefficients from
// the final k_tri_coef); nothing when the iteration budget ran out.
__global__ __launch_bounds__(kRedThreads) void k_final_out(const double2* __restrict__ V,
                                                           int64_t n, int64_t vstride,
                                                           const double2* __restrict__ coef,
                                                           const double* __restrict__ sig,
                                                           const double2* __restrict__ res0,
                                                           const double2* __restrict__ res1,
                                                           double2* __restrict__ out,
                                                           const int* st) {
  const int status = final_kind(st);
  if (status == 1) {
    const double2* src = st[KS_RESIDX] ? res1 : res0;
    for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedThreads)
      out[i] = src[i];
  } else if (status == 2 || status == 3) {
    const int nv = st[KS_MFINAL];
    for (int64_t i = blockIdx.x * (int64_t)kRedThreads + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kRedThreads) {
      double xr = 0.0, xi = 0.0;
      for (int j = 0; j < nv; ++j) {
        const double2 c = make_double2(coef[j].x * sig[j], coef[j].y * sig[j]);
        const double2 v = V[(int64_t)j * vstride + i];
        xr += v.x * c.x - v.y * c.y;
        xi += v.x * c.y + v.y * c.x;
      }
      out[i] = make_double2(xr, xi);
    }
  }
}

// info[0] = matvecs used; info[1] = 1 finished by the reference's rule,
// 0 iteration budget exhausted (out not written), -1 zero starting vector
__global__ void k_info(const int* st, int queued, int* info) {
  const int kind = final_kind(st);
  info[0] = kind == 0 ? queued : (kind == 1 ? st[KS_CONV] : st[KS_NMATVEC]);
  info[1] = kind == 0 ? 0 : (kind < 0 ? -1 : 1);
}

// Precision switch point of the mixed entry point, measured on this call's
// result y = sum_k c_k v_k (v_k normalised): the smallest j >= 2 such that
// every coefficient from c_{j-1} on has |c_k| < tau |v0|.  A matvec H v_j
// whose error is delta contributes about |c_{j-1}| delta to the result, so
// from that iteration on the FP32 product is as good as the FP64 one at
// the tolerance tau * 1e-7.  info[2] = 1 << 30 when the last coefficient is
// not below tau (no iteration of this call would have been low precision).
__global__ void k_switch_point(const int* st, const double2* __restrict__ coef_all, int kcap,
                               const double* __restrict__ scal, double tau, int* info) {
  const int kind = final_kind(st);
  int m = 0;
  const double2* c = nullptr;
  if (kind == 1) {
    m = st[KS_CONV];
    c = coef_all + (size_t)st[KS_RESIDX] * kcap;
  } else if (kind == 2 || kind == 3) {
    m = st[KS_MFINAL];
    c = coef_all + (size_t)2 * kcap;
  }
  int js = 1 << 30;
  if (c && m > 0) {
    const double thr = tau * scal[0];
    int last_big = -1;
    for (int k = m - 1; k >= 0; --k)
      if (!(sqrt(c[k].x * c[k].x + c[k].y * c[k].y) < thr)) {
        last_big = k;
        break;
      }
    if (last_big < m - 1) js = max(2, last_big + 2);
  }
  info[2] = js;
}

struct KrylovLayout {
  size_t off_wl, off_wl_lo, off_t1, off_t2, off_split, off_V, off_w, off_r0, off_r1, off_part,
      off_part2, off_scal, off_coef;
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
  // the FP32 copy of the permuted MPO for the low-precision iterations of
  // the mixed entry point (a few KB)
  k.off_wl_lo = o;
  if (op->nsite == 1)
    o += align_up((size_t)op->Dl * op->d * op->d * op->Dr * sizeof(float));
  k.off_t1 = o; o += h.t1_bytes;
  k.off_t2 = o; o += h.t2_bytes;
  k.off_split = o; o += h.part_bytes;
  k.split_bytes = h.part_bytes;
  k.off_V = o; o += vb * (size_t)kcap;
  k.off_w = o; o += vb;
  k.off_r0 = o; o += vb;
  k.off_r1 = o; o += vb;
  k.off_part = o; o += align_up(kMaxPartials * sizeof(double));
  k.off_part2 = o; o += align_up(2 * kMaxPartials * sizeof(double));
  k.off_scal = o; o += align_up((size_t)(3 * kcap + 17) * sizeof(double));
  k.off_coef = o; o += align_up((size_t)3 * kcap * 16);  // two checks, the final
  k.total = o;
  return k;
}

// Per-thread pinned host copy of the device state block (synchronous entry
// point) and the event it waits on.
struct HostStage {
  int* pin = nullptr;
  hipEvent_t ev = nullptr;
  ~HostStage() {
    if (pin) (void)hipHostFree(pin);
    if (ev) (void)hipEventDestroy(ev);
  }
};
static thread_local HostStage t_stage;

// One call's device buffers and launch parameters.
struct KrylovRun {
  const reno_hop_desc* op;
  // mixed entry point: iterations j >= j_switch apply op_lo (complex64
  // environments, float32 MPO, FP32 matrix cores) instead of op
  const reno_hop_desc* op_lo = nullptr;
  int64_t j_switch = INT64_MAX;
  void* wl_lo = nullptr;
  hipStream_t s;
  int64_t n;
  int kcap;
  KrylovLayout L;
  void *wl, *t1, *t2, *split;
  double2* V;
  size_t vstride;
  double2* w;
  double2* res[2];
  double* part;
  double* part2;
  double* scal;
  int* st;
  double* d_alpha;
  double* d_beta;
  double* d_sig;  // kcap + 1 entries
  double2* d_coef;
  int np, ew_blocks;
  double thr, dt_re, dt_im;
};

static int kcap_for(const reno_hop_desc* op, size_t ws_bytes) {
  int kcap = 2;
  while (kcap < kMaxKrylov && krylov_layout(op, kcap + 1).total <= ws_bytes) ++kcap;
  return kcap;
}

static void bind_run(KrylovRun& r, const reno_hop_desc* op, hipStream_t s, void* ws, int kcap,
                     double dt_re, double dt_im) {
  r.op = op;
  r.s = s;
  r.n = hop_vec_len(op);
  r.kcap = kcap;
  r.L = krylov_layout(op, kcap);
  char* base = (char*)ws;
  r.wl = base + r.L.off_wl;
  r.wl_lo = base + r.L.off_wl_lo;
  r.t1 = base + r.L.off_t1;
  r.t2 = base + r.L.off_t2;
  r.split = base + r.L.off_split;
  r.V = (double2*)(base + r.L.off_V);
  r.vstride = align_up((size_t)r.n * 16) / 16;
  r.w = (double2*)(base + r.L.off_w);
  r.res[0] = (double2*)(base + r.L.off_r0);
  r.res[1] = (double2*)(base + r.L.off_r1);
  r.part = (double*)(base + r.L.off_part);
  r.part2 = (double*)(base + r.L.off_part2);
  r.scal = (double*)(base + r.L.off_scal);  // [0] |v0|, [1] unused, [2..5] state ints
  r.st = (int*)(r.scal + 2);
  r.d_alpha = r.scal + 8;
  r.d_beta = r.scal + 8 + kcap;
  r.d_sig = r.scal + 8 + 2 * kcap;
  r.d_coef = (double2*)(base + r.L.off_coef);
  r.np = n_partials(r.n);
  r.ew_blocks = (int)((r.n + 255) / 256 < 2048 ? (r.n + 255) / 256 : 2048);
  r.thr = 100.0 * (double)r.n * 2.220446049250313e-16;  // krylov.py:72
  r.dt_re = dt_re;
  r.dt_im = dt_im;
}

// |v0|, state reset, V[0] = v0 / |v0| (krylov.py:42-45)
static int queue_start(KrylovRun& r, const void* v0) {
  if (r.op->nsite == 1)
    RENO_TRY(permute_w(r.op->prec, r.op->W, r.wl, r.op->Dl, r.op->d, r.op->Dr, r.s));
  if (r.op_lo && r.op_lo->nsite == 1)
    RENO_TRY(permute_w(r.op_lo->prec, r.op_lo->W, r.wl_lo, r.op_lo->Dl, r.op_lo->d, r.op_lo->Dr,
                       r.s));
  hipLaunchKernelGGL(k_norm2, dim3(r.np), dim3(kRedThreads), 0, r.s, (const double2*)v0, r.n,
                     r.part);
  hipLaunchKernelGGL(k_kry_init, dim3(1), dim3(kRedThreads), 0, r.s, (const double*)r.part, r.np,
                     r.scal, r.d_sig, r.st);
  hipLaunchKernelGGL(k_div, dim3(r.ew_blocks), dim3(256), 0, r.s, (const double2*)v0, r.V,
                     (const double*)r.scal, r.n, (const int*)r.st);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

// coefficient buffer: 0 / 1 for the checks (alternating, so a check sees
// the previous one's coefficients), 2 for the final combination
static double2* coef_buf(KrylovRun& r, int which) { return r.d_coef + (size_t)which * r.kcap; }

// k_tri_coef instantiation for up to m entries
static void launch_tri_coef(KrylovRun& r, int m, int final, double2* coef) {
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(1), dim3(64), 0, r.s, (const double*)r.d_alpha,
                       (const double*)r.d_beta, (const double*)r.scal, r.dt_re, r.dt_im, m, final,
                       (const int*)r.st, coef);
  };
  if (m <= 64) go(k_tri_coef<1>);
  else if (m <= 128) go(k_tri_coef<2>);
  else if (m <= 256) go(k_tri_coef<4>);
  else if (m <= 512) go(k_tri_coef<8>);
  else go(k_tri_coef<16>);
}

// iteration j (krylov.py:55-81): matvec, alpha_j; unless j is the last
// possible one: recurrence, beta_j with the breakdown test, V[j+1]; at even
// j > 3 the combination and (from the second check on) the allclose test.
static int queue_iter(KrylovRun& r, int64_t j) {
  const int* stop = r.st + KS_STOP;
  double2* vj = r.V + (size_t)j * r.vstride;
  // matvec with the dot partials of Re<w, V[j]> fused into its last stage
  const bool lo = r.op_lo && j >= r.j_switch;
  RENO_TRY(hop_apply_core(lo ? r.op_lo : r.op, lo ? r.wl_lo : r.wl, vj, r.w, r.t1, r.t2, r.s, stop,
                          r.split, r.L.split_bytes, vj, r.part, r.np));
  if (j == r.n - 1) {
    hipLaunchKernelGGL(k_fold_alpha, dim3(1), dim3(kRedThreads), 0, r.s, (const double*)r.part,
                       (const double*)(r.part2 + (size_t)(j & 1) * kMaxPartials), r.np, r.d_alpha,
                       r.d_beta, r.d_sig, (int)j, r.thr, r.st);
    hipLaunchKernelGGL(k_mark_full, dim3(1), dim3(1), 0, r.s, (int)r.n, r.st);
    RENO_CHECK_LAUNCH();
    return RENO_OK;
  }
  // no slot for V[j+1] (the last queued iteration of a full workspace): the
  // recurrence still runs for beta_j and the breakdown test, into w
  double2* vnext = j + 1 < r.kcap ? r.V + (size_t)(j + 1) * r.vstride : r.w;
  // the norm partials of V[j+1] go to the other half of part2 (the current
  // half holds those of V[j], still being folded by every block)
  double* nprev = r.part2 + (size_t)(j & 1) * kMaxPartials;
  double* nnext = r.part2 + (size_t)((j + 1) & 1) * kMaxPartials;
  hipLaunchKernelGGL(k_recur, dim3(r.np), dim3(kRedThreads), 0, r.s, (const double2*)r.w,
                     (const double2*)vj, j > 0 ? (const double2*)(vj - r.vstride) : (const double2*)nullptr,
                     vnext, (const double*)r.part, (const double*)nprev, r.np, r.d_alpha, r.d_beta,
                     r.d_sig, (int)j, r.n, r.thr, nnext, r.st);
  if (j > 3 && j % 2 == 0) {
    const int m = (int)j + 1;
    const int cb = (int)((j / 2) & 1);
    launch_tri_coef(r, m, 0, coef_buf(r, cb));
    if (j > 4) {  // the first check has nothing to compare with
      const int nprobe = (int)((r.n + 4095) / 4096);
      hipLaunchKernelGGL(k_check, dim3(nprobe), dim3(kRedThreads), 0, r.s, (const double2*)r.V, r.n,
                         (int64_t)r.vstride, m, (const double2*)coef_buf(r, cb),
                         (const double2*)coef_buf(r, cb ^ 1), (const double*)r.d_sig, r.res[cb],
                         (int)j, cb, 1, r.st);
      hipLaunchKernelGGL(k_check, dim3(r.np), dim3(kRedThreads), 0, r.s, (const double2*)r.V, r.n,
                         (int64_t)r.vstride, m, (const double2*)coef_buf(r, cb),
                         (const double2*)coef_buf(r, cb ^ 1), (const double*)r.d_sig, r.res[cb],
                         (int)j, cb, 0, r.st);
    }
  }
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

// final combination / copy into out; m_max bounds MFINAL
static int queue_final(KrylovRun& r, void* out, int m_max) {
  launch_tri_coef(r, m_max, 1, coef_buf(r, 2));
  hipLaunchKernelGGL(k_final_out, dim3(r.np), dim3(kRedThreads), 0, r.s, (const double2*)r.V,
                     r.n, (int64_t)r.vstride, (const double2*)coef_buf(r, 2), (const double*)r.d_sig,
                     (const double2*)r.res[0], (const double2*)r.res[1], (double2*)out,
                     (const int*)r.st);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

static int check_args(const reno_hop_desc* op, const void* v0, const void* out) {
  RENO_TRY(check_hop(op));
  if (!v0 || !out) {
    set_error("expm_krylov: null argument");
    return RENO_EINVAL;
  }
  return RENO_OK;
}

}  // namespace reno

using namespace reno;

extern "C" size_t reno_krylov_workspace_bytes(const reno_hop_desc* op, int kcap) {
  if (check_hop(op) != RENO_OK || kcap < 2 || kcap > kMaxKrylov) return 0;
  return krylov_layout(op, kcap).total;
}

extern "C" int reno_expm_krylov(const reno_hop_desc* op, double dt_re, double dt_im,
                                const void* v0, void* out, int* n_matvec, size_t* needed,
                                void* ws, size_t ws_bytes, void* stream) {
  RENO_TRY(check_args(op, v0, out));
  if (!n_matvec) {
    set_error("expm_krylov: null n_matvec");
    return RENO_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int kcap = kcap_for(op, ws_bytes);
  if (!ws || krylov_layout(op, kcap).total > ws_bytes) {
    if (needed) *needed = krylov_layout(op, 16).total;
    set_error("expm_krylov: workspace %zu too small", ws_bytes);
    return RENO_ENOSPACE;
  }
  if (!t_stage.pin) {
    if (hipHostMalloc((void**)&t_stage.pin, KS_INTS * sizeof(int), hipHostMallocDefault) !=
        hipSuccess) {
      t_stage.pin = nullptr;
      set_error("expm_krylov: pinned host buffer allocation failed");
      return RENO_EHIP;
    }
  }
  if (!t_stage.ev) RENO_HIP_CHECK(hipEventCreateWithFlags(&t_stage.ev, hipEventDisableTiming));
  KrylovRun r;
  bind_run(r, op, s, ws, kcap, dt_re, dt_im);
  RENO_TRY(queue_start(r, v0));
  // chunks of iterations; the host reads the device state once per chunk
  int64_t j = 0;
  int64_t chunk = 12;
  const int64_t jmax = std::min<int64_t>(r.n, kcap);
  while (true) {
    const int64_t jend = std::min<int64_t>(j + chunk, jmax);
    for (; j < jend; ++j) RENO_TRY(queue_iter(r, j));
    RENO_HIP_CHECK(hipMemcpyAsync(t_stage.pin, r.st, KS_INTS * sizeof(int),
                                  hipMemcpyDeviceToHost, s));
    RENO_HIP_CHECK(hipEventRecord(t_stage.ev, s));
    RENO_HIP_CHECK(hipEventSynchronize(t_stage.ev));
    const int kind = final_kind(t_stage.pin);
    if (kind < 0) {
      set_error("expm_krylov: zero starting vector (krylov.py:44 assert)");
      return RENO_EINVAL;
    }
    if (kind > 0) {
      *n_matvec = kind == 1 ? t_stage.pin[KS_CONV] : t_stage.pin[KS_NMATVEC];
      return queue_final(r, out, (int)jend);
    }
    if (j >= jmax) {
      if (needed) *needed = krylov_layout(op, std::min(2 * kcap, kMaxKrylov)).total;
      set_error("expm_krylov: %d Krylov vectors do not fit the workspace", kcap);
      return kcap >= kMaxKrylov ? RENO_EINVAL : RENO_ENOSPACE;
    }
    chunk = 8;
  }
}

extern "C" int reno_expm_krylov_async(const reno_hop_desc* op, double dt_re, double dt_im,
                                      const void* v0, void* out, int max_matvec, int* info,
                                      void* ws, size_t ws_bytes, void* stream) {
  RENO_TRY(check_args(op, v0, out));
  if (!info || max_matvec < 1) {
    set_error("expm_krylov_async: null info or max_matvec < 1");
    return RENO_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int kcap = kcap_for(op, ws_bytes);
  const int64_t jmax = std::min<int64_t>(hop_vec_len(op), max_matvec);
  if (!ws || krylov_layout(op, kcap).total > ws_bytes || jmax > kcap) {
    set_error("expm_krylov_async: workspace %zu holds %d Krylov vectors, %lld requested",
              ws_bytes, kcap, (long long)jmax);
    return RENO_ENOSPACE;
  }
  KrylovRun r;
  bind_run(r, op, s, ws, kcap, dt_re, dt_im);
  RENO_TRY(queue_start(r, v0));
  for (int64_t j = 0; j < jmax; ++j) RENO_TRY(queue_iter(r, j));
  RENO_TRY(queue_final(r, out, (int)jmax));
  hipLaunchKernelGGL(k_info, dim3(1), dim3(1), 0, s, (const int*)r.st, (int)jmax, info);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" int reno_expm_krylov_mixed_async(const reno_hop_desc* op, const reno_hop_desc* op_lo,
                                            int j_switch, double tau, double dt_re, double dt_im,
                                            const void* v0, void* out, int max_matvec, int* info,
                                            void* ws, size_t ws_bytes, void* stream) {
  RENO_TRY(check_args(op, v0, out));
  RENO_TRY(check_hop(op_lo));
  if (op->prec != RENO_PREC_FP64 || op_lo->prec != RENO_PREC_MIXED || op->nsite != op_lo->nsite ||
      op->Ml != op_lo->Ml || op->d != op_lo->d || op->Mr != op_lo->Mr || op->Dl != op_lo->Dl ||
      op->Dr != op_lo->Dr) {
    set_error("expm_krylov_mixed: op must be FP64 and op_lo the same operator in the mixed mode");
    return RENO_EINVAL;
  }
  if (!info || max_matvec < 1 || !(tau >= 0.0)) {
    set_error("expm_krylov_mixed: null info, max_matvec < 1 or negative tau");
    return RENO_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  const int kcap = kcap_for(op, ws_bytes);
  const int64_t jmax = std::min<int64_t>(hop_vec_len(op), max_matvec);
  if (!ws || krylov_layout(op, kcap).total > ws_bytes || jmax > kcap) {
    set_error("expm_krylov_mixed: workspace %zu holds %d Krylov vectors, %lld requested",
              ws_bytes, kcap, (long long)jmax);
    return RENO_ENOSPACE;
  }
  KrylovRun r;
  bind_run(r, op, s, ws, kcap, dt_re, dt_im);
  r.op_lo = op_lo;
  r.j_switch = std::max(j_switch, 1);
  RENO_TRY(queue_start(r, v0));
  for (int64_t j = 0; j < jmax; ++j) RENO_TRY(queue_iter(r, j));
  RENO_TRY(queue_final(r, out, (int)jmax));
  hipLaunchKernelGGL(k_info, dim3(1), dim3(1), 0, s, (const int*)r.st, (int)jmax, info);
  hipLaunchKernelGGL(k_switch_point, dim3(1), dim3(1), 0, s, (const int*)r.st,
                     (const double2*)r.d_coef, r.kcap, (const double*)r.scal, tau, info);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}
