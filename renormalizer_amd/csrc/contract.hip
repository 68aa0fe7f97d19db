// Effective-Hamiltonian matvec (hop_expr) and environment update
// (contract_one_site) as chains of batched complex GEMMs.
//
// Index names follow the reference einsum strings (hop_expr.py:63-79,
// lib.py:200-248).  Every stage is one launch of the GEMM in gemm.hip; the MPO
// stage is a batched GEMM with a real A operand (the site MPO, permuted once
// into "Wl" layout for the left-type contractions).
#include <algorithm>

#include "contract.h"

namespace reno {

// Wl[(dd*Dr + f), (b*d + e)] = W[b, dd, e, f]
template <typename T>
__global__ void k_permute_w(const T* __restrict__ W, T* __restrict__ Wl, int Dl, int d,
                            int Dr) {
  const int64_t total = (int64_t)Dl * d * d * Dr;
  for (int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t r = idx;
    const int f = r % Dr;
    r /= Dr;
    const int e = r % d;
    r /= d;
    const int dd = r % d;
    const int b = (int)(r / d);
    Wl[((int64_t)dd * Dr + f) * ((int64_t)Dl * d) + (int64_t)b * d + e] = W[idx];
  }
}

int permute_w(int prec, const void* W, void* Wl, int Dl, int d, int Dr, hipStream_t s) {
  const int64_t total = (int64_t)Dl * d * d * Dr;
  if (total == 0) return RENO_OK;
  const int threads = 256;
  const int blocks = (int)((total + threads - 1) / threads < 1024 ? (total + threads - 1) / threads : 1024);
  if (prec == RENO_PREC_FP64)
    hipLaunchKernelGGL(k_permute_w<double>, dim3(blocks), dim3(threads), 0, s,
                       (const double*)W, (double*)Wl, Dl, d, Dr);
  else
    hipLaunchKernelGGL(k_permute_w<float>, dim3(blocks), dim3(threads), 0, s,
                       (const float*)W, (float*)Wl, Dl, d, Dr);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

static inline int cplx_kind(int prec) { return prec == RENO_PREC_FP64 ? ST_Z128 : ST_C64; }
static inline int real_kind(int prec) { return prec == RENO_PREC_FP64 ? ST_F64 : ST_F32; }
static inline size_t real_bytes(int prec) { return prec == RENO_PREC_FP64 ? 8 : 4; }

int check_hop(const reno_hop_desc* op) {
  if (!op) {
    set_error("hop: null descriptor");
    return RENO_EINVAL;
  }
  if (op->nsite != 0 && op->nsite != 1) {
    set_error("hop: nsite %d not supported (0 or 1)", op->nsite);
    return RENO_EINVAL;
  }
  if (op->prec != RENO_PREC_FP64 && op->prec != RENO_PREC_MIXED) {
    set_error("hop: unknown precision mode %d", op->prec);
    return RENO_EINVAL;
  }
  if (op->Ml <= 0 || op->Mr <= 0 || op->d <= 0 || op->Dl <= 0 || op->Dr <= 0) {
    set_error("hop: non-positive dimension Ml=%d d=%d Mr=%d Dl=%d Dr=%d", op->Ml, op->d,
              op->Mr, op->Dl, op->Dr);
    return RENO_EINVAL;
  }
  if (op->nsite == 0 && (op->d != 1 || op->Dl != op->Dr)) {
    set_error("hop: nsite 0 needs d == 1 and Dl == Dr");
    return RENO_EINVAL;
  }
  if (!op->L || !op->R || (op->nsite == 1 && !op->W)) {
    set_error("hop: null L/R/W");
    return RENO_EINVAL;
  }
  return RENO_OK;
}

HopLayout hop_layout(const reno_hop_desc* op) {
  HopLayout h{};
  const size_t ce = elem_bytes(op->prec);
  const int64_t Ml = op->Ml, d = op->d, Mr = op->Mr, Dl = op->Dl, Dr = op->Dr;
  const int p = op->prec;
  if (op->nsite == 1) {
    h.wl_bytes = align_up((size_t)(Dl * d * d * Dr) * real_bytes(op->prec));
    h.t1_bytes = align_up((size_t)(Ml * Dl * d * Mr) * ce);
    h.t2_bytes = align_up((size_t)(Ml * d * Dr * Mr) * ce);
    h.part_bytes = std::max({splitk_bytes(p, Ml * Dl, d * Mr, Ml, 1),
                             splitk_bytes(p, d * Dr, Mr, Dl * d, Ml),
                             splitk_bytes(p, Ml * d, Mr, Dr * Mr, 1)});
  } else {
    h.wl_bytes = 0;
    h.t1_bytes = align_up((size_t)(Ml * Dl * Mr) * ce);
    h.t2_bytes = 0;
    h.part_bytes = std::max(splitk_bytes(p, Ml * Dl, Mr, Ml, 1), splitk_bytes(p, Ml, Mr, Dl * Mr, 1));
  }
  h.part_bytes = align_up(h.part_bytes);
  h.total = h.wl_bytes + h.t1_bytes + h.t2_bytes + h.part_bytes;
  return h;
}

// out = H_eff * cin with pre-permuted Wl (nsite 1) and scratch t1/t2.
int hop_apply_core(const reno_hop_desc* op, const void* Wl, const void* cin, void* cout,
                   void* t1, void* t2, hipStream_t s, const int* skip, void* part,
                   size_t part_bytes, const double2* dot_v, double* dot_part, int dot_np) {
  const int prec = op->prec;
  const int ck = cplx_kind(prec);
  const int64_t Ml = op->Ml, d = op->d, Mr = op->Mr, Dl = op->Dl, Dr = op->Dr;
  if (op->nsite == 1) {
    // S1: T1[(a,b),(e,k)] = L[(a,b),c] C[c,(e,k)]
    GemmCall g1{};
    g1.M = (int)(Ml * Dl); g1.N = (int)(d * Mr); g1.K = (int)Ml; g1.batch = 1; g1.prec = prec;
    g1.A = op->L; g1.a_kind = ck; g1.sAm = Ml; g1.sAk = 1;
    g1.B = cin; g1.b_kind = ST_Z128; g1.sBk = d * Mr; g1.sBn = 1;
    g1.C = t1; g1.c_kind = ck; g1.sCm = d * Mr; g1.sCn = 1;
    g1.skip = skip;
    g1.part = part;
    g1.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g1, s));
    // S2: T2[a,(dd,f),k] = Wl[(dd,f),(b,e)] T1[a,(b,e),k]   (batched over a)
    GemmCall g2{};
    g2.M = (int)(d * Dr); g2.N = (int)Mr; g2.K = (int)(Dl * d); g2.batch = (int)Ml; g2.prec = prec;
    g2.A = Wl; g2.a_kind = real_kind(prec); g2.sAm = Dl * d; g2.sAk = 1; g2.sAb = 0;
    g2.B = t1; g2.b_kind = ck; g2.sBk = Mr; g2.sBn = 1; g2.sBb = Dl * d * Mr;
    g2.C = t2; g2.c_kind = ck; g2.sCm = Mr; g2.sCn = 1; g2.sCb = d * Dr * Mr;
    g2.skip = skip;
    g2.part = part;
    g2.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g2, s));
    // S3: out[(a,dd),l] = T2[(a,dd),(f,k)] R[l,(f,k)]
    GemmCall g3{};
    g3.M = (int)(Ml * d); g3.N = (int)Mr; g3.K = (int)(Dr * Mr); g3.batch = 1; g3.prec = prec;
    g3.A = t2; g3.a_kind = ck; g3.sAm = Dr * Mr; g3.sAk = 1;
    g3.B = op->R; g3.b_kind = ck; g3.sBk = 1; g3.sBn = Dr * Mr;
    g3.C = cout; g3.c_kind = ST_Z128; g3.sCm = Mr; g3.sCn = 1;
    g3.skip = skip;
    g3.part = part;
    g3.part_bytes = part_bytes;
    g3.dot_v = dot_v;
    g3.dot_part = dot_part;
    g3.dot_np = dot_np;
    RENO_TRY(launch_gemm(g3, s));
  } else {
    const int64_t D = Dl;
    // S1: T[(a,b),k] = L[(a,b),c] S[c,k]
    GemmCall g1{};
    g1.M = (int)(Ml * D); g1.N = (int)Mr; g1.K = (int)Ml; g1.batch = 1; g1.prec = prec;
    g1.A = op->L; g1.a_kind = ck; g1.sAm = Ml; g1.sAk = 1;
    g1.B = cin; g1.b_kind = ST_Z128; g1.sBk = Mr; g1.sBn = 1;
    g1.C = t1; g1.c_kind = ck; g1.sCm = Mr; g1.sCn = 1;
    g1.skip = skip;
    g1.part = part;
    g1.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g1, s));
    // S2: out[a,l] = T[a,(b,k)] R[l,(b,k)]
    GemmCall g2{};
    g2.M = (int)Ml; g2.N = (int)Mr; g2.K = (int)(D * Mr); g2.batch = 1; g2.prec = prec;
    g2.A = t1; g2.a_kind = ck; g2.sAm = D * Mr; g2.sAk = 1;
    g2.B = op->R; g2.b_kind = ck; g2.sBk = 1; g2.sBn = D * Mr;
    g2.C = cout; g2.c_kind = ST_Z128; g2.sCm = Mr; g2.sCn = 1;
    g2.skip = skip;
    g2.part = part;
    g2.part_bytes = part_bytes;
    g2.dot_v = dot_v;
    g2.dot_part = dot_part;
    g2.dot_np = dot_np;
    RENO_TRY(launch_gemm(g2, s));
  }
  return RENO_OK;
}

static size_t env_part_bytes(int p, int64_t Ml, int64_t d, int64_t Mr, int64_t Dl, int64_t Dr) {
  return align_up(std::max({splitk_bytes(p, Ml * Dl, d * Mr, Ml, 1), splitk_bytes(p, d * Dr, Mr, Dl * d, Ml),
                            splitk_bytes(p, Mr, Dr * Mr, Ml * d, 1), splitk_bytes(p, Ml * d, Mr, Mr, Dr),
                            splitk_bytes(p, Dl * d, Mr, d * Dr, Ml), splitk_bytes(p, Ml, Ml, d * Mr, Dl)}));
}

size_t env_ws_bytes(int prec, int64_t Ml, int64_t d, int64_t Mr, int64_t Dl, int64_t Dr) {
  const size_t ce = elem_bytes(prec);
  size_t wl = align_up((size_t)(Dl * d * d * Dr) * real_bytes(prec));
  size_t y = align_up((size_t)(Ml * d * Mr * (Dl > Dr ? Dl : Dr)) * ce);
  return wl + 2 * y + env_part_bytes(prec, Ml, d, Mr, Dl, Dr);
}

int env_update_core(int domain, int prec, int64_t Ml, int64_t d, int64_t Mr, int64_t Dl,
                    int64_t Dr, const void* env_in, const void* ms, const void* ms_conj,
                    const void* W, void* env_out, void* ws, hipStream_t s) {
  const int ck = cplx_kind(prec);
  const size_t ce = elem_bytes(prec);
  char* p = (char*)ws;
  void* wl = p;
  p += align_up((size_t)(Dl * d * d * Dr) * real_bytes(prec));
  void* y = p;
  p += align_up((size_t)(Ml * d * Mr * (Dl > Dr ? Dl : Dr)) * ce);
  void* z = p;
  p += align_up((size_t)(Ml * d * Mr * (Dl > Dr ? Dl : Dr)) * ce);
  void* part = p;
  const size_t part_bytes = env_part_bytes(prec, Ml, d, Mr, Dl, Dr);
  // bra operand: conj(ms) unless an explicit (already conjugated) bra is given
  const void* bra = ms_conj ? ms_conj : ms;
  const int conj_bra = ms_conj ? 0 : 1;
  if (domain == 0) {
    RENO_TRY(permute_w(prec, W, wl, (int)Dl, (int)d, (int)Dr, s));
    // S1: Y[(a,b),(e,h)] = L[(a,b),c] ms[c,(e,h)]
    GemmCall g1{};
    g1.M = (int)(Ml * Dl); g1.N = (int)(d * Mr); g1.K = (int)Ml; g1.batch = 1; g1.prec = prec;
    g1.A = env_in; g1.a_kind = ck; g1.sAm = Ml; g1.sAk = 1;
    g1.B = ms; g1.b_kind = ST_Z128; g1.sBk = d * Mr; g1.sBn = 1;
    g1.C = y; g1.c_kind = ck; g1.sCm = d * Mr; g1.sCn = 1;
    g1.part = part;
    g1.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g1, s));
    // S2: Z[a,(dd,g),h] = Wl[(dd,g),(b,e)] Y[a,(b,e),h]
    GemmCall g2{};
    g2.M = (int)(d * Dr); g2.N = (int)Mr; g2.K = (int)(Dl * d); g2.batch = (int)Ml; g2.prec = prec;
    g2.A = wl; g2.a_kind = real_kind(prec); g2.sAm = Dl * d; g2.sAk = 1; g2.sAb = 0;
    g2.B = y; g2.b_kind = ck; g2.sBk = Mr; g2.sBn = 1; g2.sBb = Dl * d * Mr;
    g2.C = z; g2.c_kind = ck; g2.sCm = Mr; g2.sCn = 1; g2.sCb = d * Dr * Mr;
    g2.part = part;
    g2.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g2, s));
    // S3: out[f,(g,h)] = bra*[(a,dd),f] Z[(a,dd),(g,h)]
    GemmCall g3{};
    g3.M = (int)Mr; g3.N = (int)(Dr * Mr); g3.K = (int)(Ml * d); g3.batch = 1; g3.prec = prec;
    g3.A = bra; g3.a_kind = ST_Z128; g3.sAm = 1; g3.sAk = Mr; g3.conjA = conj_bra;
    g3.B = z; g3.b_kind = ck; g3.sBk = Dr * Mr; g3.sBn = 1;
    g3.C = env_out; g3.c_kind = ck; g3.sCm = Dr * Mr; g3.sCn = 1;
    g3.part = part;
    g3.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g3, s));
  } else {
    // S1: Y[(h,e), b, a] = ms[(h,e),c] R[a,b,c]   (batched over b)
    GemmCall g1{};
    g1.M = (int)(Ml * d); g1.N = (int)Mr; g1.K = (int)Mr; g1.batch = (int)Dr; g1.prec = prec;
    g1.A = ms; g1.a_kind = ST_Z128; g1.sAm = Mr; g1.sAk = 1; g1.sAb = 0;
    g1.B = env_in; g1.b_kind = ck; g1.sBk = 1; g1.sBn = Dr * Mr; g1.sBb = Mr;
    g1.C = y; g1.c_kind = ck; g1.sCm = Dr * Mr; g1.sCn = 1; g1.sCb = Mr;
    g1.part = part;
    g1.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g1, s));
    // S2: Z[h,(g,dd),a] = W[(g,dd),(e,b)] Y[h,(e,b),a]   (batched over h;
    //     W viewed as the (Dl*d) x (d*Dr) matrix it already is)
    GemmCall g2{};
    g2.M = (int)(Dl * d); g2.N = (int)Mr; g2.K = (int)(d * Dr); g2.batch = (int)Ml; g2.prec = prec;
    g2.A = W; g2.a_kind = real_kind(prec); g2.sAm = d * Dr; g2.sAk = 1; g2.sAb = 0;
    g2.B = y; g2.b_kind = ck; g2.sBk = Mr; g2.sBn = 1; g2.sBb = d * Dr * Mr;
    g2.C = z; g2.c_kind = ck; g2.sCm = Mr; g2.sCn = 1; g2.sCb = Dl * d * Mr;
    g2.part = part;
    g2.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g2, s));
    // S3: out[f,g,h] = bra*[f,(dd,a)] Z[h,g,(dd,a)]   (batched over g)
    GemmCall g3{};
    g3.M = (int)Ml; g3.N = (int)Ml; g3.K = (int)(d * Mr); g3.batch = (int)Dl; g3.prec = prec;
    g3.A = bra; g3.a_kind = ST_Z128; g3.sAm = d * Mr; g3.sAk = 1; g3.sAb = 0; g3.conjA = conj_bra;
    g3.B = z; g3.b_kind = ck; g3.sBk = 1; g3.sBn = Dl * d * Mr; g3.sBb = d * Mr;
    g3.C = env_out; g3.c_kind = ck; g3.sCm = Dl * Ml; g3.sCn = 1; g3.sCb = Ml;
    g3.part = part;
    g3.part_bytes = part_bytes;
    RENO_TRY(launch_gemm(g3, s));
  }
  return RENO_OK;
}

// out[g] = sum_{f,h} X[f,g,h] Y[f,h]: one workgroup per g, fixed-order tree
// reduction (deterministic).  X, Y complex128 (kind 0) or complex64 (kind 1).
template <typename TX, typename TY>
__global__ __launch_bounds__(256) void k_env_close(const TX* __restrict__ X,
                                                   const TY* __restrict__ Y, int64_t F,
                                                   int64_t G, int64_t H, double2* out) {
  __shared__ double sr[256], si[256];
  const int64_t g = blockIdx.x;
  double ar = 0.0, ai = 0.0;
  for (int64_t t = threadIdx.x; t < F * H; t += 256) {
    const int64_t f = t / H, h = t - (t / H) * H;
    const TX x = X[(f * G + g) * H + h];
    const TY y = Y[f * H + h];
    ar += (double)x.x * (double)y.x - (double)x.y * (double)y.y;
    ai += (double)x.x * (double)y.y + (double)x.y * (double)y.x;
  }
  sr[threadIdx.x] = ar;
  si[threadIdx.x] = ai;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if (threadIdx.x < off) {
      sr[threadIdx.x] += sr[threadIdx.x + off];
      si[threadIdx.x] += si[threadIdx.x + off];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[g] = make_double2(sr[0], si[0]);
}

}  // namespace reno

using namespace reno;

extern "C" int reno_env_close(int prec, int F, int G, int H, const void* X, const void* Y,
                              void* out, void* stream) {
  if (F <= 0 || G <= 0 || H <= 0 || !X || !Y || !out || G > 65535) {
    set_error("env_close: bad arguments F=%d G=%d H=%d", F, G, H);
    return RENO_EINVAL;
  }
  hipStream_t s = (hipStream_t)stream;
  if (prec == RENO_PREC_FP64)
    hipLaunchKernelGGL((k_env_close<double2, double2>), dim3(G), dim3(256), 0, s,
                       (const double2*)X, (const double2*)Y, (int64_t)F, (int64_t)G, (int64_t)H,
                       (double2*)out);
  else
    hipLaunchKernelGGL((k_env_close<float2, float2>), dim3(G), dim3(256), 0, s,
                       (const float2*)X, (const float2*)Y, (int64_t)F, (int64_t)G, (int64_t)H,
                       (double2*)out);
  RENO_CHECK_LAUNCH();
  return RENO_OK;
}

extern "C" size_t reno_hop_workspace_bytes(const reno_hop_desc* op) {
  if (check_hop(op) != RENO_OK) return 0;
  return hop_layout(op).total;
}

extern "C" int reno_hop_apply(const reno_hop_desc* op, const void* c_in, void* c_out,
                              void* ws, size_t ws_bytes, void* stream) {
  RENO_TRY(check_hop(op));
  if (!c_in || !c_out || c_in == c_out) {
    set_error("hop_apply: null or aliased vectors");
    return RENO_EINVAL;
  }
  const HopLayout h = hop_layout(op);
  if (ws_bytes < h.total || (h.total && !ws)) {
    set_error("hop_apply: workspace %zu < %zu", ws_bytes, h.total);
    return RENO_ENOSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* p = (char*)ws;
  void* wl = p;
  void* t1 = p + h.wl_bytes;
  void* t2 = p + h.wl_bytes + h.t1_bytes;
  void* part = p + h.wl_bytes + h.t1_bytes + h.t2_bytes;
  if (op->nsite == 1) RENO_TRY(permute_w(op->prec, op->W, wl, op->Dl, op->d, op->Dr, s));
  return hop_apply_core(op, wl, c_in, c_out, t1, t2, s, nullptr, part, h.part_bytes);
}

extern "C" int reno_autotune(const reno_hop_desc* op, const void* c_in, void* c_out, void* ws,
                             size_t ws_bytes, int reps, reno_tile_cfg* out, int max_out,
                             int* n_out, void* stream) {
  if (n_out) *n_out = 0;
  RENO_TRY(check_hop(op));
  if (!c_in || !c_out || c_in == c_out) {
    set_error("autotune: null or aliased vectors");
    return RENO_EINVAL;
  }
  if (reps < 1 || reps > 1000) {
    set_error("autotune: reps %d outside [1, 1000]", reps);
    return RENO_EINVAL;
  }
  const HopLayout h = hop_layout(op);
  if (ws_bytes < h.total || (h.total && !ws)) {
    set_error("autotune: workspace %zu < %zu", ws_bytes, h.total);
    return RENO_ENOSPACE;
  }
  hipStream_t s = (hipStream_t)stream;
  char* p = (char*)ws;
  void* wl = p;
  void* t1 = p + h.wl_bytes;
  void* t2 = p + h.wl_bytes + h.t1_bytes;
  void* part = p + h.wl_bytes + h.t1_bytes + h.t2_bytes;
  // the stages' GEMM calls, as the matvec would launch them
  std::vector<GemmCall> calls;
  gemm_record_begin(&calls);
  const int rec = hop_apply_core(op, wl, c_in, c_out, t1, t2, s, nullptr, part, h.part_bytes);
  gemm_record_end();
  RENO_TRY(rec);
  bool all = true;
  for (const GemmCall& g : calls) all = all && gemm_is_tuned(g);
  if (all) return RENO_OK;
  // one real matvec first, so every stage reads its true operands
  if (op->nsite == 1) RENO_TRY(permute_w(op->prec, op->W, wl, op->Dl, op->d, op->Dr, s));
  RENO_TRY(hop_apply_core(op, wl, c_in, c_out, t1, t2, s, nullptr, part, h.part_bytes));
  int n = 0;
  for (const GemmCall& g : calls) {
    if (gemm_is_tuned(g)) continue;
    reno_tile_cfg cfg{};
    RENO_TRY(gemm_tune(g, s, reps, &cfg));
    if (out && n < max_out) out[n] = cfg;
    ++n;
  }
  if (n_out) *n_out = n < max_out ? n : max_out;
  return RENO_OK;
}

extern "C" int reno_autotune_reset(void) { return gemm_tune_reset(); }

extern "C" int reno_autotune_entries(void) { return gemm_tune_entries(); }

extern "C" size_t reno_env_workspace_bytes(int prec, int Ml, int d, int Mr, int Dl, int Dr) {
  return env_ws_bytes(prec, Ml, d, Mr, Dl, Dr);
}

extern "C" int reno_env_update(int domain, int prec, int Ml, int d, int Mr, int Dl, int Dr,
                               const void* env_in, const void* ms, const void* ms_conj,
                               const void* W, void* env_out, void* ws, size_t ws_bytes,
                               void* stream) {
  if (domain != 0 && domain != 1) {
    set_error("env_update: domain %d (0 = L, 1 = R)", domain);
    return RENO_EINVAL;
  }
  if (prec != RENO_PREC_FP64 && prec != RENO_PREC_MIXED) {
    set_error("env_update: unknown precision mode %d", prec);
    return RENO_EINVAL;
  }
  if (Ml <= 0 || d <= 0 || Mr <= 0 || Dl <= 0 || Dr <= 0 || !env_in || !ms || !W || !env_out) {
    set_error("env_update: bad arguments");
    return RENO_EINVAL;
  }
  const size_t need = env_ws_bytes(prec, Ml, d, Mr, Dl, Dr);
  if (ws_bytes < need || !ws) {
    set_error("env_update: workspace %zu < %zu", ws_bytes, need);
    return RENO_ENOSPACE;
  }
  return env_update_core(domain, prec, Ml, d, Mr, Dl, Dr, env_in, ms, ms_conj, W, env_out, ws,
                         (hipStream_t)stream);
}
