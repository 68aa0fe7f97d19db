/*
 * reno_hip.h -- C ABI of libreno_hip.so, the MI355X (gfx950) kernels behind the
 * one-site TDVP-PS / DMRG local update of Renormalizer.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, returns an
 * int status (RENO_OK == 0, negative on error) and never throws.  Device
 * buffers are caller-owned; the library launches asynchronously on the HIP
 * stream it is given and allocates no device memory of its own except inside
 * rocSOLVER (block QR/RQ), whose handle keeps a per-device pool.
 *
 * Tensor layouts are those of the reference (row-major numpy):
 *   MPS site          (M_l, d, M_r)            complex
 *   MPO site W        (D_l, d_out, d_in, D_r)  real
 *   environment L/R   (bra, mpo, ket)          complex
 *   bond matrix       (M_l, M_r)               complex
 * Complex numbers are interleaved (re, im) pairs: complex128 in FP64 mode,
 * complex64 for environments / MPO in the mixed mode (see reno_prec).
 */
#ifndef RENO_HIP_H
#define RENO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define RENO_OK 0
#define RENO_EINVAL (-1)   /* bad argument / shape mismatch (reference: AssertionError) */
#define RENO_EHIP (-2)     /* HIP runtime error */
#define RENO_ENOSPACE (-3) /* caller workspace too small; *needed bytes reported */
#define RENO_ESOLVER (-4)  /* rocSOLVER failure */
#define RENO_EQN (-5)      /* no valid quantum-number block (reference: ValueError) */

/* precision modes */
#define RENO_PREC_FP64 0  /* complex128 everywhere, FP64 MFMA */
#define RENO_PREC_MIXED 1 /* complex64 environments + FP32 MFMA contractions,
                             complex128 Krylov vectors, FP64 reductions */

/* Effective-Hamiltonian operator, reference hop_expr (hop_expr.py:7-116).
 * nsite = 1: out[a,d,l] = sum L[a,b,c] W[b,d,e,f] R[l,f,k] C[c,e,k]  (hop_expr.py:68-79)
 * nsite = 0: out[a,l]   = sum L[a,b,c] R[l,b,k] S[c,k]               (hop_expr.py:57-66)
 *            (d = 1 and W unused for nsite = 0; Dl = Dr = D)
 * L is (Ml, Dl, Ml), R is (Mr, Dr, Mr).  In RENO_PREC_FP64 L/R are complex128
 * and W float64; in RENO_PREC_MIXED L/R are complex64 and W float32.  Vectors
 * (C in / out) are complex128 in both modes. */
typedef struct {
  int nsite;
  int prec;
  int Ml, d, Mr, Dl, Dr;
  const void* L;
  const void* W;
  const void* R;
} reno_hop_desc;

/* One quantum-number block of svd_qn: rows `lset` x cols `rset` of the
 * coefficient matrix (device int64 index arrays). */
typedef struct {
  int nrows;
  int ncols;
  const int64_t* rows; /* device pointer, nrows entries */
  const int64_t* cols; /* device pointer, ncols entries */
} reno_qn_block;

/* ---- library ---------------------------------------------------------- */
int reno_version(void);
const char* reno_status_string(int status);
/* last HIP / rocSOLVER error text recorded by this thread */
const char* reno_last_error(void);

/* ---- workspace sizes (bytes) ---------------------------------------- */
size_t reno_hop_workspace_bytes(const reno_hop_desc* op);
/* workspace for expm_krylov with room for `kcap` Krylov vectors */
size_t reno_krylov_workspace_bytes(const reno_hop_desc* op, int kcap);
size_t reno_env_workspace_bytes(int prec, int Ml, int d, int Mr, int Dl, int Dr);
size_t reno_qr_workspace_bytes(int rows, int cols, int nblocks, const reno_qn_block* blocks);

/* ---- hop_expr: one matvec (replaces hop_expr.py:57-91 expr(C)) ------ */
int reno_hop_apply(const reno_hop_desc* op, const void* c_in, void* c_out,
                   void* ws, size_t ws_bytes, void* stream);

/* ---- hop_apply without allocation --------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  Computes
 * exactly what reno_hop_apply computes, with the same kernels, but its
 * scratch lives in a static device array of 96 MiB (one per device, held
 * until the process ends) instead of a caller workspace; it calls no
 * allocation routine.  Operators whose reno_hop_workspace_bytes exceeds
 * 96 MiB are rejected with RENO_EINVAL (the FP64 one-site matvec at
 * M = 256, d = 10, D = 3 needs 90 MiB).  All calls on one device share the
 * array: calls that overlap on a device race and must be serialised. */
int reno_hop_apply_noalloc(const reno_hop_desc* op, const void* c_in, void* c_out,
                           void* stream);

/* ---- tile autotuner ------------------------------------------------------
 * No code counterpart in the reference (its contractions go through
 * opt_einsum / cuBLAS, oe_contract_wrap.py:24-58, which pick their own
 * kernels); the proposed reno_autotune of the design.  Runs the matvec of
 * `op` once (c_in -> c_out, workspace as for reno_hop_apply), then times
 * every tile / split-K candidate of each of its GEMM stages on `stream`
 * with HIP events (`reps` launches per candidate) and keeps the fastest in
 * a process-wide table, guarded by a mutex, that every later launch of the
 * same GEMM signature (precision, shape, storage kinds) consults, from any
 * entry point.  Stages already in the table are not re-timed.  Writes one
 * record per newly tuned stage into out[0..max_out) and their number into
 * *n_out.  Synchronises the stream.  Tuned splits change the summation
 * order of the split-K reduction (results differ at rounding level from
 * the untuned planner's). */
typedef struct {
  int M, N, K, batch, prec;
  int tile, nsplit;                 /* chosen: tile edge 64 or 32, K chunks */
  int planned_tile, planned_nsplit; /* the built-in planner's choice */
  float us_tuned, us_planned;       /* measured microseconds per launch */
} reno_tile_cfg;

int reno_autotune(const reno_hop_desc* op, const void* c_in, void* c_out, void* ws,
                  size_t ws_bytes, int reps, reno_tile_cfg* out, int max_out, int* n_out,
                  void* stream);
/* forget every tuned choice (the planner decides again) */
int reno_autotune_reset(void);
/* number of GEMM signatures in the tuned table */
int reno_autotune_entries(void);

/* ---- contract_one_site (replaces lib.py:169-250) --------------------
 * domain 0 = "L": out[f,g,h] = sum env[a,b,c] conj(ms[a,d,f]) W[b,d,e,g] ms[c,e,h]
 *   env_in (Ml, Dl, Ml) -> env_out (Mr, Dr, Mr)
 * domain 1 = "R": out[f,g,h] = sum conj(ms[f,d,a]) env[a,b,c] W[g,d,e,b] ms[h,e,c]
 *   env_in (Mr, Dr, Mr) -> env_out (Ml, Dl, Ml)
 * ms is complex128 in both modes; env/W follow `prec` as in reno_hop_desc.
 * ms_conj may be NULL (then conj(ms) is used, lib.py:184-185); if given it is
 * a complex128 tensor of ms's shape and conj(ms_conj) is NOT taken: the
 * reference passes the already-conjugated bra tensor. */
int reno_env_update(int domain, int prec, int Ml, int d, int Mr, int Dl, int Dr,
                    const void* env_in, const void* ms, const void* ms_conj,
                    const void* W, void* env_out, void* ws, size_t ws_bytes,
                    void* stream);

/* ---- closing contraction of two environments (the last step of
 * Mps.expectation, mps.py:516-520, and of calc_1site_rdm, mps.py:1587-1596):
 *   out[g] = sum_{f,h} X[f,g,h] * Y[f,h]
 * X (F, G, H) and Y (F, H) complex128 (prec 0) or complex64 (prec 1);
 * out (G) complex128. */
int reno_env_close(int prec, int F, int G, int H, const void* X, const void* Y,
                   void* out, void* stream);

/* ---- expm_krylov (replaces lib/krylov/krylov.py:27-84) --------------
 * out = expm(dt * H) v0 by Lanczos without re-orthogonalisation, with the
 * reference's stopping rule: breakdown beta_j < 100 n eps, full-space exit,
 * and at every even j > 3 allclose(res, new_res, rtol 1e-5, atol 1e-8).
 * dt = dt_re + i dt_im.  v0/out are complex128 of n = Ml*d*Mr elements and
 * may alias.  *n_matvec receives the number of matvecs (reference j+1).
 * Returns RENO_ENOSPACE with *needed = bytes if ws holds too few Krylov
 * vectors; nothing is written to out then and the call may be repeated. */
int reno_expm_krylov(const reno_hop_desc* op, double dt_re, double dt_im,
                     const void* v0, void* out, int* n_matvec, size_t* needed,
                     void* ws, size_t ws_bytes, void* stream);

/* ---- expm_krylov without host waits (same computation as
 * reno_expm_krylov; the stopping rule is evaluated on the device) --------
 * Queues at most max_matvec Lanczos iterations (the workspace must hold that
 * many Krylov vectors: reno_krylov_workspace_bytes(op, max_matvec)) and the
 * final combination into out, then returns without waiting.  info is a
 * device int[2]: info[0] = matvecs used, info[1] = 1 when the reference's
 * rule (allclose / breakdown / full space) ended the iteration, 0 when
 * max_matvec iterations ran without it (out is then NOT the result: repeat
 * with reno_expm_krylov), -1 for a zero starting vector. */
int reno_expm_krylov_async(const reno_hop_desc* op, double dt_re, double dt_im,
                           const void* v0, void* out, int max_matvec, int* info,
                           void* ws, size_t ws_bytes, void* stream);

/* ---- expm_krylov with adaptive precision (the mixed mode) --------------
 * The same computation as reno_expm_krylov_async (reference krylov.py:27-84,
 * stopping rule included), but Lanczos iteration j applies op (FP64:
 * complex128 L/R, float64 W) for j < j_switch and op_lo (the same operator
 * held in RENO_PREC_MIXED: complex64 L/R, float32 W, FP32 matrix cores) for
 * j >= j_switch.  Vectors, recurrence, reductions, the tridiagonal
 * exponential and the stopping rule stay FP64.  A matvec whose relative
 * error is delta perturbs the result by about |c_{j-1}| delta, c_k the
 * result's coefficient of the k-th Lanczos vector, so the low-precision
 * tail costs nothing in accuracy once those coefficients are small.
 * info is a device int[3]: info[0..1] as reno_expm_krylov_async, info[2] =
 * the switch point measured on this call's result with threshold tau (the
 * first j >= 2 with |c_k| < tau |v0| for every k >= j - 1; 1 << 30 if the
 * last coefficient is not below tau), which the caller hands to the same
 * local step of the next time step.  j_switch >= 1 << 30 gives the FP64
 * path exactly.  Workspace: reno_krylov_workspace_bytes(op, max_matvec).
 * No reference counterpart: the reference's only low-precision knob is the
 * global complex64 switch (utils/backend.py:111-114). */
int reno_expm_krylov_mixed_async(const reno_hop_desc* op, const reno_hop_desc* op_lo,
                                 int j_switch, double tau, double dt_re, double dt_im,
                                 const void* v0, void* out, int max_matvec, int* info,
                                 void* ws, size_t ws_bytes, void* stream);

/* ---- batched sweep: many small states, one launch per sweep position ------
 * One state's site update of the one-site TDVP-PS sweep (reference
 * mps/mps.py:1296-1397, FP64): expm(-i dt/2 H_eff) of the site tensor v0 by
 * Lanczos with the reference's stopping rule (krylov.py:27-84) into kout; for
 * a site that is not the last of the half-sweep, the quantum-number blocked
 * QR (dir 1, left to right: (Ml d) x Mr, u -> site_out (Ml, d, kt), v^T ->
 * bond (kt x Mr)) or RQ (dir 0, right to left: Ml x (d Mr), u -> bond
 * (Ml x kt), v^T -> site_out (kt, d, Mr)) of kout as svd_qn(QR=True)
 * computes it (svd_qn.py:99-240; blocks in the caller's order), the
 * environment update of site_out (contract_one_site, lib.py:169-250: env_in
 * L[i-1] -> env_out L[i], or R[i+1] -> R[i]), expm(+i dt/2 H_bond) of the
 * bond matrix with the environments L0 / R0 (MPO bond D0) into bond_out, and
 * the bond matrix absorbed into the neighbouring site nbr (dir 0: carry =
 * nbr (nbr_a x Ml) bond_out; dir 1: carry = bond_out nbr (Mr x nbr_b)).
 * All pointers are device pointers; Wl is the site MPO permuted to
 * Wl[(dd, f), (b, e)] = W[b, dd, e, f].  scratch holds
 * reno_batch_scratch_bytes(Ml, d, Mr, Dl, Dr, kcap) bytes; counts receives
 * the matvecs of the two Krylov calls (-1: more than kcap were needed, the
 * outputs are void; -2: zero vector). */
typedef struct {
  int active, dir, last;
  int Ml, d, Mr, Dl, Dr;
  int kt, nblocks;
  int nbr_a, nbr_b;
  int D0, kcap;
  int pad0, pad1;
  const void* L;
  const void* R;
  const void* Wl;
  const void* W;
  const void* v0;
  void* kout;
  void* site_out;
  void* bond;
  const void* env_in;
  void* env_out;
  const void* L0;
  const void* R0;
  void* bond_out;
  const void* nbr;
  void* carry;
  const int64_t* brows[4];
  const int64_t* bcols[4];
  int bnr[4];
  int bnc[4];
  void* scratch;
  int* counts;
} reno_batch_site;

size_t reno_batch_scratch_bytes(int Ml, int d, int Mr, int Dl, int Dr, int kcap);
/* No reference counterpart (the reference evolves one state at a time):
 * the site updates of nstates states at one sweep position, one workgroup
 * per state, in a single launch; sites is a device array of descriptors
 * (inactive entries do nothing).  Replaces, per state, the hop_expr /
 * expm_krylov / svd_qn / contract_one_site sequence of mps.py:1296-1397. */
int reno_batch_site_step(int nstates, const reno_batch_site* sites, double dt, void* stream);
/* Scale each state's centre site (device pointers, len[b] complex128
 * entries) to unit Frobenius norm, the norm into norms[b] (device): the
 * normalize("mps_only") of a canonical state (mps.py:2025-2059). */
int reno_batch_normalize(int nstates, void* const* centre, const int64_t* len, double* norms,
                         void* stream);

/* ---- svd_qn QR / RQ branch (replaces svd_qn.py:99-240, QR=True,
 * full_matrices=False) ------------------------------------------------
 * coef is the complex128 (rows x cols) coefficient matrix.  For each block
 * (in the caller's order == reference `set` iteration order, svd_qn.py:177):
 * system 0 = "L": block = Q R   (scipy.linalg.qr, economic)
 * system 1 = "R": block = R Q   (scipy.linalg.rq, economic)
 * u (rows x kt) and v (cols x kt), kt = sum_b min(nrows_b, ncols_b), are
 * zero-filled and receive each block's factors in its row/col index set at
 * column offset sum of previous dims:
 *   "L": u[rows_b, off:off+k_b] = Q_b,  v[cols_b, off:off+k_b] = R_b^T
 *   "R": u[rows_b, off:off+k_b] = R_b,  v[cols_b, off:off+k_b] = Q_b^T
 * exactly as blockappend/blockrecover (svd_qn.py:66-96) lay them out.
 * With v_transposed != 0, v receives v^T (kt x cols) instead, the layout the
 * sweep stores as the next site / bond matrix (mps.py:1328-1331). */
int reno_block_qr(int system, int rows, int cols, const void* coef,
                  int nblocks, const reno_qn_block* blocks, void* u, void* v,
                  int v_transposed, void* ws, size_t ws_bytes, void* stream);

/* ---- svd_qn SVD branch (replaces svd_qn.py:99-240 with QR=False,
 * full_matrices=True, the decomposition of mp.py:693 / _update_mps) ----
 * Per block (caller's order == reference `set` iteration order), B_b =
 * U_b S_b V_b^H with complete bases (rocSOLVER zgesvd).  u is (rows x
 * sum_b nrows_b) and v (cols x sum_b ncols_b), complex128 row-major,
 * zero-filled: first every block's min(nrows_b, ncols_b) singular columns,
 * block by block, then every block's completion columns, block by block
 * (blockappend of the full-matrices branch); v holds V_b without
 * conjugation (the reference's v = vt.T).  s (device, float64) receives the
 * singular values of every block in the same order. */
size_t reno_svd_workspace_bytes(int rows, int cols, int nblocks, const reno_qn_block* blocks);
int reno_block_svd(int rows, int cols, const void* coef, int nblocks, const reno_qn_block* blocks,
                   void* u, void* v, void* s, void* ws, size_t ws_bytes, void* stream);

/* ---- eigh_qn (replaces svd_qn.py:243-303, the reduced-density-matrix
 * decomposition of the state-averaged update, mp.py:769-826) -----------
 * dm is the complex128 (n x n) density matrix.  For each block (square:
 * rows == cols, a quantum-number sector, caller's order) the Hermitian part
 * (B + B^H) / 2 is diagonalised (rocSOLVER zheevdj; a block whose result is
 * not finite is solved again with zheev, so the call waits for the stream
 * once per block): u (n x sum_b nrows_b,
 * row-major, zero-filled) receives the eigenvectors of every block at its
 * row set and column offset, w (device, float64) the eigenvalues in
 * ascending order within each block. */
size_t reno_eigh_workspace_bytes(int n, int nblocks, const reno_qn_block* blocks);
int reno_block_eigh(int n, const void* dm, int nblocks, const reno_qn_block* blocks, void* u, void* w,
                    void* ws, size_t ws_bytes, void* stream);

/* ---- DMRG local eigenproblem (FP64: L/R complex128, W float64) --------
 * Diagonal of the one-site H_eff, the Davidson preconditioner (replaces the
 * hdiag contraction of gs.py:424-436):
 *   out[a,d,l] = sum_{b,f} L[a,b,a] W[b,d,d,f] R[l,f,l]      (complex128)
 * Two-site problems pass the two MPO sites merged into one (d = d0*d1). */
int reno_heff_diag(int Ml, int d, int Mr, int Dl, int Dr, const void* L,
                   const void* W, const void* R, void* out, void* stream);
/* Dense H_eff on the allowed quantum-number entries, for the direct
 * diagonalisation of small sites (replaces get_ham_direct, gs.py:302-330):
 *   out[i][j] = sum_{b,f} L[a,b,c] W[b,d,e,f] R[l,f,k],
 *   (a,d,l) = idx[i], (c,e,k) = idx[j]  (flat row-major site indices;
 *   idx is a device int64 array of nidx entries; out nidx x nidx complex128) */
int reno_heff_dense(int Ml, int d, int Mr, int Dl, int Dr, const void* L,
                    const void* W, const void* R, const int64_t* idx, int nidx,
                    void* out, void* stream);

/* ---- Davidson vector algebra (replaces the numpy vector work of the
 * reference's Davidson-Liu solver, lib/davidson/davidson.py:154-441, run on
 * the DMRG local problem at gs.py:533-538).  Basis vectors are rows of
 * length n, row i at V + i * ldv (ldv >= n), complex128; at most 64 rows.
 * reno_dav_orth: x <- x - V^T (conj(V) x) twice against rows 0..k-1 (the
 * projections of davidson.py:399-402 and _qr, 453-477, with one more
 * pass), nrm2[0] (device double) = |x|^2, and x <- x / |x| when
 * nrm2[0] >= thresh (the lindep test of davidson.py:405-411).
 * reno_dav_dots: out[i] = sum_r conj(V[i][r]) y[r], i < k (device
 * complex128): a column of the subspace matrix (_fill_heff_hermitian,
 * davidson.py:54-70).
 * reno_dav_ritz: with U (m x k, row-major complex128) the lowest k
 * eigenvectors of the m-dimensional subspace matrix and e (device float64)
 * their values, X[j] = sum_i U[i][j] V[i], AX[j] = sum_i U[i][j] AV[i],
 * R[j] = AX[j] - e[j] X[j] (rows of length n, contiguous) and rnorm2[j] =
 * |R[j]|^2 (davidson.py:331-345, _gen_x0 479-487); m <= 64, m * k <= 1024.
 * reno_dav_precond: out = r / (hdiag - e + 1e-4) where mask (one byte per
 * entry) is set, 0 elsewhere (gs.py:534 on the allowed entries).
 * Sums run as fixed-order per-block partials and one fold: the results are
 * reproducible bit for bit. */
size_t reno_dav_workspace_bytes(int64_t n, int k);
int reno_dav_orth(int64_t n, int k, const void* V, int64_t ldv, void* x, double thresh, void* nrm2,
                  void* ws, size_t ws_bytes, void* stream);
int reno_dav_dots(int64_t n, int k, const void* V, int64_t ldv, const void* y, void* out, void* ws,
                  size_t ws_bytes, void* stream);
int reno_dav_ritz(int64_t n, int m, int k, const void* V, const void* AV, int64_t ldv, const void* U,
                  const void* e, void* X, void* AX, void* R, void* rnorm2, void* ws, size_t ws_bytes,
                  void* stream);
int reno_dav_precond(int64_t n, const void* r, const void* hdiag, const void* mask, double e, void* out,
                     void* stream);

/* ---- tensordot (replaces matrix.py:210-211 for the 2-D contractions of
 * the sweep: absorbing the evolved bond into the neighbouring site) ------
 * C (M x N) = A (M x K) . B (K x N), all row-major complex128. */
int reno_zgemm(int M, int N, int K, const void* A, const void* B, void* C,
               void* stream);

/* ---- launch timing of the contraction kernel (no reference counterpart;
 * the benchmark's roofline figure).  reno_prof_enable(period), period >= 1,
 * clears the record and starts it: every subsequent GEMM launch is recorded
 * (kernel instantiation, algorithmic FLOP, and a device-side marker that
 * tells a launch that ran from one cancelled on the device), and a
 * pseudo-random 1-in-period sample of them is timed by two HIP events bound
 * to the dispatch (hipExtLaunchKernel: kernel execution only, as a profiler
 * reports it; period 1: every launch; timing every launch slows the stream,
 * a sample keeps the measured run close to an unmeasured one).  reno_prof_enable(0) stops recording;
 * reno_prof_enable(-period) resumes it without clearing (work between a stop
 * and a resume is not recorded).
 * reno_prof_collect waits for the events and fills, per kernel
 * instantiation k (see reno_prof_kind_info): over the timed sample, the
 * count of executed launches, their summed GPU milliseconds and summed
 * algorithmic FLOP; over all recorded launches, the count and summed FLOP
 * of the executed ones and the count of all, including those cancelled on
 * the device (Lanczos iterations queued past convergence, which return at
 * once).  Any output pointer may be NULL. */
int reno_prof_enable(int period);
int reno_prof_collect(int nkinds, double* ms, double* flop_timed, int* timed, double* flop,
                      int* executed, int* queued);
/* The event-timed launches of the sample that were cancelled on the device
 * (they return at once), per instantiation: count and summed event
 * milliseconds -- the part of a measured launch that is the event bracket's
 * own (queue latency), not the kernel. */
int reno_prof_cancelled(int nkinds, double* ms, int* count);
/* Number of GEMM kernel instantiations; for 0 <= kind < that number, fills
 * info[0..6] = precision, A k-contiguous, B k-contiguous, the storage kinds
 * of A, B, C (0 complex128, 1 complex64, 2 float64, 3 float32) and the
 * square output tile (64, or 32 for products with few 64 x 64 tiles). */
int reno_prof_kind_info(int kind, int* info);

/* ---- test hook: one v_mfma 16x16x4 tile, used by the lane-layout check.
 * a (16x4), b (4x16), c (16x16) row-major; prec selects f64 or f32. */
int reno_probe_mfma16(int prec, const void* a, const void* b, void* c,
                      void* stream);

#ifdef __cplusplus
}
#endif
#endif /* RENO_HIP_H */
