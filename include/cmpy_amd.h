/*
 * cmpy_amd.h -- C ABI of the MI355X exact-diagonalisation engine (libcmpy_amd.so).
 *
 * This is the boundary the Python surface (cmpy_amd) binds with ctypes, and the
 * one any other host language would bind.  Types are plain C: integers, doubles,
 * device pointers and an opaque operator handle.  No torch type appears here.
 *
 * Conventions
 *   - Every function returns an int status: ED_OK (0) or a negative ED_ERR_*.
 *     ed_last_error() returns a thread-local message for the last failure.
 *     No C++ exception crosses this boundary.
 *   - Pointers named *_dev are device (HBM) pointers; pointers without the
 *     suffix are host pointers.  The caller owns every buffer it passes.  The
 *     library owns what sits behind an ed_op handle (state lists, hop tables,
 *     diagonal tables) or an ed_batch handle, freed by ed_op_destroy /
 *     ed_batch_destroy; a few calls (continued fraction, time sum, dense
 *     eigensolvers, partially reorthogonalised Lanczos coefficients) allocate
 *     scratch and free it before returning; pair sums and occupations take
 *     theirs from the stream-ordered allocator (hipMallocAsync/hipFreeAsync on
 *     `stream`) and return without waiting for the device.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *     Device work is enqueued on it; functions that return host results
 *     (ed_lanczos_*, ed_batch_*) synchronise that stream before returning.
 *   - Sector vectors are FP64, index k = iu * num_dn + id with iu/id the
 *     positions of the spin-up/spin-down states in ascending integer order
 *     (cmpy/basis.py:570-575, 655-666).  A block of nvec vectors is stored
 *     vector after vector ([nvec][D]).
 *   - Complex numbers are interleaved (re, im) doubles.
 *   - Distinct handles may be used from distinct threads; one handle must not
 *     be used by two threads at once.
 */
#ifndef CMPY_AMD_H
#define CMPY_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ED_OK 0
#define ED_ERR_INVALID (-1)  /* bad argument (shape, range, bond order, ...) */
#define ED_ERR_HIP (-2)      /* a HIP runtime call failed                     */
#define ED_ERR_NOMEM (-3)    /* device allocation failed                      */
#define ED_ERR_NUMERIC (-4)  /* NaN/Inf met, or a solve did not converge       */
#define ED_ERR_LIMIT (-5)    /* size beyond what the library supports         */

#define ED_UP 1 /* cmpy/basis.py:40 */
#define ED_DN 2

#define ED_SIGN_REFERENCE 0     /* no Jordan-Wigner string (cmpy/operators.py:652-703) */
#define ED_SIGN_JORDAN_WIGNER 1 /* within-spin string (-1)^popcount(s & (2^pos - 1))  */

#define ED_MAX_SITES 30

typedef struct ed_op ed_op;

/* ------------------------------------------------------------------------ */
/* Library                                                                   */
/* ------------------------------------------------------------------------ */

const char* ed_last_error(void);
int ed_version(void);
/* Number of visible HIP devices (0 without a GPU); never fails. */
int ed_device_count(void);

/* ------------------------------------------------------------------------ */
/* Basis -- replaces Basis.generate_states / get_states (cmpy/basis.py:655-676)
 * and the bisect_left index lookup (cmpy/operators.py:276-299).             */
/* ------------------------------------------------------------------------ */

/* C(n, k) for 0 <= n <= 64; 0 outside the triangle. */
int64_t ed_binom(int n, int k);

/* Host: all num_sites-bit states with popcount n, ascending (Gosper's
 * next-combination).  `out` holds ed_binom(num_sites, n) entries.           */
int ed_sector_states_host(int num_sites, int n, uint32_t* out);

/* Device: same list, state r computed by combinadic unranking of r.        */
int ed_sector_states(int num_sites, int n, uint32_t* out_dev, void* stream);

/* Device: combinadic rank r(s) = sum_k C(p_k, k) of each state (its position
 * in the ascending list of its popcount); -1 where popcount(s) != n.        */
int ed_sector_rank(int num_sites, int n, const uint32_t* states_dev, int64_t count,
                   int64_t* ranks_dev, void* stream);

/* ------------------------------------------------------------------------ */
/* Projectors -- replaces _hopping_sign / _compute_hopping_term
 * (cmpy/operators.py:425-460).  One entry per state of the (num_sites, n)
 * spin sector, in ascending state order: target = index of the hopped state
 * (or -1 when the bond cannot hop from this state), value = sign * hop.
 * `width` is the bit width of the sign string: num_sites for the Hubbard
 * model (cmpy/models/hubbard.py:22), 0 for the Anderson impurity model
 * (cmpy/models/anderson.py:149, all signs +1).                              */
/* ------------------------------------------------------------------------ */
int ed_hop_terms(int num_sites, int n, int width, int site1, int site2, double hop,
                 int64_t* target_dev, double* value_dev, void* stream);

/* ------------------------------------------------------------------------ */
/* Sector Hamiltonian -- replaces AbstractManyBodyModel.hamilton_operator +
 * HamiltonOperator (cmpy/models/abc.py:242-256, cmpy/operators.py:614-646)
 * with a matrix-free operator over the up (x) dn product structure:
 *   H = sum_i eps_i (n_i,up + n_i,dn) + sum_i u_i n_i,up n_i,dn
 *     + sum_b hop_b sum_sigma (c+_{j,sigma} c_{i,sigma} + h.c.) * sign       */
/* ------------------------------------------------------------------------ */

typedef struct ed_model_desc {
  int32_t num_sites;     /* <= ED_MAX_SITES                                  */
  int32_t n_up, n_dn;    /* sector fillings                                  */
  int32_t num_bonds;
  const int32_t* bonds;  /* host [num_bonds][2], bonds[b][0] < bonds[b][1]   */
  const double* hop;     /* host [num_bonds]                                 */
  const double* eps;     /* host [num_sites], on-site energy minus mu        */
  const double* u;       /* host [num_sites], on-site interaction            */
  int32_t width;         /* hopping-sign width (see ed_hop_terms)            */
} ed_model_desc;

int ed_op_create(ed_op** op, const ed_model_desc* desc, void* stream);
int ed_op_destroy(ed_op* op);
int ed_op_shape(const ed_op* op, int64_t* num_up, int64_t* num_dn);

/* y = H x for nvec vectors (cmpy/operators.py:626-630).  x and y must not
 * overlap.                                                                  */
int ed_op_matvec(ed_op* op, const double* x_dev, double* y_dev, int64_t nvec, void* stream);

/* The same H.v on vectors stored row by row with a row stride of ld doubles
 * (num_dn <= ld <= num_dn + 4096; [nvec][num_up][ld]): entry (iu, id) of a
 * vector sits at iu * ld + id.  x's entries past num_dn in a row are never
 * read; y's are written as zeros.  No reference counterpart: the reference's
 * vectors are contiguous (cmpy/operators.py:626-630, ld = num_dn here).
 * Padding rows to whole 128-byte lines keeps the column pass's slabs of x
 * line-aligned (see ed_op_vector_ld).                                       */
int ed_op_matvec_ld(ed_op* op, const double* x_dev, double* y_dev, int64_t ld, int64_t nvec, void* stream);

/* Row stride the library's own Lanczos vectors use for this operator:
 * num_dn rounded up to 16 doubles where the two-pass sweep runs, else num_dn
 * (environment CMPY_AMD_HV_PITCH=0: always num_dn).                         */
int ed_op_vector_ld(const ed_op* op, int64_t* ld);

/* y = H x for H given as triples sorted by target row (CSR by target):
 * y[t] = sum_{k=ptr[t]}^{ptr[t+1]-1} val[k] * x[src[k]], in stored order.
 * The COO HamiltonOperator path (cmpy/operators.py:614-630).              */
int ed_csr_matvec(const int64_t* ptr_dev, const int64_t* src_dev, const double* val_dev, int64_t n,
                  const double* x_dev, double* y_dev, void* stream);

/* Diagonal of H (onsite + interaction, cmpy/operators.py:305-422).          */
int ed_op_diagonal(ed_op* op, double* diag_dev, void* stream);

/* Dense H, row-major D x D (HamiltonOperator.toarray, operators.py:632-635). */
int ed_op_dense(ed_op* op, double* h_dev, void* stream);

/* Per-spin hop table of the operator in its ELL form, for inspection:
 * cnt[D_sigma], col/val[D_sigma][max_hops] (row-major, valid entries first,
 * in bond order).  max_hops is returned through *max_hops (call with NULL
 * buffers to query it).                                                     */
int ed_op_hop_table(const ed_op* op, int sigma, int32_t* max_hops, int32_t* cnt_dev,
                    int32_t* col_dev, double* val_dev, void* stream);

/* ------------------------------------------------------------------------ */
/* Creation / annihilation -- replaces _apply_{creation,annihilation}_{up,dn},
 * CreationOperator, AnnihilationOperator (cmpy/operators.py:652-791).
 * x lives in sector (n_up, n_dn); y in the sector one particle above (cdag)
 * or below (c).  y is fully written (zeros where nothing maps).            */
/* ------------------------------------------------------------------------ */
int ed_apply_cdag(int num_sites, int n_up, int n_dn, int pos, int sigma, int sign_mode,
                  const double* x_dev, double* y_dev, int64_t nvec, void* stream);
int ed_apply_c(int num_sites, int n_up, int n_dn, int pos, int sigma, int sign_mode,
               const double* x_dev, double* y_dev, int64_t nvec, void* stream);

/* ------------------------------------------------------------------------ */
/* Lanczos -- replaces compute_groundstate's eigsh call for one sector
 * (cmpy/exactdiag.py:24-43) and iter_lanczos_coeffs / lanczos_ground_state
 * (cmpy/exactdiag.py:324-375).  FP64 throughout, deterministic reductions.  */
/* ------------------------------------------------------------------------ */

typedef struct ed_lanczos_cfg {
  int32_t max_iter;     /* maximum Lanczos steps                              */
  int32_t min_iter;     /* no convergence test before this many steps         */
  int32_t check_every;  /* steps between convergence tests                    */
  int32_t reorth;       /* 0: three-vector recurrence (two passes for the vector);
                           1: full reorthogonalisation; 2: partial (Simon's
                           estimate, Gram-Schmidt only where it passes sqrt(eps));
                           1 and 2 keep the basis, so the vector takes one pass */
  int32_t want_vector;  /* 1: also return the ground-state vector             */
  int32_t fixed_iter;   /* 1: run exactly max_iter steps (no early stop)      */
  double tol;           /* stop when the Ritz residual |H psi - E psi| <= tol  */
  uint64_t seed;        /* start vector seed (used when v0_dev is NULL)       */
  int32_t time_matvec;  /* 1: time every H.v sweep with HIP events into
                           matvec_ms (each pair of events costs a few us of
                           GPU time per step); 0: matvec_ms = 0               */
  int32_t pad;
} ed_lanczos_cfg;

typedef struct ed_lanczos_result {
  double e0;             /* lowest Ritz value                                  */
  double ritz_residual;  /* beta_{m+1} |s_m| of the lowest Ritz pair           */
  double residual;       /* explicit |H psi - e0 psi| (want_vector only)        */
  int32_t iterations;    /* Lanczos steps taken                                */
  int32_t converged;
  int32_t breakdown;     /* invariant subspace reached                         */
  int32_t pad;
  double matvec_ms;      /* summed H.v kernel time (HIP events; time_matvec)   */
  double total_ms;       /* wall time of the call                              */
} ed_lanczos_result;

/* Workspace (bytes) that ed_lanczos_groundstate needs for this op/config.   */
int ed_lanczos_workspace(const ed_op* op, const ed_lanczos_cfg* cfg, int64_t* bytes);

int ed_lanczos_groundstate(ed_op* op, const ed_lanczos_cfg* cfg, const double* v0_dev,
                           double* psi_dev, void* work_dev, int64_t work_bytes,
                           ed_lanczos_result* res, void* stream);

/* Lanczos coefficients from a given start vector (the continued-fraction
 * input): alpha[0..m-1], beta[0..m-2] (beta[k] couples steps k and k+1),
 * *norm2 = |v0|^2, *steps = m <= nsteps (fewer on breakdown).  reorth = 1
 * orthogonalises every new vector against the whole basis (two Gram-Schmidt
 * passes; nsteps <= 4096), which keeps real-axis continued fractions exact up
 * to the Krylov dimension; reorth = 2 does so only at the steps where Simon's
 * orthogonality estimate passes sqrt(eps) (partial reorthogonalisation: the
 * basis stays semi-orthogonal, the coefficients as accurate as their own
 * conditioning allows; a small scratch buffer is allocated for the estimate).
 * Optionally the normalised basis vectors are written to basis_dev
 * ([nsteps][D]).                                                            */
int ed_lanczos_coeffs_workspace(const ed_op* op, int32_t nsteps, int32_t reorth, int64_t* bytes);
int ed_lanczos_coeffs(ed_op* op, const double* v0_dev, int32_t nsteps, int32_t reorth, double* alpha,
                      double* beta, double* norm2, int32_t* steps, double* basis_dev,
                      void* work_dev, int64_t work_bytes, void* stream);

/* ------------------------------------------------------------------------ */
/* Green's functions                                                         */
/* ------------------------------------------------------------------------ */

/* Continued fraction (no reference counterpart; cmpy has no CF):
 *   g[k] (+)= sign * norm2 / (sign z_k + shift - a_0 - b_0^2/(sign z_k + shift - a_1 - ...))
 * sign=+1, shift=E0: particle part from sector N+1;
 * sign=-1, shift=E0: hole part from sector N-1.
 * alpha/beta are host arrays (m and m-1 entries).                           */
int ed_cf_eval(const double* alpha, const double* beta, int32_t m, double norm2, double shift,
               int32_t sign, const double* z_dev, int64_t nz, double* g_dev, int32_t accumulate,
               void* stream);

/* Real-time Green's function from Lanczos coefficients: replaces the
 * expm_multiply propagation and overlap of gf_greater / gf_lesser
 * (cmpy/exactdiag.py:248-301) by the Gauss quadrature of phi's spectral
 * measure, exact for polynomials of degree 2m - 1 in H:
 *   sign=+1: g[k] (+)= -i norm2 sum_j w_j exp(-i (theta_j - e0) t_k)
 *            (greater part, phi = c+ psi0 in sector N+1)
 *   sign=-1: g[k] (+)= +i norm2 sum_j w_j exp(+i (theta_j - e0) t_k)
 *            (lesser part, phi = c psi0 in sector N-1)
 * theta_j / w_j: eigenvalues / squared first eigenvector components of the
 * tridiagonal (alpha[m], beta[m-1], host arrays, m <= 8192).
 * t_dev: [nt] times; g_dev: [nt] complex interleaved.                      */
int ed_tevo_eval(const double* alpha, const double* beta, int32_t m, double norm2, double e0,
                 int32_t sign, const double* t_dev, int64_t nt, double* g_dev, int32_t accumulate,
                 void* stream);

/* T = 0 G(z) of c = c_{pos,sigma} in one call, from the ground state psi0
 * (device, contiguous) of sector (n_up, n_dn) of num_sites sites, energy e0:
 *   G(z) = <psi0| c (z + e0 - H)^-1 c+ |psi0> + <psi0| c+ (z - e0 + H)^-1 c |psi0>,
 * the T = 0 limit of the sum accumulated by accumulate_gf / _accumulate_sum
 * (cmpy/exactdiag.py:110-134).  op_p / op_m: operators of the sectors one
 * particle above / below in spin sigma, NULL where that sector does not
 * exist (its part is then 0).  z_dev, g_dev: [nz] complex interleaved; g is
 * overwritten.  sign_mode: ED_SIGN_* of the ladder operators.               */

/* Continued fractions: particle and hole Lanczos runs of nsteps each (capped
 * at the sector size; reorth as ed_lanczos_coeffs) started at c+ psi0 and
 * c psi0 -- the recurrence of iter_lanczos_coeffs (cmpy/exactdiag.py:324-348)
 * -- evaluated as ed_cf_eval does.  parts (optional) receives per part
 * (0 particle, 1 hole) the step count, |phi|^2 and, where the caller gives
 * host arrays of nsteps doubles, alpha_0.. and beta_1.. .                   */
typedef struct ed_cf_parts {
  double* alpha[2];   /* in: host [nsteps] or NULL                         */
  double* beta[2];    /* in: host [nsteps] or NULL (steps - 1 written)      */
  double norm2[2];    /* out: |c+ psi0|^2, |c psi0|^2                        */
  int32_t steps[2];   /* out: Lanczos steps (0: empty part)                 */
} ed_cf_parts;

int ed_greens_cf_workspace(const ed_op* op_p, const ed_op* op_m, int32_t nsteps, int32_t reorth, int64_t* bytes);
int ed_greens_cf(ed_op* op_p, ed_op* op_m, int32_t num_sites, int32_t n_up, int32_t n_dn, int32_t pos,
                 int32_t sigma, int32_t sign_mode, const double* psi_dev, double e0, int32_t nsteps, int32_t reorth,
                 const double* z_dev, int64_t nz, double* g_dev, ed_cf_parts* parts, void* work_dev,
                 int64_t work_bytes, void* stream);

/* Dense Lehmann sums: both neighbour sectors diagonalised on the device
 * (ed_op_dense + ed_dense_eigh, at most 65,536 states), overlaps <m|c+|psi0>
 * and <n|c|psi0>, then ed_pole_sum -- greens_lehmann_t0 of the Python layer,
 * the ground-state term of gf_lehmann (cmpy/exactdiag.py:215-245).          */
int ed_greens_lehmann_t0_workspace(const ed_op* op_p, const ed_op* op_m, int64_t* bytes);
int ed_greens_lehmann_t0(ed_op* op_p, ed_op* op_m, int32_t num_sites, int32_t n_up, int32_t n_dn, int32_t pos,
                         int32_t sigma, int32_t sign_mode, const double* psi_dev, double e0, const double* z_dev,
                         int64_t nz, double* g_dev, void* work_dev, int64_t work_bytes, void* stream);

/* Real time to complex frequency, the transform of fourier_t2z
 * (cmpy/exactdiag.py:311-316, which calls gftool.fourier.tt2z):
 *   out[j] = int_{t_0}^{t_{nt-1}} e^{i z_j t} G(t) dt,  G linear between the
 * samples (the piecewise-linear rule, each interval integrated exactly).
 * t_dev: [nt] increasing times; g_dev: [nt] and z_dev, out_dev: [nz] complex
 * interleaved.                                                              */
int ed_fourier_t2z(const double* t_dev, const double* g_dev, int64_t nt, const double* z_dev, int64_t nz,
                   double* out_dev, void* stream);

/* Pole sum g[k] (+)= sum_m w_m / (z_k - p_m): the T = 0 Lehmann sum of
 * _accumulate_sum (cmpy/exactdiag.py:110-129) restricted to the ground state. */
int ed_pole_sum(const double* w_dev, const double* p_dev, int64_t m, const double* z_dev,
                int64_t nz, double* g_dev, int32_t accumulate, void* stream);

/* Finite-temperature pair sum of _accumulate_sum (cmpy/exactdiag.py:110-129):
 *   g[k] += sum_{m,n} |ov[m,n]|^2 (e^{-beta(E_n-emin)} + e^{-beta(E_m-emin)})
 *                      / (z_k - E_m + E_n)
 * with ov = V_p1^T C (overlap_dev holds ov[m][n], row-major M x N).
 * beta = +inf gives unit weights, as the reference does.  At finite beta a
 * pair whose two Boltzmann factors are both below 1e-20 is skipped (for
 * normalised overlaps the skipped terms add up to at most
 * (N + M) 1e-20 / Im z); environment CMPY_AMD_PAIR_SCREEN=0 sums every pair. */
int ed_lehmann_pairs(const double* overlap_dev, const double* evals_dev, int64_t n,
                     const double* evals_p1_dev, int64_t m, double beta, double emin,
                     const double* z_dev, int64_t nz, double* g_dev, void* stream);

/* ed_lehmann_pairs without any allocation.  No counterpart in the reference:
 * this entry point is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Its scratch (per-group partials and the screening bounds) is a fixed
 * 4 MiB device array per GPU (524,288 doubles, held until the process ends),
 * so it needs 2 * ceil(n * m / 8192) * nz + 2 <= 524,288 and returns
 * ED_ERR_LIMIT beyond that.  Same arguments, kernels and results as
 * ed_lehmann_pairs.  Every call on one GPU uses the same array: calls that
 * overlap on one GPU (different streams or threads) race.  Each call writes
 * all the scratch it reads, so nothing has to be cleared between calls.   */
int ed_lehmann_pairs_noalloc(const double* overlap_dev, const double* evals_dev, int64_t n,
                             const double* evals_p1_dev, int64_t m, double beta, double emin,
                             const double* z_dev, int64_t nz, double* g_dev, void* stream);

/* Off-diagonal pair sum: as ed_lehmann_pairs with |ov[m,n]|^2 replaced by
 * ov_i[m,n] ov_j[m,n], where ov_x = V_p1^T (c+_x evecs), i.e.
 *   g[k] += sum_{m,n} <n|c_i|m><m|c+_j|n> (e^{-beta(E_n-emin)} + e^{-beta(E_m-emin)})
 *                     / (z_k - E_m + E_n)
 * for real eigenvectors.  This is the sum lehmann_full.py:19-44 sets out to
 * form (its code keeps only the c_i factor, squared).                        */
int ed_lehmann_pairs_ij(const double* overlap_i_dev, const double* overlap_j_dev, const double* evals_dev,
                        int64_t n, const double* evals_p1_dev, int64_t m, double beta, double emin,
                        const double* z_dev, int64_t nz, double* g_dev, void* stream);

/* C[M x N] = A B^T with A: [M][K], B: [N][K] (vectors stored one after
 * another): the overlap product evecs_p1^T (c+ evecs) of cmpy/exactdiag.py:112.
 * Small products run on an FP64 matrix-core kernel, large ones through
 * rocblas_dgemm (environment CMPY_AMD_GEMM = 0 VALU kernel, 1 matrix-core
 * kernel, 2 rocBLAS).  K = 0 gives C = 0 (a_dev, b_dev may then be NULL).   */
int ed_gemm_nt(const double* a_dev, const double* b_dev, double* c_dev, int64_t m, int64_t n,
               int64_t k, void* stream);

/* Thermal occupations of occupation_up/dn and double_occupation
 * (cmpy/exactdiag.py:60-107):  *out = sum_k w_k sum_{i in S} |V_k[i]|^2 with
 * V: [nvec][D] sector vectors of (num_sites, n_up, n_dn) and S the indices
 * whose up (kind 0), dn (kind 1) or both (kind 2) spins occupy `pos`.
 * w_dev: [nvec] weights; out_dev: one double.                               */
int ed_occupation(const double* vecs_dev, int64_t nvec, int num_sites, int n_up, int n_dn, int pos,
                  int kind, const double* w_dev, double* out_dev, void* stream);

/* The same sum over caller-given spin-state lists, the signature of
 * occupation_up / occupation_dn / double_occupation (cmpy/exactdiag.py:60-107,
 * which take up_states and dn_states arrays): up_states_dev [num_up] and
 * dn_states_dev [num_dn] (device, uint32, any order), V: [nvec][num_up*num_dn]
 * with the flat index up_idx*num_dn + dn_idx (the index the reference's
 * running `idx` stands for, exactdiag.py:95-107).  pos < 32.                 */
int ed_occupation_states(const double* vecs_dev, int64_t nvec, const uint32_t* up_states_dev, int64_t num_up,
                         const uint32_t* dn_states_dev, int64_t num_dn, int pos, int kind,
                         const double* w_dev, double* out_dev, void* stream);

/* Dense symmetric eigensolver (LAPACK dsyevd semantics, ascending values):
 * replaces np.linalg.eigh of solve_sector (cmpy/exactdiag.py:46-57).
 * a_dev (n x n symmetric) is overwritten by the eigenvectors stored vector
 * after vector: a[k*n + i] = component i of eigenvector k.                */
int ed_dense_eigh(double* a_dev, int64_t n, double* w_dev, void* stream);

/* ed_dense_eigh over `batch` matrices of one size n stored one after another
 * (a_dev: [batch][n][n], w_dev: [batch][n]): the sectors of equal dimension
 * that gf_lehmann diagonalises (cmpy/exactdiag.py:46-57, 215-245) in one
 * solver call.                                                             */
int ed_dense_eigh_batched(double* a_dev, int64_t n, int64_t batch, double* w_dev, void* stream);

/* ------------------------------------------------------------------------ */
/* Batched Anderson impurity solver.  No reference counterpart: cmpy solves one
 * SingleImpurityAndersonModel per call (cmpy/models/anderson.py:163-164).
 * One persistent launch (replayed from a captured HIP graph) solves every
 * model of the batch: ground state of its (n_up, n_dn) sector by Lanczos to
 * Ritz residual tol, then the impurity Green's function at Matsubara
 * frequencies, G(i w_n) = particle + hole continued fraction from
 * c+_{0,up} psi0 and c_{0,up} psi0 (Lanczos until G at the lowest frequency
 * has converged to cf_tol, at most cf_steps steps each).  Each model
 * is built as SingleImpurityAndersonModel does (anderson.py:147-158): site 0
 * the impurity with eps_imp - mu and U, bath sites 1..nb with eps_bath (not
 * shifted), bonds (0, k) with hopping v_k, hopping-sign width sign_width
 * (0 = the reference's all-positive signs).                                 */
/* ------------------------------------------------------------------------ */
#define ED_BATCH_MAX_BATH 15

typedef struct ed_batch_desc {
  int64_t num_models;
  const int32_t* num_bath;  /* host [num_models], 1..ED_BATCH_MAX_BATH          */
  const int32_t* n_up;      /* host [num_models]                                */
  const int32_t* n_dn;      /* host [num_models]                                */
  const double* u;          /* host [num_models]                                */
  const double* eps_imp;    /* host [num_models], impurity level minus mu       */
  const double* eps_bath;   /* host [num_models][max_bath] (first num_bath used) */
  const double* v;          /* host [num_models][max_bath]                      */
  int32_t max_bath;         /* row stride of eps_bath and v                     */
  int32_t sign_width;       /* 0: reference signs; >= num_sites: fermionic      */
  int32_t max_iter;         /* ground-state Lanczos steps cap (1..2000)         */
  int32_t cf_steps;         /* continued-fraction depth cap (1..2000; a fraction
                               that reaches it has not converged, and
                               iters[3 m + 1..2] then equals cf_steps)          */
  int32_t check_every;      /* steps between ground-state convergence tests     */
  int32_t keep_basis;       /* 1: keep psi and the Krylov bases for ed_batch_backward
                               (needs cf_steps <= 1024; kept bases take
                               8 cf_steps D(N+-1) bytes per model)              */
  double tol;               /* ground-state Ritz residual tolerance              */
  uint64_t seed;            /* Lanczos start-vector seed                        */
  double cf_tol;            /* continued-fraction stop: relative change of G at
                               the lowest frequency over two 8-step checks      */
} ed_batch_desc;

typedef struct ed_batch ed_batch;

int ed_batch_create(ed_batch** b, const ed_batch_desc* desc, void* stream);
int ed_batch_destroy(ed_batch* b);

/* Launch shape: models whose sector has at most wave_max_dim states run one
 * per wave, in workgroups [0, wave_blocks); the rest run one per workgroup.
 * Without this call the first ed_batch_solve autotunes the shape (it times
 * candidate splits and grid sizes on the batch itself).                    */
int ed_batch_config(ed_batch* b, int64_t wave_max_dim, int32_t grid_blocks, int32_t wave_blocks, void* stream);
int ed_batch_info(const ed_batch* b, int64_t* wave_max_dim, int32_t* grid_blocks, int32_t* wave_blocks,
                  double* tune_ms);

/* Solve every model.  wn: host [nw] Matsubara frequencies w_n (z = i w_n).
 * Host outputs (any may be NULL): e0[num_models]; g[num_models][nw] complex
 * interleaved; iters[num_models][3] Lanczos steps (ground state, particle,
 * hole continued fraction);
 * resid[num_models][2] = (Ritz residual, explicit |H psi - e0 psi|);
 * *elapsed_ms = device time of the launch.                                 */
int ed_batch_solve(ed_batch* b, const double* wn, int32_t nw, double* e0, double* g, int32_t* iters,
                   double* resid, double* elapsed_ms, void* stream);

/* Replace every model's bath (host [num_models][max_bath] eps_bath and v, the
 * first num_bath entries used); the launch shape and the captured graph are
 * kept, so a fit loop solves again without re-tuning.                       */
int ed_batch_set_bath(ed_batch* b, const double* eps_bath, const double* v, void* stream);

/* Adjoint backward of the last ed_batch_solve (needs keep_basis).  No
 * reference counterpart: it turns the bath fit of cmpy/dmft into a gradient
 * method.  Input: dL/dRe G and dL/dIm G, host [num_models][nw] each, for a
 * real loss L.  Output (host, [num_models][max_bath]): dL/d eps_bath and
 * dL/d v; cg_iters[num_models]: iterations of the projected ground-state
 * response solve.  Uses Hellmann-Feynman for dE0, the kept Krylov bases for
 * the resolvents and conjugate gradients on (H - E0) over the complement of
 * the ground state for the state response.                                 */
int ed_batch_backward(ed_batch* b, const double* dl_dre, const double* dl_dim, double* grad_eps, double* grad_v,
                      int32_t* cg_iters, double* elapsed_ms, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CMPY_AMD_H */
