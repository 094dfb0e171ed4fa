/*
 * chemtools_amd.h -- C ABI of the MI355X density-descriptor evaluator.
 *
 * Plain pointers and sizes only.  Every function returns CT_OK (0) on success
 * or a negative CT_E* status; the message of the last failure on the calling
 * thread is available from ct_last_error().  Pointer arguments for points and
 * outputs may be host memory or device (HBM) memory: the library detects
 * which.  ct_eval with host arrays runs its chunks as a pipeline: inputs are
 * copied (several host threads) into pinned staging and uploaded, outputs
 * downloaded into pinned staging and copied out, on streams of their own,
 * overlapping the kernels of the neighbouring chunks.  ct_esp and
 * ct_backward copy host arrays in and out around their kernels.
 *
 * Each ct_mol owns its device copies of the shell table and of P, its own HIP
 * stream and workspace, so handles are independent and reentrant (the
 * reference keeps the basis in one global __constant__ bank,
 * src/basis_to_gpu.cu:13, which makes every call non-reentrant).  Calls on
 * one handle from several threads must be serialised by the caller.
 *
 * Reference interfaces replaced (paths relative to qtchem/ChemToolsCUDA):
 *   ct_mol_create            chemtools::Molecule::Molecule + IOData
 *                            (src/pymolecule.cu:18-23, include/iodata.h:18-101)
 *   ct_mol_create_from_fchk  get_molecular_basis_from_fchk (src/iodata.cpp:15-211)
 *   ct_mol_create_from_file  the same, for every format its iodata.load_one call
 *                            (src/iodata.cpp:17-20) reads that this library supports:
 *                            .fchk, AIM .wfn / .wfx, Molden
 *   ct_eval                  evaluate_electron_density_on_any_grid (src/evaluate_density.cu:663-840),
 *                            evaluate_electron_density_gradient (src/evaluate_gradient.cu:1184-1371),
 *                            evaluate_electron_density_hessian (src/evaluate_hessian.cu:1893-2307),
 *                            evaluate_laplacian (src/evaluate_laplacian.cu:843-853),
 *                            evaluate_positive_definite_kinetic_density and
 *                            evaluate_general_kinetic_energy_density (src/evaluate_kinetic_dens.cu:182-214),
 *                            compute_reduced_density_gradient, compute_weizsacker_ked,
 *                            compute_thomas_fermi_energy_density, compute_ked_gradient_expansion_general,
 *                            compute_general_ked, compute_shannon_information_density
 *                            (src/evaluate_densbased.cu:44-223)
 *   ct_eval_cubic            evaluate_electron_density_on_cubic (src/evaluate_density.cu:425-660),
 *                            Molecule::compute_electron_density_cubic (src/pymolecule.cu:59-71)
 *   ct_norm_rows             compute_norm_of_3d_vector (src/evaluate_densbased.cu:10-41)
 *   ct_esp                   compute_electrostatic_potential_over_points (src/evaluate_electrostatic.cu:10161-10372)
 *   ct_boys                  boys_function<M> (include/boys_functions.cuh:54-72)
 *   ct_backward              (no reference counterpart: adjoint of rho w.r.t. P and the centres)
 *   ct_eval_noalloc          (no reference counterpart: ct_eval for callers that cannot allocate)
 */
#ifndef CHEMTOOLS_AMD_H_
#define CHEMTOOLS_AMD_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CT_OK 0
#define CT_E_INVALID -1    /* bad argument (shape, null pointer, unsupported shell) */
#define CT_E_HIP -2        /* HIP runtime error */
#define CT_E_NOMEM -3      /* device allocation failed even at the smallest chunk */
#define CT_E_IO -4         /* file could not be read or parsed */
#define CT_E_UNSUPPORTED -5

/* Shell table.  shell_l: 0..4 Cartesian, -2..-4 pure (reference convention,
 * src/iodata.cpp:76-87); generalised / SP contractions are passed as separate
 * shells sharing exponents.  coefs are contraction coefficients of
 * L2-normalised primitives.  Basis functions are numbered shell by shell, fchk
 * component order within a shell. */
typedef struct {
  int32_t nshell;
  const int32_t* shell_atom;
  const int32_t* shell_l;
  const int32_t* shell_nprim;
  const int32_t* shell_prim_off;
  const double* exps;
  const double* coefs;
} ct_basis_desc;

typedef struct ct_mol ct_mol;

/* Quantity bits for ct_eval. */
#define CT_Q_RHO (1u << 0)           /* rho                           out->rho      (N)      */
#define CT_Q_GRAD (1u << 1)          /* grad rho                      out->grad     (N,3)    */
#define CT_Q_HESS (1u << 2)          /* Hessian of rho                out->hess     (N,3,3)  */
#define CT_Q_LAP (1u << 3)           /* Laplacian of rho              out->lap      (N)      */
#define CT_Q_TAU_PD (1u << 4)        /* 1/2 sum_k d_k phi P d_k phi   out->tau_pd   (N)      */
#define CT_Q_RDG (1u << 5)           /* reduced density gradient      out->rdg      (N)      */
#define CT_Q_SHANNON (1u << 6)       /* rho ln rho                    out->shannon  (N)      */
#define CT_Q_WEIZSACKER (1u << 7)    /* |grad rho|^2 / (8 rho)        out->weizsacker        */
#define CT_Q_THOMAS_FERMI (1u << 8)  /* 0.3 (3 pi^2)^{2/3} rho^{5/3}  out->thomas_fermi      */
#define CT_Q_GGE (1u << 9)           /* TF + a W + b lap              out->gge  (params a,b) */
#define CT_Q_GENERAL_KED (1u << 10)  /* tau_PD + (a-1)/4 lap          out->general_ked (a)   */
#define CT_Q_KED_ALPHA (1u << 11)    /* tau_PD + alpha lap            out->ked_alpha (alpha) */

/* Layouts of points (ct_eval, ct_esp, ct_backward) and of the grad / hess
 * outputs (ct_eval_params.out_layout).  ROW: points (n,3) row-major, x y z of
 * a point together (the reference's Molecule methods, include/pymolecule.cuh:20,
 * and compute_electrostatic_potential_over_points); grad (n,3), hess (n,3,3).
 * COL: points (3,n), x[0..n) then y then z (the reference's free functions,
 * e.g. evaluate_electron_density_on_any_grid, src/evaluate_density.cu:663-664,
 * 745-752); grad (3,n) and hess (9,n) with d_k d_l rho at row 3k+l (the
 * reference's return_row = false, src/evaluate_gradient.cu:1399-1413,
 * src/evaluate_hessian.cu:2244-2279). */
#define CT_LAYOUT_ROW 0
#define CT_LAYOUT_COL 1

/* Precision. */
#define CT_PREC_FP64 0        /* everything in FP64 (parity mode) */
#define CT_PREC_MIXED 1       /* FP32 AOs, split-BF16 MFMA contraction with FP32 accumulation,
                                 FP64 per-point sums/epilogues, FP64 re-evaluation of points
                                 whose error estimate exceeds params.mixed_tol */

typedef struct {
  double gge_a, gge_b;     /* CT_Q_GGE */
  double ked_a;            /* CT_Q_GENERAL_KED */
  double ked_alpha;        /* CT_Q_KED_ALPHA */
  double mixed_tol;        /* relative error target of the mixed path (0 => 1e-6) */
  double rho_floor;        /* points with rho below this are not re-evaluated (0 => 1e-10) */
  int32_t route;           /* CT_ROUTE_AUTO (0), CT_ROUTE_DENSITY, CT_ROUTE_FACTOR (both precisions) */
  int32_t out_layout;      /* grad / hess outputs: CT_LAYOUT_ROW (0) or CT_LAYOUT_COL */
  /* Mixed path, optional (NULL: off): a target density per point (n
   * entries, host or device).  The error of rho is then bounded by
   * mixed_tol * |rho - rho_target| instead of mixed_tol * |rho| (points that
   * miss it are re-evaluated in FP64): for losses on the residual
   * rho - rho_target, whose gradients need the residual, not rho, to
   * mixed_tol.  Points with |rho| below rho_floor are still accepted. */
  const double* rho_target;
} ct_eval_params;

/* Routes of the contraction (every density-family quantity).  DENSITY
 * contracts the AOs with P itself, as the reference does (rho = phi^T P phi).
 * FACTOR contracts them with a pivoted Cholesky factor P = L L^T computed on
 * the host when P is set (psi = L^T phi, rho = psi^T psi, grad rho =
 * 2 psi^T grad psi, ...): rank(P) columns instead of nbasis, available when P
 * is positive semidefinite of rank at most ~half the basis (SCF densities).
 * AUTO takes FACTOR when it is available and needs fewer matrix-core
 * products (all modes, both precisions).  In FP64 the two routes agree to
 * rounding; in the mixed path both meet the same mixed_tol through an FP64
 * re-evaluation of flagged points (on the route that flagged them). */
#define CT_ROUTE_AUTO 0
#define CT_ROUTE_DENSITY 1
#define CT_ROUTE_FACTOR 2

typedef struct {
  double* rho;
  double* grad;
  double* hess;
  double* lap;
  double* tau_pd;
  double* rdg;
  double* shannon;
  double* weizsacker;
  double* thomas_fermi;
  double* gge;
  double* general_ked;
  double* ked_alpha;
} ct_outputs;

typedef struct {
  int64_t n_points;
  int64_t n_reevaluated;   /* mixed path: points recomputed in FP64 */
  int64_t n_chunks;
  double max_error_estimate;
  int64_t factor_rank;     /* rank of the factor if the FACTOR route ran, else -1 */
  double active_fraction;  /* FACTOR route with screening: fraction of the (point tile, k-step)
                              blocks contracted; 1 otherwise */
} ct_eval_report;

/* Lifetime ---------------------------------------------------------------- */
int ct_mol_create(const ct_basis_desc* basis, const double* atom_xyz, const double* charges, int32_t natom,
                  const double* P, int32_t nbasis, ct_mol** out);
int ct_mol_create_from_fchk(const char* path, ct_mol** out);
/* Format from the extension: .fchk/.fch, .wfn, .wfx, .molden/.molden.input;
 * CT_E_UNSUPPORTED for other files. */
int ct_mol_create_from_file(const char* path, ct_mol** out);

/* Host-only access to what a file holds (no device work): the shell table in
 * ct_basis_desc conventions, P (nbasis x nbasis, row-major), coordinates
 * (bohr), nuclear charges and atomic numbers.  Read, query the sizes, copy
 * into caller arrays (any pointer may be NULL), destroy.  Same formats and
 * reference interface as ct_mol_create_from_file. */
typedef struct ct_wavefunction ct_wavefunction;
int ct_wavefunction_read(const char* path, ct_wavefunction** out);
int ct_wavefunction_sizes(const ct_wavefunction* wfn, int32_t* nshell, int32_t* nprim, int32_t* nbasis,
                          int32_t* natom);
int ct_wavefunction_copy(const ct_wavefunction* wfn, int32_t* shell_atom, int32_t* shell_l, int32_t* shell_nprim,
                         int32_t* shell_prim_off, double* exps, double* coefs, double* P, double* atom_xyz,
                         double* charges, int32_t* numbers);
int ct_wavefunction_destroy(ct_wavefunction* wfn);
int ct_mol_destroy(ct_mol* mol);

int ct_mol_nbasis(const ct_mol* mol);
int ct_mol_natom(const ct_mol* mol);
/* Copies of the molecule's data (host pointers). */
int ct_mol_get_coordinates(const ct_mol* mol, double* xyz /* natom*3 */);
int ct_mol_get_numbers(const ct_mol* mol, int32_t* numbers /* natom */);
int ct_mol_get_charges(const ct_mol* mol, double* charges /* natom */);
int ct_mol_get_density_matrix(const ct_mol* mol, double* P /* nbasis^2 */);
/* Replace centres / P (host or device pointers). */
int ct_mol_set_coordinates(ct_mol* mol, const double* atom_xyz);
int ct_mol_set_density_matrix(ct_mol* mol, const double* P);
/* Device-memory budget for the per-chunk workspace (bytes; default a quarter
 * of the device's memory, at most 48 GiB). */
int ct_mol_set_workspace_limit(ct_mol* mol, int64_t bytes);
/* Process-wide cap on the device memory the library's own buffers hold (all
 * handles together; bytes, -1 = no cap, the default).  A workspace that would
 * exceed it fails as a failed hipMalloc does: evaluation halves its chunk and
 * retries, down to 256 points, then returns CT_E_NOMEM.  (The reference
 * asserts instead when its chunk would be empty, src/evaluate_density.cu:690-705.) */
int ct_set_device_allocation_limit(int64_t bytes);
/* Device memory the library's buffers hold now (all handles; bytes). */
int ct_device_allocation_held(int64_t* bytes);

/* Rates this device sustains, measured now on the current device with
 * short probe kernels at the clocks and power of a real run (no reference
 * counterpart; a measuring aid for roofline fractions): bf16 matrix-core
 * TFLOP/s (v_mfma_f32_16x16x32_bf16), FP64 matrix-core TFLOP/s
 * (v_mfma_f64_16x16x4f64), FP64 vector FMA TFLOP/s, HBM copy GB/s (read +
 * write).  Allocates 4 GiB for the copy for the duration of the call. */
int ct_measure_peaks(double* bf16_mfma_tflops, double* f64_mfma_tflops, double* f64_valu_tflops, double* hbm_gbs);

/* Spatial screening of the FACTOR route (no counterpart in the
 * reference, which evaluates every basis function at every point,
 * src/evaluate_density.cu:25-411).  Points are taken in Morton order and,
 * per tile of 16 (rho+grad+lap+tau), 32 (grad, Hessian) or 64 (rho) of them,
 * the 32-function blocks of the basis whose
 * values, first and second derivatives are all at most eps on the tile are
 * skipped; 4 eps ||c||_2 sqrt(rho) (c_j = ||L_j||_1, P = L L^T) is added to
 * each point's error estimate, so mixed_tol still holds for rho.  eps = 0 turns it off; default
 * 1e-15.  The other quantities change by at most about eps times their
 * derivative order's AO scale.  The FP64 path screens at eps * 1e-7: what it
 * skips stays below FP64 rounding of rho wherever rho is above ~1e-12. */
int ct_mol_set_screening(ct_mol* mol, double eps);
/* Distance from the centre beyond which every function of one shell (l,
 * exponents and contraction coefficients as in ct_basis_desc) and its first
 * and second derivatives are at most eps in magnitude: the cutoff the
 * screening uses.  Host only. */
int ct_shell_screening_radius(int32_t l, int32_t nprim, const double* exps, const double* coefs, double eps,
                              double* radius);

/* Evaluation --------------------------------------------------------------- */
/* points: (n,3) row-major (layout CT_LAYOUT_ROW) or (3,n) (CT_LAYOUT_COL).
 * Output arrays named by the quantity bits must be non-null; others are
 * ignored.  params and report may be NULL. */
int ct_eval(ct_mol* mol, const double* points, int64_t n, int32_t layout, uint32_t quantities, int32_t precision,
            const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report);

/* ct_eval without any allocation, for callers that cannot allocate device or
 * pinned memory (work under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No reference counterpart.  Points and outputs must
 * be device memory, n < 2^31; the handle (shell table, P, its factor and the
 * screening tables) is made beforehand.  All scratch lives in one
 * file-scope device array of CT_NOALLOC_SCRATCH_BYTES per GPU, held by that
 * GPU until the process ends: points are processed in chunks that fit it, and
 * flagged points of the mixed path are re-evaluated chunk by chunk.  A call
 * whose smallest chunk (256 points) does not fit returns CT_E_NOMEM with the
 * sizes in the message.  Every call on one GPU shares that GPU's array: calls
 * that overlap on one GPU (other threads, other handles) race and must be
 * serialised by the caller.  Results: FP64 the same bits as ct_eval; mixed
 * within mixed_tol (the screened route groups points by chunk). */
#define CT_NOALLOC_SCRATCH_BYTES (256u << 20)
int ct_eval_noalloc(ct_mol* mol, const double* points, int64_t n, int32_t layout, uint32_t quantities,
                    int32_t precision, const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report);

/* The same quantities on a uniform grid generated on the device (no point
 * upload): point (i, j, k) = origin + i axes[0..2] + j axes[3..5] + k axes[6..8]
 * for 0 <= i < shape[0], 0 <= j < shape[1], 0 <= k < shape[2]; outputs are
 * laid out with k fastest, (shape[0], shape[1], shape[2], width).
 * Replaces evaluate_electron_density_on_cubic (src/evaluate_density.cu:425-660)
 * and Molecule::compute_electron_density_cubic (src/pymolecule.cu:59-71). */
int ct_eval_cubic(ct_mol* mol, const double* origin, const double* axes, const int64_t* shape, uint32_t quantities,
                  int32_t precision, const ct_eval_params* params, const ct_outputs* out, ct_eval_report* report);

/* |v| of each row of an (n,3) array. */
int ct_norm_rows(const double* v, int64_t n, double* out);

/* Molecular electrostatic potential at (n,3) row-major (CT_LAYOUT_ROW) or
 * (3,n) (CT_LAYOUT_COL) points.  CT_PREC_FP64:
 * all in FP64.  CT_PREC_MIXED: the pair geometry, the Boys function and the
 * sums over primitive pairs in FP64, each pair's Hermite polynomial in FP32. */
int ct_esp(ct_mol* mol, const double* points, int64_t n, int32_t layout, int32_t precision, double* out);

/* Primitive pairs of the molecule's ESP pre-contraction, by total angular
 * momentum L = 0..8 (pairs_by_L: 9 entries), after screening: the units of
 * work of ct_esp per point.  (No reference counterpart: the reference
 * evaluates every shell pair's integrals per point.) */
int ct_mol_esp_pairs(ct_mol* mol, int64_t* pairs_by_L);

/* The device Boys function F_0..F_m(T) at n values of T, as the ESP kernels
 * evaluate it for primitive pairs of total angular momentum up to lmax_pairs
 * (which selects the table grid); 0 <= m <= lmax_pairs <= 8; F is (n, m+1)
 * row-major.  T >= 0; host or device pointers. */
int ct_boys(const double* T, int64_t n, int32_t m, int32_t lmax_pairs, double* F);

/* Adjoint of rho: given dL/drho at the points ((n,3) or (3,n) by layout),
 * dL/dP (nbasis x nbasis, row-major, accumulated as sum_g w_g phi_i phi_j)
 * and dL/dR (natom x 3).  Either output may be NULL.  dL/dP is the same bits
 * on every call; dL/dR is summed with FP64 atomics and agrees from call to
 * call to the last bits. */
int ct_backward(ct_mol* mol, const double* points, int64_t n, int32_t layout, const double* dL_drho,
                int32_t precision, double* dL_dP, double* dL_dR);

/* Synchronise the handle's stream. */
int ct_mol_synchronize(ct_mol* mol);

/* Kernel timing with HIP events recorded on the handle's own stream around
 * every launch (enable resets the counters). */
typedef struct {
  double ao_ms, contract_ms, finalize_ms, esp_ms, backward_ms, reeval_ms;
  int64_t ao_calls, contract_calls, finalize_calls, esp_calls, backward_calls, reeval_calls;
  double screen_ms;        /* point sort + tile activity of the screened factor route */
  int64_t screen_calls;
} ct_kernel_stats;
int ct_mol_enable_profiling(ct_mol* mol, int enable);
int ct_mol_kernel_stats(ct_mol* mol, ct_kernel_stats* out);

const char* ct_last_error(void);
const char* ct_version(void);

#ifdef __cplusplus
}
#endif

#endif /* CHEMTOOLS_AMD_H_ */
