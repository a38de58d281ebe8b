/*
 * multiscat_amd.h -- C ABI of libmultiscat_amd.so, the MI355X (gfx950) close-coupling
 * solver. Plain C types only: ints, doubles and pointers. Complex arrays are interleaved
 * (re, im) doubles. No call aborts the process; every entry point returns a status.
 *
 * Two groups of entry points:
 *
 *  (1) The drop-in group. It replaces the f2py module `multiscat_fortran` one for one:
 *      same functions, same argument meaning, same array layouts (the Fortran side's
 *      column-major 2-D/3-D arrays and C-ravel-ordered 1-D channel waves), same ierr
 *      codes 0/1/2. Each declaration cites the reference routine it replaces (paths
 *      relative to the reference repository root). All arrays are host memory.
 *
 *  (2) The batched group (no counterpart in the reference): a potential table held in
 *      HBM and a batch of incidence conditions that share it, solved together on the
 *      GPU with either the reference's Gauss-Seidel split or the all-channel
 *      block-Jacobi (eigenbasis) preconditioner, in FP64 or in mixed precision (FP32
 *      inner solves with FP64 residual refinement), plus the adjoint solve that gives
 *      d(loss)/dV_G(z).
 */
#ifndef MULTISCAT_AMD_H
#define MULTISCAT_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 0..2 keep the reference's ierr meanings
 * (lib/multiscat_fortran/fortran/multiscat_python_bindings.f90:42-45,58-60). Conditions the
 * reference ends with `error stop` (process abort) get codes of their own here. */
enum {
    MSX_OK = 0,
    MSX_ERR_DIMS = 1,          /* a dimension is <= 0                                        */
    MSX_ERR_NOT_CONVERGED = 2, /* GMRES did not reach the tolerance within maxiter           */
    MSX_ERR_INVALID = 3,       /* null pointer, bad option, inconsistent sizes               */
    MSX_ERR_NO_SPECULAR = 4,   /* specular channel (0,0) not in the channel list             */
    MSX_ERR_EIGEN = 5,         /* eigensolver did not converge ("precon 2" in the reference) */
    MSX_ERR_DEVICE = 6,        /* HIP runtime error or no GPU                                */
    MSX_ERR_OUT_OF_MEMORY = 7  /* device allocation failed                                   */
};

/* Human-readable text of the last error on the calling thread ("" if none). */
const char *msx_last_error(void);

/* Library/device identity: writes the gfx target name (e.g. "gfx950") into buf. */
int msx_device_info(char *buf, int buflen, int *device_count);

/* Waits for all work the library queued on the current device (timing helper; every
 * solve entry point already returns with its results complete). */
int msx_synchronize(void);

/* ---------------------------------------------------------------------------------------
 * (1) Drop-in group: mirrors lib/multiscat_fortran/src/multiscat_fortran/__init__.py:39-169
 *     and the f2py interface lib/multiscat_fortran/fortran/multiscat_python_bindings.pyf:3-97.
 * ------------------------------------------------------------------------------------- */

/* Replaces run_multiscat_fortran (multiscat_python_bindings.f90:1-61 -> multiscat.f90:20-62
 * -> multiscat_gmres.f90:105-618). Solves the close-coupled system for the scattered state
 * with GMRES preconditioned by the block Gauss-Seidel split (I + L^-1 U), left-preconditioned
 * by (I - L^-1 U) when gmres_preconditioner_flag == 1, tolerance 0.5*10^-nsig on the
 * preconditioned residual, 3-iteration window, at most 2000 iterations, x0 = 0, source
 * wave_b[0] at the last node of the specular channel. The arithmetic runs on the GPU in FP64.
 *   potential_values        complex[nkx*nky*nz]  column-major (nkx, nky, nz)
 *   perpendicular_kinetic_difference double[nkx*nky] column-major (nkx, nky)
 *   wave_a, wave_b, wave_c  complex[nkx*nky]      channel idx = ix*nky + iy
 *   parallel_kinetic_energy double[nz*nz]         column-major (nz, nz)
 *   scattered_state_dense   complex[nkx*nky*nz]   column-major (nkx, nky, nz), output
 * Returns the ierr: 0, 1 (bad dims), 2 (not converged; the state is still written). */
int msx_run_multiscat_fortran(int gmres_preconditioner_flag, int convergence_significant_figures,
                              int nkx, int nky, int nz, const double *potential_values,
                              const double *perpendicular_kinetic_difference,
                              const double *wave_a, const double *wave_b, const double *wave_c,
                              const double *parallel_kinetic_energy,
                              double *scattered_state_dense);

/* As msx_run_multiscat_fortran, also returning the iteration count and the final relative
 * (preconditioned) residual estimate |g_{k+1}|/|g_1| (the reference computes both but does
 * not return them, multiscat_gmres.f90:590,608). */
int msx_run_multiscat_fortran_ex(int gmres_preconditioner_flag,
                                 int convergence_significant_figures, int nkx, int nky, int nz,
                                 const double *potential_values,
                                 const double *perpendicular_kinetic_difference,
                                 const double *wave_a, const double *wave_b,
                                 const double *wave_c, const double *parallel_kinetic_energy,
                                 double *scattered_state_dense, int *iterations,
                                 double *residual);

/* Replaces get_perpendicular_kinetic_difference (multiscat_python_bindings.f90:63-124 ->
 * scatsub.f90:54-86): e(ix,iy) = |K + G|^2 - |k|^2 over FFT-ordered G; out column-major. */
int msx_get_perpendicular_kinetic_difference(double incident_kx, double incident_ky,
                                             double incident_kz, int nx, int ny,
                                             double unit_cell_ax, double unit_cell_ay,
                                             double unit_cell_bx, double unit_cell_by,
                                             double *perpendicular_kinetic_difference);

/* Replaces get_parallel_kinetic_energy (multiscat_python_bindings.f90:126-154 ->
 * scatsub.f90:134-226): the Lobatto kinetic matrix of an (nz+1)-node rule on [zmin, zmax]
 * with its first node dropped, nz x nz column-major. */
int msx_get_parallel_kinetic_energy(double zmin, double zmax, int nz,
                                    double *parallel_kinetic_energy);

/* Replaces get_lobatto_weights (multiscat_python_bindings.f90:156-187 -> scatsub.f90:88-100):
 * node_count-point Gauss-Lobatto nodes x on [zmin, zmax] and sqrt of the weights w. */
int msx_get_lobatto_weights(double zmin, double zmax, int node_count, double *w, double *x);

/* Replaces get_abc_arrays (multiscat_python_bindings.f90:189-231 -> scatsub.f90:102-132,
 * 307-335): a, b, c boundary waves of an (n_z_points+1)-node rule; channel idx = ix*ny+iy. */
int msx_get_abc_arrays(double zmin, double zmax, int nx, int ny,
                       const double *perpendicular_kinetic_difference, int n_z_points,
                       double *wave_a, double *wave_b, double *wave_c);

/* Replaces debug_build_preconditioner_fortran (multiscat_python_bindings.f90:233-271 ->
 * multiscat_gmres.f90:226-317): ascending eigenvalues of H0 = T + diag(Re V_00), the factors
 * f_j = (H0 + e_j)^-1 e_N as (nz, nkx*nky) column-major, and the eigenvectors (nz, nz)
 * column-major (eigenvector signs are arbitrary, as with any eigensolver). Runs on the GPU. */
int msx_debug_build_preconditioner(int nkx, int nky, int nz, const double *potential_values,
                                   const double *perpendicular_kinetic_difference,
                                   const double *parallel_kinetic_energy, double *eigenvalues,
                                   double *preconditioner_factors, double *eigenvectors);

/* Replaces debug_apply_upper_block_fortran (multiscat_python_bindings.f90:273-302 ->
 * multiscat_gmres.f90:338-382,620-639): state_out_j = sum_{i>j} V[G_j-G_i] state_in_i over
 * the full grid in storage order; states are (nz, nkx*nky) column-major. Runs on the GPU. */
int msx_debug_apply_upper_block(int nkx, int nky, int nz, const double *potential_values,
                                const double *state_in, double *state_out);

/* Replaces debug_solve_lower_block_fortran (multiscat_python_bindings.f90:304-342 ->
 * multiscat_gmres.f90:384-450,641-687): state_out = L^-1 state_in. Runs on the GPU. */
int msx_debug_solve_lower_block(int nkx, int nky, int nz, const double *potential_values,
                                const double *wave_c,
                                const double *perpendicular_kinetic_difference,
                                const double *parallel_kinetic_energy, const double *state_in,
                                double *state_out);

/* Debug face of the channel coupling in the style of the debug group above (extension):
 * y_j = sum_i V[G_j - G_i] x_i over i > j (mode 0), i != j (mode 1) or i < j (mode 2) for an
 * explicit channel list; precision 0 = FP64, 1 = FP32 (x, y still passed as FP64);
 * algorithm 0 = direct sum, 1 = DFT form (mode 1 only; 2, 3, 4: the DFT form with 1, 2 or 8
 * nodes per wavefront instead of the built-in 4; 5: the DFT form on split-FP16 operands, FP32,
 * square tables up to 32 x 32). V is C order (nkx, nky, nz) here, states are (nc, nz)
 * row-major. */
int msx_debug_apply_coupling(int nkx, int nky, int nz, const double *potential, int nc,
                             const int *chx, const int *chy, const double *state_in,
                             double *state_out, int mode, int precision, int algorithm);

/* Debug face of the diagonal blocks (extension): y_j = D_j^-1 x_j (inverse = 1) or D_j x_j
 * (inverse = 0) with D_j = T + diag(Re V_00) + e_j - c_j e_N e_N^T, for nc channels
 * (multiscat/multiscat/_scipy.py:249-354). potential00: complex[nz] (V_00(z)); kinetic:
 * nz x nz row-major; states (nc, nz) row-major complex. variant 0 = matrix-core GEMMs
 * (nz <= 320), 1 = vector ALU, 2 = matrix-core GEMMs with 8 channels per wavefront (FP64,
 * nz <= 160; variant 0 takes 16 there), 3 and 4 = matrix-core GEMMs on split-FP16 operands
 * with 16 and 8 channels per wavefront (FP32 inverse only, nz <= 224; above 160 the planes are
 * read from L2 and both take 8 channels per wavefront); precision 0 = FP64, 1 = FP32. */
int msx_debug_apply_diagonal(int nz, const double *potential00, const double *kinetic, int nc,
                             const double *e, const double *wave_c, const double *state_in,
                             double *state_out, int inverse, int precision, int variant);

/* ---------------------------------------------------------------------------------------
 * (2) Batched group (extension; no counterpart in the reference).
 * ------------------------------------------------------------------------------------- */

typedef struct msx_table msx_table; /* a potential table resident in HBM */
typedef struct msx_batch msx_batch; /* incidence conditions sharing one table */

enum { MSX_METHOD_GAUSS_SEIDEL = 0, MSX_METHOD_JACOBI = 1 };
enum { MSX_PRECISION_FP64 = 0, MSX_PRECISION_MIXED = 1 };

typedef struct {
    int method;       /* MSX_METHOD_*                                                       */
    int precision;    /* MSX_PRECISION_*                                                    */
    int neumann;      /* 0: none; 1: left preconditioner (I - M^-1 N), the reference's flag;
                         p >= 2 (block Jacobi only): the von Neumann series to degree p,
                         sum_{i<=p} (-D^-1 V)^i, the reference's `order`
                         (multiscat/multiscat/_scipy_von_neumann.py:40-58)                   */
    double eps;       /* stop when |g_{k+1}| < eps |g_1| for `window` consecutive iterations */
    int maxiter;      /* iteration budget per condition (reference: 2000)                   */
    int restart;      /* Krylov dimension before a restart; 0 = unrestarted up to maxiter while
                         the basis and Hessenberg matrices fit in a quarter of the free device
                         memory, else restarted at the largest dimension that fits (fixed at a
                         batch's first solve with these options)                            */
    int window;       /* reference: 3                                                       */
    double inner_eps; /* mixed precision: relative tolerance of each FP32 inner solve;
                         0 = sqrt(eps)/3 clamped to [1e-5, 2.5e-4] (two sweeps reach eps) */
    int max_refine;   /* mixed precision: FP64 refinement sweeps allowed                    */
    int coupling;     /* full coupling: 0 = DFT form on the matrix cores when the table fits
                         (<= 80 x 80; the gradient map of the adjoint: <= 32 x 32), 1 = direct
                         sum                                                                 */
    int diag;         /* all-channel D^-1: 0 = auto (matrix cores when nz <= 320), 1 = vector, 2 = matrix cores */
    int warm_start;   /* 1: GMRES starts from the batch's current state (msx_batch_set_state, or the
                         last solve) instead of 0; tolerance still relative to the source, as an
                         initial guess does in the reference's SciPy path (_gmres.py:106-114) */
    int inner_orth;   /* Gram-Schmidt passes of the FP32 inner Arnoldi (mixed precision): 0 = default
                         (Neumann degree <= 2: one classical pass, and a second for the conditions
                         whose new vector kept a projection on the basis above 1e-4 of its norm;
                         higher degrees: two), 1 = classical single pass, 2 = classical with one
                         reorthogonalisation */
    int autotune;     /* 1 (block Jacobi): at the first solve of a batch shape, time the launch
                         shapes of the DFT coupling (nodes per wavefront), the Arnoldi vector
                         kernels (block length) and the FP64 diagonal blocks (channels per
                         wavefront) on the batch and keep the fastest (cached per shape; the file
                         named by MSX_TUNE_FILE keeps them across processes); 0: built-in shapes */
    int graphs;       /* 1: each Arnoldi iteration's fixed kernel sequence (operator, Gram-Schmidt
                         passes, Givens update) is captured once as a HIP graph per iteration index
                         and launch shape, and replayed by later iterations and solves of the
                         batch; 0: every kernel launched one by one. Off while profiling.      */
    int inner_products; /* mixed precision: operands of the FP32 inner operator's products on
                         the matrix cores (D^-1 and the DFT coupling). 0 = auto (split FP16
                         where the kernels cover the shape: node-major vectors; D^-1 nz <= 224
                         with LDS for the tile scan, the coupling square tables up to 32 x 32;
                         else FP32),
                         1 = FP32 operands, 2 = split FP16 (an error where the D^-1 kernel does
                         not cover the shape). Split FP16: each FP32 operand as a pair hi + lo
                         of FP16 values under a power-of-two scale (22 significant bits), three
                         FP16 products accumulated in FP32.                                 */
} msx_solve_options;

/* Fills `opt` with the reference's settings for a given nsig (GS, FP64, window 3, 2000). */
void msx_default_options(msx_solve_options *opt, int convergence_significant_figures);

/* Uploads a potential table V (complex, C order (nkx, nky, nz), natural units) and the
 * kinetic matrix T (nz x nz, symmetric) and computes the eigensystem of
 * H0 = T + diag(Re V_00) on the GPU in FP64. */
int msx_table_create(int nkx, int nky, int nz, const double *potential, const double *kinetic,
                     msx_table **out);
/* Recomputes H0 and its eigensystem from the table already in HBM (the per-potential set-up
 * of multiscat_gmres.f90:226-273, done once for every batch that shares the table). */
int msx_table_prepare(msx_table *table);
/* Replaces the table's potential with `potential` (complex, C order (nkx, nky, nz), natural
 * units; host memory, or device memory of this process when on_device = 1) and prepares the
 * table again; batches on the table use it from their next solve (the channel factors are
 * rebuilt at every solve). For callers that update the potential in a loop (gradient descent
 * through the autograd bridge) without a new table and batch per step. */
int msx_table_set_potential(msx_table *table, const double *potential, int on_device);
int msx_table_destroy(msx_table *table);
/* Copies the eigensystem to the host: lam[nz] ascending, Q row-major (Q[z*nz + m]). */
int msx_table_eigensystem(const msx_table *table, double *lam, double *Q);
/* Diagnostics of the last eigensolve, ms[7]: durations (ms) of its four phases
 * (tridiagonalisation, accumulation of the Householder reflectors, solution of the tridiagonal
 * problem, back-transformation or reordering), then for the QL solver (nz > 1024) the time (ms)
 * one lane spent chasing bulges, the number of sweeps and of rotations (zero otherwise). */
int msx_table_eig_phases(const msx_table *table, double *ms);
/* Eigensystem of a symmetric n x n matrix A (row-major, host) with the table's GPU solver:
 * lam[n] ascending, Q row-major (A = Q diag(lam) Q^T). method 0: divide and conquer for
 * 2 <= n <= 1024, QL above; 1: Householder + QL; 2: Householder + divide and conquer. Replaces
 * the reference's diagonalisation routine (lib/multiscat_fortran/fortran/diagsub.f90:5-284)
 * for callers that need it outside a table. */
int msx_symmetric_eigensystem(int n, const double *A, int method, double *lam, double *Q);

/* Largest n msx_symmetric_eigensystem_noalloc takes. */
#define MSX_NOALLOC_EIG_MAX_N 512

/* No counterpart in the reference. For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the eigensystem of
 * msx_symmetric_eigensystem (Householder tridiagonalisation + divide and conquer, FP64) on
 * device memory the caller owns, with no allocation of any kind. All pointers are device
 * pointers: A (n x n row-major, symmetric, read only), lam[n] (ascending), Q (n x n row-major,
 * A = Q diag(lam) Q^T), info (one int: 0 on success, nonzero if an eigenvalue is not finite).
 * stream: a hipStream_t, or null for the null stream. The work is only queued: the results
 * and info are ready once the stream has reached this point. 2 <= n <= MSX_NOALLOC_EIG_MAX_N
 * (512), else MSX_ERR_DIMS. Intermediates live in one static device array of about 10 MiB
 * per GPU (held until the process ends); calls that overlap on one GPU race on it. */
int msx_symmetric_eigensystem_noalloc(int n, const double *A, double *lam, double *Q, int *info,
                                      void *stream);

/* Builds a batch of n_cond conditions on `table`. Channel lists are concatenated:
 * condition b owns channels [off_b, off_b + n_channels[b]) with off_b the prefix sum.
 *   chx, chy  storage indices (0 <= chx < nkx, 0 <= chy < nky) in solve order
 *   e         channel energies e_G (natural units)
 *   wave_b, wave_c  complex per channel;  wave_a0 complex per condition (specular a_0)
 * The specular channel of each condition is its first (0,0) entry; the source is b at
 * that channel's last node (multiscat.f90:49). */
int msx_batch_create(msx_table *table, int n_cond, const int *n_channels, const int *chx,
                     const int *chy, const double *e, const double *wave_b,
                     const double *wave_c, const double *wave_a0, msx_batch **out);
int msx_batch_destroy(msx_batch *batch);

/* Solves every condition; the state stays in HBM. Returns MSX_ERR_NOT_CONVERGED if any
 * condition missed the tolerance (see msx_batch_stats). */
int msx_batch_solve(msx_batch *batch, const msx_solve_options *opt);

/* Results. psi: complex, concatenated (n_channels[b] x nz) rows; P: per channel
 * |2i b psi(z_N) + a0 delta_spec|^2 (computed on the GPU in FP64);
 * iterations/info per condition; residual: final relative preconditioned residual. */
int msx_batch_state(const msx_batch *batch, double *psi);
/* Sets the state (same layout as msx_batch_state) that a solve with warm_start = 1 starts from
 * (the reference's ScatteringCondition.initial_state, multiscat/config.py:115-121). */
int msx_batch_set_state(msx_batch *batch, const double *psi);
int msx_batch_intensities(const msx_batch *batch, double *P);
/* The autotuner's choices for this batch and the times they were chosen from, as a JSON
 * object ("{}" before a solve with autotune = 1). MSX_ERR_INVALID if buflen is too short. */
int msx_batch_tuning(const msx_batch *batch, char *buf, int buflen);
int msx_batch_stats(const msx_batch *batch, int *iterations, int *info, double *residual);

/* The lower block operator of the Gauss-Seidel split in each condition's channel order,
 * L = D + V_{i<j} (D_j = T + diag(Re V_00) + e_j - c_j e_N e_N^T): state_out = L state_in
 * (inverse = 0) or L^-1 state_in (inverse = 1), host arrays in the msx_batch_state layout.
 * Replaces the `lower` and `inverse_lower` operators of the reference's SciPy path
 * (multiscat/multiscat/_scipy.py:499-527, applied per channel at :263-296 and :366-388): the
 * reference's get_scattering_state returns L psi (multiscat/multiscat/_gmres.py:127), its
 * get_scattering_matrix_from_state applies L^-1 to a state (_multiscat.py:81-88) and an
 * initial state x enters GMRES as L^-1 x (_gmres.py:106-108). */
int msx_batch_apply_lower(msx_batch *batch, int inverse, const double *state_in,
                          double *state_out);

/* Device pointers for callers that keep data on the GPU (e.g. a torch autograd bridge):
 * the state psi (complex FP64, same layout as msx_batch_state) and the intensities. */
int msx_batch_device_state(const msx_batch *batch, void **psi_device);
int msx_batch_total_channels(const msx_batch *batch, long long *total);

/* Adjoint of the intensities: given dL/dP per channel (host, same layout as
 * msx_batch_intensities), solves A^H lambda = dL/dpsi on the GPU in FP64 and accumulates
 * dL/dV into grad (host, complex C order (nkx, nky, nz)), summed over the batch. V enters
 * through the off-diagonal coupling and through Re V_00 in the diagonal blocks. */
int msx_batch_adjoint(msx_batch *batch, const msx_solve_options *opt, const double *dL_dP,
                      double *grad_potential);

/* Instrumentation (extension): when enabled, the main kernel launches are bracketed by HIP
 * events on the library's stream and carry their algorithmic work (flops, bytes) counted on
 * the host. msx_profile_read sums them per kernel name; msx_profile_names lists the names
 * seen (comma separated). */
int msx_profile_enable(int on);
int msx_profile_reset(void);
int msx_profile_read(const char *kernel, long long *count, double *total_ms, double *flops,
                     double *bytes);
int msx_profile_names(char *buf, int buflen);
/* Restricts the instrumentation to the kernel names in a comma-separated list (NULL or ""
 * instruments every kernel again): the others launch with no events around them. */
int msx_profile_select(const char *names);

#ifdef __cplusplus
}
#endif
#endif /* MULTISCAT_AMD_H */
