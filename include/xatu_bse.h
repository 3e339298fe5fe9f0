/*
 * xatu_bse.h -- C-ABI of the MI355X (gfx950) Bethe-Salpeter kernel builder and
 * exciton solver. This is the drop-in boundary for the hot path of Xatu's
 * ExcitonTB: the body of the Fourier part of ExcitonTB::initializeResultsH0
 * (src/ExcitonTB.cpp:707-787), ExcitonTB::BShamiltonian (src/ExcitonTB.cpp:840-921)
 * and ExcitonTB::diagonalizeRaw (src/ExcitonTB.cpp:930-959, with
 * src/davidson.cpp:5-61). The single-particle band problem (src/ExcitonTB.cpp:737-758)
 * stays with the caller, which passes the eigenvalue/eigenvector stacks in.
 *
 * Conventions
 *  - Plain C types only. Complex numbers are interleaved (re, im) pairs of
 *    doubles (or floats where stated), laid out exactly as Armadillo stores
 *    cx_mat / cx_cube (column-major), so a reference caller can pass
 *    eigvecKStack.memptr() and receive HBS into HBS.memptr().
 *  - Every call is synchronous on the context's own HIP stream; host buffers
 *    are owned by the caller; device buffers are owned by the context.
 *  - Return value 0 (XATU_OK) on success, otherwise an XATU_ERR_* code, with
 *    a message from xatu_last_error(). The reference throws
 *    std::invalid_argument / std::logic_error at the same points.
 */
#ifndef XATU_BSE_H
#define XATU_BSE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xatu_ctx xatu_ctx;

/* Status codes */
#define XATU_OK 0
#define XATU_ERR_INVALID_ARGUMENT 1 /* reference: std::invalid_argument */
#define XATU_ERR_LOGIC 2            /* reference: std::logic_error, call order */
#define XATU_ERR_DEVICE 3           /* HIP / rocBLAS / rocSOLVER failure */
#define XATU_ERR_OUT_OF_MEMORY 4
#define XATU_ERR_NOT_CONVERGED 5
#define XATU_ERR_NOT_HERMITIAN 6    /* reference: davidson_method invalid_argument */

/* Arithmetic of a build or of a solve */
#define XATU_FP64 0
#define XATU_FP32 1
/* xatu_build_bse mode that prepares matrix-free products without storing the matrix */
#define XATU_MATRIX_FREE 2
/* Matrix-free product forms (xatu_solve_options.matrix_free, xatu_bse_matvec) */
#define XATU_MATRIX_FREE_CONV 1
#define XATU_MATRIX_FREE_TILES 2

/* Interaction modes (ExcitonTB::setMode) */
#define XATU_MODE_REALSPACE 0
#define XATU_MODE_RECIPROCAL 1

/* Interaction potentials (src/ExcitonTB.cpp:347-399) */
#define XATU_POTENTIAL_KELDYSH 0
#define XATU_POTENTIAL_COULOMB 1

/* Solve methods (src/ExcitonTB.cpp:940-954) */
#define XATU_SOLVE_DIAG 0     /* dense: all eigenpairs, ascending ("diag") */
#define XATU_SOLVE_DAVIDSON 1 /* block Davidson, lowest nstates ("davidson"; also serves "sparse") */

/* Hellmann-Feynman parameters */
#define XATU_PARAM_R0 0
#define XATU_PARAM_EPS_M 1
#define XATU_PARAM_EPS_S 2

/* Screened interaction set-up: ExcitonTB attributes used by initializeResultsH0. */
typedef struct xatu_interaction {
    double eps_m;          /* dielectric constant of the embedding medium */
    double eps_s;          /* dielectric constant of the substrate */
    double r0;             /* screening length (Angstrom), must be non-zero */
    double regularization; /* V(0) is V(regularization); <= 1e-10 selects |a1| (ExcitonTB.cpp:803-805) */
    double cutoff;         /* cutoff_ in units of |a1|: cells with |R| < |a1|*cutoff + 1e-5 are summed,
                              and V(r) = 0 for r > |a1|*cutoff + 1e-5 (ExcitonTB.cpp:350,710) */
    int ncell;             /* cells per axis of the centred supercell the cells are drawn from */
    int total_cells;       /* normalisation of the lattice sums, (ncell*factor)^ndim (ExcitonTB.cpp:808) */
    int potential;          /* XATU_POTENTIAL_* for the direct term */
    int exchange_potential; /* XATU_POTENTIAL_* for the exchange term */
    int exchange;           /* non-zero: also compute V(Q) for the exchange term */
    double Q[3];            /* centre-of-mass momentum (1/Angstrom, Cartesian) */
    int mode;               /* XATU_MODE_REALSPACE (lattice sums) or XATU_MODE_RECIPROCAL (sums over G,
                               ExcitonTB.cpp:408-426, 562-621) */
    int n_reciprocal;       /* reciprocal mode: G vectors from the centred n_reciprocal^ndim combinations
                               with |G| < cutoff*|b1| + 1e-5 (Lattice.cpp:349-370) */
} xatu_interaction;

typedef struct xatu_solve_options {
    int method;      /* XATU_SOLVE_DIAG or XATU_SOLVE_DAVIDSON */
    int precision;   /* XATU_FP64: products with the FP64 matrix.
                        XATU_FP32: products with the FP32 matrix, then FP64 refinement */
    int matrix_free; /* 0: resident matrix. XATU_MATRIX_FREE_CONV (1): matrix-free, the direct term as a
                        k-space convolution (FFT) and the exchange term along the upper triangle;
                        XATU_MATRIX_FREE_TILES (2): matrix-free, tiles of H recomputed inside each
                        product on the matrix cores. The reciprocal-space mode has no periodic
                        table to convolve with: its matrix-free products always use the tiles */
    int nstates;     /* number of lowest eigenpairs wanted */
    int block;       /* Davidson block size; 0 selects max(2*nstates, 8) */
    int max_basis;   /* Davidson basis size before restart; 0 selects a default */
    int max_iter;    /* 0 selects a default */
    double tol;      /* FP64 residual 2-norm tolerance in eV (0 selects 1e-5); the FP32 form iterates
                        to max(tol/2, 2e-5) with FP32 products, then refines with FP64 products */
    int fp32_arith;  /* products with the FP32 matrix: XATU_ARITH_SPLIT16 (0, default) runs them on the
                        FP16 matrix cores with every FP32 operand split into two FP16 parts (22-bit
                        operands, FP32 accumulation) when N is even, else on the FP32 matrix cores;
                        XATU_ARITH_FP32 (1) always uses the FP32 matrix cores */
} xatu_solve_options;

#define XATU_ARITH_SPLIT16 0
#define XATU_ARITH_FP32 1

typedef struct xatu_solve_report {
    int iterations;          /* Davidson iterations (both phases) */
    int products;            /* matrix-vector products (vectors) */
    int restarts;
    int refine_iterations;   /* FP64 refinement iterations after the FP32 phase */
    int converged;           /* 1 if every wanted residual met tol */
    double max_residual;     /* FP64 residual ||H x - e x|| max over the wanted states */
    double ms_total;         /* wall time of the solve (ms) */
    double ms_products;      /* time in the matrix-vector products (ms, HIP events) */
    double ms_rayleigh_ritz; /* time in the host Rayleigh-Ritz eigenproblems (ms) */
    double ms_projections;   /* time in the basis projections and Ritz-vector products (ms, HIP events) */
} xatu_solve_report;

typedef struct xatu_timings {
    double ms_fourier;      /* last xatu_fourier_sums (HIP events) */
    double ms_build;        /* last xatu_build_bse: all its kernels (HIP events) */
    double ms_build_kernel; /* last xatu_build_bse: the assembly kernel alone (HIP events) */
    double ms_matvec_kernel;/* average duration of the dense product kernel over the last solve */
    int n_matvec_kernel;    /* number of dense product kernel launches in the last solve */
    int n_matvec_vectors;   /* vectors multiplied by those launches */
} xatu_timings;

/* ---- context ---- */
int xatu_create(xatu_ctx** ctx, int device);
void xatu_destroy(xatu_ctx* ctx);
const char* xatu_last_error(const xatu_ctx* ctx);

/* Lattice and k-mesh geometry (Lattice.cpp:32-47,140-162).
 * bravais: ndim x 3 row-major (row = lattice vector), Angstrom.
 * motif: natoms x 3 row-major atomic positions (1 <= natoms <= 16); norb_per_atom: orbitals of each atom
 * (the basis orders orbitals atom by atom, SystemTB/ExcitonTB.cpp:535-543).
 * mesh_n: points per axis of the full Monkhorst-Pack mesh the lattice sums are taken on
 * (ncell*factor, Lattice.cpp:172-197). */
int xatu_set_lattice(xatu_ctx* ctx, int ndim, const double* bravais, int natoms, const double* motif,
                     const int* norb_per_atom, int mesh_n);

/* Band stacks (ExcitonTB.cpp:717-758): nb = nv + nc bands per k, valence first.
 * eigval_k, eigval_kq: nb x nk column-major (eigvalKStack, eigvalKQStack).
 * eigvec_k, eigvec_kq: complex norb x nb x nk (eigvecKStack, eigvecKQStack), phase-fixed.
 * k_offsets: nk x ndim row-major integers, the offset of each k-point from k-point 0 in
 * units of the mesh spacing 1/mesh_n of each reciprocal vector; k' - k maps to the mesh
 * point (k'-k) mod mesh as Lattice::findEquivalentPointBZ does (Lattice.cpp:251-279). */
int xatu_set_bands(xatu_ctx* ctx, int nk, int nv, int nc, const double* eigval_k, const double* eigval_kq,
                   const double* eigvec_k, const double* eigvec_kq, const int* k_offsets);

/* Lattice Fourier sums V_ab(q) on every mesh point and, if exchange, V_ab(Q)
 * (ExcitonTB.cpp:438-476, 695-697, 761-786). The potential table V_ab(R) is evaluated
 * once in FP64 (Keldysh: Struve H0 minus Neumann Y0, ExcitonTB.cpp:306-366). */
int xatu_fourier_sums(xatu_ctx* ctx, const xatu_interaction* inter);

/* Assemble the BSE matrix in the given arithmetic and keep it resident in HBM
 * (ExcitonTB.cpp:840-921). scissor in eV. If hbs_out is non-null it receives the full
 * N x N matrix, column-major complex (double pairs for FP64, float pairs for FP32).
 * Without hbs_out an FP32 build at even N assembles only the block-lower half that the
 * split-FP16 products read (the rest follows on the first call that needs it:
 * xatu_get_bse, xatu_compare_builds or a solve with XATU_ARITH_FP32).
 * With precision XATU_MATRIX_FREE nothing is stored: the call only prepares the
 * exchange factors and the diagonal so that solves and products can run matrix-free. */
int xatu_build_bse(xatu_ctx* ctx, int precision, double scissor, void* hbs_out);

/* Same, for an explicit subset basis (ExcitonTB::BShamiltonian(const arma::imat&),
 * src/ExcitonTB.cpp:840-921): basis is nbasis x 3 row-major {valence column, conduction
 * column, k index}, columns indexing the band stacks (bandToIndex). FP64 only. The
 * nbasis x nbasis matrix stays resident as the operator of the following calls (its
 * dimension replaces N): xatu_solve (DIAG, or DAVIDSON with FP64 products on the resident
 * matrix), xatu_get_bse(FP64), xatu_bse_matvec(FP64); the calls that need the full basis
 * (FP32 and matrix-free products, gradients, oscillator strengths, build comparison)
 * return XATU_ERR_LOGIC until the next xatu_build_bse, xatu_set_bands or
 * xatu_fourier_sums restores it. hbs_out (nullable) receives a copy. */
int xatu_build_bse_basis(xatu_ctx* ctx, int nbasis, const int* basis, double scissor, double* hbs_out);

/* Copy a resident matrix to the host (same layouts as xatu_build_bse). */
int xatu_get_bse(xatu_ctx* ctx, int precision, void* hbs_out);
/* Release a resident matrix. */
int xatu_free_bse(xatu_ctx* ctx, int precision);

/* Diagonalise (ExcitonTB.cpp:930-959). eigval_out: nstates (all N for DIAG) ascending.
 * eigvec_out (nullable): N x nstates complex128 column-major. */
int xatu_solve(xatu_ctx* ctx, const xatu_solve_options* opt, double* eigval_out, double* eigvec_out,
               xatu_solve_report* report);

/* y = H x for nvec vectors (complex128, N x nvec column-major) with the resident matrix of
 * the given precision, or matrix-free (matrix_free: XATU_MATRIX_FREE_CONV or _TILES).
 * FP32 products use the default arithmetic (XATU_ARITH_SPLIT16). */
int xatu_bse_matvec(xatu_ctx* ctx, int precision, int matrix_free, int nvec, const double* x, double* y);

/* No counterpart in the reference: y = H x with the resident FP64 matrix for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). x and y are DEVICE pointers (N x nvec complex128,
 * column-major, caller-allocated); the call allocates nothing: its staged vectors and
 * partial sums live in a 64 MiB scratch array that each GPU holds from xatu_create until
 * the process ends, so N is limited to 65536 (XATU_ERR_INVALID_ARGUMENT beyond). Every call
 * on one GPU shares that GPU's scratch: calls that overlap on one GPU (other threads, other
 * contexts) race. Synchronous on the context's stream; same kernels as xatu_bse_matvec. */
int xatu_bse_matvec_noalloc(xatu_ctx* ctx, int nvec, const double* x, double* y);

/* Hellmann-Feynman gradients dE_n/dtheta = psi_n^H (dH/dtheta) psi_n for the screening
 * parameters (XATU_PARAM_*), without forming dH. eigvec: N x nstates complex128 (null: the
 * eigenvectors of the last solve). grad_out: nparams x nstates, state index fastest.
 * Real-space mode only. */
int xatu_hf_gradients(xatu_ctx* ctx, int nstates, const double* eigvec, int nparams, const int* params,
                      double* grad_out);

/* Sensitivities of exciton energies to the band data, the adjoint of the BSE assembly
 * (no reference counterpart; used for gradients with respect to the tight-binding
 * hopping amplitudes together with first-order band perturbation theory on the host).
 * For each state n (eigvec: N x nstates complex128, null: the last solve's vectors):
 *   w_out: dE_n/d eps_b(k), nstates x 2 x nk x nb doubles (state slowest; stack 0 is
 *          eigvalKStack, stack 1 eigvalKQStack, band fastest),
 *   g_out: dE_n/d conj(C_b(k)) (Wirtinger derivative of the real E_n with respect to
 *          the eigenvector coefficients), nstates x 2 stacks of norb x nb x nk complex128
 *          (the eigvecKStack / eigvecKQStack layout),
 * so that a change of the band data changes E_n by sum [w d(eps) + 2 Re(conj(g) dC)].
 * Only the columns the BSE uses are non-zero: valence columns of the k stack and
 * conduction columns of the k+Q stack. */
int xatu_band_gradients(xatu_ctx* ctx, int nstates, const double* eigvec, double* w_out, double* g_out);

/* Exciton oscillator strengths (the exciton sum of ResultTB::excitonOscillatorStrength,
 * src/ResultTB.cpp:242-298 / src/skubo_w.f90:300-323): out[j + 3 n] = sum_i X_n(i) v_j(i)
 * for j = x, y, z, with the pair velocity matrix elements velocity[i + N j] (complex128,
 * atomic units, computed on the host from the band stacks) and the state vectors X
 * (eigvec: N x nstates complex128, null: the last solve's vectors); no conjugation of X,
 * as in the reference. out: 3 x nstates complex128. */
int xatu_oscillator_strengths(xatu_ctx* ctx, int nstates, const double* eigvec, const double* velocity,
                              double* out);

/* Read-backs for parity checks.
 * ft_out: natoms x natoms x nmesh complex128 (ftMotifStack); ftq_out: natoms x natoms (ftMotifQ).
 * Real-space mode only (the reciprocal mode has no lattice-sum stack).
 * cells: number of lattice vectors nR and (if coefs non-null) nR x ndim integer coefficients. */
int xatu_get_fourier(xatu_ctx* ctx, double* ft_out, double* ftq_out);
int xatu_get_cells(xatu_ctx* ctx, int* n_cells, int* coefs);
/* Per-block error of the resident FP32 matrix against the resident FP64 matrix: the
 * matrix is cut into blocks of kblock k-points per side; block_err (nblk x nblk,
 * column-major) receives max |H32 - H64| per block, max_abs the overall maximum. */
int xatu_compare_builds(xatu_ctx* ctx, int kblock, double* block_err, double* max_abs);
int xatu_get_timings(const xatu_ctx* ctx, xatu_timings* out);

#ifdef __cplusplus
}
#endif

#endif /* XATU_BSE_H */
