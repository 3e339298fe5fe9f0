// Internal state of an xatu_ctx and helpers shared by the device sources.
#pragma once

#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdint>
#include <stdexcept>
#include <memory>
#include <string>
#include <vector>

#include "../../../include/xatu_bse.h"

namespace xatu_dev {

// Error carrying an XATU_ERR_* status through the C-ABI layer.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define XATU_HIP(call)                                                                          \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            throw ::xatu_dev::Error(e_ == hipErrorOutOfMemory ? XATU_ERR_OUT_OF_MEMORY          \
                                                              : XATU_ERR_DEVICE,                \
                                    std::string(#call) + ": " + hipGetErrorString(e_));         \
        }                                                                                       \
    } while (0)

#define XATU_BLAS(call)                                                                         \
    do {                                                                                        \
        rocblas_status s_ = (call);                                                             \
        if (s_ != rocblas_status_success) {                                                     \
            throw ::xatu_dev::Error(XATU_ERR_DEVICE, std::string(#call) + ": rocblas status " + \
                                                         std::to_string(static_cast<int>(s_))); \
        }                                                                                       \
    } while (0)

// Owning device allocation.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        if (count == n && p) return;
        release();
        if (count == 0) return;
        XATU_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    // grow only: keeps an allocation of at least count entries (its contents are undefined
    // after a regrow), so that repeated solves and steps allocate and free nothing
    void reserve(size_t count) {
        if (count <= n && p) return;
        alloc(count);
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Owning pinned host allocation, grown on demand: small transfers to and from it need no
// staging copy. reserve_mapped: fine-grained (coherent) memory mapped into the device's
// address space at dev, which kernels store into directly (visible to the host once the
// kernel has completed, with no copy).
template <typename T>
struct HostBuf {
    T* p = nullptr;
    T* dev = nullptr;
    size_t n = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf() { release(); }
    void release() {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        dev = nullptr;
        n = 0;
    }
    void reserve(size_t count) {
        if (count <= n && p && !dev) return;
        release();
        XATU_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
        n = count;
    }
    void reserve_mapped(size_t count) {
        if (count <= n && p && dev) return;
        release();
        XATU_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocMapped | hipHostMallocCoherent));
        n = count;
        XATU_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev), p, 0));
    }
};

// Device workspace of the block Davidson solver: basis V, products W = H V (N x mmax), Ritz
// block X, AX, residuals R, scratch block Tb (N x b), projections P, Ritz coefficients Y
// (mmax x b), Ritz values theta (b), residual partial sums, the projected matrix Tm (mmax x
// mmax) and its reflectors (device Rayleigh-Ritz). Held by the context between solves.
struct SolverWorkspace {
    DevBuf<double2> V, W, X, AX, R, Tb;
    DevBuf<double2> P, Y;
    DevBuf<double> theta, partial;
    DevBuf<double2> Tm, refl;
};

// Geometry and lattice-sum description shared with the kernels (by value).
struct LatticeDesc {
    int ndim;
    int natoms;
    int norb;
    int mesh_n;   // points per axis of the full mesh
    int nmesh;    // mesh_n^ndim
    double bravais[9];
    double tau[16 * 3];  // atomic positions (natoms <= kMaxAtoms)
    int orb_begin[17];   // orbital offset of each atom, orb_begin[natoms] = norb
    // interaction table: real-space mode indexes it by the mesh point of k_b - k_a
    // (tab_n = mesh_n); reciprocal mode by the raw offset difference k_a - k_b + (mesh_n - 1)
    // per axis (tab_n = 2 mesh_n - 1), because its sums over G are not periodic in q
    int recip;
    int tab_n;
    int ntab;           // tab_n^ndim
    double recip_vec[9];  // reciprocal lattice vectors (rows), 1/Angstrom
};

constexpr int kMaxAtoms = 16;
static_assert(sizeof(LatticeDesc::tau) == 3 * kMaxAtoms * sizeof(double), "motif storage");
static_assert(sizeof(LatticeDesc::orb_begin) == (kMaxAtoms + 1) * sizeof(int), "orbital offsets");

// Potential parameters for the device evaluation of V(r).
struct PotDesc {
    int kind;         // XATU_POTENTIAL_*
    double eps_bar;   // (eps_m + eps_s) / 2
    double r0;
    double reg;       // regularization distance
    double cutoff;    // |a1|*cutoff_ + 1e-5
};

// State of the convolution (FFT) form of the matrix-free product (convolution.hip).
struct ConvState;
void destroy_conv_state(ConvState* s);

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    rocblas_handle blas = nullptr;
    std::string err;

    // lattice
    bool have_lattice = false;
    LatticeDesc lat{};
    double a1norm = 0.0;

    // bands
    bool have_bands = false;
    int nk = 0, nv = 0, nc = 0, nb = 0;
    int64_t N = 0;
    std::vector<double> h_eval_k, h_eval_kq;  // nb x nk
    std::vector<int> h_koff;                  // nk x ndim
    DevBuf<double> d_eval_k, d_eval_kq;       // nb x nk
    DevBuf<double2> d_evec_k, d_evec_kq;      // norb x nb x nk
    DevBuf<int> d_kq;                         // nk: mesh linear index of each k (offset coords)
    DevBuf<int> d_kcoord;                     // nk x 3 mesh coordinates (padded to 3)

    // interaction
    bool have_ft = false;
    xatu_interaction inter{};
    PotDesc pot_d{}, pot_x{};
    int nR = 0;
    std::vector<int> h_cells;   // nR x ndim integer coefficients
    DevBuf<int> d_cells;        // nR x 3 (padded)
    DevBuf<double> d_vtab;      // npairs x nR direct potential table
    DevBuf<double> d_vtabx;     // npairs x nR exchange potential table
    DevBuf<double2> d_twiddle;  // 2*mesh_n roots of unity
    DevBuf<double2> d_ft;       // natoms x natoms x nmesh
    mutable DevBuf<float2> d_ft32;  // its FP32 copy for the FP32 build (rows padded to an even length)
    DevBuf<double2> d_ftq;      // natoms x natoms
    DevBuf<double2> d_e1, d_e2; // exchange factors: nk x natoms x (nc*nv)

    // subset basis (xatu_build_bse_basis): the resident FP64 matrix and N are those of the
    // subset until the next full build, set_bands or fourier_sums
    bool subset = false;
    void require_full_basis(const char* what) const {
        if (subset) {
            throw Error(XATU_ERR_LOGIC, std::string(what) +
                                            " needs the full exciton basis; a subset-basis Hamiltonian is resident "
                                            "(call xatu_build_bse)");
        }
    }

    // resident matrices
    double scissor = 0.0;
    bool built64 = false, built32 = false;
    bool prepared = false;  // exchange factors and diagonal ready for products
    int prep_gen = 0;       // incremented by every preparation of the operator
    std::unique_ptr<ConvState, void (*)(ConvState*)> conv{nullptr, destroy_conv_state};
    DevBuf<double2> d_H64;
    DevBuf<float2> d_H32;
    std::vector<double> h_diag;  // diagonal of H (FP64), N
    DevBuf<double> d_diag;

    // scratch for products and projections
    DevBuf<double2> s_xt64, s_yp64;   // staged vectors / split partial sums, FP64 dense product
    DevBuf<float2> s_xt32, s_yp32;    // same, FP32 dense product (partial sums also of the split product)
    bool h32_lower = false;           // the FP32 matrix holds only its block-lower part (split product)
    bool h32_presplit = false;        // ... and holds it as scaled split FP16 parts (the product's load order)
    DevBuf<float4> s_xs16;            // vectors staged as split-FP16 fragments
    DevBuf<float4> s_xt16;            // same, in the row order of the transposed-block products
    DevBuf<float2> s_pt32;            // partial sums of the transposed-block products (split product)
    DevBuf<int2> d_herm_items;        // (row block, column chunk) work items of the split product
    int herm_items = 0;
    int64_t herm_n = 0;               // matrix dimension the work items were laid out for
    DevBuf<float> s_colmax, s_xscale; // column maxima and power-of-two scales of the staged vectors
    DevBuf<double> s_xinv;            // inverse combined scales applied by the split-product reduction
    DevBuf<double> s_bounds;          // per-block partial bounds of the split product's scales
    float h16_scale = 0.0f;           // power-of-two scale of H for the split product (0: unavailable)
    bool fp32_arith_only = false;     // products with the FP32 matrix stay on the FP32 matrix cores
    DevBuf<double2> s_gram;           // split partial sums of the Gram kernel
    // pinned host staging of the Davidson solver's per-iteration transfers: projections,
    // Ritz coefficients and values, residual partial sums, block whitening coefficients
    HostBuf<double2> h_proj, h_coef, h_whiten;
    HostBuf<double> h_theta, h_part;
    HostBuf<double> h_gather;  // mapped: the per-iteration results of the device solver
    // solver workspaces, kept (grow only) between solves: [0] the FP64 solve or refinement,
    // [1] the FP32 phase of the mixed form, which both exist during a mixed solve
    SolverWorkspace solver_ws[2];

    // last solve
    int last_nstates = 0;
    int64_t last_n = 0;          // dimension of the operator the last solve ran on
    DevBuf<double2> d_last_vec;  // N x last_nstates (allocation kept, grow only)
    std::vector<double> last_val;
    // forget the last solve's eigenpairs: they belong to an operator that has changed (the
    // buffer stays allocated for the next solve; last_vectors refuses it until then)
    void drop_solution() {
        last_nstates = 0;
        last_n = 0;
        last_val.clear();
    }
    // eigenvectors of the last solve for nstates states of the current operator
    const double2* last_vectors(int nstates) const {
        if (last_nstates < nstates || last_n != N || !d_last_vec.p) {
            throw Error(XATU_ERR_LOGIC, "no eigenvectors of the current BSE Hamiltonian from a previous solve");
        }
        return d_last_vec.p;
    }

    // timings
    xatu_timings tim{};

    void require_lattice() const {
        if (!have_lattice) throw Error(XATU_ERR_LOGIC, "lattice not set (call xatu_set_lattice first)");
    }
    void require_bands() const {
        if (!have_bands) throw Error(XATU_ERR_LOGIC, "Error: BZ mesh must be initialized first");
    }
    void require_ft() const {
        if (!have_ft) throw Error(XATU_ERR_LOGIC, "lattice Fourier sums not computed (call xatu_fourier_sums)");
    }
};

// Simple event timer on the context stream.
struct EventTimer {
    hipEvent_t a = nullptr, b = nullptr;
    hipStream_t s;
    explicit EventTimer(hipStream_t st) : s(st) {
        XATU_HIP(hipEventCreate(&a));
        XATU_HIP(hipEventCreate(&b));
    }
    ~EventTimer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    void start() { XATU_HIP(hipEventRecord(a, s)); }
    void stop() { XATU_HIP(hipEventRecord(b, s)); }
    double ms() {
        XATU_HIP(hipEventSynchronize(b));
        float t = 0.f;
        XATU_HIP(hipEventElapsedTime(&t, a, b));
        return static_cast<double>(t);
    }
};

// Kernel launchers implemented in the .hip translation units.
void launch_potential_table(const Context& c, const PotDesc& pot, double* out, hipStream_t s);
void launch_twiddles(int mesh_n, double2* out, hipStream_t s);
void launch_fourier_mesh(const Context& c, hipStream_t s);
void launch_fourier_q(const Context& c, hipStream_t s);
// Reciprocal-mode tables: M_ab(q) over the raw mesh differences into d_ft and, with
// exchange, M_ab(Q) into d_ftq; d_gvec holds ng reciprocal lattice vectors (x, y, z).
void launch_reciprocal_table(const Context& c, const double* d_gvec, int ng, double pref, hipStream_t s);
void launch_exchange_factors(const Context& c, hipStream_t s);
void launch_build_tiles64(const Context& c, double2* H, hipStream_t s);
// lower: only the tiles holding entries (k, i) with k >= kSplitRowBlock floor(i / kSplitRowBlock)
// (the block-lower part of column-major H read by the split product)
// presplit: write the block-lower part in the split FP16 layout the product on the FP16
// matrix cores reads (see presplit_build_available).
void launch_build_tiles32(const Context& c, float2* H, bool lower, bool presplit, hipStream_t s);
// The FP32 build can write the split FP16 layout: N and the tile rows multiples of 32.
bool presplit_build_available(const Context& c);
// Fills the block-upper part of a FP32 matrix built block-lower (Hermitian mirror).
void complete_matrix32(Context& c);
// Before a block-lower build: zeroes the never-built entries that the split product reads
// past the end of a column.
void clear_upper_heads(Context& c);
// The split product's layout is usable for the current operator (even N, finite scale).
bool split_layout_available(const Context& c);
void launch_build_basis(const Context& c, int nbasis, const int* d_basis, double scissor, double2* H, hipStream_t s);
void launch_diag_extract(const Context& c, const double2* H, double* diag, hipStream_t s);
void launch_diag_direct(const Context& c, double* diag, hipStream_t s);
// The (natoms, nc, nv) shape has specialised tile kernels (natoms <= 4, nc == nv <= 3);
// every other shape runs on the element-wise kernels.
bool shape_specialised(const Context& c);
void launch_compare(const Context& c, int kblock, double* d_blockerr, hipStream_t s);

// Dense products Y = H X (N x nvec, complex FP64 vectors) with the resident FP64 or
// FP32 matrix (the FP32 product stages and accumulates in FP32).
void launch_matvec64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
// The FP64 product with a caller-provided scratch (staged vectors, then the partial sums
// of as many column splits as fit, at most max_splits): no allocation.
// matvec64_scratch_elems is the scratch the product uses with at most max_splits splits.
int64_t matvec64_scratch_elems(int64_t N, int max_splits);
void launch_matvec64_scratch(const Context& c, const double2* X, double2* Y, int nvec, double2* scratch,
                             int64_t scratch_elems, hipStream_t s);
void launch_matvec32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
// Products with the resident FP32 matrix on the FP16 matrix cores, operands split into
// hi + lo FP16 parts (split_product.hip). split_product_scale sets h16_scale from bounds
// on |H| (after the tables, bands and diagonal are ready); the split product needs an
// even N (16-byte aligned columns) and the FP32 matrix with 32 zero entries of padding.
void split_product_scale(Context& c);
bool split_product_available(const Context& c);
void launch_matvec_split16(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
// C (m x n, leading dimension ldc) = A^H B over the N rows of A (lda) and B (ldb), FP64.
void launch_gram(Context& c, const double2* A, int lda, const double2* B, int ldb, int m, int n, double2* C, int ldc,
                 hipStream_t s);
// Matrix-free products Y = H X by k-space convolution of the direct term (FFT) plus
// the exchange term along the upper triangle, in FP64 / FP32 arithmetic.
void launch_conv64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
void launch_conv32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
// S[(n nmesh + q) NA^2 + a + NA b] = sum_k G_ab(k, k+q) for ns <= 32 states (FFT autocorrelation).
void conv_pair_density(Context& c, const double2* psi, int ns, double2* S);
// Matrix-free products Y = H X (complex FP64 vectors) with tiles of H recomputed in
// FP64 / FP32 arithmetic inside the product (FP32: FP32 tiles, FP32 accumulation).
void launch_matfree64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
void launch_matfree32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s);
// Vector blocks of the products: Xt[row][32] (zero padded beyond nvec <= 32) in the
// product's precision, and the fixed-order sum Y[row + n N] = sum_s Yp[s][row][n].
void launch_stage_vectors(const double2* X, int64_t N, int nvec, double2* Xt, hipStream_t s);
void launch_stage_vectors(const double2* X, int64_t N, int nvec, float2* Xt, hipStream_t s);
// C (N x nb) = alpha A B + beta C, A: N x m (lda), B: m x nb (ldb), FP64 complex, in
// launches of 32 columns; beta = 0 does not read C. C must not overlap A or B.
void launch_block_gemm(const double2* A, int lda, const double2* B, int ldb, int64_t N, int m, int nb, double alpha,
                       double beta, double2* C, int ldc, hipStream_t s);
void launch_reduce_partials(const double2* Yp, int64_t N, int nvec, int splits, double2* Y, hipStream_t s);
void launch_reduce_partials(const float2* Yp, int64_t N, int nvec, int splits, double2* Y, hipStream_t s);

// Solver entry points (davidson.cpp).
void solve_dense_diag(Context& c, int nstates, double* eigval_out, double* eigvec_out);
void solve_davidson(Context& c, const xatu_solve_options& opt, double* eigval_out, double* eigvec_out,
                    xatu_solve_report* rep);
void apply_operator(Context& c, int precision, int matrix_free, const double2* X, double2* Y, int nvec,
                    double* ms_kernel);
void launch_operator(Context& c, int precision, int matrix_free, const double2* X, double2* Y, int nvec);
void record_product(Context& c, int matrix_free, double ms, int nvec);

// Hellmann-Feynman gradients (hf.hip / capi.cpp).
void hf_gradients(Context& c, int nstates, const double2* d_vec, int nparams, const int* params,
                  double* grad_out);
// out[j + 3 n] = sum_i X[i + n N] vel[i + j N] (oscillator strengths, vec_kernels.hip).
void launch_oscillator(const double2* X, int64_t N, int nstates, const double2* vel, double2* out, hipStream_t s);
// Band-data adjoint of the BSE assembly (band_adjoint.hip).
void band_gradients(Context& c, int nstates, const double2* d_vec, double* w_out, double* g_out);

}  // namespace xatu_dev
