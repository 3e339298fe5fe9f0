// C-ABI entry points (include/xatu_bse.h). Each call validates its inputs the way
// the reference does, runs on the context's stream and converts exceptions into
// status codes plus a per-context message.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>

#include "context.hpp"
#include "mfma.hpp"

using namespace xatu_dev;

// Scratch of the no-allocation product (xatu_bse_matvec_noalloc): staged vectors and the
// partial sums of the column splits. Every GPU the library runs on holds its own copy at
// its own address until the process ends; host code reaches it only through the address
// hipGetSymbolAddress returns for the current device (set up by xatu_create).
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
constexpr int64_t kNoAllocScratchElems = static_cast<int64_t>(kNoAllocScratchBytes / sizeof(double2));
static __device__ __align__(256) double2 scratch[kNoAllocScratchElems];

struct xatu_ctx {
    Context c;
};

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2* g_scratch_ptr[kMaxDevices] = {};

// Device address of the no-allocation scratch on the current device, looked up once per
// device (the device must be current).
void setup_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(XATU_ERR_INVALID_ARGUMENT, "device index too large");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return;
    void* p = nullptr;
    XATU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) throw Error(XATU_ERR_DEVICE, "no device address for the no-allocation scratch");
    g_scratch_ptr[device] = static_cast<double2*>(p);
}

double2* scratch_address(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    double2* p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
    if (!p) throw Error(XATU_ERR_LOGIC, "no-allocation scratch not set up for this device");
    return p;
}

template <typename F>
int guarded(xatu_ctx* ctx, F&& f) {
    if (!ctx) return XATU_ERR_INVALID_ARGUMENT;
    try {
        XATU_HIP(hipSetDevice(ctx->c.device));
        f(ctx->c);
        ctx->c.err.clear();
        return XATU_OK;
    } catch (const Error& e) {
        ctx->c.err = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->c.err = "host allocation failed";
        return XATU_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        ctx->c.err = e.what();
        return XATU_ERR_DEVICE;
    }
}

// Centred supercell combinations within a sphere (Lattice::generateCombinations with
// centered=true, Lattice::truncateSupercell and Lattice::truncateReciprocalSupercell,
// src/Lattice.cpp:293-370): coefficients run over [0, n) - n/2 with the first one
// fastest; a combination m of the rows of vecs is kept if |sum_j m_j vecs_j| < radius + 1e-5.
std::vector<int> truncate_supercell(const double* vecs, int ndim, int ncell, double radius) {
    std::vector<int> out;
    const int shift = ncell / 2;
    long long total = 1;
    for (int j = 0; j < ndim; ++j) total *= ncell;
    for (long long i = 0; i < total; ++i) {
        int m[3] = {0, 0, 0};
        long long rem = i;
        for (int j = 0; j < ndim; ++j) {
            m[j] = static_cast<int>(rem % ncell) - shift;
            rem /= ncell;
        }
        double v[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < ndim; ++j) {
            v[0] += m[j] * vecs[3 * j + 0];
            v[1] += m[j] * vecs[3 * j + 1];
            v[2] += m[j] * vecs[3 * j + 2];
        }
        const double d = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (d < radius + 1E-5) {
            for (int j = 0; j < ndim; ++j) out.push_back(m[j]);
        }
    }
    return out;
}

// Reciprocal lattice vectors (rows): b_i . a_j = 2 pi delta_ij within the first ndim
// Cartesian components (1D: b = 2 pi a / |a|^2), as Lattice::calculateReciprocalLattice
// (src/Lattice.cpp:86-113) for lattices in the x(y) plane.
void reciprocal_lattice(LatticeDesc& L) {
    for (double& x : L.recip_vec) x = 0.0;
    const double* a = L.bravais;
    const double two_pi = 2.0 * M_PI;
    if (L.ndim == 1) {
        const double n2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
        for (int x = 0; x < 3; ++x) L.recip_vec[x] = two_pi * a[x] / n2;
        return;
    }
    const int n = L.ndim;
    double A[3][3] = {}, inv[3][3] = {};
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) A[i][j] = a[3 * i + j];
    if (n == 2) {
        const double det = A[0][0] * A[1][1] - A[0][1] * A[1][0];
        if (std::fabs(det) < 1e-300) throw Error(XATU_ERR_INVALID_ARGUMENT, "singular Bravais basis");
        inv[0][0] = A[1][1] / det;
        inv[0][1] = -A[0][1] / det;
        inv[1][0] = -A[1][0] / det;
        inv[1][1] = A[0][0] / det;
    } else {
        const double det = A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) -
                           A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
                           A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
        if (std::fabs(det) < 1e-300) throw Error(XATU_ERR_INVALID_ARGUMENT, "singular Bravais basis");
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
                inv[i][j] = (A[i1][j1] * A[i2][j2] - A[i1][j2] * A[i2][j1]) / det;
            }
    }
    // A b_i = 2 pi e_i: b_i is column i of 2 pi A^-1
    for (int i = 0; i < n; ++i)
        for (int x = 0; x < n; ++x) L.recip_vec[3 * i + x] = two_pi * inv[x][i];
}

// Unit cell length / area / volume (Lattice::computeUnitCellArea, src/Lattice.cpp:119-131).
double unit_cell_measure(const LatticeDesc& L) {
    const double* a = L.bravais;
    if (L.ndim == 1) return std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    const double cx = a[1] * a[5] - a[2] * a[4];
    const double cy = a[2] * a[3] - a[0] * a[5];
    const double cz = a[0] * a[4] - a[1] * a[3];
    if (L.ndim == 2) return std::sqrt(cx * cx + cy * cy + cz * cz);
    return cx * a[6] + cy * a[7] + cz * a[8];
}

PotDesc make_pot(const Context& c, int kind) {
    if (kind != XATU_POTENTIAL_KELDYSH && kind != XATU_POTENTIAL_COULOMB) {
        throw Error(XATU_ERR_INVALID_ARGUMENT, "selectPotential(): potential must be either 'keldysh' or 'coulomb'");
    }
    PotDesc p{};
    p.kind = kind;
    p.eps_bar = (c.inter.eps_m + c.inter.eps_s) / 2;
    p.r0 = c.inter.r0;
    p.reg = (c.inter.regularization < 1E-10) ? c.a1norm : c.inter.regularization;
    p.cutoff = c.a1norm * c.inter.cutoff + 1E-5;
    return p;
}

// Interaction tables of the reciprocal-space mode: the sums over the reciprocal
// vectors G of the centred n_reciprocal^ndim combinations with |G| < cutoff |b1| + 1e-5
// (ExcitonTB::reciprocalInteractionTerm, src/ExcitonTB.cpp:562-583).
void reciprocal_tables(Context& c) {
    LatticeDesc& L = c.lat;
    const xatu_interaction& in = c.inter;
    if (in.n_reciprocal < 0) {
        throw Error(XATU_ERR_INVALID_ARGUMENT, "setReciprocalVectors(): given number must be positive");
    }
    long long nt = 1;
    for (int j = 0; j < L.ndim; ++j) nt *= 2LL * L.mesh_n - 1;
    if (nt > (1LL << 30)) throw Error(XATU_ERR_INVALID_ARGUMENT, "mesh too large for the reciprocal-space mode");
    L.recip = 1;
    L.tab_n = 2 * L.mesh_n - 1;
    L.ntab = static_cast<int>(nt);
    const double* b1 = L.recip_vec;
    const double b1n = std::sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]);
    const std::vector<int> m = truncate_supercell(L.recip_vec, L.ndim, in.n_reciprocal, in.cutoff * b1n);
    const int ng = static_cast<int>(m.size() / L.ndim);
    std::vector<double> g(3 * static_cast<size_t>(std::max(ng, 1)), 0.0);
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < L.ndim; ++j)
            for (int x = 0; x < 3; ++x) g[3 * i + x] += m[i * L.ndim + j] * L.recip_vec[3 * j + x];
    DevBuf<double> d_g;
    d_g.alloc(g.size());
    XATU_HIP(hipMemcpyAsync(d_g.p, g.data(), g.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    const int na2 = L.natoms * L.natoms;
    c.d_ft.alloc(static_cast<size_t>(na2) * L.ntab);
    c.d_ftq.alloc(na2);
    const double kEc = 1.6021766E-19, kEps0 = 8.8541878E-12;  // include/xatu/ExcitonTB.hpp:10-14
    const double eps_bar = (in.eps_m + in.eps_s) / 2;
    const double pref = kEc * 1E10 / (2 * kEps0 * eps_bar * unit_cell_measure(L) * in.total_cells);
    EventTimer ev(c.stream);
    ev.start();
    launch_reciprocal_table(c, d_g.p, ng, pref, c.stream);
    if (!in.exchange) XATU_HIP(hipMemsetAsync(c.d_ftq.p, 0, na2 * sizeof(double2), c.stream));
    ev.stop();
    c.tim.ms_fourier = ev.ms();  // synchronizes the stream before d_g is released
    c.nR = 0;
    c.have_ft = true;
    c.built64 = c.built32 = c.prepared = false;
}

// Back from a subset basis to the full one: the subset matrix is dropped.
void leave_subset(Context& c) {
    if (!c.subset) return;
    XATU_HIP(hipStreamSynchronize(c.stream));
    c.subset = false;
    c.N = static_cast<int64_t>(c.nk) * c.nv * c.nc;
    c.d_H64.release();
    c.built64 = c.built32 = c.prepared = false;
    c.drop_solution();
}

// Exchange factors and the FP64 diagonal (Davidson preconditioner), needed by
// every product whether or not a matrix is resident.
void prepare_operator(Context& c) {
    if (c.inter.exchange) {
        const size_t ne = static_cast<size_t>(c.nk) * c.lat.natoms * c.nc * c.nv;
        c.d_e1.alloc(ne);
        c.d_e2.alloc(ne);
        launch_exchange_factors(c, c.stream);
    } else {
        c.d_e1.alloc(1);
        c.d_e2.alloc(1);
    }
    c.d_diag.alloc(c.N);
    launch_diag_direct(c, c.d_diag.p, c.stream);
    c.h_diag.resize(c.N);
    XATU_HIP(hipMemcpyAsync(c.h_diag.data(), c.d_diag.p, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    XATU_HIP(hipStreamSynchronize(c.stream));
    split_product_scale(c);
    c.prepared = true;
    ++c.prep_gen;
}

}  // namespace

namespace xatu_dev {

// Product with the resident matrix of the given precision (or matrix-free), launched
// on the context stream without waiting.
void launch_operator(Context& c, int precision, int matrix_free, const double2* X, double2* Y, int nvec) {
    if (matrix_free) {
        if (!c.prepared) throw Error(XATU_ERR_LOGIC, "operator not prepared (call xatu_build_bse)");
        // the convolution form needs a table periodic on the mesh: the reciprocal-space
        // mode runs its matrix-free products on recomputed tiles
        if (matrix_free == XATU_MATRIX_FREE_TILES || c.lat.recip) {
            if (precision == XATU_FP64) launch_matfree64(c, X, Y, nvec, c.stream);
            else launch_matfree32(c, X, Y, nvec, c.stream);
        } else {
            if (precision == XATU_FP64) launch_conv64(c, X, Y, nvec, c.stream);
            else launch_conv32(c, X, Y, nvec, c.stream);
        }
    } else if (precision == XATU_FP64) {
        if (!c.built64) throw Error(XATU_ERR_LOGIC, "FP64 BSE matrix is not resident");
        launch_matvec64(c, X, Y, nvec, c.stream);
    } else {
        if (!c.built32) throw Error(XATU_ERR_LOGIC, "FP32 BSE matrix is not resident");
        launch_matvec32(c, X, Y, nvec, c.stream);
    }
}

// Product statistics (dense products only).
void record_product(Context& c, int matrix_free, double ms, int nvec) {
    if (matrix_free) return;
    c.tim.ms_matvec_kernel += ms;
    c.tim.n_matvec_kernel += 1;
    c.tim.n_matvec_vectors += nvec;
}

// Timed product (waits for completion).
void apply_operator(Context& c, int precision, int matrix_free, const double2* X, double2* Y, int nvec,
                    double* ms_kernel) {
    EventTimer ev(c.stream);
    ev.start();
    launch_operator(c, precision, matrix_free, X, Y, nvec);
    ev.stop();
    const double ms = ev.ms();
    if (ms_kernel) *ms_kernel = ms;
    record_product(c, matrix_free, ms, nvec);
}

}  // namespace xatu_dev

extern "C" {

int xatu_create(xatu_ctx** out, int device) {
    if (!out) return XATU_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    xatu_ctx* ctx = new (std::nothrow) xatu_ctx();
    if (!ctx) return XATU_ERR_OUT_OF_MEMORY;
    ctx->c.device = device;
    int rc = guarded(ctx, [&](Context& c) {
        int ndev = 0;
        XATU_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(XATU_ERR_INVALID_ARGUMENT, "no such HIP device");
        XATU_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        XATU_BLAS(rocblas_create_handle(&c.blas));
        XATU_BLAS(rocblas_set_stream(c.blas, c.stream));
        setup_scratch(device);  // guarded() made the device current
    });
    if (rc != XATU_OK) {
        xatu_destroy(ctx);
        return rc;
    }
    *out = ctx;
    return XATU_OK;
}

void xatu_destroy(xatu_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    if (ctx->c.blas) rocblas_destroy_handle(ctx->c.blas);
    hipStream_t s = ctx->c.stream;
    delete ctx;  // frees device buffers
    if (s) (void)hipStreamDestroy(s);
}

const char* xatu_last_error(const xatu_ctx* ctx) {
    if (!ctx) return "null context";
    return ctx->c.err.c_str();
}

int xatu_set_lattice(xatu_ctx* ctx, int ndim, const double* bravais, int natoms, const double* motif,
                     const int* norb_per_atom, int mesh_n) {
    return guarded(ctx, [&](Context& c) {
        if (ndim < 1 || ndim > 3) throw Error(XATU_ERR_INVALID_ARGUMENT, "ndim must be 1, 2 or 3");
        if (natoms < 1 || natoms > kMaxAtoms) throw Error(XATU_ERR_INVALID_ARGUMENT, "natoms must be in [1, 16]");
        if (!bravais || !motif || !norb_per_atom) throw Error(XATU_ERR_INVALID_ARGUMENT, "null pointer");
        if (mesh_n < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "mesh_n must be positive");
        LatticeDesc L{};
        L.ndim = ndim;
        L.natoms = natoms;
        L.mesh_n = mesh_n;
        long long nm = 1;
        for (int j = 0; j < ndim; ++j) nm *= mesh_n;
        if (nm > (1LL << 30)) throw Error(XATU_ERR_INVALID_ARGUMENT, "mesh too large");
        L.nmesh = static_cast<int>(nm);
        for (int i = 0; i < 3 * ndim; ++i) L.bravais[i] = bravais[i];
        for (int i = 0; i < 3 * natoms; ++i) L.tau[i] = motif[i];
        L.orb_begin[0] = 0;
        for (int a = 0; a < natoms; ++a) {
            if (norb_per_atom[a] < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "every atom needs >= 1 orbital");
            L.orb_begin[a + 1] = L.orb_begin[a] + norb_per_atom[a];
        }
        L.norb = L.orb_begin[natoms];
        reciprocal_lattice(L);
        L.recip = 0;
        L.tab_n = mesh_n;
        L.ntab = L.nmesh;
        c.lat = L;
        c.a1norm = std::sqrt(bravais[0] * bravais[0] + bravais[1] * bravais[1] + bravais[2] * bravais[2]);
        c.have_lattice = true;
        c.have_bands = c.have_ft = c.built64 = c.built32 = c.prepared = false;
        c.drop_solution();
    });
}

int xatu_set_bands(xatu_ctx* ctx, int nk, int nv, int nc, const double* eigval_k, const double* eigval_kq,
                   const double* eigvec_k, const double* eigvec_kq, const int* k_offsets) {
    return guarded(ctx, [&](Context& c) {
        c.require_lattice();
        if (nv < 1 || nc < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "Error: Exciton object must have some bands");
        if (nk < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "Error: BZ mesh must be initialized first");
        if (!eigval_k || !eigvec_k || !k_offsets) throw Error(XATU_ERR_INVALID_ARGUMENT, "null pointer");
        if (!eigval_kq) eigval_kq = eigval_k;
        if (!eigvec_kq) eigvec_kq = eigvec_k;
        const int nb = nv + nc;
        if (nb > c.lat.norb) {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "Error: Number of bands cannot be higher than actual material bands");
        }
        c.subset = false;
        c.nk = nk;
        c.nv = nv;
        c.nc = nc;
        c.nb = nb;
        c.N = static_cast<int64_t>(nk) * nv * nc;
        const size_t ne = static_cast<size_t>(nb) * nk;
        const size_t nvec = static_cast<size_t>(c.lat.norb) * nb * nk;
        c.h_eval_k.assign(eigval_k, eigval_k + ne);
        c.h_eval_kq.assign(eigval_kq, eigval_kq + ne);
        c.d_eval_k.alloc(ne);
        c.d_eval_kq.alloc(ne);
        c.d_evec_k.alloc(nvec);
        c.d_evec_kq.alloc(nvec);
        XATU_HIP(hipMemcpyAsync(c.d_eval_k.p, eigval_k, ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
        XATU_HIP(hipMemcpyAsync(c.d_eval_kq.p, eigval_kq, ne * sizeof(double), hipMemcpyHostToDevice, c.stream));
        XATU_HIP(hipMemcpyAsync(c.d_evec_k.p, eigvec_k, nvec * sizeof(double2), hipMemcpyHostToDevice, c.stream));
        XATU_HIP(hipMemcpyAsync(c.d_evec_kq.p, eigvec_kq, nvec * sizeof(double2), hipMemcpyHostToDevice, c.stream));
        std::vector<int> coord(static_cast<size_t>(nk) * 3, 0);
        c.h_koff.assign(k_offsets, k_offsets + static_cast<size_t>(nk) * c.lat.ndim);
        for (int k = 0; k < nk; ++k)
            for (int j = 0; j < c.lat.ndim; ++j) coord[3 * k + j] = k_offsets[k * c.lat.ndim + j];
        c.d_kcoord.alloc(coord.size());
        XATU_HIP(hipMemcpyAsync(c.d_kcoord.p, coord.data(), coord.size() * sizeof(int), hipMemcpyHostToDevice,
                                c.stream));
        XATU_HIP(hipStreamSynchronize(c.stream));
        c.have_bands = true;
        c.built64 = c.built32 = c.prepared = false;
        c.drop_solution();
    });
}

int xatu_fourier_sums(xatu_ctx* ctx, const xatu_interaction* inter) {
    return guarded(ctx, [&](Context& c) {
        c.require_lattice();
        if (!inter) throw Error(XATU_ERR_INVALID_ARGUMENT, "null interaction");
        if (inter->r0 == 0) throw Error(XATU_ERR_INVALID_ARGUMENT, "Error: r0 must be non-zero");
        if (inter->ncell < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "ncell must be a positive number");
        if (inter->total_cells < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "total_cells must be positive");
        if (inter->mode != XATU_MODE_REALSPACE && inter->mode != XATU_MODE_RECIPROCAL) {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "setMode(): mode must be either realspace or reciprocalspace");
        }
        leave_subset(c);
        c.drop_solution();
        c.have_ft = false;
        c.built64 = c.built32 = c.prepared = false;
        c.inter = *inter;
        c.pot_d = make_pot(c, inter->potential);
        c.pot_x = inter->exchange ? make_pot(c, inter->exchange_potential) : c.pot_d;
        if (inter->mode == XATU_MODE_RECIPROCAL) {
            reciprocal_tables(c);
            return;
        }
        c.lat.recip = 0;
        c.lat.tab_n = c.lat.mesh_n;
        c.lat.ntab = c.lat.nmesh;
        c.h_cells = truncate_supercell(c.lat.bravais, c.lat.ndim, inter->ncell, c.a1norm * inter->cutoff);
        c.nR = static_cast<int>(c.h_cells.size() / c.lat.ndim);
        if (c.nR == 0) throw Error(XATU_ERR_INVALID_ARGUMENT, "no lattice vectors inside the cutoff");
        std::vector<int> padded(static_cast<size_t>(c.nR) * 3, 0);
        for (int r = 0; r < c.nR; ++r)
            for (int j = 0; j < c.lat.ndim; ++j) padded[3 * r + j] = c.h_cells[r * c.lat.ndim + j];
        c.d_cells.alloc(padded.size());
        XATU_HIP(hipMemcpyAsync(c.d_cells.p, padded.data(), padded.size() * sizeof(int), hipMemcpyHostToDevice,
                                c.stream));
        const int npairs = c.lat.natoms * (c.lat.natoms + 1) / 2;
        const int na2 = c.lat.natoms * c.lat.natoms;
        c.d_vtab.alloc(static_cast<size_t>(npairs) * c.nR);
        c.d_twiddle.alloc(2 * c.lat.mesh_n);
        c.d_ft.alloc(static_cast<size_t>(na2) * c.lat.nmesh);
        c.d_ftq.alloc(na2);
        EventTimer ev(c.stream);
        ev.start();
        launch_potential_table(c, c.pot_d, c.d_vtab.p, c.stream);
        launch_twiddles(c.lat.mesh_n, c.d_twiddle.p, c.stream);
        launch_fourier_mesh(c, c.stream);
        if (inter->exchange) {
            c.d_vtabx.alloc(static_cast<size_t>(npairs) * c.nR);
            launch_potential_table(c, c.pot_x, c.d_vtabx.p, c.stream);
            launch_fourier_q(c, c.stream);
        } else {
            XATU_HIP(hipMemsetAsync(c.d_ftq.p, 0, na2 * sizeof(double2), c.stream));
        }
        ev.stop();
        c.tim.ms_fourier = ev.ms();
        c.have_ft = true;
        c.built64 = c.built32 = c.prepared = false;
    });
}

int xatu_build_bse(xatu_ctx* ctx, int precision, double scissor, void* hbs_out) {
    return guarded(ctx, [&](Context& c) {
        c.require_lattice();
        c.require_bands();
        c.require_ft();
        if (precision != XATU_FP64 && precision != XATU_FP32 && precision != XATU_MATRIX_FREE) {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "bad precision");
        }
        if (precision == XATU_MATRIX_FREE && hbs_out) {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "a matrix-free build has no matrix to copy out");
        }
        leave_subset(c);
        if (c.prepared && c.scissor != scissor) {
            // a resident matrix built with another scissor no longer matches the operator
            c.built64 = c.built32 = false;
        }
        c.scissor = scissor;
        c.drop_solution();
        const size_t nn = static_cast<size_t>(c.N) * static_cast<size_t>(c.N);
        EventTimer ev(c.stream), evk(c.stream);
        ev.start();
        prepare_operator(c);
        evk.start();
        if (precision == XATU_FP64) {
            c.built64 = false;
            c.d_H64.alloc(nn);
            launch_build_tiles64(c, c.d_H64.p, c.stream);
        } else if (precision == XATU_FP32) {
            c.built32 = false;
            // 32 zero entries past the end: the split product reads up to 31 entries past
            // the last column (multiplied by zero vector padding)
            c.d_H32.alloc(nn + 32);
            XATU_HIP(hipMemsetAsync(c.d_H32.p + nn, 0, 32 * sizeof(float2), c.stream));
            // only the block-lower part when the split product will read it (and nothing is
            // copied out); completed on demand for the FP32 matrix-core product and copies
            c.h32_lower = !hbs_out && split_layout_available(c);
            // ... and in the split FP16 layout the product on the FP16 matrix cores loads
            c.h32_presplit = c.h32_lower && presplit_build_available(c);
            if (c.h32_lower) clear_upper_heads(c);
            launch_build_tiles32(c, c.d_H32.p, c.h32_lower, c.h32_presplit, c.stream);
        }
        evk.stop();
        ev.stop();
        c.tim.ms_build = ev.ms();
        c.tim.ms_build_kernel = evk.ms();
        if (precision == XATU_FP64) c.built64 = true;
        if (precision == XATU_FP32) c.built32 = true;
        if (hbs_out) {
            const size_t bytes = nn * (precision == XATU_FP64 ? sizeof(double2) : sizeof(float2));
            const void* src = precision == XATU_FP64 ? static_cast<const void*>(c.d_H64.p)
                                                     : static_cast<const void*>(c.d_H32.p);
            XATU_HIP(hipMemcpyAsync(hbs_out, src, bytes, hipMemcpyDeviceToHost, c.stream));
            XATU_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int xatu_build_bse_basis(xatu_ctx* ctx, int nbasis, const int* basis, double scissor, double* hbs_out) {
    return guarded(ctx, [&](Context& c) {
        c.require_lattice();
        c.require_bands();
        c.require_ft();
        if (nbasis < 1 || !basis) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad basis");
        for (int i = 0; i < nbasis; ++i) {
            const int v = basis[3 * i], cc = basis[3 * i + 1], k = basis[3 * i + 2];
            if (v < 0 || v >= c.nb || cc < 0 || cc >= c.nb || k < 0 || k >= c.nk) {
                throw Error(XATU_ERR_INVALID_ARGUMENT, "basis entry out of range");
            }
        }
        // the exchange factors of the full basis are not used: the element kernel forms them
        XATU_HIP(hipStreamSynchronize(c.stream));
        c.built64 = c.built32 = c.prepared = false;
        c.drop_solution();
        c.d_H32.release();
        DevBuf<int> d_basis;
        d_basis.alloc(static_cast<size_t>(nbasis) * 3);
        XATU_HIP(hipMemcpyAsync(d_basis.p, basis, d_basis.bytes(), hipMemcpyHostToDevice, c.stream));
        c.d_H64.alloc(static_cast<size_t>(nbasis) * nbasis);
        launch_build_basis(c, nbasis, d_basis.p, scissor, c.d_H64.p, c.stream);
        // the subset matrix is now the operator: dense and FP64 Davidson solves run on it
        c.subset = true;
        c.N = nbasis;
        c.d_diag.alloc(c.N);
        launch_diag_extract(c, c.d_H64.p, c.d_diag.p, c.stream);
        c.h_diag.resize(c.N);
        XATU_HIP(hipMemcpyAsync(c.h_diag.data(), c.d_diag.p, c.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        if (hbs_out) {
            XATU_HIP(hipMemcpyAsync(hbs_out, c.d_H64.p, c.d_H64.bytes(), hipMemcpyDeviceToHost, c.stream));
        }
        XATU_HIP(hipStreamSynchronize(c.stream));
        c.built64 = true;
        c.prepared = true;
        ++c.prep_gen;
    });
}

int xatu_get_bse(xatu_ctx* ctx, int precision, void* hbs_out) {
    return guarded(ctx, [&](Context& c) {
        if (!hbs_out) throw Error(XATU_ERR_INVALID_ARGUMENT, "null output");
        const size_t nn = static_cast<size_t>(c.N) * static_cast<size_t>(c.N);
        if (precision == XATU_FP64) {
            if (!c.built64) throw Error(XATU_ERR_LOGIC, "FP64 BSE matrix is not resident");
            XATU_HIP(hipMemcpyAsync(hbs_out, c.d_H64.p, nn * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
            XATU_HIP(hipStreamSynchronize(c.stream));
        } else {
            if (!c.built32) throw Error(XATU_ERR_LOGIC, "FP32 BSE matrix is not resident");
            complete_matrix32(c);  // launched on c.stream: the copy must follow it there
            XATU_HIP(hipMemcpyAsync(hbs_out, c.d_H32.p, nn * sizeof(float2), hipMemcpyDeviceToHost, c.stream));
            XATU_HIP(hipStreamSynchronize(c.stream));
        }
    });
}

int xatu_free_bse(xatu_ctx* ctx, int precision) {
    return guarded(ctx, [&](Context& c) {
        XATU_HIP(hipStreamSynchronize(c.stream));
        if (precision == XATU_FP64) {
            c.d_H64.release();
            c.built64 = false;
        } else if (precision == XATU_FP32) {
            c.d_H32.release();
            c.built32 = false;
        } else {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "bad precision");
        }
    });
}

int xatu_solve(xatu_ctx* ctx, const xatu_solve_options* opt, double* eigval_out, double* eigvec_out,
               xatu_solve_report* report) {
    return guarded(ctx, [&](Context& c) {
        if (!opt || !eigval_out) throw Error(XATU_ERR_INVALID_ARGUMENT, "null argument");
        if (opt->fp32_arith != XATU_ARITH_SPLIT16 && opt->fp32_arith != XATU_ARITH_FP32) {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "fp32_arith must be XATU_ARITH_SPLIT16 or XATU_ARITH_FP32");
        }
        if (c.subset && opt->method == XATU_SOLVE_DAVIDSON && (opt->precision != XATU_FP64 || opt->matrix_free)) {
            throw Error(XATU_ERR_LOGIC, "a subset-basis Hamiltonian is solved with its resident FP64 matrix");
        }
        xatu_solve_report rep{};
        c.tim.ms_matvec_kernel = 0.0;
        c.tim.n_matvec_kernel = 0;
        c.tim.n_matvec_vectors = 0;
        if (opt->method == XATU_SOLVE_DIAG) {
            solve_dense_diag(c, opt->nstates, eigval_out, eigvec_out);
            rep.converged = 1;
        } else if (opt->method == XATU_SOLVE_DAVIDSON) {
            if (!opt->matrix_free) {
                if (opt->precision == XATU_FP64 && !c.built64) {
                    throw Error(XATU_ERR_LOGIC, "diagonalizeRaw(): BSE Hamiltonian is not initialized.");
                }
                if (opt->precision == XATU_FP32 && !c.built32) {
                    throw Error(XATU_ERR_LOGIC, "diagonalizeRaw(): BSE Hamiltonian is not initialized.");
                }
            }
            if (!c.prepared) throw Error(XATU_ERR_LOGIC, "diagonalizeRaw(): BSE Hamiltonian is not initialized.");
            solve_davidson(c, *opt, eigval_out, eigvec_out, &rep);
        } else {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "unknown solve method");
        }
        if (c.tim.n_matvec_kernel > 0) c.tim.ms_matvec_kernel /= c.tim.n_matvec_kernel;
        if (report) *report = rep;
    });
}

int xatu_bse_matvec(xatu_ctx* ctx, int precision, int matrix_free, int nvec, const double* x, double* y) {
    return guarded(ctx, [&](Context& c) {
        if (nvec < 1 || !x || !y) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad vectors");
        c.require_bands();
        if (precision != XATU_FP64 || matrix_free) c.require_full_basis("this product");
        const size_t n = static_cast<size_t>(c.N) * nvec;
        DevBuf<double2> X, Y;
        X.alloc(n);
        Y.alloc(n);
        XATU_HIP(hipMemcpyAsync(X.p, x, n * sizeof(double2), hipMemcpyHostToDevice, c.stream));
        apply_operator(c, precision, matrix_free, X.p, Y.p, nvec, nullptr);
        XATU_HIP(hipMemcpyAsync(y, Y.p, n * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
        XATU_HIP(hipStreamSynchronize(c.stream));
    });
}

int xatu_bse_matvec_noalloc(xatu_ctx* ctx, int nvec, const double* x, double* y) {
    return guarded(ctx, [&](Context& c) {
        if (nvec < 1 || !x || !y) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad vectors");
        c.require_bands();
        // the staged vectors and at least one split of partial sums must fit
        if (c.N > kNoAllocScratchElems / (2 * kBlockVectors)) {
            throw Error(XATU_ERR_INVALID_ARGUMENT,
                        "BSE dimension " + std::to_string(c.N) + " exceeds the no-allocation product's scratch (N <= " +
                            std::to_string(kNoAllocScratchElems / (2 * kBlockVectors)) + ")");
        }
        if (!c.built64) throw Error(XATU_ERR_LOGIC, "FP64 BSE matrix is not resident");
        launch_matvec64_scratch(c, reinterpret_cast<const double2*>(x), reinterpret_cast<double2*>(y), nvec,
                                scratch_address(c.device), kNoAllocScratchElems, c.stream);
        XATU_HIP(hipStreamSynchronize(c.stream));
    });
}

int xatu_hf_gradients(xatu_ctx* ctx, int nstates, const double* eigvec, int nparams, const int* params,
                      double* grad_out) {
    return guarded(ctx, [&](Context& c) {
        c.require_ft();
        c.require_full_basis("xatu_hf_gradients");
        if (c.lat.recip) {
            throw Error(XATU_ERR_LOGIC, "interaction-parameter gradients are available in the real-space mode only");
        }
        if (nstates < 1 || nparams < 1 || !params || !grad_out) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad args");
        for (int i = 0; i < nparams; ++i) {
            if (params[i] < XATU_PARAM_R0 || params[i] > XATU_PARAM_EPS_S) {
                throw Error(XATU_ERR_INVALID_ARGUMENT, "unknown gradient parameter");
            }
        }
        DevBuf<double2> V;
        const double2* dv;
        if (eigvec) {
            V.alloc(static_cast<size_t>(c.N) * nstates);
            XATU_HIP(hipMemcpyAsync(V.p, eigvec, V.bytes(), hipMemcpyHostToDevice, c.stream));
            dv = V.p;
        } else {
            dv = c.last_vectors(nstates);
        }
        hf_gradients(c, nstates, dv, nparams, params, grad_out);
    });
}

int xatu_band_gradients(xatu_ctx* ctx, int nstates, const double* eigvec, double* w_out, double* g_out) {
    return guarded(ctx, [&](Context& c) {
        c.require_ft();
        c.require_full_basis("xatu_band_gradients");
        if (nstates < 1 || !w_out || !g_out) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad args");
        DevBuf<double2> V;
        const double2* dv;
        if (eigvec) {
            V.alloc(static_cast<size_t>(c.N) * nstates);
            XATU_HIP(hipMemcpyAsync(V.p, eigvec, V.bytes(), hipMemcpyHostToDevice, c.stream));
            dv = V.p;
        } else {
            dv = c.last_vectors(nstates);
        }
        band_gradients(c, nstates, dv, w_out, g_out);
    });
}

int xatu_oscillator_strengths(xatu_ctx* ctx, int nstates, const double* eigvec, const double* velocity,
                              double* out) {
    return guarded(ctx, [&](Context& c) {
        c.require_bands();
        c.require_full_basis("xatu_oscillator_strengths");
        if (nstates < 1 || !velocity || !out) throw Error(XATU_ERR_INVALID_ARGUMENT, "bad args");
        DevBuf<double2> V, vel, o;
        const double2* dv;
        if (eigvec) {
            V.alloc(static_cast<size_t>(c.N) * nstates);
            XATU_HIP(hipMemcpyAsync(V.p, eigvec, V.bytes(), hipMemcpyHostToDevice, c.stream));
            dv = V.p;
        } else {
            dv = c.last_vectors(nstates);
        }
        vel.alloc(static_cast<size_t>(c.N) * 3);
        o.alloc(static_cast<size_t>(3) * nstates);
        XATU_HIP(hipMemcpyAsync(vel.p, velocity, vel.bytes(), hipMemcpyHostToDevice, c.stream));
        launch_oscillator(dv, c.N, nstates, vel.p, o.p, c.stream);
        XATU_HIP(hipMemcpyAsync(out, o.p, o.bytes(), hipMemcpyDeviceToHost, c.stream));
        XATU_HIP(hipStreamSynchronize(c.stream));
    });
}

int xatu_get_fourier(xatu_ctx* ctx, double* ft_out, double* ftq_out) {
    return guarded(ctx, [&](Context& c) {
        c.require_ft();
        if (c.lat.recip) throw Error(XATU_ERR_LOGIC, "the reciprocal-space mode has no lattice Fourier sums");
        const size_t na2 = static_cast<size_t>(c.lat.natoms) * c.lat.natoms;
        if (ft_out) {
            XATU_HIP(hipMemcpyAsync(ft_out, c.d_ft.p, na2 * c.lat.nmesh * sizeof(double2), hipMemcpyDeviceToHost,
                                    c.stream));
        }
        if (ftq_out) XATU_HIP(hipMemcpyAsync(ftq_out, c.d_ftq.p, na2 * sizeof(double2), hipMemcpyDeviceToHost, c.stream));
        XATU_HIP(hipStreamSynchronize(c.stream));
    });
}

int xatu_get_cells(xatu_ctx* ctx, int* n_cells, int* coefs) {
    return guarded(ctx, [&](Context& c) {
        c.require_ft();
        if (n_cells) *n_cells = c.nR;
        if (coefs) std::memcpy(coefs, c.h_cells.data(), c.h_cells.size() * sizeof(int));
    });
}

int xatu_compare_builds(xatu_ctx* ctx, int kblock, double* block_err, double* max_abs) {
    return guarded(ctx, [&](Context& c) {
        c.require_full_basis("xatu_compare_builds");
        if (!c.built64 || !c.built32) throw Error(XATU_ERR_LOGIC, "both FP64 and FP32 matrices must be resident");
        if (kblock < 1) throw Error(XATU_ERR_INVALID_ARGUMENT, "kblock must be positive");
        complete_matrix32(c);
        const int rows = kblock * c.nc * c.nv;
        const int nblk = static_cast<int>((c.N + rows - 1) / rows);
        DevBuf<double> d;
        d.alloc(static_cast<size_t>(nblk) * nblk);
        launch_compare(c, kblock, d.p, c.stream);
        std::vector<double> h(d.n);
        XATU_HIP(hipMemcpyAsync(h.data(), d.p, d.bytes(), hipMemcpyDeviceToHost, c.stream));
        XATU_HIP(hipStreamSynchronize(c.stream));
        if (block_err) std::memcpy(block_err, h.data(), d.bytes());
        if (max_abs) *max_abs = *std::max_element(h.begin(), h.end());
    });
}

int xatu_get_timings(const xatu_ctx* ctx, xatu_timings* out) {
    if (!ctx || !out) return XATU_ERR_INVALID_ARGUMENT;
    *out = ctx->c.tim;
    return XATU_OK;
}

}  // extern "C"
