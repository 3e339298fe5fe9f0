// Eigensolvers for the resident (or matrix-free) BSE operator.
//
// Dense path ("diag", reference src/ExcitonTB.cpp:940-943 -> LAPACK zheevd):
// rocSOLVER zheevd on a device copy of the FP64 matrix, all eigenpairs ascending.
//
// Block Davidson ("davidson", reference src/davidson.cpp:5-61, redesigned):
// the basis V, its image W = H V, Ritz vectors and residuals live in HBM in FP64;
// products H V use the FP64 or the FP32 matrix (or the matrix-free product);
// the projected matrix V^H W is formed incrementally (Gram kernel, FP64) and
// diagonalised on the device (Rayleigh-Ritz, projected_eigen.hip; on the host for
// bases above 128 or blocks above 64); corrections use the diagonal
// preconditioner r / (theta - diag H), are orthogonalised twice against V
// (classical Gram-Schmidt, FP64) and orthonormalised by an eigen-decomposition
// of their Gram matrix, dropping dependent directions; the block products with the
// small coefficient matrices run on a tall-by-small kernel. The basis restarts from
// the current Ritz block when it would exceed max_basis.
// Unlike the reference, the guess is the unit vectors on the smallest diagonal
// entries (the reference uses the first 2*nstates unit vectors, :19), no dense
// (M - theta I) is ever formed (:37), and convergence is on residual norms
// rather than on the change of the eigenvalues (:43).
// With the device Rayleigh-Ritz an iteration synchronises with the host once: the
// residual norms, the Ritz values and the correction block's Gram matrix arrive together
// (the corrections are formed before the convergence test; the block is whitened on the
// host, a 32 x 32 eigenproblem).
// The FP32 form runs this loop with FP32 products to a loose tolerance and then
// refines in FP64: it restarts from the FP32 Ritz block and iterates with FP64
// products until the FP64 residuals meet the tolerance.
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <cmath>
#include <complex>
#include <numeric>

#include "../common/heev.hpp"
#include "context.hpp"

namespace xatu_dev {

void launch_residual(const double2* X, const double2* AX, const double* theta, int64_t N, int nvec, double2* R,
                     double* partial, int* nblocks, hipStream_t s);
void launch_precondition(const double2* R, const double* theta, const double* diag, int64_t N, int nvec, double2* T,
                         hipStream_t s);
void launch_unit_vectors(const int* idx, int64_t N, int nvec, double2* V, hipStream_t s);
void launch_copy_vectors(const double2* src, int64_t n, double2* dst, hipStream_t s);
void launch_gather3(const double* a, int na, const double* b, int nb, const double* c, int nc, double* dst,
                    hipStream_t s);
int projected_eigen_max_dim();
int projected_eigen_max_want();
// fp32_reduction: the Householder reduction in FP32 (the FP32 phase of the mixed form)
void launch_projected_eigen(const double2* T, int ldt, int m, int nwant, double2* refl, double* theta, double2* Y,
                            int ldy, bool fp32_reduction, hipStream_t s);
void launch_extend_projected(const double2* P, int m0, int nb, double2* T, int ldt, hipStream_t s);
void launch_restart_projected(double2* T, int ld, const double* theta, int nb, hipStream_t s);

namespace {

using cd = std::complex<double>;


inline rocblas_double_complex* bz(double2* p) { return reinterpret_cast<rocblas_double_complex*>(p); }
inline const rocblas_double_complex* bz(const double2* p) {
    return reinterpret_cast<const rocblas_double_complex*>(p);
}

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

using Workspace = SolverWorkspace;

class Davidson {
public:
    Davidson(Context& c, Workspace& ws, int precision, int matrix_free, int nev, int block, int mmax,
             xatu_solve_report* rep)
        : c_(c), prec_(precision), mf_(matrix_free), nev_(nev), b_(block), mmax_(mmax), ws_(ws), rep_(rep) {
        N_ = c.N;
        ws_.V.reserve(static_cast<size_t>(N_) * mmax_);
        ws_.W.reserve(static_cast<size_t>(N_) * mmax_);
        ws_.X.reserve(static_cast<size_t>(N_) * b_);
        ws_.AX.reserve(static_cast<size_t>(N_) * b_);
        ws_.R.reserve(static_cast<size_t>(N_) * b_);
        ws_.Tb.reserve(static_cast<size_t>(N_) * b_);
        ws_.P.reserve(static_cast<size_t>(mmax_) * b_);
        ws_.Y.reserve(static_cast<size_t>(mmax_) * b_);
        ws_.theta.reserve(b_);
        ws_.partial.reserve(static_cast<size_t>(b_) * ((N_ + 255) / 256));
        T_.assign(static_cast<size_t>(mmax_) * mmax_, cd(0.0, 0.0));
        c_.h_proj.reserve(static_cast<size_t>(mmax_) * b_);
        c_.h_coef.reserve(static_cast<size_t>(mmax_) * b_);
        c_.h_whiten.reserve_mapped(static_cast<size_t>(b_) * b_);
        c_.h_theta.reserve(b_);
        c_.h_part.reserve(static_cast<size_t>(b_) * ((N_ + 255) / 256));
        c_.h_gather.reserve_mapped(2 * static_cast<size_t>(mmax_) * b_ + static_cast<size_t>(b_) * ((N_ + 255) / 256) + b_);
        dev_rr_ = mmax_ <= projected_eigen_max_dim() && b_ <= projected_eigen_max_want();
        if (dev_rr_) {
            const int md = projected_eigen_max_dim();
            ws_.Tm.reserve(static_cast<size_t>(mmax_) * mmax_);
            ws_.refl.reserve(static_cast<size_t>(md) * md);
            XATU_HIP(hipMemsetAsync(ws_.Tm.p, 0, ws_.Tm.bytes(), c_.stream));
        }
    }

    // Start from the unit vectors on the b smallest diagonal entries.
    void init_unit() {
        // the b smallest diagonal entries, ties by index (the order a stable sort gives)
        std::vector<int> idx(N_);
        std::iota(idx.begin(), idx.end(), 0);
        std::partial_sort(idx.begin(), idx.begin() + b_, idx.end(), [&](int a, int b) {
            return c_.h_diag[a] < c_.h_diag[b] || (c_.h_diag[a] == c_.h_diag[b] && a < b);
        });
        idx.resize(b_);
        DevBuf<int> d_idx;
        d_idx.alloc(b_);
        XATU_HIP(hipMemcpyAsync(d_idx.p, idx.data(), b_ * sizeof(int), hipMemcpyHostToDevice, c_.stream));
        XATU_HIP(hipMemsetAsync(ws_.V.p, 0, static_cast<size_t>(N_) * b_ * sizeof(double2), c_.stream));
        launch_unit_vectors(d_idx.p, N_, b_, ws_.V.p, c_.stream);
        m_ = 0;
        append_block(b_);
        XATU_HIP(hipStreamSynchronize(c_.stream));
    }

    // Start from b vectors given on the device (N x b), orthonormalised (they come from a
    // solve whose Rayleigh-Ritz step may have been FP32): V = X0 S.
    void init_from(const double2* X0) {
        XATU_HIP(hipMemcpyAsync(ws_.V.p, X0, static_cast<size_t>(N_) * b_ * sizeof(double2),
                                hipMemcpyDeviceToDevice, c_.stream));
        m_ = 0;
        const int kept = orthonormalise_slot(b_);
        if (kept < b_) throw Error(XATU_ERR_NOT_CONVERGED, "Davidson: the refinement's starting block is rank deficient");
        append_block(b_);
    }

    // Iterate to convergence; returns true if converged.
    bool run(double tol, int maxit) {
        if (dev_rr_) return run_device(tol, maxit);
        for (int it = 0; it < maxit; ++it) {
            rayleigh_ritz();
            ritz_and_residuals();
            rep_->iterations++;
            double worst = 0.0;
            for (int i = 0; i < nev_; ++i) worst = std::max(worst, res_[i]);
            last_worst_ = worst;
            if (worst < tol) return true;
            if (m_ + b_ > mmax_) restart();
            if (!add_corrections()) return false;  // no new direction: stagnated
        }
        return false;
    }

    void set_precision(int precision, int matrix_free) {
        prec_ = precision;
        mf_ = matrix_free;
    }

    double worst() const { return last_worst_; }
    const std::vector<double>& theta() const { return theta_; }
    const double2* ritz_vectors() const { return ws_.X.p; }
    double products_ms() const { return prod_ms_; }
    double rr_ms() const { return rr_ms_; }
    double projections_ms() const { return proj_ms_; }

private:
    // Event pairs read after the next synchronisation (reading them adds no wait).
    struct Pending {
        std::unique_ptr<EventTimer> ev;
        int kind;  // 0 product (nvec), 1 projection, 2 Rayleigh-Ritz
        int nvec;
    };
    EventTimer* timer(int kind, int nvec = 0) {
        pending_.push_back({std::make_unique<EventTimer>(c_.stream), kind, nvec});
        return pending_.back().ev.get();
    }
    // Reads the event pairs; with completed_only, only those already complete (the others
    // stay pending), so that it can run while the device works.
    void drain_timers(bool completed_only = false) {
        std::vector<Pending> later;
        for (Pending& p : pending_) {
            if (completed_only && hipEventQuery(p.ev->b) != hipSuccess) {
                later.push_back(std::move(p));
                continue;
            }
            const double ms = p.ev->ms();
            if (p.kind == 0) {
                prod_ms_ += ms;
                record_product(c_, mf_, ms, p.nvec);
            } else if (p.kind == 1) {
                proj_ms_ += ms;
            } else {
                rr_ms_ += ms;
            }
        }
        pending_ = std::move(later);
    }

    // One iteration per synchronisation: Rayleigh-Ritz on the device, Ritz block and
    // residuals, restart when the basis is full, preconditioned corrections and their
    // Gram matrix against the basis; then residual norms, Ritz values and the Gram matrix
    // to the host together.
    bool run_device(double tol, int maxit) {
        const int n = static_cast<int>(N_);
        for (int it = 0; it < maxit; ++it) {
            EventTimer* er = timer(2);
            er->start();
            launch_projected_eigen(ws_.Tm.p, mmax_, m_, b_, ws_.refl.p, ws_.theta.p, ws_.Y.p, m_, prec_ == XATU_FP32,
                                   c_.stream);
            er->stop();
            EventTimer* ep = timer(1);
            ep->start();
            launch_block_gemm(ws_.V.p, n, ws_.Y.p, m_, N_, m_, b_, 1.0, 0.0, ws_.X.p, n, c_.stream);
            launch_block_gemm(ws_.W.p, n, ws_.Y.p, m_, N_, m_, b_, 1.0, 0.0, ws_.AX.p, n, c_.stream);
            ep->stop();
            int nblk = 0;
            launch_residual(ws_.X.p, ws_.AX.p, ws_.theta.p, N_, b_, ws_.R.p, ws_.partial.p, &nblk, c_.stream);
            rep_->iterations++;
            if (m_ + b_ > mmax_) {
                launch_copy_vectors(ws_.X.p, static_cast<int64_t>(N_) * b_, ws_.V.p, c_.stream);
                launch_copy_vectors(ws_.AX.p, static_cast<int64_t>(N_) * b_, ws_.W.p, c_.stream);
                launch_restart_projected(ws_.Tm.p, mmax_, ws_.theta.p, b_, c_.stream);
                m_ = b_;
                rep_->restarts++;
            }
            const int nb = std::min(b_, mmax_ - m_);
            const int np = nb > 0 ? 2 * (m_ + nb) * nb : 0;  // doubles of the block's Gram rows
            if (nb > 0) {
                launch_precondition(ws_.R.p, ws_.theta.p, c_.d_diag.p, N_, nb, ws_.V.p + static_cast<size_t>(m_) * N_,
                                    c_.stream);
                ortho_gram(nb);
            }
            // Gram rows, residual partial sums and Ritz values stored into mapped host memory by
            // one kernel (three device-to-host copies each waited tens of microseconds here)
            launch_gather3(reinterpret_cast<const double*>(ws_.P.p), np, ws_.partial.p, nblk * b_, ws_.theta.p, b_,
                           c_.h_gather.dev, c_.stream);
            XATU_HIP(hipStreamSynchronize(c_.stream));
            const double* got = c_.h_gather.p;
            theta_.assign(got + np + nblk * b_, got + np + nblk * b_ + b_);
            residual_norms(nblk, got + np);
            double worst = 0.0;
            for (int i = 0; i < nev_; ++i) worst = std::max(worst, res_[i]);
            last_worst_ = worst;
            if (worst < tol || nb <= 0) {
                drain_timers();
                return worst < tol;
            }
            const int kept = ortho_whiten(nb, 0, reinterpret_cast<const cd*>(got));
            if (kept == 0) {
                drain_timers();
                return false;
            }
            append_block(kept);
            drain_timers(true);  // the timings of this iteration, while the next one runs
        }
        drain_timers();
        return false;
    }

    // part: the residual kernel's per-block partial sums on the host (nblk x b_)
    void residual_norms(int nblk, const double* part = nullptr) {
        if (!part) part = c_.h_part.p;
        res_.assign(b_, 0.0);
        for (int j = 0; j < b_; ++j) {
            double s = 0.0;
            for (int i = 0; i < nblk; ++i) s += part[i + static_cast<size_t>(j) * nblk];
            res_[j] = std::sqrt(s);
        }
    }

    // W[:, m:m+nb] = H V[:, m:m+nb], then extend the projected matrix.
    void append_block(int nb) {
        double2* Vn = ws_.V.p + static_cast<size_t>(m_) * N_;
        double2* Wn = ws_.W.p + static_cast<size_t>(m_) * N_;
        if (dev_rr_) {
            EventTimer* ev = timer(0, nb);
            ev->start();
            launch_operator(c_, prec_, mf_, Vn, Wn, nb);
            ev->stop();
            rep_->products += nb;
            const int mn = m_ + nb;
            EventTimer* eg = timer(1);
            eg->start();
            launch_gram(c_, ws_.V.p, static_cast<int>(N_), Wn, static_cast<int>(N_), mn, nb, ws_.P.p, mn, c_.stream);
            eg->stop();
            launch_extend_projected(ws_.P.p, m_, nb, ws_.Tm.p, mmax_, c_.stream);
            m_ = mn;
            return;
        }
        // event pairs are read after the synchronisation below, so that timing adds no waits
        EventTimer evp(c_.stream), evg(c_.stream);
        evp.start();
        launch_operator(c_, prec_, mf_, Vn, Wn, nb);
        evp.stop();
        rep_->products += nb;
        const int mn = m_ + nb;
        // P = V[:, 0:mn]^H W_new  (mn x nb)
        evg.start();
        launch_gram(c_, ws_.V.p, static_cast<int>(N_), Wn, static_cast<int>(N_), mn, nb, ws_.P.p, mn, c_.stream);
        evg.stop();
        const cd* P = reinterpret_cast<const cd*>(c_.h_proj.p);
        XATU_HIP(hipMemcpyAsync(c_.h_proj.p, ws_.P.p, static_cast<size_t>(mn) * nb * sizeof(cd), hipMemcpyDeviceToHost,
                                c_.stream));
        XATU_HIP(hipStreamSynchronize(c_.stream));
        const double pms = evp.ms();
        prod_ms_ += pms;
        record_product(c_, mf_, pms, nb);
        proj_ms_ += evg.ms();
        // rows of the existing basis
        for (int j = 0; j < nb; ++j) {
            for (int i = 0; i < m_; ++i) {
                const cd v = P[i + static_cast<size_t>(j) * mn];
                T_[i + static_cast<size_t>(m_ + j) * mmax_] = v;
                T_[(m_ + j) + static_cast<size_t>(i) * mmax_] = std::conj(v);
            }
        }
        // new diagonal block, symmetrised: (P[m+i, j] + conj(P[m+j, i])) / 2
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i <= j; ++i) {
                const cd pij = P[(m_ + i) + static_cast<size_t>(j) * mn];
                const cd pji = P[(m_ + j) + static_cast<size_t>(i) * mn];
                cd avg = 0.5 * (pij + std::conj(pji));
                if (i == j) avg = cd(avg.real(), 0.0);
                T_[(m_ + i) + static_cast<size_t>(m_ + j) * mmax_] = avg;
                T_[(m_ + j) + static_cast<size_t>(m_ + i) * mmax_] = std::conj(avg);
            }
        m_ = mn;
    }

    void rayleigh_ritz() {
        const double t0 = now_ms();
        std::vector<double> w(m_);
        cd* Yb = reinterpret_cast<cd*>(c_.h_coef.p);
        std::fill(Yb, Yb + static_cast<size_t>(m_) * b_, cd(0.0, 0.0));
        xatu_linalg::heev_lowest(m_, T_.data(), mmax_, b_, w.data(), Yb);
        theta_.assign(w.begin(), w.begin() + b_);
        rr_ms_ += now_ms() - t0;
    }

    void ritz_and_residuals() {
        std::copy(theta_.begin(), theta_.end(), c_.h_theta.p);
        XATU_HIP(hipMemcpyAsync(ws_.Y.p, c_.h_coef.p, static_cast<size_t>(m_) * b_ * sizeof(cd), hipMemcpyHostToDevice,
                                c_.stream));
        XATU_HIP(hipMemcpyAsync(ws_.theta.p, c_.h_theta.p, b_ * sizeof(double), hipMemcpyHostToDevice, c_.stream));
        const int n = static_cast<int>(N_);
        EventTimer ev(c_.stream);
        ev.start();
        launch_block_gemm(ws_.V.p, n, ws_.Y.p, m_, N_, m_, b_, 1.0, 0.0, ws_.X.p, n, c_.stream);
        launch_block_gemm(ws_.W.p, n, ws_.Y.p, m_, N_, m_, b_, 1.0, 0.0, ws_.AX.p, n, c_.stream);
        ev.stop();
        int nblk = 0;
        launch_residual(ws_.X.p, ws_.AX.p, ws_.theta.p, N_, b_, ws_.R.p, ws_.partial.p, &nblk, c_.stream);
        XATU_HIP(hipMemcpyAsync(c_.h_part.p, ws_.partial.p, static_cast<size_t>(nblk) * b_ * sizeof(double),
                                hipMemcpyDeviceToHost, c_.stream));
        XATU_HIP(hipStreamSynchronize(c_.stream));
        proj_ms_ += ev.ms();
        residual_norms(nblk);
    }

    void restart() {
        launch_copy_vectors(ws_.X.p, static_cast<int64_t>(N_) * b_, ws_.V.p, c_.stream);
        launch_copy_vectors(ws_.AX.p, static_cast<int64_t>(N_) * b_, ws_.W.p, c_.stream);
        std::fill(T_.begin(), T_.end(), cd(0.0, 0.0));
        for (int i = 0; i < b_; ++i) T_[i + static_cast<size_t>(i) * mmax_] = cd(theta_[i], 0.0);
        m_ = b_;
        rep_->restarts++;
    }

    // Orthonormalise the nb vectors in the basis slot V[:, m:m+nb] against V[:, 0:m] and
    // among themselves; returns the number kept, written to V[:, m:m+kept].
    // Two passes of classical Gram-Schmidt against the basis; the Gram kernel of each
    // pass forms [V, T]^H T, so the second pass also yields the block's own Gram matrix
    // T^H T (its change by the second, O(eps) projection is O(eps^2)). The block is then
    // whitened through the eigen-decomposition of its column-normalised Gram matrix,
    // dropping dependent directions. An ill-conditioned block (kept eigenvalues spanning
    // more than 1e8) repeats the whole procedure once on the whitened vectors.
    int orthonormalise_slot(int nb, int round = 0) {
        ortho_gram(nb);
        XATU_HIP(hipMemcpyAsync(c_.h_proj.p, ws_.P.p, static_cast<size_t>(m_ + nb) * nb * sizeof(cd),
                                hipMemcpyDeviceToHost, c_.stream));
        XATU_HIP(hipStreamSynchronize(c_.stream));
        drain_timers();
        return ortho_whiten(nb, round);
    }

    // The device part: two classical Gram-Schmidt passes; P = [V, T]^H T of the second.
    void ortho_gram(int nb) {
        const int n = static_cast<int>(N_);
        const int mn = m_ + nb;
        double2* T = ws_.V.p + static_cast<size_t>(m_) * N_;
        EventTimer* ev = timer(1);
        ev->start();
        for (int pass = 0; pass < 2; ++pass) {
            launch_gram(c_, ws_.V.p, n, T, n, mn, nb, ws_.P.p, mn, c_.stream);
            if (m_ > 0) {
                launch_block_gemm(ws_.V.p, n, ws_.P.p, mn, N_, m_, nb, -1.0, 1.0, T, n, c_.stream);
            }
        }
        ev->stop();
    }

    // The host part, with P on the host (Ph, else the pinned copy): whiten the block (its Gram
    // matrix rows m..m+nb).
    int ortho_whiten(int nb, int round, const cd* Ph = nullptr) {
        const int n = static_cast<int>(N_);
        const int mn = m_ + nb;
        double2* T = ws_.V.p + static_cast<size_t>(m_) * N_;
        const cd* P = Ph ? Ph : reinterpret_cast<const cd*>(c_.h_proj.p);
        // column-normalised Gram matrix of the block
        std::vector<cd> G(static_cast<size_t>(nb) * nb);
        std::vector<double> dsc(nb);
        for (int j = 0; j < nb; ++j) {
            const double g = P[(m_ + j) + static_cast<size_t>(j) * mn].real();
            dsc[j] = g > 0.0 ? 1.0 / std::sqrt(g) : 0.0;
        }
        for (int j = 0; j < nb; ++j)
            for (int i = 0; i < nb; ++i) G[i + static_cast<size_t>(j) * nb] = P[(m_ + i) + static_cast<size_t>(j) * mn] * dsc[i] * dsc[j];
        // Well-conditioned blocks (every Cholesky pivot^2 of the unit-diagonal G above 1e-6, so
        // cond(G) stays below ~1e6): S = D^-1/2 L^-H from G = L L^H, a tenth of the cost of the
        // eigen-decomposition below, which handles the rest (dependent directions dropped).
        {
            std::vector<cd> L(G);
            bool ok = true;
            for (int j = 0; j < nb && ok; ++j) {
                double d = L[j + static_cast<size_t>(j) * nb].real();
                for (int k = 0; k < j; ++k) d -= std::norm(L[j + static_cast<size_t>(k) * nb]);
                if (!(d > 1e-6)) {
                    ok = false;
                    break;
                }
                const double ljj = std::sqrt(d);
                L[j + static_cast<size_t>(j) * nb] = ljj;
                for (int i = j + 1; i < nb; ++i) {
                    cd sum = L[i + static_cast<size_t>(j) * nb];
                    for (int k = 0; k < j; ++k) sum -= L[i + static_cast<size_t>(k) * nb] * std::conj(L[j + static_cast<size_t>(k) * nb]);
                    L[i + static_cast<size_t>(j) * nb] = sum / ljj;
                }
            }
            if (ok) {
                std::vector<cd> Li(static_cast<size_t>(nb) * nb, cd(0.0, 0.0));  // L^-1, lower
                for (int j = 0; j < nb; ++j) {
                    Li[j + static_cast<size_t>(j) * nb] = 1.0 / L[j + static_cast<size_t>(j) * nb].real();
                    for (int i = j + 1; i < nb; ++i) {
                        cd sum(0.0, 0.0);
                        for (int k = j; k < i; ++k) sum += L[i + static_cast<size_t>(k) * nb] * Li[k + static_cast<size_t>(j) * nb];
                        Li[i + static_cast<size_t>(j) * nb] = -sum / L[i + static_cast<size_t>(i) * nb].real();
                    }
                }
                cd* S = reinterpret_cast<cd*>(c_.h_whiten.p);
                for (int j = 0; j < nb; ++j)
                    for (int i = 0; i < nb; ++i)
                        S[i + static_cast<size_t>(j) * nb] = i <= j ? dsc[i] * std::conj(Li[j + static_cast<size_t>(i) * nb]) : cd(0.0, 0.0);
                launch_copy_vectors(c_.h_whiten.dev, static_cast<int64_t>(nb) * nb, ws_.P.p, c_.stream);
                launch_block_gemm(T, n, ws_.P.p, nb, N_, nb, nb, 1.0, 0.0, ws_.Tb.p, n, c_.stream);
                launch_copy_vectors(ws_.Tb.p, static_cast<int64_t>(N_) * nb, T, c_.stream);
                return nb;
            }
        }
        std::vector<double> lam(nb);
        std::vector<cd> U(static_cast<size_t>(nb) * nb);
        xatu_linalg::heev(nb, G.data(), nb, lam.data(), U.data());
        const double lmax = std::max(lam[nb - 1], 0.0);
        std::vector<cd> S;
        int kept = 0;
        double lmin_kept = lmax;
        for (int j = nb - 1; j >= 0; --j) {
            if (lam[j] > 1e-20 && lam[j] > 1e-14 * lmax) {
                const double sc = 1.0 / std::sqrt(lam[j]);
                for (int i = 0; i < nb; ++i) S.push_back(U[i + static_cast<size_t>(j) * nb] * (dsc[i] * sc));
                lmin_kept = lam[j];
                ++kept;
            }
        }
        if (kept == 0) return 0;
        // V[:, m:m+kept] = T S (through the scratch block: the product may not overlap its input)
        std::copy(S.begin(), S.end(), reinterpret_cast<cd*>(c_.h_whiten.p));
        launch_copy_vectors(c_.h_whiten.dev, static_cast<int64_t>(S.size()), ws_.P.p, c_.stream);
        launch_block_gemm(T, n, ws_.P.p, nb, N_, nb, kept, 1.0, 0.0, ws_.Tb.p, n, c_.stream);
        launch_copy_vectors(ws_.Tb.p, static_cast<int64_t>(N_) * kept, T, c_.stream);
        if (round == 0 && lmin_kept < 1e-8 * lmax) return orthonormalise_slot(kept, 1);
        return kept;
    }

    bool add_corrections() {
        const int nb = std::min(b_, mmax_ - m_);
        if (nb <= 0) return false;
        double2* T = ws_.V.p + static_cast<size_t>(m_) * N_;
        launch_precondition(ws_.R.p, ws_.theta.p, c_.d_diag.p, N_, nb, T, c_.stream);
        const int kept = orthonormalise_slot(nb);
        if (kept == 0) return false;
        append_block(kept);
        return true;
    }

    Context& c_;
    int prec_, mf_, nev_, b_, mmax_;
    Workspace& ws_;  // the context's: kept between solves
    bool dev_rr_ = false;
    std::vector<Pending> pending_;
    xatu_solve_report* rep_;
    int64_t N_ = 0;
    int m_ = 0;
    std::vector<cd> T_;
    std::vector<double> theta_, res_;
    double prod_ms_ = 0.0, rr_ms_ = 0.0, proj_ms_ = 0.0, last_worst_ = 0.0;
};

}  // namespace

void solve_dense_diag(Context& c, int nstates, double* eigval_out, double* eigvec_out) {
    if (!c.built64) throw Error(XATU_ERR_LOGIC, "diagonalizeRaw(): BSE Hamiltonian is not initialized.");
    const int n = static_cast<int>(c.N);
    DevBuf<double2> A;
    A.alloc(static_cast<size_t>(n) * n);
    XATU_HIP(hipMemcpyAsync(A.p, c.d_H64.p, A.bytes(), hipMemcpyDeviceToDevice, c.stream));
    DevBuf<double> D, E;
    DevBuf<int> info;
    D.alloc(n);
    E.alloc(n);
    info.alloc(1);
    XATU_BLAS(rocsolver_zheevd(c.blas, rocblas_evect_original, rocblas_fill_lower, n, bz(A.p), n, D.p, E.p,
                               info.p));
    int hinfo = 0;
    XATU_HIP(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    XATU_HIP(hipStreamSynchronize(c.stream));
    if (hinfo != 0) throw Error(XATU_ERR_NOT_CONVERGED, "zheevd failed to converge");
    const int ns = (nstates <= 0 || nstates > n) ? n : nstates;
    XATU_HIP(hipMemcpy(eigval_out, D.p, ns * sizeof(double), hipMemcpyDeviceToHost));
    if (eigvec_out) {
        XATU_HIP(hipMemcpy(eigvec_out, A.p, static_cast<size_t>(n) * ns * sizeof(double2), hipMemcpyDeviceToHost));
    }
    c.last_nstates = ns;
    c.last_n = c.N;
    c.d_last_vec.reserve(static_cast<size_t>(n) * ns);
    XATU_HIP(hipMemcpyAsync(c.d_last_vec.p, A.p, static_cast<size_t>(n) * ns * sizeof(double2),
                            hipMemcpyDeviceToDevice, c.stream));
    c.last_val.assign(eigval_out, eigval_out + ns);
    XATU_HIP(hipStreamSynchronize(c.stream));
}

void solve_davidson(Context& c, const xatu_solve_options& opt, double* eigval_out, double* eigvec_out,
                    xatu_solve_report* rep) {
    const int64_t N = c.N;
    const int nev = opt.nstates;
    if (nev <= 0 || nev > N) throw Error(XATU_ERR_INVALID_ARGUMENT, "nstates must be in [1, N]");
    int b = opt.block > 0 ? opt.block : std::max(2 * nev, 8);
    b = static_cast<int>(std::min<int64_t>(b, N));
    if (b < nev) throw Error(XATU_ERR_INVALID_ARGUMENT, "Davidson block must be >= nstates");
    int mmax = opt.max_basis > 0 ? opt.max_basis : 4 * b;
    mmax = static_cast<int>(std::min<int64_t>(mmax, N));
    mmax = std::max(mmax, std::min<int>(2 * b, static_cast<int>(N)));
    if (mmax < b + 1 && N > b) mmax = b + 1;
    const int maxit = opt.max_iter > 0 ? opt.max_iter : 400;
    const double tol = opt.tol > 0 ? opt.tol : 1e-5;

    const double t0 = now_ms();
    // arithmetic of the products with the FP32 matrix, for this solve only
    struct ArithScope {
        Context& c;
        bool saved;
        ArithScope(Context& cc, bool fp32) : c(cc), saved(cc.fp32_arith_only) { c.fp32_arith_only = fp32; }
        ~ArithScope() { c.fp32_arith_only = saved; }
    } arith(c, opt.fp32_arith == XATU_ARITH_FP32);
    std::unique_ptr<Davidson> d32;
    std::unique_ptr<Davidson> dav;
    bool ok;
    if (opt.precision == XATU_FP32) {
        // FP32 products to a tolerance above the FP32 noise floor ...
        d32 = std::make_unique<Davidson>(c, c.solver_ws[1], XATU_FP32, opt.matrix_free, nev, b, mmax, rep);
        d32->init_unit();
        // With a resident FP32 matrix, 6e-4 eV: past that the FP32 iterations, at the full
        // block and basis, cost more than the FP64 refinement iterations that replace them
        // (N = 36864, 18 orbitals: 2e-5 -> 3e-4 took the step from 67.1 to 63.9 ms at the same
        // 23 iterations, 2 -> 6 of them refinement; later, with cheaper iterations, 3e-4 ->
        // 6e-4 took it from 56.4 to 55.3 ms, 6 -> 8 refinement, and 2e-3 back to 57.0 ms).
        // Matrix-free, an FP64 product costs as much as an FP32 one, so the FP32 phase goes on
        // to 2e-5 (N = 129600 at 3e-4: 0.439 s with 8 refinement iterations, against 0.409 s
        // with 2).
        d32->run(std::max(0.5 * tol, opt.matrix_free ? 2e-5 : 6e-4), maxit);
        const int it32 = rep->iterations;
        // ... then FP64 refinement from the lowest nev Ritz vectors, with FP64 products
        // (matrix-free when no FP64 matrix is resident), until the FP64 residuals meet tol.
        // Block nev, basis 6 nev up to the device Rayleigh-Ritz limit (N = 36864, nev = 16:
        // basis 96 took 7 refinement iterations against 8 with 64, 53.5 against 55.0 ms per
        // step; 128 also 7, at 54.0 ms; blocks of 24 or 32 cost more per product than they save)
        const int bref = nev;
        const int mref = static_cast<int>(std::min<int64_t>(
            N, std::max({4 * bref, std::min(6 * bref, projected_eigen_max_dim()), bref + 1})));
        dav = std::make_unique<Davidson>(c, c.solver_ws[0], XATU_FP64, opt.matrix_free ? opt.matrix_free : (c.built64 ? 0 : 1), nev, bref, mref, rep);
        dav->init_from(d32->ritz_vectors());
        ok = dav->run(tol, maxit);
        rep->refine_iterations = rep->iterations - it32;
        rep->ms_products = d32->products_ms();
        rep->ms_rayleigh_ritz = d32->rr_ms();
        rep->ms_projections = d32->projections_ms();
    } else {
        dav = std::make_unique<Davidson>(c, c.solver_ws[0], opt.precision, opt.matrix_free, nev, b, mmax, rep);
        dav->init_unit();
        ok = dav->run(tol, maxit);
    }
    rep->converged = ok ? 1 : 0;
    rep->max_residual = dav->worst();
    rep->ms_products += dav->products_ms();
    rep->ms_rayleigh_ritz += dav->rr_ms();
    rep->ms_projections += dav->projections_ms();
    const Davidson& final_d = *dav;

    const std::vector<double>& th = final_d.theta();
    std::copy(th.begin(), th.begin() + nev, eigval_out);
    c.last_nstates = nev;
    c.last_n = N;
    c.last_val.assign(th.begin(), th.begin() + nev);
    c.d_last_vec.reserve(static_cast<size_t>(N) * nev);
    XATU_HIP(hipMemcpyAsync(c.d_last_vec.p, final_d.ritz_vectors(), static_cast<size_t>(N) * nev * sizeof(double2),
                            hipMemcpyDeviceToDevice, c.stream));
    if (eigvec_out) {
        XATU_HIP(hipMemcpyAsync(eigvec_out, final_d.ritz_vectors(), static_cast<size_t>(N) * nev * sizeof(double2),
                                hipMemcpyDeviceToHost, c.stream));
    }
    XATU_HIP(hipStreamSynchronize(c.stream));
    rep->ms_total = now_ms() - t0;
}

}  // namespace xatu_dev
