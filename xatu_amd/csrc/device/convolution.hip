// Matrix-free BSE products by k-space convolution.
//
// The direct term of the reference's matrix (src/ExcitonTB.cpp:520-548, 840-921),
//   D(kcv, k'c'v') = sum_ab rho^c_a M_ab(k'-k) rho^v_b,
//   rho^c_a = sum_{al in a} conj(C_kc,al) C_k'c',al,  rho^v_b = sum_{be in b} conj(V_k'v',be) V_kv,be,
// depends on the pair (k, k') only through the lattice-sum table M(k'-k), so its
// product with a vector factorises:
//   z_{al be}(k')  = sum_{c'v'} C_k'c',al conj(V_k'v',be) x(k'c'v')
//   w_{al be}(k)   = sum_k' M_{a(al) b(be)}(k'-k) z_{al be}(k')        (a correlation on the mesh)
//   (D x)(kcv)     = sum_{al be} conj(C_kc,al) V_kv,be w_{al be}(k).
// The correlation is evaluated with FFTs on the mesh of the lattice sums (mesh_n points
// per axis; k-points sit at their integer offsets, so reduced and shifted meshes work
// unchanged): w = IFFT( P_ab . FFT(z) ) with P_ab the backward transform of the table
// M_ab(Delta) = ft[q(Delta)] (q the same mesh index the assembly uses), scaled by 1/mesh.
// Cost per vector: norb^2 FFT pairs of mesh size plus O(nk norb^2 nc) contractions,
// instead of O(nk^2) pair evaluations for a recomputed tile.
//
// Off the diagonal the reference fills only the upper triangle and mirrors it. The
// direct formula is Hermitian under the exchange of the pair because the table obeys
// M(-q) = conj M(q) (including the conjugated diagonal the reference stores), so the
// full correlation equals the mirrored matrix up to rounding. The exchange term
// X(i,j) = sum_a E1_a(i) E2_a(j) is not Hermitian when V(Q) has a complex diagonal, so
// it follows the triangle: (X_tri x)_i = sum_a [E1_a(i) sum_{j>i} E2_a(j) x_j
// + conj(E2_a(i)) sum_{j<i} conj(E1_a(j)) x_j] + Re(sum_a E1_a(i) E2_a(i)) x_i, with the
// partial sums from a blocked scan in a fixed order. The diagonal energies enter as
// e_i x_i, so H x = e x - D x + X_tri x exactly as the assembled matrix.
//
// FFTs: rocFFT (batched, in place, interleaved complex); plans are cached per
// precision and batch size. The exchange part runs in FP64 for both precisions.
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include <rocfft/rocfft.h>

#include "context.hpp"
#include "cplx.hpp"
#include "mfma.hpp"

#define XATU_FFT(call)                                                                                    \
    do {                                                                                                  \
        rocfft_status s_ = (call);                                                                        \
        if (s_ != rocfft_status_success) {                                                                \
            throw ::xatu_dev::Error(XATU_ERR_DEVICE,                                                      \
                                    std::string(#call) + ": rocfft status " + std::to_string(static_cast<int>(s_))); \
        }                                                                                                 \
    } while (0)

namespace xatu_dev {

namespace {

constexpr int kScanRows = 256;

// Layout of the correlation buffer G (batch entries b of M mesh points g = o0 + mesh_n o1).
// Entry-major: G[b M + g]. Row-interleaved (2-D meshes): G[(o1 batch + b) mesh_n + o0],
// so that the FFTs along o0 (unit stride, rows of mesh_n) and along o1 (stride
// batch mesh_n, consecutive transforms adjacent) are each one uniform batched 1-D pass
// with no transposition between them.
struct GLayout {
    int64_t bstride;    // distance between consecutive batch entries
    int rowlen;         // 0: entry-major; mesh_n: row-interleaved
    int64_t rowstride;  // row-interleaved: batch mesh_n
    __device__ int64_t at(int64_t b, int g) const {
        return b * bstride + (rowlen ? static_cast<int64_t>(g / rowlen) * rowstride + g % rowlen : g);
    }
};

__device__ inline int atom_of(const LatticeDesc& lat, int orb) {
    int a = 0;
    while (a + 1 < lat.natoms && orb >= lat.orb_begin[a + 1]) ++a;
    return a;
}

// Mt[ab][g] = ft[q(g) NA^2 + ab], g = sum_j d_j m^j, q = sum_j ((d_j + m/2) mod m) m^j.
__global__ void table_grid_kernel(LatticeDesc lat, const double2* __restrict__ ft, double2* __restrict__ Mt) {
    const int64_t M = lat.nmesh;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int na2 = lat.natoms * lat.natoms;
    if (idx >= M * na2) return;
    const int ab = static_cast<int>(idx / M);
    int64_t g = idx % M;
    const int m = lat.mesh_n;
    int64_t q = 0, mul = 1;
    for (int j = 0; j < lat.ndim; ++j) {
        const int d = static_cast<int>(g % m);
        g /= m;
        q += static_cast<int64_t>((d + m / 2) % m) * mul;
        mul *= m;
    }
    Mt[idx] = ft[q * na2 + ab];
}

template <typename R>
__global__ void scale_spectrum_kernel(const double2* __restrict__ in, int64_t n, double scale, cx_t<R>* __restrict__ out) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = mk(static_cast<R>(in[i].x * scale), static_cast<R>(in[i].y * scale));
}

// Band stacks transposed for coalesced access across k: Ct[(c norb + al) nk + k] =
// C^{k+Q}_{nv+c}(k)[al] (conduction) and Vt[(v norb + be) nk + k] = C^{k}_{v}(k)[be] (valence).
template <typename R>
__global__ void transpose_stacks_kernel(int nk, int nv, int nc, int norb, const double2* __restrict__ evk,
                                        const double2* __restrict__ evkq, cx_t<R>* __restrict__ Ct,
                                        cx_t<R>* __restrict__ Vt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;  // 0..nc*norb-1 conduction, then nv*norb valence
    if (k >= nk) return;
    const int nb = nv + nc;
    if (col < nc * norb) {
        const int c = col / norb, al = col % norb;
        const double2 v = evkq[(static_cast<size_t>(k) * nb + nv + c) * norb + al];
        Ct[static_cast<size_t>(col) * nk + k] = mk(static_cast<R>(v.x), static_cast<R>(v.y));
    } else {
        const int cv = col - nc * norb;
        const int v = cv / norb, be = cv % norb;
        const double2 w = evk[(static_cast<size_t>(k) * nb + v) * norb + be];
        Vt[static_cast<size_t>(cv) * nk + k] = mk(static_cast<R>(w.x), static_cast<R>(w.y));
    }
}

// orbitals al per thread in the z contraction: each group re-reads u (FP64 at 18 orbitals:
// 6 took the product's z step from 176 to 158 us against 4; FP32 at 36 orbitals: 12 made it
// slower, 1.92 against 1.77 ms, the G stores of more orbitals per thread outweighing the
// fewer passes over u)
template <typename R>
constexpr int conv_orb_group() { return sizeof(R) == 8 ? 6 : 4; }

// vectors per thread in the u and y contractions (register budget: FP64 halves it)
template <typename R>
constexpr int conv_vec_group() { return sizeof(R) == 8 ? 2 : 4; }
// vectors per thread of the y contraction: one, for enough threads to cover the latency of
// its norb^2 loads per vector (the band coefficients it re-reads hit the caches)
template <typename R>
constexpr int conv_y_group() { return 1; }
// orbitals be per pass over al in the y contraction (their partial sums stay in registers)
template <typename R>
constexpr int conv_y_be_block() { return sizeof(R) == 8 ? 6 : 12; }

// u[((n NC + c) norb + be) nk + k] = sum_v conj(V_k[be, v]) x[k NCV + c NV + v, n]
// thread per k, blockIdx.y = group of vectors
template <typename R, int NC, int NV>
__global__ __launch_bounds__(256) void conv_u_kernel(int nk, int norb, int nvec, const cx_t<R>* __restrict__ Vt,
                                                     const double2* __restrict__ X, int64_t N, cx_t<R>* __restrict__ U) {
    constexpr int VG = conv_vec_group<R>();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = blockIdx.y * VG;
    if (k >= nk) return;
    constexpr int ncv = NC * NV;
    cx_t<R> x[VG][NC][NV];
#pragma unroll
    for (int g = 0; g < VG; ++g)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const int n = n0 + g;
                const double2 xv = n < nvec ? X[static_cast<size_t>(k) * ncv + c * NV + v + n * N] : mk(0.0, 0.0);
                x[g][c][v] = mk(static_cast<R>(xv.x), static_cast<R>(xv.y));
            }
    for (int be = 0; be < norb; ++be) {
        cx_t<R> vb[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) vb[v] = Vt[(static_cast<size_t>(v) * norb + be) * nk + k];
#pragma unroll
        for (int g = 0; g < VG; ++g) {
            const int n = n0 + g;
            if (n < nvec) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    cx_t<R> s = mk(R(0), R(0));
#pragma unroll
                    for (int v = 0; v < NV; ++v) cfma_conj(s, vb[v], x[g][c][v]);
                    U[((static_cast<size_t>(n) * NC + c) * norb + be) * nk + k] = s;
                }
            }
        }
    }
}

// G[entry (al norb + be) nvec + n][g(k)] = sum_c C_k[al, c] u[n][c][be][k]
// thread per k, blockIdx.y = group of conv_orb_group<R>() orbitals al, blockIdx.z = vector n
template <typename R, int NC>
__global__ __launch_bounds__(256) void conv_z_kernel(int nk, int norb, int nvec, GLayout lay, const int* __restrict__ gk,
                                                     const cx_t<R>* __restrict__ Ct, const cx_t<R>* __restrict__ U,
                                                     cx_t<R>* __restrict__ G) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    constexpr int AG = conv_orb_group<R>();
    const int al0 = blockIdx.y * AG;
    const int n = blockIdx.z;
    if (k >= nk) return;
    cx_t<R> Ca[AG][NC];
#pragma unroll
    for (int a = 0; a < AG; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c)
            Ca[a][c] = al0 + a < norb ? Ct[(static_cast<size_t>(c) * norb + al0 + a) * nk + k] : mk(R(0), R(0));
    const int g = gk[k];
    for (int be = 0; be < norb; ++be) {
        cx_t<R> ub[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) ub[c] = U[((static_cast<size_t>(n) * NC + c) * norb + be) * nk + k];
#pragma unroll
        for (int a = 0; a < AG; ++a) {
            if (al0 + a < norb) {
                cx_t<R> s = mk(R(0), R(0));
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(s, Ca[a][c], ub[c]);
                G[lay.at((static_cast<int64_t>(al0 + a) * norb + be) * nvec + n, g)] = s;
            }
        }
    }
}

// G[b][xi] *= P[a(al) + NA b(be)][xi], batch b = (al norb + be) nvec + n
template <typename R>
__global__ void conv_multiply_kernel(LatticeDesc lat, int nvec, int64_t M, GLayout lay, const cx_t<R>* __restrict__ P,
                                     cx_t<R>* __restrict__ G) {
    const int64_t xi = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int b = blockIdx.y;
    if (xi >= M) return;
    const int ab = b / nvec;
    const int al = ab / lat.norb, be = ab % lat.norb;
    const int pa = atom_of(lat, al) + lat.natoms * atom_of(lat, be);
    const int64_t i = lay.at(b, static_cast<int>(xi));
    G[i] = cmul(G[i], P[static_cast<int64_t>(pa) * M + xi]);
}

// ---- fused transform along the second mesh axis (row-interleaved layout) ----
// For 2-D meshes whose side n = A B factors into two of the lengths 1, 2, 3, 4, 5, 6, 8, 9,
// 10, 12, 15, 16 (n <= kColMax), the transform along o1, the product with the spectrum and
// the inverse transform along o1 run in one kernel: a workgroup holds 32 adjacent columns
// u = b mesh_n + xi0 (each a length-n vector of stride batch mesh_n) and does a four-step
// transform in registers with two LDS exchanges, so G is read and written once instead of
// twice per direction with a transposition in between.
constexpr int kColMax = 128;

// In-register DFT of length RD; sg = -1 forward (e^{-2 pi i / RD}), +1 inverse.
template <typename R, int RD>
__device__ inline void small_dft(cx_t<R> (&v)[RD], R sg) {
    using C = cx_t<R>;
    auto isc = [&](C z, R f) { return mk(-sg * f * z.y, sg * f * z.x); };  // i sg f z
    if constexpr (RD == 2) {
        const C a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (RD == 3) {
        const R h3 = R(0.86602540378443864676);  // sin(2 pi / 3)
        const C a = cadd(v[1], v[2]), d = csub(v[1], v[2]);
        const C m = mk(v[0].x - R(0.5) * a.x, v[0].y - R(0.5) * a.y);
        const C t = isc(d, h3);
        v[0] = cadd(v[0], a);
        v[1] = cadd(m, t);
        v[2] = csub(m, t);
    } else if constexpr (RD == 4) {
        const C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]), t2 = cadd(v[1], v[3]), t3 = isc(csub(v[1], v[3]), R(1));
        v[0] = cadd(t0, t2);
        v[1] = cadd(t1, t3);
        v[2] = csub(t0, t2);
        v[3] = csub(t1, t3);
    } else {
        static_assert(RD == 5, "radix");
        // cos and sin of 2 pi / 5 and 4 pi / 5
        const R c1 = R(0.30901699437494742410), s1 = R(0.95105651629515357212);
        const R c2 = R(-0.80901699437494742410), s2 = R(0.58778525229247312917);
        const C a1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]), a2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const C p1 = mk(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
        const C p2 = mk(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
        const C q1 = isc(mk(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y), R(1));
        const C q2 = isc(mk(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y), R(1));
        v[0] = cadd(v[0], cadd(a1, a2));
        v[1] = cadd(p1, q1);
        v[4] = csub(p1, q1);
        v[2] = cadd(p2, q2);
        v[3] = csub(p2, q2);
    }
}

// Twiddle e^{sg 2 pi i k / n} from the table of e^{-2 pi i k / n}.
template <typename R>
__device__ inline cx_t<R> twiddle(const cx_t<R>* tw, int k, R sg) {
    const cx_t<R> w = tw[k];
    return mk(w.x, sg < R(0) ? w.y : -w.y);
}

// DFT of length A B in registers as two Stockham stages (radix A, then radix B with the
// twiddles e^{sg 2 pi i j r / (A B)} read from the n-point table).
template <typename R, int A, int B>
__device__ inline void composite_dft(cx_t<R> (&v)[A * B], R sg, const cx_t<R>* tw, int n) {
    using C = cx_t<R>;
    constexpr int RD = A * B;
    C t[RD];
#pragma unroll
    for (int j = 0; j < B; ++j) {
        C x[A];
#pragma unroll
        for (int r = 0; r < A; ++r) x[r] = v[j + r * B];
        small_dft<R, A>(x, sg);
#pragma unroll
        for (int r = 0; r < A; ++r) t[j * A + r] = x[r];
    }
    const int unit = n / RD;
#pragma unroll
    for (int j = 0; j < A; ++j) {
        C x[B];
#pragma unroll
        for (int r = 0; r < B; ++r) x[r] = (j * r) ? cmul(t[j + r * A], twiddle(tw, j * r * unit, sg)) : t[j + r * A];
        small_dft<R, B>(x, sg);
#pragma unroll
        for (int r = 0; r < B; ++r) v[j + r * A] = x[r];
    }
}

// DFT of one of the lengths 1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16.
template <typename R, int RD>
__device__ inline void dft_len(cx_t<R> (&a)[RD], R sg, const cx_t<R>* tw, int n) {
    if constexpr (RD == 1) {
    } else if constexpr (RD <= 5) {
        small_dft<R, RD>(a, sg);
    } else if constexpr (RD == 6) {
        composite_dft<R, 3, 2>(a, sg, tw, n);
    } else if constexpr (RD == 8) {
        composite_dft<R, 4, 2>(a, sg, tw, n);
    } else if constexpr (RD == 9) {
        composite_dft<R, 3, 3>(a, sg, tw, n);
    } else if constexpr (RD == 10) {
        composite_dft<R, 5, 2>(a, sg, tw, n);
    } else if constexpr (RD == 12) {
        composite_dft<R, 3, 4>(a, sg, tw, n);
    } else if constexpr (RD == 15) {
        composite_dft<R, 5, 3>(a, sg, tw, n);
    } else {
        static_assert(RD == 16, "length");
        composite_dft<R, 4, 4>(a, sg, tw, n);
    }
}

constexpr int kFourStepCols = 32; // columns per workgroup (512 threads: 16 per column)

// Stage 1 (thread s < B): rows B j2 + s, A-point DFT, twiddle, Y[s][k1] to LDS.
template <typename R, int A, int T>
__device__ inline void four_step_in(const cx_t<R>* gcol, bool live, int64_t count, int B, int s, int c, int n,
                                    const cx_t<R>* tw, cx_t<R>* ex) {
    cx_t<R> v[A];
#pragma unroll
    for (int j2 = 0; j2 < A; ++j2) v[j2] = live ? gcol[static_cast<int64_t>(B * j2 + s) * count] : mk(R(0), R(0));
    dft_len<R, A>(v, R(-1), tw, n);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) ex[(s * A + k1) * T + c] = k1 ? cmul(v[k1], tw[s * k1]) : v[k1];
}

// Stage 2 (threads with act, s < A): Y[j1][s], B-point DFT (xi1 = s + A k2), times the
// spectrum column pc, inverse B-point DFT, twiddle back, Z[s][j1] to LDS. Every thread of
// the workgroup calls it (B is uniform) and passes its barrier between the reads and the
// writes of the shared buffer.
template <typename R, int B, int T>
__device__ inline void four_step_mid(bool act, const cx_t<R>* pc, int A, int s, int c, int n, const cx_t<R>* tw,
                                     cx_t<R>* ex) {
    cx_t<R> v[B];
    if (act) {
#pragma unroll
        for (int j1 = 0; j1 < B; ++j1) v[j1] = ex[(j1 * A + s) * T + c];
    }
    __syncthreads();
    if (!act) return;
    dft_len<R, B>(v, R(-1), tw, n);
    if (pc) {
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k2] = cmul(v[k2], pc[static_cast<int64_t>(n) * (s + A * k2)]);
    }
    dft_len<R, B>(v, R(1), tw, n);
#pragma unroll
    for (int j1 = 0; j1 < B; ++j1) {
        const cx_t<R> w = tw[j1 * s];
        ex[(s * B + j1) * T + c] = (j1 * s) ? cmul(v[j1], mk(w.x, -w.y)) : v[j1];
    }
}

// Stage 3 (thread s < B): Z[k1][s], inverse A-point DFT, rows B j2 + s to G.
template <typename R, int A, int T>
__device__ inline void four_step_out(cx_t<R>* gcol, bool live, int64_t count, int B, int s, int c, int n,
                                     const cx_t<R>* tw, const cx_t<R>* ex) {
    cx_t<R> v[A];
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) v[k1] = ex[(k1 * B + s) * T + c];
    dft_len<R, A>(v, R(1), tw, n);
    if (live) {
#pragma unroll
        for (int j2 = 0; j2 < A; ++j2) gcol[static_cast<int64_t>(B * j2 + s) * count] = v[j2];
    }
}

#define XATU_FOUR_STEP_LENGTHS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(8) X(9) X(10) X(12) X(15) X(16)

// The same fused column transform for n = A B (A >= B, both lengths of dft_len) as a
// four-step transform in registers: thread (s, c) of column c loads rows B j2 + s (j2 < A)
// and does the A-point DFTs, twiddles by e^{-2 pi i s k1 / n} and stores Y[s][k1] to LDS;
// thread (s, c) then takes Y[j1][s] (j1 < B), does the B-point DFT (outputs xi1 = s + A k2),
// multiplies by the spectrum, inverts the B-point DFT, twiddles back and stores Z[s][j1];
// thread (s, c) finally takes Z[k1][s] (k1 < A), inverts the A-point DFT and stores rows
// B j2 + s. Two LDS exchanges instead of a pass per radix. grid ceil(count / 32), block 512.
template <typename R>
__global__ __launch_bounds__(512) void conv_columns4_kernel(LatticeDesc lat, int n, int A, int B, int64_t count,
                                                            int nvec, int64_t M, const cx_t<R>* __restrict__ P,
                                                            cx_t<R>* __restrict__ G) {
    using C = cx_t<R>;
    constexpr int T = kFourStepCols;
    __shared__ C ex[kColMax * T];
    __shared__ C tw[kColMax];
    const int c = threadIdx.x % T, s = threadIdx.x / T;
    const int64_t u = static_cast<int64_t>(blockIdx.x) * T + c;
    const bool live = u < count;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        double sn, cs;
        sincospi(2.0 * k / n, &sn, &cs);
        tw[k] = mk(static_cast<R>(cs), static_cast<R>(-sn));  // e^{-2 pi i k / n}
    }
    __syncthreads();
    C* gcol = G + u;
    if (s < B) {
        switch (A) {
#define XATU_IN(L) \
    case L: four_step_in<R, L, T>(gcol, live, count, B, s, c, n, tw, ex); break;
            XATU_FOUR_STEP_LENGTHS(XATU_IN)
#undef XATU_IN
        }
    }
    __syncthreads();
    const C* pc = nullptr;
    if (live) {
        const int64_t b = u / n;
        const int xi0 = static_cast<int>(u - b * n);
        const int ab = static_cast<int>(b / nvec);
        const int al = ab / lat.norb, be = ab % lat.norb;
        pc = P + static_cast<int64_t>(atom_of(lat, al) + lat.natoms * atom_of(lat, be)) * M + xi0;
    }
    // B is the same for the whole workgroup: every thread takes the same case and passes the
    // barrier inside four_step_mid once
    switch (B) {
#define XATU_MID(L) \
    case L: four_step_mid<R, L, T>(s < A, pc, A, s, c, n, tw, ex); break;
        XATU_FOUR_STEP_LENGTHS(XATU_MID)
#undef XATU_MID
    }
    __syncthreads();
    if (s < B) {
        switch (A) {
#define XATU_OUT(L) \
    case L: four_step_out<R, L, T>(gcol, live, count, B, s, c, n, tw, ex); break;
            XATU_FOUR_STEP_LENGTHS(XATU_OUT)
#undef XATU_OUT
        }
    }
}

// Whole-plane form of the same transform, for 2-D meshes whose n x n plane (FP64 up to n = 100,
// FP32 up to 128) fits in LDS: one workgroup per batch entry b loads its plane once,
// transforms the rows, transforms the columns with the spectrum product between the forward
// and inverse column transforms, inverts the rows and stores the plane, so G is read and
// written once per product instead of once per pass (row transform, column kernel, inverse
// row transform). Every stage is a four-step stage done in place: a task reads the positions
// B j + s (or B t + j) of its row or column and writes the same positions, so only the
// barriers between stages are needed. The forward A-point stage stores its output k1 at
// B k1 + s; after the forward B-point stage, position B t + k2 of a row holds the frequency
// t + A k2 (the columns are transformed in that order and the spectrum read accordingly).
// Rows are ldp = n + 1 entries apart in LDS so that tasks on consecutive rows use different
// banks. grid batch, block plane_threads<R>() (FP64: 512, whose 256 registers per thread hold
// the 16-point stages without spilling; one FP64 plane takes most of the LDS anyway),
// dynamic LDS (kColMax + n (n + 1)) entries.
template <typename R>
constexpr int plane_threads() { return sizeof(R) == 8 ? 512 : 1024; }

// A-point stage, forward (from positions B j2 + s to B k1 + s, times e^{-2 pi i s k1 / n}) or
// inverse (from B k1 + s to B j2 + s, no twiddle); in and out may be the same.
template <typename R, int A>
__device__ inline void plane_stage_a(const cx_t<R>* in, int istr, cx_t<R>* out, int ostr, int B, int s, int n,
                                     const cx_t<R>* tw, bool forward) {
    cx_t<R> v[A];
#pragma unroll
    for (int j = 0; j < A; ++j) v[j] = in[static_cast<int64_t>(B * j + s) * istr];
    dft_len<R, A>(v, forward ? R(-1) : R(1), tw, n);
#pragma unroll
    for (int k = 0; k < A; ++k)
        out[static_cast<int64_t>(B * k + s) * ostr] = (forward && k) ? cmul(v[k], tw[s * k]) : v[k];
}

// B-point stage in place at positions B t + j (j < B): forward (output k2 at B t + k2), or
// inverse then times e^{+2 pi i j t / n}; with spec, forward, times spec[(t + A k2) n],
// inverse and twiddle back (the column stage between the forward and inverse transforms).
template <typename R, int B>
__device__ inline void plane_stage_b(cx_t<R>* base, int str, int t, int A, int n, const cx_t<R>* tw, int mode,
                                     const cx_t<R>* spec) {
    using C = cx_t<R>;
    C v[B];
#pragma unroll
    for (int j = 0; j < B; ++j) v[j] = base[(B * t + j) * str];
    if (mode >= 0) dft_len<R, B>(v, R(-1), tw, n);  // 0: forward; 1: forward, spectrum, inverse
    if (mode == 1) {
#pragma unroll
        for (int k = 0; k < B; ++k) v[k] = cmul(v[k], spec[static_cast<int64_t>(n) * (t + A * k)]);
    }
    if (mode != 0) dft_len<R, B>(v, R(1), tw, n);  // -1: inverse only
#pragma unroll
    for (int j = 0; j < B; ++j) {
        const C w = tw[j * t];
        base[(B * t + j) * str] = (mode != 0 && j * t) ? cmul(v[j], mk(w.x, -w.y)) : v[j];
    }
}

template <typename R, int NT>
__global__ __launch_bounds__(NT) void conv_plane_kernel(LatticeDesc lat, int n, int A, int B, int batch,
                                                                   int nvec, int64_t M, const cx_t<R>* __restrict__ P,
                                                                   cx_t<R>* __restrict__ G) {
    using C = cx_t<R>;
    extern __shared__ unsigned char plane_raw[];
    C* tw = reinterpret_cast<C*>(plane_raw);
    C* pl = tw + kColMax;
    const int ldp = n + 1;
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    for (int k = tid; k < n; k += NT) {
        double sn, cs;
        sincospi(2.0 * k / n, &sn, &cs);
        tw[k] = mk(static_cast<R>(cs), static_cast<R>(-sn));  // e^{-2 pi i k / n}
    }
    const int ab = b / nvec;
    const int al = ab / lat.norb, be = ab % lat.norb;
    const C* pp = P + static_cast<int64_t>(atom_of(lat, al) + lat.natoms * atom_of(lat, be)) * M;
    C* g = G + static_cast<int64_t>(b) * n;  // row o1 of the plane at g + o1 gs
    const int64_t gs = static_cast<int64_t>(batch) * n;
    __syncthreads();
    // A and B are the same for the whole workgroup: every thread takes the same case
#define XATU_SWITCH_A switch (A) { XATU_FOUR_STEP_LENGTHS(XATU_CASE_A) }
#define XATU_SWITCH_B switch (B) { XATU_FOUR_STEP_LENGTHS(XATU_CASE_B) }
    // rows, forward A-point stage from G (s fastest: a row's 16-byte entries together)
    for (int task = tid; task < n * B; task += NT) {
        const int s = task % B, o1 = task / B;
#define XATU_CASE_A(L) \
    case L: plane_stage_a<R, L>(g + o1 * gs, 1, pl + o1 * ldp, 1, B, s, n, tw, true); break;
        XATU_SWITCH_A
#undef XATU_CASE_A
    }
    __syncthreads();
    // rows, forward B-point stage (o1 fastest: rows ldp entries apart)
    for (int task = tid; task < n * A; task += NT) {
        const int o1 = task % n, t = task / n;
#define XATU_CASE_B(L) \
    case L: plane_stage_b<R, L>(pl + o1 * ldp, 1, t, A, n, tw, 0, nullptr); break;
        XATU_SWITCH_B
#undef XATU_CASE_B
    }
    __syncthreads();
    // columns (position p of every row), forward A-point stage
    for (int task = tid; task < n * B; task += NT) {
        const int p = task % n, s = task / n;
#define XATU_CASE_A(L) \
    case L: plane_stage_a<R, L>(pl + p, ldp, pl + p, ldp, B, s, n, tw, true); break;
        XATU_SWITCH_A
#undef XATU_CASE_A
    }
    __syncthreads();
    // columns: forward B-point stage, spectrum of row frequency p / B + A (p % B), inverse
    for (int task = tid; task < n * A; task += NT) {
        const int p = task % n, t = task / n;
        const C* spec = pp + (p / B + A * (p % B));
#define XATU_CASE_B(L) \
    case L: plane_stage_b<R, L>(pl + p, ldp, t, A, n, tw, 1, spec); break;
        XATU_SWITCH_B
#undef XATU_CASE_B
    }
    __syncthreads();
    // columns, inverse A-point stage
    for (int task = tid; task < n * B; task += NT) {
        const int p = task % n, s = task / n;
#define XATU_CASE_A(L) \
    case L: plane_stage_a<R, L>(pl + p, ldp, pl + p, ldp, B, s, n, tw, false); break;
        XATU_SWITCH_A
#undef XATU_CASE_A
    }
    __syncthreads();
    // rows, inverse B-point stage
    for (int task = tid; task < n * A; task += NT) {
        const int o1 = task % n, t = task / n;
#define XATU_CASE_B(L) \
    case L: plane_stage_b<R, L>(pl + o1 * ldp, 1, t, A, n, tw, -1, nullptr); break;
        XATU_SWITCH_B
#undef XATU_CASE_B
    }
    __syncthreads();
    // rows, inverse A-point stage to G
    for (int task = tid; task < n * B; task += NT) {
        const int s = task % B, o1 = task / B;
#define XATU_CASE_A(L) \
    case L: plane_stage_a<R, L>(pl + o1 * ldp, 1, g + o1 * gs, 1, B, s, n, tw, false); break;
        XATU_SWITCH_A
#undef XATU_CASE_A
    }
#undef XATU_SWITCH_A
#undef XATU_SWITCH_B
}

// LDS bytes of conv_plane_kernel for mesh side n.
template <typename R>
size_t conv_plane_lds(int n) {
    return (static_cast<size_t>(kColMax) + static_cast<size_t>(n) * (n + 1)) * sizeof(cx_t<R>);
}

// Factors (A, B) of the four-step column transform for mesh side n <= kColMax, A >= B, both
// lengths of dft_len, B as large as possible; A = 0 if none.
void four_step_factors(int n, int& A, int& B) {
    static const int len[] = {1, 2, 3, 4, 5, 6, 8, 9, 10, 12, 15, 16};
    A = B = 0;
    if (n < 1 || n > kColMax) return;  // the LDS tile and twiddle table hold kColMax rows
    for (int b : len)
        for (int a : len)
            if (a >= b && a * b == n && b > B) {
                A = a;
                B = b;
            }
}

// Y[k NCV + c NV + v, n] = e(kcv) x - sum_be V_k[be, v] sum_al conj(C_k[al, c]) w[(al norb + be) nvec + n][g(k)]
// thread per k, blockIdx.y = group of vectors (the band coefficients are loaded once per group)
template <typename R, int NC, int NV>
__global__ __launch_bounds__(256) void conv_y_kernel(int nk, int norb, int nvec, GLayout lay, const int* __restrict__ gk,
                                                     const cx_t<R>* __restrict__ Ct, const cx_t<R>* __restrict__ Vt,
                                                     const double* __restrict__ eval_k, const double* __restrict__ eval_kq,
                                                     double scissor, const cx_t<R>* __restrict__ W,
                                                     const double2* __restrict__ X, int64_t N, double2* __restrict__ Y) {
    constexpr int VG = conv_y_group<R>();
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = blockIdx.y * VG;
    if (k >= nk) return;
    constexpr int nb = NV + NC, ncv = NC * NV;
    const int ng = min(VG, nvec - n0);
    const int g = gk[k];
    cx_t<R> acc[VG][NC][NV];
#pragma unroll
    for (int q = 0; q < VG; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[q][c][v] = mk(R(0), R(0));
    // BE orbitals be per pass over al, so that C_k[al, c] is loaded once per pass instead of
    // once per be; every t(be) = sum_al conj(C) w and acc += V t(be) keeps its order (al, then
    // be increasing), so the result does not depend on BE
    constexpr int BE = conv_y_be_block<R>();
    for (int be0 = 0; be0 < norb; be0 += BE) {
        const int nbe = min(BE, norb - be0);
        cx_t<R> t[VG][BE][NC];
#pragma unroll
        for (int q = 0; q < VG; ++q)
#pragma unroll
            for (int j = 0; j < BE; ++j)
#pragma unroll
                for (int c = 0; c < NC; ++c) t[q][j][c] = mk(R(0), R(0));
        for (int al = 0; al < norb; ++al) {
            cx_t<R> ca[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) ca[c] = Ct[(static_cast<size_t>(c) * norb + al) * nk + k];
#pragma unroll
            for (int j = 0; j < BE; ++j) {
                if (j < nbe) {
                    const cx_t<R>* w = W + lay.at((static_cast<int64_t>(al) * norb + be0 + j) * nvec + n0, g);
#pragma unroll
                    for (int q = 0; q < VG; ++q) {
                        if (q < ng) {
                            const cx_t<R> wq = w[q * lay.bstride];
#pragma unroll
                            for (int c = 0; c < NC; ++c) cfma_conj(t[q][j][c], ca[c], wq);
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int j = 0; j < BE; ++j) {
            if (j < nbe) {
                const int be = be0 + j;
                cx_t<R> vb[NV];
#pragma unroll
                for (int v = 0; v < NV; ++v) vb[v] = Vt[(static_cast<size_t>(v) * norb + be) * nk + k];
#pragma unroll
                for (int q = 0; q < VG; ++q)
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int v = 0; v < NV; ++v) cfma(acc[q][c][v], vb[v], t[q][j][c]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < VG; ++q) {
        if (q < ng) {
            const int n = n0 + q;
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const int64_t row = static_cast<int64_t>(k) * ncv + c * NV + v;
                    const double e = (scissor + eval_kq[static_cast<size_t>(k) * nb + NV + c]) -
                                     eval_k[static_cast<size_t>(k) * nb + v];
                    const double2 x = X[row + n * N];
                    Y[row + n * N] = mk(e * x.x - static_cast<double>(acc[q][c][v].x),
                                        e * x.y - static_cast<double>(acc[q][c][v].y));
                }
        }
    }
}

// Launch the contraction kernels for the context's band counts (nc, nv in 1..4).
template <typename R, typename Fn>
void with_bands(const Context& c, Fn&& fn) {
#define XATU_CONV_BANDS(NC, NV)                          \
    if (c.nc == NC && c.nv == NV) {                       \
        fn(std::integral_constant<int, NC>{}, std::integral_constant<int, NV>{}); \
        return;                                          \
    }
    XATU_CONV_BANDS(1, 1) XATU_CONV_BANDS(1, 2) XATU_CONV_BANDS(1, 3) XATU_CONV_BANDS(1, 4)
    XATU_CONV_BANDS(2, 1) XATU_CONV_BANDS(2, 2) XATU_CONV_BANDS(2, 3) XATU_CONV_BANDS(2, 4)
    XATU_CONV_BANDS(3, 1) XATU_CONV_BANDS(3, 2) XATU_CONV_BANDS(3, 3) XATU_CONV_BANDS(3, 4)
    XATU_CONV_BANDS(4, 1) XATU_CONV_BANDS(4, 2) XATU_CONV_BANDS(4, 3) XATU_CONV_BANDS(4, 4)
#undef XATU_CONV_BANDS
    throw Error(XATU_ERR_INVALID_ARGUMENT, "convolution products support 1 <= nc, nv <= 4");
}

// ---- exchange term along the upper triangle (FP64) ----
// Block sums of A_a(i) = E2_a(i) x_i and B_a(i) = conj(E1_a(i)) x_i over kScanRows rows:
// part[((blk NA + a) 2 + s) nvec + n]
__global__ __launch_bounds__(kScanRows) void xscan_partial_kernel(int na, int ncv, const double2* __restrict__ e1,
                                                                  const double2* __restrict__ e2,
                                                                  const double2* __restrict__ X, int64_t N, int nvec,
                                                                  double2* __restrict__ part) {
    __shared__ double2 s[kScanRows];
    const int n = blockIdx.y;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kScanRows + threadIdx.x;
    const bool ok = i < N;
    const double2 x = ok ? X[i + n * N] : mk(0.0, 0.0);
    const int64_t k = ok ? i / ncv : 0;
    const int cv = ok ? static_cast<int>(i % ncv) : 0;
    for (int a = 0; a < na; ++a) {
        for (int sidx = 0; sidx < 2; ++sidx) {
            double2 v = mk(0.0, 0.0);
            if (ok) {
                const size_t off = (static_cast<size_t>(k) * na + a) * ncv + cv;
                v = sidx == 0 ? cmul(e2[off], x) : cmul(cconj(e1[off]), x);
            }
            s[threadIdx.x] = v;
            __syncthreads();
            for (int st = kScanRows / 2; st > 0; st >>= 1) {
                if (threadIdx.x < st) s[threadIdx.x] = cadd(s[threadIdx.x], s[threadIdx.x + st]);
                __syncthreads();
            }
            if (threadIdx.x == 0) part[((static_cast<size_t>(blockIdx.x) * na + a) * 2 + sidx) * nvec + n] = s[0];
            __syncthreads();
        }
    }
}

// Exclusive prefix over blocks, in place, one thread per series (a, s, n); also the totals.
__global__ void xscan_offsets_kernel(int nblk, int nseries, double2* __restrict__ part, double2* __restrict__ total) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nseries) return;
    double2 run = mk(0.0, 0.0);
    for (int b = 0; b < nblk; ++b) {
        const size_t i = static_cast<size_t>(b) * nseries + t;
        const double2 v = part[i];
        part[i] = run;
        run = cadd(run, v);
    }
    total[t] = run;
}

// y_i += sum_a [E1_a(i) (A_tot - sum_{j<=i} A_a(j)) + conj(E2_a(i)) sum_{j<i} B_a(j)] + Re(sum_a E1_a E2_a) x_i
__global__ __launch_bounds__(kScanRows) void xscan_apply_kernel(int na, int ncv, const double2* __restrict__ e1,
                                                                const double2* __restrict__ e2,
                                                                const double2* __restrict__ X, int64_t N, int nvec,
                                                                const double2* __restrict__ part,
                                                                const double2* __restrict__ total,
                                                                double2* __restrict__ Y) {
    __shared__ double2 s[kScanRows];
    const int n = blockIdx.y;
    const int64_t i = static_cast<int64_t>(blockIdx.x) * kScanRows + threadIdx.x;
    const bool ok = i < N;
    const double2 x = ok ? X[i + n * N] : mk(0.0, 0.0);
    const int64_t k = ok ? i / ncv : 0;
    const int cv = ok ? static_cast<int>(i % ncv) : 0;
    const int nseries = na * 2 * nvec;
    double2 y = mk(0.0, 0.0);
    double diag = 0.0;
    for (int a = 0; a < na; ++a) {
        const size_t off = (static_cast<size_t>(k) * na + a) * ncv + cv;
        const double2 E1 = ok ? e1[off] : mk(0.0, 0.0);
        const double2 E2 = ok ? e2[off] : mk(0.0, 0.0);
        diag += E1.x * E2.x - E1.y * E2.y;
        for (int sidx = 0; sidx < 2; ++sidx) {
            const double2 v = sidx == 0 ? cmul(E2, x) : cmul(cconj(E1), x);
            // inclusive scan within the block (Hillis-Steele, fixed order)
            s[threadIdx.x] = v;
            __syncthreads();
            for (int st = 1; st < kScanRows; st <<= 1) {
                const double2 add = threadIdx.x >= st ? s[threadIdx.x - st] : mk(0.0, 0.0);
                __syncthreads();
                s[threadIdx.x] = cadd(s[threadIdx.x], add);
                __syncthreads();
            }
            const size_t ser = (static_cast<size_t>(a) * 2 + sidx) * nvec + n;
            const double2 before = part[static_cast<size_t>(blockIdx.x) * nseries + ser];  // sum over earlier blocks
            const double2 incl = cadd(before, s[threadIdx.x]);
            if (sidx == 0) {
                const double2 after = csub(total[ser], incl);  // sum_{j>i} A_a(j)
                cfma(y, E1, after);
            } else {
                const double2 excl = csub(incl, v);            // sum_{j<i} B_a(j)
                cfma(y, cconj(E2), excl);
            }
            __syncthreads();
        }
    }
    if (ok) {
        double2 o = Y[i + n * N];
        o.x += y.x + diag * x.x;
        o.y += y.y + diag * x.y;
        Y[i + n * N] = o;
    }
}

// Pw[(n NA^2 + a + NA b) M + xi] = sum_{al in a, be in b} |G[((al norb + be) ns + n) M + xi]|^2
__global__ void power_sum_kernel(LatticeDesc lat, int ns, int64_t M, const double2* __restrict__ G,
                                 double2* __restrict__ Pw) {
    const int64_t xi = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int ab = blockIdx.y, n = blockIdx.z;
    if (xi >= M) return;
    const int a = ab % lat.natoms, b = ab / lat.natoms;
    double sum = 0.0;
    for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al)
        for (int be = lat.orb_begin[b]; be < lat.orb_begin[b + 1]; ++be) {
            const double2 z = G[((static_cast<int64_t>(al) * lat.norb + be) * ns + n) * M + xi];
            sum += z.x * z.x + z.y * z.y;
        }
    Pw[(static_cast<int64_t>(n) * lat.natoms * lat.natoms + ab) * M + xi] = mk(sum, 0.0);
}

// S[(n nmesh + q) NA^2 + ab] = W[(n NA^2 + ab) M + g(Delta(q))] / M with Delta_j = (q_j - m/2) mod m
__global__ void scatter_pair_density_kernel(LatticeDesc lat, int ns, const double2* __restrict__ W,
                                            double2* __restrict__ S) {
    const int64_t M = lat.nmesh;
    const int na2 = lat.natoms * lat.natoms;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= M * na2 * ns) return;
    const int ab = static_cast<int>(idx % na2);
    const int64_t q = (idx / na2) % M;
    const int n = static_cast<int>(idx / (na2 * M));
    const int m = lat.mesh_n;
    int64_t g = 0, mul = 1, rem = q;
    for (int j = 0; j < lat.ndim; ++j) {
        const int d = static_cast<int>(rem % m);
        rem /= m;
        g += static_cast<int64_t>(((d - m / 2) % m + m) % m) * mul;
        mul *= m;
    }
    const double2 w = W[(static_cast<int64_t>(n) * na2 + ab) * M + g];
    const double inv = 1.0 / static_cast<double>(M);
    S[idx] = mk(w.x * inv, w.y * inv);
}

template <typename R>
constexpr rocfft_precision fft_precision() {
    return sizeof(R) == 8 ? rocfft_precision_double : rocfft_precision_single;
}

}  // namespace

struct ConvState {
    int gen = -1;
    int64_t M = 0;
    DevBuf<int> d_gk;
    DevBuf<double2> P64;
    DevBuf<float2> P32;
    DevBuf<double2> G64, U64, Ct64, Vt64;
    DevBuf<float2> G32, U32, Ct32, Vt32;
    DevBuf<double2> part, total;
    DevBuf<double2> Mt;  // interaction table on the mesh (set-up scratch, kept between set-ups)
    size_t lds_limit = 0;  // LDS bytes a workgroup may use on this device (0: not queried yet)
    // A transform over the whole mesh: one rocFFT plan per direction (entry-major layout),
    // or for the row-interleaved layout two batched 1-D plans per direction (along o0,
    // then along o1)
    struct Plan {
        rocfft_plan fwd = nullptr, bwd = nullptr, fwd2 = nullptr, bwd2 = nullptr;
        rocfft_execution_info info = nullptr;
        void* work = nullptr;
    };
    std::map<std::tuple<int, int, int>, Plan> plans;  // (bytes per real, batch, row-interleaved)
    ~ConvState() {
        for (auto& kv : plans) {
            for (rocfft_plan q : {kv.second.fwd, kv.second.bwd, kv.second.fwd2, kv.second.bwd2})
                if (q) rocfft_plan_destroy(q);
            if (kv.second.info) rocfft_execution_info_destroy(kv.second.info);
            if (kv.second.work) (void)hipFree(kv.second.work);
        }
    }
};

void destroy_conv_state(ConvState* s) { delete s; }

// LDS a workgroup may use on the current device (queried once per state).
size_t conv_lds_limit(ConvState& cs) {
    if (cs.lds_limit == 0) {
        int dev = 0, v = 0;
        XATU_HIP(hipGetDevice(&dev));
        XATU_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cs.lds_limit = static_cast<size_t>(std::max(v, 1));
    }
    return cs.lds_limit;
}

namespace {

void fft_setup_once() {
    static std::once_flag flag;
    static rocfft_status st = rocfft_status_success;
    std::call_once(flag, [] { st = rocfft_setup(); });
    if (st != rocfft_status_success) throw Error(XATU_ERR_DEVICE, "rocfft_setup failed");
}

// Batched 1-D plan of length n with the given stride and distance.
template <typename R>
rocfft_plan strided_plan(rocfft_transform_type dir, size_t n, size_t count, size_t stride, size_t dist) {
    rocfft_plan_description d = nullptr;
    XATU_FFT(rocfft_plan_description_create(&d));
    size_t st[1] = {stride};
    rocfft_status r = rocfft_plan_description_set_data_layout(d, rocfft_array_type_complex_interleaved,
                                                             rocfft_array_type_complex_interleaved, nullptr, nullptr,
                                                             1, st, dist, 1, st, dist);
    rocfft_plan q = nullptr;
    if (r == rocfft_status_success) r = rocfft_plan_create(&q, rocfft_placement_inplace, dir, fft_precision<R>(), 1, &n, count, d);
    rocfft_plan_description_destroy(d);
    XATU_FFT(r);
    return q;
}

// Plans over `batch` mesh transforms; interleaved (2-D meshes only): the row-interleaved
// layout of GLayout.
template <typename R>
ConvState::Plan& get_plan(Context& c, ConvState& cs, int batch, bool interleaved = false) {
    const auto key = std::make_tuple(static_cast<int>(sizeof(R)), batch, interleaved ? 1 : 0);
    auto it = cs.plans.find(key);
    if (it != cs.plans.end()) return it->second;
    fft_setup_once();
    ConvState::Plan p;
    const size_t m = static_cast<size_t>(c.lat.mesh_n);
    size_t lengths[3] = {m, m, m};
    if (interleaved) {
        const size_t count = static_cast<size_t>(batch) * m;
        p.fwd = strided_plan<R>(rocfft_transform_type_complex_forward, m, count, 1, m);
        p.fwd2 = strided_plan<R>(rocfft_transform_type_complex_forward, m, count, count, 1);
        p.bwd = strided_plan<R>(rocfft_transform_type_complex_inverse, m, count, 1, m);
        p.bwd2 = strided_plan<R>(rocfft_transform_type_complex_inverse, m, count, count, 1);
    } else {
        XATU_FFT(rocfft_plan_create(&p.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward,
                                    fft_precision<R>(), c.lat.ndim, lengths, batch, nullptr));
        XATU_FFT(rocfft_plan_create(&p.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse,
                                    fft_precision<R>(), c.lat.ndim, lengths, batch, nullptr));
    }
    size_t w = 0;
    for (rocfft_plan q : {p.fwd, p.bwd, p.fwd2, p.bwd2}) {
        size_t wq = 0;
        if (q) XATU_FFT(rocfft_plan_get_work_buffer_size(q, &wq));
        w = std::max(w, wq);
    }
    XATU_FFT(rocfft_execution_info_create(&p.info));
    XATU_FFT(rocfft_execution_info_set_stream(p.info, c.stream));
    if (w > 0) {
        XATU_HIP(hipMalloc(&p.work, w));
        XATU_FFT(rocfft_execution_info_set_work_buffer(p.info, p.work, w));
    }
    return cs.plans.emplace(key, p).first->second;
}

// Kernel spectra P_ab = (1/M) backward FFT of the table on the mesh, and the grid
// index of every k-point; rebuilt after each preparation of the operator.
void ensure_conv(Context& c) {
    if (!c.conv) c.conv = std::unique_ptr<ConvState, void (*)(ConvState*)>(new ConvState, destroy_conv_state);
    ConvState& cs = *c.conv;
    if (cs.gen == c.prep_gen) return;
    if (c.lat.ndim < 1 || c.lat.ndim > 2) throw Error(XATU_ERR_INVALID_ARGUMENT, "convolution products need ndim 1 or 2");
    const int64_t M = c.lat.nmesh;
    cs.M = M;
    const int m = c.lat.mesh_n;
    std::vector<int> gk(c.nk);
    for (int k = 0; k < c.nk; ++k) {
        int64_t g = 0, mul = 1;
        for (int j = 0; j < c.lat.ndim; ++j) {
            int o = c.h_koff[static_cast<size_t>(k) * c.lat.ndim + j] % m;
            if (o < 0) o += m;
            g += o * mul;
            mul *= m;
        }
        gk[k] = static_cast<int>(g);
    }
    cs.d_gk.alloc(c.nk);
    XATU_HIP(hipMemcpyAsync(cs.d_gk.p, gk.data(), gk.size() * sizeof(int), hipMemcpyHostToDevice, c.stream));
    const int norb = c.lat.norb;
    cs.Ct64.alloc(static_cast<size_t>(c.nc) * norb * c.nk);
    cs.Vt64.alloc(static_cast<size_t>(c.nv) * norb * c.nk);
    cs.Ct32.alloc(cs.Ct64.n);
    cs.Vt32.alloc(cs.Vt64.n);
    const dim3 tg((c.nk + 255) / 256, (c.nc + c.nv) * norb);
    hipLaunchKernelGGL((transpose_stacks_kernel<double>), tg, dim3(256), 0, c.stream, c.nk, c.nv, c.nc, norb,
                       c.d_evec_k.p, c.d_evec_kq.p, cs.Ct64.p, cs.Vt64.p);
    hipLaunchKernelGGL((transpose_stacks_kernel<float>), tg, dim3(256), 0, c.stream, c.nk, c.nv, c.nc, norb,
                       c.d_evec_k.p, c.d_evec_kq.p, cs.Ct32.p, cs.Vt32.p);
    XATU_HIP(hipGetLastError());
    const int na2 = c.lat.natoms * c.lat.natoms;
    DevBuf<double2>& Mt = cs.Mt;
    Mt.alloc(static_cast<size_t>(na2) * M);
    hipLaunchKernelGGL(table_grid_kernel, dim3((Mt.n + 255) / 256), dim3(256), 0, c.stream, c.lat, c.d_ft.p, Mt.p);
    XATU_HIP(hipGetLastError());
    ConvState::Plan& p = get_plan<double>(c, cs, na2);
    void* buf[1] = {Mt.p};
    XATU_FFT(rocfft_execute(p.bwd, buf, nullptr, p.info));
    cs.P64.alloc(Mt.n);
    cs.P32.alloc(Mt.n);
    const double scale = 1.0 / static_cast<double>(M);
    hipLaunchKernelGGL((scale_spectrum_kernel<double>), dim3((Mt.n + 255) / 256), dim3(256), 0, c.stream, Mt.p,
                       static_cast<int64_t>(Mt.n), scale, cs.P64.p);
    hipLaunchKernelGGL((scale_spectrum_kernel<float>), dim3((Mt.n + 255) / 256), dim3(256), 0, c.stream, Mt.p,
                       static_cast<int64_t>(Mt.n), scale, cs.P32.p);
    XATU_HIP(hipGetLastError());
    cs.gen = c.prep_gen;
}

template <typename R>
void conv_product(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s, DevBuf<cx_t<R>>& G,
                  DevBuf<cx_t<R>>& U, const cx_t<R>* P, const cx_t<R>* Ct, const cx_t<R>* Vt) {
    ConvState& cs = *c.conv;
    const int norb = c.lat.norb;
    const int64_t M = cs.M;
    const int64_t N = c.N;
    const bool full = (static_cast<int64_t>(c.nk) == M);
    for (int v0 = 0; v0 < nvec; v0 += kBlockVectors) {
        const int nv = std::min(kBlockVectors, nvec - v0);
        const double2* Xb = X + static_cast<int64_t>(v0) * N;
        double2* Yb = Y + static_cast<int64_t>(v0) * N;
        const int batch = norb * norb * nv;
        // 2-D meshes: row-interleaved layout, two 1-D passes without transpositions
        const bool inter = c.lat.ndim == 2;
        const GLayout lay = inter ? GLayout{c.lat.mesh_n, c.lat.mesh_n, static_cast<int64_t>(batch) * c.lat.mesh_n}
                                  : GLayout{M, 0, 0};
        G.alloc(std::max<size_t>(G.n, static_cast<size_t>(batch) * M));
        U.alloc(std::max<size_t>(U.n, static_cast<size_t>(nv) * c.nc * norb * c.nk));
        if (!full) XATU_HIP(hipMemsetAsync(G.p, 0, static_cast<size_t>(batch) * M * sizeof(cx_t<R>), s));
        const dim3 kb((c.nk + 255) / 256);
        const int vg = (nv + conv_vec_group<R>() - 1) / conv_vec_group<R>();
        with_bands<R>(c, [&](auto ncc, auto nvc) {
            constexpr int NC = decltype(ncc)::value, NV = decltype(nvc)::value;
            hipLaunchKernelGGL((conv_u_kernel<R, NC, NV>), dim3(kb.x, vg), dim3(256), 0, s, c.nk, norb, nv, Vt, Xb, N,
                               U.p);
            hipLaunchKernelGGL((conv_z_kernel<R, NC>), dim3(kb.x, (norb + conv_orb_group<R>() - 1) / conv_orb_group<R>(), nv),
                               dim3(256), 0, s, c.nk, norb, nv, lay, cs.d_gk.p, Ct, U.p, G.p);
        });
        XATU_HIP(hipGetLastError());
        int fa = 0, fb = 0;
        four_step_factors(c.lat.mesh_n, fa, fb);
        const size_t plane_lds = conv_plane_lds<R>(c.lat.mesh_n);
        if (inter && fa > 0 && plane_lds <= conv_lds_limit(cs)) {
            // whole planes in LDS: G read and written once
            hipLaunchKernelGGL((conv_plane_kernel<R, plane_threads<R>()>), dim3(batch), dim3(plane_threads<R>()),
                               plane_lds, s, c.lat,
                               c.lat.mesh_n, fa, fb, batch, nv, M, P, G.p);
        } else {
            ConvState::Plan& p = get_plan<R>(c, cs, batch, inter);
            void* buf[1] = {G.p};
            XATU_FFT(rocfft_execute(p.fwd, buf, nullptr, p.info));
            if (inter && fa > 0) {
                const int64_t count = static_cast<int64_t>(batch) * c.lat.mesh_n;
                hipLaunchKernelGGL((conv_columns4_kernel<R>),
                                   dim3(static_cast<unsigned>((count + kFourStepCols - 1) / kFourStepCols)), dim3(512), 0,
                                   s, c.lat, c.lat.mesh_n, fa, fb, count, nv, M, P, G.p);
            } else {
                if (p.fwd2) XATU_FFT(rocfft_execute(p.fwd2, buf, nullptr, p.info));
                hipLaunchKernelGGL((conv_multiply_kernel<R>), dim3((M + 255) / 256, batch), dim3(256), 0, s, c.lat, nv,
                                   M, lay, P, G.p);
                if (p.bwd2) XATU_FFT(rocfft_execute(p.bwd2, buf, nullptr, p.info));
            }
            XATU_HIP(hipGetLastError());
            XATU_FFT(rocfft_execute(p.bwd, buf, nullptr, p.info));
        }
        XATU_HIP(hipGetLastError());
        with_bands<R>(c, [&](auto ncc, auto nvc) {
            constexpr int NC = decltype(ncc)::value, NV = decltype(nvc)::value;
            const int yg = (nv + conv_y_group<R>() - 1) / conv_y_group<R>();
            hipLaunchKernelGGL((conv_y_kernel<R, NC, NV>), dim3(kb.x, yg), dim3(256), 0, s, c.nk, norb, nv, lay,
                               cs.d_gk.p, Ct, Vt, c.d_eval_k.p, c.d_eval_kq.p, c.scissor, G.p, Xb, N, Yb);
        });
        XATU_HIP(hipGetLastError());
        if (c.inter.exchange) {
            const int na = c.lat.natoms;
            const int ncv = c.nc * c.nv;
            const int nblk = static_cast<int>((N + kScanRows - 1) / kScanRows);
            const int nseries = na * 2 * nv;
            cs.part.alloc(std::max<size_t>(cs.part.n, static_cast<size_t>(nblk) * nseries));
            cs.total.alloc(std::max<size_t>(cs.total.n, static_cast<size_t>(nseries)));
            hipLaunchKernelGGL(xscan_partial_kernel, dim3(nblk, nv), dim3(kScanRows), 0, s, na, ncv, c.d_e1.p,
                               c.d_e2.p, Xb, N, nv, cs.part.p);
            hipLaunchKernelGGL(xscan_offsets_kernel, dim3((nseries + 255) / 256), dim3(256), 0, s, nblk, nseries,
                               cs.part.p, cs.total.p);
            hipLaunchKernelGGL(xscan_apply_kernel, dim3(nblk, nv), dim3(kScanRows), 0, s, na, ncv, c.d_e1.p, c.d_e2.p,
                               Xb, N, nv, cs.part.p, cs.total.p, Yb);
            XATU_HIP(hipGetLastError());
        }
    }
}

}  // namespace

// Pair-density adjoint of the direct term for the Hellmann-Feynman gradients:
//   S_ab(q) = sum_k G_ab(k, k+q),  G_ab(k,k') = sum_{al in a, be in b} conj(z_ab(k)) z_ab(k'),
// with z_{al be}(k) = sum_{cv} C_kc,al conj(V_kv,be) psi_kcv (the vector the convolution
// product forms from x = psi). The sum over k is an autocorrelation on the mesh:
// S_ab(Delta) = (1/M) backward( sum_{al in a, be in b} |forward(z_{al be})|^2 )(Delta).
// Output layout S[(n nmesh + q) NA^2 + a + NA b] for the ns states of psi (N x ns).
void conv_pair_density(Context& c, const double2* psi, int ns, double2* S) {
    ensure_conv(c);
    ConvState& cs = *c.conv;
    const int norb = c.lat.norb;
    const int na2 = c.lat.natoms * c.lat.natoms;
    const int64_t M = cs.M;
    const int64_t N = c.N;
    hipStream_t s = c.stream;
    if (ns < 1 || ns > kBlockVectors) throw Error(XATU_ERR_INVALID_ARGUMENT, "pair density: 1..32 states per call");
    const int batch = norb * norb * ns;
    cs.G64.alloc(std::max<size_t>(cs.G64.n, static_cast<size_t>(batch) * M));
    cs.U64.alloc(std::max<size_t>(cs.U64.n, static_cast<size_t>(ns) * c.nc * norb * c.nk));
    if (static_cast<int64_t>(c.nk) != M) XATU_HIP(hipMemsetAsync(cs.G64.p, 0, static_cast<size_t>(batch) * M * sizeof(double2), s));
    const dim3 kb((c.nk + 255) / 256);
    with_bands<double>(c, [&](auto ncc, auto nvc) {
        constexpr int NC = decltype(ncc)::value, NV = decltype(nvc)::value;
        constexpr int VG = conv_vec_group<double>();
        hipLaunchKernelGGL((conv_u_kernel<double, NC, NV>), dim3(kb.x, (ns + VG - 1) / VG), dim3(256), 0, s, c.nk, norb,
                           ns, cs.Vt64.p, psi, N, cs.U64.p);
        hipLaunchKernelGGL((conv_z_kernel<double, NC>), dim3(kb.x, (norb + conv_orb_group<double>() - 1) / conv_orb_group<double>(), ns),
                           dim3(256), 0, s, c.nk, norb, ns, GLayout{M, 0, 0}, cs.d_gk.p, cs.Ct64.p, cs.U64.p,
                           cs.G64.p);
    });
    XATU_HIP(hipGetLastError());
    ConvState::Plan& pf = get_plan<double>(c, cs, batch);
    void* buf[1] = {cs.G64.p};
    XATU_FFT(rocfft_execute(pf.fwd, buf, nullptr, pf.info));
    DevBuf<double2> Pw;
    Pw.alloc(static_cast<size_t>(ns) * na2 * M);
    hipLaunchKernelGGL(power_sum_kernel, dim3((M + 255) / 256, na2, ns), dim3(256), 0, s, c.lat, ns, M, cs.G64.p, Pw.p);
    XATU_HIP(hipGetLastError());
    ConvState::Plan& pb = get_plan<double>(c, cs, na2 * ns);
    void* buf2[1] = {Pw.p};
    XATU_FFT(rocfft_execute(pb.bwd, buf2, nullptr, pb.info));
    hipLaunchKernelGGL(scatter_pair_density_kernel, dim3((Pw.n + 255) / 256), dim3(256), 0, s, c.lat, ns, Pw.p, S);
    XATU_HIP(hipGetLastError());
    XATU_HIP(hipStreamSynchronize(s));  // Pw goes out of scope
}

void launch_conv64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    ensure_conv(c);
    conv_product<double>(c, X, Y, nvec, s, c.conv->G64, c.conv->U64, c.conv->P64.p, c.conv->Ct64.p, c.conv->Vt64.p);
}

void launch_conv32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    ensure_conv(c);
    conv_product<float>(c, X, Y, nvec, s, c.conv->G32, c.conv->U32, c.conv->P32.p, c.conv->Ct32.p, c.conv->Vt32.p);
}

}  // namespace xatu_dev
