// Bethe-Salpeter matrix assembly (reference: ExcitonTB::BShamiltonian and
// realSpaceInteractionTerm, src/ExcitonTB.cpp:520-548, 840-921).
//
// Basis row r = k*(nc*nv) + c*nv + v (Exciton.tpp:128-147). For rows i <= j
//   D(i,j) = sum_ab rho^c_a M_ab(k'-k) rho^v_b,
//     rho^c_a = sum_{alpha in a} conj(C_{k+Q}[alpha,c]) C_{k'+Q}[alpha,c'],
//     rho^v_b = sum_{beta in b}  conj(V_{k'}[beta,v'])   V_k[beta,v],
//   X(i,j) = sum_ab rho1_a MQ_ab rho2_b = sum_a E1_k[a](c,v) E2_{k'}[a](c',v'),
//   H(i,j) = -(D - X) (i < j),  H(i,i) = scissor + E_c(k+Q) - E_v(k) - Re(D - X),
//   H(j,i) = conj(H(i,j)).
// The structured kernel works on tiles of TK x TK k-point pairs: the tile's
// coefficient vectors are staged in LDS; each thread owns one pair (k_row, k_col)
// and produces its nc*nv x nc*nv block. A pair with k_row > k_col evaluates the
// block of (k_col, k_row) with the very same instruction sequence as the thread
// that owns (k_col, k_row) and stores its conjugate transpose, so the stored matrix
// is exactly Hermitian while every store stays coalesced along the rows.
#include "context.hpp"
#include "cplx.hpp"

namespace xatu_dev {

namespace {

__device__ inline int mesh_index(const int* __restrict__ kcoord, int ka, int kb, int ndim, int n) {
    int q = 0, mul = 1;
    const int half = n / 2;
    for (int j = 0; j < ndim; ++j) {
        int d = kcoord[3 * kb + j] - kcoord[3 * ka + j] + half;
        d %= n;
        if (d < 0) d += n;
        q += d * mul;
        mul *= n;
    }
    return q;
}

// E1_k[a](c,v) = sum_{alpha in a} conj(CKQ_k[alpha,c]) VK_k[alpha,v]
// E2_k[a](c,v) = sum_b MQ_ab sum_{beta in b} conj(VK_k[beta,v]) CKQ_k[beta,c]
// layout: [(k*natoms + a)*(nc*nv) + c*nv + v]
__global__ void exchange_factors_kernel(LatticeDesc lat, int nk, int nv, int nc, const double2* __restrict__ evk,
                                        const double2* __restrict__ evkq, const double2* __restrict__ ftq,
                                        double2* __restrict__ e1, double2* __restrict__ e2) {
    const int ncv = nc * nv;
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= nk * ncv) return;
    const int k = idx / ncv;
    const int cv = idx % ncv;
    const int c = cv / nv, v = cv % nv;
    const int nb = nv + nc;
    const double2* vk = evk + static_cast<size_t>(k) * nb * lat.norb + static_cast<size_t>(v) * lat.norb;
    const double2* ck = evkq + static_cast<size_t>(k) * nb * lat.norb + static_cast<size_t>(nv + c) * lat.norb;
    double2 rho2[kMaxAtoms];
    for (int a = 0; a < lat.natoms; ++a) {
        double2 s1 = mk(0.0, 0.0), s2 = mk(0.0, 0.0);
        for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) {
            cfma_conj(s1, ck[al], vk[al]);
            cfma_conj(s2, vk[al], ck[al]);
        }
        e1[(static_cast<size_t>(k) * lat.natoms + a) * ncv + cv] = s1;
        rho2[a] = s2;
    }
    for (int a = 0; a < lat.natoms; ++a) {
        double2 s = mk(0.0, 0.0);
        for (int b = 0; b < lat.natoms; ++b) cfma(s, ftq[a + lat.natoms * b], rho2[b]);
        e2[(static_cast<size_t>(k) * lat.natoms + a) * ncv + cv] = s;
    }
}

template <typename R>
__device__ inline cx_t<R> to_cx(double2 v);
template <>
__device__ inline double2 to_cx<double>(double2 v) {
    return v;
}
template <>
__device__ inline float2 to_cx<float>(double2 v) {
    return make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
}

template <typename R>
__device__ inline void store_h(cx_t<R>* H, int64_t idx, cx_t<R> v) {
    H[idx] = v;
}

// Structured tile kernel. R: arithmetic type of the assembly (double or float).
template <typename R, int NA, int NC, int NV, int TK>
__global__ __launch_bounds__(TK* TK) void bse_tile_kernel(LatticeDesc lat, int nk, const double2* __restrict__ evk,
                                                          const double2* __restrict__ evkq,
                                                          const double* __restrict__ eval_k,
                                                          const double* __restrict__ eval_kq,
                                                          const int* __restrict__ kcoord,
                                                          const double2* __restrict__ ft,
                                                          const double2* __restrict__ e1,
                                                          const double2* __restrict__ e2, int exchange,
                                                          double scissor, cx_t<R>* __restrict__ H) {
    using C = cx_t<R>;
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    extern __shared__ unsigned char smem_raw[];
    C* s_coef = reinterpret_cast<C*>(smem_raw);  // [2*TK][norb][NB]: conduction first, then valence
    const int norb = lat.norb;
    const int Kr = blockIdx.x, Kc = blockIdx.y;
    const int t = threadIdx.x;

    // stage coefficients of the TK row k-points and the TK column k-points
    const int total = 2 * TK * norb * NB;
    for (int e = t; e < total; e += TK * TK) {
        const int al = e % norb;
        const int rest = e / norb;
        const int band = rest % NB;
        const int slot = rest / NB;
        const int k = (slot < TK) ? (Kr * TK + slot) : (Kc * TK + slot - TK);
        C val = mk(R(0), R(0));
        if (k < nk) {
            // band < NC: conduction column NV+band of the k+Q stack; else valence column band-NC of the k stack
            const double2 g = (band < NC)
                                  ? evkq[(static_cast<size_t>(k) * NB + (NV + band)) * norb + al]
                                  : evk[(static_cast<size_t>(k) * NB + (band - NC)) * norb + al];
            val = to_cx<R>(g);
        }
        s_coef[(static_cast<size_t>(slot) * norb + al) * NB + band] = val;
    }
    __syncthreads();

    const int tr = t % TK, tc = t / TK;
    const int kr = Kr * TK + tr, kc = Kc * TK + tc;
    if (kr >= nk || kc >= nk) return;
    const bool sw = kr > kc;
    const int ka = sw ? kc : kr;
    const int kb = sw ? kr : kc;
    const int sa = sw ? (TK + tc) : tr;
    const int sb = sw ? tr : (TK + tc);
    const C* ca_ = s_coef + static_cast<size_t>(sa) * norb * NB;
    const C* cb_ = s_coef + static_cast<size_t>(sb) * norb * NB;

    const int q = mesh_index(kcoord, ka, kb, lat.ndim, lat.mesh_n);
    C M[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b) M[a][b] = to_cx<R>(ft[static_cast<size_t>(q) * NA * NA + a + NA * b]);

    // rho^v_b(vb, va) = sum_{beta in b} conj(V_kb[beta, vb]) V_ka[beta, va]
    C rv[NA][NV][NV];
#pragma unroll
    for (int b = 0; b < NA; ++b) {
#pragma unroll
        for (int vb = 0; vb < NV; ++vb)
#pragma unroll
            for (int va = 0; va < NV; ++va) rv[b][vb][va] = mk(R(0), R(0));
        for (int be = lat.orb_begin[b]; be < lat.orb_begin[b + 1]; ++be) {
#pragma unroll
            for (int vb = 0; vb < NV; ++vb) {
                const C xb = cb_[be * NB + NC + vb];
#pragma unroll
                for (int va = 0; va < NV; ++va) cfma_conj(rv[b][vb][va], xb, ca_[be * NB + NC + va]);
            }
        }
    }

    const int64_t N = static_cast<int64_t>(nk) * NCV;
    const bool diagblk = (ka == kb);
#pragma unroll
    for (int cA = 0; cA < NC; ++cA) {
        // rho^c_a(cB) = sum_{alpha in a} conj(C_ka[alpha, cA]) C_kb[alpha, cB]
        C rc[NA][NC];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int cB = 0; cB < NC; ++cB) rc[a][cB] = mk(R(0), R(0));
            for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) {
                const C xa = ca_[al * NB + cA];
#pragma unroll
                for (int cB = 0; cB < NC; ++cB) cfma_conj(rc[a][cB], xa, cb_[al * NB + cB]);
            }
        }
        // T_b(cB) = sum_a rho^c_a(cB) M_ab
        C T[NA][NC];
#pragma unroll
        for (int b = 0; b < NA; ++b)
#pragma unroll
            for (int cB = 0; cB < NC; ++cB) {
                C s = mk(R(0), R(0));
#pragma unroll
                for (int a = 0; a < NA; ++a) cfma(s, rc[a][cB], M[a][b]);
                T[b][cB] = s;
            }
#pragma unroll
        for (int vA = 0; vA < NV; ++vA) {
            C ex1[NA];
            if (exchange) {
#pragma unroll
                for (int a = 0; a < NA; ++a)
                    ex1[a] = to_cx<R>(e1[(static_cast<size_t>(ka) * NA + a) * NCV + cA * NV + vA]);
            }
#pragma unroll
            for (int cB = 0; cB < NC; ++cB)
#pragma unroll
                for (int vB = 0; vB < NV; ++vB) {
                    C D = mk(R(0), R(0));
#pragma unroll
                    for (int b = 0; b < NA; ++b) cfma(D, T[b][cB], rv[b][vB][vA]);
                    C X = mk(R(0), R(0));
                    if (exchange) {
#pragma unroll
                        for (int a = 0; a < NA; ++a)
                            cfma(X, ex1[a], to_cx<R>(e2[(static_cast<size_t>(kb) * NA + a) * NCV + cB * NV + vB]));
                    }
                    const C DX = csub(D, X);
                    const int ia = cA * NV + vA, ib = cB * NV + vB;
                    const int64_t i = static_cast<int64_t>(ka) * NCV + ia;
                    const int64_t j = static_cast<int64_t>(kb) * NCV + ib;
                    if (diagblk) {
                        if (ia < ib) {
                            const C f = mk(-DX.x, -DX.y);
                            H[i + j * N] = f;
                            H[j + i * N] = cconj(f);
                        } else if (ia == ib) {
                            const double e = (scissor + eval_kq[static_cast<size_t>(ka) * NB + NV + cA]) -
                                             eval_k[static_cast<size_t>(ka) * NB + vA];
                            const double h = e / 2. - static_cast<double>(DX.x) / 2.;
                            H[i + i * N] = mk(R(h + h), R(0));
                        }
                    } else if (!sw) {
                        H[i + j * N] = mk(-DX.x, -DX.y);
                    } else {
                        H[j + i * N] = mk(-DX.x, DX.y);
                    }
                }
        }
    }
}

// Generic per-element kernel for an explicit basis list {vcol, ccol, k} (FP64).
__global__ void bse_basis_kernel(LatticeDesc lat, int nb, int nbasis, const int* __restrict__ basis,
                                 const double2* __restrict__ evk, const double2* __restrict__ evkq,
                                 const double* __restrict__ eval_k, const double* __restrict__ eval_kq,
                                 const int* __restrict__ kcoord, const double2* __restrict__ ft,
                                 const double2* __restrict__ ftq, int exchange, double scissor,
                                 double2* __restrict__ H) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    const int col = blockIdx.y;
    if (r >= nbasis) return;
    const int i = min(r, col), j = max(r, col);
    const int v = basis[3 * i + 0], c = basis[3 * i + 1], k = basis[3 * i + 2];
    const int v2 = basis[3 * j + 0], c2 = basis[3 * j + 1], k2 = basis[3 * j + 2];
    const int norb = lat.norb;
    const double2* cK = evkq + (static_cast<size_t>(k) * nb + c) * norb;     // coefsKQ
    const double2* vK = evk + (static_cast<size_t>(k) * nb + v) * norb;      // coefsK
    const double2* cK2 = evkq + (static_cast<size_t>(k2) * nb + c2) * norb;  // coefsK2Q
    const double2* vK2 = evk + (static_cast<size_t>(k2) * nb + v2) * norb;   // coefsK2
    const int q = mesh_index(kcoord, k, k2, lat.ndim, lat.mesh_n);
    const int na = lat.natoms;
    double2 r1[kMaxAtoms], r2[kMaxAtoms], x1[kMaxAtoms], x2[kMaxAtoms];
#pragma unroll
    for (int a = 0; a < kMaxAtoms; ++a) {
        r1[a] = r2[a] = x1[a] = x2[a] = mk(0.0, 0.0);
        if (a < na) {
            for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) {
                cfma_conj(r1[a], cK[al], cK2[al]);
                cfma_conj(r2[a], vK2[al], vK[al]);
                cfma_conj(x1[a], cK[al], vK[al]);
                cfma_conj(x2[a], vK2[al], cK2[al]);
            }
        }
    }
    const double2* M = ft + static_cast<size_t>(q) * na * na;
    double2 D = mk(0.0, 0.0), X = mk(0.0, 0.0);
#pragma unroll
    for (int a = 0; a < kMaxAtoms; ++a) {
        if (a < na) {
            double2 y = mk(0.0, 0.0), yx = mk(0.0, 0.0);
#pragma unroll
            for (int b = 0; b < kMaxAtoms; ++b) {
                if (b < na) {
                    cfma(y, M[a + na * b], r2[b]);
                    if (exchange) cfma(yx, ftq[a + na * b], x2[b]);
                }
            }
            cfma(D, r1[a], y);
            cfma(X, x1[a], yx);
        }
    }
    const double2 DX = csub(D, X);
    double2 out;
    if (i == j) {
        const double e = (scissor + eval_kq[static_cast<size_t>(k) * nb + c]) - eval_k[static_cast<size_t>(k) * nb + v];
        const double h = e / 2. - DX.x / 2.;
        out = mk(h + h, 0.0);
    } else {
        out = mk(-DX.x, -DX.y);
        if (r > col) out = cconj(out);
    }
    H[static_cast<int64_t>(r) + static_cast<int64_t>(col) * nbasis] = out;
}

// diag[i] = Re H(i,i) read from the resident FP64 matrix.
__global__ void diag_extract_kernel(const double2* __restrict__ H, int64_t N, double* __restrict__ diag) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < N) diag[i] = H[i + i * N].x;
}

// max |H32 - H64| per block of kb x kb rows/cols; one workgroup per block.
__global__ void compare_kernel(const double2* __restrict__ H64, const float2* __restrict__ H32, int64_t N, int rows,
                               int nblk, double* __restrict__ out) {
    __shared__ double s[256];
    const int bi = blockIdx.x, bj = blockIdx.y;
    const int64_t r0 = static_cast<int64_t>(bi) * rows, c0 = static_cast<int64_t>(bj) * rows;
    double m = 0.0;
    for (int64_t c = c0; c < min<int64_t>(c0 + rows, N); ++c) {
        for (int64_t r = r0 + threadIdx.x; r < min<int64_t>(r0 + rows, N); r += blockDim.x) {
            const double2 a = H64[r + c * N];
            const float2 b = H32[r + c * N];
            const double dr = a.x - static_cast<double>(b.x), di = a.y - static_cast<double>(b.y);
            m = fmax(m, sqrt(dr * dr + di * di));
        }
    }
    s[threadIdx.x] = m;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + st]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[bi + static_cast<size_t>(nblk) * bj] = s[0];
}

template <typename R, int NA, int NC, int NV>
bool try_launch(const Context& c, cx_t<R>* H, hipStream_t s) {
    if (c.lat.natoms != NA || c.nc != NC || c.nv != NV) return false;
    constexpr int NB = NC + NV;
    const size_t bytes16 = static_cast<size_t>(2 * 16) * c.lat.norb * NB * sizeof(cx_t<R>);
    if (bytes16 <= 64 * 1024) {
        constexpr int TK = 16;
        const int nT = (c.nk + TK - 1) / TK;
        hipLaunchKernelGGL((bse_tile_kernel<R, NA, NC, NV, TK>), dim3(nT, nT), dim3(TK * TK), bytes16, s, c.lat, c.nk,
                           c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p,
                           c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, H);
    } else {
        constexpr int TK = 8;
        const size_t bytes8 = static_cast<size_t>(2 * TK) * c.lat.norb * NB * sizeof(cx_t<R>);
        if (bytes8 > 160 * 1024) throw Error(XATU_ERR_INVALID_ARGUMENT, "too many orbitals for the tile kernel");
        const int nT = (c.nk + TK - 1) / TK;
        hipLaunchKernelGGL((bse_tile_kernel<R, NA, NC, NV, TK>), dim3(nT, nT), dim3(TK * TK), bytes8, s, c.lat, c.nk,
                           c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p,
                           c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, H);
    }
    XATU_HIP(hipGetLastError());
    return true;
}

template <typename R, int NA>
bool try_atoms(const Context& c, cx_t<R>* H, hipStream_t s) {
    return try_launch<R, NA, 1, 1>(c, H, s) || try_launch<R, NA, 2, 2>(c, H, s) || try_launch<R, NA, 3, 3>(c, H, s);
}

template <typename R>
bool dispatch_tiles(const Context& c, cx_t<R>* H, hipStream_t s) {
    return try_atoms<R, 1>(c, H, s) || try_atoms<R, 2>(c, H, s) || try_atoms<R, 3>(c, H, s) ||
           try_atoms<R, 4>(c, H, s);
}

}  // namespace

void launch_exchange_factors(const Context& c, hipStream_t s) {
    const int n = c.nk * c.nc * c.nv;
    hipLaunchKernelGGL(exchange_factors_kernel, dim3((n + 255) / 256), dim3(256), 0, s, c.lat, c.nk, c.nv, c.nc,
                       c.d_evec_k.p, c.d_evec_kq.p, c.d_ftq.p, c.d_e1.p, c.d_e2.p);
    XATU_HIP(hipGetLastError());
}

// Full product basis in the structured ordering, expressed as an explicit list.
static void full_basis(const Context& c, std::vector<int>& basis) {
    basis.resize(static_cast<size_t>(c.N) * 3);
    size_t r = 0;
    for (int k = 0; k < c.nk; ++k)
        for (int ic = 0; ic < c.nc; ++ic)
            for (int iv = 0; iv < c.nv; ++iv) {
                basis[3 * r + 0] = iv;
                basis[3 * r + 1] = c.nv + ic;
                basis[3 * r + 2] = k;
                ++r;
            }
}

void launch_build_basis(const Context& c, int nbasis, const int* d_basis, double2* H, hipStream_t s) {
    dim3 grid((nbasis + 255) / 256, nbasis);
    hipLaunchKernelGGL(bse_basis_kernel, grid, dim3(256), 0, s, c.lat, c.nb, nbasis, d_basis, c.d_evec_k.p,
                       c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p, c.d_ftq.p,
                       c.inter.exchange, c.scissor, H);
    XATU_HIP(hipGetLastError());
}

void launch_build_tiles64(const Context& c, double2* H, hipStream_t s) {
    if (dispatch_tiles<double>(c, H, s)) return;
    // Shapes without a specialised tile kernel use the generic element kernel.
    std::vector<int> basis;
    full_basis(c, basis);
    DevBuf<int> d_basis;
    d_basis.alloc(basis.size());
    XATU_HIP(hipMemcpyAsync(d_basis.p, basis.data(), basis.size() * sizeof(int), hipMemcpyHostToDevice, s));
    if (c.N > (1 << 30)) throw Error(XATU_ERR_INVALID_ARGUMENT, "basis too large for the generic kernel");
    launch_build_basis(c, static_cast<int>(c.N), d_basis.p, H, s);
    XATU_HIP(hipStreamSynchronize(s));
}

void launch_build_tiles32(const Context& c, float2* H, hipStream_t s) {
    if (dispatch_tiles<float>(c, H, s)) return;
    throw Error(XATU_ERR_INVALID_ARGUMENT,
                "FP32 assembly supports natoms <= 4 and nv == nc <= 3 (use the FP64 build for other shapes)");
}

void launch_diag_extract(const Context& c, const double2* H, double* diag, hipStream_t s) {
    hipLaunchKernelGGL(diag_extract_kernel, dim3((c.N + 255) / 256), dim3(256), 0, s, H, c.N, diag);
    XATU_HIP(hipGetLastError());
}

void launch_compare(const Context& c, int kblock, double* d_blockerr, hipStream_t s) {
    const int rows = kblock * c.nc * c.nv;
    const int nblk = static_cast<int>((c.N + rows - 1) / rows);
    hipLaunchKernelGGL(compare_kernel, dim3(nblk, nblk), dim3(256), 0, s, c.d_H64.p, c.d_H32.p, c.N, rows, nblk,
                       d_blockerr);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
