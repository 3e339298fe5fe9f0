// Adjoint of the BSE assembly with respect to the band data: for an exciton state
// psi with energy E = psi^H H psi, the sensitivities
//   w_b(k) = dE / d eps_b(k)            (band energies of the stacks)
//   g_b(k) = dE / d conj(C_b(k))        (Wirtinger derivative, eigenvector coefficients)
// so that a perturbation of the band data changes E by
//   dE = sum [ w d(eps) + 2 Re( g^H dC ) ].
// With first-order perturbation theory of the tight-binding bands this gives the
// gradient with respect to every hopping amplitude (done by the host facade).
//
// The reference builds H from its upper triangle (src/ExcitonTB.cpp:840-921):
//   H_ij = -F_ij (i < j), H_ii = e_i - Re F_ii, H_ji = conj(H_ij), F = D - X,
// so E = sum_i |psi_i|^2 e_i - Re sum_{i<=j} p_ij F_ij with p_ij = w_ij conj(psi_i) psi_j,
// w_ii = 1, w_ij = 2. For a canonical pair (row k_a, column k_b, k_a <= k_b) with
// coefficient vectors A = C^{k+Q}_{c}(k_a), Q = C^{k}_{v}(k_a), B = C^{k+Q}_{c'}(k_b),
// P = C^{k}_{v'}(k_b):
//   D = sum_ab rho^c_a M_ab rho^v_b,  rho^c_a = sum_{al in a} conj(A_al) B_al,
//       rho^v_b = sum_{be in b} conj(P_be) Q_be,
//   X = sum_a E1_a(k_a; c, v) E2_a(k_b; c', v'),
// F depends on conj(A), conj(P) and on B, Q. With the pair tensors
//   Y_a(c,c')  = sum_{v,v'} p sum_b M_ab rho^v_b(v',v)
//   S_b(v',v)  = sum_{c,c'} p sum_a rho^c_a(c,c') M_ab
//   U_a(c,v)   = sum_{c',v'} p E2_a(k_b; c',v')
//   T_b(v',c') = sum_{c,v} p sum_a E1_a(k_a; c,v) MQ_ab
// the Wirtinger derivatives of -Re(sum p F) are, for orbital al of atom t:
//   g_A(c,al)   += -1/2 [ sum_c' Y_t(c,c') B_c'al      - sum_v  U_t(c,v) Q_v,al ]
//   g_P(v',al)  += -1/2 [ sum_v  S_t(v',v) Q_v,al      - sum_c' T_t(v',c') B_c'al ]
//   g_B(c',al)  += -1/2 [ sum_c  conj Y_t(c,c') A_c,al - sum_v' conj T_t(v',c') P_v',al ]
//   g_Q(v,al)   += -1/2 [ sum_v' conj S_t(v',v) P_v',al - sum_c conj U_t(c,v) A_c,al ]
// (A and P enter F conjugated, so their derivative is -1/2 p dF/d conj; B and Q enter
// holomorphically, so theirs is -1/2 conj(p dF/dB)). Every canonical pair is handled by
// one thread, which adds its contributions to the four coefficient vectors with
// hardware FP64 atomics. Conduction vectors belong to the k+Q stack, valence vectors
// to the k stack, as in the assembly.
#include <vector>

#include "context.hpp"
#include "cplx.hpp"
#include "kmesh.hpp"

namespace xatu_dev {

namespace {

__device__ inline void atomic_add_cx(double2* p, double2 v) {
    unsafeAtomicAdd(&p->x, v.x);
    unsafeAtomicAdd(&p->y, v.y);
}

template <int NA, int NC, int NV, int TK>
__global__ __launch_bounds__(TK* TK) void band_adjoint_kernel(LatticeDesc lat, int nk, const double2* __restrict__ evk,
                                                              const double2* __restrict__ evkq,
                                                              const int* __restrict__ kcoord,
                                                              const double2* __restrict__ ft,
                                                              const double2* __restrict__ ftq,
                                                              const double2* __restrict__ e1,
                                                              const double2* __restrict__ e2, int exchange,
                                                              const double2* __restrict__ psi,
                                                              double2* __restrict__ gK, double2* __restrict__ gKQ) {
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    extern __shared__ double2 s_coef[];  // 2 TK slots [norb][NB]: conduction (k+Q stack), then valence (k stack)
    const int Kr = blockIdx.x, Kc = blockIdx.y;
    if (Kr > Kc) return;  // canonical pairs only (uniform over the workgroup)
    const int norb = lat.norb;
    const int total = 2 * TK * norb * NB;
    for (int e = threadIdx.x; e < total; e += TK * TK) {
        const int al = e % norb;
        const int rest = e / norb;
        const int band = rest % NB;
        const int slot = rest / NB;
        const int k = (slot < TK) ? (Kr * TK + slot) : (Kc * TK + slot - TK);
        double2 val = mk(0.0, 0.0);
        if (k < nk) {
            val = (band < NC) ? evkq[(static_cast<size_t>(k) * NB + (NV + band)) * norb + al]
                              : evk[(static_cast<size_t>(k) * NB + (band - NC)) * norb + al];
        }
        s_coef[(static_cast<size_t>(slot) * norb + al) * NB + band] = val;
    }
    __syncthreads();
    const int tr = threadIdx.x % TK, tc = threadIdx.x / TK;
    const int ka = Kr * TK + tr, kb = Kc * TK + tc;
    if (ka >= nk || kb >= nk || ka > kb) return;
    const double2* ca = s_coef + static_cast<size_t>(tr) * norb * NB;
    const double2* cb = s_coef + static_cast<size_t>(TK + tc) * norb * NB;
    // A_{c,al} = ca[al NB + c], Q_{v,al} = ca[al NB + NC + v], B_{c,al} = cb[al NB + c], P_{v,al} = cb[al NB + NC + v]

    const int q = table_index(lat, kcoord, ka, kb);
    double2 M[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b) M[a][b] = ft[static_cast<size_t>(q) * NA * NA + a + NA * b];

    // p(cA vA, cB vB) = w conj(psi_i) psi_j over the canonical triangle
    double2 p[NCV][NCV];
#pragma unroll
    for (int ia = 0; ia < NCV; ++ia) {
        const double2 pi = cconj(psi[static_cast<size_t>(ka) * NCV + ia]);
#pragma unroll
        for (int ib = 0; ib < NCV; ++ib) {
            const double w = (ka < kb || ia < ib) ? 2.0 : (ia == ib ? 1.0 : 0.0);
            const double2 v = cmul(pi, psi[static_cast<size_t>(kb) * NCV + ib]);
            p[ia][ib] = mk(w * v.x, w * v.y);
        }
    }

    // Y_a(c,c') = sum_{v,v'} p sum_b M_ab rho^v_b(v',v);  rho^v_b(v',v) = sum_be conj(P_v'be) Q_vbe
    double2 Y[NA][NC][NC], S[NA][NV][NV];
    {
        double2 MV[NA][NV][NV];
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int vb = 0; vb < NV; ++vb)
#pragma unroll
                for (int va = 0; va < NV; ++va) MV[a][vb][va] = mk(0.0, 0.0);
#pragma unroll
        for (int b = 0; b < NA; ++b) {
            double2 rv[NV][NV];
#pragma unroll
            for (int vb = 0; vb < NV; ++vb)
#pragma unroll
                for (int va = 0; va < NV; ++va) rv[vb][va] = mk(0.0, 0.0);
            for (int be = lat.orb_begin[b]; be < lat.orb_begin[b + 1]; ++be) {
#pragma unroll
                for (int vb = 0; vb < NV; ++vb) {
                    const double2 xb = cb[be * NB + NC + vb];
#pragma unroll
                    for (int va = 0; va < NV; ++va) cfma_conj(rv[vb][va], xb, ca[be * NB + NC + va]);
                }
            }
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int vb = 0; vb < NV; ++vb)
#pragma unroll
                    for (int va = 0; va < NV; ++va) cfma(MV[a][vb][va], M[a][b], rv[vb][va]);
        }
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                for (int cB = 0; cB < NC; ++cB) {
                    double2 s = mk(0.0, 0.0);
#pragma unroll
                    for (int vA = 0; vA < NV; ++vA)
#pragma unroll
                        for (int vB = 0; vB < NV; ++vB) cfma(s, p[cA * NV + vA][cB * NV + vB], MV[a][vB][vA]);
                    Y[a][cA][cB] = s;
                }
    }
    // S_b(v',v) = sum_{c,c'} p sum_a rho^c_a(c,c') M_ab;  rho^c_a(c,c') = sum_al conj(A_cal) B_c'al
    {
        double2 CM[NA][NC][NC];
#pragma unroll
        for (int b = 0; b < NA; ++b)
#pragma unroll
            for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                for (int cB = 0; cB < NC; ++cB) CM[b][cA][cB] = mk(0.0, 0.0);
#pragma unroll
        for (int a = 0; a < NA; ++a) {
            double2 rc[NC][NC];
#pragma unroll
            for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                for (int cB = 0; cB < NC; ++cB) rc[cA][cB] = mk(0.0, 0.0);
            for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) {
#pragma unroll
                for (int cA = 0; cA < NC; ++cA) {
                    const double2 xa = ca[al * NB + cA];
#pragma unroll
                    for (int cB = 0; cB < NC; ++cB) cfma_conj(rc[cA][cB], xa, cb[al * NB + cB]);
                }
            }
#pragma unroll
            for (int b = 0; b < NA; ++b)
#pragma unroll
                for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                    for (int cB = 0; cB < NC; ++cB) cfma(CM[b][cA][cB], rc[cA][cB], M[a][b]);
        }
#pragma unroll
        for (int b = 0; b < NA; ++b)
#pragma unroll
            for (int vB = 0; vB < NV; ++vB)
#pragma unroll
                for (int vA = 0; vA < NV; ++vA) {
                    double2 s = mk(0.0, 0.0);
#pragma unroll
                    for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                        for (int cB = 0; cB < NC; ++cB) cfma(s, p[cA * NV + vA][cB * NV + vB], CM[b][cA][cB]);
                    S[b][vB][vA] = s;
                }
    }
    // exchange tensors U_a(c,v) and T_b(v',c')
    double2 U[NA][NC][NV], T[NA][NV][NC];
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
        for (int cA = 0; cA < NC; ++cA)
#pragma unroll
            for (int vA = 0; vA < NV; ++vA) U[a][cA][vA] = mk(0.0, 0.0);
#pragma unroll
        for (int vB = 0; vB < NV; ++vB)
#pragma unroll
            for (int cB = 0; cB < NC; ++cB) T[a][vB][cB] = mk(0.0, 0.0);
    }
    if (exchange) {
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                for (int vA = 0; vA < NV; ++vA) {
                    double2 s = mk(0.0, 0.0);
#pragma unroll
                    for (int ib = 0; ib < NCV; ++ib)
                        cfma(s, p[cA * NV + vA][ib], e2[(static_cast<size_t>(kb) * NA + a) * NCV + ib]);
                    U[a][cA][vA] = s;
                }
#pragma unroll
        for (int ia = 0; ia < NCV; ++ia) {
            double2 e1m[NA];  // sum_a E1_a(k_a; ia) MQ_ab
#pragma unroll
            for (int b = 0; b < NA; ++b) e1m[b] = mk(0.0, 0.0);
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                const double2 x = e1[(static_cast<size_t>(ka) * NA + a) * NCV + ia];
#pragma unroll
                for (int b = 0; b < NA; ++b) cfma(e1m[b], x, ftq[a + NA * b]);
            }
#pragma unroll
            for (int b = 0; b < NA; ++b)
#pragma unroll
                for (int vB = 0; vB < NV; ++vB)
#pragma unroll
                    for (int cB = 0; cB < NC; ++cB) cfma(T[b][vB][cB], p[ia][cB * NV + vB], e1m[b]);
        }
    }

    // contributions to the coefficient gradients
    double2* gA = gKQ + static_cast<size_t>(ka) * NB * norb;  // conduction columns NV.. of the k+Q stack
    double2* gB = gKQ + static_cast<size_t>(kb) * NB * norb;
    double2* gQ = gK + static_cast<size_t>(ka) * NB * norb;   // valence columns 0.. of the k stack
    double2* gP = gK + static_cast<size_t>(kb) * NB * norb;
#pragma unroll
    for (int t = 0; t < NA; ++t) {
        for (int al = lat.orb_begin[t]; al < lat.orb_begin[t + 1]; ++al) {
            double2 Aal[NC], Bal[NC], Qal[NV], Pal[NV];
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                Aal[c] = ca[al * NB + c];
                Bal[c] = cb[al * NB + c];
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                Qal[v] = ca[al * NB + NC + v];
                Pal[v] = cb[al * NB + NC + v];
            }
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                double2 s = mk(0.0, 0.0);
#pragma unroll
                for (int c2 = 0; c2 < NC; ++c2) cfma(s, Y[t][c][c2], Bal[c2]);
#pragma unroll
                for (int v = 0; v < NV; ++v) s = csub(s, cmul(U[t][c][v], Qal[v]));
                atomic_add_cx(&gA[(NV + c) * norb + al], mk(-0.5 * s.x, -0.5 * s.y));
            }
#pragma unroll
            for (int c2 = 0; c2 < NC; ++c2) {
                double2 s = mk(0.0, 0.0);
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(s, cconj(Y[t][c][c2]), Aal[c]);
#pragma unroll
                for (int v2 = 0; v2 < NV; ++v2) s = csub(s, cmul(cconj(T[t][v2][c2]), Pal[v2]));
                atomic_add_cx(&gB[(NV + c2) * norb + al], mk(-0.5 * s.x, -0.5 * s.y));
            }
#pragma unroll
            for (int v2 = 0; v2 < NV; ++v2) {
                double2 s = mk(0.0, 0.0);
#pragma unroll
                for (int v = 0; v < NV; ++v) cfma(s, S[t][v2][v], Qal[v]);
#pragma unroll
                for (int c2 = 0; c2 < NC; ++c2) s = csub(s, cmul(T[t][v2][c2], Bal[c2]));
                atomic_add_cx(&gP[v2 * norb + al], mk(-0.5 * s.x, -0.5 * s.y));
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                double2 s = mk(0.0, 0.0);
#pragma unroll
                for (int v2 = 0; v2 < NV; ++v2) cfma(s, cconj(S[t][v2][v]), Pal[v2]);
#pragma unroll
                for (int c = 0; c < NC; ++c) s = csub(s, cmul(cconj(U[t][c][v]), Aal[c]));
                atomic_add_cx(&gQ[v * norb + al], mk(-0.5 * s.x, -0.5 * s.y));
            }
        }
    }
}

template <int NA, int NC, int NV>
bool launch_shape(const Context& c, const double2* psi, double2* gK, double2* gKQ, hipStream_t s) {
    if (c.lat.natoms != NA || c.nc != NC || c.nv != NV) return false;
    constexpr int TK = 8;
    const size_t lds = static_cast<size_t>(2 * TK) * c.lat.norb * (NC + NV) * sizeof(double2);
    if (lds > 160 * 1024) throw Error(XATU_ERR_INVALID_ARGUMENT, "too many orbitals for the band adjoint kernel");
    const int nT = (c.nk + TK - 1) / TK;
    hipLaunchKernelGGL((band_adjoint_kernel<NA, NC, NV, TK>), dim3(nT, nT), dim3(TK * TK), lds, s, c.lat, c.nk,
                       c.d_evec_k.p, c.d_evec_kq.p, c.d_kcoord.p, c.d_ft.p, c.d_ftq.p, c.d_e1.p, c.d_e2.p,
                       c.inter.exchange, psi, gK, gKQ);
    XATU_HIP(hipGetLastError());
    return true;
}

bool launch_band_adjoint(const Context& c, const double2* psi, double2* gK, double2* gKQ, hipStream_t s) {
#define XATU_BA_SHAPE(NA, NC, NV) \
    if (launch_shape<NA, NC, NV>(c, psi, gK, gKQ, s)) return true;
    XATU_BA_SHAPE(1, 1, 1)
    XATU_BA_SHAPE(1, 2, 2)
    XATU_BA_SHAPE(2, 1, 1)
    XATU_BA_SHAPE(2, 2, 2)
    XATU_BA_SHAPE(3, 1, 1)
    XATU_BA_SHAPE(3, 2, 2)
    XATU_BA_SHAPE(4, 1, 1)
    XATU_BA_SHAPE(4, 2, 2)
    XATU_BA_SHAPE(4, 3, 3)
#undef XATU_BA_SHAPE
    return false;
}

}  // namespace

void band_gradients(Context& c, int nstates, const double2* d_vec, double* w_out, double* g_out) {
    if (!c.prepared) throw Error(XATU_ERR_LOGIC, "operator not prepared (call xatu_build_bse)");
    const size_t stack = static_cast<size_t>(c.lat.norb) * c.nb * c.nk;  // complex entries per stack
    const int64_t N = c.N;
    hipStream_t s = c.stream;
    DevBuf<double2> g;
    g.alloc(2 * stack);
    std::vector<double2> psi(static_cast<size_t>(N) * nstates);
    XATU_HIP(hipMemcpyAsync(psi.data(), d_vec, psi.size() * sizeof(double2), hipMemcpyDeviceToHost, s));
    for (int st = 0; st < nstates; ++st) {
        XATU_HIP(hipMemsetAsync(g.p, 0, g.bytes(), s));
        if (!launch_band_adjoint(c, d_vec + static_cast<int64_t>(st) * N, g.p, g.p + stack, s)) {
            throw Error(XATU_ERR_INVALID_ARGUMENT,
                        "band gradients support natoms <= 4 and nv == nc <= 2 (nv == nc == 3 with 4 atoms)");
        }
        XATU_HIP(hipMemcpyAsync(g_out + 2 * (2 * stack) * st, g.p, g.bytes(), hipMemcpyDeviceToHost, s));
    }
    XATU_HIP(hipStreamSynchronize(s));
    // energy sensitivities: dE/dE_c(k+Q) = sum_v |psi|^2, dE/dE_v(k) = -sum_c |psi|^2
    const int nb = c.nb, nv = c.nv, nc = c.nc, ncv = nc * nv;
    for (int st = 0; st < nstates; ++st) {
        double* wk = w_out + static_cast<size_t>(st) * 2 * nb * c.nk;
        double* wkq = wk + static_cast<size_t>(nb) * c.nk;
        for (size_t i = 0; i < static_cast<size_t>(2) * nb * c.nk; ++i) wk[i] = 0.0;
        for (int k = 0; k < c.nk; ++k)
            for (int ic = 0; ic < nc; ++ic)
                for (int iv = 0; iv < nv; ++iv) {
                    const double2 z = psi[static_cast<size_t>(st) * N + static_cast<size_t>(k) * ncv + ic * nv + iv];
                    const double a2 = z.x * z.x + z.y * z.y;
                    wkq[static_cast<size_t>(k) * nb + nv + ic] += a2;
                    wk[static_cast<size_t>(k) * nb + iv] -= a2;
                }
    }
}

}  // namespace xatu_dev
