// Hellmann-Feynman gradients of exciton energies with respect to the screening
// parameters (r0, eps_m, eps_s), without forming dH.
//
// For a normalised eigenvector psi, dE/dtheta = psi^H (dH/dtheta) psi with
// dH = -(dD - dX). The direct term's formula is Hermitian under swapping the pair
// (the lattice sums satisfy M(-q) = conj M(q) with the reference's conjugation
// convention), so psi^H dD psi is the full double sum
//   sum_{k,k'} sum_ab dM_ab(k'-k) G_ab(k,k'),
//   G_ab(k,k') = sum_{c v c' v'} conj(psi_kcv) rho^c_a(ck,c'k') rho^v_b(v'k',vk) psi_k'c'v'.
// Grouping the pairs by q = k'-k gives the adjoint of the pair densities,
// S_ab(q) = sum_k G_ab(k, k+q) (one workgroup per mesh point q, deterministic);
// the adjoint of the lattice sum then maps S to the cell table,
//   Shat_p(R) = sum_q [S_ab(q) e^{iqR} + S_ba(q) e^{-iqR}]  (a < b),
//   Shat_p(R) = sum_q S_aa(q) e^{-iqR}                        (a = b),
// and psi^H dD psi = (1/T) sum_p sum_R dV_p(R) Shat_p(R) for any parameter, with
// dV from the analytic derivative of the potential (FP64). The exchange formula
// X(i,j) = sum_ab f_a(i) MQ_ab g_b(j) is not Hermitian when V(Q) has a complex
// diagonal (Q != 0 with a cell set that is not inversion-symmetric), so it follows
// the reference's construction H = A + A^H from the upper triangle:
//   psi^H dX psi = Re sum_ab dMQ_ab (2 P_ab + G_ab),
//   P_ab = sum_j (sum_{i<j} f_a(i)) g_b(j),  G_ab = sum_i f_a(i) g_b(i),
// with f_a(i) = conj(psi_i) rho1_a(i), g_b(j) = rho2_b(j) psi_j (rows on the GPU,
// the prefix sums on the host).
#include <cmath>
#include <vector>

#include "context.hpp"
#include "cplx.hpp"
#include "potential_dev.hpp"

namespace xatu_dev {

namespace {

constexpr int kHfThreads = 256;
constexpr int kHfStates = 4;  // states per pass of the pair-density adjoint

// dV/dtheta for pair p and cell R. which: 0 r0, 1 eps (m or s: identical).
__global__ void dpotential_table_kernel(LatticeDesc lat, PotDesc pot, const int* __restrict__ cells, int nR, int npairs,
                                        int which, const double* __restrict__ gl, double* __restrict__ out) {
    const int R = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (R >= nR || p >= npairs) return;
    int a = 0, rem = p;
    while (rem >= lat.natoms - a) {
        rem -= lat.natoms - a;
        ++a;
    }
    const int b = a + rem;
    const double r = cell_distance(lat, cells + 3 * R, a, b);
    double d = 0.0;
    if (pot.kind == XATU_POTENTIAL_KELDYSH) {
        d = (which == 0) ? keldysh_dr0(r, pot, gl) : -potential(r, pot) / (2.0 * pot.eps_bar);
    }
    out[static_cast<size_t>(p) * nR + R] = d;
}

// S[s][q][a + na b] for states [s0, s0 + ns): one workgroup per mesh point q.
// kmap[linear mesh offset] = k index or -1; psi: N x nstates column-major.
__global__ __launch_bounds__(kHfThreads) void pair_adjoint_kernel(LatticeDesc lat, int nk, int nv, int nc,
                                                                  const double2* __restrict__ evk,
                                                                  const double2* __restrict__ evkq,
                                                                  const int* __restrict__ kcoord,
                                                                  const int* __restrict__ kmap,
                                                                  const double2* __restrict__ psi, int64_t N, int s0,
                                                                  int ns, double2* __restrict__ S) {
    const int q = blockIdx.x;
    const int n = lat.mesh_n;
    const int na = lat.natoms;
    const int nb = nv + nc;
    const int ncv = nc * nv;
    const int norb = lat.norb;
    // mesh coordinates of q -> offset delta (q = (delta + n/2) mod n per axis)
    int dq[3] = {0, 0, 0};
    {
        int rem = q;
        for (int j = 0; j < lat.ndim; ++j) {
            dq[j] = rem % n - n / 2;
            rem /= n;
        }
    }
    __shared__ double2 red[kHfThreads];
    for (int a = 0; a < na; ++a) {
        for (int b = 0; b < na; ++b) {
            double2 acc[kHfStates];
#pragma unroll
            for (int st = 0; st < kHfStates; ++st) acc[st] = mk(0.0, 0.0);
            for (int k = threadIdx.x; k < nk; k += blockDim.x) {
                int lin = 0, mul = 1;
                for (int j = 0; j < lat.ndim; ++j) {
                    int o = (kcoord[3 * k + j] + dq[j]) % n;
                    if (o < 0) o += n;
                    lin += o * mul;
                    mul *= n;
                }
                const int k2 = kmap[lin];
                if (k2 < 0) continue;
                const double2* Ck = evkq + static_cast<size_t>(k) * nb * norb;
                const double2* Vk = evk + static_cast<size_t>(k) * nb * norb;
                const double2* Ck2 = evkq + static_cast<size_t>(k2) * nb * norb;
                const double2* Vk2 = evk + static_cast<size_t>(k2) * nb * norb;
                // G_ab = sum_{c,c'} rho^c_a(c,c') sum_{v,v'} conj(psi_kcv) rho^v_b(v',v) psi_k'c'v'
                for (int c = 0; c < nc; ++c) {
                    for (int c2 = 0; c2 < nc; ++c2) {
                        double2 rc = mk(0.0, 0.0);
                        for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al)
                            cfma_conj(rc, Ck[(nv + c) * norb + al], Ck2[(nv + c2) * norb + al]);
                        for (int v = 0; v < nv; ++v) {
                            for (int v2 = 0; v2 < nv; ++v2) {
                                double2 rv = mk(0.0, 0.0);
                                for (int be = lat.orb_begin[b]; be < lat.orb_begin[b + 1]; ++be)
                                    cfma_conj(rv, Vk2[v2 * norb + be], Vk[v * norb + be]);
                                const double2 w = cmul(rc, rv);
                                const int64_t i = static_cast<int64_t>(k) * ncv + c * nv + v;
                                const int64_t j = static_cast<int64_t>(k2) * ncv + c2 * nv + v2;
#pragma unroll
                                for (int st = 0; st < kHfStates; ++st) {
                                    if (st < ns) {
                                        const double2 pi = psi[i + static_cast<int64_t>(s0 + st) * N];
                                        const double2 pj = psi[j + static_cast<int64_t>(s0 + st) * N];
                                        double2 t = mk(0.0, 0.0);
                                        cfma_conj(t, pi, pj);
                                        cfma(acc[st], t, w);
                                    }
                                }
                            }
                        }
                    }
                }
            }
            // block reduction, one state at a time
            for (int st = 0; st < ns; ++st) {
                double2 mine = mk(0.0, 0.0);
#pragma unroll
                for (int u = 0; u < kHfStates; ++u)
                    if (u == st) mine = acc[u];
                red[threadIdx.x] = mine;
                __syncthreads();
                for (int stp = kHfThreads / 2; stp > 0; stp >>= 1) {
                    if (threadIdx.x < stp) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + stp]);
                    __syncthreads();
                }
                if (threadIdx.x == 0) S[(static_cast<size_t>(st) * lat.nmesh + q) * na * na + a + na * b] = red[0];
                __syncthreads();
            }
        }
    }
}

// Shat[s][p][R] from S (adjoint of the lattice sum), for ns states.
// grid: (ceil(nR/256), npairs, ns)
__global__ void adjoint_fourier_kernel(LatticeDesc lat, const int* __restrict__ cells, int nR, int npairs,
                                       const double2* __restrict__ S, const double2* __restrict__ tw,
                                       double* __restrict__ shat) {
    const int R = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    const int s = blockIdx.z;
    if (R >= nR) return;
    int a = 0, rem = p;
    while (rem >= lat.natoms - a) {
        rem -= lat.natoms - a;
        ++a;
    }
    const int b = a + rem;
    const int na = lat.natoms;
    const int n = lat.mesh_n;
    const int two_n = 2 * n;
    double sum = 0.0;
    for (int q = 0; q < lat.nmesh; ++q) {
        long long ph = 0;
        int rem2 = q;
        for (int j = 0; j < lat.ndim; ++j) {
            const int cj = rem2 % n;
            rem2 /= n;
            ph += static_cast<long long>(2 * cj - n + (n & 1)) * cells[3 * R + j];
        }
        int idx = static_cast<int>(ph % two_n);
        if (idx < 0) idx += two_n;
        const double2 w = tw[idx];  // exp(i q.R)
        const double2* Sq = S + (static_cast<size_t>(s) * lat.nmesh + q) * na * na;
        if (a < b) {
            const double2 sab = Sq[a + na * b], sba = Sq[b + na * a];
            // Re[S_ab e^{iqR} + S_ba e^{-iqR}]
            sum += (sab.x * w.x - sab.y * w.y) + (sba.x * w.x + sba.y * w.y);
        } else {
            const double2 saa = Sq[a + na * a];
            sum += saa.x * w.x + saa.y * w.y;  // Re[S_aa e^{-iqR}]
        }
    }
    shat[(static_cast<size_t>(s) * npairs + p) * nR + R] = sum;
}

// Exchange rows: fg[(s*N + i)*2*na + a] = f_a(i), [.. + na + a] = g_a(i) for state s.
// grid: (ceil(N/256), nstates)
__global__ void exchange_rows_kernel(LatticeDesc lat, int nk, int nv, int nc, const double2* __restrict__ evk,
                                     const double2* __restrict__ evkq, const double2* __restrict__ e1,
                                     const double2* __restrict__ psi, int64_t N, double2* __restrict__ fg) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    if (i >= N) return;
    const int na = lat.natoms;
    const int nb = nv + nc, ncv = nc * nv;
    const int k = static_cast<int>(i / ncv), cv = static_cast<int>(i % ncv);
    const int c = cv / nv, v = cv % nv;
    const double2 p = psi[i + static_cast<int64_t>(s) * N];
    const double2* vk = evk + (static_cast<size_t>(k) * nb + v) * lat.norb;
    const double2* ck = evkq + (static_cast<size_t>(k) * nb + nv + c) * lat.norb;
    double2* out = fg + (static_cast<size_t>(s) * N + i) * 2 * na;
    for (int a = 0; a < na; ++a) {
        double2 f = mk(0.0, 0.0);
        cfma_conj(f, p, e1[(static_cast<size_t>(k) * na + a) * ncv + cv]);
        double2 r2 = mk(0.0, 0.0);
        for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) cfma_conj(r2, vk[al], ck[al]);
        out[a] = f;
        out[na + a] = cmul(r2, p);
    }
}

}  // namespace

void hf_gradients(Context& c, int nstates, const double2* d_vec, int nparams, const int* params, double* grad_out) {
    if (!c.prepared) throw Error(XATU_ERR_LOGIC, "operator not prepared (call xatu_build_bse)");
    const int na = c.lat.natoms;
    const int npairs = na * (na + 1) / 2;
    const int nR = c.nR;
    const double T = static_cast<double>(c.inter.total_cells);
    hipStream_t s = c.stream;

    // k -> mesh-offset map
    const int n = c.lat.mesh_n;
    std::vector<int> kmap(c.lat.nmesh, -1);
    for (int k = 0; k < c.nk; ++k) {
        int lin = 0, mul = 1;
        for (int j = 0; j < c.lat.ndim; ++j) {
            int o = c.h_koff[k * c.lat.ndim + j] % n;
            if (o < 0) o += n;
            lin += o * mul;
            mul *= n;
        }
        kmap[lin] = k;
    }
    DevBuf<int> d_kmap;
    d_kmap.alloc(kmap.size());
    XATU_HIP(hipMemcpyAsync(d_kmap.p, kmap.data(), kmap.size() * sizeof(int), hipMemcpyHostToDevice, s));

    // pair-density adjoint S and its lattice adjoint Shat, kHfStates states at a time
    DevBuf<double2> S;
    S.alloc(static_cast<size_t>(kHfStates) * c.lat.nmesh * na * na);
    DevBuf<double> shat;
    shat.alloc(static_cast<size_t>(nstates) * npairs * nR);
    for (int s0 = 0; s0 < nstates; s0 += kHfStates) {
        const int ns = std::min(kHfStates, nstates - s0);
        if (c.lat.ndim <= 2) {
            // autocorrelation on the mesh with FFTs: O(norb^2 M log M) instead of O(nk^2) pairs
            conv_pair_density(c, d_vec + static_cast<int64_t>(s0) * c.N, ns, S.p);
        } else {
            hipLaunchKernelGGL(pair_adjoint_kernel, dim3(c.lat.nmesh), dim3(kHfThreads), 0, s, c.lat, c.nk, c.nv,
                               c.nc, c.d_evec_k.p, c.d_evec_kq.p, c.d_kcoord.p, d_kmap.p, d_vec, c.N, s0, ns, S.p);
        }
        hipLaunchKernelGGL(adjoint_fourier_kernel, dim3((nR + 255) / 256, npairs, ns), dim3(256), 0, s, c.lat,
                           c.d_cells.p, nR, npairs, S.p, c.d_twiddle.p,
                           shat.p + static_cast<size_t>(s0) * npairs * nR);
        XATU_HIP(hipGetLastError());
    }
    std::vector<double> h_shat(shat.n);
    XATU_HIP(hipMemcpyAsync(h_shat.data(), shat.p, shat.bytes(), hipMemcpyDeviceToHost, s));

    // exchange rows f, g (host prefix sums below)
    std::vector<double2> h_fg;
    if (c.inter.exchange) {
        DevBuf<double2> fg;
        fg.alloc(static_cast<size_t>(nstates) * c.N * 2 * na);
        hipLaunchKernelGGL(exchange_rows_kernel, dim3((c.N + 255) / 256, nstates), dim3(256), 0, s, c.lat, c.nk, c.nv,
                           c.nc, c.d_evec_k.p, c.d_evec_kq.p, c.d_e1.p, d_vec, c.N, fg.p);
        XATU_HIP(hipGetLastError());
        h_fg.resize(fg.n);
        XATU_HIP(hipMemcpyAsync(h_fg.data(), fg.p, fg.bytes(), hipMemcpyDeviceToHost, s));
        XATU_HIP(hipStreamSynchronize(s));
    }
    XATU_HIP(hipStreamSynchronize(s));
    std::vector<double> cellsR(static_cast<size_t>(nR) * 3, 0.0);
    for (int r = 0; r < nR; ++r)
        for (int j = 0; j < c.lat.ndim; ++j)
            for (int x = 0; x < 3; ++x) cellsR[3 * r + x] += c.h_cells[r * c.lat.ndim + j] * c.lat.bravais[3 * j + x];

    // W_ab = 2 P_ab + G_ab per state (triangular exchange sums, FP64 on the host)
    std::vector<double2> Wx(static_cast<size_t>(nstates) * na * na, mk(0.0, 0.0));
    if (c.inter.exchange) {
        for (int st = 0; st < nstates; ++st) {
            std::vector<double> Fr(na, 0.0), Fi(na, 0.0);
            std::vector<double> Pr(na * na, 0.0), Pi(na * na, 0.0), Gr(na * na, 0.0), Gi(na * na, 0.0);
            for (int64_t i = 0; i < c.N; ++i) {
                const double2* row = h_fg.data() + (static_cast<size_t>(st) * c.N + i) * 2 * na;
                for (int a = 0; a < na; ++a)
                    for (int b = 0; b < na; ++b) {
                        const double2 g = row[na + b];
                        Pr[a + na * b] += Fr[a] * g.x - Fi[a] * g.y;
                        Pi[a + na * b] += Fr[a] * g.y + Fi[a] * g.x;
                        const double2 f = row[a];
                        Gr[a + na * b] += f.x * g.x - f.y * g.y;
                        Gi[a + na * b] += f.x * g.y + f.y * g.x;
                    }
                for (int a = 0; a < na; ++a) {
                    Fr[a] += row[a].x;
                    Fi[a] += row[a].y;
                }
            }
            for (int e = 0; e < na * na; ++e)
                Wx[static_cast<size_t>(st) * na * na + e] = mk(2.0 * Pr[e] + Gr[e], 2.0 * Pi[e] + Gi[e]);
        }
    }
    DevBuf<double> dv;
    dv.alloc(static_cast<size_t>(npairs) * nR);
    std::vector<double> h_dv(dv.n), h_dvx(dv.n);
    // quadrature rule of the dV/dr0 evaluation
    static const std::vector<double> h_gl = [] {
        std::vector<double> r(2 * kGlNodes);
        gauss_legendre_rule(r.data());
        return r;
    }();
    DevBuf<double> gl;
    gl.alloc(h_gl.size());
    XATU_HIP(hipMemcpyAsync(gl.p, h_gl.data(), gl.bytes(), hipMemcpyHostToDevice, s));
    for (int ip = 0; ip < nparams; ++ip) {
        const int which = (params[ip] == XATU_PARAM_R0) ? 0 : 1;
        hipLaunchKernelGGL(dpotential_table_kernel, dim3((nR + 255) / 256, npairs), dim3(256), 0, s, c.lat, c.pot_d,
                           c.d_cells.p, nR, npairs, which, gl.p, dv.p);
        XATU_HIP(hipGetLastError());
        XATU_HIP(hipMemcpyAsync(h_dv.data(), dv.p, dv.bytes(), hipMemcpyDeviceToHost, s));
        if (c.inter.exchange) {
            XATU_HIP(hipStreamSynchronize(s));
            hipLaunchKernelGGL(dpotential_table_kernel, dim3((nR + 255) / 256, npairs), dim3(256), 0, s, c.lat,
                               c.pot_x, c.d_cells.p, nR, npairs, which, gl.p, dv.p);
            XATU_HIP(hipGetLastError());
            XATU_HIP(hipMemcpyAsync(h_dvx.data(), dv.p, dv.bytes(), hipMemcpyDeviceToHost, s));
        }
        XATU_HIP(hipStreamSynchronize(s));
        for (int st = 0; st < nstates; ++st) {
            // direct: psi^H dD psi = (1/T) sum_p sum_R dV_p(R) Shat_p(R)
            double dd = 0.0;
            const double* sh = h_shat.data() + static_cast<size_t>(st) * npairs * nR;
            for (size_t e = 0; e < static_cast<size_t>(npairs) * nR; ++e) dd += h_dv[e] * sh[e];
            dd /= T;
            // exchange: Re sum_ab dMQ_ab W_ab, with dMQ_ab = (1/T) sum_R dV_p(R) e^{iQR} (a < b),
            // dMQ_ba and dMQ_aa with e^{-iQR} (lower triangle and diagonal by conjugation)
            double dx = 0.0;
            if (c.inter.exchange) {
                const double2* W = Wx.data() + static_cast<size_t>(st) * na * na;
                int p = 0;
                for (int a = 0; a < na; ++a)
                    for (int b = a; b < na; ++b, ++p) {
                        double mr = 0.0, mi = 0.0;  // sum_R dV e^{iQR}
                        for (int r = 0; r < nR; ++r) {
                            const double* R = cellsR.data() + 3 * r;
                            const double ph = c.inter.Q[0] * R[0] + c.inter.Q[1] * R[1] + c.inter.Q[2] * R[2];
                            const double d = h_dvx[static_cast<size_t>(p) * nR + r];
                            mr += d * std::cos(ph);
                            mi += d * std::sin(ph);
                        }
                        if (a < b) {
                            const double2 wab = W[a + na * b], wba = W[b + na * a];
                            dx += (mr * wab.x - mi * wab.y) + (mr * wba.x + mi * wba.y);
                        } else {
                            const double2 waa = W[a + na * a];
                            dx += mr * waa.x + mi * waa.y;
                        }
                    }
                dx /= T;
            }
            grad_out[static_cast<size_t>(ip) * nstates + st] = -(dd - dx);
        }
    }
}

}  // namespace xatu_dev
