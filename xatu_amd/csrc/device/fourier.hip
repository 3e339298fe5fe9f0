// Lattice Fourier sums of the screened interaction (reference: motifFourierTransform /
// motifFTMatrix / initializeMotifFT, src/ExcitonTB.cpp:438-476, 695-697, 761-786).
//
//   V_ab(q) = (1/T) sum_R V(|R + tau_a - tau_b|) exp(i q.R),   a < b,
//   V_ba(q) = conj(V_ab(q)),  and V_aa(q) = conj of its sum (the reference fills the
//   lower triangle by conjugation, overwriting the diagonal with its conjugate)
//
// Design: the potential is evaluated once per (pair, R) into a table V_p(R) instead
// of once per (q, pair, R). On the Monkhorst-Pack mesh q.R = pi (sum_j t_j m_j) / n
// with integers t_j = 2 c_j - n + (n odd), so the phase is an exact 2n-th root of
// unity read from a table staged in LDS. One workgroup owns one mesh point; its
// threads stream R (coalesced table reads), accumulate with Neumaier compensation,
// and reduce pairwise through LDS.
#include "context.hpp"
#include "potential_dev.hpp"

namespace xatu_dev {

namespace {

constexpr int kFtThreads = 256;
constexpr int kFtPairs = 10;  // pairs handled per workgroup (natoms <= 4 -> one chunk)

__global__ void potential_table_kernel(LatticeDesc lat, PotDesc pot, const int* __restrict__ cells, int nR,
                                       int npairs, double* __restrict__ out) {
    const int R = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (R >= nR || p >= npairs) return;
    // pair index -> (a, b), a <= b, enumerated row by row
    int a = 0, rem = p;
    while (rem >= lat.natoms - a) {
        rem -= lat.natoms - a;
        ++a;
    }
    const int b = a + rem;
    out[static_cast<size_t>(p) * nR + R] = potential(cell_distance(lat, cells + 3 * R, a, b), pot);
}

__global__ void twiddle_kernel(int n, double2* __restrict__ tw) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= 2 * n) return;
    double s, c;
    sincospi(static_cast<double>(p) / static_cast<double>(n), &s, &c);
    tw[p] = make_double2(c, s);
}

__device__ inline void neumaier_add(double& sum, double& comp, double x) {
    const double t = sum + x;
    if (fabs(sum) >= fabs(x)) {
        comp += (sum - t) + x;
    } else {
        comp += (x - t) + sum;
    }
    sum = t;
}

// grid: (nmesh, ceil(npairs / kFtPairs)); block: kFtThreads
__global__ __launch_bounds__(kFtThreads) void fourier_mesh_kernel(LatticeDesc lat, const int* __restrict__ cells,
                                                                  int nR, int npairs,
                                                                  const double* __restrict__ vtab,
                                                                  const double2* __restrict__ tw,
                                                                  double inv_total, double2* __restrict__ ft) {
    __shared__ double2 s_tw[2 * 512];
    __shared__ double s_red[2][kFtThreads];
    const int q = blockIdx.x;
    const int p0 = blockIdx.y * kFtPairs;
    const int np = min(kFtPairs, npairs - p0);
    const int n = lat.mesh_n;
    const int two_n = 2 * n;
    for (int i = threadIdx.x; i < two_n; i += blockDim.x) s_tw[i] = tw[i];

    // mesh coordinates of q (first coordinate fastest) -> t_j
    int t[3] = {0, 0, 0};
    {
        int rem = q;
        for (int j = 0; j < lat.ndim; ++j) {
            const int cj = rem % n;
            rem /= n;
            t[j] = 2 * cj - n + (n & 1);
        }
    }
    __syncthreads();

    double sr[kFtPairs], si[kFtPairs], cr[kFtPairs], ci[kFtPairs];
#pragma unroll
    for (int k = 0; k < kFtPairs; ++k) sr[k] = si[k] = cr[k] = ci[k] = 0.0;

    for (int R = threadIdx.x; R < nR; R += blockDim.x) {
        long long ph = 0;
        for (int j = 0; j < lat.ndim; ++j) ph += static_cast<long long>(t[j]) * cells[3 * R + j];
        int idx = static_cast<int>(ph % two_n);
        if (idx < 0) idx += two_n;
        const double2 w = s_tw[idx];
#pragma unroll
        for (int k = 0; k < kFtPairs; ++k) {
            if (k < np) {
                const double v = vtab[static_cast<size_t>(p0 + k) * nR + R];
                neumaier_add(sr[k], cr[k], v * w.x);
                neumaier_add(si[k], ci[k], v * w.y);
            }
        }
    }

    for (int k = 0; k < np; ++k) {
        s_red[0][threadIdx.x] = sr[k] + cr[k];
        s_red[1][threadIdx.x] = si[k] + ci[k];
        __syncthreads();
        for (int stride = kFtThreads / 2; stride > 0; stride >>= 1) {
            if (threadIdx.x < stride) {
                s_red[0][threadIdx.x] += s_red[0][threadIdx.x + stride];
                s_red[1][threadIdx.x] += s_red[1][threadIdx.x + stride];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const int p = p0 + k;
            int a = 0, rem = p;
            while (rem >= lat.natoms - a) {
                rem -= lat.natoms - a;
                ++a;
            }
            const int b = a + rem;
            const double re = s_red[0][0] * inv_total;
            const double im = s_red[1][0] * inv_total;
            const size_t base = static_cast<size_t>(q) * lat.natoms * lat.natoms;
            // (a,b) holds the sum and (b,a) its conjugate; for a == b the conjugate is
            // what remains, as the reference assigns (b,a) after (a,b).
            if (a != b) ft[base + a + static_cast<size_t>(lat.natoms) * b] = make_double2(re, im);
            ft[base + b + static_cast<size_t>(lat.natoms) * a] = make_double2(re, -im);
        }
        __syncthreads();
    }
}

// V_ab(Q) for the exchange term at an arbitrary Q: phases from Cartesian Q.R.
// grid: ceil(npairs / kFtPairs); block: kFtThreads
__global__ __launch_bounds__(kFtThreads) void fourier_q_kernel(LatticeDesc lat, const int* __restrict__ cells, int nR,
                                                               int npairs, const double* __restrict__ vtab,
                                                               double qx, double qy, double qz, double inv_total,
                                                               double2* __restrict__ ftq) {
    __shared__ double s_red[2][kFtThreads];
    const int p0 = blockIdx.x * kFtPairs;
    const int np = min(kFtPairs, npairs - p0);
    double sr[kFtPairs], si[kFtPairs], cr[kFtPairs], ci[kFtPairs];
#pragma unroll
    for (int k = 0; k < kFtPairs; ++k) sr[k] = si[k] = cr[k] = ci[k] = 0.0;
    for (int R = threadIdx.x; R < nR; R += blockDim.x) {
        double v[3] = {0.0, 0.0, 0.0};
        for (int j = 0; j < lat.ndim; ++j) {
            const double m = static_cast<double>(cells[3 * R + j]);
            v[0] += m * lat.bravais[3 * j + 0];
            v[1] += m * lat.bravais[3 * j + 1];
            v[2] += m * lat.bravais[3 * j + 2];
        }
        double s, c;
        sincos(qx * v[0] + qy * v[1] + qz * v[2], &s, &c);
#pragma unroll
        for (int k = 0; k < kFtPairs; ++k) {
            if (k < np) {
                const double val = vtab[static_cast<size_t>(p0 + k) * nR + R];
                neumaier_add(sr[k], cr[k], val * c);
                neumaier_add(si[k], ci[k], val * s);
            }
        }
    }
    for (int k = 0; k < np; ++k) {
        s_red[0][threadIdx.x] = sr[k] + cr[k];
        s_red[1][threadIdx.x] = si[k] + ci[k];
        __syncthreads();
        for (int stride = kFtThreads / 2; stride > 0; stride >>= 1) {
            if (threadIdx.x < stride) {
                s_red[0][threadIdx.x] += s_red[0][threadIdx.x + stride];
                s_red[1][threadIdx.x] += s_red[1][threadIdx.x + stride];
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const int p = p0 + k;
            int a = 0, rem = p;
            while (rem >= lat.natoms - a) {
                rem -= lat.natoms - a;
                ++a;
            }
            const int b = a + rem;
            const double re = s_red[0][0] * inv_total;
            const double im = s_red[1][0] * inv_total;
            if (a != b) ftq[a + lat.natoms * b] = make_double2(re, im);
            ftq[b + lat.natoms * a] = make_double2(re, -im);
        }
        __syncthreads();
    }
}

// Reciprocal-space mode (src/ExcitonTB.cpp:408-426, 562-621). The direct term of a
// pair of k-points is
//   D = sum_ab rho^c_a M_ab(k - k') conj(rho^v'_b),
//   M_ab(q) = sum_G W(|q + G|) exp(i (q + G).(tau_a - tau_b)),
//   W(q) = 0 for q < |b1| / T, else pref / (q (1 + r0 q))     (keldyshFT),
// the exchange term uses M_ab(Q). M is not periodic on the mesh (the G set is cut
// off by |G|), so the table holds every raw difference k - k' of mesh points: offset
// d_j - (m - 1) per axis, in units of b_j / m. M_ba = conj(M_ab) because W is real.
struct RecipDesc {
    double pref;  // e 1e10 / (2 eps0 eps_bar A T)
    double r0;
    double qmin;  // |b1| / T
};

__device__ inline double2 reciprocal_sum(const LatticeDesc& lat, const double* q, int a, int b,
                                         const double* __restrict__ gvec, int ng, const RecipDesc& rd) {
    const double dx = lat.tau[3 * a + 0] - lat.tau[3 * b + 0];
    const double dy = lat.tau[3 * a + 1] - lat.tau[3 * b + 1];
    const double dz = lat.tau[3 * a + 2] - lat.tau[3 * b + 2];
    double re = 0.0, im = 0.0;
    for (int g = 0; g < ng; ++g) {
        const double x = q[0] + gvec[3 * g + 0];
        const double y = q[1] + gvec[3 * g + 1];
        const double z = q[2] + gvec[3 * g + 2];
        const double qn = sqrt(x * x + y * y + z * z);
        if (qn < rd.qmin) continue;
        const double w = rd.pref / (qn * (1.0 + rd.r0 * qn));
        double sn, cs;
        sincos(x * dx + y * dy + z * dz, &sn, &cs);
        re += w * cs;
        im += w * sn;
    }
    return make_double2(re, im);
}

__device__ inline void pair_of(int p, int natoms, int& a, int& b) {
    a = 0;
    int rem = p;
    while (rem >= natoms - a) {
        rem -= natoms - a;
        ++a;
    }
    b = a + rem;
}

// grid: (ceil(ntab / 256), npairs); block: 256
__global__ void reciprocal_table_kernel(LatticeDesc lat, const double* __restrict__ gvec, int ng, RecipDesc rd,
                                        int npairs, double2* __restrict__ ft) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int p = blockIdx.y;
    if (t >= lat.ntab || p >= npairs) return;
    int a, b;
    pair_of(p, lat.natoms, a, b);
    double q[3] = {0.0, 0.0, 0.0};
    int rem = t;
    for (int j = 0; j < lat.ndim; ++j) {
        const double f = static_cast<double>(rem % lat.tab_n - (lat.mesh_n - 1)) / lat.mesh_n;
        rem /= lat.tab_n;
        q[0] += f * lat.recip_vec[3 * j + 0];
        q[1] += f * lat.recip_vec[3 * j + 1];
        q[2] += f * lat.recip_vec[3 * j + 2];
    }
    const double2 m = reciprocal_sum(lat, q, a, b, gvec, ng, rd);
    const size_t base = static_cast<size_t>(t) * lat.natoms * lat.natoms;
    ft[base + a + static_cast<size_t>(lat.natoms) * b] = m;
    if (a != b) ft[base + b + static_cast<size_t>(lat.natoms) * a] = make_double2(m.x, -m.y);
}

// M_ab(Q) for the exchange term. grid: 1; block: 64 (npairs <= 36)
__global__ void reciprocal_q_kernel(LatticeDesc lat, const double* __restrict__ gvec, int ng, RecipDesc rd, int npairs,
                                    double qx, double qy, double qz, double2* __restrict__ ftq) {
    const int p = threadIdx.x;
    if (p >= npairs) return;
    int a, b;
    pair_of(p, lat.natoms, a, b);
    const double q[3] = {qx, qy, qz};
    const double2 m = reciprocal_sum(lat, q, a, b, gvec, ng, rd);
    ftq[a + lat.natoms * b] = m;
    if (a != b) ftq[b + lat.natoms * a] = make_double2(m.x, -m.y);
}

}  // namespace

void launch_reciprocal_table(const Context& c, const double* d_gvec, int ng, double pref, hipStream_t s) {
    const int npairs = c.lat.natoms * (c.lat.natoms + 1) / 2;
    RecipDesc rd{};
    rd.pref = pref;
    rd.r0 = c.inter.r0;
    const double* b1 = c.lat.recip_vec;
    rd.qmin = sqrt(b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2]) / static_cast<double>(c.inter.total_cells);
    dim3 grid((c.lat.ntab + 255) / 256, npairs);
    hipLaunchKernelGGL(reciprocal_table_kernel, grid, dim3(256), 0, s, c.lat, d_gvec, ng, rd, npairs, c.d_ft.p);
    XATU_HIP(hipGetLastError());
    if (c.inter.exchange) {
        hipLaunchKernelGGL(reciprocal_q_kernel, dim3(1), dim3(64), 0, s, c.lat, d_gvec, ng, rd, npairs, c.inter.Q[0],
                           c.inter.Q[1], c.inter.Q[2], c.d_ftq.p);
        XATU_HIP(hipGetLastError());
    }
}

void launch_potential_table(const Context& c, const PotDesc& pot, double* out, hipStream_t s) {
    const int npairs = c.lat.natoms * (c.lat.natoms + 1) / 2;
    dim3 grid((c.nR + 255) / 256, npairs);
    hipLaunchKernelGGL(potential_table_kernel, grid, dim3(256), 0, s, c.lat, pot, c.d_cells.p, c.nR, npairs, out);
    XATU_HIP(hipGetLastError());
}

void launch_twiddles(int mesh_n, double2* out, hipStream_t s) {
    hipLaunchKernelGGL(twiddle_kernel, dim3((2 * mesh_n + 255) / 256), dim3(256), 0, s, mesh_n, out);
    XATU_HIP(hipGetLastError());
}

void launch_fourier_mesh(const Context& c, hipStream_t s) {
    if (c.lat.mesh_n > 512) throw Error(XATU_ERR_INVALID_ARGUMENT, "mesh_n above 512 is not supported");
    const int npairs = c.lat.natoms * (c.lat.natoms + 1) / 2;
    dim3 grid(c.lat.nmesh, (npairs + kFtPairs - 1) / kFtPairs);
    const double inv_total = 1.0 / static_cast<double>(c.inter.total_cells);
    hipLaunchKernelGGL(fourier_mesh_kernel, grid, dim3(kFtThreads), 0, s, c.lat, c.d_cells.p, c.nR, npairs,
                       c.d_vtab.p, c.d_twiddle.p, inv_total, c.d_ft.p);
    XATU_HIP(hipGetLastError());
}

void launch_fourier_q(const Context& c, hipStream_t s) {
    const int npairs = c.lat.natoms * (c.lat.natoms + 1) / 2;
    const double inv_total = 1.0 / static_cast<double>(c.inter.total_cells);
    hipLaunchKernelGGL(fourier_q_kernel, dim3((npairs + kFtPairs - 1) / kFtPairs), dim3(kFtThreads), 0, s, c.lat,
                       c.d_cells.p, c.nR, npairs, c.d_vtabx.p, c.inter.Q[0], c.inter.Q[1], c.inter.Q[2], inv_total,
                       c.d_ftq.p);
    XATU_HIP(hipGetLastError());
}

}  // namespace xatu_dev
