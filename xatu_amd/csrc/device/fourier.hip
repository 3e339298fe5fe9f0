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
    double v[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < lat.ndim; ++j) {
        const double m = static_cast<double>(cells[3 * R + j]);
        v[0] += m * lat.bravais[3 * j + 0];
        v[1] += m * lat.bravais[3 * j + 1];
        v[2] += m * lat.bravais[3 * j + 2];
    }
    const double dx = v[0] + lat.tau[3 * a + 0] - lat.tau[3 * b + 0];
    const double dy = v[1] + lat.tau[3 * a + 1] - lat.tau[3 * b + 1];
    const double dz = v[2] + lat.tau[3 * a + 2] - lat.tau[3 * b + 2];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    out[static_cast<size_t>(p) * nR + R] = potential(r, pot);
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

}  // namespace

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
