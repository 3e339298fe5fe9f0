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
#include "kmesh.hpp"
#include "mfma.hpp"

namespace xatu_dev {

namespace {

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

// Stage the coefficient vectors of TK row k-points (slots [0,TK)) and TK column
// k-points (slots [TK,2TK)) in LDS, slot fastest: entry (orbital al, band) of slot s at
// s_coef[(al NB + band) 2 TK + s], conduction bands first (from the k+Q stack), then the
// valence bands (from the k stack). The threads of a wave own consecutive row slots, so
// their reads of one entry fall on consecutive LDS words.
template <typename R, int NC, int NV, int TK>
__device__ inline void stage_coefficients(cx_t<R>* s_coef, int norb, int nk, int Kr, int Kc,
                                          const double2* __restrict__ evk, const double2* __restrict__ evkq,
                                          int nthreads) {
    constexpr int NB = NC + NV;
    const int total = 2 * TK * norb * NB;
    for (int e = threadIdx.x; e < total; e += nthreads) {
        const int al = e % norb;
        const int rest = e / norb;
        const int band = rest % NB;
        const int slot = rest / NB;
        const int k = (slot < TK) ? (Kr * TK + slot) : (Kc * TK + slot - TK);
        cx_t<R> val = mk(R(0), R(0));
        if (k < nk) {
            const double2 g = (band < NC) ? evkq[(static_cast<size_t>(k) * NB + (NV + band)) * norb + al]
                                          : evk[(static_cast<size_t>(k) * NB + (band - NC)) * norb + al];
            val = to_cx<R>(g);
        }
        s_coef[(static_cast<size_t>(al) * NB + band) * (2 * TK) + slot] = val;
    }
}

template <typename R>
__device__ inline cx_t<R> table_cx(double2 v) { return to_cx<R>(v); }
template <typename R>
__device__ inline cx_t<R> table_cx(float2 v) { return mk(static_cast<R>(v.x), static_cast<R>(v.y)); }

// The nc*nv x nc*nv block of (D - X) for the k-point pair (ka, kb), ka <= kb, from the
// staged coefficients ca_ (ka) and cb_ (kb): entry (orbital al, band) of a k-point at
// ptr[(al NB + band) cs]. sink(cA, vA, cB, vB, DX) receives every element, (cB, vB) fastest,
// or with COLS_OUTER (cA, vA) fastest (the order in which a pair's rows of one column of H
// are stored together). Each element's instruction sequence depends only on the data, not
// on that order, so any two threads evaluating the same pair produce bit-identical values.
// ft: the lattice-sum table, entry (a, b) of table row q at ft[q fts + a + NA b] (FP64, or the
// FP32 copy of the FP32 build with rows padded to an even length).
template <typename R, int NA, int NC, int NV, bool COLS_OUTER = false, typename FT, typename Sink>
__device__ inline void pair_block(const LatticeDesc& lat, const cx_t<R>* ca_, const cx_t<R>* cb_, int cs, int ka, int kb,
                                  const FT* __restrict__ ft, int fts, const int* __restrict__ kcoord,
                                  const double2* __restrict__ e1, const double2* __restrict__ e2, int exchange,
                                  Sink&& sink) {
    using C = cx_t<R>;
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    const int q = table_index(lat, kcoord, ka, kb);
    C M[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int b = 0; b < NA; ++b) M[a][b] = table_cx<R>(ft[static_cast<size_t>(q) * fts + a + NA * b]);

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
                const C xb = cb_[(be * NB + NC + vb) * cs];
#pragma unroll
                for (int va = 0; va < NV; ++va) cfma_conj(rv[b][vb][va], xb, ca_[(be * NB + NC + va) * cs]);
            }
        }
    }
    // T_b(cA, cB) = sum_a rho^c_a(cA, cB) M_ab, rho^c_a(cA, cB) = sum_{alpha in a} conj(C_ka[alpha, cA]) C_kb[alpha, cB]
    auto t_of = [&](int cA, C (&T)[NA][NC]) {
        C rc[NA][NC];
#pragma unroll
        for (int a = 0; a < NA; ++a) {
#pragma unroll
            for (int cB = 0; cB < NC; ++cB) rc[a][cB] = mk(R(0), R(0));
            for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) {
                const C xa = ca_[(al * NB + cA) * cs];
#pragma unroll
                for (int cB = 0; cB < NC; ++cB) cfma_conj(rc[a][cB], xa, cb_[(al * NB + cB) * cs]);
            }
        }
#pragma unroll
        for (int b = 0; b < NA; ++b)
#pragma unroll
            for (int cB = 0; cB < NC; ++cB) {
                C s = mk(R(0), R(0));
#pragma unroll
                for (int a = 0; a < NA; ++a) cfma(s, rc[a][cB], M[a][b]);
                T[b][cB] = s;
            }
    };
    auto element = [&](const C (&T)[NA][NC], const C* ex1, int cA, int vA, int cB, int vB) {
        C D = mk(R(0), R(0));
#pragma unroll
        for (int b = 0; b < NA; ++b) cfma(D, T[b][cB], rv[b][vB][vA]);
        C X = mk(R(0), R(0));
        if (exchange) {
#pragma unroll
            for (int a = 0; a < NA; ++a)
                cfma(X, ex1[a], to_cx<R>(e2[(static_cast<size_t>(kb) * NA + a) * NCV + cB * NV + vB]));
        }
        sink(cA, vA, cB, vB, csub(D, X));
    };
    auto ex1_of = [&](int cA, int vA, C* ex1) {
        if (exchange) {
#pragma unroll
            for (int a = 0; a < NA; ++a) ex1[a] = to_cx<R>(e1[(static_cast<size_t>(ka) * NA + a) * NCV + cA * NV + vA]);
        }
    };
    if constexpr (COLS_OUTER) {
        C T[NC][NA][NC];
        C ex1[NC][NV][NA];
#pragma unroll
        for (int cA = 0; cA < NC; ++cA) {
            t_of(cA, T[cA]);
#pragma unroll
            for (int vA = 0; vA < NV; ++vA) ex1_of(cA, vA, ex1[cA][vA]);
        }
#pragma unroll
        for (int cB = 0; cB < NC; ++cB)
#pragma unroll
            for (int vB = 0; vB < NV; ++vB)
#pragma unroll
                for (int cA = 0; cA < NC; ++cA)
#pragma unroll
                    for (int vA = 0; vA < NV; ++vA) element(T[cA], ex1[cA][vA], cA, vA, cB, vB);
    } else {
#pragma unroll
        for (int cA = 0; cA < NC; ++cA) {
            C T[NA][NC];
            t_of(cA, T);
#pragma unroll
            for (int vA = 0; vA < NV; ++vA) {
                C ex1[NA];
                ex1_of(cA, vA, ex1);
#pragma unroll
                for (int cB = 0; cB < NC; ++cB)
#pragma unroll
                    for (int vB = 0; vB < NV; ++vB) element(T, ex1, cA, vA, cB, vB);
            }
        }
    }
}

// Diagonal element: scissor + E_c(k+Q) - E_v(k) - Re(D - X), computed as the reference
// does, ((scissor + Ec - Ev)/2 - Re(D - X)/2) doubled.
__device__ inline double diag_value(const double* __restrict__ eval_k, const double* __restrict__ eval_kq, int nb,
                                    int nv, int k, int c, int v, double scissor, double dx_re) {
    const double e = (scissor + eval_kq[static_cast<size_t>(k) * nb + nv + c]) - eval_k[static_cast<size_t>(k) * nb + v];
    const double h = e / 2. - dx_re / 2.;
    return h + h;
}

// LDS layout of the staged tile: column c at c * tile_pitch, its row r at position
// (r mod NCV) TK + r / NCV, so that the threads of a wave (consecutive pair rows tr, then
// pair columns tc) write consecutive slots for each element of their blocks; the pitch is
// the smallest >= ROWS that moves the next pair column by one run of TK slots modulo the
// 64 LDS banks (2 TK banks for 8-byte entries), which spreads a wave's 64 writes evenly.
constexpr int tile_pitch(int tk, int ncv) {
    const int rows = tk * ncv;
    for (int l = rows; l < rows + 64; ++l)
        if ((2 * l * ncv - 2 * tk) % 64 == 0) return l;
    return rows + 1;
}
__host__ __device__ constexpr int tile_slot(int r, int tk, int ncv) { return (r % ncv) * tk + r / ncv; }

// Structured build kernel: tile of TK x TK k-point pairs, one pair per thread. With
// STAGE the (TK NCV)^2 tile of H is assembled in LDS (column-major, rows permuted as
// tile_slot gives) and then written column by column, consecutive threads on
// consecutive rows, so that every store instruction writes whole contiguous column
// segments; without it (the FP64 build, whose tile would cost occupancy) each pair
// stores its block directly.
// R: arithmetic type of the assembly (double or float).
// PART: 0 every tile; 1 the tiles above the block diagonal only, their pairs' elements
// produced column by column (rows together, for the direct stores); 2 the other tiles.
template <typename R, int NA, int NC, int NV, int TK, bool STAGE, int PART = 0>
__global__ __launch_bounds__(TK* TK) void bse_tile_kernel(LatticeDesc lat, int nk, const double2* __restrict__ evk,
                                                          const double2* __restrict__ evkq,
                                                          const double* __restrict__ eval_k,
                                                          const double* __restrict__ eval_kq,
                                                          const int* __restrict__ kcoord,
                                                          const cx_t<R>* __restrict__ ft, int fts,
                                                          const double2* __restrict__ e1,
                                                          const double2* __restrict__ e2, int exchange,
                                                          double scissor, int lower_rows, float hsplit,
                                                          cx_t<R>* __restrict__ H) {
    using C = cx_t<R>;
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    constexpr int ROWS = TK * NCV;
    constexpr int LDT = tile_pitch(TK, NCV);
    extern __shared__ unsigned char smem_raw[];
    const int norb = lat.norb;
    C* s_coef = reinterpret_cast<C*>(smem_raw);
    C* s_tile = s_coef + static_cast<size_t>(2 * TK) * norb * NB;
    const int Kr = blockIdx.x, Kc = blockIdx.y;
    if ((PART == 1 && Kr >= Kc) || (PART == 2 && Kr < Kc)) return;
    // block-lower build: skip tiles whose last row lies above the row block of their first
    // column (no entry of theirs is read)
    if (lower_rows > 0) {
        const int64_t last_row = static_cast<int64_t>(Kr + 1) * ROWS - 1;
        const int64_t first_col = static_cast<int64_t>(Kc) * ROWS;
        if (last_row < first_col / lower_rows * lower_rows) return;
    }
    stage_coefficients<R, NC, NV, TK>(s_coef, norb, nk, Kr, Kc, evk, evkq, TK * TK);
    __syncthreads();

    const int t = threadIdx.x;
    const int tr = t % TK, tc = t / TK;
    const int kr = Kr * TK + tr, kc = Kc * TK + tc;
    if (kr < nk && kc < nk) {
        const bool sw = kr > kc;
        const int ka = sw ? kc : kr;
        const int kb = sw ? kr : kc;
        const C* ca_ = s_coef + (sw ? (TK + tc) : tr);
        const C* cb_ = s_coef + (sw ? tr : (TK + tc));
        const bool diagblk = (ka == kb);
        const int64_t N = static_cast<int64_t>(nk) * NCV;
        // element (r, c) of this pair's block: to the LDS tile, or straight to H
        const int r0 = tr * NCV, c0 = tc * NCV;
        const int64_t g0 = static_cast<int64_t>(Kr) * ROWS, h0 = static_cast<int64_t>(Kc) * ROWS;
        auto put = [&](int r, int c, C v) {
            if (STAGE) {
                s_tile[c * LDT + tile_slot(r, TK, NCV)] = v;
            } else {
                H[(g0 + r) + (h0 + c) * N] = v;
            }
        };
        auto sink = [&](int cA, int vA, int cB, int vB, C DX) {
                                      const int ia = cA * NV + vA, ib = cB * NV + vB;
                                      if (diagblk) {
                                          if (ia < ib) {
                                              const C f = mk(-DX.x, -DX.y);
                                              put(r0 + ia, c0 + ib, f);
                                              put(r0 + ib, c0 + ia, cconj(f));
                                          } else if (ia == ib) {
                                              const double h = diag_value(eval_k, eval_kq, NB, NV, ka, cA, vA, scissor,
                                                                          static_cast<double>(DX.x));
                                              put(r0 + ia, c0 + ia, mk(R(h), R(0)));
                                          }
                                      } else if (!sw) {
                                          put(r0 + ia, c0 + ib, mk(-DX.x, -DX.y));
                                      } else {
                                          put(r0 + ib, c0 + ia, mk(-DX.x, DX.y));
                                      }
                                  };
        // above the diagonal (PART 1) a pair's direct stores go column by column, its rows
        // together (their 128-byte lines filled by consecutive stores); below it the
        // transposed stores are in that order already
        pair_block<R, NA, NC, NV, PART == 1>(lat, ca_, cb_, 2 * TK, ka, kb, ft, fts, kcoord, e1, e2, exchange, sink);
    }
    if (STAGE) {
        __syncthreads();
        const int64_t N = static_cast<int64_t>(nk) * NCV;
        const int64_t row0 = static_cast<int64_t>(Kr) * ROWS, col0 = static_cast<int64_t>(Kc) * ROWS;
        constexpr bool kSplittable = sizeof(R) == 4 && ROWS % 32 == 0;
        if (kSplittable && hsplit > 0.0f) {
            // split FP16 layout of the product on the FP16 matrix cores (ROWS a multiple of
            // 32, N a multiple of 32): 32 entries x = a + i b of a column and a 32-row step
            // occupy the same 256 bytes as in FP32, as 16 pieces of 8 halves, piece
            // (part q, group g) at byte 16 (4 q + g) holding part q (a hi, a lo, b hi, b lo of
            // hsplit x) of the rows 8 (e >> 1) + 2 g + (e & 1), e = 0..7
            constexpr int kSteps = kSplittable ? ROWS / 32 : 1;
            for (int e = t; e < ROWS * kSteps * 16; e += TK * TK) {
                const int piece = e & 15, step = (e >> 4) % kSteps, c = (e >> 4) / kSteps;
                const int q = piece >> 2, g = piece & 3;
                // a step lies wholly inside or outside the matrix (N and row0 multiples of 32)
                if (col0 + c >= N || row0 + 32 * step >= N) continue;
                f16x8 out;
#pragma unroll
                for (int el = 0; el < 8; ++el) {
                    const int r = 32 * step + 8 * (el >> 1) + 2 * g + (el & 1);
                    const C v = s_tile[c * LDT + tile_slot(r, TK, NCV)];
                    const float x = static_cast<float>((q < 2) ? v.x : v.y) * hsplit;
                    const _Float16 hi = static_cast<_Float16>(x);
                    out[el] = (q & 1) ? static_cast<_Float16>(x - static_cast<float>(hi)) : hi;
                }
                f16x8* dst = reinterpret_cast<f16x8*>(H + (row0 + 32 * step) + (col0 + c) * N) + piece;
                *dst = out;
            }
        } else {
            for (int e = t; e < ROWS * ROWS; e += TK * TK) {
                const int r = e % ROWS, c = e / ROWS;
                if (row0 + r < N && col0 + c < N) H[(row0 + r) + (col0 + c) * N] = s_tile[c * LDT + tile_slot(r, TK, NCV)];
            }
        }
    }
}

// Diagonal of H only (Davidson preconditioner without a resident matrix): the
// diagonal tiles of the build grid, where the threads with k_row == k_col evaluate
// their (k, k) block.
template <typename R, int NA, int NC, int NV, int TK>
__global__ __launch_bounds__(TK* TK) void bse_diag_kernel(LatticeDesc lat, int nk, const double2* __restrict__ evk,
                                                          const double2* __restrict__ evkq,
                                                          const double* __restrict__ eval_k,
                                                          const double* __restrict__ eval_kq,
                                                          const int* __restrict__ kcoord,
                                                          const double2* __restrict__ ft,
                                                          const double2* __restrict__ e1,
                                                          const double2* __restrict__ e2, int exchange,
                                                          double scissor, double* __restrict__ diag) {
    using C = cx_t<R>;
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    extern __shared__ unsigned char smem_raw[];
    C* s_coef = reinterpret_cast<C*>(smem_raw);
    const int norb = lat.norb;
    const int K = blockIdx.x;
    stage_coefficients<R, NC, NV, TK>(s_coef, norb, nk, K, K, evk, evkq, TK * TK);
    __syncthreads();
    const int t = threadIdx.x;
    const int tr = t % TK, tc = t / TK;
    const int k = K * TK + tr;
    if (tr != tc || k >= nk) return;
    const C* ca_ = s_coef + tr;
    const C* cb_ = s_coef + TK + tc;
    pair_block<R, NA, NC, NV>(lat, ca_, cb_, 2 * TK, k, k, ft, NA * NA, kcoord, e1, e2, exchange,
                              [&](int cA, int vA, int cB, int vB, C DX) {
                                  if (cA == cB && vA == vB) {
                                      diag[static_cast<int64_t>(k) * NCV + cA * NV + vA] =
                                          diag_value(eval_k, eval_kq, NB, NV, k, cA, vA, scissor,
                                                     static_cast<double>(DX.x));
                                  }
                              });
}

// Matrix-free product Y = H X on the matrix cores. The grid is (row tiles) x (column
// splits) in one dimension: consecutive workgroup ids go round-robin over the XCDs,
// so split = id % splits keeps a column range, and its slice of the staged vectors,
// on one XCD's L2. A workgroup stages the coefficients of its TK row k-points once;
// for every column tile of its range it stages the TK column k-points, recomputes the
// (TK*NCV) x (TK*NCV) tile of H into LDS (column-major, one k-point pair per thread,
// the instruction sequence of the build kernel, so the tile equals the stored matrix
// bit for bit), and multiplies it with the vectors Xt[col][32] on 16x16x4 MFMAs (four
// real products per complex tile pair). Each wave owns a fixed set of (16-row tile,
// 16-vector tile) outputs. Partial sums per split go to Yp[split][row][32] and are
// added in a fixed order by launch_reduce_partials.
constexpr int matfree_threads(int tk) { return tk * tk >= 256 ? 256 : 64; }

// Stage the coefficient vectors of `count` k-points from kbase into slots of stride
// `ss` elements, per slot [norb][NB]: conduction columns (k+Q stack), then valence (k).
template <typename R, int NC, int NV>
__device__ inline void stage_slots(cx_t<R>* dst, int ss, int norb, int nk, int kbase, int count,
                                   const double2* __restrict__ evk, const double2* __restrict__ evkq, int nthreads) {
    constexpr int NB = NC + NV;
    const int total = count * norb * NB;
    for (int e = threadIdx.x; e < total; e += nthreads) {
        const int al = e % norb;
        const int rest = e / norb;
        const int band = rest % NB;
        const int slot = rest / NB;
        const int k = kbase + slot;
        cx_t<R> val = mk(R(0), R(0));
        if (k < nk) {
            const double2 g = (band < NC) ? evkq[(static_cast<size_t>(k) * NB + (NV + band)) * norb + al]
                                          : evk[(static_cast<size_t>(k) * NB + (band - NC)) * norb + al];
            val = to_cx<R>(g);
        }
        dst[static_cast<size_t>(slot) * ss + al * NB + band] = val;
    }
}

template <typename R, int NA, int NC, int NV, int TK>
__global__ __launch_bounds__(matfree_threads(TK)) void bse_matfree_kernel(
    LatticeDesc lat, int nk, const double2* __restrict__ evk, const double2* __restrict__ evkq,
    const double* __restrict__ eval_k, const double* __restrict__ eval_kq, const int* __restrict__ kcoord,
    const double2* __restrict__ ft, const double2* __restrict__ e1, const double2* __restrict__ e2, int exchange,
    double scissor, const cx_t<R>* __restrict__ Xt, int nvec, int splits, int tiles_per_split,
    cx_t<R>* __restrict__ Yp) {
    using C = cx_t<R>;
    using M = Mfma<R>;
    using acc_t = typename M::acc_t;
    constexpr int NB = NC + NV;
    constexpr int NCV = NC * NV;
    constexpr int ROWS = TK * NCV;
    constexpr int ROWS16 = (ROWS + 15) / 16 * 16;  // MFMA-padded tile size
    constexpr int LDH = ROWS16 + 1;                // column-major tile, one element of padding
    constexpr int NT = matfree_threads(TK);
    constexpr int WPB = NT / 64;
    constexpr int RT = ROWS16 / 16;
    constexpr int MAXC = (2 * RT + WPB - 1) / WPB;  // output tiles per wave (up to 2 vector tiles)
    extern __shared__ unsigned char smem_raw[];
    const int norb = lat.norb;
    const int ss = norb * NB + 1;  // slot stride, padded against LDS bank conflicts
    C* s_row = reinterpret_cast<C*>(smem_raw);
    C* s_col = s_row + static_cast<size_t>(TK) * ss;
    C* s_h = s_col + static_cast<size_t>(TK) * ss;

    const int split = blockIdx.x % splits;
    const int Kr = blockIdx.x / splits;
    const int nT = (nk + TK - 1) / TK;
    const int kc0 = split * tiles_per_split;
    const int kc1 = min(nT, kc0 + tiles_per_split);
    const int64_t N = static_cast<int64_t>(nk) * NCV;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ntiles = RT * ((nvec + 15) / 16);

    stage_slots<R, NC, NV>(s_row, ss, norb, nk, Kr * TK, TK, evk, evkq, NT);
    // padding rows/columns are never written by the build step: start from zero
    for (int e = t; e < ROWS16 * LDH; e += NT) s_h[e] = mk(R(0), R(0));
    acc_t accR[MAXC], accI[MAXC];
#pragma unroll
    for (int i = 0; i < MAXC; ++i) accR[i] = accI[i] = acc_t{0, 0, 0, 0};

    for (int Kc = kc0; Kc < kc1; ++Kc) {
        __syncthreads();  // the previous tile has been consumed
        stage_slots<R, NC, NV>(s_col, ss, norb, nk, Kc * TK, TK, evk, evkq, NT);
        __syncthreads();
        for (int p = t; p < TK * TK; p += NT) {
            const int tr = p % TK, tc = p / TK;
            const int kr = Kr * TK + tr, kc = Kc * TK + tc;
            if (kr < nk && kc < nk) {
                const bool sw = kr > kc;
                const int ka = sw ? kc : kr;
                const int kb = sw ? kr : kc;
                const C* ca_ = sw ? s_col + static_cast<size_t>(tc) * ss : s_row + static_cast<size_t>(tr) * ss;
                const C* cb_ = sw ? s_row + static_cast<size_t>(tr) * ss : s_col + static_cast<size_t>(tc) * ss;
                const bool diagblk = (ka == kb);
                pair_block<R, NA, NC, NV>(lat, ca_, cb_, 1, ka, kb, ft, NA * NA, kcoord, e1, e2, exchange,
                                          [&](int cA, int vA, int cB, int vB, C DX) {
                                              const int ia = cA * NV + vA, ib = cB * NV + vB;
                                              // element (row, col) of the tile at s_h[col * LDH + row]
                                              const int r0 = tr * NCV, c0 = tc * NCV;
                                              if (diagblk) {
                                                  if (ia < ib) {
                                                      s_h[(c0 + ib) * LDH + r0 + ia] = mk(-DX.x, -DX.y);
                                                      s_h[(c0 + ia) * LDH + r0 + ib] = mk(-DX.x, DX.y);
                                                  } else if (ia == ib) {
                                                      const double h = diag_value(eval_k, eval_kq, NB, NV, ka, cA, vA,
                                                                                  scissor, static_cast<double>(DX.x));
                                                      s_h[(c0 + ia) * LDH + r0 + ia] = mk(R(h), R(0));
                                                  }
                                              } else if (!sw) {
                                                  s_h[(c0 + ib) * LDH + r0 + ia] = mk(-DX.x, -DX.y);
                                              } else {
                                                  s_h[(c0 + ia) * LDH + r0 + ib] = mk(-DX.x, DX.y);
                                              }
                                          });
            }
        }
        __syncthreads();
        const int64_t cbase = static_cast<int64_t>(Kc) * ROWS;
#pragma unroll
        for (int i = 0; i < MAXC; ++i) {
            const int tile = wave + WPB * i;
            if (tile < ntiles) {
                const int rt = tile % RT, vt = tile / RT;
                acc_t aR = accR[i], aI = accI[i];
#pragma unroll
                for (int kk = 0; kk < ROWS16 / 4; ++kk) {
                    const int cl = 4 * kk + lk;
                    const C a = s_h[cl * LDH + 16 * rt + li];
                    const int64_t col = cbase + cl;
                    // columns past the tile or the matrix meet zero vectors (tile entries there are finite)
                    const C b = (cl < ROWS && col < N) ? Xt[col * kBlockVectors + 16 * vt + li] : mk(R(0), R(0));
                    aR = M::mma(a.x, b.x, aR);
                    aR = M::mma(a.y, -b.y, aR);
                    aI = M::mma(a.x, b.y, aI);
                    aI = M::mma(a.y, b.x, aI);
                }
                accR[i] = aR;
                accI[i] = aI;
            }
        }
    }
    C* out = Yp + static_cast<int64_t>(split) * N * kBlockVectors;
#pragma unroll
    for (int i = 0; i < MAXC; ++i) {
        const int tile = wave + WPB * i;
        if (tile < ntiles) {
            const int rt = tile % RT, vt = tile / RT;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int rl = 16 * rt + M::out_row(lk, v);
                const int64_t row = static_cast<int64_t>(Kr) * ROWS + rl;
                if (rl < ROWS && row < N) out[row * kBlockVectors + 16 * vt + li] = mk(accR[i][v], accI[i][v]);
            }
        }
    }
}

// ---- generic shapes: any natoms <= kMaxAtoms, any nv and nc ----
// One element per thread, coefficients read straight from the stacks. The element
// (row, col) is evaluated for the canonical pair i = min(row, col) <= j = max(row, col)
// with an instruction sequence that depends only on (i, j), so H(j, i) = conj H(i, j)
// holds bit for bit. R is the arithmetic of the assembly: coefficients, lattice sums
// and exchange factors are rounded to R first, as in the specialised kernels.
template <typename R>
__device__ inline cx_t<R> generic_dx(const LatticeDesc& lat, int nv, int nc, int64_t i, int64_t j,
                                     const double2* __restrict__ evk, const double2* __restrict__ evkq,
                                     const int* __restrict__ kcoord, const double2* __restrict__ ft,
                                     const double2* __restrict__ e1, const double2* __restrict__ e2, int exchange) {
    using C = cx_t<R>;
    const int ncv = nc * nv, nb = nc + nv, norb = lat.norb, na = lat.natoms;
    const int ka = static_cast<int>(i / ncv), ia = static_cast<int>(i % ncv);
    const int kb = static_cast<int>(j / ncv), ib = static_cast<int>(j % ncv);
    const int cA = ia / nv, vA = ia % nv, cB = ib / nv, vB = ib % nv;
    const double2* Ca = evkq + (static_cast<size_t>(ka) * nb + nv + cA) * norb;
    const double2* Va = evk + (static_cast<size_t>(ka) * nb + vA) * norb;
    const double2* Cb = evkq + (static_cast<size_t>(kb) * nb + nv + cB) * norb;
    const double2* Vb = evk + (static_cast<size_t>(kb) * nb + vB) * norb;
    const double2* M = ft + static_cast<size_t>(table_index(lat, kcoord, ka, kb)) * na * na;
    // rho^v_b = sum_{beta in b} conj(V_kb[beta, vB]) V_ka[beta, vA]
    C rv[kMaxAtoms];
    for (int b = 0; b < na; ++b) {
        C sacc = mk(R(0), R(0));
        for (int be = lat.orb_begin[b]; be < lat.orb_begin[b + 1]; ++be) cfma_conj(sacc, to_cx<R>(Vb[be]), to_cx<R>(Va[be]));
        rv[b] = sacc;
    }
    C D = mk(R(0), R(0)), X = mk(R(0), R(0));
    for (int a = 0; a < na; ++a) {
        // rho^c_a = sum_{alpha in a} conj(C_ka[alpha, cA]) C_kb[alpha, cB]
        C rc = mk(R(0), R(0));
        for (int al = lat.orb_begin[a]; al < lat.orb_begin[a + 1]; ++al) cfma_conj(rc, to_cx<R>(Ca[al]), to_cx<R>(Cb[al]));
        C y = mk(R(0), R(0));
        for (int b = 0; b < na; ++b) cfma(y, to_cx<R>(M[a + na * b]), rv[b]);
        cfma(D, rc, y);
        if (exchange) {
            cfma(X, to_cx<R>(e1[(static_cast<size_t>(ka) * na + a) * ncv + ia]),
                 to_cx<R>(e2[(static_cast<size_t>(kb) * na + a) * ncv + ib]));
        }
    }
    return csub(D, X);
}

// Element (row, col) of H from the canonical (D - X): diagonal rule on the diagonal,
// -(D - X) above it and its conjugate below.
template <typename R>
__device__ inline cx_t<R> generic_element(const LatticeDesc& lat, int nv, int nc, int64_t row, int64_t col,
                                          const double2* __restrict__ evk, const double2* __restrict__ evkq,
                                          const double* __restrict__ eval_k, const double* __restrict__ eval_kq,
                                          const int* __restrict__ kcoord, const double2* __restrict__ ft,
                                          const double2* __restrict__ e1, const double2* __restrict__ e2,
                                          int exchange, double scissor) {
    const int64_t i = row < col ? row : col, j = row < col ? col : row;
    const cx_t<R> DX = generic_dx<R>(lat, nv, nc, i, j, evk, evkq, kcoord, ft, e1, e2, exchange);
    if (i == j) {
        const int ncv = nc * nv;
        const int k = static_cast<int>(i / ncv), ia = static_cast<int>(i % ncv);
        return mk(R(diag_value(eval_k, eval_kq, nc + nv, nv, k, ia / nv, ia % nv, scissor, static_cast<double>(DX.x))),
                  R(0));
    }
    return row < col ? mk(-DX.x, -DX.y) : mk(-DX.x, DX.y);
}

// Whole matrix (or, with lower_rows > 0, the block-lower part read by the split product:
// entries (row, col) with row >= lower_rows floor(col / lower_rows)), column-major.
template <typename R>
__global__ __launch_bounds__(256) void bse_generic_kernel(LatticeDesc lat, int nk, int nv, int nc,
                                                          const double2* __restrict__ evk,
                                                          const double2* __restrict__ evkq,
                                                          const double* __restrict__ eval_k,
                                                          const double* __restrict__ eval_kq,
                                                          const int* __restrict__ kcoord,
                                                          const double2* __restrict__ ft,
                                                          const double2* __restrict__ e1,
                                                          const double2* __restrict__ e2, int exchange,
                                                          double scissor, int lower_rows, cx_t<R>* __restrict__ H) {
    const int64_t N = static_cast<int64_t>(nk) * nc * nv;
    const int64_t row = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= N) return;
    for (int64_t col = blockIdx.y; col < N; col += gridDim.y) {
        if (lower_rows > 0 && row < col / lower_rows * lower_rows) continue;
        H[row + col * N] = generic_element<R>(lat, nv, nc, row, col, evk, evkq, eval_k, eval_kq, kcoord, ft, e1, e2,
                                              exchange, scissor);
    }
}

// Diagonal of H (FP64), one row per thread.
__global__ __launch_bounds__(256) void bse_generic_diag_kernel(LatticeDesc lat, int nk, int nv, int nc,
                                                               const double2* __restrict__ evk,
                                                               const double2* __restrict__ evkq,
                                                               const double* __restrict__ eval_k,
                                                               const double* __restrict__ eval_kq,
                                                               const int* __restrict__ kcoord,
                                                               const double2* __restrict__ ft,
                                                               const double2* __restrict__ e1,
                                                               const double2* __restrict__ e2, int exchange,
                                                               double scissor, double* __restrict__ diag) {
    const int64_t N = static_cast<int64_t>(nk) * nc * nv;
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= N) return;
    diag[r] = generic_element<double>(lat, nv, nc, r, r, evk, evkq, eval_k, eval_kq, kcoord, ft, e1, e2, exchange,
                                      scissor)
                  .x;
}

// Matrix-free product with recomputed tiles for generic shapes: a workgroup owns 64 rows
// and, for each 64-column tile of its split, evaluates the tile element by element into
// LDS (column-major, padded) and multiplies it with the staged vectors Xt[col][32] on the
// 16x16x4 matrix cores; partial sums per split go to Yp[split][row][32].
constexpr int kGenRows = 64;
template <typename R>
__global__ __launch_bounds__(256) void bse_matfree_generic_kernel(
    LatticeDesc lat, int nk, int nv, int nc, const double2* __restrict__ evk, const double2* __restrict__ evkq,
    const double* __restrict__ eval_k, const double* __restrict__ eval_kq, const int* __restrict__ kcoord,
    const double2* __restrict__ ft, const double2* __restrict__ e1, const double2* __restrict__ e2, int exchange,
    double scissor, const cx_t<R>* __restrict__ Xt, int nvec, int splits, int tiles_per_split,
    cx_t<R>* __restrict__ Yp) {
    using C = cx_t<R>;
    using M = Mfma<R>;
    using acc_t = typename M::acc_t;
    constexpr int LDH = kGenRows + 1;
    constexpr int RT = kGenRows / 16;
    __shared__ C s_h[kGenRows * LDH];
    const int64_t N = static_cast<int64_t>(nk) * nc * nv;
    const int split = blockIdx.x % splits;
    const int64_t rbase = static_cast<int64_t>(blockIdx.x / splits) * kGenRows;
    const int nT = static_cast<int>((N + kGenRows - 1) / kGenRows);
    const int t0 = split * tiles_per_split;
    const int t1 = min(nT, t0 + tiles_per_split);
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int ntiles = RT * ((nvec + 15) / 16);  // at most 8: two per wave
    acc_t accR[2], accI[2];
    accR[0] = accI[0] = accR[1] = accI[1] = acc_t{0, 0, 0, 0};
    for (int T = t0; T < t1; ++T) {
        const int64_t cbase = static_cast<int64_t>(T) * kGenRows;
        __syncthreads();  // the previous tile has been consumed
        for (int e = t; e < kGenRows * kGenRows; e += 256) {
            const int rl = e % kGenRows, cl = e / kGenRows;
            const int64_t row = rbase + rl, col = cbase + cl;
            C h = mk(R(0), R(0));
            if (row < N && col < N) {
                h = generic_element<R>(lat, nv, nc, row, col, evk, evkq, eval_k, eval_kq, kcoord, ft, e1, e2, exchange,
                                       scissor);
            }
            s_h[cl * LDH + rl] = h;
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int tile = wave + 4 * i;
            if (tile < ntiles) {
                const int rt = tile % RT, vt = tile / RT;
                acc_t aR = accR[i], aI = accI[i];
#pragma unroll 4
                for (int kk = 0; kk < kGenRows / 4; ++kk) {
                    const int cl = 4 * kk + lk;
                    const C a = s_h[cl * LDH + 16 * rt + li];
                    const int64_t col = cbase + cl;
                    const C b = (col < N) ? Xt[col * kBlockVectors + 16 * vt + li] : mk(R(0), R(0));
                    aR = M::mma(a.x, b.x, aR);
                    aR = M::mma(a.y, -b.y, aR);
                    aI = M::mma(a.x, b.y, aI);
                    aI = M::mma(a.y, b.x, aI);
                }
                accR[i] = aR;
                accI[i] = aI;
            }
        }
    }
    C* out = Yp + static_cast<int64_t>(split) * N * kBlockVectors;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int tile = wave + 4 * i;
        if (tile < ntiles) {
            const int rt = tile % RT, vt = tile / RT;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int64_t row = rbase + 16 * rt + M::out_row(lk, v);
                if (row < N) out[row * kBlockVectors + 16 * vt + li] = mk(accR[i][v], accI[i][v]);
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
    const int q = table_index(lat, kcoord, k, k2);
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

// out[q fts + j] = ft[q na2 + j] rounded to FP32 (j < na2), zero padding up to fts.
__global__ void table_to_fp32_kernel(const double2* __restrict__ ft, int64_t rows, int na2, int fts,
                                     float2* __restrict__ out) {
    const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (e >= rows * fts) return;
    const int64_t q = e / fts;
    const int j = static_cast<int>(e - q * fts);
    if (j < na2) {
        const double2 v = ft[q * na2 + j];
        out[e] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
    } else {
        out[e] = make_float2(0.0f, 0.0f);
    }
}

// The FP32 build stages its output tile in LDS for coalesced stores; the FP64 tile would
// halve the workgroups per CU.
template <typename R>
constexpr bool kStageTile = sizeof(R) == 4;

template <typename R, int TK>
size_t coef_bytes(const Context& c, int nb) {
    return static_cast<size_t>(2 * TK) * c.lat.norb * nb * sizeof(cx_t<R>);
}

// LDS of the tile kernel: staged coefficients and the padded output tile.
template <typename R, int TK>
size_t tile_bytes(const Context& c, int nb, int ncv) {
    const size_t rows = static_cast<size_t>(TK) * ncv;
    return coef_bytes<R, TK>(c, nb) + (kStageTile<R> ? rows * static_cast<size_t>(tile_pitch(TK, ncv)) * sizeof(cx_t<R>) : 0);
}

// Launchers for one (natoms, nc, nv) shape.
template <typename R, int NA, int NC, int NV>
struct Shape {
    static constexpr int NB = NC + NV;
    static constexpr int NCV = NC * NV;
    static bool matches(const Context& c) { return c.lat.natoms == NA && c.nc == NC && c.nv == NV; }

    template <int TK>
    static void build_tk(const Context& c, cx_t<R>* H, int lower_rows, float hsplit, hipStream_t s) {
        const int nT = (c.nk + TK - 1) / TK;
        const size_t lds = tile_bytes<R, TK>(c, NB, NCV);
        if (hsplit > 0.0f && ((TK * NCV) % 32 != 0 || c.N % 32 != 0 || !kStageTile<R>)) {
            throw Error(XATU_ERR_LOGIC, "split FP16 layout needs 32-row tiles");
        }
        const cx_t<R>* ft;
        int fts;
        if constexpr (sizeof(R) == 4) {
            // FP32 copy of the table, rows padded to an even length (16-byte aligned rows)
            constexpr int na2 = NA * NA;
            fts = (na2 + 1) & ~1;
            const int64_t rows = static_cast<int64_t>(c.d_ft.n / na2);
            c.d_ft32.alloc(static_cast<size_t>(rows) * fts);
            hipLaunchKernelGGL(table_to_fp32_kernel, dim3(static_cast<unsigned>((rows * fts + 255) / 256)), dim3(256), 0, s,
                               c.d_ft.p, rows, na2, fts, c.d_ft32.p);
            ft = c.d_ft32.p;
        } else {
            fts = NA * NA;
            ft = c.d_ft.p;
        }
        if constexpr (kStageTile<R>) {
            hipLaunchKernelGGL((bse_tile_kernel<R, NA, NC, NV, TK, true>), dim3(nT, nT), dim3(TK * TK), lds, s, c.lat, c.nk,
                               c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, ft, fts, c.d_e1.p,
                               c.d_e2.p, c.inter.exchange, c.scissor, lower_rows, hsplit, H);
        } else {
            // the tiles above the diagonal and the others in two launches, one element order each
            hipLaunchKernelGGL((bse_tile_kernel<R, NA, NC, NV, TK, false, 1>), dim3(nT, nT), dim3(TK * TK), lds, s, c.lat,
                               c.nk, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, ft, fts,
                               c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, lower_rows, hsplit, H);
            hipLaunchKernelGGL((bse_tile_kernel<R, NA, NC, NV, TK, false, 2>), dim3(nT, nT), dim3(TK * TK), lds, s, c.lat,
                               c.nk, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, ft, fts,
                               c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, lower_rows, hsplit, H);
        }
    }
    // The tile size the build uses (0: none fits).
    static int build_tile(const Context& c) {
        constexpr size_t kLds = 160 * 1024;
        if (tile_bytes<R, 16>(c, NB, NCV) <= kLds) return 16;
        if (tile_bytes<R, 8>(c, NB, NCV) <= kLds) return 8;
        if (tile_bytes<R, 4>(c, NB, NCV) <= kLds) return 4;
        return 0;
    }
    static void build(const Context& c, cx_t<R>* H, hipStream_t s, int lower_rows = 0, float hsplit = 0.0f) {
        const int tk = build_tile(c);
        if (tk == 16) {
            build_tk<16>(c, H, lower_rows, hsplit, s);
        } else if (tk == 8) {
            build_tk<8>(c, H, lower_rows, hsplit, s);
        } else if (tk == 4) {
            build_tk<4>(c, H, lower_rows, hsplit, s);
        } else {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "too many orbitals for the tile kernel");
        }
        XATU_HIP(hipGetLastError());
    }

    static void diag(const Context& c, double* d, hipStream_t s) {
        constexpr int TK = 8;
        if (coef_bytes<R, TK>(c, NB) > 160 * 1024) throw Error(XATU_ERR_INVALID_ARGUMENT, "too many orbitals");
        const int nT = (c.nk + TK - 1) / TK;
        const size_t lds = coef_bytes<R, TK>(c, NB);
        hipLaunchKernelGGL((bse_diag_kernel<R, NA, NC, NV, TK>), dim3(nT), dim3(TK * TK), lds, s, c.lat, c.nk, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p,
                           c.d_ft.p, c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, d);
        XATU_HIP(hipGetLastError());
    }

    template <int TK>
    static size_t matfree_bytes(const Context& c) {
        constexpr int ROWS16 = (TK * NCV + 15) / 16 * 16;
        return (static_cast<size_t>(2 * TK) * (c.lat.norb * NB + 1) + static_cast<size_t>(ROWS16) * (ROWS16 + 1)) *
               sizeof(cx_t<R>);
    }
    template <int TK>
    static void matfree_tk(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s,
                           DevBuf<cx_t<R>>& xt, DevBuf<cx_t<R>>& yp) {
        const int nT = (c.nk + TK - 1) / TK;
        const size_t lds = matfree_bytes<TK>(c);
        // about 1024 workgroups or more; splits a multiple of the XCD count
        int splits = std::max(1, (1024 + nT - 1) / nT);
        splits = std::min((splits + kXcds - 1) / kXcds * kXcds, std::max(kXcds, nT));
        const int tiles_per_split = (nT + splits - 1) / splits;
        xt.alloc(std::max<size_t>(xt.n, static_cast<size_t>(c.N) * kBlockVectors));
        yp.alloc(std::max<size_t>(yp.n, static_cast<size_t>(splits) * c.N * kBlockVectors));
        for (int v0 = 0; v0 < nvec; v0 += kBlockVectors) {
            const int nv = std::min(kBlockVectors, nvec - v0);
            launch_stage_vectors(X + static_cast<int64_t>(v0) * c.N, c.N, nv, xt.p, s);
            hipLaunchKernelGGL((bse_matfree_kernel<R, NA, NC, NV, TK>), dim3(nT * splits), dim3(matfree_threads(TK)),
                               lds, s, c.lat, c.nk, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p,
                               c.d_kcoord.p, c.d_ft.p, c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, xt.p, nv,
                               splits, tiles_per_split, yp.p);
            launch_reduce_partials(yp.p, c.N, nv, splits, Y + static_cast<int64_t>(v0) * c.N, s);
        }
    }
    static void matfree(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s,
                        DevBuf<cx_t<R>>& xt, DevBuf<cx_t<R>>& yp) {
        constexpr size_t kLds = 160 * 1024;
        if (matfree_bytes<16>(c) <= kLds) {
            matfree_tk<16>(c, X, Y, nvec, s, xt, yp);
        } else if (matfree_bytes<8>(c) <= kLds) {
            matfree_tk<8>(c, X, Y, nvec, s, xt, yp);
        } else if (matfree_bytes<4>(c) <= kLds) {
            matfree_tk<4>(c, X, Y, nvec, s, xt, yp);
        } else {
            throw Error(XATU_ERR_INVALID_ARGUMENT, "too many orbitals for the matrix-free kernel");
        }
        XATU_HIP(hipGetLastError());
    }
};

// Call fn(Shape<R, NA, NC, NV>{}) for the context's shape; false if none matches.
template <typename R, typename Fn>
bool for_shape(const Context& c, Fn&& fn) {
#define XATU_SHAPE(NA, NC, NV)                 \
    if (Shape<R, NA, NC, NV>::matches(c)) {    \
        fn(Shape<R, NA, NC, NV>{});            \
        return true;                           \
    }
    XATU_SHAPE(1, 1, 1)
    XATU_SHAPE(1, 2, 2)
    XATU_SHAPE(1, 3, 3)
    XATU_SHAPE(2, 1, 1)
    XATU_SHAPE(2, 2, 2)
    XATU_SHAPE(2, 3, 3)
    XATU_SHAPE(3, 1, 1)
    XATU_SHAPE(3, 2, 2)
    XATU_SHAPE(3, 3, 3)
    XATU_SHAPE(4, 1, 1)
    XATU_SHAPE(4, 2, 2)
    XATU_SHAPE(4, 3, 3)
#undef XATU_SHAPE
    return false;
}

}  // namespace

void launch_exchange_factors(const Context& c, hipStream_t s) {
    const int n = c.nk * c.nc * c.nv;
    hipLaunchKernelGGL(exchange_factors_kernel, dim3((n + 255) / 256), dim3(256), 0, s, c.lat, c.nk, c.nv, c.nc,
                       c.d_evec_k.p, c.d_evec_kq.p, c.d_ftq.p, c.d_e1.p, c.d_e2.p);
    XATU_HIP(hipGetLastError());
}

void launch_build_basis(const Context& c, int nbasis, const int* d_basis, double scissor, double2* H, hipStream_t s) {
    dim3 grid((nbasis + 255) / 256, nbasis);
    hipLaunchKernelGGL(bse_basis_kernel, grid, dim3(256), 0, s, c.lat, c.nb, nbasis, d_basis, c.d_evec_k.p,
                       c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p, c.d_ftq.p,
                       c.inter.exchange, scissor, H);
    XATU_HIP(hipGetLastError());
}

namespace {

template <typename R>
void build_generic(const Context& c, cx_t<R>* H, int lower_rows, hipStream_t s) {
    const unsigned gx = static_cast<unsigned>((c.N + 255) / 256);
    const unsigned gy = static_cast<unsigned>(std::min<int64_t>(c.N, 65535));
    hipLaunchKernelGGL((bse_generic_kernel<R>), dim3(gx, gy), dim3(256), 0, s, c.lat, c.nk, c.nv, c.nc, c.d_evec_k.p,
                       c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p, c.d_e1.p, c.d_e2.p,
                       c.inter.exchange, c.scissor, lower_rows, H);
    XATU_HIP(hipGetLastError());
}

template <typename R>
void matfree_generic(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s, DevBuf<cx_t<R>>& xt,
                     DevBuf<cx_t<R>>& yp) {
    const int nT = static_cast<int>((c.N + kGenRows - 1) / kGenRows);
    // about 1024 workgroups or more; splits a multiple of the XCD count
    int splits = std::max(1, (1024 + nT - 1) / nT);
    splits = std::min((splits + kXcds - 1) / kXcds * kXcds, std::max(kXcds, nT));
    const int tiles_per_split = (nT + splits - 1) / splits;
    xt.alloc(std::max<size_t>(xt.n, static_cast<size_t>(c.N) * kBlockVectors));
    yp.alloc(std::max<size_t>(yp.n, static_cast<size_t>(splits) * c.N * kBlockVectors));
    for (int v0 = 0; v0 < nvec; v0 += kBlockVectors) {
        const int nv = std::min(kBlockVectors, nvec - v0);
        launch_stage_vectors(X + static_cast<int64_t>(v0) * c.N, c.N, nv, xt.p, s);
        hipLaunchKernelGGL((bse_matfree_generic_kernel<R>), dim3(nT * splits), dim3(256), 0, s, c.lat, c.nk, c.nv,
                           c.nc, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p, c.d_ft.p,
                           c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, xt.p, nv, splits, tiles_per_split, yp.p);
        XATU_HIP(hipGetLastError());
        launch_reduce_partials(yp.p, c.N, nv, splits, Y + static_cast<int64_t>(v0) * c.N, s);
    }
}

}  // namespace

bool shape_specialised(const Context& c) {
    return for_shape<double>(c, [](auto) {});
}

void launch_build_tiles64(const Context& c, double2* H, hipStream_t s) {
    if (for_shape<double>(c, [&](auto sh) { decltype(sh)::build(c, H, s); })) return;
    build_generic<double>(c, H, 0, s);
}

void launch_build_tiles32(const Context& c, float2* H, bool lower, bool presplit, hipStream_t s) {
    const int lower_rows = lower ? kSplitRowBlock : 0;
    const float hsplit = presplit ? c.h16_scale : 0.0f;
    if (for_shape<float>(c, [&](auto sh) { decltype(sh)::build(c, H, s, lower_rows, hsplit); })) return;
    if (presplit) throw Error(XATU_ERR_LOGIC, "split FP16 layout needs a specialised tile kernel");
    build_generic<float>(c, H, lower_rows, s);
}

bool presplit_build_available(const Context& c) {
    if (c.N % 32 != 0) return false;
    bool ok = false;
    for_shape<float>(c, [&](auto sh) {
        using S = decltype(sh);
        const int tk = S::build_tile(c);
        ok = tk > 0 && (tk * S::NCV) % 32 == 0;
    });
    return ok;
}

void launch_diag_direct(const Context& c, double* diag, hipStream_t s) {
    if (for_shape<double>(c, [&](auto sh) { decltype(sh)::diag(c, diag, s); })) return;
    hipLaunchKernelGGL(bse_generic_diag_kernel, dim3(static_cast<unsigned>((c.N + 255) / 256)), dim3(256), 0, s, c.lat,
                       c.nk, c.nv, c.nc, c.d_evec_k.p, c.d_evec_kq.p, c.d_eval_k.p, c.d_eval_kq.p, c.d_kcoord.p,
                       c.d_ft.p, c.d_e1.p, c.d_e2.p, c.inter.exchange, c.scissor, diag);
    XATU_HIP(hipGetLastError());
}

void launch_matfree64(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    if (for_shape<double>(c, [&](auto sh) { decltype(sh)::matfree(c, X, Y, nvec, s, c.s_xt64, c.s_yp64); })) return;
    matfree_generic<double>(c, X, Y, nvec, s, c.s_xt64, c.s_yp64);
}

void launch_matfree32(Context& c, const double2* X, double2* Y, int nvec, hipStream_t s) {
    if (for_shape<float>(c, [&](auto sh) { decltype(sh)::matfree(c, X, Y, nvec, s, c.s_xt32, c.s_yp32); })) return;
    matfree_generic<float>(c, X, Y, nvec, s, c.s_xt32, c.s_yp32);
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
