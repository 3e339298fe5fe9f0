// Index of the mesh point equivalent to k_b - k_a on a Monkhorst-Pack mesh of n points
// per axis: Lattice::findEquivalentPointBZ (src/Lattice.cpp:251-279) reduces, on such
// meshes, to integer offsets taken modulo n (the reference rounds reciprocal
// coordinates of k_b - k_a and wraps them; kcoord holds the integer offsets of every
// k-point from k-point 0, so the difference is exact).
#pragma once
#include <hip/hip_runtime.h>

namespace xatu_dev {

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

// Index of the interaction table entry for the pair (k_a, k_b): the mesh point of
// k_b - k_a in real-space mode; the raw difference k_a - k_b (offsets shifted by
// mesh_n - 1 per axis) in reciprocal mode, whose table is not periodic on the mesh.
template <typename Lat>
__device__ inline int table_index(const Lat& lat, const int* __restrict__ kcoord, int ka, int kb) {
    if (!lat.recip) return mesh_index(kcoord, ka, kb, lat.ndim, lat.mesh_n);
    int q = 0, mul = 1;
    for (int j = 0; j < lat.ndim; ++j) {
        q += (kcoord[3 * ka + j] - kcoord[3 * kb + j] + lat.mesh_n - 1) * mul;
        mul *= lat.tab_n;
    }
    return q;
}

}  // namespace xatu_dev
