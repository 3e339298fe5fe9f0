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

}  // namespace xatu_dev
