// CIC field gather (device helpers shared by the gather kernels).
#pragma once
#include "common.hpp"

namespace opalsc {

// IntCICImpl<3>::gather (ippl/src/Particle/IntCIC.h:933-977): sum over the 8
// corners in reference order of ((a*b)*c)*E; corners outside the grid read 0.
// pz: z faces periodic (the guard layers hold the opposite cells).
template <typename T>
__device__ inline void gather_E(const typename vec4_of<T>::type *Eg, const int l[3], const double d[3],
                                int n0, int n1, int n2, double Ef[3], bool pz = false) {
    double w[8];
    cic_weights(d, w);
    T a0 = T(0), a1 = T(0), a2 = T(0);
#pragma unroll
    for (int c = 0; c < 8; ++c) {
        const int i = l[0] + (c & 1), j = l[1] + ((c >> 1) & 1), k = corner_z(l[2] + ((c >> 2) & 1), n2, pz);
        typename vec4_of<T>::type e;
        if (i >= 0 && j >= 0 && k >= 0 && i < n0 && j < n1) e = Eg[((size_t)k * n1 + j) * n0 + i];
        else e.x = e.y = e.z = e.w = T(0);
        const T wc = (T)w[c];
        if (c == 0) {
            a0 = wc * e.x;
            a1 = wc * e.y;
            a2 = wc * e.z;
        } else {
            a0 = a0 + wc * e.x;
            a1 = a1 + wc * e.y;
            a2 = a2 + wc * e.z;
        }
    }
    Ef[0] = (double)a0;
    Ef[1] = (double)a1;
    Ef[2] = (double)a2;
}

// Mixed-precision gather: the same stencil with FP32 weights and FMAs.
// PZ (a compile-time choice: the periodic corner lookup in the edge path
// measurably slows the common case): z faces periodic.
template <bool PZ = false>
__device__ inline void gather_E_fast(const float4 *Eg, const int l[3], const double d[3], int n0, int n1,
                                     int n2, double Ef[3]) {
    const float b0 = (float)d[0], b1 = (float)d[1], b2 = (float)d[2];
    const float a0 = 1.0f - b0, a1 = 1.0f - b1, a2 = 1.0f - b2;
    const float w[8] = {a0 * a1 * a2, b0 * a1 * a2, a0 * b1 * a2, b0 * b1 * a2,
                        a0 * a1 * b2, b0 * a1 * b2, a0 * b1 * b2, b0 * b1 * b2};
    const bool inside = l[0] >= 0 && l[1] >= 0 && l[2] >= 0 && l[0] + 1 < n0 && l[1] + 1 < n1 && l[2] + 1 < n2;
    float e0 = 0.f, e1 = 0.f, e2 = 0.f;
    if (inside) {
        const float4 *base = Eg + ((size_t)l[2] * n1 + l[1]) * n0 + l[0];
        const size_t sy = n0, sz = (size_t)n0 * n1;
        const float4 c[8] = {base[0], base[1], base[sy], base[sy + 1],
                             base[sz], base[sz + 1], base[sz + sy], base[sz + sy + 1]};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            e0 = fmaf(w[k], c[k].x, e0);
            e1 = fmaf(w[k], c[k].y, e1);
            e2 = fmaf(w[k], c[k].z, e2);
        }
    } else {
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = l[0] + (k & 1), j = l[1] + ((k >> 1) & 1);
            const int kk = PZ ? corner_z(l[2] + ((k >> 2) & 1), n2, true) : l[2] + ((k >> 2) & 1);
            if (i >= 0 && j >= 0 && kk >= 0 && i < n0 && j < n1 && kk < n2) {
                const float4 c = Eg[((size_t)kk * n1 + j) * n0 + i];
                e0 = fmaf(w[k], c.x, e0);
                e1 = fmaf(w[k], c.y, e1);
                e2 = fmaf(w[k], c.z, e2);
            }
        }
    }
    Ef[0] = e0;
    Ef[1] = e1;
    Ef[2] = e2;
}

}  // namespace opalsc
