// Per-point weights of the ordered tetrapyd (device functions shared by the
// Gram kernels).
#pragma once
#include <hip/hip_runtime.h>

namespace mcmb {

__device__ __forceinline__ double permsix_d(int i, int j, int k)
{
    return (i == j && j == k) ? 1.0 : ((i == j || j == k || i == k) ? 3.0 : 6.0);
}

// Exact h^2 = (2l1+1)(2l2+1)(2l3+1)/(4 pi) * f(g-l1) f(g-l2) f(g-l3)/((2g+1) f(g)),
// f(a) = prod_{k<=a}(2k-1)/(2k), g = (l1+l2+l3)/2 (even-parity triangle only).
// ft interleaves ft[2a] = f(a) and ft[2a+1] = 1/((2a+1) f(a)) (rounded once
// from long double), so no division is made per point.  The product of the
// three odd factors is exact in a double for l < 1e5.
__device__ __forceinline__ double h2_exact(int i, int j, int k, const double* __restrict__ ft)
{
    int g = (i + j + k) >> 1;
    const long odd = (long)(2 * i + 1) * (2 * j + 1) * (2 * k + 1);
    const double pre = (double)odd * 0.079577471545947667884;  // 1/(4 pi)
    return pre * (ft[2 * (g - i)] * ft[2 * (g - j)] * ft[2 * (g - k)]) * ft[2 * g + 1];
}

// calculate_geometric.c:59-92 (the reference's asymptotic h^2)
__device__ __forceinline__ double h2_asym(int l1, int l2, int l3)
{
    double L = (double)(l1 + l2 + l3);
    double a1 = (double)(l2 + l3 - l1), a2 = (double)(l3 + l1 - l2), a3 = (double)(l1 + l2 - l3);
    double pref = 0.31830988618379067154 * 0.31830988618379067154 / 2.0;
    double p1 = pref * ((2.0 * l1 + 1.0) * (2.0 * l2 + 1.0) * (2.0 * l3 + 1.0)) / (L + 1.0);
    double p2 = (L + 1.0 / 3.0) / ((a1 + 1.0 / 3.0) * (a2 + 1.0 / 3.0) * (a3 + 1.0 / 3.0));
    double p3 = sqrt((a1 + 1.0 / 6.0) * (a2 + 1.0 / 6.0) * (a3 + 1.0 / 6.0) / (L + 1.0 / 6.0));
    return p1 * p2 * p3;
}

__device__ __forceinline__ double lspace_weight(int i, int j, int k, int asym,
                                                const double* __restrict__ f,
                                                const double* __restrict__ v)
{
    double h2 = asym ? h2_asym(i, j, k) : h2_exact(i, j, k, f);
    return permsix_d(i, j, k) * h2 * v[i] * v[j] * v[k];
}

// The exact l-space weight in the separable Gram's l3 contraction, regrouped
// as a row factor, an l3 factor and the 3j part:
//   w = [permsix (1/4pi) (2i+1)(2j+1) v_i v_j] [(2k+1) v_k]
//       [f(g-i) f(g-j)] [f(g-k) / ((2g+1) f(g))].
// The l3 factor (2k+1) v_k depends on k alone, so the contraction takes it
// with the basis (its l3-stage table holds (2k+1) v_k U(k)), and the weight
// it forms is the rest: sep_weight = row [f(g-i) f(g-j)] [f(g-k) rg], 3
// multiplies per point where lspace_weight makes 9.  It differs from
// lspace_weight by a few ulps (another rounding order); every form of the
// contraction (staged or fused weights, MFMA or VALU) uses it, so they agree
// bit for bit.
__device__ __forceinline__ double sep_row_factor(int i, int j, double pk, const double* __restrict__ v)
{
    return pk * 0.079577471545947667884 * (double)((2 * i + 1) * (2 * j + 1)) * v[i] * v[j];
}
__device__ __forceinline__ double sep_l3_factor(int k, const double* __restrict__ v)
{
    return (double)(2 * k + 1) * v[k];
}
__device__ __forceinline__ double sep_weight(double row, double f1, double f2, double f3, double rg)
{
    return row * ((f1 * f2) * (f3 * rg));
}

// k-space weight of the ordered GL point (a, b, c): permsix w_a w_b w_c / K
__device__ __forceinline__ double kspace_weight(int a, int b, int c, const double* __restrict__ kn,
                                                const double* __restrict__ kw)
{
    return permsix_d(a, b, c) * kw[a] * kw[b] * kw[c] / (kn[a] + kn[b] + kn[c]);
}

}  // namespace mcmb
