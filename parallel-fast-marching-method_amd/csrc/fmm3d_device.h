// fmm3d_device.h -- device-side numerics shared by the HIP kernels.
//
// The octant solve restates the reference's local update (fouds2_all /
// fouds1_all, fm3dt_parallel.f90:535-1004) in a generic form: per axis an
// order (0 absent, 1 first, 2 second), the first and second upwind
// neighbours and the spacing; the unknown is shifted to the leading axis
// (first first-order axis, else first second-order axis) and the quadratic
// is multiplied through by the squared spacings.  The operation order is
// identical to the CPU oracle's so the float64 path reproduces it bit for
// bit (all HIP code is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmm3d {

template <typename R> struct Lim;
template <> struct Lim<double> { static __device__ __host__ constexpr double inf() { return __builtin_huge_val(); } };
template <> struct Lim<float> { static __device__ __host__ constexpr float inf() { return __builtin_huge_valf(); } };

// One axis side of a node's stencil, prepared once per node update.
template <typename R> struct Side {
    int ord;   // -1 octant invalid (no first neighbour), 0 absent, 1, 2
    R t1, t2;
};

// Solve one octant.  ordr/ordx/ordz in {0,1,2}, not all zero.
template <typename R>
__device__ __forceinline__ R octant_solve(const Side<R>& sr, const Side<R>& sx, const Side<R>& sz,
                                          R hr, R hx, R hz, R slown)
{
    const int ord[3] = { sr.ord, sx.ord, sz.ord };
    const R t1[3] = { sr.t1, sx.t1, sz.t1 };
    const R t2[3] = { sr.t2, sx.t2, sz.t2 };
    const R h[3] = { hr, hx, hz };
    int L = -1;
#pragma unroll
    for (int a = 2; a >= 0; --a) if (ord[a] == 1) L = a;
    if (L < 0) {
#pragma unroll
        for (int a = 2; a >= 0; --a) if (ord[a] == 2) L = a;
    }
    // remaining axes: (theta, r) when phi leads, natural order otherwise
    int o0 = -1, o1 = -1;
    if (L == 2) {
        if (ord[1]) o0 = 1;
        if (ord[0]) { if (o0 < 0) o0 = 0; else o1 = 0; }
    } else {
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (a != L && ord[a]) { if (o0 < 0) o0 = a; else o1 = a; }
    }
    const R cL = ord[L] == 2 ? R(3) : R(1);
    const R DL = ord[L] == 2 ? R(4) * t1[L] - t2[L] : t1[L];
    const R HL = ord[L] == 2 ? R(2) * h[L] : h[L];
    const R u2 = HL * HL, s2 = slown * slown;
    R a, b, c;
    if (o0 < 0) {
        a = R(1); b = R(0); c = -(u2 * s2);
    } else {
        const R ca0 = ord[o0] == 2 ? R(3) : R(1);
        const R Da0 = ord[o0] == 2 ? R(4) * t1[o0] - t2[o0] : t1[o0];
        const R Ha0 = ord[o0] == 2 ? R(2) * h[o0] : h[o0];
        const R m0 = ca0 / cL;
        const R e0 = m0 * DL - Da0;
        const R H20 = Ha0 * Ha0;
        if (o1 < 0) {
            a = H20 + m0 * m0 * u2;
            b = R(2) * m0 * e0 * u2;
            c = u2 * (e0 * e0 - s2 * H20);
        } else {
            const R ca1 = ord[o1] == 2 ? R(3) : R(1);
            const R Da1 = ord[o1] == 2 ? R(4) * t1[o1] - t2[o1] : t1[o1];
            const R Ha1 = ord[o1] == 2 ? R(2) * h[o1] : h[o1];
            const R m1 = ca1 / cL;
            const R e1 = m1 * DL - Da1;
            const R H21 = Ha1 * Ha1;
            a = H20 * H21 + m0 * m0 * u2 * H21 + m1 * m1 * u2 * H20;
            b = R(2) * u2 * (m0 * e0 * H21 + m1 * e1 * H20);
            c = u2 * (e0 * e0 * H21 + e1 * e1 * H20 - s2 * H20 * H21);
        }
    }
    R disc = b * b - R(4) * a * c;
    if (disc < R(0)) disc = R(0);
    const R X = (-b + sqrt(disc)) / (R(2) * a);
    return (DL + X) / cL;
}

// Minimum over the valid octants of a node whose six sides are prepared.
template <typename R>
__device__ __forceinline__ R octant_min(const Side<R> s[3][2], R hr, R hx, R hz, R slown)
{
    R best = Lim<R>::inf();
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                const Side<R>& a = s[0][i];
                const Side<R>& b = s[1][j];
                const Side<R>& c = s[2][k];
                if (a.ord < 0 || b.ord < 0 || c.ord < 0) continue;
                if (a.ord == 0 && b.ord == 0 && c.ord == 0) continue;
                R T = octant_solve<R>(a, b, c, hr, hx, hz, slown);
                if (T < best) best = T;
            }
    return best;
}

// uniform cubic B-spline basis at fraction t (fm3dt_parallel.f90:7000-7003)
__host__ __device__ __forceinline__ void bspline4(double t, double w[4])
{
    double s = 1.0 - t, t2 = t * t, t3 = t2 * t;
    w[0] = s * s * s / 6.0;
    w[1] = (4.0 - 6.0 * t2 + 3.0 * t3) / 6.0;
    w[2] = (1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3) / 6.0;
    w[3] = t3 / 6.0;
}

}  // namespace fmm3d
