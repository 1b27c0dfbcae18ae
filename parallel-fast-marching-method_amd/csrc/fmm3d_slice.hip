// fmm3d_slice.hip -- velocity-perturbation slices of a vertex model on the
// device: the GPU form of the reference's slicer (gmtplot/gmtslicet.f90:
// 4-540) behind fmm3d_slice_velocity.
//
// gmtslicet subtracts a reference vertex model from the true one, takes the
// cubic B-spline of the difference on the propagation grid with a one-node
// border in theta and phi (velv, lines 273-306), and samples one plane of
// it -- a depth, N-S or E-W slice -- with a second B-spline diced d1 x d2
// per grid cell (lines 327-531).  Here one kernel evaluates only the four
// planes of velv the slice's B-spline reaches (one thread per node), and a
// second evaluates every diced value (one thread each) from them; both sum
// in the reference's order, in float64 without contraction, so the values
// are the reference's.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/fmm3d.h"

namespace {

constexpr double PI_REF = (double)3.1415926535898f;   // gmtslicet.f90:29, a default-real literal

struct SliceGrid {
    int nvr, nvx, nvz, nnr, nnx, nnz;
    double gor, gox, goz, dvr, dvx, dvz;   // vertex origin / spacing (km, rad)
    double rgor, rgox, rgoz, dnr, dnx, dnz;
};

// B-spline basis as gmtslicet writes it (lines 283-300, 346-375)
__host__ __device__ inline void slice_basis(double u, double w[4])
{
    w[0] = (1.0 - u) * (1.0 - u) * (1.0 - u) / 6.0;
    w[1] = (4.0 - 6.0 * (u * u) + 3.0 * (u * u * u)) / 6.0;
    w[2] = (1.0 + 3.0 * u + 3.0 * (u * u) - 3.0 * (u * u * u)) / 6.0;
    w[3] = u * u * u / 6.0;
}

// velv(k, j, i) for k in [k0, k0+nk), j in [j0, j0+nj), i in [i0, i0+ni)
// (k = 1..nnr, j = 0..nnx+1, i = 0..nnz+1, gmtslicet's indices), stored
// compactly with k fastest; vv is the vertex model difference, r fastest.
// A node whose taps fall outside the vertex array sets *bad.
__global__ void k_slice_velv(SliceGrid g, const double* __restrict__ vv, int k0, int nk, int j0, int nj, int i0,
                             int ni, double* __restrict__ velv, int* __restrict__ bad)
{
    const long total = (long)nk * nj * ni;
    const int NR = g.nvr + 2, NX = g.nvx + 2, NZ = g.nvz + 2;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int k = k0 + (int)(q % nk), j = j0 + (int)((q / nk) % nj), i = i0 + (int)(q / ((long)nk * nj));
        const int ri = (int)((g.gor - g.rgor + (k - 1) * g.dnr) / g.dvr) + 1;
        const double distr = -(g.rgor - (k - 1) * g.dnr) + (g.gor - (ri - 1) * g.dvr);
        const int xi = (int)((g.rgox + (j - 1) * g.dnx - g.gox) / g.dvx) + 1;
        const double distx = g.rgox + (j - 1) * g.dnx - g.gox - (xi - 1) * g.dvx;
        const int zi = (int)((g.rgoz + (i - 1) * g.dnz - g.goz) / g.dvz) + 1;
        const double distz = g.rgoz + (i - 1) * g.dnz - g.goz - (zi - 1) * g.dvz;
        if (ri - 1 < 0 || ri + 2 >= NR || xi - 1 < 0 || xi + 2 >= NX || zi - 1 < 0 || zi + 2 >= NZ) {
            atomicOr(bad, 1);
            velv[q] = 0.0;
            continue;
        }
        double ai[4], bi[4], ci[4];
        slice_basis(distr / g.dvr, ai);
        slice_basis(distx / g.dvx, bi);
        slice_basis(distz / g.dvz, ci);
        double sumi = 0.0;
        for (int i1 = 0; i1 < 4; ++i1) {
            double sumj = 0.0;
            for (int j1 = 0; j1 < 4; ++j1) {
                const double* col = vv + ((size_t)(zi - 1 + i1) * NX + (xi - 1 + j1)) * NR;
                double sumk = 0.0;
                for (int k1 = 0; k1 < 4; ++k1) sumk = sumk + ai[k1] * col[ri - 1 + k1];
                sumj = sumj + bi[j1] * sumk;
            }
            sumi = sumi + ci[i1] * sumj;
        }
        velv[q] = sumi;
    }
}

struct SliceArgs {
    int kind;               // 0 depth, 1 N-S, 2 E-W
    int d1, d2;             // dicing along the first and second slice axes
    int c1, c2;             // cells along them (the last cell takes one value more)
    int n1, n2;             // values along them
    int rnode;              // the velv plane the slice is centred on
    double wi[4];           // weights across the plane
    const double* ui;       // (d1+1) x 4
    const double* vi;       // (d2+1) x 4
    const double* velv;     // the four planes, k fastest
    int k0, nk, j0, nj, i0; // their origin and extent
};

__device__ __forceinline__ double velv_at(const SliceArgs& a, int k, int j, int i)
{
    return a.velv[((size_t)(i - a.i0) * a.nj + (j - a.j0)) * a.nk + (k - a.k0)];
}

// One diced value per thread, in gmtslicet's summation order; out in its
// write order (outer loop over the second axis, inner over the first from
// its far end).
__global__ void k_slice_dice(SliceArgs a, double* __restrict__ out)
{
    const long total = (long)a.n1 * a.n2;
    for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total; q += (long)gridDim.x * blockDim.x) {
        const int s1 = (int)(q % a.n1) + 1, s2 = (int)(q / a.n1) + 1;   // 1-based positions along the axes
        // cell and offset in it: the cells before the last take d values, the last d+1
        const int c1 = min((s1 - 1) / a.d1 + 1, a.c1), o1 = s1 - a.d1 * (c1 - 1);
        const int c2 = min((s2 - 1) / a.d2 + 1, a.c2), o2 = s2 - a.d2 * (c2 - 1);
        const double* u = a.ui + 4 * (o1 - 1);
        const double* v = a.vi + 4 * (o2 - 1);
        double sumi = 0.0;
        if (a.kind == 0) {          // cells (j, i) = (c1, c2); lines 358-377
            const int j = c1, i = c2;
            for (int i1 = 1; i1 <= 4; ++i1) {
                double sumj = 0.0;
                for (int j1 = 1; j1 <= 4; ++j1) {
                    double sumk = 0.0;
                    for (int k1 = 1; k1 <= 4; ++k1)
                        sumk = sumk + a.wi[k1 - 1] * velv_at(a, a.rnode - 2 + k1, j - 2 + j1, i - 2 + i1);
                    sumj = sumj + u[j1 - 1] * sumk;
                }
                sumi = sumi + v[i1 - 1] * sumj;
            }
        } else if (a.kind == 1) {   // cells (k, j) = (c1 + 1, c2); lines 434-452
            const int k = c1 + 1, j = c2;
            for (int i1 = 1; i1 <= 4; ++i1) {
                double sumj = 0.0;
                for (int j1 = 1; j1 <= 4; ++j1) {
                    double sumk = 0.0;
                    for (int k1 = 1; k1 <= 4; ++k1)
                        sumk = sumk + u[k1 - 1] * velv_at(a, k - 2 + k1, j - 2 + j1, a.rnode - 2 + i1);
                    sumj = sumj + v[j1 - 1] * sumk;
                }
                sumi = sumi + a.wi[i1 - 1] * sumj;
            }
        } else {                    // cells (k, i) = (c1 + 1, c2); lines 506-523
            const int k = c1 + 1, i = c2;
            for (int i1 = 1; i1 <= 4; ++i1) {
                double sumj = 0.0;
                for (int j1 = 1; j1 <= 4; ++j1) {
                    double sumk = 0.0;
                    for (int k1 = 1; k1 <= 4; ++k1)
                        sumk = sumk + u[k1 - 1] * velv_at(a, k - 2 + k1, a.rnode - 2 + j1, i - 2 + i1);
                    sumj = sumj + a.wi[j1 - 1] * sumk;
                }
                sumi = sumi + v[i1 - 1] * sumj;
            }
        }
        out[(size_t)(s2 - 1) * a.n1 + (a.n1 - s1)] = sumi;
    }
}

thread_local std::string g_slice_err;

int slice_fail(int code, const char* msg)
{
    g_slice_err = msg;
    return code;
}

}  // namespace

extern "C" const char* fmm3d_slice_last_error(void) { return g_slice_err.c_str(); }

extern "C" int fmm3d_slice_velocity(const fmm3d_grid_desc* d, int device, const double* velv, const double* velv_ref,
                                    int kind, double where, int dice1, int dice2, double* out, int64_t cap, int* n1,
                                    int* n2)
{
    if (!d || !velv || !n1 || !n2) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: null argument");
    if (kind < 0 || kind > 2) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: kind must be 0, 1 or 2");
    if (dice1 < 1 || dice2 < 1) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: dicing must be >= 1");
    if (d->nnr < 4 || d->nnx < 2 || d->nnz < 2 || d->nvr < 1 || d->nvx < 1 || d->nvz < 1)
        return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: grid too small");
    SliceGrid g;
    g.nvr = d->nvr; g.nvx = d->nvx; g.nvz = d->nvz; g.nnr = d->nnr; g.nnx = d->nnx; g.nnz = d->nnz;
    // gmtslicet.f90:264-272: degrees to radians with its float32-rounded pi
    g.gor = d->gor; g.dvr = d->dvr;
    g.dvx = d->dvxd * PI_REF / 180.0; g.dvz = d->dvzd * PI_REF / 180.0;
    g.gox = (90.0 - d->goxd) * PI_REF / 180.0; g.goz = d->gozd * PI_REF / 180.0;
    g.rgor = d->rgor; g.dnr = d->dnr;
    g.dnx = d->dnxd * PI_REF / 180.0; g.dnz = d->dnzd * PI_REF / 180.0;
    g.rgox = (90.0 - d->rgoxd) * PI_REF / 180.0; g.rgoz = d->rgozd * PI_REF / 180.0;
    SliceArgs a{};
    a.kind = kind;
    a.d1 = dice1; a.d2 = dice2;
    double u;
    if (kind == 0) {        // depth: theta x phi (lines 327-345)
        a.rnode = (int)((d->rgor - where) / d->dnr) + 1;
        u = fabs(d->rgor - where - (a.rnode - 1) * d->dnr) / d->dnr;
        a.c1 = d->nnx - 1; a.c2 = d->nnz - 1;
    } else if (kind == 1) { // N-S at a longitude: r x theta (lines 398-432)
        a.rnode = (int)((where - d->rgozd) / d->dnzd) + 1;
        u = fabs(where - d->rgozd - (a.rnode - 1) * d->dnzd) / d->dnzd;
        a.c1 = d->nnr - 3; a.c2 = d->nnx - 1;
    } else {                // E-W at a latitude: r x phi (lines 470-504)
        a.rnode = (int)((d->rgoxd - where) / d->dnxd) + 1;
        u = fabs(d->rgoxd - where - (a.rnode - 1) * d->dnxd) / d->dnxd;
        a.c1 = d->nnr - 3; a.c2 = d->nnz - 1;
    }
    slice_basis(u, a.wi);
    a.n1 = a.c1 * dice1 + 1;
    a.n2 = a.c2 * dice2 + 1;
    *n1 = a.n1;
    *n2 = a.n2;
    if ((int64_t)a.n1 * a.n2 > cap) return slice_fail(FMM3D_E_CAPACITY, "fmm3d_slice_velocity: output buffer too small");
    if (!out) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: null output");
    // the four velv planes the slice reaches, and the range of the other two axes it reads
    if (kind == 0) {
        a.k0 = a.rnode - 1; a.nk = 4; a.j0 = 0; a.nj = d->nnx + 2; a.i0 = 0;
        if (a.k0 < 1 || a.k0 + 3 > d->nnr) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: depth slice outside the grid");
    } else if (kind == 1) {
        a.k0 = 1; a.nk = d->nnr; a.j0 = 0; a.nj = d->nnx + 2; a.i0 = a.rnode - 1;
        if (a.i0 < 0 || a.i0 + 3 > d->nnz + 1) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: N-S slice outside the grid");
    } else {
        a.k0 = 1; a.nk = d->nnr; a.j0 = a.rnode - 1; a.nj = 4; a.i0 = 0;
        if (a.j0 < 0 || a.j0 + 3 > d->nnx + 1) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: E-W slice outside the grid");
    }
    const int ni = kind == 1 ? 4 : d->nnz + 2;
    std::vector<double> ui((size_t)(dice1 + 1) * 4), vi((size_t)(dice2 + 1) * 4);
    for (int i = 1; i <= dice1 + 1; ++i) slice_basis((double)(i - 1) / (double)dice1, ui.data() + 4 * (i - 1));
    for (int i = 1; i <= dice2 + 1; ++i) slice_basis((double)(i - 1) / (double)dice2, vi.data() + 4 * (i - 1));
    const size_t nv = (size_t)(d->nvr + 2) * (d->nvx + 2) * (d->nvz + 2);
    std::vector<double> diff(velv, velv + nv);
    if (velv_ref)
        for (size_t q = 0; q < nv; ++q) diff[q] = velv[q] - velv_ref[q];
    const size_t nplane = (size_t)a.nk * a.nj * ni, nout = (size_t)a.n1 * a.n2;
    if (hipSetDevice(device) != hipSuccess) return slice_fail(FMM3D_E_DEVICE, "fmm3d_slice_velocity: hipSetDevice failed");
    hipStream_t st = nullptr;
    double *dvv = nullptr, *dvelv = nullptr, *dout = nullptr, *dui = nullptr, *dvi = nullptr;
    int* dbad = nullptr;
    int bad = 0;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&dvv, nv * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dvelv, nplane * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dout, nout * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dui, ui.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dvi, vi.size() * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&dbad, sizeof(int));
    if (e == hipSuccess) e = hipMemcpyAsync(dvv, diff.data(), nv * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dui, ui.data(), ui.size() * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dvi, vi.data(), vi.size() * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipMemsetAsync(dbad, 0, sizeof(int), st);
    if (e == hipSuccess) {
        const int blocks = (int)std::min<size_t>((nplane + 255) / 256, 65536);
        k_slice_velv<<<blocks, 256, 0, st>>>(g, dvv, a.k0, a.nk, a.j0, a.nj, a.i0, ni, dvelv, dbad);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        a.ui = dui; a.vi = dvi; a.velv = dvelv;
        const int blocks = (int)std::min<size_t>((nout + 255) / 256, 65536);
        k_slice_dice<<<blocks, 256, 0, st>>>(a, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, nout * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipMemcpyAsync(&bad, dbad, sizeof(int), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(dvv); (void)hipFree(dvelv); (void)hipFree(dout); (void)hipFree(dui); (void)hipFree(dvi); (void)hipFree(dbad);
    if (st) (void)hipStreamDestroy(st);
    if (e != hipSuccess) {
        g_slice_err = std::string("fmm3d_slice_velocity: ") + hipGetErrorString(e);
        return e == hipErrorOutOfMemory ? FMM3D_E_NOMEM : FMM3D_E_DEVICE;
    }
    if (bad) return slice_fail(FMM3D_E_ARG, "fmm3d_slice_velocity: B-spline taps outside the vertex grid");
    return FMM3D_OK;
}
