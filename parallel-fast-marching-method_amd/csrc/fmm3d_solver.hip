// fmm3d_solver.hip -- libfmm3d_hip.so: gridder, eikonal solver, receiver
// times and rays/Frechet rows on one MI355X, behind the C ABI of fmm3d.h.
//
// Solver.  The reference marches a binary heap per MPI rank (local_fmm,
// fm3dt_parallel.f90:78-155) and relaxes rank faces iteratively
// (order_overlap, F:157-416).  Here the whole grid lives in HBM and is
// solved by tile-parallel relaxation of the same discrete upwind system
// (the octant quadratic of fouds2_all / fouds1_all, F:535-1004) in three
// phases that together reproduce the heap march's result on one domain:
//   A  first-order updates, a neighbour counts iff its value is below the
//      node's (the value-based form of F:1124); a node left with no
//      counting neighbour loses its value;
//   B  mixed-order updates with the stencil choice (which neighbours count,
//      which axes go second order) frozen from the phase-A field;
//   C  as B with the choice frozen from the phase-B field, repeated until a
//      pass changes nothing.
// Each phase iterates in place over 3-D tiles; a tile is visited in
// iteration k iff it or a face neighbour changed in iteration k-1, and a
// phase ends at the first iteration with no change.
//
// Layouts on the device: fields and slowness are r-fastest,
// idx = (iz*nnx + ix)*nnr + ir (0-based); the node velocity used by rays
// keeps the reference gridder's halo (-1..n+1 per axis), r-fastest.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/fmm3d.h"
#include "fmm3d_device.h"

using namespace fmm3d;

namespace {

constexpr double PI_REF = (double)3.1415926535898f;   // fm3dt_parallel.f90:61

struct Grid {
    int nvr, nvx, nvz;
    double gor, gox, goz, dvr, dvx, dvz;
    int nnr, nnx, nnz;
    double rgor, rgox, rgoz, dnr, dnx, dnz;
    double earth;
};

// ------------------------------------------------------------ gridder

// node velocity on nodes -1..n+1 per axis (fm3dt_parallel.f90:6989-7038);
// taps beyond the vertex array are clamped into it.  Also writes slowness
// (float64 and float32) for the interior nodes in solver layout.
__global__ void k_gridder(Grid g, const double* __restrict__ velv, double* __restrict__ veln,
                          double* __restrict__ slo64, float* __restrict__ slo32)
{
    const int MR = g.nnr + 3, MX = g.nnx + 3, MZ = g.nnz + 3;
    const long total = (long)MR * MX * MZ;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        int kr = (int)(i % MR) - 1;
        int kx = (int)((i / MR) % MX) - 1;
        int kz = (int)(i / ((long)MR * MX)) - 1;
        const int NR = g.nvr + 2, NX = g.nvx + 2, NZ = g.nvz + 2;
        double h = g.rgor - (kr - 1) * g.dnr;
        int ri = (int)floor((g.gor - g.rgor + (kr - 1) * g.dnr) / g.dvr) + 1;
        double wr[4], wx[4], wz[4];
        bspline4(((g.gor - (ri - 1) * g.dvr) - h) / g.dvr, wr);
        double th = g.rgox + (kx - 1) * g.dnx;
        int xi = (int)floor((th - g.gox) / g.dvx) + 1;
        bspline4((th - g.gox - (xi - 1) * g.dvx) / g.dvx, wx);
        double ph = g.rgoz + (kz - 1) * g.dnz;
        int zi = (int)floor((ph - g.goz) / g.dvz) + 1;
        bspline4((ph - g.goz - (zi - 1) * g.dvz) / g.dvz, wz);
        double acc_z = 0.0;
        for (int a = 0; a < 4; ++a) {
            int iz = min(max(zi - 1 + a, 0), NZ - 1);
            double acc_x = 0.0;
            for (int b = 0; b < 4; ++b) {
                int ix = min(max(xi - 1 + b, 0), NX - 1);
                const double* col = velv + ((size_t)iz * NX + ix) * NR;
                double acc_r = 0.0;
                for (int c = 0; c < 4; ++c) acc_r += wr[c] * col[min(max(ri - 1 + c, 0), NR - 1)];
                acc_x += wx[b] * acc_r;
            }
            acc_z += wz[a] * acc_x;
        }
        veln[i] = acc_z;
        if (kr >= 1 && kr <= g.nnr && kx >= 1 && kx <= g.nnx && kz >= 1 && kz <= g.nnz) {
            size_t s = ((size_t)(kz - 1) * g.nnx + (kx - 1)) * g.nnr + (kr - 1);
            double sl = 1.0 / acc_z;
            slo64[s] = sl;
            slo32[s] = (float)sl;
        }
    }
}

// ------------------------------------------------------------ solver sweeps

template <typename R> struct SweepArgs {
    R* T;                   // fields, slot s at T + s*N
    const R* P;             // frozen ordering fields (phases B/C)
    const R* slo;           // slowness, N
    const R* hx;            // per ir: ri*dnx
    const R* hz;            // per (ix, ir): ri*sin(theta)*dnz, index ix*nnr + ir
    R hr;                   // dnr
    const uint32_t* fixbits;// per slot: N bits, fixed (Dirichlet) nodes
    int* stamp;             // per slot per tile: last iteration with a change
    int* lastchg;           // per slot: last iteration with a change
    const int* slot_on;     // per slot: 1 while iterating
    int nnr, nnx, nnz;
    int ntr, ntx, ntz;      // tiles per axis
    int nslots;
    int iter;               // iteration number k (>= 1)
    int phase;              // 0: A (first order, value-based); 1: B/C (frozen ordering)
    int fom;                // second order allowed in phase 1
    int64_t* updates;       // node updates counter (per launch accumulation)
};

template <typename R, bool SHIFT>
__device__ __forceinline__ R solve_node(const SweepArgs<R>& a, const R* __restrict__ T, const R* __restrict__ P,
                                        int ir, int ix, int iz, size_t c, R tc)
{
    const int n[3] = { a.nnr, a.nnx, a.nnz };
    const int pos[3] = { ir, ix, iz };
    const size_t stride[3] = { 1, (size_t)a.nnr, (size_t)a.nnr * a.nnx };
    Side<R> s[3][2];
    const R inf = Lim<R>::inf();
    const R pc = a.phase ? P[c] : R(0);
    R tb = inf;
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) {
#pragma unroll
        for (int sd = 0; sd < 2; ++sd) {
            const int d = sd ? 1 : -1;
            Side<R>& S = s[ax][sd];
            S.ord = 0; S.t1 = R(0); S.t2 = R(0);
            const int p1 = pos[ax] + d;
            if (p1 < 0 || p1 >= n[ax]) { S.ord = -1; continue; }
            const size_t i1 = d > 0 ? c + stride[ax] : c - stride[ax];
            const R v1 = T[i1];
            if (a.phase == 0) {
                if (!(v1 < tc)) continue;
                S.ord = 1; S.t1 = v1;
            } else {
                const R q1 = P[i1];
                if (!(q1 < pc) || !(v1 < inf)) continue;
                S.ord = 1; S.t1 = v1;
                const int p2 = pos[ax] + 2 * d;
                if (a.fom && p2 >= 0 && p2 < n[ax]) {
                    const size_t i2 = d > 0 ? i1 + stride[ax] : i1 - stride[ax];
                    const R q2 = P[i2], v2 = T[i2];
                    if (q2 < inf && v2 < inf && q1 > q2) { S.ord = 2; S.t2 = v2; }
                }
            }
            if (SHIFT && S.t1 < tb) tb = S.t1;
        }
    }
    if (SHIFT) {
        if (!(tb < inf)) return inf;
#pragma unroll
        for (int ax = 0; ax < 3; ++ax)
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) { s[ax][sd].t1 -= tb; s[ax][sd].t2 -= tb; }
    }
    const R hx = a.hx[ir];
    const R hz = a.hz[(size_t)ix * a.nnr + ir];
    R best = octant_min<R>(s, a.hr, hx, hz, a.slo[c]);
    if (SHIFT && best < inf) best = tb + best;
    return best;
}

template <typename R, int TR, int TX, int TZ, bool SHIFT>
__global__ void __launch_bounds__(TR * TX * TZ)
k_sweep(SweepArgs<R> a)
{
    const int ntiles = a.ntr * a.ntx * a.ntz;
    const int slot = blockIdx.x / ntiles;
    const int tile = blockIdx.x - slot * ntiles;
    if (slot >= a.nslots || !a.slot_on[slot]) return;
    const int tr = tile % a.ntr, tx = (tile / a.ntr) % a.ntx, tz = tile / (a.ntr * a.ntx);
    int* st = a.stamp + (size_t)slot * ntiles;
    // activity: this tile or a face neighbour changed in the previous iteration
    __shared__ int s_act, s_chg;
    if (threadIdx.x == 0) {
        int m = st[tile];
        if (tr > 0) m = max(m, st[tile - 1]);
        if (tr < a.ntr - 1) m = max(m, st[tile + 1]);
        if (tx > 0) m = max(m, st[tile - a.ntr]);
        if (tx < a.ntx - 1) m = max(m, st[tile + a.ntr]);
        if (tz > 0) m = max(m, st[tile - a.ntr * a.ntx]);
        if (tz < a.ntz - 1) m = max(m, st[tile + a.ntr * a.ntx]);
        s_act = (m >= a.iter - 1);
        s_chg = 0;
    }
    __syncthreads();
    if (!s_act) return;
    const size_t N = (size_t)a.nnr * a.nnx * a.nnz;
    R* T = a.T + (size_t)slot * N;
    const R* P = a.phase ? a.P + (size_t)slot * N : nullptr;
    const uint32_t* fb = a.fixbits + (size_t)slot * ((N + 31) / 32);
    const int lr = threadIdx.x % TR, lx = (threadIdx.x / TR) % TX, lz = threadIdx.x / (TR * TX);
    const int ir = tr * TR + lr, ix = tx * TX + lx, iz = tz * TZ + lz;
    int changed = 0;
    if (ir < a.nnr && ix < a.nnx && iz < a.nnz) {
        const size_t c = ((size_t)iz * a.nnx + ix) * a.nnr + ir;
        if (!((fb[c >> 5] >> (c & 31)) & 1u)) {
            const R tc = T[c];
            const R nv = solve_node<R, SHIFT>(a, T, P, ir, ix, iz, c, tc);
            if (nv != tc) { T[c] = nv; changed = 1; }
        }
    }
    if (changed) s_chg = 1;   // benign race: all writers store 1
    __syncthreads();
    if (threadIdx.x == 0 && s_chg) {
        st[tile] = a.iter;
        atomicMax(a.lastchg + slot, a.iter);
    }
}

// ------------------------------------------------------------ receivers

template <typename R>
__device__ __forceinline__ double field_at(const R* T, double toff, size_t i) { return toff + (double)T[i]; }
template <>
__device__ __forceinline__ double field_at<double>(const double* T, double, size_t i) { return T[i]; }

// trilinear receiver time with a float32 accumulator (fm3dt_parallel.f90:7063-7106)
template <typename R>
__global__ void k_receivers(Grid g, const R* __restrict__ T, const double* __restrict__ toff, int nslots,
                            int nrc, const double* __restrict__ rcv, float* __restrict__ out)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nslots * nrc) return;
    int s = gid / nrc, i = gid % nrc;
    const size_t N = (size_t)g.nnr * g.nnx * g.nnz;
    const R* F = T + (size_t)s * N;
    double to = toff[s];
    double rr = rcv[3 * i], rx = rcv[3 * i + 1], rz = rcv[3 * i + 2];
    int irr = (int)floor((g.rgor - rr) / g.dnr) + 1;
    int irx = (int)floor((rx - g.rgox) / g.dnx) + 1;
    int irz = (int)floor((rz - g.rgoz) / g.dnz) + 1;
    if (irx == g.nnx) irx--;
    if (irz == g.nnz) irz--;
    if (irr < 1 || irr >= g.nnr || irx < 1 || irx >= g.nnx || irz < 1 || irz >= g.nnz) {
        out[gid] = __builtin_nanf("");
        return;
    }
    double drx = (rx - g.rgox) - (irx - 1) * g.dnx;
    double drz = (rz - g.rgoz) - (irz - 1) * g.dnz;
    double drr = (g.rgor - rr) - (irr - 1) * g.dnr;
    float acc = 0.0f;
    for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l) {
                double w = (1.0 - fabs((l * g.dnz - drz) / g.dnz)) * (1.0 - fabs((k * g.dnx - drx) / g.dnx));
                w = w * (1.0 - fabs((j * g.dnr - drr) / g.dnr));
                size_t idx = ((size_t)(irz - 1 + l) * g.nnx + (irx - 1 + k)) * g.nnr + (irr - 1 + j);
                acc = (float)((double)acc + field_at<R>(F, to, idx) * w);
            }
    out[gid] = acc;
}

// ------------------------------------------------------------ rays + Frechet

__device__ __forceinline__ double tri_vel(const Grid& g, const double* __restrict__ veln, int ipr, int ipx, int ipz,
                                          double drr, double drx, double drz)
{
    const int MR = g.nnr + 3, MX = g.nnx + 3;
    double vel = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l)
            for (int m = 0; m < 2; ++m) {
                double w = 1.0 - fabs((m * g.dnz - drz) / g.dnz);
                w = w * (1.0 - fabs((l * g.dnx - drx) / g.dnx));
                w = w * (1.0 - fabs((k * g.dnr - drr) / g.dnr));
                int zr = ipr + k, zx = ipx + l, zz = ipz + m;
                if (zz <= g.nnz + 1 && zx <= g.nnx + 1 && zr <= g.nnr + 1 && zr >= -1 && zx >= -1 && zz >= -1)
                    vel += veln[((size_t)(zz + 1) * MX + (zx + 1)) * MR + (zr + 1)] * w;
            }
    return vel;
}

struct RayOut {
    int32_t* cols;      // staging, ray r at cols + r*cap
    float* vals;
    int32_t* count;     // entries per ray (-1 = start outside the grid, -2 = overflow)
    int64_t cap;
};

// One wavefront per ray.  The ray is stepped redundantly by all 64 lanes;
// each lane owns the vertices whose (r, theta, phi) indices are congruent to
// its (lane/16, lane/4 % 4, lane % 4) mod 4, so the 64 taps of every
// sub-segment land on 64 different lanes and every vertex is only ever
// accumulated by one lane (no cross-lane ordering is needed).  The dense
// per-worker Frechet scratch is scanned over the touched box at the end.
template <typename R>
__global__ void __launch_bounds__(64)
k_rays(Grid g, const R* __restrict__ T, const double* __restrict__ toff, const double* __restrict__ veln,
       int nslots, int nrc, const double* __restrict__ rcv, int ray0, int nrays, float* __restrict__ scratch,
       RayOut out)
{
    const int lane = threadIdx.x;
    const int NVR = g.nvr + 2, NVX = g.nvx + 2, NVZ = g.nvz + 2;
    const size_t NV = (size_t)NVR * NVX * NVZ;
    float* fdm = scratch + (size_t)blockIdx.x * NV;
    const int own_r = lane / 16, own_x = (lane / 4) % 4, own_z = lane % 4;
    for (int rid = ray0 + blockIdx.x; rid < ray0 + nrays; rid += gridDim.x) {
        const int s = rid / nrc, i = rid % nrc;
        const size_t N = (size_t)g.nnr * g.nnx * g.nnz;
        const R* F = T + (size_t)s * N;
        const double to = toff[s];
        auto TT = [&](int r, int x, int z) -> double {   // 1-based node access
            return field_at<R>(F, to, ((size_t)(z - 1) * g.nnx + (x - 1)) * g.nnr + (r - 1));
        };
        const int64_t orow = (int64_t)(rid - ray0) * out.cap;
        double r0 = rcv[3 * i], x0 = rcv[3 * i + 1], z0 = rcv[3 * i + 2];
        int ipr = (int)floor((g.rgor - r0) / g.dnr) + 1;
        int ipx = (int)floor((x0 - g.rgox) / g.dnx) + 1;
        int ipz = (int)floor((z0 - g.rgoz) / g.dnz) + 1;
        int sw = 0;
        if (ipr < 1 || ipr >= g.nnr) sw = 1;
        if (ipx < 1 || ipx >= g.nnx) sw = 1;
        if (ipz < 1 || ipz >= g.nnz) sw = 1;
        if (ipr == g.nnr) { ipr--; sw = 0; }
        if (ipx == g.nnx) { ipx--; sw = 0; }
        if (ipz == g.nnz) { ipz--; sw = 0; }
        if (sw) { if (lane == 0) out.count[rid - ray0] = -1; continue; }
        int bmin[3] = { 1 << 30, 1 << 30, 1 << 30 }, bmax[3] = { -1, -1, -1 };
        double pr = r0, px = x0, pz = z0;
        const double dpl = g.dnr;
        for (int step = 1; step <= 100000; ++step) {
            double dtr = 0.0;
            int cnt = 0;
            const int cz[4] = { 0, 1, 1, 0 }, cx[4] = { 0, 0, 1, 1 };
            for (int q = 0; q < 4; ++q) {
                double lo = TT(ipr + 1, ipx + cx[q], ipz + cz[q]);
                double hi = TT(ipr, ipx + cx[q], ipz + cz[q]);
                if (lo < hi) { dtr = dtr + lo - hi; cnt++; }
            }
            if (cnt) dtr = dtr / (cnt * g.dnr);
            double dtx = TT(ipr, ipx + 1, ipz) - TT(ipr, ipx, ipz);
            dtx = dtx + TT(ipr, ipx + 1, ipz + 1) - TT(ipr, ipx, ipz + 1);
            dtx = dtx + TT(ipr + 1, ipx + 1, ipz + 1) - TT(ipr + 1, ipx, ipz + 1);
            dtx = dtx + TT(ipr + 1, ipx + 1, ipz) - TT(ipr + 1, ipx, ipz);
            double rcell = g.earth + g.rgor - (ipr - 1) * g.dnr;
            dtx = dtx / (4.0 * rcell * g.dnx);
            double dtz = TT(ipr, ipx, ipz + 1) - TT(ipr, ipx, ipz);
            dtz = dtz + TT(ipr + 1, ipx + 1, ipz + 1) - TT(ipr + 1, ipx + 1, ipz);
            dtz = dtz + TT(ipr, ipx + 1, ipz + 1) - TT(ipr, ipx + 1, ipz);
            dtz = dtz + TT(ipr + 1, ipx, ipz + 1) - TT(ipr + 1, ipx, ipz);
            const double sx = sin(px);
            dtz = dtz / (4.0 * rcell * sx * g.dnz);
            double crad = g.earth + pr;
            double gn = sqrt(dtr * dtr + dtx * dtx + dtz * dtz);
            double nr_ = pr + dpl * dtr / gn;
            double nx_ = px - dpl * dtx / (crad * gn);
            double nz_ = pz - dpl * dtz / (crad * sx * gn);
            int ipro = ipr, ipxo = ipx, ipzo = ipz;
            ipr = (int)floor((g.rgor - nr_) / g.dnr) + 1;
            ipx = (int)floor((nx_ - g.rgox) / g.dnx) + 1;
            ipz = (int)floor((nz_ - g.rgoz) / g.dnz) + 1;
            int exr = (ipr < 1 || ipr >= g.nnr);
            int exx = (ipx < 1 || ipx >= g.nnx);
            int exz = (ipz < 1 || ipz >= g.nnz);
            int done = 0;
            if (exr || exx || exz) {
                // clip to the first face crossed (F:7463-7538)
                int which = 0;
                double atio = 0.0;
                if (exr) {
                    double rb;
                    if (ipr < 1) { rb = g.rgor; ipr = 1; }
                    else { rb = g.rgor - (g.nnr - 1) * g.dnr; ipr = g.nnr - 1; }
                    atio = (rb - pr) / (nr_ - pr);
                    which = 1;
                }
                if (exx) {
                    double xb;
                    if (ipx < 1) { xb = g.rgox; ipx = 1; }
                    else { xb = g.rgox + (g.nnx - 1) * g.dnx; ipx = g.nnx - 1; }
                    double q = (xb - px) / (nx_ - px);
                    if (which == 1) { if (q < atio) { atio = q; which = 2; } }
                    else { atio = q; which = 2; }
                }
                if (exz) {
                    double zb;
                    if (ipz < 1) { zb = g.rgoz; ipz = 1; }
                    else { zb = g.rgoz + (g.nnz - 1) * g.dnz; ipz = g.nnz - 1; }
                    double q = (zb - pz) / (nz_ - pz);
                    if (which != 0) { if (q < atio) { atio = q; which = 3; } }
                    else { atio = q; which = 3; }
                }
                if (which == 1) { ipx = ipxo; ipz = ipzo; }
                else if (which == 2) { ipr = ipro; ipz = ipzo; }
                else if (which == 3) { ipr = ipro; ipx = ipxo; }
                nz_ = pz + atio * (nz_ - pz);
                nx_ = px + atio * (nx_ - px);
                nr_ = pr + atio * (nr_ - pr);
                done = 1;
            }
            // Frechet contribution of this step (F:7542-7737)
            int ivr = (int)floor((g.gor - nr_) / g.dvr) + 1;
            int ivx = (int)floor((nx_ - g.gox) / g.dvx) + 1;
            int ivz = (int)floor((nz_ - g.goz) / g.dvz) + 1;
            int ivro = (int)floor((g.gor - pr) / g.dvr) + 1;
            int ivxo = (int)floor((px - g.gox) / g.dvx) + 1;
            int ivzo = (int)floor((pz - g.goz) / g.dvz) + 1;
            int nhp = 0;
            double vrat[4];
            int chp[4];
            if (ivr != ivro) {
                double rb = (ivr > ivro) ? g.gor - (ivr - 1) * g.dvr : g.gor - ivr * g.dvr;
                vrat[nhp] = (rb - pr) / (nr_ - pr); chp[nhp] = 1; nhp++;
            }
            if (ivx != ivxo) {
                double xb = (ivx > ivxo) ? g.gox + (ivx - 1) * g.dvx : g.gox + ivx * g.dvx;
                double q = (xb - px) / (nx_ - px);
                int p = nhp;
                while (p > 0 && q < vrat[p - 1]) { vrat[p] = vrat[p - 1]; chp[p] = chp[p - 1]; p--; }
                vrat[p] = q; chp[p] = 2; nhp++;
            }
            if (ivz != ivzo) {
                double zb = (ivz > ivzo) ? g.goz + (ivz - 1) * g.dvz : g.goz + ivz * g.dvz;
                double q = (zb - pz) / (nz_ - pz);
                int p = nhp;
                while (p > 0 && q < vrat[p - 1]) { vrat[p] = vrat[p - 1]; chp[p] = chp[p - 1]; p--; }
                vrat[p] = q; chp[p] = 3; nhp++;
            }
            vrat[nhp] = 1.0; chp[nhp] = 0; nhp++;
            double vel = tri_vel(g, veln, ipro, ipxo, ipzo,
                                 (g.rgor - pr) - (ipro - 1) * g.dnr,
                                 (px - g.rgox) - (ipxo - 1) * g.dnx,
                                 (pz - g.rgoz) - (ipzo - 1) * g.dnz);
            double ui[4], vi[4], wi[4];
            bspline4(((g.gor - pr) - (ivro - 1) * g.dvr) / g.dvr, ui);
            bspline4(((px - g.gox) - (ivxo - 1) * g.dvx) / g.dvx, vi);
            bspline4(((pz - g.goz) - (ivzo - 1) * g.dvz) / g.dvz, wi);
            int ivrt = ivro, ivxt = ivxo, ivzt = ivzo;
            for (int k = 0; k < nhp; ++k) {
                double velo = vel, uio[4], vio[4], wio[4];
                for (int q = 0; q < 4; ++q) { uio[q] = ui[q]; vio[q] = vi[q]; wio[q] = wi[q]; }
                if (k > 0) {
                    if (chp[k - 1] == 1) ivrt = ivr;
                    else if (chp[k - 1] == 2) ivxt = ivx;
                    else if (chp[k - 1] == 3) ivzt = ivz;
                }
                double sz_ = pz + vrat[k] * (nz_ - pz);
                double sx_ = px + vrat[k] * (nx_ - px);
                double sr_ = pr + vrat[k] * (nr_ - pr);
                int iprt = (int)floor((g.rgor - sr_) / g.dnr) + 1;
                int ipxt = (int)floor((sx_ - g.rgox) / g.dnx) + 1;
                int ipzt = (int)floor((sz_ - g.rgoz) / g.dnz) + 1;
                vel = tri_vel(g, veln, iprt, ipxt, ipzt,
                              (g.rgor - sr_) - (iprt - 1) * g.dnr,
                              (sx_ - g.rgox) - (ipxt - 1) * g.dnx,
                              (sz_ - g.rgoz) - (ipzt - 1) * g.dnz);
                bspline4(((g.gor - sr_) - (ivrt - 1) * g.dvr) / g.dvr, ui);
                bspline4(((sx_ - g.gox) - (ivxt - 1) * g.dvx) / g.dvx, vi);
                bspline4(((sz_ - g.goz) - (ivzt - 1) * g.dvz) / g.dvz, wi);
                double dinc = (k == 0) ? vrat[0] * dpl : (vrat[k] - vrat[k - 1]) * dpl;
                // this lane's tap: vertex (ivt-1+n) with n the lane's residue offset
                int rr = ivrt - 1 + (((own_r - (ivrt - 1)) % 4 + 4) % 4);
                int xx = ivxt - 1 + (((own_x - (ivxt - 1)) % 4 + 4) % 4);
                int zz = ivzt - 1 + (((own_z - (ivzt - 1)) % 4 + 4) % 4);
                int n = rr - (ivrt - 1), m = xx - (ivxt - 1), l = zz - (ivzt - 1);
                double e1 = ui[n] * vi[m] * wi[l] / (vel * vel);
                double e0 = uio[n] * vio[m] * wio[l] / (velo * velo);
                double inc = -(e1 + e0) * dinc / 2.0;
                if (zz >= 0 && zz < NVZ && xx >= 0 && xx < NVX && rr >= 0 && rr < NVR) {
                    float* p = fdm + ((size_t)zz * NVX + xx) * NVR + rr;
                    *p = (float)(inc + (double)*p);
                }
                bmin[0] = min(bmin[0], max(ivrt - 1, 0)); bmax[0] = max(bmax[0], min(ivrt + 2, NVR - 1));
                bmin[1] = min(bmin[1], max(ivxt - 1, 0)); bmax[1] = max(bmax[1], min(ivxt + 2, NVX - 1));
                bmin[2] = min(bmin[2], max(ivzt - 1, 0)); bmax[2] = max(bmax[2], min(ivzt + 2, NVZ - 1));
            }
            pr = nr_; px = nx_; pz = nz_;
            if (done) break;
        }
        __syncthreads();
        // sparsify the touched box in column order (F:7739-7764)
        const int BR = bmax[0] - bmin[0] + 1, BX = bmax[1] - bmin[1] + 1, BZ = bmax[2] - bmin[2] + 1;
        const int nbox = (bmax[0] >= 0) ? BR * BX * BZ : 0;
        int64_t written = 0;
        bool overflow = false;
        for (int base = 0; base < nbox; base += 64) {
            int e = base + lane;
            float v = 0.0f;
            size_t vi_ = 0;
            if (e < nbox) {
                int br = e % BR, bx = (e / BR) % BX, bz = e / (BR * BX);
                vi_ = ((size_t)(bmin[2] + bz) * NVX + (bmin[1] + bx)) * NVR + (bmin[0] + br);
                v = fdm[vi_];
                if (v != 0.0f) fdm[vi_] = 0.0f;
            }
            bool keep = (e < nbox) && (fabsf(v) >= 1.0e-6f);
            uint64_t mask = __ballot(keep);
            int before = __popcll(mask & ((1ull << lane) - 1ull));
            if (keep) {
                int64_t pos = written + before;
                if (pos < out.cap) {
                    out.cols[orow + pos] = (int32_t)(vi_ + 1);
                    out.vals[orow + pos] = v;
                } else {
                    overflow = true;
                }
            }
            written += __popcll(mask);
        }
        uint64_t ov = __ballot(overflow);
        if (lane == 0) out.count[rid - ray0] = ov ? -2 : (int32_t)written;
        __syncthreads();
    }
}

// ------------------------------------------------------------ context

struct Slots {
    int cap = 0;           // allocated slots
    int n = 0;             // resident sources
    int prec = FMM3D_PREC_FP64;
    void* T = nullptr;     // fields
    void* P = nullptr;     // ordering fields
    uint32_t* fix = nullptr;
    int* stamp = nullptr;
    int* lastchg = nullptr;
    int* on = nullptr;
    double* toff = nullptr;
};

}  // namespace

struct fmm3d_ctx {
    Grid g;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::string err;
    size_t N = 0;
    double* veln = nullptr;      // (nnr+3)(nnx+3)(nnz+3)
    double* slo64 = nullptr;
    float* slo32 = nullptr;
    double* hx64 = nullptr;
    double* hz64 = nullptr;
    float* hx32 = nullptr;
    float* hz32 = nullptr;
    bool have_model = false;
    Slots sl;
    int tile_r = 16, tile_x = 4, tile_z = 4;
    fmm3d_stats stats{};
    int64_t* d_updates = nullptr;
    float* ray_scratch = nullptr;
    size_t ray_scratch_workers = 0;
};

static std::string g_create_err;

static int fail(fmm3d_ctx* c, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->err = buf; else g_create_err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), e_ == hipErrorOutOfMemory ? FMM3D_E_NOMEM : FMM3D_E_DEVICE,   \
                        "%s: %s", #expr, hipGetErrorString(e_));                           \
    } while (0)

template <typename P> static hipError_t dalloc(P** p, size_t bytes)
{
    return hipMalloc((void**)p, bytes ? bytes : 16);
}

extern "C" const char* fmm3d_last_error(const fmm3d_ctx* c)
{
    return c ? c->err.c_str() : g_create_err.c_str();
}

extern "C" int fmm3d_create(const fmm3d_grid_desc* d, int device, fmm3d_ctx** out)
{
    if (!d || !out) return fail(nullptr, FMM3D_E_ARG, "fmm3d_create: null argument");
    *out = nullptr;
    if (d->nnr < 3 || d->nnx < 3 || d->nnz < 3)
        return fail(nullptr, FMM3D_E_ARG, "fmm3d_create: propagation grid must have >= 3 nodes per axis");
    if (d->nvr < 2 || d->nvx < 2 || d->nvz < 2)
        return fail(nullptr, FMM3D_E_ARG, "fmm3d_create: vertex grid must have >= 2 vertices per axis");
    if (!(d->dnr > 0 && d->dnxd > 0 && d->dnzd > 0 && d->dvr > 0 && d->dvxd > 0 && d->dvzd > 0 && d->earth > 0))
        return fail(nullptr, FMM3D_E_ARG, "fmm3d_create: spacings and earth radius must be positive");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, FMM3D_E_DEVICE, "fmm3d_create: no HIP device");
    if (device < 0 || device >= ndev)
        return fail(nullptr, FMM3D_E_ARG, "fmm3d_create: device %d out of range (%d devices)", device, ndev);
    fmm3d_ctx* c = new fmm3d_ctx();
    Grid& g = c->g;
    g.nvr = d->nvr; g.nvx = d->nvx; g.nvz = d->nvz;
    g.gor = d->gor; g.dvr = d->dvr;
    g.dvx = d->dvxd * PI_REF / 180.0;
    g.dvz = d->dvzd * PI_REF / 180.0;
    g.gox = (90.0 - d->goxd) * PI_REF / 180.0;
    g.goz = d->gozd * PI_REF / 180.0;
    g.nnr = d->nnr; g.nnx = d->nnx; g.nnz = d->nnz;
    g.rgor = d->rgor; g.dnr = d->dnr;
    g.dnx = d->dnxd * PI_REF / 180.0;
    g.dnz = d->dnzd * PI_REF / 180.0;
    g.rgox = (90.0 - d->rgoxd) * PI_REF / 180.0;
    g.rgoz = d->rgozd * PI_REF / 180.0;
    g.earth = d->earth;
    c->device = device;
    c->N = (size_t)g.nnr * g.nnx * g.nnz;
    int rc;
    if (hipSetDevice(device) != hipSuccess) { delete c; return fail(nullptr, FMM3D_E_DEVICE, "hipSetDevice failed"); }
#define CREATECHK(expr)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (expr);                                                                \
        if (e_ != hipSuccess) {                                                                \
            rc = fail(nullptr, e_ == hipErrorOutOfMemory ? FMM3D_E_NOMEM : FMM3D_E_DEVICE,    \
                      "%s: %s", #expr, hipGetErrorString(e_));                                 \
            fmm3d_destroy(c);                                                                  \
            return rc;                                                                         \
        }                                                                                      \
    } while (0)
    CREATECHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CREATECHK(hipEventCreate(&c->ev0));
    CREATECHK(hipEventCreate(&c->ev1));
    size_t NH = (size_t)(g.nnr + 3) * (g.nnx + 3) * (g.nnz + 3);
    CREATECHK(dalloc(&c->veln, NH * sizeof(double)));
    CREATECHK(dalloc(&c->slo64, c->N * sizeof(double)));
    CREATECHK(dalloc(&c->slo32, c->N * sizeof(float)));
    CREATECHK(dalloc(&c->d_updates, sizeof(int64_t)));
    // metric tables in the reference's arithmetic (F:541-542)
    std::vector<double> hx(g.nnr), hz((size_t)g.nnx * g.nnr);
    std::vector<float> hx32(g.nnr), hz32((size_t)g.nnx * g.nnr);
    for (int ir = 0; ir < g.nnr; ++ir) {
        double ri = g.rgor - (ir) * g.dnr + g.earth;
        hx[ir] = ri * g.dnx;
        hx32[ir] = (float)hx[ir];
        for (int ix = 0; ix < g.nnx; ++ix) {
            double sth = sin(g.rgox + (ix) * g.dnx);
            hz[(size_t)ix * g.nnr + ir] = ri * sth * g.dnz;
            hz32[(size_t)ix * g.nnr + ir] = (float)hz[(size_t)ix * g.nnr + ir];
        }
    }
    CREATECHK(dalloc(&c->hx64, hx.size() * sizeof(double)));
    CREATECHK(dalloc(&c->hz64, hz.size() * sizeof(double)));
    CREATECHK(dalloc(&c->hx32, hx32.size() * sizeof(float)));
    CREATECHK(dalloc(&c->hz32, hz32.size() * sizeof(float)));
    CREATECHK(hipMemcpy(c->hx64, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->hz64, hz.data(), hz.size() * sizeof(double), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->hx32, hx32.data(), hx32.size() * sizeof(float), hipMemcpyHostToDevice));
    CREATECHK(hipMemcpy(c->hz32, hz32.data(), hz32.size() * sizeof(float), hipMemcpyHostToDevice));
#undef CREATECHK
    const char* ts = getenv("FMM3D_TILE");
    if (ts) {
        int a, b, e;
        if (sscanf(ts, "%dx%dx%d", &a, &b, &e) == 3) { c->tile_r = a; c->tile_x = b; c->tile_z = e; }
    }
    *out = c;
    return FMM3D_OK;
}

static void free_slots(Slots& s)
{
    (void)hipFree(s.T); (void)hipFree(s.P); (void)hipFree(s.fix); (void)hipFree(s.stamp);
    (void)hipFree(s.lastchg); (void)hipFree(s.on); (void)hipFree(s.toff);
    s = Slots();
}

extern "C" int fmm3d_destroy(fmm3d_ctx* c)
{
    if (!c) return FMM3D_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_slots(c->sl);
    (void)hipFree(c->veln); (void)hipFree(c->slo64); (void)hipFree(c->slo32);
    (void)hipFree(c->hx64); (void)hipFree(c->hz64); (void)hipFree(c->hx32); (void)hipFree(c->hz32);
    (void)hipFree(c->d_updates); (void)hipFree(c->ray_scratch);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return FMM3D_OK;
}

extern "C" int fmm3d_set_model_vtx(fmm3d_ctx* c, const double* velv)
{
    if (!c || !velv) return fail(c, FMM3D_E_ARG, "fmm3d_set_model_vtx: null argument");
    HIPCHK(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    size_t nv = (size_t)(g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2);
    for (size_t i = 0; i < nv; ++i)
        if (!(velv[i] > 0.0) || !isfinite(velv[i]))
            return fail(c, FMM3D_E_ARG, "fmm3d_set_model_vtx: vertex %zu has non-positive velocity %g", i, velv[i]);
    double* d = nullptr;
    HIPCHK(c, dalloc(&d, nv * sizeof(double)));
    hipError_t e = hipMemcpyAsync(d, velv, nv * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        long total = (long)(g.nnr + 3) * (g.nnx + 3) * (g.nnz + 3);
        int blocks = (int)std::min<long>((total + 255) / 256, 65536);
        k_gridder<<<blocks, 256, 0, c->stream>>>(g, d, c->veln, c->slo64, c->slo32);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "gridder: %s", hipGetErrorString(e));
    c->have_model = true;
    return FMM3D_OK;
}

static int ensure_slots(fmm3d_ctx* c, int nsrc, int prec)
{
    Slots& s = c->sl;
    if (s.cap >= nsrc && s.prec == prec) return FMM3D_OK;
    free_slots(s);
    const size_t N = c->N;
    const size_t es = prec == FMM3D_PREC_FP32 ? sizeof(float) : sizeof(double);
    // stamps for the largest tile count of any supported shape (4x4x4 is the smallest tile)
    const size_t ntiles = (size_t)((c->g.nnr + 3) / 4) * ((c->g.nnx + 3) / 4) * ((c->g.nnz + 3) / 4);
    HIPCHK(c, dalloc(&s.T, (size_t)nsrc * N * es));
    HIPCHK(c, dalloc(&s.P, (size_t)nsrc * N * es));
    HIPCHK(c, dalloc(&s.fix, (size_t)nsrc * ((N + 31) / 32) * sizeof(uint32_t)));
    // tile stamps for the largest tile count any supported shape gives
    HIPCHK(c, dalloc(&s.stamp, (size_t)nsrc * ntiles * sizeof(int)));
    HIPCHK(c, dalloc(&s.lastchg, nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.on, nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.toff, nsrc * sizeof(double)));
    s.cap = nsrc;
    s.prec = prec;
    return FMM3D_OK;
}

// ------------------------------------------------------------ solve driver

__global__ void k_init_tele64(int nnr, int nnx, int nnz, int nslots, const double* faces, double* T,
                              uint32_t* fix, double* toff)
{
    const size_t N = (size_t)nnr * nnx * nnz;
    const size_t NW = (N + 31) / 32;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N * nslots; i += (size_t)gridDim.x * blockDim.x) {
        size_t s = i / N, c = i % N;
        int ir = (int)(c % nnr);
        size_t col = c / nnr;          // iz*nnx + ix
        int ix = (int)(col % nnx), iz = (int)(col / nnx);
        double v = __builtin_huge_val();
        if (ir == nnr - 1) {
            double f = faces[s * (size_t)nnx * nnz + (size_t)ix * nnz + iz];
            if (f > 0.0) v = f;
        }
        T[i] = v;
        (void)fix; (void)NW; (void)toff;
    }
}

__global__ void k_init_tele32(int nnr, int nnx, int nnz, int nslots, const double* faces, float* T,
                              const double* toff)
{
    const size_t N = (size_t)nnr * nnx * nnz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < N * nslots; i += (size_t)gridDim.x * blockDim.x) {
        size_t s = i / N, c = i % N;
        int ir = (int)(c % nnr);
        size_t col = c / nnr;
        int ix = (int)(col % nnx), iz = (int)(col / nnx);
        float v = __builtin_huge_valf();
        if (ir == nnr - 1) {
            double f = faces[s * (size_t)nnx * nnz + (size_t)ix * nnz + iz];
            if (f > 0.0) v = (float)(f - toff[s]);
        }
        T[i] = v;
    }
}

// fixed-node bits: bottom-face nodes holding a positive boundary time
__global__ void k_fixbits_tele(int nnr, int nnx, int nnz, int nslots, const double* faces, uint32_t* fix)
{
    const size_t N = (size_t)nnr * nnx * nnz;
    const size_t NW = (N + 31) / 32;
    for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < NW * nslots; w += (size_t)gridDim.x * blockDim.x) {
        size_t s = w / NW, w0 = w % NW;
        uint32_t bits = 0;
        for (int b = 0; b < 32; ++b) {
            size_t c = w0 * 32 + b;
            if (c >= N) break;
            int ir = (int)(c % nnr);
            if (ir != nnr - 1) continue;
            size_t col = c / nnr;
            int ix = (int)(col % nnx), iz = (int)(col / nnx);
            if (faces[s * (size_t)nnx * nnz + (size_t)ix * nnz + iz] > 0.0) bits |= 1u << b;
        }
        fix[w] = bits;
    }
}

// Phase-start tile stamps: -1 everywhere (phase A: tiles holding fixed
// nodes 0), or 0 everywhere (phases B/C).
__global__ void k_stamp_init(int nslots, int ntiles, int ntr, int all, int* stamp)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nslots * ntiles; i += (size_t)gridDim.x * blockDim.x) {
        int t = (int)(i % ntiles);
        stamp[i] = all ? 0 : (((t % ntr) == ntr - 1) ? 0 : -1);
    }
}

__global__ void k_fill_int(int* p, int n, int v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

template <typename R>
__global__ void k_copy(const R* __restrict__ a, R* __restrict__ b, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) b[i] = a[i];
}

template <typename R, bool SHIFT>
static hipError_t launch_sweep(fmm3d_ctx* c, const SweepArgs<R>& a, int blocks)
{
    const int tr = c->tile_r, tx = c->tile_x, tz = c->tile_z;
#define FMM3D_TILE_CASE(A, B, C)                                                                 \
    if (tr == A && tx == B && tz == C) {                                                         \
        k_sweep<R, A, B, C, SHIFT><<<blocks, A * B * C, 0, c->stream>>>(a);                       \
        return hipGetLastError();                                                                \
    }
    FMM3D_TILE_CASE(16, 4, 4)
    FMM3D_TILE_CASE(8, 8, 4)
    FMM3D_TILE_CASE(8, 8, 8)
    FMM3D_TILE_CASE(4, 4, 4)
    FMM3D_TILE_CASE(8, 4, 4)
    FMM3D_TILE_CASE(32, 4, 2)
#undef FMM3D_TILE_CASE
    return hipErrorInvalidValue;
}

template <typename R, bool SHIFT>
static int run_phases(fmm3d_ctx* c, int nsrc, int order, int* sweeps_out)
{
    Slots& s = c->sl;
    const Grid& g = c->g;
    const int ntr = (g.nnr + c->tile_r - 1) / c->tile_r;
    const int ntx = (g.nnx + c->tile_x - 1) / c->tile_x;
    const int ntz = (g.nnz + c->tile_z - 1) / c->tile_z;
    const int ntiles = ntr * ntx * ntz;
    SweepArgs<R> a;
    a.T = (R*)s.T;
    a.P = (const R*)s.P;
    a.slo = std::is_same<R, double>::value ? (const R*)c->slo64 : (const R*)c->slo32;
    a.hx = std::is_same<R, double>::value ? (const R*)c->hx64 : (const R*)c->hx32;
    a.hz = std::is_same<R, double>::value ? (const R*)c->hz64 : (const R*)c->hz32;
    a.hr = (R)g.dnr;
    a.fixbits = s.fix;
    a.stamp = s.stamp;
    a.lastchg = s.lastchg;
    a.slot_on = s.on;
    a.nnr = g.nnr; a.nnx = g.nnx; a.nnz = g.nnz;
    a.ntr = ntr; a.ntx = ntx; a.ntz = ntz;
    a.nslots = nsrc;
    a.fom = order == FMM3D_ORDER_MIXED;
    a.updates = c->d_updates;
    const int blocks = nsrc * ntiles;
    const int maxit = 40 * (g.nnr + g.nnx + g.nnz) + 1000;
    const int CHECK = 16;
    std::vector<int> lc(nsrc);
    int total_sweeps = 0;
    int nphases = order == FMM3D_ORDER_MIXED ? 5 : 1;   // A, B, C (+ up to 2 repeats of C)
    for (int ph = 0; ph < nphases; ++ph) {
        a.phase = ph == 0 ? 0 : 1;
        if (ph >= 1) {
            // freeze the ordering: P <- T
            size_t n = (size_t)nsrc * c->N;
            k_copy<R><<<4096, 256, 0, c->stream>>>((const R*)s.T, (R*)s.P, n);
            HIPCHK(c, hipGetLastError());
        }
        k_stamp_init<<<std::min(4096, (blocks + 255) / 256), 256, 0, c->stream>>>(nsrc, ntiles, ntr, ph >= 1, s.stamp);
        k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.lastchg, nsrc, 0);
        k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.on, nsrc, 1);
        HIPCHK(c, hipGetLastError());
        int k = 1;
        int last_k_with_change = 0;
        bool done = false;
        while (!done) {
            for (int j = 0; j < CHECK; ++j, ++k) {
                a.iter = k;
                HIPCHK(c, (launch_sweep<R, SHIFT>(c, a, blocks)));
                c->stats.kernel_launches++;
            }
            HIPCHK(c, hipMemcpyAsync(lc.data(), s.lastchg, nsrc * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            int mx = 0;
            for (int q = 0; q < nsrc; ++q) mx = std::max(mx, lc[q]);
            last_k_with_change = mx;
            if (mx < k - 1) done = true;          // the last iteration changed nothing
            if (!done && k > maxit)
                return fail(c, FMM3D_E_NOCONV, "solver: phase %d did not converge in %d sweeps", ph, maxit);
        }
        total_sweeps += last_k_with_change + 1;
        if (ph >= 2 && last_k_with_change == 0) break;   // C changed nothing: the ordering is consistent
    }
    *sweeps_out = total_sweeps;
    return FMM3D_OK;
}

extern "C" int fmm3d_solve_tele(fmm3d_ctx* c, int nsrc, const double* faces, int order, int precision, int flags)
{
    if (!c) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: null context");
    if (nsrc <= 0 || !faces) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: need nsrc > 0 and faces");
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: bad order %d", order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: bad precision %d", precision);
    if (flags != 0) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: flags must be 0");
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "fmm3d_solve_tele: no velocity model (call fmm3d_set_model_vtx)");
    HIPCHK(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    const size_t F = (size_t)g.nnx * g.nnz;
    // boundary minimum per source: the float32 offset (and a sanity check)
    std::vector<double> toff(nsrc, 0.0);
    for (int s = 0; s < nsrc; ++s) {
        double m = __builtin_huge_val();
        for (size_t i = 0; i < F; ++i) {
            double v = faces[(size_t)s * F + i];
            if (isnan(v)) return fail(c, FMM3D_E_ARG, "fmm3d_solve_tele: NaN boundary time (source %d)", s);
            if (v > 0.0 && v < m) m = v;
        }
        toff[s] = isfinite(m) ? m : 0.0;
    }
    int rc = ensure_slots(c, nsrc, precision);
    if (rc) return rc;
    Slots& s = c->sl;
    s.n = 0;
    double* dfaces = nullptr;
    HIPCHK(c, dalloc(&dfaces, (size_t)nsrc * F * sizeof(double)));
    HIPCHK(c, hipMemcpyAsync(dfaces, faces, (size_t)nsrc * F * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(s.toff, toff.data(), nsrc * sizeof(double), hipMemcpyHostToDevice, c->stream));
    if (precision == FMM3D_PREC_FP64)
        k_init_tele64<<<4096, 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, (double*)s.T, s.fix, s.toff);
    else
        k_init_tele32<<<4096, 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, (float*)s.T, s.toff);
    k_fixbits_tele<<<4096, 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, s.fix);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) { (void)hipFree(dfaces); return fail(c, FMM3D_E_DEVICE, "init: %s", hipGetErrorString(le)); }
    c->stats = fmm3d_stats{};
    c->stats.tile_r = c->tile_r; c->stats.tile_x = c->tile_x; c->stats.tile_z = c->tile_z;
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int sweeps = 0;
    if (precision == FMM3D_PREC_FP64) rc = run_phases<double, false>(c, nsrc, order, &sweeps);
    else rc = run_phases<float, true>(c, nsrc, order, &sweeps);
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    (void)hipFree(dfaces);
    if (rc) return rc;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->stats.solve_ms = ms;
    c->stats.sweeps = sweeps;
    s.n = nsrc;
    return FMM3D_OK;
}

extern "C" int fmm3d_get_field(fmm3d_ctx* c, int src, double* ttn)
{
    if (!c || !ttn) return fail(c, FMM3D_E_ARG, "fmm3d_get_field: null argument");
    if (src < 0 || src >= c->sl.n) return fail(c, FMM3D_E_STATE, "fmm3d_get_field: source %d not resident", src);
    HIPCHK(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    const size_t N = c->N;
    std::vector<double> d(N);
    double to = 0.0;
    if (c->sl.prec == FMM3D_PREC_FP64) {
        HIPCHK(c, hipMemcpy(d.data(), (double*)c->sl.T + (size_t)src * N, N * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(N);
        HIPCHK(c, hipMemcpy(f.data(), (float*)c->sl.T + (size_t)src * N, N * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&to, c->sl.toff + src, sizeof(double), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < N; ++i) d[i] = to + (double)f[i];
    }
    // device r-fastest -> host phi-fastest; unknown nodes come back as 0 (the reference's marker)
    for (int iz = 0; iz < g.nnz; ++iz)
        for (int ix = 0; ix < g.nnx; ++ix)
            for (int ir = 0; ir < g.nnr; ++ir) {
                double v = d[((size_t)iz * g.nnx + ix) * g.nnr + ir];
                ttn[((size_t)ir * g.nnx + ix) * g.nnz + iz] = isfinite(v) ? v : 0.0;
            }
    return FMM3D_OK;
}

extern "C" int fmm3d_receivers(fmm3d_ctx* c, int nrc, const double* rcv, float* t_out)
{
    if (!c || (nrc > 0 && (!rcv || !t_out)) || nrc < 0) return fail(c, FMM3D_E_ARG, "fmm3d_receivers: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_receivers: no resident solution");
    if (nrc == 0) return FMM3D_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const int ns = c->sl.n;
    double* drcv = nullptr;
    float* dout = nullptr;
    HIPCHK(c, dalloc(&drcv, (size_t)nrc * 3 * sizeof(double)));
    HIPCHK(c, dalloc(&dout, (size_t)ns * nrc * sizeof(float)));
    HIPCHK(c, hipMemcpyAsync(drcv, rcv, (size_t)nrc * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int total = ns * nrc;
    if (c->sl.prec == FMM3D_PREC_FP64)
        k_receivers<double><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const double*)c->sl.T, c->sl.toff, ns, nrc, drcv, dout);
    else
        k_receivers<float><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const float*)c->sl.T, c->sl.toff, ns, nrc, drcv, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(t_out, dout, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    (void)hipFree(drcv); (void)hipFree(dout);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "receivers: %s", hipGetErrorString(e));
    return FMM3D_OK;
}

extern "C" int fmm3d_frechet(fmm3d_ctx* c, int nrc, const double* rcv, int64_t* row_ptr, int32_t* cols,
                             float* vals, int64_t cap)
{
    if (!c || nrc < 0 || (nrc > 0 && (!rcv || !row_ptr)) || cap < 0) return fail(c, FMM3D_E_ARG, "fmm3d_frechet: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_frechet: no resident solution");
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "fmm3d_frechet: no velocity model");
    HIPCHK(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    const int ns = c->sl.n;
    const int nrays = ns * nrc;
    row_ptr[0] = 0;
    if (nrays == 0) return FMM3D_OK;
    const size_t NV = (size_t)(g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2);
    // ray workers: one wavefront each, with a dense float32 scratch of NV
    const size_t workers = std::min<size_t>(4096, (size_t)nrays);
    if (c->ray_scratch_workers < workers) {
        (void)hipFree(c->ray_scratch);
        c->ray_scratch = nullptr;
        c->ray_scratch_workers = 0;
        HIPCHK(c, dalloc(&c->ray_scratch, workers * NV * sizeof(float)));
        HIPCHK(c, hipMemsetAsync(c->ray_scratch, 0, workers * NV * sizeof(float), c->stream));
        c->ray_scratch_workers = workers;
    }
    // rays go in chunks so the staging (NV entries per ray) stays bounded
    const int64_t rcap = (int64_t)NV;
    const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(nrays, (int64_t)(256ll << 20) / (rcap * 8)));
    double* drcv = nullptr;
    int32_t* dcols = nullptr;
    float* dvals = nullptr;
    int32_t* dcount = nullptr;
    HIPCHK(c, dalloc(&drcv, (size_t)nrc * 3 * sizeof(double)));
    HIPCHK(c, dalloc(&dcols, (size_t)chunk * rcap * sizeof(int32_t)));
    HIPCHK(c, dalloc(&dvals, (size_t)chunk * rcap * sizeof(float)));
    HIPCHK(c, dalloc(&dcount, (size_t)chunk * sizeof(int32_t)));
    hipError_t e = hipMemcpyAsync(drcv, rcv, (size_t)nrc * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    std::vector<int32_t> cnt(chunk);
    std::vector<int32_t> hc;
    std::vector<float> hv;
    int64_t total = 0;
    bool short_cap = false;
    for (int r0 = 0; r0 < nrays && e == hipSuccess; r0 += chunk) {
        int nr = std::min(chunk, nrays - r0);
        RayOut o{ dcols, dvals, dcount, rcap };
        int grid = (int)std::min<size_t>(workers, (size_t)nr);
        if (c->sl.prec == FMM3D_PREC_FP64)
            k_rays<double><<<grid, 64, 0, c->stream>>>(g, (const double*)c->sl.T, c->sl.toff, c->veln, ns, nrc, drcv, r0, nr, c->ray_scratch, o);
        else
            k_rays<float><<<grid, 64, 0, c->stream>>>(g, (const float*)c->sl.T, c->sl.toff, c->veln, ns, nrc, drcv, r0, nr, c->ray_scratch, o);
        e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(cnt.data(), dcount, nr * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) break;
        int64_t need = 0;
        for (int q = 0; q < nr; ++q) {
            if (cnt[q] == -2) { e = hipErrorUnknown; break; }
            need = std::max<int64_t>(need, cnt[q]);
        }
        if (e != hipSuccess) break;
        hc.resize(need); hv.resize(need);
        for (int q = 0; q < nr; ++q) {
            int64_t n = std::max(cnt[q], 0);
            row_ptr[r0 + q + 1] = row_ptr[r0 + q] + n;
            if (n == 0) continue;
            if (total + n > cap) { short_cap = true; total += n; continue; }
            e = hipMemcpy(cols + total, dcols + (int64_t)q * rcap, n * sizeof(int32_t), hipMemcpyDeviceToHost);
            if (e == hipSuccess) e = hipMemcpy(vals + total, dvals + (int64_t)q * rcap, n * sizeof(float), hipMemcpyDeviceToHost);
            if (e != hipSuccess) break;
            total += n;
        }
    }
    (void)hipFree(drcv); (void)hipFree(dcols); (void)hipFree(dvals); (void)hipFree(dcount);
    if (e == hipErrorUnknown) return fail(c, FMM3D_E_DEVICE, "frechet: row staging overflow");
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "frechet: %s", hipGetErrorString(e));
    if (short_cap) return fail(c, FMM3D_E_CAPACITY, "fmm3d_frechet: capacity %lld < %lld entries", (long long)cap, (long long)row_ptr[nrays]);
    return FMM3D_OK;
}

extern "C" int fmm3d_get_stats(const fmm3d_ctx* c, fmm3d_stats* out)
{
    if (!c || !out) return FMM3D_E_ARG;
    *out = c->stats;
    return FMM3D_OK;
}
