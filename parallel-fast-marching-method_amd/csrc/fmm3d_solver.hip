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
// Layouts on the device: fields and slowness are phi-fastest, one radial
// layer after another, idx = (ir*nnx + ix)*nnz + iz (0-based) -- the
// reference's ttn(iz,ix,ir) order -- so a layer is contiguous and lanes map
// to phi; the node velocity used by rays keeps the reference gridder's halo
// (-1..n+1 per axis) in the same order.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <mutex>
#include <string>
#include <type_traits>
#include <thread>
#include <vector>

#include "../../include/fmm3d.h"
#include "fmm3d_device.h"
#include "fmm3d_engine.h"

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
                          double* __restrict__ slo64, float* __restrict__ slo32, int onx, int onz)
{
    const int MR = g.nnr + 3, MX = g.nnx + 3, MZ = g.nnz + 3;
    const long total = (long)MR * MX * MZ;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        int kz = (int)(i % MZ) - 1;
        int kx = (int)((i / MZ) % MX) - 1;
        int kr = (int)(i / ((long)MZ * MX)) - 1;
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
            // slowness rows of onx x onz nodes: g's own, or a larger container's
            size_t s = ((size_t)(kr - 1) * onx + (kx - 1)) * onz + (kz - 1);
            double sl = 1.0 / acc_z;
            slo64[s] = sl;
            slo32[s] = (float)sl;
        }
    }
}

// Mean slowness of each layer (one block per layer): the vertical profile
// of the float32 teleseismic reference.
__global__ void __launch_bounds__(256) k_layer_mean(const double* __restrict__ slo, size_t NL, double* __restrict__ out)
{
    __shared__ double sm[256];
    const double* p = slo + (size_t)blockIdx.x * NL;
    double acc = 0.0;
    for (size_t i = threadIdx.x; i < NL; i += 256) acc += p[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sm[threadIdx.x] += sm[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0] / (double)NL;
}

// ------------------------------------------------------------ solver sweeps

template <typename R> struct SweepArgs {
    R* T;                   // fields, slot s at T + s*N
    const R* P;             // frozen ordering fields (phases B/C)
    const R* slo;           // slowness, N
    const R* hx;            // per ir: ri*dnx
    const R* hz;            // per (ir, ix): ri*sin(theta)*dnz, index ir*nnx + ix
    R hr;                   // dnr
    const uint32_t* fixbits;// per slot: N bits, fixed (Dirichlet) nodes
    int* stamp;             // per slot per brick: last sweep with a change
    int* lastchg;           // per slot: last sweep with a change
    const int* slot_on;     // per slot: 1 while iterating
    int nnr, nnx, nnz;
    int nbr, nbx, nbz;      // bricks per axis
    int sl;                 // layers per brick
    int kin;                // lateral relaxations per layer
    int nslots;
    int iter;               // sweep number k (>= 1)
    int dir;                // +1: layers bottom to top, -1: top to bottom
    int phase;              // 0: A (first order, value-based); 1: B/C (frozen ordering)
    int fom;                // second order allowed in phase 1
    unsigned long long* visits;  // active bricks counter
    double tol;             // relative change below which a brick counts as settled
    int* valchg;            // per slot, 1 if a sweep changed a value by more than tol
    int dirmode;            // 0: sweep direction alternates; 1: per brick, along increasing time
    unsigned long long* updates;  // node updates performed
};

// One node update from gathered stencil values.  Sides are ordered
// (r-, r+, x-, x+, z-, z+); in1/in2 say whether the first/second neighbour
// exists.  Phase 0 counts a neighbour iff its value is below the node's
// (float32: not above, so a value rounded onto its neighbour keeps its
// support); phase 1 counts by the frozen ordering field and goes second
// order iff the second neighbour is ordered before the first.
template <typename R, bool SHIFT>
__device__ __forceinline__ R node_update(int phase, int fom, R tc, R pc, const R v1[6], const R v2[6],
                                         const R q1[6], const R q2[6], const bool in1[6], const bool in2[6],
                                         R hr, R hx, R hz, R slo)
{
    const R inf = Lim<R>::inf();
    Side<R> sd[3][2];
    R tb = inf;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        Side<R>& S = sd[q / 2][q % 2];
        S.ord = 0; S.t1 = R(0); S.t2 = R(0);
        if (!in1[q]) { S.ord = -1; continue; }
        if (phase == 0) {
            const bool cnt = SHIFT ? (v1[q] <= tc && v1[q] < inf) : (v1[q] < tc);
            if (!cnt) continue;
            S.ord = 1; S.t1 = v1[q];
        } else {
            if (!(q1[q] < pc) || !(v1[q] < inf)) continue;
            S.ord = 1; S.t1 = v1[q];
            if (fom && in2[q] && q2[q] < inf && v2[q] < inf && q1[q] > q2[q]) { S.ord = 2; S.t2 = v2[q]; }
        }
        if (SHIFT && S.t1 < tb) tb = S.t1;
    }
    if (SHIFT) {
        if (!(tb < inf)) return inf;
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) { sd[a][b].t1 -= tb; sd[a][b].t2 -= tb; }
    }
    R best = octant_min<R>(sd, hr, hx, hz, slo);
    if (SHIFT && best < inf) best = tb + best;
    return best;
}

// Brick sweep.  A block owns a CX x CZ column patch of one brick of `sl`
// layers and walks the layers in the sweep direction; each layer is staged
// in LDS with a two-node lateral halo and relaxed `kin` times, the radial
// neighbours come from HBM (the layer just finished by this block is
// visible after the barrier).  Updates are in place; a brick runs in sweep
// k iff it or a face neighbour changed in sweep k-1.
template <typename R, int CX, int CZ, bool SHIFT>
__global__ void __launch_bounds__(CX * CZ)
k_brick(SweepArgs<R> a)
{
    constexpr int LX = CX + 4, LZ = CZ + 4, NT = CX * CZ;
    __shared__ R cur[LX * LZ];
    __shared__ R pcu[LX * LZ];
    __shared__ int s_act, s_chg;
    const int nb = a.nbr * a.nbx * a.nbz;
    const int slot = blockIdx.x / nb;
    const int b = blockIdx.x - slot * nb;
    if (slot >= a.nslots || !a.slot_on[slot]) return;
    const int br = b / (a.nbx * a.nbz), bx = (b / a.nbz) % a.nbx, bz = b % a.nbz;
    int* st = a.stamp + (size_t)slot * nb;
    if (threadIdx.x == 0) {
        int m = st[b];
        if (br > 0) m = max(m, st[b - a.nbx * a.nbz]);
        if (br < a.nbr - 1) m = max(m, st[b + a.nbx * a.nbz]);
        if (bx > 0) m = max(m, st[b - a.nbz]);
        if (bx < a.nbx - 1) m = max(m, st[b + a.nbz]);
        if (bz > 0) m = max(m, st[b - 1]);
        if (bz < a.nbz - 1) m = max(m, st[b + 1]);
        s_act = (m >= a.iter - 1);
        s_chg = 0;
    }
    __syncthreads();
    if (!s_act) return;
    if (threadIdx.x == 0) atomicAdd(a.visits, 1ull);
    const size_t NL = (size_t)a.nnx * a.nnz;
    const size_t N = NL * a.nnr;
    R* T = a.T + (size_t)slot * N;
    const R* P = a.phase ? a.P + (size_t)slot * N : nullptr;
    const uint32_t* fb = a.fixbits + (size_t)slot * ((N + 31) / 32);
    const R inf = Lim<R>::inf();
    const int lx = threadIdx.x / CZ, lz = threadIdx.x % CZ;
    const int ix = bx * CX + lx, iz = bz * CZ + lz;
    const bool valid = ix < a.nnx && iz < a.nnz;
    const int own = (lx + 2) * LZ + (lz + 2);
    const int r_lo = br * a.sl, r_hi = min(r_lo + a.sl, a.nnr) - 1;
    const int nl = r_hi - r_lo + 1;
    // Sweep direction of this brick: along increasing time, judged on the
    // patch's middle column (the ordering field in phases B/C, the current
    // field in phase A); the sweep's alternating default where unknown.
    // Relaxing second-order updates against the dependency direction only
    // feeds them stale inputs, which they amplify.
    __shared__ int s_dir;
    if (threadIdx.x == 0) {
        const int mx = min(bx * CX + CX / 2, a.nnx - 1), mz = min(bz * CZ + CZ / 2, a.nnz - 1);
        const R* D = a.phase ? P : T;
        const R top = D[(size_t)r_lo * NL + (size_t)mx * a.nnz + mz];
        const R bot = D[(size_t)r_hi * NL + (size_t)mx * a.nnz + mz];
        int d = a.dir;
        if (a.dirmode && top < inf && bot < inf && top != bot) d = top > bot ? 1 : -1;
        s_dir = d;
    }
    __syncthreads();
    const int dir = s_dir;
    // radial side behind the sweep (already relaxed in this sweep) and ahead of it
    const int qb = dir > 0 ? 1 : 0, qa = 1 - qb;
    const long sb = dir > 0 ? (long)NL : -(long)NL;      // offset of the node behind
    // the two nodes behind the first layer belong to the neighbouring brick
    R beh1 = inf, beh2 = inf;
    {
        const int ir0 = dir > 0 ? r_hi : r_lo;
        const int i1 = ir0 + (dir > 0 ? 1 : -1), i2 = ir0 + (dir > 0 ? 2 : -2);
        if (valid) {
            if (i1 >= 0 && i1 < a.nnr) beh1 = T[(size_t)i1 * NL + (size_t)ix * a.nnz + iz];
            if (i2 >= 0 && i2 < a.nnr) beh2 = T[(size_t)i2 * NL + (size_t)ix * a.nnz + iz];
        }
    }
    bool any = false;
    unsigned nupd = 0;
    for (int qq = 0; qq < nl; ++qq) {
        const int ir = dir > 0 ? r_hi - qq : r_lo + qq;
        const size_t lay = (size_t)ir * NL;
        for (int e = threadIdx.x; e < LX * LZ; e += NT) {
            const int hx_ = bx * CX + e / LZ - 2, hz_ = bz * CZ + e % LZ - 2;
            const bool in = hx_ >= 0 && hx_ < a.nnx && hz_ >= 0 && hz_ < a.nnz;
            const size_t gi = lay + (size_t)hx_ * a.nnz + hz_;
            cur[e] = in ? T[gi] : inf;
            if (a.phase) pcu[e] = in ? P[gi] : inf;
        }
        __syncthreads();
        const size_t c = lay + (size_t)ix * a.nnz + iz;
        const bool upd = valid && !((fb[c >> 5] >> (c & 31)) & 1u);
        R v1[6], v2[6], q1[6], q2[6];
        bool in1[6], in2[6];
        R slo = R(0), hx = R(0), hz = R(0), pc = R(0);
        if (upd) {
            slo = a.slo[c];
            hx = a.hx[ir];
            hz = a.hz[(size_t)ir * a.nnx + ix];
            pc = a.phase ? pcu[own] : R(0);
            in1[0] = ir - 1 >= 0; in2[0] = ir - 2 >= 0;
            in1[1] = ir + 1 < a.nnr; in2[1] = ir + 2 < a.nnr;
            // behind: this thread's own column, relaxed in the previous steps
            v1[qb] = in1[qb] ? beh1 : inf;
            v2[qb] = in2[qb] ? beh2 : inf;
            // ahead: not yet visited in this sweep
            v1[qa] = in1[qa] ? T[c - sb] : inf;
            v2[qa] = in2[qa] ? T[c - 2 * sb] : inf;
            if (a.phase) {
                q1[0] = in1[0] ? P[c - NL] : inf; q2[0] = in2[0] ? P[c - 2 * NL] : inf;
                q1[1] = in1[1] ? P[c + NL] : inf; q2[1] = in2[1] ? P[c + 2 * NL] : inf;
            }
            in1[2] = ix - 1 >= 0; in2[2] = ix - 2 >= 0;
            in1[3] = ix + 1 < a.nnx; in2[3] = ix + 2 < a.nnx;
            in1[4] = iz - 1 >= 0; in2[4] = iz - 2 >= 0;
            in1[5] = iz + 1 < a.nnz; in2[5] = iz + 2 < a.nnz;
        }
        for (int k = 0; k < a.kin; ++k) {
            R nv = R(0), tc = R(0);
            bool moved = false;
            if (upd) {
                tc = cur[own];
                v1[2] = cur[own - LZ]; v2[2] = cur[own - 2 * LZ];
                v1[3] = cur[own + LZ]; v2[3] = cur[own + 2 * LZ];
                v1[4] = cur[own - 1]; v2[4] = cur[own - 2];
                v1[5] = cur[own + 1]; v2[5] = cur[own + 2];
                if (a.phase) {
                    q1[2] = pcu[own - LZ]; q2[2] = pcu[own - 2 * LZ];
                    q1[3] = pcu[own + LZ]; q2[3] = pcu[own + 2 * LZ];
                    q1[4] = pcu[own - 1]; q2[4] = pcu[own - 2];
                    q1[5] = pcu[own + 1]; q2[5] = pcu[own + 2];
                }
                nv = node_update<R, SHIFT>(a.phase, a.fom, tc, pc, v1, v2, q1, q2, in1, in2, a.hr, hx, hz, slo);
                nupd++;
            }
            __syncthreads();
            if (upd && nv != tc) {
                cur[own] = nv;
                T[c] = nv;
                // a change of at most `tol` (relative to the larger value) does
                // not wake the neighbourhood; a node gaining or losing a value does
                const bool fin_n = nv < inf, fin_t = tc < inf;
                if (fin_n != fin_t || !(fabs((double)nv - (double)tc) <= a.tol * fmax(fabs((double)nv), fabs((double)tc)))) {
                    any = true;
                    moved = true;
                }
            }
            // relax the layer again only while something in the patch still moves
            if (!__syncthreads_or(moved)) break;
        }
        __syncthreads();
        if (valid) { beh2 = beh1; beh1 = cur[own]; }
        __syncthreads();
    }
    if (any) s_chg = 1;
    // node updates of this block: wave-level sum, one atomic per wave
    {
        unsigned long long w = nupd;
        for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off, 64);
        if ((threadIdx.x & 63) == 0 && w) atomicAdd(a.updates, w);
    }
    __syncthreads();
    if (threadIdx.x == 0 && s_chg) {
        st[b] = a.iter;
        atomicMax(a.lastchg + slot, a.iter);
        atomicOr(a.valchg + slot, 1);
    }
}

// Column walk (single phase).  A one-wavefront block owns a PX x PZ patch of
// every layer and walks the column from the bottom layer to the top.  Each
// layer is staged in LDS (values with a two-node lateral halo, the two
// layers above and the two just finished below, slowness with fixed nodes
// marked negative) and relaxed by four ordered line passes: +phi, -phi,
// +theta, -theta.  In a phi pass every lane owns an S-node segment of one
// theta row (S = PX*PZ/64) and steps along it; in a theta pass a lane owns
// an S-node segment of one phi column.  The lanes step together, so the
// nodes of one step update from the values the previous steps left.  The
// layer goes back to HBM once, after its passes.
//
// Node updates use the value-based mixed-order predicate on the node's own
// value (a neighbour counts iff its value is below the node's -- float32:
// not above; second order iff the second neighbour is known and below the
// first) and drop a value no neighbour supports.  Processed in this
// near-causal order the relaxation converges to the heap march's field
// itself; the lateral halo is read from HBM when the layer is staged
// (neighbouring patches walk the same layers concurrently), so a patch
// column runs again in the next sweep while it or a lateral neighbour
// changed.
// Flow entries at or above FLOW_CENTER (alternating walks: point sources)
// hold the (theta, phi) node index of the slot's earliest fixed node plus
// FLOW_CENTER; the column walk then orders that slot's patches outward from
// the patch holding it.  Smaller entries are lateral flow codes (+-1, +-2).
constexpr int FLOW_CENTER = 1 << 24;

template <typename R> struct ColArgs {
    R* T;
    const R* slo;
    const R* hx;
    const R* hz;
    R hr;
    const uint32_t* fixbits;
    int* stamp;             // per slot per patch column
    int* lastchg;
    int* valchg;
    const int* slot_on;
    int nnr, nnx, nnz;
    int npx, npz;           // patches per axis
    int nslots;
    int iter;
    int fom;
    int maxpass;            // line passes per layer at most (cycle +phi, -phi, +theta, -theta)
    double tol;
    unsigned long long* visits;
    unsigned long long* updates;
    unsigned long long* nodes;   // free nodes staged
    int* prog;              // per slot per patch: iter*pstride + layers (steps) finished, or null (no ordering)
    int pstride;            // progress counts per sweep (both column walks: nnr + 16)
    int ordered_until;      // sweeps up to this one wait on their upstream patches (0: all)
    const int* flow;        // per slot: lateral flow signs (fx, fz), or null
    int spin_limit;         // polls before a wait on a neighbour is given up
    int down;               // 1: walk top to bottom (when walkalt is null)
    const int* iterp;       // sweep number from device memory (graph replays), or null: iter
    const int* walkalt;     // per slot: 1 = alternate the walk (down on even sweeps), or null
    int px, pz;             // patch shape (nodes)
    int verify;             // check the fixed point after this sweep (0: never)
    int verify_patches;     // the check also narrows the next sweep to the moving patches
    int* pend;              // per slot per patch: a node would move (fixed-point check)
    int* slotpend;          // per slot: some node would move
    const int* born;        // per slot: its first sweep, checked after it (the serving engine), or null
    size_t geo_n, geo_hx, geo_hz;   // per-slot strides of slo, hx, hz (slots on grids of their own), or 0 (shared)
    const int* sdims;       // per slot (nnr, nnx, nnz) of its own grid inside the nnr x nnx x nnz
                            // container (strides stay the container's), or null: the container's
    unsigned long long* lrng;   // alternating walks: per slot per patch, two entries (by sweep parity) of
                                // (sweep << 32 | first changed layer << 16 | last changed layer), or null
    const double* ref;      // float32 fields: per slot nnr + 2 reference values (A[ir], gx, gz), or null
    double tfloor;          // slots with a reference: a change counts above tol * max(|value|, tfloor)
    float qscale;           // slots with a reference: updated values are rounded to multiples of 1/qscale (0: not)
    float ptol;             // slots with a reference: a layer's passes stop after one that moves no node by
                            // more (s), and the fixed-point check counts only such moves (0: exact)
    int* wlost;             // per slot: a wait on a neighbour was given up in a sweep
    int lguess;             // slots with a reference: a node not yet reached starts its layer's passes
                            // from the polynomial through the lguess nodes walked just before it in its
                            // column (1: the node before; up to 4: the cubic), 0: from infinity
    int vfilter;            // the check after a slot's first sweep evaluates only the nodes an input
                            // relaxed after them could move (below)
};

// Reference value of node (ir, ix, iz) from a slot's reference row
// (A[0..nnr-1], gx, gz); 0 without one.
__device__ __forceinline__ double ref_at(const double* __restrict__ rs, int nnr, size_t ir, size_t ix, size_t iz)
{
    return rs ? rs[ir] + rs[nnr] * (double)ix + rs[nnr + 1] * (double)iz : 0.0;
}

// An updated value stored against a reference, rounded to the nearest
// multiple of 1/qs (qs a power of two; 0: as computed).  The reference's
// steps are multiples of 2^-20 s, so with qs >= 2^20 a stored value plus a
// step is exact while under 2^24/qs s in magnitude and the update's
// neighbour comparisons agree from both sides; the fixed grid also bounds
// how fine a change the relaxation resolves.  inf stays inf.
template <typename R> __device__ __forceinline__ R ref_round(R v, R qs)
{
    return qs != R(0) ? rint(v * qs) * (R(1) / qs) : v;
}

// A value every lane of a wave reads at the same address, from memory no
// kernel writes while it runs: read through the constant address space, so
// the load is a scalar one and the value takes no vector registers.
template <typename P> __device__ __forceinline__ P uniform_load(const P* p, size_t i)
{
    return ((const __attribute__((address_space(4))) P*)p)[i];
}

// Steps of a slot's reference from node (ir, ix, iz) to its stencil
// neighbours: a float32 neighbour value plus its step is the neighbour's
// time relative to the centre node's reference.  Every step is a multiple
// of 2^-20 s small enough to be exact in float32.  float64 fields carry no
// reference and take no steps.
template <typename R> struct RefSteps {
    R up1 = R(0), up2 = R(0), dn1 = R(0), dn2 = R(0);   // layers ir-1, ir-2, ir+1, ir+2
    R gx = R(0), gz = R(0);                             // one node along theta, phi
    __device__ __forceinline__ void lateral(const double* rs, int nnr)
    {
        if (std::is_same<R, float>::value && rs) { gx = (R)rs[nnr]; gz = (R)rs[nnr + 1]; }
    }
    // layer ir in two steps, so that other loads can be issued between
    // them: fetch issues five independent loads at indices inside the row
    // (no branch between them: one memory round trip), settle forms the steps
    double f0 = 0.0, fm1 = 0.0, fm2 = 0.0, fp1 = 0.0, fp2 = 0.0;
    __device__ __forceinline__ void fetch(const double* rs, int ir, int snr)
    {
        if (std::is_same<R, float>::value && rs) {
            f0 = uniform_load(rs, ir); fm1 = uniform_load(rs, max(ir - 1, 0)); fm2 = uniform_load(rs, max(ir - 2, 0));
            fp1 = uniform_load(rs, min(ir + 1, snr - 1)); fp2 = uniform_load(rs, min(ir + 2, snr - 1));
        }
    }
    __device__ __forceinline__ void settle(const double* rs, int ir, int snr)
    {
        if (std::is_same<R, float>::value && rs) {
            up1 = ir >= 1 ? (R)(fm1 - f0) : R(0);
            up2 = ir >= 2 ? (R)(fm2 - f0) : R(0);
            dn1 = ir + 1 < snr ? (R)(fp1 - f0) : R(0);
            dn2 = ir + 2 < snr ? (R)(fp2 - f0) : R(0);
        }
    }
    __device__ __forceinline__ void layer(const double* rs, int ir, int snr)
    {
        fetch(rs, ir, snr);
        settle(rs, ir, snr);
    }
    // v1/v2 in side order r-, r+, x-, x+, z-, z+
    __device__ __forceinline__ void apply(R v1[6], R v2[6]) const
    {
        if (std::is_same<R, float>::value) {
            v1[0] += up1; v2[0] += up2; v1[1] += dn1; v2[1] += dn2;
            v1[2] -= gx; v2[2] -= gx + gx; v1[3] += gx; v2[3] += gx + gx;
            v1[4] -= gz; v2[4] -= gz + gz; v1[5] += gz; v2[5] += gz + gz;
        }
    }
};

// A layer-range entry: the sweep it belongs to and the layers lo..hi that changed.
__host__ __device__ __forceinline__ unsigned long long lrng_pack(int sweep, int lo, int hi)
{
    return ((unsigned long long)(unsigned)sweep << 32) | ((unsigned long long)(lo & 0xffff) << 16) | (unsigned)(hi & 0xffff);
}

// float32 helpers for the axis-sum octant form: fused multiply-subtract,
// hardware square root and reciprocal (about 1 ulp each).
template <typename R> __device__ __forceinline__ R fused_mul_sub(R a, R b, R c) { return fma(a, b, -c); }
template <typename R> __device__ __forceinline__ R fast_sqrt(R x) { return sqrt(x); }
template <> __device__ __forceinline__ float fast_sqrt<float>(float x) { return __builtin_amdgcn_sqrtf(x); }
template <typename R> __device__ __forceinline__ R fast_rcp(R x) { return R(1) / x; }
template <> __device__ __forceinline__ float fast_rcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Octant minimum in the axis-sum form (FAST): for the counted sides of an
// octant sum (c*T - D)^2 / H^2 = s^2 with c, D, H = (1, t1, h) first order
// and (3, 4 t1 - t2, 2 h) second order, i.e. A T^2 - 2 B T + C = 0 with
// A = sum (c/H)^2, B = sum (c/H)(D/H), C = sum (D/H)^2 - s^2; the larger
// root with the discriminant clamped at zero.  This is the reference's
// octant quadratic after the substitution T = (X + D_L)/c_L on its leading
// axis (fm3dt_parallel.f90:535-1004), so the two agree up to rounding.
// Octants through a missing side are skipped, octants with no counted side
// give nothing, as in the reference: the first carry A = NaN, the second
// A = B = 0 and so T = 0 * inf = NaN, and the running minimum is fmin,
// which passes over a NaN.  Values arrive shifted by the smallest counted
// neighbour.
// One node update from the gathered stencil (sides r-, r+, x-, x+, z-, z+).
// FAST selects the axis-sum octant form; otherwise the reference's
// leading-axis form in the CPU oracle's operation order.
template <typename R, bool FAST, bool SHIFT>
__device__ __forceinline__ R col_update(int fom, R tc, const R v1[6], const R v2[6], const bool in1[6],
                                        const bool in2[6], const R h[3], const R ih[3], R slown)
{
    const R inf = Lim<R>::inf();
    bool cnt[6], sec[6];
    R tb = inf;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        cnt[q] = in1[q] && (SHIFT ? (v1[q] <= tc && v1[q] < inf) : (v1[q] < tc));
        sec[q] = cnt[q] && fom && in2[q] && v2[q] < inf && v1[q] > v2[q];
        if (SHIFT && cnt[q] && v1[q] < tb) tb = v1[q];
    }
    if (SHIFT && !(tb < inf)) return inf;
    const R sh = SHIFT ? tb : R(0);
    R best;
    if (FAST) {
        // per side (a, b, c) = (alpha^2, alpha*beta, beta^2); a missing side
        // carries a = NaN so every octant through it fails the A > 0 test
        const R nan = __builtin_nan("");
        R ta[6], tbb[6], tc2[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            // selects on plain values keep this branch-free
            const R t1 = v1[q] - sh, t2 = v2[q] - sh;
            const R num = sec[q] ? fma(R(-0.5), t2, t1 + t1) : t1;
            const R al = ih[q / 2] * (cnt[q] ? (sec[q] ? R(1.5) : R(1)) : R(0));
            const R be = cnt[q] ? num * ih[q / 2] : R(0);
            ta[q] = in1[q] ? al * al : nan;
            tbb[q] = al * be;
            tc2[q] = be * be;
        }
        const R s2 = slown * slown;
        tc2[0] -= s2; tc2[1] -= s2;
        best = inf;
#pragma unroll
        for (int j = 2; j < 4; ++j)
#pragma unroll
            for (int k = 4; k < 6; ++k) {
                const R Axz = ta[j] + ta[k], Bxz = tbb[j] + tbb[k], Cxz = tc2[j] + tc2[k];
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const R A = ta[i] + Axz, B = tbb[i] + Bxz, C = tc2[i] + Cxz;
                    R disc = fused_mul_sub(B, B, A * C);
                    disc = disc > R(0) ? disc : R(0);
                    best = fmin(best, (B + fast_sqrt(disc)) * fast_rcp(A));
                }
            }
    } else {
        Side<R> sd[3][2];
#pragma unroll
        for (int q = 0; q < 6; ++q) {
            Side<R>& S = sd[q / 2][q % 2];
            S.ord = !in1[q] ? -1 : (sec[q] ? 2 : (cnt[q] ? 1 : 0));
            S.t1 = cnt[q] ? v1[q] - sh : R(0);
            S.t2 = sec[q] ? v2[q] - sh : R(0);
        }
        best = octant_min<R>(sd, h[0], h[1], h[2], slown);
    }
    if (SHIFT && best < inf) best = tb + best;
    return best;
}

// Patch index of launch rank r on an axis of n patches ordered outward from
// patch s: s, s+1, s-1, s+2, s-2, ..., then the rest of the longer side.
// A patch's neighbour towards s always has the lower rank.
__device__ __forceinline__ int centre_rank(int r, int s, int n)
{
    const int lo = s, hi = n - 1 - s, m = min(lo, hi);
    if (r == 0) return s;
    if (r <= 2 * m) {
        const int d = (r + 1) >> 1;
        return (r & 1) ? s + d : s - d;
    }
    const int d = r - m;
    return hi > lo ? s + d : s - d;
}

template <typename R, int PX, int PZ, bool FAST, bool SHIFT>
__global__ void __launch_bounds__(64)
k_column(ColArgs<R> a)
{
    constexpr int NP = PX * PZ, S = NP / 64;
    static_assert(NP % 64 == 0 && PX % S == 0 && PZ % S == 0, "segments must tile the patch");
    static_assert(PX <= 64, "one lane per theta row of the patch stages its spacing");
    constexpr int LX = PX + 4, LZ = (PZ + 4) | 1;   // odd row stride: phi passes spread over banks
    __shared__ R cur[LX * LZ];
    __shared__ R b1[NP], b2[NP], u1[NP], u2[NP], sl[NP];
    __shared__ R b3[NP], b4[NP];   // the third and fourth layers walked before (layer start only)
    __shared__ R shz[PX], sihz[PX];
    __shared__ int s_act;
    __shared__ int s_win[2];
    const int np = a.npx * a.npz;
    // Block -> patch.  A row of patches (one slot, one theta index, every phi
    // index) runs on one XCD -- workgroups go round-robin over the 8 XCDs --
    // so patches that share cache lines share an L2; rows go round-robin
    // over the XCDs.  Inside a slot, rows and the patches of a row follow
    // the slot's lateral flow (fx, fz), so a patch's upstream neighbours have
    // lower block indices.
    const int xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int row = (k / a.npz) * 8 + xcd;
    if (row >= a.nslots * a.npx) return;
    const int slot = row / a.npx;
    if (!a.slot_on[slot]) return;
    // sweep number: an argument, or a device counter (graph replays)
    const int iter = a.iterp ? __hip_atomic_load(a.iterp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : a.iter;
    const int f0 = a.flow ? a.flow[2 * slot] : 1, f1 = a.flow ? a.flow[2 * slot + 1] : 1;
    const int rx = row - slot * a.npx, rz = k % a.npz;
    // fx, fz: the direction a patch's wave arrives from along each axis
    // (0: none, the patch is upstream of everything on that axis)
    int bx, bz, fx, fz;
    const bool centred = f0 >= FLOW_CENTER;
    if (centred) {
        // outward from the patch holding the slot's earliest fixed node
        const int sx = min((f0 - FLOW_CENTER) / a.px, a.npx - 1), sz = min((f1 - FLOW_CENTER) / a.pz, a.npz - 1);
        bx = centre_rank(rx, sx, a.npx);
        bz = centre_rank(rz, sz, a.npz);
        fx = bx > sx ? 1 : (bx < sx ? -1 : 0);
        fz = bz > sz ? 1 : (bz < sz ? -1 : 0);
    } else {
        fx = f0 < 0 ? -1 : 1;
        fz = f1 < 0 ? -1 : 1;
        bx = fx > 0 ? rx : a.npx - 1 - rx;
        bz = fz > 0 ? rz : a.npz - 1 - rz;
    }
    const int b = bx * a.npz + bz;
    int* st = a.stamp + (size_t)slot * np;
    const R* slo = a.slo + slot * a.geo_n;
    const R* hxs = a.hx + slot * a.geo_hx;
    const R* hzs = a.hz + slot * a.geo_hz;
    const bool alt = a.walkalt && a.walkalt[slot];
    const bool down = a.walkalt ? (alt && !(iter & 1)) : (a.down != 0);
    // alternating walks wait on their upstream patches only when ordered outward from their source
    int* pg = (a.prog && (!alt || centred) && (a.ordered_until == 0 || iter <= a.ordered_until)) ? a.prog + (size_t)slot * np : nullptr;
    const int pbase = iter * a.pstride;
    const int t = threadIdx.x;
    // this slot's own grid: nodes beyond it read as outside the grid
    const int snr = a.sdims ? a.sdims[3 * slot] : a.nnr;
    const int snx = a.sdims ? a.sdims[3 * slot + 1] : a.nnx;
    const int snz = a.sdims ? a.sdims[3 * slot + 2] : a.nnz;
    if (t == 0) {
        int m = st[b];
        if (bx > 0) m = max(m, st[b - a.npz]);
        if (bx < a.npx - 1) m = max(m, st[b + a.npz]);
        if (bz > 0) m = max(m, st[b - 1]);
        if (bz < a.npz - 1) m = max(m, st[b + 1]);
        s_act = (m >= iter - 1) && bx * PX < snx && bz * PZ < snz;
        // an idle patch is finished with every layer of this sweep
        if (!s_act && pg) __hip_atomic_store(pg + b, pbase + a.nnr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        // layer window (alternating walks): the layers within two of one
        // that changed, in this patch or a lateral neighbour, in the sweep
        // before or so far in this one.  Every node outside it was relaxed
        // after its inputs last changed; a change inside the walk carries
        // it on past the window (below).
        int lo = 0, hi = snr - 1;
        if (s_act && alt && a.lrng) {
            const unsigned long long* lr = a.lrng + (size_t)slot * np * 2;
            int l = 0x7fffffff, h = -1;
            const int nb[5] = { b, bx > 0 ? b - a.npz : -1, bx < a.npx - 1 ? b + a.npz : -1,
                                bz > 0 ? b - 1 : -1, bz < a.npz - 1 ? b + 1 : -1 };
            for (int j = 0; j < 5; ++j) {
                if (nb[j] < 0) continue;
                for (int par = 0; par < 2; ++par) {
                    const unsigned long long v = __hip_atomic_load(lr + 2 * nb[j] + par, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    const int tag = (int)(unsigned)(v >> 32);
                    if (tag == iter || tag == iter - 1) {
                        l = min(l, (int)((v >> 16) & 0xffff));
                        h = max(h, (int)(v & 0xffff));
                    }
                }
            }
            if (h < 0) {   // nothing changed near it: no layer to walk
                lo = snr;
                hi = snr - 1;
            } else {
                lo = max(0, l - 2);
                hi = min(snr - 1, h + 2);
            }
        }
        s_win[0] = lo;
        s_win[1] = hi;
    }
    __syncthreads();
    if (!s_act) return;
    if (t == 0) atomicAdd(a.visits, 1ull);
    // upstream neighbours along the flow, -1 where none
    const int ux = (fx != 0 && bx - fx >= 0 && bx - fx < a.npx) ? b - fx * a.npz : -1;
    const int uz = (fz != 0 && bz - fz >= 0 && bz - fz < a.npz) ? b - fz : -1;
    bool wx = pg && ux >= 0, wz = pg && uz >= 0;   // still waiting on them (given up after a timeout)
    int seen_x = 0, seen_z = 0;                     // their progress as last read (lane 0)
    const size_t NL = (size_t)a.nnx * a.nnz;
    const size_t N = NL * a.nnr;
    R* T = a.T + (size_t)slot * N;
    const uint32_t* fb = a.fixbits + (size_t)slot * ((N + 31) / 32);
    const R inf = Lim<R>::inf();
    const int x0 = bx * PX, z0 = bz * PZ;
    const R hr = a.hr, ihr = R(1) / a.hr;
    // the lateral halo ring (two nodes deep round the patch): this lane's
    // entries as an LDS index and an offset within a layer (-1: outside the
    // grid or beyond the ring).  Ring index i: two rows above the patch, two
    // below, then the two columns each side of its PX rows.
    constexpr int HW = PZ + 4, RING = 4 * (PX + PZ) + 16, NR = (RING + 63) / 64;
    auto ring = [&](int k, int& idx, int& off) {
        const int i = t + 64 * k;
        int hx_ = 0, hz_ = 0;
        if (i < 2 * HW) { hx_ = i / HW; hz_ = i % HW; }
        else if (i < 4 * HW) { hx_ = PX + 2 + (i - 2 * HW) / HW; hz_ = (i - 2 * HW) % HW; }
        else { const int j = i - 4 * HW, m = j & 3; hx_ = 2 + (j >> 2); hz_ = m < 2 ? m : PZ + m; }
        const int gx = x0 + hx_ - 2, gz = z0 + hz_ - 2;
        idx = i < RING ? hx_ * LZ + hz_ : -1;
        off = (i < RING && gx >= 0 && gx < snx && gz >= 0 && gz < snz) ? gx * a.nnz + gz : -1;
    };
    const double* rs = a.ref ? a.ref + (size_t)slot * (a.nnr + 2) : nullptr;
    RefSteps<R> rst;
    rst.lateral(rs, a.nnr);
    // a change counts above tol times the larger magnitude; with a reference
    // the stored values sit near zero and carry rounding at the scale of the
    // reference's steps, so the magnitude is taken as at least tfloor
    const bool hasref = rs && rs[0] != 0.0;
    const double tfloor = hasref ? a.tfloor : 0.0;
    const R rq = hasref ? (R)a.qscale : R(0);
    const R ptl = hasref ? (R)a.ptol : R(0);
    const bool lguess = hasref && a.lguess;
    bool anyv = false;
    bool capped = false;   // a layer's passes hit the cap with nodes still moving
    unsigned long long nupd = 0, nvis = 0;   // wave-uniform counts (scalar registers)
    R orig[S];
    // walk: bottom to top, or top to bottom when a.down; "prev" layers are the
    // two this block relaxed just before (b1, b2), "next" ones come from HBM (u1, u2)
    const int dn = down ? 1 : -1;
    // walk steps: from the window's first layer in walk order, on past its
    // last while a layer within the two before changed
    const bool win = alt && a.lrng;
    const int qs = down ? s_win[0] : snr - 1 - s_win[1];
    const int qw = down ? s_win[1] : snr - 1 - s_win[0];
    int lastq = -1000, chg_lo = 0x7fffffff, chg_hi = -1;
    if (pg && t == 0 && qs > 0) __hip_atomic_store(pg + b, pbase + qs, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int q = qs; q < snr; ++q) {
        if (q > qw && q - lastq > 2) break;
        const int ir = down ? q : snr - 1 - q;
        const size_t lay = (size_t)ir * NL;
        unsigned nfree = 0;
        // wait (bounded) until the upstream patches have finished this layer;
        // progress read earlier that already covers it saves the poll (an
        // upstream patch that runs ahead covers many layers with one read)
        if (wx || wz) {
            if (t == 0) {
                const int need = pbase + q + 1;
                int spins = 0;
                while ((wx && seen_x < need) || (wz && seen_z < need)) {
                    // both neighbours' counters in flight together
                    const int vx = (wx && seen_x < need) ? __hip_atomic_load(pg + ux, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : seen_x;
                    const int vz = (wz && seen_z < need) ? __hip_atomic_load(pg + uz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : seen_z;
                    seen_x = vx; seen_z = vz;
                    if ((!wx || seen_x >= need) && (!wz || seen_z >= need)) break;
                    if (++spins > a.spin_limit) {
                        if (seen_x < need) wx = false;
                        if (seen_z < need) wz = false;
                        if (a.wlost) __hip_atomic_store(a.wlost + slot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(2);
                }
                s_act = (wx ? 1 : 0) | (wz ? 2 : 0);
            }
            __syncthreads();
            wx = s_act & 1; wz = s_act & 2;
        }
        // stage.  Every load of the layer is issued before the first LDS
        // write needs one, so the stage costs one memory round trip: the halo
        // ring, then the interior (own values kept for the write-back), the
        // two layers ahead, fixed bits and slowness (read whether or not the
        // node is fixed, at an index inside the layer).  Neighbours' values
        // are device-coherent loads: another XCD may have written them this
        // sweep.
        R hv[NR];
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            int idx, off;
            ring(k, idx, off);
            hv[k] = off >= 0 ? __hip_atomic_load(T + lay + off, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : inf;
        }
        const R hzv = t < PX ? hzs[(size_t)ir * a.nnx + min(x0 + t, snx - 1)] : R(1);
        const R hxl = uniform_load(hxs, ir);
        rst.fetch(rs, ir, snr);
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int e = t + 64 * k;
            const int lx = e / PZ, lz = e % PZ;
            const int gx = x0 + lx, gz = z0 + lz;
            const bool in = gx < snx && gz < snz;
            const size_t c = lay + (size_t)gx * a.nnz + gz;
            const R v = in ? T[c] : inf;
            const int n1 = ir + dn, n2 = ir + 2 * dn;
            const R w1 = (in && n1 >= 0 && n1 < snr) ? T[c + (ptrdiff_t)dn * (ptrdiff_t)NL] : inf;
            const R w2 = (in && n2 >= 0 && n2 < snr) ? T[c + (ptrdiff_t)(2 * dn) * (ptrdiff_t)NL] : inf;
            const uint32_t fw = in ? fb[c >> 5] : 1u;
            const R sv = slo[in ? c : lay];
            // the walk's first layer: the two before it come from HBM (none at the edge)
            if (q == qs) {
                b1[e] = (in && q >= 1) ? T[c - (ptrdiff_t)dn * (ptrdiff_t)NL] : inf;
                b2[e] = (in && q >= 2) ? T[c - (ptrdiff_t)(2 * dn) * (ptrdiff_t)NL] : inf;
                b3[e] = (in && q >= 3) ? T[c - (ptrdiff_t)(3 * dn) * (ptrdiff_t)NL] : inf;
                b4[e] = (in && q >= 4) ? T[c - (ptrdiff_t)(4 * dn) * (ptrdiff_t)NL] : inf;
            }
            const bool fixed = !in || ((fw >> (c & 31)) & 1u);
            orig[k] = v;
            cur[(lx + 2) * LZ + lz + 2] = v;
            u1[e] = w1;
            u2[e] = w2;
            sl[e] = fixed ? R(-1) : sv;
            nfree += (unsigned)__popcll(__ballot(!fixed));
        }
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            int idx, off;
            ring(k, idx, off);
            if (idx >= 0) cur[idx] = hv[k];
        }
        // layer start: a free node not reached yet begins from the stored
        // values of the nodes walked before it in its column (this lane
        // staged all of them), extrapolated by the polynomial through up to
        // a.lguess of them (fewer near the walk's first layers)
        if (lguess) {
#pragma unroll
            for (int k = 0; k < S; ++k) {
                const int e = t + 64 * k;
                const R p1 = b1[e];
                if (sl[e] >= R(0) && !(orig[k] < inf) && p1 < inf) {
                    const R p2 = b2[e], p3 = b3[e], p4 = b4[e];
                    const int m = a.lguess < 2 || !(p2 < inf) ? 1 : a.lguess < 3 || !(p3 < inf) ? 2 : a.lguess < 4 || !(p4 < inf) ? 3 : 4;
                    cur[(e / PZ + 2) * LZ + e % PZ + 2] = m == 4 ? R(4) * (p1 - p2) - R(2) * (p2 - p3) + (p3 - p4) + p3
                                                        : m == 3 ? R(3) * (p1 - p2) + p3
                                                        : m == 2 ? p1 + (p1 - p2) : p1;
                }
            }
        }
        if (t < PX) {
            shz[t] = hzv;
            sihz[t] = R(1) / hzv;
        }
        const R ihx = R(1) / hxl;
        rst.settle(rs, ir, snr);
        __syncthreads();
        // the passes stop after one that changes nothing: the layer is then
        // at a fixpoint of the update and later passes would not move it.
        // Slots with a reference stop after one that moves no node by more
        // than ptl; the fixed-point check after their first sweep counts
        // only such moves too, or the residuals left below ptl would send
        // nearly every patch into the next sweep.
        int npass = 0;
#pragma unroll 1
        for (int pass = 0; pass < a.maxpass; ++pass) {
            const bool zpass = (pass & 3) < 2, rev = pass & 1;
            bool moved = false;   // a node moved by more than the pass tolerance
            const int lane_l = zpass ? t / (PZ / S) : t % PZ;          // line (row or column) index
            const int seg0 = (zpass ? t % (PZ / S) : t / PZ) * S;     // segment start along the line
#pragma unroll 1
            for (int s = 0; s < S; ++s) {
                const int p = seg0 + (rev ? S - 1 - s : s);
                const int lx = zpass ? lane_l : p, lz = zpass ? p : lane_l;
                const int pl = lx * PZ + lz;
                const R slown = sl[pl];
                if (slown >= R(0)) {
                    const int ix = x0 + lx, iz = z0 + lz;
                    const int own = (lx + 2) * LZ + lz + 2;
                    const R tc = cur[own];
                    R v1[6], v2[6];
                    bool in1[6], in2[6];
                    in1[0] = ir >= 1; in2[0] = ir >= 2;
                    in1[1] = ir + 1 < snr; in2[1] = ir + 2 < snr;
                    in1[2] = ix >= 1; in2[2] = ix >= 2;
                    in1[3] = ix + 1 < snx; in2[3] = ix + 2 < snx;
                    in1[4] = iz >= 1; in2[4] = iz >= 2;
                    in1[5] = iz + 1 < snz; in2[5] = iz + 2 < snz;
                    // side 0 is the layer above (ir-1), side 1 the one below (ir+1)
                    v1[0] = down ? b1[pl] : u1[pl]; v2[0] = down ? b2[pl] : u2[pl];
                    v1[1] = down ? u1[pl] : b1[pl]; v2[1] = down ? u2[pl] : b2[pl];
                    v1[2] = cur[own - LZ]; v2[2] = cur[own - 2 * LZ];
                    v1[3] = cur[own + LZ]; v2[3] = cur[own + 2 * LZ];
                    v1[4] = cur[own - 1]; v2[4] = cur[own - 2];
                    v1[5] = cur[own + 1]; v2[5] = cur[own + 2];
                    rst.apply(v1, v2);
                    const R h[3] = { hr, hxl, shz[lx] };
                    const R ih[3] = { ihr, ihx, sihz[lx] };
                    const R nv = ref_round(col_update<R, FAST, SHIFT>(a.fom, tc, v1, v2, in1, in2, h, ih, slown), rq);
                    // every lane has read its stencil before any lane writes (one wavefront)
                    __builtin_amdgcn_wave_barrier();
                    if (nv != tc) {
                        cur[own] = nv;
                        if (!(fabs(nv - tc) <= ptl)) moved = true;
                    }
                }
                __builtin_amdgcn_wave_barrier();
            }
            ++npass;
            if (!__ballot(moved)) break;
            if (pass == a.maxpass - 1) capped = true;
        }
        nupd += nfree * (unsigned)npass;
        nvis += nfree;
        __syncthreads();
        // write back what changed; the layer becomes the one below
        bool lay_chg = false;
#pragma unroll
        for (int k = 0; k < S; ++k) {
            const int e = t + 64 * k;
            const int lx = e / PZ, lz = e % PZ;
            const R nv = cur[(lx + 2) * LZ + lz + 2];
            const R tc = orig[k];
            if (nv != tc) {
                __hip_atomic_store(T + lay + (size_t)(x0 + lx) * a.nnz + z0 + lz, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                const bool fin_n = nv < inf, fin_t = tc < inf;
                if (fin_n != fin_t ||
                    !(fabs((double)nv - (double)tc) <= a.tol * fmax(fmax(fabs((double)nv), fabs((double)tc)), tfloor)))
                    lay_chg = true;
            }
            b4[e] = b3[e];
            b3[e] = b2[e];
            b2[e] = b1[e];
            b1[e] = nv;
        }
        anyv |= lay_chg;
        if (win && __ballot(lay_chg)) {
            lastq = q;
            chg_lo = min(chg_lo, ir);
            chg_hi = max(chg_hi, ir);
        }
        // publish the layer once its stores have completed
        __syncthreads();
        if (pg && t == 0) {
            __builtin_amdgcn_s_waitcnt(0);
            __hip_atomic_store(pg + b, pbase + q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    // layers the walk did not reach are finished too
    if (pg && t == 0) __hip_atomic_store(pg + b, pbase + a.nnr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // a layer left short of its fixpoint: the fixed-point check counts the
    // patch as moving without evaluating it (its filter assumes every layer
    // reached the fixpoint of the inputs it had)
    if (capped && t == 0 && a.pend) {
        a.pend[(size_t)slot * np + b] = 1;
        a.slotpend[slot] = 1;
    }
    if (win && t == 0 && chg_hi >= 0)
        __hip_atomic_store(a.lrng + ((size_t)slot * np + b) * 2 + (iter & 1), lrng_pack(iter, chg_lo, chg_hi),
                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    {
        if (t == 0 && nupd) atomicAdd(a.updates, nupd);
        if (t == 0 && nvis) atomicAdd(a.nodes, nvis);
    }
    const bool blk = __syncthreads_or(anyv);
    if (t == 0 && blk) {
        st[b] = iter;
        atomicMax(a.lastchg + slot, iter);
        atomicOr(a.valchg + slot, 1);
    }
}

// Fixed-point check after sweep a.verify of the column walk.  One thread
// per free node evaluates the node update on the field as it stands -- the
// walk's own col_update on the same stencil, in the same arithmetic -- and
// a node whose value would move marks its patch and its slot.  Nothing is
// written to the field.  At a fixed point no node moves whatever the order,
// so a slot with no marked node is where any further sweep would leave it.
template <typename R, bool FAST, bool SHIFT>
__global__ void __launch_bounds__(256) k_verify_mark(ColArgs<R> a)
{
    const int iter = a.iterp ? __hip_atomic_load(a.iterp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : a.iter;
    if (!a.born && iter != a.verify) return;
    // one thread per (slot, theta, phi) column walks the layers top down with
    // the radial window (two above, the node, two below) in registers, so
    // each value comes from HBM about once; lateral neighbours are loads of
    // the same layer that adjacent threads share
    const size_t NL = (size_t)a.nnx * a.nnz, N = NL * a.nnr, NW = (N + 31) / 32;
    const int np = a.npx * a.npz;
    const R inf = Lim<R>::inf();
    const R hr = a.hr, ihr = R(1) / a.hr;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)a.nslots * NL;
         i += (size_t)gridDim.x * blockDim.x) {
        const int slot = (int)(i / NL);
        const size_t f = i - (size_t)slot * NL;
        // alternating walks (point sources) are never at their fixed point
        // after one sweep: they are not checked
        if (!a.slot_on[slot] || (a.born && a.born[slot] != iter) || (a.walkalt && a.walkalt[slot])) continue;
        const int ix = (int)(f / a.nnz), iz = (int)(f - (size_t)ix * a.nnz);
        // this slot's own grid inside the container
        const int snr = a.sdims ? a.sdims[3 * slot] : a.nnr;
        const int snx = a.sdims ? a.sdims[3 * slot + 1] : a.nnx;
        const int snz = a.sdims ? a.sdims[3 * slot + 2] : a.nnz;
        if (ix >= snx || iz >= snz) continue;
        const R* T = a.T + (size_t)slot * N;
        const uint32_t* fb = a.fixbits + (size_t)slot * NW;
        const double* rs = a.ref ? a.ref + (size_t)slot * (a.nnr + 2) : nullptr;
        RefSteps<R> rst;
        rst.lateral(rs, a.nnr);
        const R qs = (rs && rs[0] != 0.0) ? (R)a.qscale : R(0);
        const R vtl = (rs && rs[0] != 0.0) ? (R)a.ptol : R(0);
        // After a slot's first sweep every patch walked every layer, each to
        // the fixpoint of the inputs it had then (a layer that hit the pass
        // cap marked its patch in the walk).  A node can then move only
        // through an input relaxed after it -- the two layers walked after
        // its own, and lateral neighbours outside its patch, except those in
        // the patch it waited for on that axis -- and only if that input is
        // not above it (the update counts no neighbour above its result).
        // Nodes with no such input are not evaluated.  late_x/late_z: per
        // offset -1, -2, +1, +2, whether that input can be one.
        const bool vf = a.vfilter && (a.born || iter == 1);
        const bool down = a.walkalt ? false : (a.down != 0);
        bool late_x[4] = { false, false, false, false }, late_z[4] = { false, false, false, false };
        if (vf) {
            int fx = 0, fz = 0;   // the waited-for patch on each axis is at patch index - f (0: none)
            const int pxo = ix / a.px, pzo = iz / a.pz;
            if (a.prog && (a.ordered_until == 0 || iter <= a.ordered_until) && !a.wlost[slot]) {
                const int f0 = a.flow ? a.flow[2 * slot] : 1, f1 = a.flow ? a.flow[2 * slot + 1] : 1;
                if (f0 >= FLOW_CENTER) {
                    const int sx = min((f0 - FLOW_CENTER) / a.px, a.npx - 1), sz = min((f1 - FLOW_CENTER) / a.pz, a.npz - 1);
                    fx = pxo > sx ? 1 : (pxo < sx ? -1 : 0);
                    fz = pzo > sz ? 1 : (pzo < sz ? -1 : 0);
                } else {
                    fx = f0 < 0 ? -1 : 1;
                    fz = f1 < 0 ? -1 : 1;
                }
            }
            const int off[4] = { -1, -2, 1, 2 };
            for (int k = 0; k < 4; ++k) {
                const int jx = ix + off[k], jz = iz + off[k];
                late_x[k] = jx >= 0 && jx / a.px != pxo && !(fx != 0 && jx / a.px == pxo - fx);
                late_z[k] = jz >= 0 && jz / a.pz != pzo && !(fz != 0 && jz / a.pz == pzo - fz);
            }
        }
        bool moved = false;
        // window w[0..4] = layers ir-2 .. ir+2 of this column
        R w[5];
        w[0] = inf; w[1] = inf;
        w[2] = T[f];
        w[3] = snr > 1 ? T[NL + f] : inf;
        w[4] = snr > 2 ? T[2 * NL + f] : inf;
        for (int ir = 0; ir < snr; ++ir) {
            const size_t c = (size_t)ir * NL + f;
            if (!((fb[c >> 5] >> (c & 31)) & 1u)) {
                R v1[6], v2[6];
                bool in1[6], in2[6];
                in1[0] = ir >= 1; in2[0] = ir >= 2;
                in1[1] = ir + 1 < snr; in2[1] = ir + 2 < snr;
                in1[2] = ix >= 1; in2[2] = ix >= 2;
                in1[3] = ix + 1 < snx; in2[3] = ix + 2 < snx;
                in1[4] = iz >= 1; in2[4] = iz >= 2;
                in1[5] = iz + 1 < snz; in2[5] = iz + 2 < snz;
                // side 0 the layer above (ir-1), side 1 the one below (ir+1);
                // out of the grid reads as inf, as the walk's staging does
                v1[0] = w[1]; v2[0] = w[0];
                v1[1] = w[3]; v2[1] = w[4];
                v1[2] = in1[2] ? T[c - a.nnz] : inf; v2[2] = in2[2] ? T[c - 2 * (size_t)a.nnz] : inf;
                v1[3] = in1[3] ? T[c + a.nnz] : inf; v2[3] = in2[3] ? T[c + 2 * (size_t)a.nnz] : inf;
                v1[4] = in1[4] ? T[c - 1] : inf; v2[4] = in2[4] ? T[c - 2] : inf;
                v1[5] = in1[5] ? T[c + 1] : inf; v2[5] = in2[5] ? T[c + 2] : inf;
                rst.layer(rs, ir, snr);
                rst.apply(v1, v2);
                // the lowest input relaxed after the node, in its frame
                R lm = down ? fmin(v1[1], v2[1]) : fmin(v1[0], v2[0]);
                if (late_x[0]) lm = fmin(lm, v1[2]);
                if (late_x[1]) lm = fmin(lm, v2[2]);
                if (late_x[2]) lm = fmin(lm, v1[3]);
                if (late_x[3]) lm = fmin(lm, v2[3]);
                if (late_z[0]) lm = fmin(lm, v1[4]);
                if (late_z[1]) lm = fmin(lm, v2[4]);
                if (late_z[2]) lm = fmin(lm, v1[5]);
                if (late_z[3]) lm = fmin(lm, v2[5]);
                if (!vf || lm <= w[2]) {
                    const R hxl = a.hx[slot * a.geo_hx + ir], hzl = a.hz[slot * a.geo_hz + (size_t)ir * a.nnx + ix];
                    const R h[3] = { hr, hxl, hzl };
                    const R ih[3] = { ihr, R(1) / hxl, R(1) / hzl };
                    const R nv = ref_round(col_update<R, FAST, SHIFT>(a.fom, w[2], v1, v2, in1, in2, h, ih, a.slo[slot * a.geo_n + c]), qs);
                    if (nv != w[2] && !(fabs(nv - w[2]) <= vtl)) moved = true;
                }
            }
            w[0] = w[1]; w[1] = w[2]; w[2] = w[3]; w[3] = w[4];
            w[4] = ir + 3 < snr ? T[c + 3 * NL] : inf;
        }
        if (moved) {
            a.pend[(size_t)slot * np + (ix / a.px) * a.npz + iz / a.pz] = 1;
            a.slotpend[slot] = 1;
        }
    }
}

// After the check: a slot with no marked node leaves the sweeps; in the
// others a patch with no marked node takes the stamp of the sweep before,
// so the next sweep relaxes the marked patches and their neighbours only
// (a patch runs when it or a neighbour changed in the sweep before).
template <typename R>
__global__ void __launch_bounds__(256) k_verify_apply(ColArgs<R> a, int* on)
{
    const int iter = a.iterp ? __hip_atomic_load(a.iterp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : a.iter;
    if (!a.born && iter != a.verify) return;
    const int np = a.npx * a.npz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)a.nslots * np;
         i += (size_t)gridDim.x * blockDim.x) {
        const int slot = (int)(i / np);
        if (!a.slot_on[slot] || (a.born && a.born[slot] != iter) || (a.walkalt && a.walkalt[slot])) continue;
        if (!a.slotpend[slot]) {
            if (i == (size_t)slot * np) on[slot] = 0;
            continue;
        }
        if (a.verify_patches && !a.pend[i]) a.stamp[i] = min(a.stamp[i], iter - 1);
    }
}

// Lateral walk direction of one axis in a sweep: flow entries are +-1 (that
// direction every sweep) or +-2 (a weak lateral gradient: that direction on
// odd sweeps, the other one on even sweeps).
__device__ __forceinline__ int flow_dir(int f, int iter)
{
    const int s = f > 0 ? 1 : -1;
    return (f == 2 || f == -2) && !(iter & 1) ? -s : s;
}

// Hyperplane column walk.  A one-wavefront block owns an 8 x 8 patch of
// every layer; lane (i, j) holds the column at flow coordinates (i, j) --
// i counts theta rows along the slot's theta flow, j phi columns along its
// phi flow -- and at step s relaxes walk layer w = s - i - j.  The 64 nodes
// of a step are independent, and every node finds its walk-previous,
// theta-upstream and phi-upstream neighbours (one and two nodes away)
// already relaxed in this sweep and the other three from the previous one:
// a sweep is one Gauss-Seidel pass over the slot in (walk layer, theta flow,
// phi flow) order that updates each free node once.  Its fixed point is the
// column walk's (and the heap march's).
//
// Each lane keeps its column's window in registers: the two layers it
// relaxed last (P1, P2), the current one as the previous sweep left it (C0)
// and the next two (N1, N2).  On the skewed schedule the lateral neighbours
// of (i, j, w) sit in the neighbouring lanes' windows -- lane (i-1, j) is one
// layer ahead, so its P1 is (i-1, j, w) relaxed; lane (i+1, j) is one layer
// behind, so its N1 is (i+1, j, w) unrelaxed; two lanes away P2 and N2 -- and
// come over by cross-lane permutes.  Only the patch-boundary halo goes
// through LDS: a ring of 16 steps x 64 cells, cell (combo, position) of a
// step being the halo value some boundary lane reads at that step.  Staging
// goes by chunks of CH steps: the own-column values (the layer three ahead,
// slowness, fixed bits, spacings) are loaded into registers at the start of
// a chunk, and the chunk's halo cells once the upstream patches have
// published the steps that produce them.  Relaxed values collect in an LDS
// ring of 16 layers and go to HBM a layer at a time, coalesced, at the step
// that completes the layer (w + 14), the changed ones only; after each
// chunk the patch publishes the steps it has completed, so a published
// count c means layers below c - 14 are in HBM (the upstream's count s + 15
// feeds this patch's step s).
template <typename R, int CH, bool FAST, bool SHIFT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(1, 5)))
k_plane(ColArgs<R> a)
{
    constexpr int P = 8, HR = 16;
    static_assert(CH >= 1 && CH + 1 <= HR, "the halo ring keeps the previous step's cells");
    constexpr int WR = 16, LAG = 2 * (P - 1);    // layer ring; steps from a layer's first node to its last
    __shared__ R hbuf[HR * 64];
    __shared__ R wbuf[WR * 64];
    __shared__ unsigned char wchg[WR * 64];
    __shared__ int s_act;
    const int np = a.npx * a.npz;
    // Block -> patch by anti-diagonals of the flow order across all slots:
    // diagonal d (patches with rx + rz = d, rx and rz counted along the
    // flow) of every slot, then diagonal d + 1, each padded to W patches.
    // A patch's upstream neighbours lie on the previous diagonal, so the
    // blocks resident together are mostly a few diagonals apart and ready
    // to run, rather than one slot's whole chain waiting on its head.
    const int W = min(a.npx, a.npz), per = a.nslots * W;
    const int dg = blockIdx.x / per, rem = blockIdx.x - dg * per;
    const int slot = rem / W;
    const int rx = max(0, dg - (a.npz - 1)) + (rem - slot * W), rz = dg - rx;
    if (dg >= a.npx + a.npz - 1 || rx >= a.npx || rz < 0) return;
    if (!a.slot_on[slot]) return;
    const int iter = a.iterp ? __hip_atomic_load(a.iterp, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : a.iter;
    const bool alt = a.walkalt && a.walkalt[slot];
    // an alternating slot (waves in every direction) cycles through the
    // eight walk orders: down/up every sweep, theta every two, phi every four
    const int fx = alt ? (((iter >> 1) & 1) ? -1 : 1) : (a.flow ? flow_dir(a.flow[2 * slot], iter) : 1);
    const int fz = alt ? (((iter >> 2) & 1) ? -1 : 1) : (a.flow ? flow_dir(a.flow[2 * slot + 1], iter) : 1);
    const int bx = fx > 0 ? rx : a.npx - 1 - rx;
    const int bz = fz > 0 ? rz : a.npz - 1 - rz;
    const int b = bx * a.npz + bz;
    int* st = a.stamp + (size_t)slot * np;
    const bool down = a.walkalt ? (alt && !(iter & 1)) : (a.down != 0);
    int* pg = (a.prog && !alt) ? a.prog + (size_t)slot * np : nullptr;
    const int nr = a.nnr, nsteps = nr + 2 * (P - 1);
    const int pbase = iter * a.pstride;
    const int t = threadIdx.x;
    if (t == 0) {
        int m = st[b];
        if (bx > 0) m = max(m, st[b - a.npz]);
        if (bx < a.npx - 1) m = max(m, st[b + a.npz]);
        if (bz > 0) m = max(m, st[b - 1]);
        if (bz < a.npz - 1) m = max(m, st[b + 1]);
        s_act = (m >= iter - 1);
        if (!s_act && pg) __hip_atomic_store(pg + b, pbase + nsteps, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if (!s_act) return;
    if (t == 0) atomicAdd(a.visits, 1ull);
    const int ux = (bx - fx >= 0 && bx - fx < a.npx) ? b - fx * a.npz : -1;
    const int uz = (bz - fz >= 0 && bz - fz < a.npz) ? b - fz : -1;
    bool wx = pg && ux >= 0, wz = pg && uz >= 0;
    const size_t NL = (size_t)a.nnx * a.nnz;
    const size_t N = NL * nr;
    R* T = a.T + (size_t)slot * N;
    const uint32_t* fb = a.fixbits + (size_t)slot * ((N + 31) / 32);
    const R inf = Lim<R>::inf();
    const int x0 = bx * P, z0 = bz * P;
    // grid theta / phi index of flow coordinate i / j (halo: -2..P+1)
    auto gxo = [&](int i) { return x0 + (fx > 0 ? i : P - 1 - i); };
    auto gzo = [&](int j) { return z0 + (fz > 0 ? j : P - 1 - j); };
    auto lay = [&](int w) { return (size_t)(down ? w : nr - 1 - w) * NL; };
    const int li = t >> 3, lj = t & 7;
    const int gx = gxo(li), gz = gzo(lj);
    const bool mine = gx < a.nnx && gz < a.nnz;
    const size_t col = mine ? (size_t)gx * a.nnz + gz : 0;
    const int hzrow = min(gx, a.nnx - 1);
    // Per-lane flags, packed in one register and rebuilt at each use (as
    // lane masks they would hold sixteen scalar register pairs through the
    // step loop): bits 0-7 which lateral stencil sides exist (flow-upstream
    // and downstream theta and phi, one and two nodes away), bits 8-15 which
    // lateral neighbours come from the halo ring, bit 16 the column is in
    // the grid.
    auto inr = [](int v, int n) { return v >= 0 && v < n; };
    const unsigned lane_flags =
        (inr(gx - fx, a.nnx) ? 1u : 0u) | (inr(gx - 2 * fx, a.nnx) ? 2u : 0u) |
        (inr(gx + fx, a.nnx) ? 4u : 0u) | (inr(gx + 2 * fx, a.nnx) ? 8u : 0u) |
        (inr(gz - fz, a.nnz) ? 16u : 0u) | (inr(gz - 2 * fz, a.nnz) ? 32u : 0u) |
        (inr(gz + fz, a.nnz) ? 64u : 0u) | (inr(gz + 2 * fz, a.nnz) ? 128u : 0u) |
        (li == 0 ? 0x100u : 0u) | (li <= 1 ? 0x200u : 0u) | (li == P - 1 ? 0x400u : 0u) | (li >= P - 2 ? 0x800u : 0u) |
        (lj == 0 ? 0x1000u : 0u) | (lj <= 1 ? 0x2000u : 0u) | (lj == P - 1 ? 0x4000u : 0u) | (lj >= P - 2 ? 0x8000u : 0u) |
        (gx < a.nnx && gz < a.nnz ? 0x10000u : 0u);
    // the halo cell this lane stages, lane = combo * 8 + position: combos
    // 0/1 theta-upstream depth 1/2, 2/3 phi-upstream, 4/5 theta-downstream,
    // 6/7 phi-downstream.  The cell of layer w is read first at step
    // w + position (upstream) or w + 6 + position (downstream), the ring
    // index of the cell.
    const int hc = t >> 3, hp = t & 7, hside = hc >> 1, hd = hc & 1;
    const int hi = hside == 0 ? -1 - hd : (hside == 2 ? P + hd : hp);
    const int hj = hside == 1 ? -1 - hd : (hside == 3 ? P + hd : hp);
    const int hlag = hside < 2 ? hp : P - 2 + hp;
    const int hgx = gxo(hi), hgz = gzo(hj);
    const bool hin = hgx >= 0 && hgx < a.nnx && hgz >= 0 && hgz < a.nnz;
    const size_t hcol = hin ? (size_t)hgx * a.nnz + hgz : 0;
    // where this lane's boundary neighbours come from: ring cell and step
    // delay (0: this step's cells, 1: the previous step's)
    const int oxm1 = 0 * 8 + lj, oxm2 = (li == 0 ? 1 : 0) * 8 + lj, dxm2 = li == 0 ? 0 : 1;
    const int oxp1 = 4 * 8 + lj, oxp2 = (li == P - 1 ? 5 : 4) * 8 + lj, dxp2 = li == P - 1 ? 1 : 0;
    const int ozm1 = 2 * 8 + li, ozm2 = (lj == 0 ? 3 : 2) * 8 + li, dzm2 = lj == 0 ? 0 : 1;
    const int ozp1 = 6 * 8 + li, ozp2 = (lj == P - 1 ? 7 : 6) * 8 + li, dzp2 = lj == P - 1 ? 1 : 0;
    const R hr = a.hr, ihr = R(1) / a.hr, tolr = (R)a.tol;
    // own-column window (unknown layers outside the grid read as +inf)
    auto old_at = [&](bool in, int w) -> R { return (in && w >= 0 && w < nr) ? T[lay(w) + col] : inf; };
    R P1 = inf, P2 = inf;
    R C0 = old_at(mine, -li - lj), N1 = old_at(mine, 1 - li - lj), N2 = old_at(mine, 2 - li - lj);
    bool anyv = false;
    unsigned nupd = 0;
    const int nchunks = (nsteps + CH - 1) / CH;
#pragma unroll 1
    for (int kc = 0; kc < nchunks; ++kc) {
        const int s0 = kc * CH, s1 = min(s0 + CH, nsteps);
        // own-column values of the chunk's steps: the layer three ahead of the
        // one relaxed, and slowness, fixed-bit word and spacings of that one
        R qn[CH], qs[CH], qhx[CH], qhz[CH];
        uint32_t qf[CH];
        unsigned lfc = lane_flags;
        asm volatile("" : "+v"(lfc));
        const bool minec = lfc & 0x10000u;
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int w = s0 + q - li - lj;
            qn[q] = old_at(minec, w + 3);
            const bool ok = minec && w >= 0 && w < nr;
            const int ir = ok ? (down ? w : nr - 1 - w) : 0;
            const size_t c = ok ? lay(w) + col : 0;
            qs[q] = ok ? a.slo[c] : R(-1);
            qf[q] = ok ? fb[c >> 5] : 0u;
            qhx[q] = a.hx[ir];
            qhz[q] = a.hz[(size_t)ir * a.nnx + hzrow];
        }
        // wait (bounded) until the upstream patches have stored the layers of
        // this chunk's upstream halo (the last read is layer s1 - 1)
        if (wx || wz) {
            if (t == 0) {
                const int need = pbase + min(s1 + LAG, nsteps);
                int spins = 0;
                while (wx && __hip_atomic_load(pg + ux, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    if (++spins > a.spin_limit) { wx = false; break; }
                    __builtin_amdgcn_s_sleep(2);
                }
                spins = 0;
                while (wz && __hip_atomic_load(pg + uz, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    if (++spins > a.spin_limit) { wz = false; break; }
                    __builtin_amdgcn_s_sleep(2);
                }
                s_act = (wx ? 1 : 0) | (wz ? 2 : 0);
            }
            __syncthreads();
            wx = s_act & 1; wz = s_act & 2;
        }
        // halo cells read first in this chunk (device-coherent loads: the
        // neighbouring patches may run on another XCD)
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int w = s0 + q - hlag;
            const R v = (hin && w >= 0 && w < nr)
                ? __hip_atomic_load(T + lay(w) + hcol, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : inf;
            hbuf[((s0 + q) & (HR - 1)) * 64 + t] = v;
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < CH; ++q) {
            const int s = s0 + q;
            if (s >= s1) break;
            const int w = s - li - lj;
            unsigned lf = lane_flags;
            asm volatile("" : "+v"(lf));
            const bool bxm1 = lf & 0x100u, bxm2 = lf & 0x200u, bxp1 = lf & 0x400u, bxp2 = lf & 0x800u;
            const bool bzm1 = lf & 0x1000u, bzm2 = lf & 0x2000u, bzp1 = lf & 0x4000u, bzp2 = lf & 0x8000u;
            // lateral neighbours from the neighbouring lanes' windows ...
            R xm1 = __shfl(P1, (t - 8) & 63, 64), xm2 = __shfl(P2, (t - 16) & 63, 64);
            R xp1 = __shfl(N1, (t + 8) & 63, 64), xp2 = __shfl(N2, (t + 16) & 63, 64);
            R zm1 = __shfl(P1, (t - 1) & 63, 64), zm2 = __shfl(P2, (t - 2) & 63, 64);
            R zp1 = __shfl(N1, (t + 1) & 63, 64), zp2 = __shfl(N2, (t + 2) & 63, 64);
            // ... or, across the patch boundary, from the halo ring (read by
            // every lane, kept by the boundary ones: no branches)
            const int r0 = (s & (HR - 1)) * 64, r1 = ((s - 1) & (HR - 1)) * 64;
            {
                const R h_xm1 = hbuf[r0 + oxm1], h_xm2 = hbuf[(dxm2 ? r1 : r0) + oxm2];
                const R h_xp1 = hbuf[r1 + oxp1], h_xp2 = hbuf[(dxp2 ? r1 : r0) + oxp2];
                const R h_zm1 = hbuf[r0 + ozm1], h_zm2 = hbuf[(dzm2 ? r1 : r0) + ozm2];
                const R h_zp1 = hbuf[r1 + ozp1], h_zp2 = hbuf[(dzp2 ? r1 : r0) + ozp2];
                xm1 = bxm1 ? h_xm1 : xm1; xm2 = bxm2 ? h_xm2 : xm2;
                xp1 = bxp1 ? h_xp1 : xp1; xp2 = bxp2 ? h_xp2 : xp2;
                zm1 = bzm1 ? h_zm1 : zm1; zm2 = bzm2 ? h_zm2 : zm2;
                zp1 = bzp1 ? h_zp1 : zp1; zp2 = bzp2 ? h_zp2 : zp2;
            }
            R cur = C0;
            bool chg = false;
            if (w >= 0 && w < nr && (lf & 0x10000u)) {
                const size_t c = lay(w) + col;
                const bool fixed = (qf[q] >> (c & 31)) & 1u;
                if (!fixed) {
                    const int ir = down ? w : nr - 1 - w;
                    R v1[6], v2[6];
                    bool in1[6], in2[6];
                    // side 0 is layer ir-1, side 1 layer ir+1; the walk reached ir from ir+1 (up) or ir-1 (down)
                    v1[0] = down ? P1 : N1; v2[0] = down ? P2 : N2;
                    v1[1] = down ? N1 : P1; v2[1] = down ? N2 : P2;
                    v1[2] = xm1; v2[2] = xm2; v1[3] = xp1; v2[3] = xp2;
                    v1[4] = zm1; v2[4] = zm2; v1[5] = zp1; v2[5] = zp2;
                    in1[0] = ir >= 1; in2[0] = ir >= 2;
                    in1[1] = ir + 1 < nr; in2[1] = ir + 2 < nr;
                    in1[2] = lf & 1u; in2[2] = lf & 2u; in1[3] = lf & 4u; in2[3] = lf & 8u;
                    in1[4] = lf & 16u; in2[4] = lf & 32u; in1[5] = lf & 64u; in2[5] = lf & 128u;
                    const R h[3] = { hr, qhx[q], qhz[q] };
                    const R ih[3] = { ihr, fast_rcp(qhx[q]), fast_rcp(qhz[q]) };
                    const R nv = col_update<R, FAST, SHIFT>(a.fom, C0, v1, v2, in1, in2, h, ih, qs[q]);
                    ++nupd;
                    if (nv != C0) {
                        cur = nv;
                        chg = true;
                        const bool fin_n = nv < inf, fin_t = C0 < inf;
                        if (fin_n != fin_t || !(fabs(nv - C0) <= tolr * fmax(fabs(nv), fabs(C0))))
                            anyv = true;
                    }
                }
            }
            if (w >= 0 && w < nr) {
                wbuf[(w & (WR - 1)) * 64 + t] = cur;
                wchg[(w & (WR - 1)) * 64 + t] = chg ? 1 : 0;
            }
            // advance the window one layer
            P2 = P1; P1 = cur; C0 = N1; N1 = N2; N2 = qn[q];
            // layer s - 14 is complete (lane (7, 7) relaxed it in this step): store its changes
            const int L = s - LAG;
            if (L >= 0 && L < nr && (lf & 0x10000u) && wchg[(L & (WR - 1)) * 64 + t])
                __hip_atomic_store(T + lay(L) + col, wbuf[(L & (WR - 1)) * 64 + t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        // publish the completed steps once their stores have landed
        if (pg) {
            __builtin_amdgcn_s_waitcnt(0);
            if (t == 0) __hip_atomic_store(pg + b, pbase + s1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    {
        unsigned long long w = nupd;
        for (int off = 32; off > 0; off >>= 1) w += __shfl_down(w, off, 64);
        if (t == 0 && w) { atomicAdd(a.updates, w); atomicAdd(a.nodes, w); }
    }
    const bool blk = __syncthreads_or(anyv);
    if (t == 0 && blk) {
        st[b] = iter;
        atomicMax(a.lastchg + slot, iter);
        atomicOr(a.valchg + slot, 1);
    }
}

// ------------------------------------------------------------ receivers

// Time at node (ir, ix, iz), index i, of a slot's field: float32 fields add
// the slot's offset and reference (rs, or null) to the stored value.
template <typename R>
__device__ __forceinline__ double field_at(const R* T, double toff, const double* rs, int nnr, int ir, int ix, int iz,
                                           size_t i)
{
    return toff + ref_at(rs, nnr, ir, ix, iz) + (double)T[i];
}
template <>
__device__ __forceinline__ double field_at<double>(const double* T, double, const double*, int, int, int, int, size_t i)
{
    return T[i];
}

// trilinear receiver time with a float32 accumulator (fm3dt_parallel.f90:7063-7106);
// NaN outside the grid
template <typename R>
__device__ __forceinline__ float receiver_time(const Grid& g, const R* __restrict__ F, double to, const double* rs,
                                               double rr, double rx, double rz)
{
    int irr = (int)floor((g.rgor - rr) / g.dnr) + 1;
    int irx = (int)floor((rx - g.rgox) / g.dnx) + 1;
    int irz = (int)floor((rz - g.rgoz) / g.dnz) + 1;
    if (irx == g.nnx) irx--;
    if (irz == g.nnz) irz--;
    if (irr < 1 || irr >= g.nnr || irx < 1 || irx >= g.nnx || irz < 1 || irz >= g.nnz) return __builtin_nanf("");
    double drx = (rx - g.rgox) - (irx - 1) * g.dnx;
    double drz = (rz - g.rgoz) - (irz - 1) * g.dnz;
    double drr = (g.rgor - rr) - (irr - 1) * g.dnr;
    float acc = 0.0f;
    for (int j = 0; j < 2; ++j)
        for (int k = 0; k < 2; ++k)
            for (int l = 0; l < 2; ++l) {
                double w = (1.0 - fabs((l * g.dnz - drz) / g.dnz)) * (1.0 - fabs((k * g.dnx - drx) / g.dnx));
                w = w * (1.0 - fabs((j * g.dnr - drr) / g.dnr));
                const int nr_ = irr - 1 + j, nx_ = irx - 1 + k, nz_ = irz - 1 + l;
                const size_t idx = ((size_t)nr_ * g.nnx + nx_) * g.nnz + nz_;
                acc = (float)((double)acc + field_at<R>(F, to, rs, g.nnr, nr_, nx_, nz_, idx) * w);
            }
    return acc;
}

template <typename R>
__global__ void k_receivers(Grid g, const R* __restrict__ T, const double* __restrict__ toff, const double* __restrict__ ref,
                            int nslots, int nrc, const double* __restrict__ rcv, float* __restrict__ out)
{
    int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nslots * nrc) return;
    int s = gid / nrc, i = gid % nrc;
    const size_t N = (size_t)g.nnr * g.nnx * g.nnz;
    const double* rs = ref ? ref + (size_t)s * (g.nnr + 2) : nullptr;
    out[gid] = receiver_time<R>(g, T + (size_t)s * N, toff[s], rs, rcv[3 * i], rcv[3 * i + 1], rcv[3 * i + 2]);
}

// Receiver times of many (slot, receiver) items at once: item i reads slot
// islot[i] at receiver rcv[3i..3i+2].
template <typename R>
__global__ void k_receivers_items(Grid g, const R* __restrict__ T, const double* __restrict__ toff,
                                  const double* __restrict__ ref, int n, const int* __restrict__ islot,
                                  const double* __restrict__ rcv, float* __restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int s = islot[i];
    const size_t N = (size_t)g.nnr * g.nnx * g.nnz;
    const double* rs = ref ? ref + (size_t)s * (g.nnr + 2) : nullptr;
    out[i] = receiver_time<R>(g, T + (size_t)s * N, toff[s], rs, rcv[3 * i], rcv[3 * i + 1], rcv[3 * i + 2]);
}

// ------------------------------------------------------------ rays + Frechet

__device__ __forceinline__ double tri_vel(const Grid& g, const double* __restrict__ veln, int ipr, int ipx, int ipz,
                                          double drr, double drx, double drz)
{
    const int MX = g.nnx + 3, MZ = g.nnz + 3;
    double vel = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l)
            for (int m = 0; m < 2; ++m) {
                double w = 1.0 - fabs((m * g.dnz - drz) / g.dnz);
                w = w * (1.0 - fabs((l * g.dnx - drx) / g.dnx));
                w = w * (1.0 - fabs((k * g.dnr - drr) / g.dnr));
                int zr = ipr + k, zx = ipx + l, zz = ipz + m;
                if (zz <= g.nnz + 1 && zx <= g.nnx + 1 && zr <= g.nnr + 1 && zr >= -1 && zx >= -1 && zz >= -1)
                    vel += veln[((size_t)(zr + 1) * MX + (zx + 1)) * MZ + (zz + 1)] * w;
            }
    return vel;
}

struct RayOut {
    int32_t* cols;      // staging, ray r at cols + r*cap
    float* vals;
    int32_t* count;     // entries per ray (-1 = start outside the grid, -2 = row overflow, -3 = path overflow)
    int64_t cap;
};

// A ray path point as k_ray_paths records it: a step's start (kind 0) or a
// sub-segment's end (kind 1, over length dinc), at (r, x, z) in vertex cell
// (ivr, ivx, ivz).  A step's start also keeps the propagation-grid cell its
// velocity is interpolated in (ipr, ipx, ipz: the ray's current cell, which
// the start and source rules can clamp); a sub-segment end's cell is the
// one its position falls in.  k_ray_frechet turns 64 points at a time, one
// per lane, into samples (fractions, velocity), so that none of that float64
// work sits on the ray's step-to-step chain.
struct RayPoint {
    double r, x, z, dinc;
    int ivr, ivx, ivz, kind;
    int ipr, ipx, ipz, pad;
};

struct RayPaths {
    RayPoint* rec;      // ray q's points at rec + q * cap
    int32_t* n;         // samples per ray; -1 start outside the grid, -3 more than cap
    int32_t* box;       // vertices touched per ray: rmin, rmax, xmin, xmax, zmin, zmax
    int64_t cap;
};

// Ray paths, one lane per ray: the reference's ray tracer (F:7347-7541) --
// the field's gradient in the current cell, a step of dpl against it,
// clipping at the first grid face crossed -- and its split of each step at
// the vertex-cell faces (F:7542-7600), recorded as points.  The lanes of a
// wavefront trace different rays; a ray's steps depend on each other, so
// the per-point work (the velocity, the B-spline fractions) is left to
// k_ray_frechet, which spreads it over the lanes.
template <typename R>
__global__ void __launch_bounds__(64)
k_ray_paths(Grid g, const R* __restrict__ T, const double* __restrict__ toff, const double* __restrict__ ref,
            int nrc, const double* __restrict__ rcv, int ray0, int nrays, RayPaths P,
            const double* __restrict__ srcpos, const int* __restrict__ islot)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nrays) return;
    const int rid = ray0 + q;
    const int NVR = g.nvr + 2, NVX = g.nvx + 2, NVZ = g.nvz + 2;
    // ray rid: receiver rid % nrc of slot rid / nrc, or (items) the slot
    // islot[rid] at receiver rcv[3 rid .. 3 rid + 2]
    const int s = islot ? islot[rid] : rid / nrc, i = islot ? rid : rid % nrc;
    const size_t N = (size_t)g.nnr * g.nnx * g.nnz;
    const R* F = T + (size_t)s * N;
    const double to = toff[s];
    const double* rs = ref ? ref + (size_t)s * (g.nnr + 2) : nullptr;
    auto TT = [&](int r, int x, int z) -> double {   // 1-based node access
        return field_at<R>(F, to, rs, g.nnr, r - 1, x - 1, z - 1, ((size_t)(r - 1) * g.nnx + (x - 1)) * g.nnz + (z - 1));
    };
    RayPoint* out = P.rec + (int64_t)q * P.cap;
    int64_t ns = 0;
    bool over = false;
    auto put = [&](double r, double x, double z, int pa, int pb, int pc, double dinc, int a, int b, int c, int kind) {
        if (ns >= P.cap) { over = true; return; }
        RayPoint S;
        S.r = r; S.x = x; S.z = z; S.dinc = dinc;
        S.ivr = a; S.ivx = b; S.ivz = c; S.kind = kind;
        S.ipr = pa; S.ipx = pb; S.ipz = pc; S.pad = 0;
        out[ns++] = S;
    };
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
    if (sw) { P.n[q] = -1; return; }
    int bmin[3] = { 1 << 30, 1 << 30, 1 << 30 }, bmax[3] = { -1, -1, -1 };
    double pr = r0, px = x0, pz = z0;
    const double dpl = g.dnr;
    // a point source ends the ray: within one step of it, the last step goes to the source
    const double sr = srcpos[3 * s], sxs = srcpos[3 * s + 1], szs = srcpos[3 * s + 2];
    const bool point = !isnan(sr);
    double xsrc[3] = { 0.0, 0.0, 0.0 };
    if (point) {
        const double rr = g.earth + sr;
        xsrc[0] = rr * sin(sxs) * cos(szs); xsrc[1] = rr * sin(sxs) * sin(szs); xsrc[2] = rr * cos(sxs);
    }
    // the vertex cell of the step's start: the previous step's end cell
    // (the same expression of the same position)
    int ivro = (int)floor((g.gor - pr) / g.dvr) + 1;
    int ivxo = (int)floor((px - g.gox) / g.dvx) + 1;
    int ivzo = (int)floor((pz - g.goz) / g.dvz) + 1;
    for (int step = 1; step <= 100000 && !over; ++step) {
        double dtr = 0.0;
        int cnt = 0;
        const int cz[4] = { 0, 1, 1, 0 }, cx[4] = { 0, 0, 1, 1 };
        for (int k = 0; k < 4; ++k) {
            double lo = TT(ipr + 1, ipx + cx[k], ipz + cz[k]);
            double hi = TT(ipr, ipx + cx[k], ipz + cz[k]);
            // teleseismic: only edges where t decreases downward (the
            // reference's rule); a point source: every edge, since rays
            // climb towards a source above them
            if (point || lo < hi) { dtr = dtr + lo - hi; cnt++; }
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
        double seg = dpl;   // length the Frechet increments use (the reference's nominal step)
        int done = 0;
        if (point) {
            const double sp = sin(px), cp = cos(px);
            const double d0 = crad * sp * cos(pz) - xsrc[0], d1 = crad * sp * sin(pz) - xsrc[1], d2 = crad * cp - xsrc[2];
            const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
            if (d <= dpl || !(gn > 0.0)) { nr_ = sr; nx_ = sxs; nz_ = szs; seg = d; done = 1; }
        }
        int ipro = ipr, ipxo = ipx, ipzo = ipz;
        ipr = (int)floor((g.rgor - nr_) / g.dnr) + 1;
        ipx = (int)floor((nx_ - g.rgox) / g.dnx) + 1;
        ipz = (int)floor((nz_ - g.rgoz) / g.dnz) + 1;
        if (done) {   // the source cell, with the last-node rule of the start point
            ipr = min(max(ipr, 1), g.nnr - 1);
            ipx = min(max(ipx, 1), g.nnx - 1);
            ipz = min(max(ipz, 1), g.nnz - 1);
        }
        int exr = !done && (ipr < 1 || ipr >= g.nnr);
        int exx = !done && (ipx < 1 || ipx >= g.nnx);
        int exz = !done && (ipz < 1 || ipz >= g.nnz);
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
                double qq = (xb - px) / (nx_ - px);
                if (which == 1) { if (qq < atio) { atio = qq; which = 2; } }
                else { atio = qq; which = 2; }
            }
            if (exz) {
                double zb;
                if (ipz < 1) { zb = g.rgoz; ipz = 1; }
                else { zb = g.rgoz + (g.nnz - 1) * g.dnz; ipz = g.nnz - 1; }
                double qq = (zb - pz) / (nz_ - pz);
                if (which != 0) { if (qq < atio) { atio = qq; which = 3; } }
                else { atio = qq; which = 3; }
            }
            if (which == 1) { ipx = ipxo; ipz = ipzo; }
            else if (which == 2) { ipr = ipro; ipz = ipzo; }
            else if (which == 3) { ipr = ipro; ipx = ipxo; }
            nz_ = pz + atio * (nz_ - pz);
            nx_ = px + atio * (nx_ - px);
            nr_ = pr + atio * (nr_ - pr);
            done = 1;
        }
        // the step's sub-segments between vertex-cell faces (F:7542-7600)
        int ivr = (int)floor((g.gor - nr_) / g.dvr) + 1;
        int ivx = (int)floor((nx_ - g.gox) / g.dvx) + 1;
        int ivz = (int)floor((nz_ - g.goz) / g.dvz) + 1;
        int nhp = 0;
        double vrat[4];
        int chp[4];
        if (ivr != ivro) {
            double rb = (ivr > ivro) ? g.gor - (ivr - 1) * g.dvr : g.gor - ivr * g.dvr;
            vrat[nhp] = (rb - pr) / (nr_ - pr); chp[nhp] = 1; nhp++;
        }
        if (ivx != ivxo) {
            double xb = (ivx > ivxo) ? g.gox + (ivx - 1) * g.dvx : g.gox + ivx * g.dvx;
            double qq = (xb - px) / (nx_ - px);
            int p = nhp;
            while (p > 0 && qq < vrat[p - 1]) { vrat[p] = vrat[p - 1]; chp[p] = chp[p - 1]; p--; }
            vrat[p] = qq; chp[p] = 2; nhp++;
        }
        if (ivz != ivzo) {
            double zb = (ivz > ivzo) ? g.goz + (ivz - 1) * g.dvz : g.goz + ivz * g.dvz;
            double qq = (zb - pz) / (nz_ - pz);
            int p = nhp;
            while (p > 0 && qq < vrat[p - 1]) { vrat[p] = vrat[p - 1]; chp[p] = chp[p - 1]; p--; }
            vrat[p] = qq; chp[p] = 3; nhp++;
        }
        vrat[nhp] = 1.0; chp[nhp] = 0; nhp++;
        put(pr, px, pz, ipro, ipxo, ipzo, 0.0, ivro, ivxo, ivzo, 0);
        int ivrt = ivro, ivxt = ivxo, ivzt = ivzo;
        for (int k = 0; k < nhp; ++k) {
            if (k > 0) {
                if (chp[k - 1] == 1) ivrt = ivr;
                else if (chp[k - 1] == 2) ivxt = ivx;
                else if (chp[k - 1] == 3) ivzt = ivz;
            }
            double sz_ = pz + vrat[k] * (nz_ - pz);
            double sx_ = px + vrat[k] * (nx_ - px);
            double sr_ = pr + vrat[k] * (nr_ - pr);
            const double dinc = (k == 0) ? vrat[0] * seg : (vrat[k] - vrat[k - 1]) * seg;
            put(sr_, sx_, sz_, 0, 0, 0, dinc, ivrt, ivxt, ivzt, 1);
            bmin[0] = min(bmin[0], max(ivrt - 1, 0)); bmax[0] = max(bmax[0], min(ivrt + 2, NVR - 1));
            bmin[1] = min(bmin[1], max(ivxt - 1, 0)); bmax[1] = max(bmax[1], min(ivxt + 2, NVX - 1));
            bmin[2] = min(bmin[2], max(ivzt - 1, 0)); bmax[2] = max(bmax[2], min(ivzt + 2, NVZ - 1));
        }
        pr = nr_; px = nx_; pz = nz_;
        ivro = ivr; ivxo = ivx; ivzo = ivz;
        if (done) break;
    }
    P.n[q] = over ? -3 : (int32_t)ns;
    int32_t* bx = P.box + (int64_t)q * 6;
    bx[0] = bmin[0]; bx[1] = bmax[0]; bx[2] = bmin[1]; bx[3] = bmax[1]; bx[4] = bmin[2]; bx[5] = bmax[2];
}

// Weight n of the uniform cubic B-spline at fraction t: bspline4's w[n] in
// the same arithmetic.
__device__ __forceinline__ double bspline_w(double t, int n)
{
    const double s = 1.0 - t, t2 = t * t, t3 = t2 * t;
    const double num = n == 0 ? s * s * s : (n == 1 ? 4.0 - 6.0 * t2 + 3.0 * t3 : (n == 2 ? 1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3 : t3));
    return num / 6.0;
}

// Lane j's double as a wavefront-uniform value.
__device__ __forceinline__ double lane_double(double v, int j)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, j);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), j);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Frechet rows from recorded ray paths, one wavefront per ray (F:7600-7764).
// The points come in 64 at a time, one per lane: the lane turns its point
// into a sample (fractions, velocity) and evaluates the sample's twelve
// B-spline weights (four per axis, bspline4's arithmetic) into LDS, once.
// The samples are then taken one after another: each lane reads the weights
// its tap uses, of this sample and of the one before, from LDS.
// Each lane owns the vertices whose (r, theta, phi) indices are congruent
// to its (lane/16, lane/4 % 4, lane % 4) mod 4, so the 64 taps of every
// sub-segment land on 64 different lanes and every vertex is only ever
// accumulated by one lane (no cross-lane ordering is needed).  A lane keeps
// its current vertex's value in a register while the ray's B-spline cell
// keeps that vertex in its support and writes it back to the dense
// per-worker scratch when the vertex changes; each increment is the
// reference's float(double(acc) + inc).  The scratch is then scanned over
// the touched box in column order, keeping |v| >= 1e-6 (F:7744).
__global__ void __launch_bounds__(64)
k_ray_frechet(Grid g, const double* __restrict__ veln, int nrays, RayPaths P, float* __restrict__ scratch, RayOut out)
{
    // the weights of the batch's samples at slots 1..64, the previous
    // batch's last sample at slot 0: [slot][axis * 4 + n]
    __shared__ double Wt[65][12];
    const int lane = threadIdx.x;
    const int NVR = g.nvr + 2, NVX = g.nvx + 2, NVZ = g.nvz + 2;
    const size_t NV = (size_t)NVR * NVX * NVZ;
    float* fdm = scratch + (size_t)blockIdx.x * NV;
    const int own_r = lane / 16, own_x = (lane / 4) % 4, own_z = lane % 4;
    for (int q = blockIdx.x; q < nrays; q += gridDim.x) {
        const int ns = P.n[q];
        if (ns < 0) { if (lane == 0) out.count[q] = ns; continue; }
        const RayPoint* S = P.rec + (int64_t)q * P.cap;
        int64_t acc_at = -1;   // scratch index of the vertex held in acc, -1 none
        float acc = 0.0f;
        double ivel2 = 1.0;
        // this lane's sample of the current batch of 64 (its weights in Wt)
        double m_ivel2 = 0.0, m_dinc = 0.0;
        int m_ivr = 0, m_ivx = 0, m_ivz = 0, m_kind = 0;
        for (int j = 0; j < ns; ++j) {
            const int jb = j & 63;
            if (jb == 0) {
                // this lane's point of the next 64: its velocity (in the
                // cell the point keeps, or the one it falls in), its
                // fractions of its vertex cell and their weights
                const int k = j + lane;
                double tc[3] = { 0.0, 0.0, 0.0 };
                if (k < ns) {
                    const RayPoint Q = S[k];
                    int ipr = Q.ipr, ipx = Q.ipx, ipz = Q.ipz;
                    if (Q.kind == 1) {
                        ipr = (int)floor((g.rgor - Q.r) / g.dnr) + 1;
                        ipx = (int)floor((Q.x - g.rgox) / g.dnx) + 1;
                        ipz = (int)floor((Q.z - g.rgoz) / g.dnz) + 1;
                    }
                    const double vel = tri_vel(g, veln, ipr, ipx, ipz, (g.rgor - Q.r) - (ipr - 1) * g.dnr,
                                               (Q.x - g.rgox) - (ipx - 1) * g.dnx, (Q.z - g.rgoz) - (ipz - 1) * g.dnz);
                    tc[0] = ((g.gor - Q.r) - (Q.ivr - 1) * g.dvr) / g.dvr;
                    tc[1] = ((Q.x - g.gox) - (Q.ivx - 1) * g.dvx) / g.dvx;
                    tc[2] = ((Q.z - g.goz) - (Q.ivz - 1) * g.dvz) / g.dvz;
                    m_ivel2 = 1.0 / (vel * vel);
                    m_dinc = Q.dinc; m_ivr = Q.ivr; m_ivx = Q.ivx; m_ivz = Q.ivz; m_kind = Q.kind;
                } else {
                    m_ivel2 = m_dinc = 0.0; m_ivr = m_ivx = m_ivz = m_kind = 0;
                }
                __syncthreads();   // the previous batch's reads of Wt are done
                // (before the first sample, the weights at fraction 0)
                if (lane < 12) Wt[0][lane] = j > 0 ? Wt[64][lane] : bspline_w(0.0, lane & 3);
                __syncthreads();
#pragma unroll
                for (int a = 0; a < 3; ++a)
#pragma unroll
                    for (int n = 0; n < 4; ++n) Wt[lane + 1][a * 4 + n] = bspline_w(tc[a], n);
                __syncthreads();
            }
            const double ivel2o = ivel2;
            ivel2 = lane_double(m_ivel2, jb);
            if (__builtin_amdgcn_readlane(m_kind, jb) != 1) continue;
            const double dinc = lane_double(m_dinc, jb);
            const int ivr = __builtin_amdgcn_readlane(m_ivr, jb);
            const int ivx = __builtin_amdgcn_readlane(m_ivx, jb);
            const int ivz = __builtin_amdgcn_readlane(m_ivz, jb);
            // this lane's tap: vertex (iv - 1 + n) with n the lane's residue offset
            const int rr = ivr - 1 + (((own_r - (ivr - 1)) % 4 + 4) % 4);
            const int xx = ivx - 1 + (((own_x - (ivx - 1)) % 4 + 4) % 4);
            const int zz = ivz - 1 + (((own_z - (ivz - 1)) % 4 + 4) % 4);
            const int off[3] = { rr - (ivr - 1), xx - (ivx - 1), zz - (ivz - 1) };
            // this sample's weights and the previous sample's, at the lane's offsets
            const double* Wc = Wt[jb + 1];
            const double* Wp = Wt[jb];
            const double e1 = Wc[off[0]] * Wc[4 + off[1]] * Wc[8 + off[2]] * ivel2;
            const double e0 = Wp[off[0]] * Wp[4 + off[1]] * Wp[8 + off[2]] * ivel2o;
            const double inc = -(e1 + e0) * dinc / 2.0;
            if (zz >= 0 && zz < NVZ && xx >= 0 && xx < NVX && rr >= 0 && rr < NVR) {
                const int64_t at = ((int64_t)zz * NVX + xx) * NVR + rr;
                if (at != acc_at) {
                    if (acc_at >= 0) fdm[acc_at] = acc;
                    acc = fdm[at];
                    acc_at = at;
                }
                acc = (float)(inc + (double)acc);
            }
        }
        if (acc_at >= 0) fdm[acc_at] = acc;
        __syncthreads();
        // sparsify the touched box in column order (F:7739-7764)
        const int32_t* bx = P.box + (int64_t)q * 6;
        const int b0 = bx[0], b1 = bx[1], b2 = bx[2], b3 = bx[3], b4 = bx[4], b5 = bx[5];
        const int BR = b1 - b0 + 1, BX = b3 - b2 + 1, BZ = b5 - b4 + 1;
        const int nbox = (b1 >= 0) ? BR * BX * BZ : 0;
        const int64_t orow = (int64_t)q * out.cap;
        int64_t written = 0;
        bool overflow = false;
        for (int base = 0; base < nbox; base += 64) {
            int e = base + lane;
            float v = 0.0f;
            size_t vi_ = 0;
            if (e < nbox) {
                int br = e % BR, bxx = (e / BR) % BX, bz = e / (BR * BX);
                vi_ = ((size_t)(b4 + bz) * NVX + (b2 + bxx)) * NVR + (b0 + br);
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
        if (lane == 0) out.count[q] = ov ? -2 : (int32_t)written;
        __syncthreads();
    }
}

// ------------------------------------------------------------ context

struct Slots {
    int cap = 0;           // allocated slots
    int n = 0;             // resident sources
    int prec = FMM3D_PREC_FP64;
    size_t es = 0;         // bytes per node of T as allocated (a float64 allocation also holds float32 fields)
    void* T = nullptr;     // fields
    void* P = nullptr;     // ordering fields
    uint32_t* fix = nullptr;
    int* stamp = nullptr;
    int* lastchg = nullptr;
    int* on = nullptr;
    double* toff = nullptr;
    // float32 fields: per slot nnr + 2 values -- the vertical reference
    // A[0..nnr-1] and the lateral steps gx, gz -- so that a node (ir, ix, iz)
    // stores T - toff - (A[ir] + gx*ix + gz*iz); all zero: a plain offset
    double* ref = nullptr;
    int* valchg = nullptr;    // a sweep changed a value
    int pcap = 0;             // slots P holds (allocated by the brick strategy only)
    int* prog = nullptr;      // column walk: per slot per patch layer progress
    int* flow = nullptr;      // column walk: per slot lateral flow signs
    double* srcpos = nullptr; // per slot point-source position (height, colat, lon), NaN if none
    int* walkalt = nullptr;   // per slot: alternate the column walk direction
    int* pend = nullptr;      // column walk: per slot per patch, a node would move (fixed-point check)
    int* wlost = nullptr;     // column walk: per slot, a wait on a neighbour patch was given up
    unsigned long long* lrng = nullptr;   // column walk, alternating slots: per slot per patch, changed layers of two sweeps
    int* slotpend = nullptr;  // column walk: per slot, some node would move
    int* born = nullptr;      // serving engine: per slot, the sweep it joined at
};

}  // namespace

// A relaxation configuration the autotuner chooses between.
struct TuneChoice {
    int strategy, px, pz, passes, check;
};

struct PartState;                  // the reference-partition mode's device state (fmm3d_partition.inl)
static void part_free(PartState* ps);
static void part_model_changed(PartState* ps);

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
    // refined-block contexts: each slot on a block of its own (same shape,
    // its own origin), with slowness and spacings per slot (geo_cap slots)
    int geo_cap = 0;
    double* gslo64 = nullptr;
    float* gslo32 = nullptr;
    double* ghx64 = nullptr;
    double* ghz64 = nullptr;
    float* ghx32 = nullptr;
    float* ghz32 = nullptr;
    int* d_sdims = nullptr;        // per slot (nnr, nnx, nnz) of its block inside the context's grid
    bool have_model = false;
    Slots sl;
    int tile_x = 8, tile_z = 32;   // lateral patch of a brick (theta x phi)
    int brick_sl = 8;              // layers per brick
    int kin = 2;                   // lateral relaxations per layer (at most; stops when settled)
    int dirmode = 0;               // sweep direction policy
    int strategy = 1;              // 0: brick sweeps in three phases; 1: column walk; 2: hyperplane walk
    int col_px = 8, col_pz = 8;    // column-walk patch (theta x phi)
    int col_passes = 16;           // line passes per layer at most
    int plane_ch = 4;              // hyperplane walk: steps per staging chunk (2, 4 or 8)
    int plane_sweeps = 0;          // hyperplane walk: sweeps before the column walk takes over (0: all)
    int use_graph = 1;             // replay sweep chunks from a captured hipGraph
    int ordered_until = 0;         // column walk: sweeps up to this one wait on upstream patches (0: all)
    int check_every = 4;           // column walk: sweeps per convergence check (one graph replay)
    int verify = 1;                // column walk: fixed-point check after this sweep (0: none)
    // the check narrows the next sweep to the moving patches (-1: float32
    // only; in float64 the narrower sweep slowed propagation, 6 -> 17 sweeps
    // on 50 plane waves at 201^3)
    int verify_patches = -1;
    int verify_filter = 1;         // the check after a first sweep skips nodes no later input can move
    int layer_guess = 4;           // fields with a reference: unreached nodes start from the cubic through the four nodes walked before
    int autotune = 1;              // time candidate configurations at the first solve of a kind
    bool tune_fixed = false;       // the caller chose the strategy, patch, pass cap or check interval
    std::string tune_arch;         // device identity for the tuning cache (arch name / compute units)
    int tele_ref = 1;              // float32 teleseismic fields relative to a 1-D plane-wave reference
    double tol_floor = 64.0;       // those fields: a change counts above tol32 * max(|value|, tol_floor)
    int ref_qbits = 22;            // those fields: updates rounded to multiples of 2^-ref_qbits s (0: not)
    double pass_tol = 5e-7;        // those fields: a layer's passes stop after one that moves no node by more (s)
    std::vector<double> layer_slo; // mean slowness of each layer of the current model
    std::vector<std::pair<std::string, TuneChoice>> tuned;   // tuning key -> configuration (this context)
    int* d_iter = nullptr;         // device sweep counter (graph replays)
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<char> g_sig;       // kernel arguments, grid and type the graph was captured with
    int walk_alt = 0;              // this solve: alternate the column walk direction per sweep (2: per slot)
    std::vector<int> walk_slots;   // per slot walk flags when walk_alt == 2 (mixed batches)
    std::vector<double> velv_host; // the vertex model (for refined sub-grids)
    fmm3d_ctx* fine = nullptr;     // refined source neighbourhood (point sources)
    std::vector<fmm3d_ctx*> fines; // the same for batches of point sources: a pool kept by block shape
    int point_workers = 4;         // threads running a batch's refined-block solves
    int point_group_mnodes = 128;  // context nodes (millions) of one group of refined blocks at most
    // continuous-batching engine
    bool eng_open = false;
    int eng_k = 1, eng_order = FMM3D_ORDER_MIXED, eng_pending = 0;
    std::vector<int> eng_active, eng_born;
    int* eng_lc = nullptr;         // pinned host copy of the per-slot last-change sweep
    int64_t* pin_rp = nullptr;     // fmm3d_frechet_pinned's rows: pinned host arrays
    int32_t* pin_cols = nullptr;
    float* pin_vals = nullptr;
    size_t pin_rows = 0;           // entries pin_rp holds
    int64_t pin_cap = 0;           // entries pin_cols / pin_vals hold
    int32_t* pin_dcols = nullptr;  // the device's addresses of pin_cols / pin_vals
    float* pin_dvals = nullptr;
    int32_t* pin_cnt = nullptr;    // Frechet chunks: row entry counts (pinned host)
    size_t pin_cnt_n = 0;
    double* d_faces = nullptr;     // boundary faces of the last teleseismic solve
    size_t d_faces_n = 0;
    // staging of fixed node sets (grow only: a hipFree would wait for every
    // stream of the device, including other threads' refined solves)
    int64_t* st_offs = nullptr;
    int64_t* st_node = nullptr;
    double* st_tv = nullptr;
    size_t st_offs_n = 0, st_n = 0;
    double* d_velv = nullptr;      // the vertex model on the device (gridder input)
    size_t d_velv_n = 0;
    double* d_layer = nullptr;     // per-layer mean slowness (nnr)
    double* d_gather = nullptr;    // refined contexts: the coarse nodes of one block
    size_t d_gather_n = 0;
    float* d_warm = nullptr;       // refined contexts: float32 fields of a warm start
    double* d_warm_off = nullptr;  // and their offsets
    size_t d_warm_n = 0;
    int* d_bad = nullptr;          // first source with a NaN boundary time (or the source count)
    double* d_sides = nullptr;     // side faces of the last teleseismic solve: theta faces, then phi faces
    int* d_side_of = nullptr;      // per source: theta side, phi side (-1 none, 0 low index, 1 high)
    size_t d_sides_n = 0, d_side_of_n = 0;
    // teleseismic sources staged in d_faces (and d_sides): by fmm3d_stage_tele_sides or by a solve
    int stg_n = 0;
    bool stg_sides = false;
    double* d_stg_toff = nullptr;  // per source: the boundary minimum
    double* d_stg_ref = nullptr;   // per source: the plane-wave reference row (zero without a fitted plane)
    double* d_stg_off = nullptr;   // per source: the reference's offset (NaN without a fitted plane)
    int* d_stg_flow = nullptr;     // per source: the lateral flow
    size_t stg_cap = 0;            // sources those arrays hold
    int64_t* eng_node = nullptr;   // device staging of an admission's fixed set
    double* eng_tv = nullptr;
    size_t eng_stage = 0;
    double* eng_faces = nullptr;   // per slot bottom face of a teleseismic admission
    std::vector<int> eng_flow;     // per slot lateral flow (host copy the uploads read from)
    std::vector<double> eng_ref;   // per slot reference row and offset (host copies the uploads read from)
    std::vector<double> eng_off;
    double* eng_rcv = nullptr;     // receiver staging (grows; no free on the serving path)
    float* eng_out = nullptr;
    int* eng_islot = nullptr;
    size_t eng_rcv_n = 0;
    std::vector<double> eng_hrcv;  // host staging of a batched receiver read
    std::vector<int> eng_hslot;
    std::vector<float> eng_hout;
    int ordered = 1;               // column walk: patches wait for their upstream neighbours
    int spin_limit = 20000;        // polls (~64 cycles each) before such a wait is given up
    int max_sweeps = 0;            // column walk: stop after this many sweeps (0: at convergence)
    bool walk_settled = true;      // the last column-walk solve ended at its fixed point (not at max_sweeps)
    int point_order = 1;           // alternating walks: patches outward from the earliest fixed node, waiting on their inner neighbours
    int point_window = 1;          // alternating walks: each patch walks the layers near the last sweeps' changes only
    int point_warm = 1;            // refined blocks: a float32 solve first, then float64 sweeps from its field
    int trace = 0;                 // per-sweep counters to stderr (FMM3D_TRACE)
    int timing = 0;                // time every relaxation launch
    hipEvent_t kev[2 * 8] = {};    // per-launch events (timing)
    double tol64 = 1e-14;          // settle tolerance (relative), float64 fields
    double tol32 = 1e-6;           // settle tolerance (relative), float32 fields (~8-16 ulp)
    fmm3d_stats stats{};
    int64_t* d_updates = nullptr;
    unsigned long long* d_visits = nullptr;
    float* ray_scratch = nullptr;
    PartState* part = nullptr;     // FMM3D_REFERENCE_PARTITION solves (allocated at first use)
    void* scratch = nullptr;       // this device's fmm3d_receivers_noalloc staging (device address)
    bool part_last = false;        // the resident fields come from a partitioned solve
    size_t ray_scratch_workers = 0;
    // Frechet row staging (per ray `fr_rcap` entries), counts, offsets and the packed rows
    int32_t* fr_cols = nullptr;
    float* fr_vals = nullptr;
    size_t fr_stage = 0;           // entries the staging holds
    int32_t* fr_count = nullptr;
    int64_t* fr_off = nullptr;
    size_t fr_rays = 0;            // rays the count/offset arrays hold
    int32_t* fr_pcols = nullptr;
    float* fr_pvals = nullptr;
    size_t fr_packed = 0;          // entries the packed arrays hold
    int64_t fr_rcap = 2048;        // staging entries per ray (grows on overflow)
    RayPoint* fr_path = nullptr;   // ray path points, fr_pcap per ray
    int32_t* fr_pn = nullptr;      // samples per ray
    int32_t* fr_pbox = nullptr;    // touched vertex box per ray
    size_t fr_path_n = 0;          // samples fr_path holds
    size_t fr_path_rays = 0;       // rays fr_pn / fr_pbox hold
    int64_t fr_pcap = 0;           // path samples per ray (grows on overflow)
    double* st_rcv = nullptr;      // receiver coordinates staged for receivers / rays (grown, kept:
    size_t st_rcv_n = 0;           //   a hipFree waits for every stream of the device)
    float* st_rout = nullptr;      // receiver times staged (grown, kept)
    size_t st_rout_n = 0;
    int* st_islot = nullptr;       // per-ray slot indices staged (grown, kept)
    size_t st_islot_n = 0;
};

static thread_local std::string g_create_err;   // errors of calls without a context (per thread)

// Patch shape of the column walks (the hyperplane walk's is 8 x 8).
static inline int patch_x(const fmm3d_ctx* c) { return c->strategy == 2 ? 8 : c->col_px; }
static inline int patch_z(const fmm3d_ctx* c) { return c->strategy == 2 ? 8 : c->col_pz; }

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

// fail() into a caller's string (worker threads that share no context)
static int fail_to(std::string& err, int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    return code;
}

#define HIPCHK(c, expr)                                                                    \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail((c), e_ == hipErrorOutOfMemory ? FMM3D_E_NOMEM : FMM3D_E_DEVICE,   \
                        "%s: %s", #expr, hipGetErrorString(e_));                           \
    } while (0)

template <typename P> static hipError_t dalloc(P** p, size_t bytes);

// A device staging buffer of at least n elements, grown and kept.
template <typename P> static hipError_t staging(P** p, size_t* have, size_t n)
{
    if (*have >= n) return hipSuccess;
    (void)hipFree(*p);
    *p = nullptr;
    *have = 0;
    hipError_t e = dalloc(p, n * sizeof(P));
    if (e == hipSuccess) *have = n;
    return e;
}

template <typename P> static hipError_t dalloc(P** p, size_t bytes)
{
    return hipMalloc((void**)p, bytes ? bytes : 16);
}

extern "C" const char* fmm3d_last_error(const fmm3d_ctx* c)
{
    return c ? c->err.c_str() : g_create_err.c_str();
}

static int ctx_setup(fmm3d_ctx* c, int device, fmm3d_ctx** out);
static int upload_metrics(fmm3d_ctx* c);

// The solver's grid from the .vtx header values: degrees to radians with
// the reference's float32-rounded pi (F:61, 6103-6117).
static void grid_from_desc(const fmm3d_grid_desc* d, Grid& g)
{
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
    grid_from_desc(d, c->g);
    return ctx_setup(c, device, out);
}

// Metric tables in the reference's arithmetic (F:541-542): hx[ir] = ri*dnx,
// hz[ir][ix] = ri*sin(theta)*dnz, ri = rgor - ir*dnr + earth (0-based).
static int upload_metrics(fmm3d_ctx* c)
{
    const Grid& g = c->g;
    std::vector<double> hx(g.nnr), hz((size_t)g.nnx * g.nnr);
    std::vector<float> hx32(g.nnr), hz32((size_t)g.nnx * g.nnr);
    for (int ir = 0; ir < g.nnr; ++ir) {
        double ri = g.rgor - (ir) * g.dnr + g.earth;
        hx[ir] = ri * g.dnx;
        hx32[ir] = (float)hx[ir];
        for (int ix = 0; ix < g.nnx; ++ix) {
            double sth = sin(g.rgox + (ix) * g.dnx);
            hz[(size_t)ir * g.nnx + ix] = ri * sth * g.dnz;
            hz32[(size_t)ir * g.nnx + ix] = (float)hz[(size_t)ir * g.nnx + ix];
        }
    }
    if (!c->hx64) {
        HIPCHK(c, dalloc(&c->hx64, hx.size() * sizeof(double)));
        HIPCHK(c, dalloc(&c->hz64, hz.size() * sizeof(double)));
        HIPCHK(c, dalloc(&c->hx32, hx32.size() * sizeof(float)));
        HIPCHK(c, dalloc(&c->hz32, hz32.size() * sizeof(float)));
    }
    HIPCHK(c, hipMemcpy(c->hx64, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->hz64, hz.data(), hz.size() * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->hx32, hx32.data(), hx32.size() * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->hz32, hz32.data(), hz32.size() * sizeof(float), hipMemcpyHostToDevice));
    return FMM3D_OK;
}

// ------------------------------------------------------------ allocation-free staging

// fmm3d_receivers_noalloc stages receiver coordinates and times here rather
// than in device memory allocated per call.  Each device holds its own copy
// until the process ends; every call on a device shares it.
constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)8 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_dev[MAX_DEVICES] = {};   // device address of `scratch`, per device

// The device address of this device's scratch array, looked up once per
// device with that device current.  In host code the array's name is only
// a stand-in for the symbol: kernels and copies take this address.
static hipError_t scratch_address(int device, void** out)
{
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    if (!g_scratch_dev[device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

// Device resources of a context whose grid is set (also used for the
// refined sub-grid of point sources).
static int ctx_setup(fmm3d_ctx* c, int device, fmm3d_ctx** out)
{
    const Grid& g = c->g;
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
    CREATECHK(scratch_address(device, &c->scratch));
    CREATECHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CREATECHK(hipEventCreate(&c->ev0));
    CREATECHK(hipEventCreate(&c->ev1));
    for (auto& e : c->kev) CREATECHK(hipEventCreate(&e));
    size_t NH = (size_t)(g.nnr + 3) * (g.nnx + 3) * (g.nnz + 3);
    CREATECHK(dalloc(&c->veln, NH * sizeof(double)));
    CREATECHK(dalloc(&c->slo64, c->N * sizeof(double)));
    CREATECHK(dalloc(&c->slo32, c->N * sizeof(float)));
    CREATECHK(dalloc(&c->d_updates, sizeof(int64_t)));
    CREATECHK(dalloc(&c->d_visits, 3 * sizeof(unsigned long long)));
    CREATECHK(dalloc(&c->d_iter, sizeof(int)));
    rc = upload_metrics(c);
    if (rc) { fmm3d_destroy(c); return rc; }
#undef CREATECHK
    const char* tl = getenv("FMM3D_TOL");    // "tol64[:tol32]"
    if (tl) sscanf(tl, "%lf:%lf", &c->tol64, &c->tol32);
    const char* dm = getenv("FMM3D_DIRMODE");
    if (dm) c->dirmode = atoi(dm);
    const char* ts = getenv("FMM3D_TILE");   // "SLxCXxCZ[:K]"
    if (ts) {
        int a, b, e, k = c->kin;
        int n = sscanf(ts, "%dx%dx%d:%d", &a, &b, &e, &k);
        if (n >= 3) { c->brick_sl = a; c->tile_x = b; c->tile_z = e; c->kin = k; }
    }
    const char* od = getenv("FMM3D_ORDERED");
    if (od) c->ordered = atoi(od) != 0;
    const char* sp = getenv("FMM3D_SPIN");
    if (sp) c->spin_limit = std::max(0, atoi(sp));
    const char* at = getenv("FMM3D_AUTOTUNE");
    if (at) c->autotune = atoi(at) != 0;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, c->device) == hipSuccess)
            c->tune_arch = std::string(prop.gcnArchName) + "/" + std::to_string(prop.multiProcessorCount) + "cu";
        else
            c->tune_arch = "unknown";
    }
    const char* trf = getenv("FMM3D_TELE_REF");
    if (trf) c->tele_ref = atoi(trf) != 0;
    const char* tfl = getenv("FMM3D_TOL_FLOOR");
    if (tfl) c->tol_floor = std::max(0.0, atof(tfl));
    const char* rqb = getenv("FMM3D_REF_QBITS");
    if (rqb) c->ref_qbits = std::min(30, std::max(0, atoi(rqb)));
    const char* ptl = getenv("FMM3D_PASS_TOL");
    if (ptl) c->pass_tol = std::min(1e-3, std::max(0.0, atof(ptl)));
    const char* vfl = getenv("FMM3D_VERIFY_FILTER");
    if (vfl) c->verify_filter = atoi(vfl) != 0;
    const char* lgs = getenv("FMM3D_LAYER_GUESS");
    if (lgs) c->layer_guess = std::min(4, std::max(0, atoi(lgs)));
    const char* gv = getenv("FMM3D_GRAPH");
    if (gv) c->use_graph = atoi(gv) != 0;
    const char* vf = getenv("FMM3D_VERIFY");
    if (vf) c->verify = std::max(0, atoi(vf));
    const char* tr = getenv("FMM3D_TRACE");
    if (tr) c->trace = atoi(tr);
    const char* sg = getenv("FMM3D_STRATEGY");   // "brick" | "column[:PXxPZ[:passes]]" | "plane[:chunk]"
    if (sg) {
        c->tune_fixed = true;
        if (!strncmp(sg, "brick", 5)) c->strategy = 0;
        else if (!strncmp(sg, "plane", 5)) {
            c->strategy = 2;
            int ch = 0;
            if (sscanf(sg, "plane:%d", &ch) == 1 && (ch == 2 || ch == 4 || ch == 8)) c->plane_ch = ch;
        }
        else if (!strncmp(sg, "column", 6)) {
            c->strategy = 1;
            int a, b;
            int np_ = 0;
            if (sscanf(sg, "column:%dx%d:%d", &a, &b, &np_) == 3 && np_ >= 1 && np_ <= 16) c->col_passes = np_;
            if (sscanf(sg, "column:%dx%d", &a, &b) == 2 && (a == 8 || a == 16 || a == 32 || a == 64) &&
                (b == 8 || b == 16 || b == 32 || b == 64)) { c->col_px = a; c->col_pz = b; }
        }
    }
    *out = c;
    return FMM3D_OK;
}

static void free_slots(Slots& s)
{
    (void)hipFree(s.T); (void)hipFree(s.P); (void)hipFree(s.fix); (void)hipFree(s.stamp);
    (void)hipFree(s.lastchg); (void)hipFree(s.on); (void)hipFree(s.toff); (void)hipFree(s.ref);
    (void)hipFree(s.valchg); (void)hipFree(s.prog); (void)hipFree(s.flow); (void)hipFree(s.srcpos); (void)hipFree(s.walkalt);
    (void)hipFree(s.pend); (void)hipFree(s.wlost); (void)hipFree(s.slotpend); (void)hipFree(s.born); (void)hipFree(s.lrng);
    s = Slots();
}

extern "C" int fmm3d_destroy(fmm3d_ctx* c)
{
    if (!c) return FMM3D_OK;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->fine) { fmm3d_destroy(c->fine); c->fine = nullptr; }
    for (fmm3d_ctx* f : c->fines)
        if (f) fmm3d_destroy(f);
    c->fines.clear();
    if (c->eng_lc) (void)hipHostFree(c->eng_lc);
    if (c->pin_rp) (void)hipHostFree(c->pin_rp);
    if (c->pin_cols) (void)hipHostFree(c->pin_cols);
    if (c->pin_vals) (void)hipHostFree(c->pin_vals);
    if (c->pin_cnt) (void)hipHostFree(c->pin_cnt);
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    (void)hipFree(c->d_iter);
    (void)hipFree(c->d_faces);
    (void)hipFree(c->st_offs); (void)hipFree(c->st_node); (void)hipFree(c->st_tv); (void)hipFree(c->d_velv); (void)hipFree(c->d_bad);
    (void)hipFree(c->d_layer); (void)hipFree(c->st_rcv); (void)hipFree(c->st_rout); (void)hipFree(c->st_islot);
    (void)hipFree(c->d_sides); (void)hipFree(c->d_side_of);
    (void)hipFree(c->d_stg_toff); (void)hipFree(c->d_stg_ref); (void)hipFree(c->d_stg_off); (void)hipFree(c->d_stg_flow);
    (void)hipFree(c->eng_node); (void)hipFree(c->eng_tv); (void)hipFree(c->eng_faces);
    (void)hipFree(c->eng_rcv); (void)hipFree(c->eng_out); (void)hipFree(c->eng_islot);
    free_slots(c->sl);
    part_free(c->part);
    (void)hipFree(c->veln); (void)hipFree(c->slo64); (void)hipFree(c->slo32);
    (void)hipFree(c->gslo64); (void)hipFree(c->gslo32);
    (void)hipFree(c->ghx64); (void)hipFree(c->ghz64); (void)hipFree(c->ghx32); (void)hipFree(c->ghz32);
    (void)hipFree(c->d_sdims);
    (void)hipFree(c->d_gather);
    (void)hipFree(c->d_warm); (void)hipFree(c->d_warm_off);
    (void)hipFree(c->hx64); (void)hipFree(c->hz64); (void)hipFree(c->hx32); (void)hipFree(c->hz32);
    (void)hipFree(c->d_updates); (void)hipFree(c->d_visits); (void)hipFree(c->ray_scratch);
    (void)hipFree(c->fr_path); (void)hipFree(c->fr_pn); (void)hipFree(c->fr_pbox);
    (void)hipFree(c->fr_cols); (void)hipFree(c->fr_vals); (void)hipFree(c->fr_count); (void)hipFree(c->fr_off);
    (void)hipFree(c->fr_pcols); (void)hipFree(c->fr_pvals);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    for (auto& e : c->kev) if (e) (void)hipEventDestroy(e);
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
    c->velv_host.assign(velv, velv + nv);
    if (c->d_velv_n < nv) {
        (void)hipFree(c->d_velv);
        c->d_velv = nullptr;
        c->d_velv_n = 0;
        HIPCHK(c, dalloc(&c->d_velv, nv * sizeof(double)));
        c->d_velv_n = nv;
    }
    double* d = c->d_velv;
    hipError_t e = hipMemcpyAsync(d, velv, nv * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        long total = (long)(g.nnr + 3) * (g.nnx + 3) * (g.nnz + 3);
        int blocks = (int)std::min<long>((total + 255) / 256, 65536);
        k_gridder<<<blocks, 256, 0, c->stream>>>(g, d, c->veln, c->slo64, c->slo32, g.nnx, g.nnz);
        e = hipGetLastError();
    }
    // the per-layer mean slowness
    if (e == hipSuccess && !c->d_layer) e = dalloc(&c->d_layer, (size_t)g.nnr * sizeof(double));
    if (e == hipSuccess) {
        k_layer_mean<<<g.nnr, 256, 0, c->stream>>>(c->slo64, (size_t)g.nnx * g.nnz, c->d_layer);
        e = hipGetLastError();
        c->layer_slo.assign(g.nnr, 0.0);
        if (e == hipSuccess) e = hipMemcpyAsync(c->layer_slo.data(), c->d_layer, g.nnr * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "gridder: %s", hipGetErrorString(e));
    c->have_model = true;
    part_model_changed(c->part);
    return FMM3D_OK;
}

// relaxation launches between convergence checks
#define FMM3D_CHECK 4

// brick geometry of the current configuration
struct Bricks {
    int nbr, nbx, nbz, sl;
    int n() const { return nbr * nbx * nbz; }
};

static Bricks bricks_of(const fmm3d_ctx* c)
{
    Bricks b;
    b.sl = c->brick_sl;
    b.nbr = (c->g.nnr + b.sl - 1) / b.sl;
    b.nbx = (c->g.nnx + c->tile_x - 1) / c->tile_x;
    b.nbz = (c->g.nnz + c->tile_z - 1) / c->tile_z;
    return b;
}

static int ensure_slots(fmm3d_ctx* c, int nsrc, int prec)
{
    Slots& s = c->sl;
    c->part_last = false;   // every solve starts here
    const size_t es = prec == FMM3D_PREC_FP32 ? sizeof(float) : sizeof(double);
    if (s.cap >= nsrc && (s.prec == prec || s.es >= es)) {
        s.prec = prec;
        return FMM3D_OK;
    }
    free_slots(s);
    const size_t N = c->N;
    // stamps for the largest brick count any supported shape gives (4-layer bricks, 4x8 patches)
    const size_t nbmax = (size_t)((c->g.nnr + 3) / 4) * ((c->g.nnx + 3) / 4) * ((c->g.nnz + 7) / 8);
    HIPCHK(c, dalloc(&s.T, (size_t)nsrc * N * es));
    HIPCHK(c, dalloc(&s.fix, (size_t)nsrc * ((N + 31) / 32) * sizeof(uint32_t)));
    HIPCHK(c, dalloc(&s.stamp, (size_t)nsrc * nbmax * sizeof(int)));
    HIPCHK(c, dalloc(&s.lastchg, nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.on, nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.toff, nsrc * sizeof(double)));
    HIPCHK(c, dalloc(&s.ref, (size_t)nsrc * (c->g.nnr + 2) * sizeof(double)));
    HIPCHK(c, dalloc(&s.valchg, nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.prog, (size_t)nsrc * nbmax * sizeof(int)));
    HIPCHK(c, dalloc(&s.flow, (size_t)nsrc * 2 * sizeof(int)));
    HIPCHK(c, dalloc(&s.srcpos, (size_t)nsrc * 3 * sizeof(double)));
    HIPCHK(c, dalloc(&s.walkalt, (size_t)nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.pend, (size_t)nsrc * nbmax * sizeof(int)));
    HIPCHK(c, dalloc(&s.wlost, (size_t)nsrc * sizeof(int)));
    HIPCHK(c, hipMemset(s.wlost, 0, (size_t)nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.lrng, (size_t)nsrc * nbmax * 2 * sizeof(unsigned long long)));
    HIPCHK(c, dalloc(&s.slotpend, (size_t)nsrc * sizeof(int)));
    HIPCHK(c, dalloc(&s.born, (size_t)nsrc * sizeof(int)));
    s.cap = nsrc;
    s.prec = prec;
    s.es = es;
    return FMM3D_OK;
}

// ------------------------------------------------------------ solve driver

// Teleseismic initial state (start_fmm, F:6914-6939): bottom-face nodes with
// a positive boundary time hold it and are fixed; every other node unknown.
template <typename R>
__global__ void k_init_tele(int nnr, int nnx, int nnz, int nslots, const double* __restrict__ faces, R* __restrict__ T,
                            const double* __restrict__ toff, const double* __restrict__ ref)
{
    // one slot per grid row; the bottom layer is the index range from base
    const size_t NL = (size_t)nnx * nnz, N = NL * nnr, base = N - NL;
    constexpr int V = 16 / sizeof(R);   // values per 16-byte store
    typedef R RV __attribute__((ext_vector_type(V)));
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nth = (size_t)gridDim.x * blockDim.x;
    const R inf = Lim<R>::inf();
    for (size_t s = blockIdx.y; s < (size_t)nslots; s += gridDim.y) {
        R* Ts = T + s * N;
        // the layers above the bottom are unknown: 16-byte stores from the
        // slot's first 16-byte aligned value, single values before and after
        const size_t head = std::min(base, (size_t)((16 - ((uintptr_t)Ts & 15)) & 15) / sizeof(R));
        const size_t nv = (base - head) / V;
        RV* Tv = reinterpret_cast<RV*>(Ts + head);
        RV iv;
#pragma unroll
        for (int k = 0; k < V; ++k) iv[k] = inf;
        for (size_t i = tid; i < nv; i += nth) Tv[i] = iv;
        if (tid < head) Ts[tid] = inf;
        for (size_t c = head + nv * V + tid; c < base; c += nth) Ts[c] = inf;
        const double* fs = faces + s * NL;
        const double off = toff[s];
        const double* rs = ref ? ref + s * (nnr + 2) : nullptr;
        for (size_t c = base + tid; c < N; c += nth) {
            R v = inf;
            const double fv = fs[c - base];
            const size_t ix = (c - base) / nnz, iz = (c - base) % nnz;
            if (fv > 0.0) v = std::is_same<R, double>::value ? (R)fv : (R)(fv - off - ref_at(rs, nnr, nnr - 1, ix, iz));
            Ts[c] = v;
        }
    }
}

// Per source (one block): the smallest positive boundary time over its
// bottom face and its selected side faces -- the float32 offset, 0 when
// none is positive -- and the lowest source index holding a NaN.
__global__ void __launch_bounds__(1024) k_tele_offsets(size_t F, size_t XF, size_t ZF, const double* __restrict__ faces,
                                                      const double* __restrict__ xf, const double* __restrict__ zf,
                                                      const int* __restrict__ side, double* __restrict__ toff,
                                                      int* __restrict__ bad)
{
    __shared__ double sm[1024];
    __shared__ int sn[1024];
    const int s = blockIdx.x, t = threadIdx.x;
    double m = __builtin_huge_val();
    int nan = 0;
    auto scan = [&](const double* p, size_t n) {
        for (size_t i = t; i < n; i += blockDim.x) {
            const double v = p[i];
            nan |= (v != v);
            m = (v > 0.0 && v < m) ? v : m;
        }
    };
    scan(faces + (size_t)s * F, F);
    if (side) {
        if (side[2 * s] >= 0) scan(xf + (size_t)s * XF, XF);
        if (side[2 * s + 1] >= 0) scan(zf + (size_t)s * ZF, ZF);
    }
    sm[t] = m;
    sn[t] = nan;
    __syncthreads();
    for (int o = (int)blockDim.x / 2; o > 0; o >>= 1) {
        if (t < o) {
            sm[t] = sm[t + o] < sm[t] ? sm[t + o] : sm[t];
            sn[t] |= sn[t + o];
        }
        __syncthreads();
    }
    if (t == 0) {
        toff[s] = sm[0] < __builtin_huge_val() ? sm[0] : 0.0;
        if (sn[0]) atomicMin(bad, s);
    }
}

__global__ void k_fixbits_tele(int nnr, int nnx, int nnz, int nslots, const double* __restrict__ faces, uint32_t* __restrict__ fix)
{
    // one slot per grid row; only words from wb on reach the bottom layer
    // (nodes from base on) and hold bits
    const size_t NL = (size_t)nnx * nnz, N = NL * nnr, base = N - NL;
    const size_t NW = (N + 31) / 32, wb = base / 32;
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    const size_t tid = blockIdx.x * (size_t)blockDim.x + threadIdx.x, nth = (size_t)gridDim.x * blockDim.x;
    const int lane = threadIdx.x & 63;
    for (size_t s = blockIdx.y; s < (size_t)nslots; s += gridDim.y) {
        const double* fs = faces + s * NL;
        uint32_t* fw = fix + s * NW;
        // words below wb: zero, in 16-byte stores from the first aligned word
        const size_t head = std::min(wb, (size_t)((16 - ((uintptr_t)fw & 15)) & 15) / 4);
        const size_t nv = (wb - head) / 4;
        U4* fv = reinterpret_cast<U4*>(fw + head);
        for (size_t i = tid; i < nv; i += nth) fv[i] = U4{ 0u, 0u, 0u, 0u };
        if (tid < head) fw[tid] = 0u;
        for (size_t w = head + nv * 4 + tid; w < wb; w += nth) fw[w] = 0u;
        // words from wb: one node per lane; a wavefront's ballot is two
        // words (every wavefront starts at a multiple of 64 nodes from wb)
        const size_t nb = (NW - wb) * 32, nbr = (nb + 63) / 64 * 64;
        for (size_t j = tid; j < nbr; j += nth) {
            const size_t c = wb * 32 + j;
            const bool f = c >= base && c < N && fs[c - base] > 0.0;
            const uint64_t m = __ballot(f);
            if (lane == 0) {
                const size_t w = wb + j / 32;
                fw[w] = (uint32_t)m;
                if (w + 1 < NW) fw[w + 1] = (uint32_t)(m >> 32);
            }
        }
    }
}

// Side-face injection after the bottom face: values > 0 on a slot's theta
// face (ix = 0 or nnx-1, every ir and iz; (nnr, nnz) phi fastest) and phi
// face (iz = 0 or nnz-1, every ir and ix; (nnr, nnx) theta fastest) become
// fixed nodes, like the bottom face's (float32: minus the slot's offset and
// reference).  A
// side of -1 injects nothing.  Bits of one word come from several threads,
// hence atomicOr.
template <typename R>
__global__ void k_init_sides(int nnr, int nnx, int nnz, int nslots, const double* __restrict__ xf,
                             const double* __restrict__ zf, const int* __restrict__ side,
                             const double* __restrict__ toff, const double* __restrict__ ref, R* __restrict__ T,
                             uint32_t* __restrict__ fix)
{
    const size_t N = (size_t)nnr * nnx * nnz, NW = (N + 31) / 32;
    const size_t XF = (size_t)nnr * nnz, ZF = (size_t)nnr * nnx, per = XF + ZF;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < per * nslots; e += (size_t)gridDim.x * blockDim.x) {
        const size_t s = e / per, k = e % per;
        size_t ir, ix, iz;
        double v;
        if (k < XF) {
            const int xs = side[2 * s];
            if (xs < 0) continue;
            ir = k / nnz; iz = k % nnz; ix = xs ? nnx - 1 : 0;
            v = xf[s * XF + k];
        } else {
            const size_t k2 = k - XF;
            const int zs = side[2 * s + 1];
            if (zs < 0) continue;
            ir = k2 / nnx; ix = k2 % nnx; iz = zs ? nnz - 1 : 0;
            v = zf[s * ZF + k2];
        }
        if (!(v > 0.0)) continue;
        const size_t c = (ir * nnx + ix) * nnz + iz;
        const double rv = ref_at(ref ? ref + s * (nnr + 2) : nullptr, nnr, ir, ix, iz);
        T[s * N + c] = std::is_same<R, double>::value ? (R)v : (R)(v - toff[s] - rv);
        atomicOr(fix + s * NW + (c >> 5), 1u << (c & 31));
    }
}

// General fixed (Dirichlet) sets: every slot's field starts unknown, then
// each listed node takes its time (float32: minus the slot's offset) and its
// fixed bit.  Bits of one word can come from several entries, hence atomicOr.
template <typename R>
__global__ void k_fill_real(R* __restrict__ p, size_t n, R v)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

__global__ void k_fill_u32(uint32_t* __restrict__ p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// Free nodes of float64 fields T take the float32 fields W plus their
// slot's offset (fixed nodes keep their times): the start of a float64
// relaxation from a float32 one.
__global__ void k_warm_start(double* __restrict__ T, const float* __restrict__ W, const double* __restrict__ woff,
                             const uint32_t* __restrict__ fix, int nslots, size_t N)
{
    const size_t NW = (N + 31) / 32;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nslots * N; i += (size_t)gridDim.x * blockDim.x) {
        const size_t slot = i / N, k = i - slot * N;
        if ((fix[slot * NW + (k >> 5)] >> (k & 31)) & 1u) continue;
        const float w = W[i];
        T[i] = w < __builtin_huge_valf() ? woff[slot] + (double)w : __builtin_huge_val();
    }
}

// Layer ranges of slots whose first sweep is `first`: every layer changed
// in the sweep before (so that sweep walks whole columns), nothing in the
// other parity's entry.
__global__ void k_lrng_init(unsigned long long* __restrict__ p, size_t n, int first)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        p[2 * i + ((first - 1) & 1)] = lrng_pack(first - 1, 0, 0xffff);
        p[2 * i + (first & 1)] = lrng_pack(-1000, 0, 0);
    }
}

// The coarse nodes a refined block covers: every refine-th node of the
// block's float64 field T (laid out with row strides nnx x nnz), cr x cx x
// cz of them, into out in (r, theta, phi) order.
__global__ void k_gather_coarse(const double* __restrict__ T, int cr, int cx, int cz, int refine, int nnx, int nnz,
                                double* __restrict__ out)
{
    const size_t n = (size_t)cr * cx * cz;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int z = (int)(i % cz), x = (int)((i / cz) % cx), r = (int)(i / ((size_t)cz * cx));
        out[i] = T[((size_t)r * refine * nnx + (size_t)x * refine) * nnz + (size_t)z * refine];
    }
}

template <typename R>
__global__ void k_scatter_fixed(int nslots, size_t N, const int64_t* __restrict__ offs, const int64_t* __restrict__ node,
                                const double* __restrict__ tv, const double* __restrict__ toff, R* __restrict__ T,
                                uint32_t* __restrict__ fix)
{
    const size_t NW = (N + 31) / 32;
    const int64_t total = offs[nslots];
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        int lo = 0, hi = nslots - 1;            // slot of entry i: offs[s] <= i < offs[s+1]
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (offs[mid] <= i) lo = mid; else hi = mid - 1;
        }
        const size_t k = (size_t)node[i];
        T[(size_t)lo * N + k] = std::is_same<R, double>::value ? (R)tv[i] : (R)(tv[i] - toff[lo]);
        atomicOr(fix + (size_t)lo * NW + (k >> 5), 1u << (k & 31));
    }
}

// Phase-start brick stamps: phase A marks the bottom bricks (they hold the
// fixed nodes) as changed at sweep 0 and the rest as never; phases B/C
// mark every brick.
__global__ void k_stamp_init(int nslots, int nb, int nbxz, int nbr, int all, int* stamp)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < (size_t)nslots * nb; i += (size_t)gridDim.x * blockDim.x) {
        int t = (int)(i % nb);
        stamp[i] = all ? 0 : ((t / nbxz == nbr - 1) ? 0 : -1);
    }
}

// Offsets of a chunk's rows in the packed array: the exclusive prefix sum
// of the entry counts (a negative count -- a row to redo -- takes none).
// One block of 1024 threads, each summing a run of counts.
__global__ void __launch_bounds__(1024) k_row_offsets(int nr, const int32_t* __restrict__ count, int64_t* __restrict__ off)
{
    __shared__ int64_t part[1024];
    const int t = threadIdx.x, per = (nr + 1023) / 1024;
    const int q0 = min(nr, t * per), q1 = min(nr, q0 + per);
    int64_t run = 0;
    for (int q = q0; q < q1; ++q) run += max(count[q], 0);
    part[t] = run;
    __syncthreads();
    for (int d = 1; d < 1024; d <<= 1) {   // inclusive scan of the runs
        const int64_t v = t >= d ? part[t - d] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int64_t o = part[t] - run;
    for (int q = q0; q < q1; ++q) {
        off[q] = o;
        o += max(count[q], 0);
    }
}

// Frechet rows from the per-ray staging into one packed array: ray q's
// count[q] entries move to off[q] (one block per ray, grid-stride).
__global__ void k_pack_rows(int nr, int64_t rcap, const int32_t* __restrict__ count, const int64_t* __restrict__ off,
                            const int32_t* __restrict__ scols, const float* __restrict__ svals,
                            int32_t* __restrict__ pcols, float* __restrict__ pvals, int64_t room)
{
    for (int q = blockIdx.x; q < nr; q += gridDim.x) {
        const int n = count[q];
        if (n <= 0) continue;
        const int64_t src = (int64_t)q * rcap, dst = off[q];
        if (dst + n > room) continue;   // past the destination's end: the caller sees the total and retries
        for (int i = threadIdx.x; i < n; i += blockDim.x) {
            pcols[dst + i] = scols[src + i];
            pvals[dst + i] = svals[src + i];
        }
    }
}

// the next sweep number, for graph replays of a sweep chunk
__global__ void k_next_iter(int* p)
{
    if (threadIdx.x == 0) p[0] = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1;
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
    const int tx = c->tile_x, tz = c->tile_z;
#define FMM3D_PATCH_CASE(A, B)                                                   \
    if (tx == A && tz == B) {                                                    \
        k_brick<R, A, B, SHIFT><<<blocks, A * B, 0, c->stream>>>(a);              \
        return hipGetLastError();                                                \
    }
    FMM3D_PATCH_CASE(8, 32)
    FMM3D_PATCH_CASE(4, 32)
    FMM3D_PATCH_CASE(4, 16)
    FMM3D_PATCH_CASE(8, 16)
    FMM3D_PATCH_CASE(16, 16)
    FMM3D_PATCH_CASE(4, 64)
    FMM3D_PATCH_CASE(2, 32)
    FMM3D_PATCH_CASE(4, 8)
#undef FMM3D_PATCH_CASE
    return hipErrorInvalidValue;
}

template <typename R, bool SHIFT>
static int run_phases(fmm3d_ctx* c, int nsrc, int order, int* sweeps_out)
{
    Slots& s = c->sl;
    const Grid& g = c->g;
    const Bricks B = bricks_of(c);
    const int nb = B.n();
    const size_t nbmax = (size_t)((g.nnr + 3) / 4) * ((g.nnx + 3) / 4) * ((g.nnz + 7) / 8);
    if (B.sl < 4 || (size_t)nb > nbmax)
        return fail(c, FMM3D_E_ARG, "solver: brick shape %dx%dx%d not supported on this grid", B.sl, c->tile_x, c->tile_z);
    if (s.pcap < nsrc) {
        (void)hipFree(s.P);
        s.P = nullptr;
        s.pcap = 0;
        HIPCHK(c, dalloc(&s.P, (size_t)s.cap * c->N * sizeof(R)));
        s.pcap = s.cap;
    }
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
    a.nbr = B.nbr; a.nbx = B.nbx; a.nbz = B.nbz;
    a.sl = B.sl;
    a.kin = c->kin;
    a.nslots = nsrc;
    a.fom = order == FMM3D_ORDER_MIXED;
    a.visits = c->d_visits;
    a.tol = std::is_same<R, double>::value ? c->tol64 : c->tol32;
    a.valchg = s.valchg;
    a.dirmode = c->dirmode;
    a.updates = c->d_visits + 1;
    const int blocks = nsrc * nb;
    const int maxit = 4 * (g.nnr + g.nnx + g.nnz) + 200;
    const int CHECK = FMM3D_CHECK;
    std::vector<int> lc(nsrc);
    int total_sweeps = 0;
    // A, B, C (+ up to 2 repeats of C for float64; float32 stops after one C)
    const int nphases = order != FMM3D_ORDER_MIXED ? 1 : (std::is_same<R, double>::value ? 5 : 3);
    for (int ph = 0; ph < nphases; ++ph) {
        a.phase = ph == 0 ? 0 : 1;
        if (ph >= 1) {
            // freeze the ordering: P <- T; only fixed nodes start final
            size_t n = (size_t)nsrc * c->N;
            k_copy<R><<<4096, 256, 0, c->stream>>>((const R*)s.T, (R*)s.P, n);
            k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.valchg, nsrc, 0);
            HIPCHK(c, hipGetLastError());
        }
        k_stamp_init<<<std::min(4096, (blocks + 255) / 256), 256, 0, c->stream>>>(nsrc, nb, B.nbx * B.nbz, B.nbr, ph >= 1, s.stamp);
        k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.lastchg, nsrc, 0);
        k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.on, nsrc, 1);
        HIPCHK(c, hipGetLastError());
        int k = 1;
        int last_k_with_change = 0;
        bool done = false;
        while (!done) {
            for (int j = 0; j < CHECK; ++j, ++k) {
                a.iter = k;
                a.dir = (k & 1) ? 1 : -1;
                if (c->timing) HIPCHK(c, hipEventRecord(c->kev[2 * j], c->stream));
                {
                    hipError_t le = launch_sweep<R, SHIFT>(c, a, blocks);
                    if (le == hipErrorInvalidValue)
                        return fail(c, FMM3D_E_ARG, "solver: unsupported patch %dx%d", c->tile_x, c->tile_z);
                    HIPCHK(c, le);
                }
                if (c->timing) HIPCHK(c, hipEventRecord(c->kev[2 * j + 1], c->stream));
                c->stats.kernel_launches++;
            }
            HIPCHK(c, hipMemcpyAsync(lc.data(), s.lastchg, nsrc * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (c->timing)
                for (int j = 0; j < CHECK; ++j) {
                    float ms = 0.0f;
                    HIPCHK(c, hipEventElapsedTime(&ms, c->kev[2 * j], c->kev[2 * j + 1]));
                    c->stats.kernel_ms += ms;
                }
            int mx = 0;
            for (int q = 0; q < nsrc; ++q) mx = std::max(mx, lc[q]);
            last_k_with_change = mx;
            if (mx < k - 1) done = true;          // the last sweep changed nothing
            if (!done && k > maxit)
                return fail(c, FMM3D_E_NOCONV, "solver: phase %d did not converge in %d sweeps", ph, maxit);
        }
        total_sweeps += last_k_with_change + 1;
        c->stats.phase_sweeps[ph] = last_k_with_change + 1;
        if (ph >= 1) {
            std::vector<int> vc(nsrc);
            HIPCHK(c, hipMemcpyAsync(vc.data(), s.valchg, nsrc * sizeof(int), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            int anyv = 0;
            for (int q = 0; q < nsrc; ++q) anyv |= vc[q];
            if (ph >= 2 && !anyv) break;          // C changed nothing: the ordering is consistent
        }
    }
    *sweeps_out = total_sweeps;
    return FMM3D_OK;
}

// Kernel of sweep k: 2 the hyperplane walk, 1 the column walk.  Under the
// hyperplane strategy the first plane_sweeps sweeps of a solve are
// hyperplane sweeps (one update per node, exact along the lateral flow) and
// the column walk (lateral relaxation both ways in every layer) finishes
// what a wave crossing the flow left; the serving engine, whose slots start
// at different sweeps, keeps to one kernel.
static int sweep_kernel(const fmm3d_ctx* c, int k, bool engine)
{
    // the hyperplane walk serves batches of flow-ordered (teleseismic)
    // sources only: with alternating-walk (point) slots it crept down by
    // float32 ulps for thousands of sweeps on a 16-source batch, and the
    // serving engine may admit such slots at any time
    if (c->strategy != 2 || engine || c->walk_alt || c->geo_cap) return 1;
    if (c->plane_sweeps == 0) return 2;
    return (!engine && k <= c->plane_sweeps) ? 2 : 1;
}

template <typename R, bool FAST, bool SHIFT>
static hipError_t launch_column(fmm3d_ctx* c, const ColArgs<R>& a, int blocks, int kind)
{
    if (kind == 2) {
        // diagonal-major grid: every anti-diagonal of every slot, padded
        const int pb = (a.npx + a.npz - 1) * a.nslots * std::min(a.npx, a.npz);
        if (c->plane_ch == 2) k_plane<R, 2, FAST, SHIFT><<<pb, 64, 0, c->stream>>>(a);
        else if (c->plane_ch == 8) k_plane<R, 8, FAST, SHIFT><<<pb, 64, 0, c->stream>>>(a);
        else k_plane<R, 4, FAST, SHIFT><<<pb, 64, 0, c->stream>>>(a);
        return hipGetLastError();
    }
    if (c->strategy == 2) {   // the hyperplane strategy's patches are 8 x 8
        k_column<R, 8, 8, FAST, SHIFT><<<blocks, 64, 0, c->stream>>>(a);
        return hipGetLastError();
    }
    const int px = c->col_px, pz = c->col_pz;
#define FMM3D_COL_CASE(A, B)                                                   \
    if (px == A && pz == B) {                                                  \
        k_column<R, A, B, FAST, SHIFT><<<blocks, 64, 0, c->stream>>>(a);        \
        return hipGetLastError();                                              \
    }
    FMM3D_COL_CASE(16, 16)
    FMM3D_COL_CASE(16, 32)
    FMM3D_COL_CASE(32, 16)
    FMM3D_COL_CASE(32, 32)
    FMM3D_COL_CASE(16, 64)
    FMM3D_COL_CASE(64, 16)
    FMM3D_COL_CASE(8, 64)
    FMM3D_COL_CASE(64, 8)
    FMM3D_COL_CASE(8, 32)
    FMM3D_COL_CASE(32, 8)
    FMM3D_COL_CASE(8, 16)
    FMM3D_COL_CASE(16, 8)
    FMM3D_COL_CASE(8, 8)
#undef FMM3D_COL_CASE
    return hipErrorInvalidValue;
}

// Kernel arguments of the column walk over `nslots` resident slots.
template <typename R>
static int column_args(fmm3d_ctx* c, int nslots, int order, ColArgs<R>& a, int& blocks)
{
    Slots& s = c->sl;
    const Grid& g = c->g;
    const int npx = (g.nnx + patch_x(c) - 1) / patch_x(c), npz = (g.nnz + patch_z(c) - 1) / patch_z(c);
    const int np = npx * npz;
    const size_t nbmax = (size_t)((g.nnr + 3) / 4) * ((g.nnx + 3) / 4) * ((g.nnz + 7) / 8);
    if ((size_t)np > nbmax)
        return fail(c, FMM3D_E_ARG, "solver: %d column patches exceed the per-slot tables (%zu)", np, nbmax);
    a.T = (R*)s.T;
    a.slo = std::is_same<R, double>::value ? (const R*)c->slo64 : (const R*)c->slo32;
    a.hx = std::is_same<R, double>::value ? (const R*)c->hx64 : (const R*)c->hx32;
    a.hz = std::is_same<R, double>::value ? (const R*)c->hz64 : (const R*)c->hz32;
    a.hr = (R)g.dnr;
    a.fixbits = s.fix;
    a.stamp = s.stamp;
    a.lastchg = s.lastchg;
    a.valchg = s.valchg;
    a.slot_on = s.on;
    a.nnr = g.nnr; a.nnx = g.nnx; a.nnz = g.nnz;
    a.npx = npx; a.npz = npz;
    a.nslots = nslots;
    a.fom = order == FMM3D_ORDER_MIXED;
    a.maxpass = c->col_passes;
    a.tol = std::is_same<R, double>::value ? c->tol64 : c->tol32;
    a.visits = c->d_visits;
    a.updates = c->d_visits + 1;
    a.nodes = c->d_visits + 2;
    a.prog = c->ordered ? s.prog : nullptr;
    a.pstride = g.nnr + 16;
    a.ordered_until = c->ordered_until;
    a.walkalt = s.walkalt;
    a.flow = s.flow;
    a.spin_limit = c->spin_limit;
    a.down = 0;
    a.iter = 1;
    a.px = patch_x(c);
    a.pz = patch_z(c);
    a.verify = c->verify;
    a.verify_patches = c->verify_patches < 0 ? !std::is_same<R, double>::value : c->verify_patches;
    a.pend = s.pend;
    a.slotpend = s.slotpend;
    a.wlost = s.wlost;
    a.lguess = c->layer_guess;
    a.vfilter = c->verify_filter;
    a.born = nullptr;
    // float64, and the float32 warm-up solves of refined blocks (float64
    // sweeps follow them): float32 walks count a change only above their
    // settle tolerance (about 1e-6 relative), and windows built on those
    // changes leave sub-tolerance drift behind (up to 2.4 ms at receivers
    // of a 257^3 coarse solve)
    a.lrng = (c->point_window && (std::is_same<R, double>::value || c->geo_cap > 0)) ? s.lrng : nullptr;
    a.ref = std::is_same<R, float>::value ? s.ref : nullptr;
    a.tfloor = c->tol_floor;
    a.qscale = c->ref_qbits > 0 ? std::ldexp(1.0f, c->ref_qbits) : 0.0f;
    a.ptol = (float)c->pass_tol;
    a.geo_n = a.geo_hx = a.geo_hz = 0;
    a.sdims = nullptr;
    if (c->geo_cap > 0) {
        a.sdims = c->d_sdims;
        if (nslots > c->geo_cap) return fail(c, FMM3D_E_STATE, "solver: %d slots on a %d-block context", nslots, c->geo_cap);
        a.slo = std::is_same<R, double>::value ? (const R*)c->gslo64 : (const R*)c->gslo32;
        a.hx = std::is_same<R, double>::value ? (const R*)c->ghx64 : (const R*)c->ghx32;
        a.hz = std::is_same<R, double>::value ? (const R*)c->ghz64 : (const R*)c->ghz32;
        a.geo_n = c->N;
        a.geo_hx = (size_t)g.nnr;
        a.geo_hz = (size_t)g.nnr * g.nnx;
    }
    blocks = 8 * ((nslots * npx + 7) / 8) * npz;   // rows of patches round-robin over 8 XCDs
    return FMM3D_OK;
}

static bool column_supported(int px, int pz)
{
    static const int ok[][2] = { { 16, 16 }, { 16, 32 }, { 32, 16 }, { 32, 32 }, { 16, 64 }, { 64, 16 }, { 8, 64 },
                                 { 64, 8 }, { 8, 32 }, { 32, 8 }, { 8, 16 }, { 16, 8 }, { 8, 8 } };
    for (auto& q : ok)
        if (q[0] == px && q[1] == pz) return true;
    return false;
}

// Queue `n` sweeps starting at sweep k.  A full chunk replays a captured
// hipGraph (captured once per argument set: n column-walk launches, each
// followed by k_next_iter, the sweep number living in c->d_iter; kernel
// nodes only); other counts, and tracing, launch directly.  With timing on,
// each launch (or the whole replay, shared out evenly) is timed with events.
// The fixed-point check kernels after a column-walk launch (they act after
// sweep a.verify only, read from the same sweep counter).
template <typename R, bool FAST, bool SHIFT>
static hipError_t launch_verify(fmm3d_ctx* c, const ColArgs<R>& a)
{
    const size_t n = (size_t)a.nslots * a.nnx * a.nnz;   // one thread per column
    k_verify_mark<R, FAST, SHIFT><<<(int)std::min<size_t>(65536, (n + 255) / 256), 256, 0, c->stream>>>(a);
    const size_t m = (size_t)a.nslots * a.npx * a.npz;
    k_verify_apply<R><<<(int)std::min<size_t>(1024, (m + 255) / 256), 256, 0, c->stream>>>(a, c->sl.on);
    return hipGetLastError();
}

template <typename R, bool FAST, bool SHIFT>
static int launch_chunk(fmm3d_ctx* c, ColArgs<R> a, int blocks, int n, int k, bool engine = false)
{
    if (c->strategy != 2 && !column_supported(c->col_px, c->col_pz))
        return fail(c, FMM3D_E_ARG, "solver: unsupported column patch %dx%d", c->col_px, c->col_pz);
    const int kind = sweep_kernel(c, k, engine);
    // timing: direct launches, each relaxation kernel between its own events
    // the serving engine launches directly: its server runs refined
    // point-source solves on other threads, and HIP refuses their
    // synchronizing calls while any stream of the process captures (relaxed
    // mode or not), which also voids the capture and leaves this stream
    // stuck in it
    if (!c->use_graph || c->timing || engine || n != c->check_every || c->trace ||
        kind != sweep_kernel(c, k + n - 1, engine)) {
        a.iterp = nullptr;
        for (int j = 0; j < n; ++j) {
            a.iter = k + j;
            if (c->timing) HIPCHK(c, hipEventRecord(c->kev[2 * j], c->stream));
            const int kd = sweep_kernel(c, k + j, engine);
            hipError_t le = launch_column<R, FAST, SHIFT>(c, a, blocks, kd);
            HIPCHK(c, le);
            if (c->timing) HIPCHK(c, hipEventRecord(c->kev[2 * j + 1], c->stream));
            // engine slots join at a chunk's first sweep: their checks follow
            // it; batches of alternating walks alone are not checked
            if (kd == 1 && a.verify > 0 && (!engine || j == 0) && (engine || c->walk_alt != 1)) {
                le = launch_verify<R, FAST, SHIFT>(c, a);
                HIPCHK(c, le);
            }
            c->stats.kernel_launches++;
            if (c->trace) {
                unsigned long long v[2];
                HIPCHK(c, hipMemcpyAsync(v, c->d_visits, sizeof v, hipMemcpyDeviceToHost, c->stream));
                HIPCHK(c, hipStreamSynchronize(c->stream));
                float ms = 0.0f;
                if (c->timing) (void)hipEventElapsedTime(&ms, c->kev[2 * j], c->kev[2 * j + 1]);
                fprintf(stderr, "fmm3d trace: sweep %d visits %llu updates %llu kernel %.3f ms\n", k + j, v[0], v[1], ms);
            }
        }
        if (c->timing) {
            HIPCHK(c, hipStreamSynchronize(c->stream));
            for (int j = 0; j < n; ++j) {
                float ms = 0.0f;
                HIPCHK(c, hipEventElapsedTime(&ms, c->kev[2 * j], c->kev[2 * j + 1]));
                c->stats.kernel_ms += ms;
            }
        }
        return FMM3D_OK;
    }
    a.iter = 0;
    a.iterp = c->d_iter;
    std::vector<char> sig(sizeof a + 4 * sizeof(int));
    memcpy(sig.data(), &a, sizeof a);
    // the walk mode decides whether the graph holds the fixed-point check;
    // n is the sweeps the graph holds (a graph of fewer sweeps replayed for
    // a longer chunk would skip sweeps)
    const int tag[4] = { blocks, (int)sizeof(R) * 100 + kind * 10000000 + c->strategy * 1000000 +
                                     c->plane_ch * 100000 + c->col_px * 10000 + c->col_pz, c->walk_alt, n };
    memcpy(sig.data() + sizeof a, tag, sizeof tag);
    if (!c->gexec || sig != c->g_sig) {
        if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
        if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
        // relaxed mode; still, no other thread of this library runs device
        // work meanwhile (refined point-source workers have joined, the
        // serving engine does not capture)
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
        hipError_t e = hipSuccess;
        for (int j = 0; j < n && e == hipSuccess; ++j) {
            e = launch_column<R, FAST, SHIFT>(c, a, blocks, kind);
            if (e == hipSuccess && kind == 1 && a.verify > 0 && c->walk_alt != 1) e = launch_verify<R, FAST, SHIFT>(c, a);
            if (e == hipSuccess) {
                k_next_iter<<<1, 64, 0, c->stream>>>(c->d_iter);
                e = hipGetLastError();
            }
        }
        hipGraph_t gr = nullptr;
        hipError_t e2 = hipStreamEndCapture(c->stream, &gr);
        if (e == hipSuccess) e = e2;
        if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec, gr, nullptr, nullptr, 0);
        c->graph = gr;
        if (e != hipSuccess) {
            if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
            c->gexec = nullptr;
            return fail(c, FMM3D_E_DEVICE, "solver: graph capture: %s", hipGetErrorString(e));
        }
        c->g_sig = sig;
    }
    k_fill_int<<<1, 1, 0, c->stream>>>(c->d_iter, 1, k);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
    c->stats.kernel_launches += n;
    return FMM3D_OK;
}

template <typename R, bool FAST, bool SHIFT>
static int run_column(fmm3d_ctx* c, int nsrc, int order, int* sweeps_out)
{
    Slots& s = c->sl;
    const Grid& g = c->g;
    ColArgs<R> a;
    int blocks = 0;
    int rc0 = column_args<R>(c, nsrc, order, a, blocks);
    if (rc0) return rc0;
    const int np = a.npx * a.npz;
    const bool ordered = a.prog != nullptr;
    if (ordered) {
        k_fill_int<<<(nsrc * np + 255) / 256, 256, 0, c->stream>>>(s.prog, nsrc * np, 0);
        HIPCHK(c, hipGetLastError());
    }
    const int maxit = 4 * (g.nnr + g.nnx + g.nnz) + 200;
    std::vector<int> lc(nsrc);
    k_stamp_init<<<std::min(4096, (blocks + 255) / 256), 256, 0, c->stream>>>(nsrc, np, 1, 1, 1, s.stamp);
    k_lrng_init<<<(nsrc * np + 255) / 256, 256, 0, c->stream>>>(s.lrng, (size_t)nsrc * np, 1);
    k_fill_int<<<(nsrc * np + 255) / 256, 256, 0, c->stream>>>(s.pend, nsrc * np, 0);
    k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.slotpend, nsrc, 0);
    k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.wlost, nsrc, 0);
    k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.lastchg, nsrc, 0);
    k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.on, nsrc, 1);
    k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.valchg, nsrc, 0);
    HIPCHK(c, hipGetLastError());
    int k = 1, last = 0;
    bool done = false;
    c->walk_settled = true;
    std::vector<int> onv(nsrc, 1);
    while (!done) {
        const int chunk = c->max_sweeps > 0 ? std::min(c->check_every, c->max_sweeps - k + 1) : c->check_every;
        int rcl = launch_chunk<R, FAST, SHIFT>(c, a, blocks, chunk, k);
        if (rcl) return rcl;
        k += chunk;
        HIPCHK(c, hipMemcpyAsync(lc.data(), s.lastchg, nsrc * sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int mx = 0;
        bool leave = false;
        for (int q = 0; q < nsrc; ++q) {
            mx = std::max(mx, lc[q]);
            // sweep k-1 changed no node of slot q: it is at its fixed point
            // and leaves the remaining sweeps (its blocks return at once)
            if (onv[q] && lc[q] < k - 1) { onv[q] = 0; leave = true; }
        }
        if (leave && mx >= k - 1) {
            HIPCHK(c, hipMemcpyAsync(s.on, onv.data(), nsrc * sizeof(int), hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (c->trace) {
            fprintf(stderr, "fmm3d trace: lastchg");
            for (int q = 0; q < nsrc; ++q) fprintf(stderr, " %d", lc[q]);
            fprintf(stderr, "\n");
        }
        last = mx;
        if (mx < k - 1) done = true;
        if (!done && c->max_sweeps > 0 && k > c->max_sweeps) { c->walk_settled = false; break; }
        if (!done && k > maxit) return fail(c, FMM3D_E_NOCONV, "solver: column walk did not converge in %d sweeps", maxit);
    }
    c->stats.phase_sweeps[0] = last + 1;
    *sweeps_out = last + 1;
    return FMM3D_OK;
}

// Relax the prepared slots (fields, fixed bits, offsets and flows are set)
// and record the statistics.
static int solve_run_once(fmm3d_ctx* c, int nsrc, int order, int precision);

// ------------------------------------------------------------ autotuner

// Tuning key: device, grid, precision, order, walk mode, batch size class.
static std::string tune_key(const fmm3d_ctx* c, int nsrc, int order, int precision)
{
    int cls = 0;
    while ((1 << cls) < nsrc) ++cls;
    char b[256];
    snprintf(b, sizeof b, "%s %dx%dx%d p%d o%d w%d b%d", c->tune_arch.c_str(), c->g.nnr, c->g.nnx, c->g.nnz,
             precision, order, c->walk_alt, cls);
    return b;
}

// The tuning cache: FMM3D_TUNE_CACHE (empty: none), else
// $XDG_CACHE_HOME/fmm3d/tune.txt, else ~/.cache/fmm3d/tune.txt.  One line per
// choice, "<key> = <strategy> <patch theta> <patch phi> <pass cap> <check
// interval>"; the last line of a key wins.
static std::string tune_cache_path()
{
    const char* p = getenv("FMM3D_TUNE_CACHE");
    if (p) return p;
    const char* x = getenv("XDG_CACHE_HOME");
    if (x && *x) return std::string(x) + "/fmm3d/tune.txt";
    const char* h = getenv("HOME");
    if (h && *h) return std::string(h) + "/.cache/fmm3d/tune.txt";
    return "";
}

static bool tune_choice_ok(const TuneChoice& t)
{
    return t.strategy >= 1 && t.strategy <= 2 && column_supported(t.px, t.pz) && t.passes >= 1 &&
           t.passes <= 16 && t.check >= 1 && t.check <= 8;
}

static bool tune_cache_get(const std::string& key, TuneChoice* out)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return false;
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    const std::string pre = key + " = ";
    while (fgets(line, sizeof line, f)) {
        if (strncmp(line, pre.c_str(), pre.size())) continue;
        TuneChoice t;
        if (sscanf(line + pre.size(), "%d %d %d %d %d", &t.strategy, &t.px, &t.pz, &t.passes, &t.check) == 5 &&
            tune_choice_ok(t)) {
            *out = t;
            found = true;
        }
    }
    fclose(f);
    return found;
}

static void tune_cache_put(const std::string& key, const TuneChoice& t)
{
    const std::string path = tune_cache_path();
    if (path.empty()) return;
    // create the directories on the way (errors surface as a failed open)
    for (size_t i = 1; i < path.size(); ++i)
        if (path[i] == '/') (void)mkdir(path.substr(0, i).c_str(), 0755);
    FILE* f = fopen(path.c_str(), "a");
    if (!f) return;
    fprintf(f, "%s = %d %d %d %d %d\n", key.c_str(), t.strategy, t.px, t.pz, t.passes, t.check);
    fclose(f);
}

static TuneChoice tune_current(const fmm3d_ctx* c)
{
    return { c->strategy, c->col_px, c->col_pz, c->col_passes, c->check_every };
}

static void tune_apply(fmm3d_ctx* c, const TuneChoice& t)
{
    c->strategy = t.strategy; c->col_px = t.px; c->col_pz = t.pz; c->col_passes = t.passes; c->check_every = t.check;
}

// Autotuner.  At the first column-walk solve of a kind (the tuning key) on a
// context whose configuration the caller left alone, the choice comes from
// the tuning cache, or is timed on the real batch: the current configuration
// first, then the column walk's other patch shapes and a pass cap of 8,
// every-2-sweeps convergence checks, and the hyperplane walk (batches without
// alternating walks); each restarts from a copy of the initial fields.  A
// candidate replaces the current configuration only when at least 5 % faster;
// the choice goes to the cache.  Batches under 8 sources or grids under 2 M
// nodes keep the current configuration untimed.
static int autotune(fmm3d_ctx* c, int nsrc, int order, int precision)
{
    const std::string key = tune_key(c, nsrc, order, precision);
    for (auto& e : c->tuned)
        if (e.first == key) {
            tune_apply(c, e.second);
            return FMM3D_OK;
        }
    TuneChoice ch = tune_current(c);
    if (tune_cache_get(key, &ch)) {
        tune_apply(c, ch);
        c->tuned.push_back({ key, ch });
        return FMM3D_OK;
    }
    const TuneChoice def = tune_current(c);
    if (nsrc < 8 || c->N < ((size_t)1 << 21)) {
        c->tuned.push_back({ key, def });
        return FMM3D_OK;
    }
    std::vector<TuneChoice> cand = { def };
    static const int shapes[][3] = { { 8, 8, 16 }, { 8, 8, 8 }, { 8, 16, 16 }, { 16, 8, 16 }, { 16, 16, 16 } };
    for (auto& q : shapes) {
        const TuneChoice t = { 1, q[0], q[1], q[2], def.check };
        if (t.px != def.px || t.pz != def.pz || t.passes != def.passes || def.strategy != 1) cand.push_back(t);
    }
    cand.push_back({ 1, def.px, def.pz, def.passes, def.check == 2 ? 4 : 2 });
    if (c->walk_alt == 0 && c->geo_cap == 0) cand.push_back({ 2, 8, 8, def.passes, def.check });
    Slots& s = c->sl;
    const size_t bytes = (size_t)nsrc * c->N * (precision == FMM3D_PREC_FP64 ? 8 : 4);
    void* save = nullptr;
    HIPCHK(c, dalloc(&save, bytes));
    HIPCHK(c, hipMemcpyAsync(save, s.T, bytes, hipMemcpyDeviceToDevice, c->stream));
    std::vector<double> ms(cand.size(), 1e300);
    int rc = FMM3D_OK;
    const int max_sweeps0 = c->max_sweeps;
    for (size_t i = 0; i < cand.size(); ++i) {
        tune_apply(c, cand[i]);
        hipError_t e = hipMemcpyAsync(s.T, save, bytes, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) { rc = fail(c, FMM3D_E_DEVICE, "autotune: %s", hipGetErrorString(e)); break; }
        const int r = solve_run_once(c, nsrc, order, precision);
        // a candidate that does not settle within the sweep cap is out
        if (r == FMM3D_E_NOCONV || (r == FMM3D_OK && !c->walk_settled)) continue;
        if (r) { rc = r; break; }
        ms[i] = c->stats.solve_ms;
        // the others get at most three times the current configuration's sweeps
        if (i == 0) c->max_sweeps = 3 * c->stats.sweeps + 4;
    }
    c->max_sweeps = max_sweeps0;
    if (rc == FMM3D_OK) {
        hipError_t e = hipMemcpyAsync(s.T, save, bytes, hipMemcpyDeviceToDevice, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) rc = fail(c, FMM3D_E_DEVICE, "autotune: %s", hipGetErrorString(e));
    }
    (void)hipFree(save);
    tune_apply(c, def);
    if (rc) return rc;
    size_t bi = 0;
    for (size_t i = 1; i < cand.size(); ++i)
        if (ms[i] < ms[bi]) bi = i;
    ch = ms[bi] < 0.95 * ms[0] ? cand[bi] : def;
    tune_apply(c, ch);
    c->tuned.push_back({ key, ch });
    tune_cache_put(key, ch);
    if (c->trace) fprintf(stderr, "fmm3d autotune: %s -> %d %d %d %d %d\n", key.c_str(), ch.strategy, ch.px, ch.pz, ch.passes, ch.check);
    return FMM3D_OK;
}

static int solve_run(fmm3d_ctx* c, int nsrc, int order, int precision)
{
    if (c->autotune && !c->tune_fixed && c->strategy >= 1) {
        int rc = autotune(c, nsrc, order, precision);
        if (rc) return rc;
    }
    int rc = solve_run_once(c, nsrc, order, precision);
    if (rc == FMM3D_OK && c->strategy == 2) {
        c->stats.tile_r = c->plane_ch; c->stats.tile_x = 8; c->stats.tile_z = 8;
    } else if (rc == FMM3D_OK && (c->strategy == 1 || c->walk_alt)) {
        c->stats.tile_r = c->col_passes; c->stats.tile_x = c->col_px; c->stats.tile_z = c->col_pz;
    }
    return rc;
}

static int solve_run_once(fmm3d_ctx* c, int nsrc, int order, int precision)
{
    Slots& s = c->sl;
    k_fill_real<double><<<(3 * nsrc + 255) / 256, 256, 0, c->stream>>>(s.srcpos, (size_t)3 * nsrc, __builtin_nan(""));
    if (c->walk_alt == 2 && (int)c->walk_slots.size() == nsrc)
        HIPCHK(c, hipMemcpyAsync(s.walkalt, c->walk_slots.data(), nsrc * sizeof(int), hipMemcpyHostToDevice, c->stream));
    else
        k_fill_int<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(s.walkalt, nsrc, c->walk_alt ? 1 : 0);
    HIPCHK(c, hipGetLastError());
    c->stats = fmm3d_stats{};
    c->stats.tile_r = c->brick_sl; c->stats.tile_x = c->tile_x; c->stats.tile_z = c->tile_z;
    HIPCHK(c, hipMemsetAsync(c->d_visits, 0, 3 * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    int sweeps = 0, rc;
    if (c->strategy >= 1 || c->walk_alt) {
        if (precision == FMM3D_PREC_FP64) rc = run_column<double, false, false>(c, nsrc, order, &sweeps);
        else rc = run_column<float, true, true>(c, nsrc, order, &sweeps);
    } else {
        if (precision == FMM3D_PREC_FP64) rc = run_phases<double, false>(c, nsrc, order, &sweeps);
        else rc = run_phases<float, true>(c, nsrc, order, &sweeps);
    }
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (rc) return rc;
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->stats.solve_ms = ms;
    c->stats.sweeps = sweeps;
    unsigned long long vis[3] = { 0, 0, 0 };
    HIPCHK(c, hipMemcpy(vis, c->d_visits, sizeof vis, hipMemcpyDeviceToHost));
    c->stats.brick_visits = (int64_t)vis[0];
    c->stats.node_updates = (int64_t)vis[1];
    c->stats.node_visits = (int64_t)vis[2];
    s.n = nsrc;
    return FMM3D_OK;
}

// Lateral flow code of a slot from its (theta, phi) index gradient: the sign
// per axis; an axis whose gradient is under a quarter of the other's carries
// +-2 (the hyperplane walk then reverses it on alternate sweeps: the wave
// crosses it both ways in different parts of the model).
static void flow_code(double gx, double gz, int out[2])
{
    out[0] = gx >= 0.0 ? 1 : -1;
    out[1] = gz >= 0.0 ? 1 : -1;
    const double ax = fabs(gx), az = fabs(gz);
    if (ax < 0.25 * az) out[0] *= 2;
    else if (az < 0.25 * ax) out[1] *= 2;
}

// Lateral flow direction of a teleseismic source for the column walk's
// patch order: the signs of the (theta, phi) index gradient of a
// least-squares plane through the bottom face's known times (+1 when the
// times grow with the index, the wave then travels that way).  Only the
// order in which the relaxation visits patches depends on this.
// Least-squares plane t = c0 + gx*ix + gz*iz through a bottom face's known
// (positive) times, fitted on a sub-lattice of at most 64 x 64 nodes; false
// when fewer than three nodes or a degenerate set is known.
static bool face_plane(const Grid& g, const double* face, double* pgx, double* pgz, double* pc0)
{
    const int nx = g.nnx, nz = g.nnz;
    const int sx_ = std::max(1, nx / 64), sz_ = std::max(1, nz / 64);
    double n = 0, sx = 0, sz = 0, sxx = 0, szz = 0, sxz = 0, st = 0, sxt = 0, szt = 0;
    for (int ix = 0; ix < nx; ix += sx_)
        for (int iz = 0; iz < nz; iz += sz_) {
            const double v = face[(size_t)ix * nz + iz];
            if (!(v > 0.0)) continue;
            n += 1; sx += ix; sz += iz; sxx += (double)ix * ix; szz += (double)iz * iz; sxz += (double)ix * iz;
            st += v; sxt += ix * v; szt += iz * v;
        }
    *pgx = *pgz = 0.0;
    *pc0 = n > 0 ? st / n : 0.0;
    if (n < 3) return false;
    const double cxx = sxx - sx * sx / n, czz = szz - sz * sz / n, cxz = sxz - sx * sz / n;
    const double cxt = sxt - sx * st / n, czt = szt - sz * st / n;
    const double det = cxx * czz - cxz * cxz;
    if (!(det > 0.0)) return false;
    *pgx = (cxt * czz - czt * cxz) / det;
    *pgz = (czt * cxx - cxt * cxz) / det;
    *pc0 = (st - *pgx * sx - *pgz * sz) / n;
    return true;
}

static void lateral_flow(const Grid& g, const double* face, int out[2])
{
    double gx, gz, c0;
    face_plane(g, face, &gx, &gz, &c0);
    flow_code(gx, gz, out);
}

// The reference a float32 teleseismic field is stored against: a plane wave
// through the bottom face's fitted plane in the model's mean layer profile,
// A(ir) + gx*ix + gz*iz above the offset c0, where A climbs from 0 at the
// bottom layer by the vertical slowness sqrt(s(ir)^2 - px^2 - pz^2) of the
// layer-mean slowness s and the plane's horizontal slownesses (gx over the
// theta spacing, gz over the phi spacing of the middle column), trapezoid
// rule over dnr.  Each step and gx, gz are multiples of 2^-20 s, so every
// difference of the reference between two nodes of a stencil is exact in
// float32 and A is exact in float64 (steps under 4 s, |gx|, |gz| under 4 s).
// Writes ref[0..nnr+1] = A, gx, gz and *off = c0; false (ref zero, *off
// untouched) without a fitted plane.
static bool tele_reference(const fmm3d_ctx* c, const double* face, double* ref, double* off)
{
    const Grid& g = c->g;
    std::fill(ref, ref + g.nnr + 2, 0.0);
    if ((int)c->layer_slo.size() != g.nnr) return false;
    double gx, gz, c0;
    if (!face_plane(g, face, &gx, &gz, &c0)) return false;
    auto quant = [](double v) { return std::ldexp(std::nearbyint(std::ldexp(v, 20)), -20); };
    gx = quant(gx);
    gz = quant(gz);
    if (!(fabs(gx) < 4.0 && fabs(gz) < 4.0)) return false;
    const double sth = sin(g.rgox + (g.nnx / 2) * g.dnx);
    auto qv = [&](int ir) {
        const double ri = g.rgor - ir * g.dnr + g.earth;
        const double px = gx / (ri * g.dnx), pz = gz / (ri * sth * g.dnz), s = c->layer_slo[ir];
        return sqrt(std::max(s * s - px * px - pz * pz, 0.0));
    };
    double q1 = qv(g.nnr - 1);
    for (int ir = g.nnr - 2; ir >= 0; --ir) {
        const double q0 = qv(ir), step = quant(0.5 * g.dnr * (q0 + q1));
        if (!(step < 4.0)) {   // a stencil's two-layer difference would not be exact in float32
            std::fill(ref, ref + g.nnr + 2, 0.0);
            return false;
        }
        ref[ir] = ref[ir + 1] + step;
        q1 = q0;
    }
    ref[g.nnr] = gx;
    ref[g.nnr + 1] = gz;
    *off = c0;
    return true;
}

// Argument checks of teleseismic face arrays.  side[2s], side[2s+1]: -1
// none, 0 the low-index face, 1 the high-index one.
static int tele_check(fmm3d_ctx* c, const char* fn, int nsrc, const double* faces, const double* xf,
                      const double* zf, const int* side)
{
    if (nsrc <= 0 || !faces) return fail(c, FMM3D_E_ARG, "%s: need nsrc > 0 and faces", fn);
    if (xf && (!zf || !side)) return fail(c, FMM3D_E_ARG, "%s: need both side faces and the side indices", fn);
    if (xf)
        for (int s = 0; s < nsrc; ++s)
            if (side[2 * s] < -1 || side[2 * s] > 1 || side[2 * s + 1] < -1 || side[2 * s + 1] > 1)
                return fail(c, FMM3D_E_ARG, "%s: side indices of source %d must be -1, 0 or 1", fn, s);
    return FMM3D_OK;
}

// A staged source's offset: the reference's where a plane was fitted,
// else the boundary minimum.
__global__ void k_pick_offsets(int n, const double* __restrict__ off, double* __restrict__ toff)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && off[i] == off[i]) toff[i] = off[i];
}

// Stage teleseismic sources: their faces in device memory (d_faces,
// d_sides), and per source the boundary minimum and NaN check (on the
// device), the lateral flow and the plane-wave reference row and offset
// (fitted on the host from the face).  Solves start from the staged data
// (tele_init); fmm3d_stage_tele_sides stages them ahead of the solves.
static int tele_stage(fmm3d_ctx* c, const char* fn, int nsrc, const double* faces, const double* xf,
                      const double* zf, const int* side)
{
    const Grid& g = c->g;
    const size_t F = (size_t)g.nnx * g.nnz, XF = (size_t)g.nnr * g.nnz, ZF = (size_t)g.nnr * g.nnx;
    const size_t RS = (size_t)g.nnr + 2;
    c->stg_n = 0;
    if (c->d_faces_n < (size_t)nsrc * F) {
        (void)hipFree(c->d_faces);
        c->d_faces = nullptr;
        c->d_faces_n = 0;
        HIPCHK(c, dalloc(&c->d_faces, (size_t)nsrc * F * sizeof(double)));
        c->d_faces_n = (size_t)nsrc * F;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_faces, faces, (size_t)nsrc * F * sizeof(double), hipMemcpyHostToDevice, c->stream));
    double* dx = nullptr;
    double* dz = nullptr;
    if (xf) {
        const size_t nside = (size_t)nsrc * (XF + ZF);
        if (c->d_sides_n < nside) {
            (void)hipFree(c->d_sides);
            c->d_sides = nullptr;
            c->d_sides_n = 0;
            HIPCHK(c, dalloc(&c->d_sides, nside * sizeof(double)));
            c->d_sides_n = nside;
        }
        if (c->d_side_of_n < (size_t)nsrc * 2) {
            (void)hipFree(c->d_side_of);
            c->d_side_of = nullptr;
            c->d_side_of_n = 0;
            HIPCHK(c, dalloc(&c->d_side_of, (size_t)nsrc * 2 * sizeof(int)));
            c->d_side_of_n = (size_t)nsrc * 2;
        }
        dx = c->d_sides;
        dz = c->d_sides + (size_t)nsrc * XF;
        HIPCHK(c, hipMemcpyAsync(dx, xf, (size_t)nsrc * XF * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(dz, zf, (size_t)nsrc * ZF * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_side_of, side, (size_t)nsrc * 2 * sizeof(int), hipMemcpyHostToDevice, c->stream));
    }
    if (c->stg_cap < (size_t)nsrc) {
        (void)hipFree(c->d_stg_toff); (void)hipFree(c->d_stg_ref); (void)hipFree(c->d_stg_off); (void)hipFree(c->d_stg_flow);
        c->d_stg_toff = c->d_stg_ref = c->d_stg_off = nullptr;
        c->d_stg_flow = nullptr;
        c->stg_cap = 0;
        HIPCHK(c, dalloc(&c->d_stg_toff, (size_t)nsrc * sizeof(double)));
        HIPCHK(c, dalloc(&c->d_stg_ref, (size_t)nsrc * RS * sizeof(double)));
        HIPCHK(c, dalloc(&c->d_stg_off, (size_t)nsrc * sizeof(double)));
        HIPCHK(c, dalloc(&c->d_stg_flow, (size_t)nsrc * 2 * sizeof(int)));
        c->stg_cap = nsrc;
    }
    // the boundary minima and the NaN check, on the device
    if (!c->d_bad) HIPCHK(c, dalloc(&c->d_bad, sizeof(int)));
    k_fill_int<<<1, 1, 0, c->stream>>>(c->d_bad, 1, nsrc);
    k_tele_offsets<<<nsrc, 1024, 0, c->stream>>>(F, XF, ZF, c->d_faces, dx, dz, xf ? c->d_side_of : nullptr, c->d_stg_toff, c->d_bad);
    int bad = nsrc;
    HIPCHK(c, hipMemcpyAsync(&bad, c->d_bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    // flows and reference rows from the faces, on the host, while those run
    std::vector<int> flow((size_t)nsrc * 2);
    std::vector<double> ref((size_t)nsrc * RS, 0.0), off((size_t)nsrc, __builtin_nan(""));
    for (int q = 0; q < nsrc; ++q) {
        lateral_flow(g, faces + (size_t)q * F, &flow[2 * (size_t)q]);
        if (c->tele_ref) tele_reference(c, faces + (size_t)q * F, &ref[(size_t)q * RS], &off[q]);
    }
    HIPCHK(c, hipMemcpyAsync(c->d_stg_flow, flow.data(), flow.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_stg_ref, ref.data(), ref.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_stg_off, off.data(), off.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    // the host vectors are pageable memory: the copies finish before they go
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (bad < nsrc) return fail(c, FMM3D_E_ARG, "%s: NaN boundary time (source %d)", fn, bad);
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(c, FMM3D_E_DEVICE, "stage: %s", hipGetErrorString(le));
    c->stg_n = nsrc;
    c->stg_sides = xf != nullptr;
    return FMM3D_OK;
}

// Initial fields, fixed bits, offsets, reference rows and lateral flows of
// slots 0 .. nsrc-1 from the staged sources (slots must be allocated); all
// on the device.
static int tele_init(fmm3d_ctx* c, int nsrc, int precision)
{
    const Grid& g = c->g;
    Slots& s = c->sl;
    const size_t RS = (size_t)g.nnr + 2;
    // float32 fields on the column walk (the kernels that apply the
    // reference) are stored against each slot's plane-wave reference
    const bool useref = precision == FMM3D_PREC_FP32 && c->tele_ref && (c->strategy == 1 || (c->strategy == 0 && c->walk_alt));
    HIPCHK(c, hipMemcpyAsync(s.toff, c->d_stg_toff, nsrc * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (useref) {
        k_pick_offsets<<<(nsrc + 255) / 256, 256, 0, c->stream>>>(nsrc, c->d_stg_off, s.toff);
        HIPCHK(c, hipMemcpyAsync(s.ref, c->d_stg_ref, (size_t)nsrc * RS * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    } else {
        k_fill_real<double><<<(int)std::min<size_t>(1024, ((size_t)nsrc * RS + 255) / 256), 256, 0, c->stream>>>(s.ref, (size_t)nsrc * RS, 0.0);
    }
    HIPCHK(c, hipMemcpyAsync(s.flow, c->d_stg_flow, (size_t)nsrc * 2 * sizeof(int), hipMemcpyDeviceToDevice, c->stream));
    const double* dfaces = c->d_faces;
    if (precision == FMM3D_PREC_FP64)
        k_init_tele<double><<<dim3(256, std::min(nsrc, 65535)), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, (double*)s.T, s.toff, nullptr);
    else
        k_init_tele<float><<<dim3(256, std::min(nsrc, 65535)), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, (float*)s.T, s.toff, s.ref);
    k_fixbits_tele<<<dim3(64, std::min(nsrc, 65535)), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dfaces, s.fix);
    if (c->stg_sides) {
        const double* dx = c->d_sides;
        const double* dz = c->d_sides + (size_t)nsrc * g.nnr * g.nnz;
        if (precision == FMM3D_PREC_FP64)
            k_init_sides<double><<<4096, 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dx, dz, c->d_side_of, s.toff, nullptr, (double*)s.T, s.fix);
        else
            k_init_sides<float><<<4096, 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, nsrc, dx, dz, c->d_side_of, s.toff, s.ref, (float*)s.T, s.fix);
    }
    hipError_t le = hipGetLastError();
    if (le != hipSuccess) return fail(c, FMM3D_E_DEVICE, "init: %s", hipGetErrorString(le));
    return FMM3D_OK;
}

// Stage slots 0 .. nsrc-1 from teleseismic faces and initialize them.
static int tele_upload(fmm3d_ctx* c, const char* fn, int nsrc, const double* faces, const double* xf,
                       const double* zf, const int* side, int precision)
{
    int rc = tele_stage(c, fn, nsrc, faces, xf, zf, side);
    if (rc) return rc;
    return tele_init(c, nsrc, precision);
}

#include "fmm3d_partition.inl"

// The reference-partition teleseismic solve (bottom faces, float64).
static int solve_tele_part(fmm3d_ctx* c, const char* fn, int nsrc, const double* faces, int order)
{
    const Grid& g = c->g;
    const size_t F = (size_t)g.nnx * g.nnz;
    for (size_t i = 0; i < (size_t)nsrc * F; ++i)
        if (isnan(faces[i])) return fail(c, FMM3D_E_ARG, "%s: NaN boundary time (source %zu)", fn, i / F);
    int rc = ensure_slots(c, nsrc, FMM3D_PREC_FP64);
    if (rc) return rc;
    c->sl.n = 0;
    if (c->d_faces_n < (size_t)nsrc * F) {
        (void)hipFree(c->d_faces);
        c->d_faces = nullptr;
        c->d_faces_n = 0;
        HIPCHK(c, dalloc(&c->d_faces, (size_t)nsrc * F * sizeof(double)));
        c->d_faces_n = (size_t)nsrc * F;
    }
    c->stg_n = 0;   // the staged faces are replaced
    HIPCHK(c, hipMemcpyAsync(c->d_faces, faces, (size_t)nsrc * F * sizeof(double), hipMemcpyHostToDevice, c->stream));
    c->stats = fmm3d_stats{};
    HIPCHK(c, hipMemsetAsync(c->d_visits, 0, 3 * sizeof(unsigned long long), c->stream));
    HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    rc = part_run(c, nsrc, faces, order);
    if (rc) return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->stats.solve_ms = ms;
    unsigned long long vis[3] = { 0, 0, 0 };
    HIPCHK(c, hipMemcpy(vis, c->d_visits, sizeof vis, hipMemcpyDeviceToHost));
    c->stats.node_updates = (int64_t)vis[1];
    c->stats.tile_r = 1; c->stats.tile_x = PT_P; c->stats.tile_z = PT_P;
    c->sl.n = nsrc;
    c->part_last = true;
    return FMM3D_OK;
}

// Teleseismic solves from bottom faces and, when xf is given, one theta
// and one phi side face per source.
static int solve_tele_impl(fmm3d_ctx* c, const char* fn, int nsrc, const double* faces, const double* xf,
                           const double* zf, const int* side, int order, int precision, int flags)
{
    if (!c) return fail(c, FMM3D_E_ARG, "%s: null context", fn);
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "%s: bad order %d", fn, order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "%s: bad precision %d", fn, precision);
    if (flags != 0 && flags != FMM3D_REFERENCE_PARTITION)
        return fail(c, FMM3D_E_ARG, "%s: flags must be 0 or FMM3D_REFERENCE_PARTITION", fn);
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "%s: no velocity model (call fmm3d_set_model_vtx)", fn);
    int rc = tele_check(c, fn, nsrc, faces, xf, zf, side);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    if (flags == FMM3D_REFERENCE_PARTITION) {
        if (precision != FMM3D_PREC_FP64)
            return fail(c, FMM3D_E_ARG, "%s: FMM3D_REFERENCE_PARTITION solves in float64 only", fn);
        if (xf) return fail(c, FMM3D_E_ARG, "%s: FMM3D_REFERENCE_PARTITION takes bottom faces only", fn);
        c->walk_alt = 0;
        return solve_tele_part(c, fn, nsrc, faces, order);
    }
    rc = ensure_slots(c, nsrc, precision);
    if (rc) return rc;
    c->sl.n = 0;
    c->walk_alt = 0;
    rc = tele_upload(c, fn, nsrc, faces, xf, zf, side, precision);
    if (rc) return rc;
    return solve_run(c, nsrc, order, precision);
}

extern "C" int fmm3d_solve_tele(fmm3d_ctx* c, int nsrc, const double* faces, int order, int precision, int flags)
{
    return solve_tele_impl(c, "fmm3d_solve_tele", nsrc, faces, nullptr, nullptr, nullptr, order, precision, flags);
}

extern "C" int fmm3d_solve_tele_sides(fmm3d_ctx* c, int nsrc, const double* faces, const double* xface,
                                      const double* zface, const int* side, int order, int precision, int flags)
{
    return solve_tele_impl(c, "fmm3d_solve_tele_sides", nsrc, faces, xface, zface, side, order, precision, flags);
}

extern "C" int fmm3d_stage_tele_sides(fmm3d_ctx* c, int nsrc, const double* faces, const double* xface,
                                      const double* zface, const int* side)
{
    const char* fn = "fmm3d_stage_tele_sides";
    if (!c) return fail(c, FMM3D_E_ARG, "%s: null context", fn);
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "%s: no velocity model (call fmm3d_set_model_vtx)", fn);
    int rc = tele_check(c, fn, nsrc, faces, xface, zface, side);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    return tele_stage(c, fn, nsrc, faces, xface, zface, side);
}

extern "C" int fmm3d_solve_tele_staged(fmm3d_ctx* c, int order, int precision)
{
    const char* fn = "fmm3d_solve_tele_staged";
    if (!c) return fail(c, FMM3D_E_ARG, "%s: null context", fn);
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "%s: bad order %d", fn, order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "%s: bad precision %d", fn, precision);
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "%s: no velocity model (call fmm3d_set_model_vtx)", fn);
    if (c->stg_n <= 0) return fail(c, FMM3D_E_STATE, "%s: no staged sources (call fmm3d_stage_tele_sides)", fn);
    HIPCHK(c, hipSetDevice(c->device));
    const int nsrc = c->stg_n;
    int rc = ensure_slots(c, nsrc, precision);
    if (rc) return rc;
    c->sl.n = 0;
    c->walk_alt = 0;
    rc = tele_init(c, nsrc, precision);
    if (rc) return rc;
    return solve_run(c, nsrc, order, precision);
}

// Lateral flow signs for a general fixed set: the (theta, phi) slopes of a
// least-squares fit t = a + gx*ix + gz*iz + gr*ir through the fixed times
// over their node positions.  The depth term keeps the strong vertical
// trend of side faces (times growing upward along the whole face) out of
// the lateral slopes; a set within one layer fits (theta, phi) alone.
static void fixed_flow(const Grid& g, const int64_t* node, const double* tv, int64_t n, int out[2])
{
    const int64_t NL = (int64_t)g.nnx * g.nnz;
    double m = 0, s[3] = { 0, 0, 0 }, st = 0;
    for (int64_t i = 0; i < n; ++i) {
        const int64_t f = node[i] % NL;
        const double p[3] = { (double)(f / g.nnz), (double)(f % g.nnz), (double)(node[i] / NL) };
        m += 1; st += tv[i];
        for (int a = 0; a < 3; ++a) s[a] += p[a];
    }
    double gx = 0.0, gz = 0.0;
    if (m >= 3) {
        double C[3][3] = {}, b[3] = {};
        const double mu[3] = { s[0] / m, s[1] / m, s[2] / m }, mt = st / m;
        for (int64_t i = 0; i < n; ++i) {
            const int64_t f = node[i] % NL;
            const double d[3] = { (double)(f / g.nnz) - mu[0], (double)(f % g.nnz) - mu[1], (double)(node[i] / NL) - mu[2] };
            const double dt = tv[i] - mt;
            for (int a = 0; a < 3; ++a) {
                b[a] += d[a] * dt;
                for (int c2 = 0; c2 < 3; ++c2) C[a][c2] += d[a] * d[c2];
            }
        }
        const double det3 = C[0][0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) - C[0][1] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                            C[0][2] * (C[1][0] * C[2][1] - C[1][1] * C[2][0]);
        const double det2 = C[0][0] * C[1][1] - C[0][1] * C[1][0];
        if (C[2][2] > 0.0 && fabs(det3) > 1e-12 * fabs(C[0][0] * C[1][1] * C[2][2])) {
            // Cramer's rule on the 3 x 3 normal equations
            gx = (b[0] * (C[1][1] * C[2][2] - C[1][2] * C[2][1]) - C[0][1] * (b[1] * C[2][2] - C[1][2] * b[2]) +
                  C[0][2] * (b[1] * C[2][1] - C[1][1] * b[2])) / det3;
            gz = (C[0][0] * (b[1] * C[2][2] - C[1][2] * b[2]) - b[0] * (C[1][0] * C[2][2] - C[1][2] * C[2][0]) +
                  C[0][2] * (C[1][0] * b[2] - b[1] * C[2][0])) / det3;
        } else if (det2 > 0.0) {
            gx = (b[0] * C[1][1] - b[1] * C[0][1]) / det2;
            gz = (b[1] * C[0][0] - b[0] * C[1][0]) / det2;
        }
    }
    flow_code(gx, gz, out);
}

// Flow entries of a slot whose walk alternates: the (theta, phi) index of
// its earliest fixed node (a point source's own cell) plus FLOW_CENTER.
static void centre_flow(const Grid& g, const int64_t* node, const double* tv, int64_t n, int out[2])
{
    int64_t best = -1;
    double m = __builtin_huge_val();
    for (int64_t i = 0; i < n; ++i)
        if (tv[i] < m) { m = tv[i]; best = node[i]; }
    if (best < 0) { out[0] = out[1] = 1; return; }
    const int64_t f = best % ((int64_t)g.nnx * g.nnz);
    out[0] = FLOW_CENTER + (int)(f / g.nnz);
    out[1] = FLOW_CENTER + (int)(f % g.nnz);
}

// Argument checks of fixed node sets, and per source the float32 offset
// (the smallest fixed time) and the lateral flow code (centred: the
// earliest fixed node, for alternating walks).
static int fixed_prepare(fmm3d_ctx* c, const char* fn, int nsrc, const int64_t* offs, const int64_t* node,
                         const double* tv, std::vector<double>& toff, std::vector<int>& flow, bool centred)
{
    if (nsrc <= 0 || !offs) return fail(c, FMM3D_E_ARG, "%s: need nsrc > 0 and offsets", fn);
    if (offs[0] != 0) return fail(c, FMM3D_E_ARG, "%s: offs[0] must be 0", fn);
    if (offs[nsrc] > 0 && (!node || !tv)) return fail(c, FMM3D_E_ARG, "%s: null node or time array", fn);
    toff.assign(nsrc, 0.0);
    flow.assign((size_t)nsrc * 2, 1);
    const Grid& g = c->g;
    for (int q = 0; q < nsrc; ++q) {
        if (offs[q + 1] < offs[q]) return fail(c, FMM3D_E_ARG, "%s: offsets decrease at source %d", fn, q);
        double m = __builtin_huge_val();
        for (int64_t i = offs[q]; i < offs[q + 1]; ++i) {
            if (node[i] < 0 || (size_t)node[i] >= c->N)
                return fail(c, FMM3D_E_ARG, "%s: node %lld out of range (source %d)", fn, (long long)node[i], q);
            if (!(tv[i] >= 0.0) || !isfinite(tv[i]))
                return fail(c, FMM3D_E_ARG, "%s: time %g at node %lld must be finite and >= 0", fn, tv[i], (long long)node[i]);
            m = std::min(m, tv[i]);
        }
        toff[q] = isfinite(m) ? m : 0.0;
        if (centred) centre_flow(g, node + offs[q], tv + offs[q], offs[q + 1] - offs[q], &flow[2 * q]);
        else fixed_flow(g, node + offs[q], tv + offs[q], offs[q + 1] - offs[q], &flow[2 * q]);
    }
    return FMM3D_OK;
}

// Initial fields, fixed bits, offsets and flows of slots slot0 ..
// slot0+nsrc-1 from fixed node sets (slots must be allocated).
static int fixed_upload(fmm3d_ctx* c, const char* fn, int slot0, int nsrc, const int64_t* offs, const int64_t* node,
                        const double* tv, int precision, const std::vector<double>& toff, const std::vector<int>& flow)
{
    Slots& s = c->sl;
    const int64_t total = offs[nsrc];
    const size_t N = c->N, NW = (N + 31) / 32;
    const size_t es = precision == FMM3D_PREC_FP64 ? sizeof(double) : sizeof(float);
    void* T0 = (char*)s.T + (size_t)slot0 * N * es;
    uint32_t* fix0 = s.fix + (size_t)slot0 * NW;
    double* toff0 = s.toff + slot0;
    if (c->st_offs_n < (size_t)nsrc + 1) {
        (void)hipFree(c->st_offs);
        c->st_offs = nullptr;
        c->st_offs_n = 0;
        HIPCHK(c, dalloc(&c->st_offs, (size_t)(nsrc + 1) * sizeof(int64_t)));
        c->st_offs_n = (size_t)nsrc + 1;
    }
    if (total > 0 && c->st_n < (size_t)total) {
        (void)hipFree(c->st_node); (void)hipFree(c->st_tv);
        c->st_node = nullptr; c->st_tv = nullptr;
        c->st_n = 0;
        HIPCHK(c, dalloc(&c->st_node, (size_t)total * sizeof(int64_t)));
        HIPCHK(c, dalloc(&c->st_tv, (size_t)total * sizeof(double)));
        c->st_n = (size_t)total;
    }
    int64_t* doffs = c->st_offs;
    int64_t* dnode = c->st_node;
    double* dtv = c->st_tv;
    hipError_t e = hipMemcpyAsync(doffs, offs, (size_t)(nsrc + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && total > 0) e = hipMemcpyAsync(dnode, node, (size_t)total * sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && total > 0) e = hipMemcpyAsync(dtv, tv, (size_t)total * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(toff0, toff.data(), nsrc * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.flow + 2 * slot0, flow.data(), flow.size() * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const size_t n = (size_t)nsrc * N, nw = (size_t)nsrc * NW, RS = (size_t)c->g.nnr + 2;
        // fixed sets keep the plain offset (no reference)
        k_fill_real<double><<<(int)std::min<size_t>(4096, (nsrc * RS + 255) / 256), 256, 0, c->stream>>>(s.ref + (size_t)slot0 * RS, nsrc * RS, 0.0);
        if (precision == FMM3D_PREC_FP64) k_fill_real<double><<<4096, 256, 0, c->stream>>>((double*)T0, n, Lim<double>::inf());
        else k_fill_real<float><<<4096, 256, 0, c->stream>>>((float*)T0, n, Lim<float>::inf());
        k_fill_u32<<<4096, 256, 0, c->stream>>>(fix0, nw);
        if (total > 0) {
            const int nb = (int)std::min<int64_t>(4096, (total + 255) / 256);
            if (precision == FMM3D_PREC_FP64)
                k_scatter_fixed<double><<<nb, 256, 0, c->stream>>>(nsrc, N, doffs, dnode, dtv, toff0, (double*)T0, fix0);
            else
                k_scatter_fixed<float><<<nb, 256, 0, c->stream>>>(nsrc, N, doffs, dnode, dtv, toff0, (float*)T0, fix0);
        }
        e = hipGetLastError();
    }
    // the host vectors must outlive the copies
    const hipError_t es2 = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = es2;
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "%s: %s", fn, hipGetErrorString(e));
    return FMM3D_OK;
}

extern "C" int fmm3d_solve_fixed(fmm3d_ctx* c, int nsrc, const int64_t* offs, const int64_t* node, const double* tv,
                                 int order, int precision, int flags)
{
    const char* fn = "fmm3d_solve_fixed";
    if (!c) return fail(c, FMM3D_E_ARG, "%s: null context", fn);
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "%s: bad order %d", fn, order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "%s: bad precision %d", fn, precision);
    if (flags & ~FMM3D_WALK_ALTERNATE) return fail(c, FMM3D_E_ARG, "%s: unknown flags 0x%x", fn, flags);
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "%s: no velocity model (call fmm3d_set_model_vtx)", fn);
    std::vector<double> toff;
    std::vector<int> flow;
    int rc = fixed_prepare(c, fn, nsrc, offs, node, tv, toff, flow, (flags & FMM3D_WALK_ALTERNATE) && c->point_order);
    if (rc) return rc;
    HIPCHK(c, hipSetDevice(c->device));
    rc = ensure_slots(c, nsrc, precision);
    if (rc) return rc;
    c->sl.n = 0;
    c->walk_alt = (flags & FMM3D_WALK_ALTERNATE) ? 1 : 0;
    rc = fixed_upload(c, fn, 0, nsrc, offs, node, tv, precision, toff, flow);
    if (rc) return rc;
    return solve_run(c, nsrc, order, precision);
}

// B-spline velocity at one point, in the gridder's arithmetic (k_gridder);
// hr is its (gor - h) expression, spelled as the caller's grid spells it.
static double host_velocity(const Grid& g, const double* velv, double h, double hr, double th, double ph)
{
    const int NR = g.nvr + 2, NX = g.nvx + 2, NZ = g.nvz + 2;
    const int ri = (int)floor(hr / g.dvr) + 1;
    double wr[4], wx[4], wz[4];
    bspline4(((g.gor - (ri - 1) * g.dvr) - h) / g.dvr, wr);
    const int xi = (int)floor((th - g.gox) / g.dvx) + 1;
    bspline4((th - g.gox - (xi - 1) * g.dvx) / g.dvx, wx);
    const int zi = (int)floor((ph - g.goz) / g.dvz) + 1;
    bspline4((ph - g.goz - (zi - 1) * g.dvz) / g.dvz, wz);
    double acc_z = 0.0;
    for (int a = 0; a < 4; ++a) {
        const int iz = std::min(std::max(zi - 1 + a, 0), NZ - 1);
        double acc_x = 0.0;
        for (int b = 0; b < 4; ++b) {
            const int ix = std::min(std::max(xi - 1 + b, 0), NX - 1);
            const double* col = velv + ((size_t)iz * NX + ix) * NR;
            double acc_r = 0.0;
            for (int cc = 0; cc < 4; ++cc) acc_r += wr[cc] * col[std::min(std::max(ri - 1 + cc, 0), NR - 1)];
            acc_x += wx[b] * acc_r;
        }
        acc_z += wz[a] * acc_x;
    }
    return acc_z;
}

static void cart3(const Grid& g, double h, double th, double ph, double x[3])
{
    const double r = g.earth + h;
    x[0] = r * sin(th) * cos(ph);
    x[1] = r * sin(th) * sin(ph);
    x[2] = r * cos(th);
}

// Coarse nodes lo..hi (per axis r, theta, phi) of a point source's refined
// block: the cell holding it widened by `radius` nodes, clipped to the grid.
// False when the source lies outside the grid.
static bool point_block(const Grid& g, const double* src3, int radius, int lo[3], int hi[3])
{
    const int nn[3] = { g.nnr, g.nnx, g.nnz };
    const double u[3] = { (g.rgor - src3[0]) / g.dnr, (src3[1] - g.rgox) / g.dnx, (src3[2] - g.rgoz) / g.dnz };
    for (int a = 0; a < 3; ++a) {
        if (!(u[a] >= 0.0 && u[a] <= nn[a] - 1)) return false;
        const int cc = std::min((int)floor(u[a]), nn[a] - 2);
        lo[a] = std::max(0, cc - radius);
        hi[a] = std::min(nn[a] - 1, cc + 1 + radius);
    }
    return true;
}

// Spacings of grid f (the reference's metric arrays): per layer r dtheta,
// per layer and theta row r sin(theta) dphi.
static void grid_metrics(const Grid& g, std::vector<double>& hx, std::vector<double>& hz)
{
    hx.resize(g.nnr);
    hz.resize((size_t)g.nnx * g.nnr);
    for (int ir = 0; ir < g.nnr; ++ir) {
        const double ri = g.rgor - (ir) * g.dnr + g.earth;
        hx[ir] = ri * g.dnx;
        for (int ix = 0; ix < g.nnx; ++ix) hz[(size_t)ir * g.nnx + ix] = ri * sin(g.rgox + (ix) * g.dnx) * g.dnz;
    }
}

// Slot q of a refined-block context on block grid f (at most the
// context's grid in every dimension, laid out with the context's strides):
// its spacings, and the B-spline slowness of f's nodes from the context's
// vertex model.
static int geo_slot(fmm3d_ctx* fc, int q, const Grid& f)
{
    const Grid& cg = fc->g;
    std::vector<double> hx, hz;
    grid_metrics(f, hx, hz);
    // container rows; entries beyond f are never read (the kernels stop at f's edges)
    std::vector<double> chx(cg.nnr, 1.0), chz((size_t)cg.nnr * cg.nnx, 1.0);
    for (int ir = 0; ir < f.nnr; ++ir) {
        chx[ir] = hx[ir];
        for (int ix = 0; ix < f.nnx; ++ix) chz[(size_t)ir * cg.nnx + ix] = hz[(size_t)ir * f.nnx + ix];
    }
    std::vector<float> chx32(chx.begin(), chx.end()), chz32(chz.begin(), chz.end());
    const size_t N = fc->N;
    const int dims[3] = { f.nnr, f.nnx, f.nnz };
    HIPCHK(fc, hipMemcpyAsync(fc->ghx64 + (size_t)q * chx.size(), chx.data(), chx.size() * sizeof(double), hipMemcpyHostToDevice, fc->stream));
    HIPCHK(fc, hipMemcpyAsync(fc->ghz64 + (size_t)q * chz.size(), chz.data(), chz.size() * sizeof(double), hipMemcpyHostToDevice, fc->stream));
    HIPCHK(fc, hipMemcpyAsync(fc->ghx32 + (size_t)q * chx32.size(), chx32.data(), chx32.size() * sizeof(float), hipMemcpyHostToDevice, fc->stream));
    HIPCHK(fc, hipMemcpyAsync(fc->ghz32 + (size_t)q * chz32.size(), chz32.data(), chz32.size() * sizeof(float), hipMemcpyHostToDevice, fc->stream));
    HIPCHK(fc, hipMemcpyAsync(fc->d_sdims + 3 * q, dims, sizeof dims, hipMemcpyHostToDevice, fc->stream));
    const long total = (long)(f.nnr + 3) * (f.nnx + 3) * (f.nnz + 3);
    const int blocks = (int)std::min<long>((total + 255) / 256, 65536);
    k_gridder<<<blocks, 256, 0, fc->stream>>>(f, fc->d_velv, fc->veln, fc->gslo64 + (size_t)q * N, fc->gslo32 + (size_t)q * N,
                                              cg.nnx, cg.nnz);
    HIPCHK(fc, hipGetLastError());
    // the host vectors are pageable memory: the copies finish before they go
    HIPCHK(fc, hipStreamSynchronize(fc->stream));
    return FMM3D_OK;
}

// Per-slot geometry arrays for k blocks (grown, never shrunk).
static int geo_reserve(fmm3d_ctx* fc, int k)
{
    if (fc->geo_cap >= k) return FMM3D_OK;
    HIPCHK(fc, hipStreamSynchronize(fc->stream));
    (void)hipFree(fc->gslo64); (void)hipFree(fc->gslo32);
    (void)hipFree(fc->ghx64); (void)hipFree(fc->ghz64); (void)hipFree(fc->ghx32); (void)hipFree(fc->ghz32);
    (void)hipFree(fc->d_sdims);
    fc->gslo64 = nullptr; fc->gslo32 = nullptr;
    fc->ghx64 = fc->ghz64 = nullptr; fc->ghx32 = fc->ghz32 = nullptr;
    fc->d_sdims = nullptr;
    fc->geo_cap = 0;
    const Grid& f = fc->g;
    const size_t nhz = (size_t)f.nnr * f.nnx;
    HIPCHK(fc, dalloc(&fc->gslo64, (size_t)k * fc->N * sizeof(double)));
    HIPCHK(fc, dalloc(&fc->gslo32, (size_t)k * fc->N * sizeof(float)));
    HIPCHK(fc, dalloc(&fc->ghx64, (size_t)k * f.nnr * sizeof(double)));
    HIPCHK(fc, dalloc(&fc->ghz64, (size_t)k * nhz * sizeof(double)));
    HIPCHK(fc, dalloc(&fc->ghx32, (size_t)k * f.nnr * sizeof(float)));
    HIPCHK(fc, dalloc(&fc->ghz32, (size_t)k * nhz * sizeof(float)));
    HIPCHK(fc, dalloc(&fc->d_sdims, (size_t)k * 3 * sizeof(int)));
    fc->geo_cap = k;
    return FMM3D_OK;
}

// Node counts per axis of the refined block of one point source (0s when
// the source lies outside the grid).
static std::array<int, 3> block_shape(const Grid& g, const double* src3, int radius, int refine)
{
    int lo[3], hi[3];
    if (!point_block(g, src3, radius, lo, hi)) return { 0, 0, 0 };
    return { (hi[0] - lo[0]) * refine + 1, (hi[1] - lo[1]) * refine + 1, (hi[2] - lo[2]) * refine + 1 };
}

// The refined neighbourhoods of point sources members[0..k-1] (src and
// radius index them) solved together: one slot per source on a
// refined-block context whose grid holds every member's block (the largest
// extent per axis), each slot with its own block's extent, origin,
// spacings and B-spline slowness.  Appends each source's fixed set for the
// coarse grid to node[q] / tv[q].
static int point_group(const fmm3d_ctx* c, fmm3d_ctx** finep, std::string& err, const std::vector<int>& members,
                       const double* src, const int* radius, int refine, int order,
                       std::vector<int64_t>* node, std::vector<double>* tv)
{
    const Grid& g = c->g;
    const double* velv = c->velv_host.data();
    const int k = (int)members.size();
    std::vector<Grid> fg(k);
    std::vector<std::array<int, 3>> blo(k);
    int cn[3] = { 0, 0, 0 };
    for (int m = 0; m < k; ++m) {
        const double* s3 = src + 3 * members[m];
        int lo[3], hi[3];
        if (!point_block(g, s3, radius[members[m]], lo, hi))
            return fail_to(err, FMM3D_E_ARG, "point source (%g km, %g, %g rad) lies outside the grid", s3[0], s3[1], s3[2]);
        Grid& f = fg[m];
        f = g;
        f.nnr = (hi[0] - lo[0]) * refine + 1;
        f.nnx = (hi[1] - lo[1]) * refine + 1;
        f.nnz = (hi[2] - lo[2]) * refine + 1;
        f.rgor = g.rgor - lo[0] * g.dnr;
        f.rgox = g.rgox + lo[1] * g.dnx;
        f.rgoz = g.rgoz + lo[2] * g.dnz;
        f.dnr = g.dnr / refine; f.dnx = g.dnx / refine; f.dnz = g.dnz / refine;
        blo[m] = { lo[0], lo[1], lo[2] };
        cn[0] = std::max(cn[0], f.nnr); cn[1] = std::max(cn[1], f.nnx); cn[2] = std::max(cn[2], f.nnz);
        if ((double)f.nnr * f.nnx * f.nnz > 64.0e6)
            return fail_to(err, FMM3D_E_ARG, "refined block of %dx%dx%d nodes is too large", f.nnr, f.nnx, f.nnz);
    }
    // the context's grid: the members' largest extents (origin of the first)
    Grid cg = fg[0];
    cg.nnr = cn[0]; cg.nnx = cn[1]; cg.nnz = cn[2];
    fmm3d_ctx* fc = *finep;
    if (fc && (fc->g.nnr != cg.nnr || fc->g.nnx != cg.nnx || fc->g.nnz != cg.nnz)) {
        fmm3d_destroy(fc);
        *finep = fc = nullptr;
    }
    if (!fc) {
        fmm3d_ctx* n = new fmm3d_ctx();
        n->g = cg;
        int rc = ctx_setup(n, c->device, finep);
        if (rc) return fail_to(err, rc, "refined grid: %s", fmm3d_last_error(nullptr));
        fc = *finep;
        // direct launches (no graph capture while other refined solves run
        // on other threads); the refined solves keep the default configuration
        fc->use_graph = 0;
        fc->autotune = 0;
    }
    fc->g = cg;
    fc->point_order = c->point_order;
    fc->point_window = c->point_window;
    // the vertex model (its gridding onto the shared arrays goes unused:
    // every slot has its own block)
    int rc = fmm3d_set_model_vtx(fc, velv);
    if (rc) return fail_to(err, rc, "refined gridder: %s", fc->err.c_str());
    if ((rc = geo_reserve(fc, k))) return fail_to(err, rc, "refined blocks: %s", fc->err.c_str());
    for (int m = 0; m < k; ++m)
        if ((rc = geo_slot(fc, m, fg[m]))) return fail_to(err, rc, "refined blocks: %s", fc->err.c_str());
    // straight-ray times on the corners of the fine cell holding each
    // source, at the context grid's node indices
    std::vector<int64_t> offs(k + 1, 0), fnode(8 * (size_t)k);
    std::vector<double> ftime(8 * (size_t)k);
    for (int m = 0; m < k; ++m) {
        const Grid& f = fg[m];
        const double* s3 = src + 3 * members[m];
        const double hs = s3[0], ths = s3[1], phs = s3[2];
        int fcell[3];
        const double fu[3] = { (f.rgor - hs) / f.dnr, (ths - f.rgox) / f.dnx, (phs - f.rgoz) / f.dnz };
        const int fn[3] = { f.nnr, f.nnx, f.nnz };
        for (int a = 0; a < 3; ++a) fcell[a] = std::min(std::max((int)floor(fu[a]), 0), fn[a] - 2);
        double xs[3];
        cart3(g, hs, ths, phs, xs);
        const double ss = 1.0 / host_velocity(g, velv, hs, g.gor - hs, ths, phs);
        int e = 8 * m;
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j)
                for (int l = 0; l < 2; ++l, ++e) {
                    const int r = fcell[0] + i, x = fcell[1] + j, z = fcell[2] + l;
                    const double hn = f.rgor - r * f.dnr, thn = f.rgox + x * f.dnx, phn = f.rgoz + z * f.dnz;
                    double xn[3];
                    cart3(g, hn, thn, phn, xn);
                    const double d = sqrt((xn[0] - xs[0]) * (xn[0] - xs[0]) + (xn[1] - xs[1]) * (xn[1] - xs[1]) +
                                          (xn[2] - xs[2]) * (xn[2] - xs[2]));
                    const double sn = 1.0 / host_velocity(f, velv, hn, f.gor - f.rgor + r * f.dnr, thn, phn);
                    fnode[e] = ((int64_t)r * cg.nnx + x) * cg.nnz + z;
                    ftime[e] = d * 0.5 * (ss + sn);
                }
        offs[m + 1] = 8 * (int64_t)(m + 1);
    }
    // float64: in float32 the refined block moved the coarse field by up to
    // 1.8 ms (61^3, refinement 5 over 10 nodes).  With point_warm the
    // float64 sweeps start from a float32 solve's field.
    // A failure anywhere in the warm start (the float32 solve, its staging,
    // the float64 sweeps from its field) falls back to the cold float64 solve.
    bool warm_ok = false;
    while (c->point_warm) {
        rc = fmm3d_solve_fixed(fc, k, offs.data(), fnode.data(), ftime.data(), order, FMM3D_PREC_FP32, FMM3D_WALK_ALTERNATE);
        if (rc) break;
        const size_t n = (size_t)k * fc->N;
        if (fc->d_warm_n < n) {
            (void)hipFree(fc->d_warm); (void)hipFree(fc->d_warm_off);
            fc->d_warm = nullptr; fc->d_warm_off = nullptr; fc->d_warm_n = 0;
            if (dalloc(&fc->d_warm, n * sizeof(float)) != hipSuccess || dalloc(&fc->d_warm_off, (size_t)k * sizeof(double) + 256) != hipSuccess)
                break;
            fc->d_warm_n = n;
        }
        hipError_t e = hipMemcpyAsync(fc->d_warm, fc->sl.T, n * sizeof(float), hipMemcpyDeviceToDevice, fc->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(fc->d_warm_off, fc->sl.toff, (size_t)k * sizeof(double), hipMemcpyDeviceToDevice, fc->stream);
        if (e != hipSuccess) break;
        // the float64 fields: fixed sets as in a cold solve, free nodes from the float32 field
        std::vector<double> toff;
        std::vector<int> flow;
        if ((rc = fixed_prepare(fc, "refined solve", k, offs.data(), fnode.data(), ftime.data(), toff, flow, fc->point_order != 0)) ||
            (rc = ensure_slots(fc, k, FMM3D_PREC_FP64)))
            break;
        fc->sl.n = 0;
        fc->walk_alt = 1;
        if ((rc = fixed_upload(fc, "refined solve", 0, k, offs.data(), fnode.data(), ftime.data(), FMM3D_PREC_FP64, toff, flow)))
            break;
        k_warm_start<<<(int)std::min<size_t>((n + 255) / 256, 16384), 256, 0, fc->stream>>>(
            (double*)fc->sl.T, fc->d_warm, fc->d_warm_off, fc->sl.fix, k, fc->N);
        if ((e = hipGetLastError()) != hipSuccess) break;
        rc = solve_run(fc, k, order, FMM3D_PREC_FP64);
        warm_ok = rc == FMM3D_OK;
        break;
    }
    if (!warm_ok)
        rc = fmm3d_solve_fixed(fc, k, offs.data(), fnode.data(), ftime.data(), order, FMM3D_PREC_FP64, FMM3D_WALK_ALTERNATE);
    if (rc) return fail_to(err, rc, "refined solve: %s", fc->err.c_str());
    // the coarse nodes each block covers seed the coarse solve: gathered on
    // the device, so only they cross to the host
    std::vector<double> cv;
    for (int m = 0; m < k; ++m) {
        const int q = members[m];
        const Grid& f = fg[m];
        const int* lo = blo[m].data();
        const int cr = (f.nnr - 1) / refine + 1, cx = (f.nnx - 1) / refine + 1, cz = (f.nnz - 1) / refine + 1;
        const size_t n = (size_t)cr * cx * cz;
        if (fc->d_gather_n < n) {
            (void)hipFree(fc->d_gather);
            fc->d_gather = nullptr;
            fc->d_gather_n = 0;
            if (dalloc(&fc->d_gather, n * sizeof(double)) != hipSuccess)
                return fail_to(err, FMM3D_E_NOMEM, "refined field: no device memory for %zu coarse nodes", n);
            fc->d_gather_n = n;
        }
        cv.resize(n);
        k_gather_coarse<<<(int)std::min<size_t>((n + 255) / 256, 4096), 256, 0, fc->stream>>>(
            (const double*)fc->sl.T + (size_t)m * fc->N, cr, cx, cz, refine, cg.nnx, cg.nnz, fc->d_gather);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipMemcpyAsync(cv.data(), fc->d_gather, n * sizeof(double), hipMemcpyDeviceToHost, fc->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(fc->stream);
        if (e != hipSuccess) return fail_to(err, FMM3D_E_DEVICE, "refined field: %s", hipGetErrorString(e));
        size_t i = 0;
        for (int r = lo[0]; r < lo[0] + cr; ++r)
            for (int x = lo[1]; x < lo[1] + cx; ++x)
                for (int z = lo[2]; z < lo[2] + cz; ++z, ++i) {
                    node[q].push_back(((int64_t)r * g.nnx + x) * g.nnz + z);
                    // an unreached node reads as 0, the reference's marker
                    tv[q].push_back(isfinite(cv[i]) ? cv[i] : 0.0);
                }
    }
    return FMM3D_OK;
}

// Fixed sets of point sources, source q refined over radius[q] nodes.
// Sources are solved in groups, one slot each on a refined-block context
// whose grid holds every member's block; up to point_workers threads run
// such groups at once, each on its own context and stream, and the sets
// are joined in source order.
static int point_sets(fmm3d_ctx* c, const char* fn, int nsrc, const double* src, int refine, const int* radius,
                      int order, std::vector<int64_t>& offs, std::vector<int64_t>& node, std::vector<double>& tv)
{
    if (nsrc <= 0 || !src || !radius) return fail(c, FMM3D_E_ARG, "%s: need nsrc > 0, sources and radii", fn);
    if (refine < 1 || refine > 16) return fail(c, FMM3D_E_ARG, "%s: refine must be 1..16", fn);
    for (int q = 0; q < nsrc; ++q)
        if (radius[q] < 1 || radius[q] > 64) return fail(c, FMM3D_E_ARG, "%s: radius of source %d must be 1..64", fn, q);
    std::vector<std::vector<int64_t>> pn(nsrc);
    std::vector<std::vector<double>> pt(nsrc);
    std::vector<int> prc(nsrc, FMM3D_OK);
    std::vector<std::string> perr(nsrc);
    // groups of sources solved as one batch on one refined context, whose
    // grid holds each member's block (at most 32 sources and 128 M context
    // nodes a group); largest blocks first, each worker taking the next
    // group when it is free
    std::vector<std::pair<int64_t, int>> order_by(nsrc);
    std::vector<std::array<int, 3>> shape(nsrc);
    for (int q = 0; q < nsrc; ++q) {
        shape[q] = block_shape(c->g, src + 3 * q, radius[q], refine);
        order_by[q] = { -(int64_t)shape[q][0] * shape[q][1] * shape[q][2], q };
    }
    std::sort(order_by.begin(), order_by.end());
    std::vector<std::vector<int>> groups;
    std::vector<std::array<int, 3>> gshape;
    for (const auto& o : order_by) {
        const int q = o.second;
        const std::array<int, 3>& sh = shape[q];
        bool join = !groups.empty() && sh[0] > 0 && gshape.back()[0] > 0 && groups.back().size() < 32;
        std::array<int, 3> m{ 0, 0, 0 };
        if (join) {
            for (int a = 0; a < 3; ++a) m[a] = std::max(gshape.back()[a], sh[a]);
            join = (double)(groups.back().size() + 1) * m[0] * m[1] * m[2] <= c->point_group_mnodes * 1.0e6;
        }
        if (join) {
            groups.back().push_back(q);
            gshape.back() = m;
        } else {
            groups.push_back({ q });
            gshape.push_back(sh);
        }
    }
    const int ng = (int)groups.size();
    const int nw = std::min(ng, c->point_workers);
    // refined contexts stay in a pool between calls, matched by block
    // shape: a worker reuses an idle one of its group's shape and creates
    // one otherwise (freeing device memory would wait for the other
    // workers' streams); past 16 contexts an idle one is rebuilt instead
    std::mutex pool_m;
    std::vector<char> busy(c->fines.size(), 0);
    auto acquire = [&](const std::array<int, 3>& sh) {
        std::lock_guard<std::mutex> lk(pool_m);
        int empty = -1, spare = -1;
        for (size_t i = 0; i < c->fines.size(); ++i) {
            if (busy[i]) continue;
            const fmm3d_ctx* f = c->fines[i];
            if (f && f->g.nnr == sh[0] && f->g.nnx == sh[1] && f->g.nnz == sh[2]) { busy[i] = 1; return (int)i; }
            if (!f && empty < 0) empty = (int)i;
            if (f && spare < 0) spare = (int)i;
        }
        if (empty >= 0) { busy[empty] = 1; return empty; }
        if (c->fines.size() >= 16 && spare >= 0) { busy[spare] = 1; return spare; }
        c->fines.push_back(nullptr);
        busy.push_back(1);
        return (int)c->fines.size() - 1;
    };
    std::atomic<int> next{ 0 };
    auto work = [&]() {
        hipStreamCaptureMode mode = hipStreamCaptureModeRelaxed;
        (void)hipThreadExchangeStreamCaptureMode(&mode);
        const bool dev_ok = hipSetDevice(c->device) == hipSuccess;
        for (int t; (t = next.fetch_add(1)) < ng;) {
            std::string e;
            int rc = FMM3D_E_DEVICE;
            if (!dev_ok) {
                e = "hipSetDevice failed";
            } else {
                const int i = acquire(gshape[t]);
                fmm3d_ctx* f;
                {
                    std::lock_guard<std::mutex> lk(pool_m);
                    f = c->fines[i];
                }
                rc = point_group(c, &f, e, groups[t], src, radius, refine, order, pn.data(), pt.data());
                std::lock_guard<std::mutex> lk(pool_m);
                c->fines[i] = f;
                busy[i] = 0;
            }
            if (rc)
                for (int q : groups[t]) { prc[q] = rc; perr[q] = e; }
        }
    };
    if (nw == 1) {
        work();
    } else {
        std::vector<std::thread> th;
        for (int w = 0; w < nw; ++w) th.emplace_back(work);
        for (auto& t : th) t.join();
    }
    offs.assign(nsrc + 1, 0);
    node.clear();
    tv.clear();
    for (int q = 0; q < nsrc; ++q) {
        if (prc[q]) return fail(c, prc[q], "%s: source %d: %s", fn, q, perr[q].c_str());
        node.insert(node.end(), pn[q].begin(), pn[q].end());
        tv.insert(tv.end(), pt[q].begin(), pt[q].end());
        offs[q + 1] = (int64_t)node.size();
    }
    return FMM3D_OK;
}

extern "C" int fmm3d_solve_point(fmm3d_ctx* c, int nsrc, const double* src, int refine, int radius,
                                 int order, int precision, int flags)
{
    if (!c) return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: null context");
    if (nsrc <= 0 || !src) return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: need nsrc > 0 and sources");
    if (refine < 1 || refine > 16 || radius < 1 || radius > 64)
        return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: refine must be 1..16 and radius 1..64");
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: bad order %d", order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: bad precision %d", precision);
    if (flags != 0) return fail(c, FMM3D_E_ARG, "fmm3d_solve_point: flags must be 0");
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "fmm3d_solve_point: no velocity model (call fmm3d_set_model_vtx)");
    std::vector<int64_t> offs, node;
    std::vector<double> tv;
    const std::vector<int> rad(nsrc, radius);
    int rc = point_sets(c, "fmm3d_solve_point", nsrc, src, refine, rad.data(), order, offs, node, tv);
    if (rc) return rc;
    rc = fmm3d_solve_fixed(c, nsrc, offs.data(), node.data(), tv.data(), order, precision, FMM3D_WALK_ALTERNATE);
    if (rc) return rc;
    // rays of these sources end at them
    HIPCHK(c, hipMemcpy(c->sl.srcpos, src, (size_t)3 * nsrc * sizeof(double), hipMemcpyHostToDevice));
    return FMM3D_OK;
}

extern "C" int fmm3d_solve_mixed(fmm3d_ctx* c, int ntele, const double* faces, const double* xface, const double* zface,
                                 const int* side, int npoint, const double* src, int refine, const int* radius,
                                 int order, int precision, int flags)
{
    const char* fn = "fmm3d_solve_mixed";
    if (!c) return fail(c, FMM3D_E_ARG, "%s: null context", fn);
    if (ntele < 0 || npoint < 0 || ntele + npoint == 0) return fail(c, FMM3D_E_ARG, "%s: need ntele, npoint >= 0 and a source", fn);
    if (order != FMM3D_ORDER_FIRST && order != FMM3D_ORDER_MIXED) return fail(c, FMM3D_E_ARG, "%s: bad order %d", fn, order);
    if (precision != FMM3D_PREC_FP64 && precision != FMM3D_PREC_FP32) return fail(c, FMM3D_E_ARG, "%s: bad precision %d", fn, precision);
    if (flags != 0) return fail(c, FMM3D_E_ARG, "%s: flags must be 0", fn);
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "%s: no velocity model (call fmm3d_set_model_vtx)", fn);
    std::vector<double> ptoff;
    std::vector<int64_t> offs, node;
    std::vector<double> tv;
    std::vector<int> pflow;
    int rc;
    if (ntele > 0 && (rc = tele_check(c, fn, ntele, faces, xface, zface, side))) return rc;
    if (npoint > 0) {
        if ((rc = point_sets(c, fn, npoint, src, refine, radius, order, offs, node, tv))) return rc;
        if ((rc = fixed_prepare(c, fn, npoint, offs.data(), node.data(), tv.data(), ptoff, pflow, c->point_order != 0)))
            return rc;
    }
    HIPCHK(c, hipSetDevice(c->device));
    const int nsrc = ntele + npoint;
    if ((rc = ensure_slots(c, nsrc, precision))) return rc;
    c->sl.n = 0;
    if (ntele > 0 && (rc = tele_upload(c, fn, ntele, faces, xface, zface, side, precision))) return rc;
    if (npoint > 0 && (rc = fixed_upload(c, fn, ntele, npoint, offs.data(), node.data(), tv.data(), precision, ptoff, pflow)))
        return rc;
    c->walk_slots.assign(nsrc, 0);
    for (int q = ntele; q < nsrc; ++q) c->walk_slots[q] = 1;
    c->walk_alt = npoint == 0 ? 0 : ntele == 0 ? 1 : 2;
    rc = solve_run(c, nsrc, order, precision);
    if (rc) return rc;
    if (npoint > 0)
        HIPCHK(c, hipMemcpy(c->sl.srcpos + 3 * (size_t)ntele, src, (size_t)3 * npoint * sizeof(double), hipMemcpyHostToDevice));
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
        std::vector<double> rs((size_t)g.nnr + 2);
        HIPCHK(c, hipMemcpy(f.data(), (float*)c->sl.T + (size_t)src * N, N * sizeof(float), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(&to, c->sl.toff + src, sizeof(double), hipMemcpyDeviceToHost));
        HIPCHK(c, hipMemcpy(rs.data(), c->sl.ref + (size_t)src * rs.size(), rs.size() * sizeof(double), hipMemcpyDeviceToHost));
        // offset + reference + stored value, as the device readers form it
        const double gx = rs[g.nnr], gz = rs[g.nnr + 1];
        size_t i = 0;
        for (int ir = 0; ir < g.nnr; ++ir)
            for (int ix = 0; ix < g.nnx; ++ix)
                for (int iz = 0; iz < g.nnz; ++iz, ++i)
                    d[i] = to + (rs[ir] + gx * (double)ix + gz * (double)iz) + (double)f[i];
    }
    // same phi-fastest order as the host; unknown nodes come back as 0 (the reference's marker)
    for (size_t i = 0; i < N; ++i) ttn[i] = isfinite(d[i]) ? d[i] : 0.0;
    return FMM3D_OK;
}

extern "C" int fmm3d_receivers(fmm3d_ctx* c, int nrc, const double* rcv, float* t_out)
{
    if (!c || (nrc > 0 && (!rcv || !t_out)) || nrc < 0) return fail(c, FMM3D_E_ARG, "fmm3d_receivers: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_receivers: no resident solution");
    if (nrc == 0) return FMM3D_OK;
    HIPCHK(c, hipSetDevice(c->device));
    const int ns = c->sl.n;
    HIPCHK(c, staging(&c->st_rcv, &c->st_rcv_n, (size_t)nrc * 3));
    HIPCHK(c, staging(&c->st_rout, &c->st_rout_n, (size_t)ns * nrc));
    double* drcv = c->st_rcv;
    float* dout = c->st_rout;
    HIPCHK(c, hipMemcpyAsync(drcv, rcv, (size_t)nrc * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    int total = ns * nrc;
    if (c->sl.prec == FMM3D_PREC_FP64)
        k_receivers<double><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const double*)c->sl.T, c->sl.toff, nullptr, ns, nrc, drcv, dout);
    else
        k_receivers<float><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const float*)c->sl.T, c->sl.toff, c->sl.ref, ns, nrc, drcv, dout);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(t_out, dout, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "receivers: %s", hipGetErrorString(e));
    return FMM3D_OK;
}

// fmm3d_receivers with its staging in the device's scratch array: no
// device memory is allocated or freed.
extern "C" int fmm3d_receivers_noalloc(fmm3d_ctx* c, int nrc, const double* rcv, float* t_out)
{
    if (!c || (nrc > 0 && (!rcv || !t_out)) || nrc < 0) return fail(c, FMM3D_E_ARG, "fmm3d_receivers_noalloc: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_receivers_noalloc: no resident solution");
    if (nrc == 0) return FMM3D_OK;
    const int ns = c->sl.n;
    const size_t rb = ((size_t)nrc * 3 * sizeof(double) + 255) & ~(size_t)255;
    const size_t need = rb + (size_t)ns * nrc * sizeof(float);
    if (need > NOALLOC_SCRATCH_BYTES)
        return fail(c, FMM3D_E_ARG, "fmm3d_receivers_noalloc: %d receivers x %d sources need %zu bytes of staging, "
                    "above its maximum of %zu", nrc, ns, need, NOALLOC_SCRATCH_BYTES);
    if (!c->scratch) return fail(c, FMM3D_E_STATE, "fmm3d_receivers_noalloc: no scratch address on this device");
    HIPCHK(c, hipSetDevice(c->device));
    unsigned char* base = (unsigned char*)c->scratch;
    double* drcv = (double*)base;
    float* dout = (float*)(base + rb);
    hipError_t e = hipMemcpyAsync(drcv, rcv, (size_t)nrc * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    const int total = ns * nrc;
    if (e == hipSuccess) {
        if (c->sl.prec == FMM3D_PREC_FP64)
            k_receivers<double><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const double*)c->sl.T, c->sl.toff, nullptr, ns, nrc, drcv, dout);
        else
            k_receivers<float><<<(total + 255) / 256, 256, 0, c->stream>>>(c->g, (const float*)c->sl.T, c->sl.toff, c->sl.ref, ns, nrc, drcv, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(t_out, dout, (size_t)total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "receivers (no allocation): %s", hipGetErrorString(e));
    return FMM3D_OK;
}

// Rays ray_begin .. ray_end-1 (ray slot*nrc + i: slot's field, receiver i)
// with their Frechet rows; row_ptr has ray_end-ray_begin+1 entries.
static int frechet_range(fmm3d_ctx* c, int nrc, const double* rcv, int ray_begin, int ray_end, int64_t* row_ptr,
                         int32_t* cols, float* vals, int64_t cap, const int* islot = nullptr,
                         int32_t* dcols = nullptr, float* dvals = nullptr)
{
    HIPCHK(c, hipSetDevice(c->device));
    const Grid& g = c->g;
    const int nrays = ray_end - ray_begin;
    row_ptr[0] = 0;
    if (nrays <= 0) return FMM3D_OK;
    const size_t NV = (size_t)(g.nvr + 2) * (g.nvx + 2) * (g.nvz + 2);
    // ray workers: one wavefront each, with a dense float32 scratch of NV
    const size_t workers = std::min<size_t>(16384, (size_t)nrays);
    if (c->ray_scratch_workers < workers) {
        (void)hipFree(c->ray_scratch);
        c->ray_scratch = nullptr;
        c->ray_scratch_workers = 0;
        HIPCHK(c, dalloc(&c->ray_scratch, workers * NV * sizeof(float)));
        HIPCHK(c, hipMemsetAsync(c->ray_scratch, 0, workers * NV * sizeof(float), c->stream));
        c->ray_scratch_workers = workers;
    }
    // Rays go in chunks whose staging (fr_rcap entries per ray) and path
    // samples (fr_pcap per ray) fit fixed budgets.  Per chunk: paths, one
    // lane per ray -> Frechet rows, one wavefront per ray -> row offsets on
    // the device -> counts to the host.  With device-visible destination
    // arrays (dcols, dvals: pinned host memory) the rows are packed straight
    // into them before the counts come back; otherwise they are packed on
    // the device and copied into the caller's arrays.  A chunk with a row
    // longer than the staging runs again with room for the whole vertex
    // grid, one with a path longer than its samples with four times the
    // samples.
    const size_t budget = (size_t)512 << 20;   // bytes of staging
    const size_t pbudget = (size_t)2 << 30;    // bytes of path samples
    if (c->fr_pcap <= 0) c->fr_pcap = 2 * ((int64_t)g.nnr + g.nnx + g.nnz) + 64;
    // items (islot): one slot index and one receiver per ray, rays 0 .. nrays-1
    const size_t nco = islot ? (size_t)nrays : (size_t)nrc;
    HIPCHK(c, staging(&c->st_rcv, &c->st_rcv_n, nco * 3));
    double* drcv = c->st_rcv;
    int* dslot = nullptr;
    hipError_t e = hipMemcpyAsync(drcv, rcv, nco * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (islot) {
        HIPCHK(c, staging(&c->st_islot, &c->st_islot_n, (size_t)nrays));
        dslot = c->st_islot;
        if (e == hipSuccess) e = hipMemcpyAsync(dslot, islot, (size_t)nrays * sizeof(int), hipMemcpyHostToDevice, c->stream);
    }
    int64_t total = 0;
    bool short_cap = false;
    for (int r0 = ray_begin; r0 < ray_end && e == hipSuccess;) {
        const int64_t rcap = std::min<int64_t>(c->fr_rcap, (int64_t)NV);
        const int64_t pcap = c->fr_pcap;
        const int chunk = (int)std::max<int64_t>(1, std::min<int64_t>({ (int64_t)(ray_end - r0), (int64_t)(budget / (rcap * 8)),
                                                                        (int64_t)(pbudget / (pcap * sizeof(RayPoint))) }));
        if (c->fr_path_n < (size_t)chunk * pcap) {
            (void)hipFree(c->fr_path);
            c->fr_path = nullptr; c->fr_path_n = 0;
            e = dalloc(&c->fr_path, (size_t)chunk * pcap * sizeof(RayPoint));
            if (e != hipSuccess) break;
            c->fr_path_n = (size_t)chunk * pcap;
        }
        if (c->fr_path_rays < (size_t)chunk) {
            (void)hipFree(c->fr_pn); (void)hipFree(c->fr_pbox);
            c->fr_pn = nullptr; c->fr_pbox = nullptr; c->fr_path_rays = 0;
            e = dalloc(&c->fr_pn, (size_t)chunk * sizeof(int32_t));
            if (e == hipSuccess) e = dalloc(&c->fr_pbox, (size_t)chunk * 6 * sizeof(int32_t));
            if (e != hipSuccess) break;
            c->fr_path_rays = chunk;
        }
        if (c->fr_stage < (size_t)chunk * rcap) {
            (void)hipFree(c->fr_cols); (void)hipFree(c->fr_vals);
            c->fr_cols = nullptr; c->fr_vals = nullptr; c->fr_stage = 0;
            e = dalloc(&c->fr_cols, (size_t)chunk * rcap * sizeof(int32_t));
            if (e == hipSuccess) e = dalloc(&c->fr_vals, (size_t)chunk * rcap * sizeof(float));
            if (e != hipSuccess) break;
            c->fr_stage = (size_t)chunk * rcap;
        }
        if (c->fr_rays < (size_t)chunk) {
            (void)hipFree(c->fr_count); (void)hipFree(c->fr_off);
            c->fr_count = nullptr; c->fr_off = nullptr; c->fr_rays = 0;
            e = dalloc(&c->fr_count, (size_t)chunk * sizeof(int32_t));
            if (e == hipSuccess) e = dalloc(&c->fr_off, (size_t)chunk * sizeof(int64_t));
            if (e != hipSuccess) break;
            c->fr_rays = chunk;
        }
        const int nr = chunk;
        RayOut o{ c->fr_cols, c->fr_vals, c->fr_count, rcap };
        RayPaths pp{ c->fr_path, c->fr_pn, c->fr_pbox, pcap };
        const int pgrid = (nr + 63) / 64;
        if (c->sl.prec == FMM3D_PREC_FP64)
            k_ray_paths<double><<<pgrid, 64, 0, c->stream>>>(g, (const double*)c->sl.T, c->sl.toff, nullptr, nrc, drcv, r0, nr, pp, c->sl.srcpos, dslot);
        else
            k_ray_paths<float><<<pgrid, 64, 0, c->stream>>>(g, (const float*)c->sl.T, c->sl.toff, c->sl.ref, nrc, drcv, r0, nr, pp, c->sl.srcpos, dslot);
        e = hipGetLastError();
        if (e == hipSuccess) {
            const int grid = (int)std::min<size_t>(workers, (size_t)nr);
            k_ray_frechet<<<grid, 64, 0, c->stream>>>(g, c->veln, nr, pp, c->ray_scratch, o);
            e = hipGetLastError();
        }
        if (e == hipSuccess) {
            k_row_offsets<<<1, 1024, 0, c->stream>>>(nr, c->fr_count, c->fr_off);
            e = hipGetLastError();
        }
        if (e == hipSuccess && dcols && !short_cap) {
            k_pack_rows<<<std::min(nr, 8192), 256, 0, c->stream>>>(nr, rcap, c->fr_count, c->fr_off, c->fr_cols, c->fr_vals,
                                                                  dcols + total, dvals + total, cap - total);
            e = hipGetLastError();
        }
        if (e == hipSuccess && c->pin_cnt_n < (size_t)nr) {
            if (c->pin_cnt) (void)hipHostFree(c->pin_cnt);
            c->pin_cnt = nullptr; c->pin_cnt_n = 0;
            e = hipHostMalloc((void**)&c->pin_cnt, (size_t)nr * sizeof(int32_t), hipHostMallocDefault);
            if (e == hipSuccess) c->pin_cnt_n = nr;
        }
        if (e == hipSuccess) e = hipMemcpyAsync(c->pin_cnt, c->fr_count, nr * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) break;
        const int32_t* cnt = c->pin_cnt;
        bool over = false, pover = false;
        for (int q = 0; q < nr; ++q) { over |= cnt[q] == -2; pover |= cnt[q] == -3; }
        if (pover) {   // a path longer than its samples: four times the samples (a ray has at most 5e5)
            if (pcap >= 500000) { e = hipErrorUnknown; break; }
            c->fr_pcap = std::min<int64_t>(4 * pcap, 500000);
            continue;
        }
        if (over) {
            if (rcap >= (int64_t)NV) { e = hipErrorUnknown; break; }   // cannot happen: a row has at most NV entries
            c->fr_rcap = (int64_t)NV;
            continue;
        }
        int64_t ctot = 0;
        for (int q = 0; q < nr; ++q) {
            const int64_t n = std::max(cnt[q], 0);
            row_ptr[r0 - ray_begin + q + 1] = row_ptr[r0 - ray_begin + q] + n;
            ctot += n;
        }
        if (short_cap || total + ctot > cap) {
            short_cap = true;
        } else if (!dcols && ctot > 0) {
            if (c->fr_packed < (size_t)ctot) {
                (void)hipFree(c->fr_pcols); (void)hipFree(c->fr_pvals);
                c->fr_pcols = nullptr; c->fr_pvals = nullptr; c->fr_packed = 0;
                e = dalloc(&c->fr_pcols, (size_t)ctot * sizeof(int32_t));
                if (e == hipSuccess) e = dalloc(&c->fr_pvals, (size_t)ctot * sizeof(float));
                if (e != hipSuccess) break;
                c->fr_packed = ctot;
            }
            k_pack_rows<<<std::min(nr, 8192), 256, 0, c->stream>>>(nr, rcap, c->fr_count, c->fr_off, c->fr_cols, c->fr_vals,
                                                                  c->fr_pcols, c->fr_pvals, ctot);
            e = hipGetLastError();
            if (e == hipSuccess) e = hipMemcpyAsync(cols + total, c->fr_pcols, ctot * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(vals + total, c->fr_pvals, ctot * sizeof(float), hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e != hipSuccess) break;
        }
        total += ctot;
        r0 += nr;
    }
    if (e == hipErrorUnknown) return fail(c, FMM3D_E_DEVICE, "frechet: row staging or path overflow");
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "frechet: %s", hipGetErrorString(e));
    if (short_cap) return fail(c, FMM3D_E_CAPACITY, "fmm3d_frechet: capacity %lld < %lld entries", (long long)cap, (long long)row_ptr[nrays]);
    return FMM3D_OK;
}

extern "C" int fmm3d_frechet(fmm3d_ctx* c, int nrc, const double* rcv, int64_t* row_ptr, int32_t* cols,
                             float* vals, int64_t cap)
{
    if (!c || nrc < 0 || (nrc > 0 && (!rcv || !row_ptr)) || cap < 0) return fail(c, FMM3D_E_ARG, "fmm3d_frechet: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_frechet: no resident solution");
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "fmm3d_frechet: no velocity model");
    return frechet_range(c, nrc, rcv, 0, c->sl.n * nrc, row_ptr, cols, vals, cap);
}

// fmm3d_frechet into the context's own pinned host arrays: the copies from
// the device then run at the link's rate straight into the result, with no
// staging through bounce buffers.  The arrays grow as needed (sized from
// the previous call's entries per row, one retry at the exact size).
extern "C" int fmm3d_frechet_pinned(fmm3d_ctx* c, int nrc, const double* rcv, const int64_t** row_ptr,
                                    const int32_t** cols, const float** vals, int64_t* nnz)
{
    if (!c || nrc < 0 || (nrc > 0 && !rcv) || !row_ptr || !cols || !vals || !nnz)
        return fail(c, FMM3D_E_ARG, "fmm3d_frechet_pinned: bad arguments");
    if (c->sl.n <= 0) return fail(c, FMM3D_E_STATE, "fmm3d_frechet_pinned: no resident solution");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t rows = (size_t)c->sl.n * nrc + 1;
    if (c->pin_rows < rows) {
        if (c->pin_rp) (void)hipHostFree(c->pin_rp);
        c->pin_rp = nullptr; c->pin_rows = 0;
        HIPCHK(c, hipHostMalloc((void**)&c->pin_rp, rows * sizeof(int64_t), hipHostMallocDefault));
        c->pin_rows = rows;
    }
    int64_t want = std::max<int64_t>(c->pin_cap, (int64_t)(rows - 1) * 512 + 1024);
    for (int attempt = 0; attempt < 2; ++attempt) {
        if (c->pin_cap < want) {
            if (c->pin_cols) (void)hipHostFree(c->pin_cols);
            if (c->pin_vals) (void)hipHostFree(c->pin_vals);
            c->pin_cols = nullptr; c->pin_vals = nullptr; c->pin_cap = 0;
            HIPCHK(c, hipHostMalloc((void**)&c->pin_cols, (size_t)want * sizeof(int32_t), hipHostMallocDefault));
            HIPCHK(c, hipHostMalloc((void**)&c->pin_vals, (size_t)want * sizeof(float), hipHostMallocDefault));
            c->pin_cap = want;
            HIPCHK(c, hipHostGetDevicePointer((void**)&c->pin_dcols, c->pin_cols, 0));
            HIPCHK(c, hipHostGetDevicePointer((void**)&c->pin_dvals, c->pin_vals, 0));
        }
        if (!c->have_model) return fail(c, FMM3D_E_STATE, "fmm3d_frechet_pinned: no velocity model");
        const int rc = frechet_range(c, nrc, rcv, 0, c->sl.n * nrc, c->pin_rp, c->pin_cols, c->pin_vals, c->pin_cap, nullptr,
                                     c->pin_dcols, c->pin_dvals);
        if (rc == FMM3D_E_CAPACITY) { want = c->pin_rp[rows - 1]; continue; }
        if (rc) return rc;
        *row_ptr = c->pin_rp; *cols = c->pin_cols; *vals = c->pin_vals; *nnz = c->pin_rp[rows - 1];
        return FMM3D_OK;
    }
    return fail(c, FMM3D_E_CAPACITY, "fmm3d_frechet_pinned: capacity retry failed");
}

// ------------------------------------------------------------ engine

int fmm3d_engine_open(fmm3d_ctx* c, int cap, int precision, int order)
{
    if (!c || cap <= 0) return fail(c, FMM3D_E_ARG, "engine: bad capacity");
    if (!c->have_model) return fail(c, FMM3D_E_STATE, "engine: no model");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = ensure_slots(c, cap, precision);
    if (rc) return rc;
    Slots& s = c->sl;
    const size_t nbmax = (size_t)((c->g.nnr + 3) / 4) * ((c->g.nnx + 3) / 4) * ((c->g.nnz + 7) / 8);
    k_fill_int<<<(cap + 255) / 256, 256, 0, c->stream>>>(s.on, cap, 0);
    k_fill_int<<<(cap + 255) / 256, 256, 0, c->stream>>>(s.lastchg, cap, 0);
    k_fill_int<<<(cap + 255) / 256, 256, 0, c->stream>>>(s.valchg, cap, 0);
    k_fill_int<<<(int)((cap * nbmax + 255) / 256), 256, 0, c->stream>>>(s.prog, (int)(cap * nbmax), 0);
    k_fill_int<<<(int)((cap * nbmax + 255) / 256), 256, 0, c->stream>>>(s.stamp, (int)(cap * nbmax), -1);
    k_fill_real<double><<<(3 * cap + 255) / 256, 256, 0, c->stream>>>(s.srcpos, (size_t)3 * cap, __builtin_nan(""));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (!c->eng_lc) HIPCHK(c, hipHostMalloc((void**)&c->eng_lc, sizeof(int) * 65536, hipHostMallocDefault));
    if (cap > 65536) return fail(c, FMM3D_E_ARG, "engine: at most 65536 slots");
    c->eng_open = true;
    c->eng_k = 1;
    c->eng_order = order;
    c->eng_pending = 0;
    c->eng_active.assign(cap, 0);
    c->eng_born.assign(cap, 0);
    c->eng_flow.assign((size_t)cap * 2, 1);
    c->eng_ref.assign((size_t)cap * (c->g.nnr + 2), 0.0);
    c->eng_off.assign((size_t)cap, 0.0);
    (void)hipFree(c->eng_faces);
    c->eng_faces = nullptr;
    HIPCHK(c, dalloc(&c->eng_faces, (size_t)cap * c->g.nnx * c->g.nnz * sizeof(double)));
    s.n = cap;
    c->stats = fmm3d_stats{};
    return FMM3D_OK;
}

int fmm3d_engine_active(const fmm3d_ctx* c)
{
    int n = 0;
    for (int v : c->eng_active) n += v;
    return n;
}

int fmm3d_engine_admit(fmm3d_ctx* c, int slot, const int64_t* node, const double* tv, int64_t n, int walkalt,
                       const double* src3)
{
    if (!c || !c->eng_open || slot < 0 || slot >= c->sl.cap || c->eng_active[slot] || n < 0)
        return fail(c, FMM3D_E_ARG, "engine: bad admission");
    Slots& s = c->sl;
    double m = __builtin_huge_val();
    for (int64_t i = 0; i < n; ++i) {
        if (node[i] < 0 || (size_t)node[i] >= c->N || !(tv[i] >= 0.0) || !isfinite(tv[i]))
            return fail(c, FMM3D_E_ARG, "engine: bad fixed node %lld", (long long)node[i]);
        m = std::min(m, tv[i]);
    }
    const double toff = isfinite(m) ? m : 0.0;
    int flow[2] = { 1, 1 };
    if (walkalt && c->point_order) centre_flow(c->g, node, tv, n, flow);
    else fixed_flow(c->g, node, tv, n, flow);
    if (c->eng_stage < (size_t)n + 2) {
        (void)hipFree(c->eng_node); (void)hipFree(c->eng_tv);
        c->eng_node = nullptr; c->eng_tv = nullptr; c->eng_stage = 0;
        const size_t want = (size_t)std::max<int64_t>(n + 2, 1 << 16);
        HIPCHK(c, dalloc(&c->eng_node, want * sizeof(int64_t)));
        HIPCHK(c, dalloc(&c->eng_tv, want * sizeof(double)));
        c->eng_stage = want;
    }
    // the previous launch may still read this slot's arrays: admissions wait for it
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t N = c->N, NW = (N + 31) / 32;
    const Grid& g = c->g;
    const int npx = (g.nnx + patch_x(c) - 1) / patch_x(c), npz = (g.nnz + patch_z(c) - 1) / patch_z(c);
    const int np = npx * npz;
    const int64_t offs[2] = { 0, n };
    int64_t* doffs = c->eng_node + c->eng_stage - 2;   // the staging's last two entries hold the offsets
    hipError_t e = hipMemcpyAsync(doffs, offs, sizeof offs, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && n > 0) e = hipMemcpyAsync(c->eng_node, node, n * sizeof(int64_t), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && n > 0) e = hipMemcpyAsync(c->eng_tv, tv, n * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.toff + slot, &toff, sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(s.flow + 2 * slot, flow, sizeof flow, hipMemcpyHostToDevice, c->stream);
    double sp[3] = { __builtin_nan(""), __builtin_nan(""), __builtin_nan("") };
    if (src3) { sp[0] = src3[0]; sp[1] = src3[1]; sp[2] = src3[2]; }
    if (e == hipSuccess) e = hipMemcpyAsync(s.srcpos + 3 * slot, sp, sizeof sp, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        k_fill_real<double><<<1, 256, 0, c->stream>>>(s.ref + (size_t)slot * (g.nnr + 2), (size_t)g.nnr + 2, 0.0);
        if (s.prec == FMM3D_PREC_FP64) {
            double* T = (double*)s.T + (size_t)slot * N;
            k_fill_real<double><<<1024, 256, 0, c->stream>>>(T, N, Lim<double>::inf());
            k_fill_u32<<<1024, 256, 0, c->stream>>>(s.fix + (size_t)slot * NW, NW);
            if (n > 0) k_scatter_fixed<double><<<(int)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, c->stream>>>(
                1, N, doffs, c->eng_node, c->eng_tv, s.toff + slot, T, s.fix + (size_t)slot * NW);
        } else {
            float* T = (float*)s.T + (size_t)slot * N;
            k_fill_real<float><<<1024, 256, 0, c->stream>>>(T, N, Lim<float>::inf());
            k_fill_u32<<<1024, 256, 0, c->stream>>>(s.fix + (size_t)slot * NW, NW);
            if (n > 0) k_scatter_fixed<float><<<(int)std::min<int64_t>(1024, (n + 255) / 256), 256, 0, c->stream>>>(
                1, N, doffs, c->eng_node, c->eng_tv, s.toff + slot, T, s.fix + (size_t)slot * NW);
        }
        // every patch column of the new slot runs in the next sweep
        k_fill_int<<<(np + 255) / 256, 256, 0, c->stream>>>(s.stamp + (size_t)slot * np, np, c->eng_k - 1);
        k_fill_int<<<(np + 255) / 256, 256, 0, c->stream>>>(s.prog + (size_t)slot * np, np, 0);
        k_lrng_init<<<(np + 255) / 256, 256, 0, c->stream>>>(s.lrng + (size_t)slot * np * 2, np, c->eng_k);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.lastchg + slot, 1, c->eng_k);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.walkalt + slot, 1, walkalt ? 1 : 0);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.on + slot, 1, 1);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.born + slot, 1, c->eng_k);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.slotpend + slot, 1, 0);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.wlost + slot, 1, 0);
        k_fill_int<<<(np + 255) / 256, 256, 0, c->stream>>>(s.pend + (size_t)slot * np, np, 0);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "engine admit: %s", hipGetErrorString(e));
    c->eng_active[slot] = 1;
    c->eng_born[slot] = c->eng_k;
    return FMM3D_OK;
}

// A newly admitted slot's state: stamps (every patch column runs in sweep
// k), progress and check marks cleared, last change k, walk mode, first
// sweep k, switched on, no point-source position.
__global__ void k_slot_setup(int slot, int np, int k, int walkalt, int* __restrict__ stamp, int* __restrict__ prog,
                             int* __restrict__ pend, int* __restrict__ lastchg, int* __restrict__ walk,
                             int* __restrict__ born, int* __restrict__ slotpend, int* __restrict__ on,
                             double* __restrict__ srcpos, unsigned long long* __restrict__ lrng)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < np) {
        const size_t p = (size_t)slot * np + i;
        stamp[p] = k - 1;
        prog[p] = 0;
        pend[p] = 0;
        lrng[2 * p + ((k - 1) & 1)] = lrng_pack(k - 1, 0, 0xffff);
        lrng[2 * p + (k & 1)] = lrng_pack(-1000, 0, 0);
    }
    if (i == 0) {
        lastchg[slot] = k;
        walk[slot] = walkalt;
        born[slot] = k;
        slotpend[slot] = 0;
        on[slot] = 1;
    }
    if (i < 3) srcpos[3 * (size_t)slot + i] = __builtin_nan("");
}

// A teleseismic source into a free slot straight from its bottom face (the
// batch solve's initial state for one slot), with no host node list and no
// wait: the uploads and kernels queue behind the previous launch on the
// context's stream, and the host arrays they read (the request's face, the
// per-slot flow vector) outlive them.
int fmm3d_engine_admit_face(fmm3d_ctx* c, int slot, const double* face)
{
    if (!c || !c->eng_open || slot < 0 || slot >= c->sl.cap || c->eng_active[slot] || !face)
        return fail(c, FMM3D_E_ARG, "engine: bad admission");
    Slots& s = c->sl;
    const Grid& g = c->g;
    const size_t NL = (size_t)g.nnx * g.nnz, N = c->N, NW = (N + 31) / 32;
    const int npx = (g.nnx + patch_x(c) - 1) / patch_x(c), npz = (g.nnz + patch_z(c) - 1) / patch_z(c);
    const int np = npx * npz;
    lateral_flow(g, face, &c->eng_flow[2 * (size_t)slot]);
    double* df = c->eng_faces + (size_t)slot * NL;
    if (!c->d_bad) HIPCHK(c, dalloc(&c->d_bad, sizeof(int)));
    hipError_t e = hipMemcpyAsync(df, face, NL * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.flow + 2 * slot, &c->eng_flow[2 * (size_t)slot], 2 * sizeof(int), hipMemcpyHostToDevice, c->stream);
    // a float32 slot is stored against the face's plane-wave reference, as
    // in the batched solves (offset: the plane's; else the face minimum)
    const size_t RS = (size_t)g.nnr + 2;
    double* rrow = c->eng_ref.data() + (size_t)slot * RS;
    const bool useref = s.prec == FMM3D_PREC_FP32 && c->tele_ref && c->strategy == 1 &&
                        tele_reference(c, face, rrow, &c->eng_off[slot]);
    if (!useref) std::fill(rrow, rrow + RS, 0.0);
    if (e == hipSuccess) e = hipMemcpyAsync(s.ref + (size_t)slot * RS, rrow, RS * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess && useref)
        e = hipMemcpyAsync(s.toff + slot, &c->eng_off[slot], sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        // NaN faces are refused at submission; the flag is not read here
        if (!useref) k_tele_offsets<<<1, 1024, 0, c->stream>>>(NL, 0, 0, df, nullptr, nullptr, nullptr, s.toff + slot, c->d_bad);
        if (s.prec == FMM3D_PREC_FP64)
            k_init_tele<double><<<dim3(256, 1), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, 1, df, (double*)s.T + (size_t)slot * N, s.toff + slot, nullptr);
        else
            k_init_tele<float><<<dim3(256, 1), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, 1, df, (float*)s.T + (size_t)slot * N, s.toff + slot,
                                                                    useref ? s.ref + (size_t)slot * RS : nullptr);
        k_fixbits_tele<<<dim3(64, 1), 256, 0, c->stream>>>(g.nnr, g.nnx, g.nnz, 1, df, s.fix + (size_t)slot * NW);
        // the slot's per-patch and per-slot state in one launch; every patch
        // column of the new slot runs in the next sweep
        k_slot_setup<<<(np + 255) / 256, 256, 0, c->stream>>>(slot, np, c->eng_k, 0, s.stamp, s.prog, s.pend, s.lastchg,
                                                             s.walkalt, s.born, s.slotpend, s.on, s.srcpos, s.lrng);
        k_fill_int<<<1, 1, 0, c->stream>>>(s.wlost + slot, 1, 0);
        e = hipGetLastError();
    }
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "engine admit: %s", hipGetErrorString(e));
    c->eng_active[slot] = 1;
    c->eng_born[slot] = c->eng_k;
    return FMM3D_OK;
}

template <typename R, bool FAST, bool SHIFT>
static int engine_sweeps(fmm3d_ctx* c, int nsweeps)
{
    ColArgs<R> a;
    int blocks = 0;
    int rc = column_args<R>(c, c->sl.cap, c->eng_order, a, blocks);
    if (rc) return rc;
    a.born = c->sl.born;   // each slot's fixed-point check follows its own first sweep
    rc = launch_chunk<R, FAST, SHIFT>(c, a, blocks, nsweeps, c->eng_k, true);
    if (rc) return rc;
    c->eng_k += nsweeps;
    HIPCHK(c, hipMemcpyAsync(c->eng_lc, c->sl.lastchg, c->sl.cap * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    return FMM3D_OK;
}

int fmm3d_engine_launch(fmm3d_ctx* c, int nsweeps)
{
    if (!c || !c->eng_open || nsweeps <= 0) return fail(c, FMM3D_E_ARG, "engine: bad launch");
    if (c->eng_pending) return fail(c, FMM3D_E_STATE, "engine: collect the previous launch first");
    HIPCHK(c, hipSetDevice(c->device));
    int rc = c->sl.prec == FMM3D_PREC_FP64 ? engine_sweeps<double, false, false>(c, nsweeps)
                                           : engine_sweeps<float, true, true>(c, nsweeps);
    if (rc) return rc;
    c->eng_pending = 1;
    return FMM3D_OK;
}

int fmm3d_engine_collect(fmm3d_ctx* c, std::vector<int>& done, std::vector<int>& failed)
{
    done.clear();
    failed.clear();
    if (!c || !c->eng_open) return fail(c, FMM3D_E_ARG, "engine: not open");
    if (!c->eng_pending) return FMM3D_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->eng_pending = 0;
    const Grid& g = c->g;
    const int maxit = 4 * (g.nnr + g.nnx + g.nnz) + 200;
    const int last = c->eng_k - 1;      // the sweep just run
    for (int q = 0; q < c->sl.cap; ++q) {
        if (!c->eng_active[q]) continue;
        if (c->eng_lc[q] < last && last > c->eng_born[q]) done.push_back(q);
        else if (last - c->eng_born[q] > maxit) failed.push_back(q);
    }
    return FMM3D_OK;
}

// Receiver staging of at least n items (grows; kept across calls).
static hipError_t engine_rcv_stage(fmm3d_ctx* c, size_t n)
{
    if (c->eng_rcv_n >= n) return hipSuccess;
    (void)hipFree(c->eng_rcv); (void)hipFree(c->eng_out); (void)hipFree(c->eng_islot);
    c->eng_rcv = nullptr; c->eng_out = nullptr; c->eng_islot = nullptr; c->eng_rcv_n = 0;
    const size_t want = std::max<size_t>(n, 4096);
    hipError_t e = dalloc(&c->eng_rcv, want * 3 * sizeof(double));
    if (e == hipSuccess) e = dalloc(&c->eng_out, want * sizeof(float));
    if (e == hipSuccess) e = dalloc(&c->eng_islot, want * sizeof(int));
    if (e == hipSuccess) c->eng_rcv_n = want;
    return e;
}

int fmm3d_engine_receivers_many(fmm3d_ctx* c, int n, const int* slots, const int* nrc, const double* const* rcv,
                                float* const* out)
{
    if (!c || !c->eng_open || n < 0) return fail(c, FMM3D_E_ARG, "engine: bad receivers");
    size_t total = 0;
    for (int k = 0; k < n; ++k) {
        if (slots[k] < 0 || slots[k] >= c->sl.cap || nrc[k] < 0) return fail(c, FMM3D_E_ARG, "engine: bad receivers");
        total += (size_t)nrc[k];
    }
    if (total == 0) return FMM3D_OK;
    if (total > (size_t)INT32_MAX) return fail(c, FMM3D_E_ARG, "engine: too many receivers");
    HIPCHK(c, engine_rcv_stage(c, total));
    c->eng_hrcv.resize(total * 3);
    c->eng_hslot.resize(total);
    c->eng_hout.resize(total);
    size_t o = 0;
    for (int k = 0; k < n; ++k) {
        memcpy(c->eng_hrcv.data() + 3 * o, rcv[k], (size_t)nrc[k] * 3 * sizeof(double));
        std::fill(c->eng_hslot.begin() + o, c->eng_hslot.begin() + o + nrc[k], slots[k]);
        o += nrc[k];
    }
    hipError_t e = hipMemcpyAsync(c->eng_rcv, c->eng_hrcv.data(), total * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(c->eng_islot, c->eng_hslot.data(), total * sizeof(int), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        const int nb = (int)((total + 255) / 256);
        if (c->sl.prec == FMM3D_PREC_FP64)
            k_receivers_items<double><<<nb, 256, 0, c->stream>>>(c->g, (const double*)c->sl.T, c->sl.toff, nullptr, (int)total,
                                                                 c->eng_islot, c->eng_rcv, c->eng_out);
        else
            k_receivers_items<float><<<nb, 256, 0, c->stream>>>(c->g, (const float*)c->sl.T, c->sl.toff, c->sl.ref, (int)total,
                                                                c->eng_islot, c->eng_rcv, c->eng_out);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(c->eng_hout.data(), c->eng_out, total * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "engine receivers: %s", hipGetErrorString(e));
    o = 0;
    for (int k = 0; k < n; ++k) {
        memcpy(out[k], c->eng_hout.data() + o, (size_t)nrc[k] * sizeof(float));
        o += nrc[k];
    }
    return FMM3D_OK;
}

int fmm3d_engine_receivers(fmm3d_ctx* c, int slot, int nrc, const double* rcv, float* out)
{
    if (!c || !c->eng_open || slot < 0 || slot >= c->sl.cap || nrc < 0) return fail(c, FMM3D_E_ARG, "engine: bad receivers");
    if (nrc == 0) return FMM3D_OK;
    HIPCHK(c, engine_rcv_stage(c, (size_t)nrc));
    double* drcv = c->eng_rcv;
    float* dout = c->eng_out;
    hipError_t e = hipMemcpyAsync(drcv, rcv, (size_t)nrc * 3 * sizeof(double), hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) {
        if (c->sl.prec == FMM3D_PREC_FP64)
            k_receivers<double><<<(nrc + 255) / 256, 256, 0, c->stream>>>(c->g, (const double*)c->sl.T + (size_t)slot * c->N,
                                                                        c->sl.toff + slot, nullptr, 1, nrc, drcv, dout);
        else
            k_receivers<float><<<(nrc + 255) / 256, 256, 0, c->stream>>>(c->g, (const float*)c->sl.T + (size_t)slot * c->N,
                                                                       c->sl.toff + slot,
                                                                       c->sl.ref + (size_t)slot * (c->g.nnr + 2), 1, nrc, drcv, dout);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, (size_t)nrc * sizeof(float), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(c, FMM3D_E_DEVICE, "engine receivers: %s", hipGetErrorString(e));
    return FMM3D_OK;
}

// One slot's rays and Frechet rows for nrc receivers, into host vectors
// (rp: nrc+1 offsets); capacity grows to the exact size on a second try.
int fmm3d_engine_frechet_many(fmm3d_ctx* c, int n, const int* slots, const int* nrc, const double* const* rcv,
                              std::vector<int64_t>& rp, std::vector<int32_t>& cols, std::vector<float>& vals)
{
    if (!c || n < 0) return fail(c, FMM3D_E_ARG, "engine frechet: bad arguments");
    std::vector<int> islot;
    std::vector<double> rc;
    for (int k = 0; k < n; ++k) {
        if (slots[k] < 0 || slots[k] >= c->sl.cap || nrc[k] < 0) return fail(c, FMM3D_E_ARG, "engine frechet: bad slot");
        islot.insert(islot.end(), (size_t)nrc[k], slots[k]);
        rc.insert(rc.end(), rcv[k], rcv[k] + 3 * (size_t)nrc[k]);
    }
    const int nrays = (int)islot.size();
    rp.assign((size_t)nrays + 1, 0);
    if (nrays == 0) { cols.clear(); vals.clear(); return FMM3D_OK; }
    int64_t cap = std::max<int64_t>((int64_t)cols.capacity(), (int64_t)nrays * 512 + 1024);
    for (int attempt = 0; attempt < 2; ++attempt) {
        cols.resize(cap);
        vals.resize(cap);
        const int r = frechet_range(c, nrays, rc.data(), 0, nrays, rp.data(), cols.data(), vals.data(), cap, islot.data());
        if (r == FMM3D_OK) {
            cols.resize(rp[nrays]);
            vals.resize(rp[nrays]);
            return FMM3D_OK;
        }
        if (r != FMM3D_E_CAPACITY) return r;
        cap = rp[nrays];
    }
    return fail(c, FMM3D_E_CAPACITY, "engine frechet: capacity retry failed");
}

int fmm3d_engine_retire(fmm3d_ctx* c, int slot)
{
    if (!c || !c->eng_open || slot < 0 || slot >= c->sl.cap) return fail(c, FMM3D_E_ARG, "engine: bad retire");
    k_fill_int<<<1, 1, 0, c->stream>>>(c->sl.on + slot, 1, 0);
    HIPCHK(c, hipGetLastError());
    c->eng_active[slot] = 0;
    return FMM3D_OK;
}

fmm3d_ctx* fmm3d_engine_shadow(const fmm3d_grid_desc* d, const double* velv, int device)
{
    fmm3d_ctx* c = new fmm3d_ctx();
    grid_from_desc(d, c->g);
    c->device = device;
    c->N = (size_t)d->nnr * d->nnx * d->nnz;
    c->velv_host.assign(velv, velv + (size_t)(d->nvr + 2) * (d->nvx + 2) * (d->nvz + 2));
    c->have_model = true;
    return c;
}

void fmm3d_engine_shadow_free(fmm3d_ctx* c)
{
    delete c;   // holds host memory only
}

int fmm3d_engine_point_sets(const fmm3d_ctx* c, fmm3d_ctx** fine, std::string& err, int n, const double* src,
                            int refine, int radius, int order, std::vector<std::vector<int64_t>>& node,
                            std::vector<std::vector<double>>& tv)
{
    if (!c || !c->have_model) return fail_to(err, FMM3D_E_STATE, "engine: no model");
    if (n < 1 || !src) return fail_to(err, FMM3D_E_ARG, "engine: no point sources");
    if (refine < 1 || refine > 16 || radius < 1 || radius > 64) return fail_to(err, FMM3D_E_ARG, "engine: bad refinement");
    node.assign(n, {});
    tv.assign(n, {});
    const std::vector<int> rad(n, radius);
    // in order, groups of at most 32 sources and 128 M context nodes
    std::vector<int> members;
    std::array<int, 3> m{ 0, 0, 0 };
    for (int q = 0; q <= n; ++q) {
        std::array<int, 3> sh{ 0, 0, 0 }, mm = m;
        bool join = false;
        if (q < n) {
            sh = block_shape(c->g, src + 3 * q, radius, refine);
            for (int a = 0; a < 3; ++a) mm[a] = std::max(m[a], sh[a]);
            join = members.empty() ||
                   (members.size() < 32 && (double)(members.size() + 1) * mm[0] * mm[1] * mm[2] <= 128.0e6);
        }
        if (!join && !members.empty()) {
            const int rc = point_group(c, fine, err, members, src, rad.data(), refine, order, node.data(), tv.data());
            if (rc) return rc;
            members.clear();
            mm = sh;
        }
        if (q < n) {
            members.push_back(q);
            m = mm;
        }
    }
    return FMM3D_OK;
}

size_t fmm3d_engine_model_bytes(const fmm3d_grid_desc* d)
{
    const size_t N = (size_t)d->nnr * d->nnx * d->nnz;
    const size_t NH = (size_t)(d->nnr + 3) * (d->nnx + 3) * (d->nnz + 3);
    return NH * 8 + N * 12 + (size_t)d->nnr * d->nnx * 12 + ((size_t)64 << 20);   // + refined sub-grid allowance
}

size_t fmm3d_engine_slot_bytes(const fmm3d_grid_desc* d, int precision)
{
    const size_t N = (size_t)d->nnr * d->nnx * d->nnz;
    const size_t nbmax = (size_t)((d->nnr + 3) / 4) * ((d->nnx + 3) / 4) * ((d->nnz + 7) / 8);
    return N * (precision == FMM3D_PREC_FP64 ? 8 : 4) + (N + 31) / 32 * 4 + nbmax * 8 + 64;
}

extern "C" int fmm3d_set_option(fmm3d_ctx* c, const char* key, double v)
{
    if (!c || !key) return fail(c, FMM3D_E_ARG, "fmm3d_set_option: null argument");
    const int iv = (int)v;
    if (!strcmp(key, "brick_layers")) {
        if (iv < 4 || iv > 4096) return fail(c, FMM3D_E_ARG, "brick_layers %d out of range", iv);
        c->brick_sl = iv;
    } else if (!strcmp(key, "patch_x") || !strcmp(key, "patch_z")) {
        int px = !strcmp(key, "patch_x") ? iv : c->tile_x, pz = !strcmp(key, "patch_z") ? iv : c->tile_z;
        c->tile_x = px; c->tile_z = pz;   // validated at the next solve
    } else if (!strcmp(key, "lateral_iters")) {
        if (iv < 1 || iv > 256) return fail(c, FMM3D_E_ARG, "lateral_iters %d out of range", iv);
        c->kin = iv;
    } else if (!strcmp(key, "tol_fp64")) {
        if (!(v >= 0.0 && v < 1e-3)) return fail(c, FMM3D_E_ARG, "tol_fp64 %g out of range", v);
        c->tol64 = v;
    } else if (!strcmp(key, "tol_fp32")) {
        if (!(v >= 0.0 && v < 1e-2)) return fail(c, FMM3D_E_ARG, "tol_fp32 %g out of range", v);
        c->tol32 = v;
    } else if (!strcmp(key, "direction")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "direction must be 0 or 1");
        c->dirmode = iv;
    } else if (!strcmp(key, "strategy")) {
        if (iv < 0 || iv > 2) return fail(c, FMM3D_E_ARG, "strategy must be 0, 1 or 2");
        c->strategy = iv;
        c->tune_fixed = true;
    } else if (!strcmp(key, "plane_sweeps")) {
        if (iv < 0 || iv > 1000) return fail(c, FMM3D_E_ARG, "plane_sweeps must be in 0..1000");
        c->plane_sweeps = iv;
    } else if (!strcmp(key, "plane_chunk")) {
        if (iv != 2 && iv != 4 && iv != 8) return fail(c, FMM3D_E_ARG, "plane_chunk must be 2, 4 or 8");
        c->plane_ch = iv;
    } else if (!strcmp(key, "column_x") || !strcmp(key, "column_z")) {
        if (iv != 8 && iv != 16 && iv != 32 && iv != 64) return fail(c, FMM3D_E_ARG, "%s must be 8, 16, 32 or 64", key);
        (key[7] == 'x' ? c->col_px : c->col_pz) = iv;
        c->tune_fixed = true;
    } else if (!strcmp(key, "ordered")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "ordered must be 0 or 1");
        c->ordered = iv;
    } else if (!strcmp(key, "ordered_sweeps")) {
        if (iv < 0 || iv > 1000000) return fail(c, FMM3D_E_ARG, "ordered_sweeps out of range");
        c->ordered_until = iv;
    } else if (!strcmp(key, "layer_guess")) {
        if (iv < 0 || iv > 4) return fail(c, FMM3D_E_ARG, "layer_guess must be 0 .. 4");
        c->layer_guess = iv;
    } else if (!strcmp(key, "verify_filter")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "verify_filter must be 0 or 1");
        c->verify_filter = iv;
    } else if (!strcmp(key, "verify_patches")) {
        if (iv < -1 || iv > 1) return fail(c, FMM3D_E_ARG, "verify_patches must be -1, 0 or 1");
        c->verify_patches = iv;
    } else if (!strcmp(key, "verify")) {
        if (iv < 0 || iv > 1000) return fail(c, FMM3D_E_ARG, "verify must be in 0..1000");
        c->verify = iv;
    } else if (!strcmp(key, "check_every")) {
        // at most 8: the per-launch timing events hold one chunk of 8 launches
        if (iv < 1 || iv > 8) return fail(c, FMM3D_E_ARG, "check_every must be in 1..8");
        c->check_every = iv;
        c->tune_fixed = true;
    } else if (!strcmp(key, "spin_limit")) {
        if (iv < 0 || iv > 10000000) return fail(c, FMM3D_E_ARG, "spin_limit out of range");
        c->spin_limit = iv;
    } else if (!strcmp(key, "autotune")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "autotune must be 0 or 1");
        c->autotune = iv;
        if (!iv) c->tuned.clear();
    } else if (!strcmp(key, "ref_qbits")) {
        if (iv != 0 && (iv < 20 || iv > 30)) return fail(c, FMM3D_E_ARG, "ref_qbits must be 0 or 20..30");
        c->ref_qbits = iv;
    } else if (!strcmp(key, "pass_tol")) {
        if (!(v >= 0.0 && v <= 1e-3)) return fail(c, FMM3D_E_ARG, "pass_tol %g out of range", v);
        c->pass_tol = v;
    } else if (!strcmp(key, "tol_floor")) {
        if (!(v >= 0.0 && v <= 1e4)) return fail(c, FMM3D_E_ARG, "tol_floor %g out of range", v);
        c->tol_floor = v;
    } else if (!strcmp(key, "tele_reference")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "tele_reference must be 0 or 1");
        c->tele_ref = iv;
    } else if (!strcmp(key, "graph")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "graph must be 0 or 1");
        c->use_graph = iv;
    } else if (!strcmp(key, "point_group_mnodes")) {
        if (iv < 1 || iv > 1024) return fail(c, FMM3D_E_ARG, "point_group_mnodes must be in 1..1024");
        c->point_group_mnodes = iv;
    } else if (!strcmp(key, "point_workers")) {
        if (iv < 1 || iv > 16) return fail(c, FMM3D_E_ARG, "point_workers must be in 1..16");
        c->point_workers = iv;
    } else if (!strcmp(key, "point_warm")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "point_warm must be 0 or 1");
        c->point_warm = iv;
    } else if (!strcmp(key, "point_window")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "point_window must be 0 or 1");
        c->point_window = iv;
    } else if (!strcmp(key, "point_order")) {
        if (iv != 0 && iv != 1) return fail(c, FMM3D_E_ARG, "point_order must be 0 or 1");
        c->point_order = iv;
    } else if (!strcmp(key, "max_sweeps")) {
        if (iv < 0) return fail(c, FMM3D_E_ARG, "max_sweeps must be >= 0");
        c->max_sweeps = iv;
    } else if (!strcmp(key, "column_passes")) {
        if (iv < 1 || iv > 16) return fail(c, FMM3D_E_ARG, "column_passes must be in 1..16");
        c->tune_fixed = true;
        c->col_passes = iv;
    } else if (!strcmp(key, "timing")) {
        c->timing = iv != 0;
    } else {
        return fail(c, FMM3D_E_ARG, "fmm3d_set_option: unknown key '%s'", key);
    }
    return FMM3D_OK;
}

extern "C" int fmm3d_get_passes(const fmm3d_ctx* c, int nsrc, int* passes)
{
    if (!c || !passes) return fail(const_cast<fmm3d_ctx*>(c), FMM3D_E_ARG, "fmm3d_get_passes: null argument");
    if (!c->part_last || !c->part || (int)c->part->passes.size() != nsrc || c->sl.n != nsrc)
        return fail(const_cast<fmm3d_ctx*>(c), FMM3D_E_STATE,
                    "fmm3d_get_passes: the last solve was not a reference-partition solve of %d sources", nsrc);
    for (int s = 0; s < nsrc; ++s) passes[s] = c->part->passes[s];
    return FMM3D_OK;
}

extern "C" int fmm3d_get_stats(const fmm3d_ctx* c, fmm3d_stats* out)
{
    if (!c || !out) return FMM3D_E_ARG;
    *out = c->stats;
    return FMM3D_OK;
}
