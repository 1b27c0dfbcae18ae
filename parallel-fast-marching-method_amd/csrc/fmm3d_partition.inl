// fmm3d_partition.inl -- the reference's 2x2x2 partitioned teleseismic solve
// on the device (FMM3D_REFERENCE_PARTITION).  Included by fmm3d_solver.hip.
//
// fm3dt_parallel splits the grid over a 2x2x2 layout of ranks (F:6127-6171)
// and, per source, repeats passes (F:6397-6435) until the largest change of
// a pass is at most the float32 literal 1e-5 (F:6403):
//   1. every rank that marched in the previous pass sends its two planes next
//      to each face; a receiving rank relaxes its face nodes from them, one
//      face after the other, each face node once, in the order of a heap
//      keyed by the near ghost value (order_overlap + fouds2_<face>,
//      F:157-386, 1100-2825);
//   2. if any face node was accepted, the rank reclassifies its nodes
//      against the smallest and largest accepted value (F:387-415) and
//      marches its heap to empty (local_fmm, F:78-155), with stencils that
//      end at the rank's own faces (F:708-710).
// Pass 0 is start_fmm: the bottom face is injected and the bottom rank(s)
// with the smallest face minimum march from it (F:6894-6961).
//
// Here all eight blocks of every source live in one field per source, and
// the two sequential parts run in parallel forms that give the same values:
//   - face relaxation (k_pt_face): in heap order a face node reads the new
//     value of each in-plane neighbour (first or second) with a smaller key
//     that is in the heap, the old value of every other; so it can be
//     computed once those neighbours are done.  One workgroup per
//     (source, block) runs rounds: every ready node computes from the field
//     as it stands, then all of them apply.  Key ties go by face index.
//   - the march (k_pt_walk): the pop-order update.  The march freezes the
//     nodes alive at its start and gives every other node the minimum of its
//     starting value and the octant candidates computed each time a
//     neighbour pops, from the neighbours alive then.  A neighbour pops
//     before the node iff its value is below the node's value so far: the
//     update walks the known neighbours in increasing value, one joins while
//     it is below the running minimum, and after each join the octants
//     through it give a candidate.  The result depends on the neighbours
//     only, so Jacobi passes cannot cycle, and they settle where the march
//     does.  One wavefront per 8 x 8 patch column of a block walks the
//     block's layers, alternating direction per sweep, and sweeps repeat
//     until one changes nothing.
// The CPU oracle restates both forms (or_relax_tele_part) and checks them
// against the heap forms (or_fmm_tele_part, which reproduces the reference's
// rtravel.out bit for bit).
//
// Node velocities and metric terms use each block's own offset arithmetic,
// as the ranks compute them (F:540-543, 6989-7038), so the float64 values
// are the reference's.  Unknown times are +inf (the reference's 0); an
// unknown ghost enters a face update as 0, as in the reference.

namespace {

constexpr int PT_FACE_AXIS[6] = { 1, 1, 2, 2, 0, 0 };   // faces: left, right, in, out, up, down
constexpr int PT_FACE_SIDE[6] = { -1, 1, -1, 1, -1, 1 };
constexpr int PT_P = 8;                                 // patch edge of the block walk
constexpr int PT_FACE_THREADS = 1024;

// Block layout: per axis the first index of the second block.  The first
// block takes the extra node of an odd count (the reference's rank sizes).
struct PartGeo {
    int n[3];
    int cut[3];
    __host__ __device__ int lo(int a, int b) const { return b ? cut[a] : 0; }
    __host__ __device__ int len(int a, int b) const { return b ? n[a] - cut[a] : cut[a]; }
};

__host__ __device__ inline void pt_coord(int blk, int c[3]) { c[0] = blk >> 2; c[1] = (blk >> 1) & 1; c[2] = blk & 1; }

// the in-plane axes of face k, (outer, inner) in the reference's loop order
__host__ __device__ inline void pt_face_axes(int k, int& outer, int& inner)
{
    const int a = PT_FACE_AXIS[k];
    if (a == 1) { outer = 0; inner = 2; }
    else if (a == 2) { outer = 0; inner = 1; }
    else { outer = 1; inner = 2; }
}

static PartGeo part_geo(const Grid& g)
{
    PartGeo p;
    p.n[0] = g.nnr; p.n[1] = g.nnx; p.n[2] = g.nnz;
    for (int a = 0; a < 3; ++a) p.cut[a] = (p.n[a] + 1) / 2;
    return p;
}

}  // namespace

// Device state of the partitioned mode, owned by the context.
struct PartState {
    int cap = 0;                    // slots the arrays hold
    size_t nfmax = 0;               // largest face of a block (nodes)
    double* Tprev = nullptr;        // field at the start of the pass (ghosts, change)
    double* T0 = nullptr;           // march start: -value for heap nodes, value otherwise
    uint32_t* hbits = nullptr;      // nodes accepted by this pass's face relaxation (the heap)
    double* fnew = nullptr;         // face relaxation: per (slot, block) values of a round
    int* on = nullptr;              // per slot: passes continue
    int* dsend = nullptr;           // per slot x block: marched in the previous pass
    int* march = nullptr;           // per slot x block: marches in this pass
    double* mini = nullptr;         // per slot x block: smallest heap value
    int* stamp = nullptr;           // per slot x block x patch: last sweep that changed a value
    int* chg = nullptr;             // per slot: the sweep changed a value
    int* pstart = nullptr;          // per slot: first sweep of its current pass
    unsigned long long* err = nullptr;   // per slot: largest change of the pass (bits of a double)
    int* seed = nullptr;            // pass 0: (slot, node) pairs of the heap seeds
    int nseed_cap = 0;
    double* slo = nullptr;          // slowness, block arithmetic
    double* hx = nullptr;           // nnr
    double* hz = nullptr;           // nnr x nnx
    bool model = false;             // slo / hx / hz are those of the resident model
    int* h_chg = nullptr;           // pinned host copies
    unsigned long long* h_err = nullptr;
    std::vector<int> passes;        // per slot: passes of the last solve
};

namespace {

// Node velocity with a block's offset arithmetic (F:6989-7038 on a rank's
// local indices); slowness in solver layout.
__global__ void k_pt_gridder(Grid g, PartGeo p, const double* __restrict__ velv, double* __restrict__ slo)
{
    const long total = (long)g.nnr * g.nnx * g.nnz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total; i += (long)gridDim.x * blockDim.x) {
        const int kz = (int)(i % g.nnz), kx = (int)((i / g.nnz) % g.nnx), kr = (int)(i / ((long)g.nnz * g.nnx));
        const int k3[3] = { kr, kx, kz };
        int off[3], l[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            off[a] = k3[a] >= p.cut[a] ? p.cut[a] + 1 : 1;    // 1-based global index of local index 1
            l[a] = k3[a] - off[a] + 2;                         // 1-based local index
        }
        const double h = g.rgor - (off[0] - 1) * g.dnr - (l[0] - 1) * g.dnr;
        const double hr = g.gor - g.rgor + (off[0] - 1) * g.dnr + (l[0] - 1) * g.dnr;
        const double th = g.rgox + (off[1] - 1) * g.dnx + (l[1] - 1) * g.dnx;
        const double ph = g.rgoz + (off[2] - 1) * g.dnz + (l[2] - 1) * g.dnz;
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
            const int iz = min(max(zi - 1 + a, 0), NZ - 1);
            double acc_x = 0.0;
            for (int b = 0; b < 4; ++b) {
                const int ix = min(max(xi - 1 + b, 0), NX - 1);
                const double* col = velv + ((size_t)iz * NX + ix) * NR;
                double acc_r = 0.0;
                for (int c = 0; c < 4; ++c) acc_r += wr[c] * col[min(max(ri - 1 + c, 0), NR - 1)];
                acc_x += wx[b] * acc_r;
            }
            acc_z += wz[a] * acc_x;
        }
        slo[i] = 1.0 / acc_z;
    }
}

// Bottom-face injection (start_fmm, F:6914-6939): values > 0 known, the
// rest unknown; every other node unknown.
__global__ void k_pt_inject(int nnr, int nnx, int nnz, int nslots, const double* __restrict__ faces, double* __restrict__ T)
{
    const size_t NL = (size_t)nnx * nnz, N = NL * nnr;
    const size_t total = (size_t)nslots * N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t s = i / N, q = i - s * N;
        double v = __builtin_huge_val();
        if (q >= (size_t)(nnr - 1) * NL) {
            const double f = faces[s * NL + (q - (size_t)(nnr - 1) * NL)];
            if (f > 0.0) v = f;
        }
        T[i] = v;
    }
}

// heap membership of the pass-0 seeds: (slot, node) pairs
__global__ void k_pt_setbits(int n, const int* __restrict__ sn, size_t words, uint32_t* __restrict__ bits)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t node = (size_t)sn[2 * i + 1];
    atomicOr(bits + (size_t)sn[2 * i] * words + (node >> 5), 1u << (node & 31));
}

struct PartArgs {
    double* T;
    double* Tprev;
    double* T0;
    uint32_t* hbits;
    const double* slo;
    const double* hx;
    const double* hz;
    double hr;             // radial spacing
    PartGeo p;
    int nslots;
    int fom;
    size_t N, words;
    const int* on;
    int* dsend;
    int* march;
    double* mini;
    int* stamp;
    int* chg;
    int* pstart;
    unsigned long long* err;
    double* fnew;
    size_t nfmax;
    int npx, npz;          // patches per block per axis (the larger block's)
    int iter;              // sweep number
    int maxpass;           // Jacobi passes per layer at most
    unsigned long long* updates;
};

__device__ __forceinline__ bool pt_hbit(const uint32_t* hb, size_t q) { return (hb[q >> 5] >> (q & 31)) & 1u; }

__device__ __forceinline__ size_t pt_idx(const PartGeo& p, int r, int x, int z)
{
    return ((size_t)r * p.n[1] + x) * p.n[2] + z;
}

// ---------------------------------------------------------------- face relaxation

// One face node's value from ghost face k (F:1100-1400): the ghost side of
// the face axis always counts (second order iff near ghost > far ghost), the
// other side of that axis is not used; an in-plane neighbour counts iff known
// and below the node (or the node unknown), second order iff its second
// neighbour is known and below it; face updates are mixed order whatever the
// .in file says (F:197-383).  Neighbours beyond the block are out of bounds.
__device__ double pt_face_value(const PartArgs& a, const double* __restrict__ T, int k, const int c[3],
                                const int lo[3], const int len[3], double gn, double gf)
{
    const double inf = __builtin_huge_val();
    const int A = PT_FACE_AXIS[k], gs = PT_FACE_SIDE[k] > 0 ? 1 : 0;
    const size_t me = pt_idx(a.p, c[0], c[1], c[2]);
    const double tc = T[me];
    Side<double> sd[3][2];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax)
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            Side<double>& S = sd[ax][s];
            S.ord = 0; S.t1 = 0.0; S.t2 = 0.0;
            if (ax == A) {
                if (s != gs) { S.ord = -1; continue; }
                S.t1 = gn;
                if (gn > gf) { S.ord = 2; S.t2 = gf; } else S.ord = 1;
                continue;
            }
            const int d = s ? 1 : -1;
            int p1[3] = { c[0], c[1], c[2] }, p2[3] = { c[0], c[1], c[2] };
            p1[ax] += d; p2[ax] += 2 * d;
            if (p1[ax] < lo[ax] || p1[ax] >= lo[ax] + len[ax]) { S.ord = -1; continue; }
            const double v1 = T[pt_idx(a.p, p1[0], p1[1], p1[2])];
            if (!(v1 < inf && v1 < tc)) continue;
            S.ord = 1; S.t1 = v1;
            if (p2[ax] >= lo[ax] && p2[ax] < lo[ax] + len[ax]) {
                const double v2 = T[pt_idx(a.p, p2[0], p2[1], p2[2])];
                if (v2 < inf && v1 > v2) { S.ord = 2; S.t2 = v2; }
            }
        }
    const double hx = a.hx[c[0]], hz = a.hz[(size_t)c[0] * a.p.n[1] + c[1]];
    return octant_min<double>(sd, a.hr, hx, hz, a.slo[me]);
}

// Face k of every block whose neighbour across it marched in the previous
// pass: one workgroup per (slot, block).  Dynamic LDS: a state byte
// (0 not in the heap, 1 waiting, 2 computed this round, 3 done) and the
// dependency mask (in-plane heap neighbours of smaller key) per face node.
__global__ void __launch_bounds__(PT_FACE_THREADS) k_pt_face(PartArgs a, int k)
{
    extern __shared__ unsigned char pt_lds[];
    const int slot = blockIdx.x >> 3, blk = blockIdx.x & 7;
    if (slot >= a.nslots || !a.on[slot]) return;
    int bc[3];
    pt_coord(blk, bc);
    const int A = PT_FACE_AXIS[k], side = PT_FACE_SIDE[k];
    int nc[3] = { bc[0], bc[1], bc[2] };
    nc[A] += side;
    if (nc[A] < 0 || nc[A] > 1) return;
    const int nb = (nc[0] << 2) | (nc[1] << 1) | nc[2];
    if (!a.dsend[slot * 8 + nb]) return;
    int lo[3], len[3];
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) { lo[ax] = a.p.lo(ax, bc[ax]); len[ax] = a.p.len(ax, bc[ax]); }
    int outer, inner;
    pt_face_axes(k, outer, inner);
    const int no = len[outer], ni = len[inner], nf = no * ni;
    const int fa = side < 0 ? lo[A] : lo[A] + len[A] - 1;     // the face plane
    unsigned char* state = pt_lds;
    unsigned char* dep = pt_lds + a.nfmax;
    double* T = a.T + (size_t)slot * a.N;
    const double* P = a.Tprev + (size_t)slot * a.N;
    double* fnew = a.fnew + ((size_t)slot * 8 + blk) * a.nfmax;
    uint32_t* hb = a.hbits + (size_t)slot * a.words;
    const double inf = __builtin_huge_val();
    const int t = threadIdx.x;
    auto node_of = [&](int fi, int plane, int c[3]) {
        c[A] = plane; c[outer] = lo[outer] + fi / ni; c[inner] = lo[inner] + fi % ni;
    };
    auto key_of = [&](int fi) {
        int c[3];
        node_of(fi, fa + side, c);
        const double v = P[pt_idx(a.p, c[0], c[1], c[2])];
        return v < inf ? v : 0.0;
    };
    // the heap of this face: nodes whose near ghost is below them (or unknown nodes)
    for (int fi = t; fi < nf; fi += PT_FACE_THREADS) {
        int c[3];
        node_of(fi, fa, c);
        const double tf = T[pt_idx(a.p, c[0], c[1], c[2])];
        state[fi] = key_of(fi) < tf ? 1 : 0;
    }
    __syncthreads();
    // dependencies: in-plane heap neighbours (first and second) of smaller (key, index)
    const int NBo[8] = { -1, 1, -2, 2, 0, 0, 0, 0 }, NBi[8] = { 0, 0, 0, 0, -1, 1, -2, 2 };
    for (int fi = t; fi < nf; fi += PT_FACE_THREADS) {
        if (!state[fi]) { dep[fi] = 0; continue; }
        const int io = fi / ni, ii = fi % ni;
        const double ki = key_of(fi);
        unsigned m = 0;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const int jo = io + NBo[q], ji = ii + NBi[q];
            if (jo < 0 || jo >= no || ji < 0 || ji >= ni) continue;
            const int fj = jo * ni + ji;
            if (!state[fj]) continue;
            const double kj = key_of(fj);
            if (kj < ki || (kj == ki && fj < fi)) m |= 1u << q;
        }
        dep[fi] = (unsigned char)m;
    }
    __syncthreads();
    for (int round = 0; round <= nf; ++round) {
        int did = 0;
        for (int fi = t; fi < nf; fi += PT_FACE_THREADS) {
            if (state[fi] != 1) continue;
            const int io = fi / ni, ii = fi % ni;
            const unsigned m = dep[fi];
            bool ready = true;
#pragma unroll
            for (int q = 0; q < 8; ++q)
                if ((m >> q) & 1u) ready = ready && state[(io + NBo[q]) * ni + ii + NBi[q]] == 3;
            if (!ready) continue;
            int c[3], cg[3];
            node_of(fi, fa, c);
            node_of(fi, fa + side, cg);
            const double gn0 = P[pt_idx(a.p, cg[0], cg[1], cg[2])];
            node_of(fi, fa + 2 * side, cg);
            const double gf0 = P[pt_idx(a.p, cg[0], cg[1], cg[2])];
            fnew[fi] = pt_face_value(a, T, k, c, lo, len, gn0 < inf ? gn0 : 0.0, gf0 < inf ? gf0 : 0.0);
            state[fi] = 2;
            did = 1;
        }
        // every value of the round is computed from the field before it
        if (!__syncthreads_or(did)) break;
        for (int fi = t; fi < nf; fi += PT_FACE_THREADS) {
            if (state[fi] != 2) continue;
            int c[3];
            node_of(fi, fa, c);
            const size_t q = pt_idx(a.p, c[0], c[1], c[2]);
            const double old = T[q], best = fnew[fi];
            // acceptance as F:1378, strict on the down face (F:2817); unknown old accepts
            const bool acc = (k == 5) ? (old > best || !(old < inf)) : (old >= best || !(old < inf));
            if (acc) {
                T[q] = best;
                atomicOr(hb + (q >> 5), 1u << (q & 31));
            }
            state[fi] = 3;
        }
        __threadfence_block();
        __syncthreads();
    }
}

// Per (slot, block): whether any node is in the heap (the block marches this
// pass) and the smallest heap value (F:387-389).  Heap nodes lie on the
// block's faces.
__global__ void __launch_bounds__(256) k_pt_mark(PartArgs a)
{
    __shared__ double s_min[256];
    __shared__ int s_any[256];
    const int slot = blockIdx.x >> 3, blk = blockIdx.x & 7;
    if (slot >= a.nslots || !a.on[slot]) return;
    int bc[3], lo[3], len[3];
    pt_coord(blk, bc);
    for (int ax = 0; ax < 3; ++ax) { lo[ax] = a.p.lo(ax, bc[ax]); len[ax] = a.p.len(ax, bc[ax]); }
    const double* T = a.T + (size_t)slot * a.N;
    const uint32_t* hb = a.hbits + (size_t)slot * a.words;
    double mn = __builtin_huge_val();
    int any = 0;
    for (int f = 0; f < 6; ++f) {
        const int A = PT_FACE_AXIS[f];
        int outer, inner;
        pt_face_axes(f, outer, inner);
        const int plane = PT_FACE_SIDE[f] < 0 ? lo[A] : lo[A] + len[A] - 1;
        const int nf = len[outer] * len[inner];
        for (int fi = threadIdx.x; fi < nf; fi += 256) {
            int c[3];
            c[A] = plane; c[outer] = lo[outer] + fi / len[inner]; c[inner] = lo[inner] + fi % len[inner];
            const size_t q = pt_idx(a.p, c[0], c[1], c[2]);
            if (pt_hbit(hb, q)) { any = 1; mn = fmin(mn, T[q]); }
        }
    }
    s_min[threadIdx.x] = mn;
    s_any[threadIdx.x] = any;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            s_min[threadIdx.x] = fmin(s_min[threadIdx.x], s_min[threadIdx.x + w]);
            s_any[threadIdx.x] |= s_any[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.march[slot * 8 + blk] = s_any[0];
        a.mini[slot * 8 + blk] = s_min[0];
    }
}

// The march's starting state for the marching blocks: T0 = -value for heap
// nodes (they pop), +value otherwise (+inf unknown).
__global__ void k_pt_start(PartArgs a)
{
    const size_t total = (size_t)a.nslots * a.N;
    const size_t NL = (size_t)a.p.n[1] * a.p.n[2];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t s = i / a.N, q = i - s * a.N;
        if (!a.on[s]) continue;
        const int r = (int)(q / NL), x = (int)((q / a.p.n[2]) % a.p.n[1]), z = (int)(q % a.p.n[2]);
        const int blk = ((r >= a.p.cut[0]) << 2) | ((x >= a.p.cut[1]) << 1) | (z >= a.p.cut[2]);
        if (!a.march[s * 8 + blk]) continue;
        const double v = a.T[i];
        a.T0[i] = pt_hbit(a.hbits + s * a.words, q) ? -v : v;
    }
}

// ---------------------------------------------------------------- the march as a relaxation

// Pop-order update (see the file comment).  Sides r-, r+, x-, x+, z-, z+:
// v1/v2 first and second neighbour values (+inf unknown), in1/in2 inside the
// block; t0 the starting value (+inf unknown).  Second order on a side iff
// its second neighbour is known and below the first (F:645-707).
__device__ __forceinline__ double pt_pop_update(int fom, double t0, const double v1[6], const double v2[6],
                                                const bool in1[6], const bool in2[6], double hr, double hx,
                                                double hz, double slown)
{
    const double inf = __builtin_huge_val();
    bool sec[6];
    int rank[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) sec[q] = fom && in1[q] && in2[q] && v1[q] < inf && v2[q] < inf && v1[q] > v2[q];
    // rank of each known side in (value, side) order; unknown sides rank 6
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        int r = 0;
#pragma unroll
        for (int p = 0; p < 6; ++p)
            if (p != q) r += (in1[p] && v1[p] < inf && (v1[p] < v1[q] || (v1[p] == v1[q] && p < q))) ? 1 : 0;
        rank[q] = (in1[q] && v1[q] < inf) ? r : 6;
    }
    Side<double> sd[3][2];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
        sd[q / 2][q % 2].ord = in1[q] ? 0 : -1;
        sd[q / 2][q % 2].t1 = v1[q];
        sd[q / 2][q % 2].t2 = v2[q];
    }
    double cur = t0;
#pragma unroll 1
    for (int i = 0; i < 6; ++i) {
        int j = -1;
        double vj = inf;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (rank[q] == i) { j = q; vj = v1[q]; }
        if (j < 0 || !(vj < cur)) break;
#pragma unroll
        for (int q = 0; q < 6; ++q)
            if (q == j) sd[q / 2][q % 2].ord = sec[q] ? 2 : 1;
        // the octants through side j
#pragma unroll
        for (int i0 = 0; i0 < 2; ++i0)
#pragma unroll
            for (int i1 = 0; i1 < 2; ++i1)
#pragma unroll
                for (int i2 = 0; i2 < 2; ++i2) {
                    if (j != i0 && j != 2 + i1 && j != 4 + i2) continue;
                    const Side<double>& s0 = sd[0][i0];
                    const Side<double>& s1 = sd[1][i1];
                    const Side<double>& s2 = sd[2][i2];
                    if (s0.ord < 0 || s1.ord < 0 || s2.ord < 0) continue;
                    const double Tv = octant_solve<double>(s0, s1, s2, hr, hx, hz, slown);
                    if (Tv < cur) cur = Tv;
                }
    }
    return cur;
}

// One sweep of the march relaxation over the marching blocks: a wavefront
// per (slot, block, 8 x 8 patch column).  Per layer it stages the patch with
// a two-node ring (inside the block) and relaxes it with Jacobi passes until
// a pass changes nothing; the two layers walked just before stay in LDS.  A
// patch column runs in the first sweep of a pass, or when it or a lateral
// neighbour changed a value in the sweep before.
__global__ void __launch_bounds__(64) k_pt_walk(PartArgs a)
{
    constexpr int LX = PT_P + 4, LZ = PT_P + 5;
    __shared__ double cur[LX * LZ];
    __shared__ double b1[64], b2[64], u1[64], u2[64], t0s[64], sl[64];
    __shared__ double shz[PT_P];
    __shared__ int s_act;
    const int ppb = a.npx * a.npz;
    int w = blockIdx.x;
    const int slot = w / (8 * ppb);
    w -= slot * 8 * ppb;
    const int blk = w / ppb, pi = w - blk * ppb;
    if (slot >= a.nslots || !a.on[slot] || !a.march[slot * 8 + blk]) return;
    const int pxi = pi / a.npz, pzi = pi % a.npz;
    int bc[3], lo[3], len[3];
    pt_coord(blk, bc);
#pragma unroll
    for (int ax = 0; ax < 3; ++ax) { lo[ax] = a.p.lo(ax, bc[ax]); len[ax] = a.p.len(ax, bc[ax]); }
    const int npxb = (len[1] + PT_P - 1) / PT_P, npzb = (len[2] + PT_P - 1) / PT_P;
    if (pxi >= npxb || pzi >= npzb) return;
    const int iter = a.iter;
    const bool first = iter == a.pstart[slot];
    int* st = a.stamp + ((size_t)slot * 8 + blk) * ppb;
    const int t = threadIdx.x;
    if (t == 0) {
        int m = st[pi];
        if (pxi > 0) m = max(m, st[pi - a.npz]);
        if (pxi < npxb - 1) m = max(m, st[pi + a.npz]);
        if (pzi > 0) m = max(m, st[pi - 1]);
        if (pzi < npzb - 1) m = max(m, st[pi + 1]);
        s_act = first || m >= iter - 1;
    }
    __syncthreads();
    if (!s_act) return;
    const double inf = __builtin_huge_val();
    const size_t NL = (size_t)a.p.n[1] * a.p.n[2];
    double* T = a.T + (size_t)slot * a.N;
    const double* T0 = a.T0 + (size_t)slot * a.N;
    const double mini = a.mini[slot * 8 + blk];
    const bool down = ((iter - a.pstart[slot]) & 1) != 0;   // the first sweep of a pass walks upward
    const int dn = down ? 1 : -1;
    const int x0 = lo[1] + pxi * PT_P, z0 = lo[2] + pzi * PT_P;
    const int xe = lo[1] + len[1], ze = lo[2] + len[2];        // block ends (exclusive)
    const int r0 = lo[0], re = lo[0] + len[0];
    const int lx = t / PT_P, lz = t % PT_P;
    const int gx = x0 + lx, gz = z0 + lz;
    const bool in = gx < xe && gz < ze;
    const int own = (lx + 2) * LZ + lz + 2;
    bool anyv = false;
    unsigned long long nupd = 0;
    for (int q = 0; q < len[0]; ++q) {
        const int ir = down ? r0 + q : re - 1 - q;
        const size_t lay = (size_t)ir * NL;
        const size_t c = lay + (size_t)gx * a.p.n[2] + gz;
        const double v = in ? T[c] : inf;
        cur[own] = v;
        const double orig = v;
        {
            const int n1 = ir + dn, n2 = ir + 2 * dn;
            u1[t] = (in && n1 >= r0 && n1 < re) ? T[c + (ptrdiff_t)dn * (ptrdiff_t)NL] : inf;
            u2[t] = (in && n2 >= r0 && n2 < re) ? T[c + (ptrdiff_t)(2 * dn) * (ptrdiff_t)NL] : inf;
            const double s0 = in ? T0[c] : inf;
            // fixed: alive at the march's start (a known non-heap node at or below the heap minimum)
            const bool fixed = !in || (s0 > 0.0 && s0 < inf && s0 <= mini);
            t0s[t] = s0 < 0.0 ? -s0 : s0;
            sl[t] = fixed ? -1.0 : a.slo[c];
            if (q == 0) {
                const int p1 = ir - dn, p2 = ir - 2 * dn;
                b1[t] = (in && p1 >= r0 && p1 < re) ? T[c - (ptrdiff_t)dn * (ptrdiff_t)NL] : inf;
                b2[t] = (in && p2 >= r0 && p2 < re) ? T[c - (ptrdiff_t)(2 * dn) * (ptrdiff_t)NL] : inf;
            }
        }
        // lateral ring: neighbouring patches of the same block may be writing it
        for (int e = t; e < LX * (PT_P + 4); e += 64) {
            const int hx_ = e / (PT_P + 4), hz_ = e % (PT_P + 4);
            if (hx_ >= 2 && hx_ < PT_P + 2 && hz_ >= 2 && hz_ < PT_P + 2) continue;
            const int ggx = x0 + hx_ - 2, ggz = z0 + hz_ - 2;
            const bool ok = ggx >= lo[1] && ggx < xe && ggz >= lo[2] && ggz < ze;
            cur[hx_ * LZ + hz_] = ok ? __hip_atomic_load(T + lay + (size_t)ggx * a.p.n[2] + ggz, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT) : inf;
        }
        if (t < PT_P) shz[t] = a.hz[(size_t)ir * a.p.n[1] + min(x0 + t, xe - 1)];
        const double hxl = a.hx[ir];
        __syncthreads();
        const double slown = sl[t];
        const double t0 = t0s[t];
        bool in1[6], in2[6];
        in1[0] = ir - 1 >= r0; in2[0] = ir - 2 >= r0;
        in1[1] = ir + 1 < re; in2[1] = ir + 2 < re;
        in1[2] = gx - 1 >= lo[1]; in2[2] = gx - 2 >= lo[1];
        in1[3] = gx + 1 < xe; in2[3] = gx + 2 < xe;
        in1[4] = gz - 1 >= lo[2]; in2[4] = gz - 2 >= lo[2];
        in1[5] = gz + 1 < ze; in2[5] = gz + 2 < ze;
        int npass = 0;
#pragma unroll 1
        for (int pass = 0; pass < a.maxpass; ++pass) {
            bool moved = false;
            double nv = 0.0;
            if (slown >= 0.0) {
                double v1[6], v2[6];
                // side 0 is the layer above (ir-1), side 1 the one below (ir+1)
                v1[0] = down ? b1[t] : u1[t]; v2[0] = down ? b2[t] : u2[t];
                v1[1] = down ? u1[t] : b1[t]; v2[1] = down ? u2[t] : b2[t];
                v1[2] = cur[own - LZ]; v2[2] = cur[own - 2 * LZ];
                v1[3] = cur[own + LZ]; v2[3] = cur[own + 2 * LZ];
                v1[4] = cur[own - 1]; v2[4] = cur[own - 2];
                v1[5] = cur[own + 1]; v2[5] = cur[own + 2];
                nv = pt_pop_update(a.fom, t0, v1, v2, in1, in2, a.hr, hxl, shz[lx], slown);
                moved = nv != cur[own];
            }
            // every lane has read its stencil before any lane writes (one wavefront)
            __builtin_amdgcn_wave_barrier();
            if (moved) cur[own] = nv;
            __builtin_amdgcn_wave_barrier();
            ++npass;
            if (!__ballot(moved)) break;
        }
        nupd += (unsigned long long)npass;
        __syncthreads();
        const double nv = cur[own];
        const bool lay_chg = in && nv != orig;
        if (lay_chg) __hip_atomic_store(T + c, nv, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        b2[t] = b1[t];
        b1[t] = nv;
        anyv |= lay_chg;
        __syncthreads();
    }
    if (t == 0 && nupd) atomicAdd(a.updates, nupd * (unsigned long long)(min(PT_P, xe - x0) * min(PT_P, ze - z0)));
    const bool blkchg = __syncthreads_or(anyv);
    if (t == 0 && blkchg) {
        st[pi] = iter;
        atomicOr(a.chg + slot, 1);
    }
}

// Largest change of the pass per slot: max |T - Tprev| over the grid with
// unknown values taken as 0, the reference's error_calculate (F:6963-6987).
__global__ void __launch_bounds__(256) k_pt_err(PartArgs a)
{
    __shared__ double s_m[256];
    const int slot = blockIdx.y;
    double m = 0.0;
    if (a.on[slot]) {
        const double* T = a.T + (size_t)slot * a.N;
        const double* P = a.Tprev + (size_t)slot * a.N;
        for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < a.N; i += (size_t)gridDim.x * blockDim.x) {
            const double x = T[i] < __builtin_huge_val() ? T[i] : 0.0;
            const double y = P[i] < __builtin_huge_val() ? P[i] : 0.0;
            m = fmax(m, fabs(x - y));
        }
    }
    s_m[threadIdx.x] = m;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) s_m[threadIdx.x] = fmax(s_m[threadIdx.x], s_m[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0 && s_m[0] > 0.0)
        atomicMax(a.err + slot, (unsigned long long)__double_as_longlong(s_m[0]));
}

// per slot: copy the field to Tprev and clear the heap bits (passes continuing)
__global__ void k_pt_pass_begin(PartArgs a)
{
    const size_t total = (size_t)a.nslots * a.N;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        const size_t s = i / a.N;
        if (!a.on[s]) continue;
        a.Tprev[i] = a.T[i];
        const size_t q = i - s * a.N;
        if (q < a.words) a.hbits[s * a.words + q] = 0u;
    }
}

// per slot: the blocks that marched this pass send next pass; the pass's
// first sweep number; the change flags cleared
__global__ void k_pt_pass_end(PartArgs a, int next_iter)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nslots * 8) return;
    const int s = i >> 3;
    if (!a.on[s]) return;
    a.dsend[i] = a.march[i];
    if ((i & 7) == 0) {
        a.pstart[s] = next_iter;
        a.err[s] = 0ull;
    }
}

__global__ void k_pt_fill_int(int* p, size_t n, int v)
{
    const size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_pt_fill_u32(uint32_t* p, size_t n)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

}  // namespace

// ---------------------------------------------------------------- host driver

// the block-arithmetic slowness is recomputed before the next partitioned solve
static void part_model_changed(PartState* ps)
{
    if (ps) ps->model = false;
}

static void part_free(PartState* ps)
{
    if (!ps) return;
    (void)hipFree(ps->Tprev); (void)hipFree(ps->T0); (void)hipFree(ps->hbits); (void)hipFree(ps->fnew);
    (void)hipFree(ps->on); (void)hipFree(ps->dsend); (void)hipFree(ps->march); (void)hipFree(ps->mini);
    (void)hipFree(ps->stamp); (void)hipFree(ps->chg); (void)hipFree(ps->pstart); (void)hipFree(ps->err);
    (void)hipFree(ps->seed); (void)hipFree(ps->slo); (void)hipFree(ps->hx); (void)hipFree(ps->hz);
    (void)hipHostFree(ps->h_chg); (void)hipHostFree(ps->h_err);
    delete ps;
}

// Metric tables with each block's offset arithmetic (F:540-543): a node at
// global index k of block b (local index l = k - off + 2, off its first
// global index) has ri = rgor - (off-1) dnr - (l-1) dnr + earth.
static void part_metrics(const Grid& g, const PartGeo& p, std::vector<double>& hx, std::vector<double>& hz)
{
    hx.assign(g.nnr, 0.0);
    hz.assign((size_t)g.nnr * g.nnx, 0.0);
    for (int kr = 0; kr < g.nnr; ++kr) {
        const int off0 = kr >= p.cut[0] ? p.cut[0] + 1 : 1, l0 = kr - off0 + 2;
        const double ri = g.rgor - (off0 - 1) * g.dnr - (l0 - 1) * g.dnr + g.earth;
        hx[kr] = ri * g.dnx;
        for (int kx = 0; kx < g.nnx; ++kx) {
            const int off1 = kx >= p.cut[1] ? p.cut[1] + 1 : 1, l1 = kx - off1 + 2;
            const double sth = sin(g.rgox + (off1 - 1) * g.dnx + (l1 - 1) * g.dnx);
            hz[(size_t)kr * g.nnx + kx] = ri * sth * g.dnz;
        }
    }
}

static int part_prepare(fmm3d_ctx* c, int nsrc)
{
    const Grid& g = c->g;
    const PartGeo p = part_geo(g);
    if (!c->part) c->part = new PartState();
    PartState& ps = *c->part;
    const size_t N = c->N, words = (N + 31) / 32;
    if (!ps.model) {
        if (!ps.slo) {
            HIPCHK(c, dalloc(&ps.slo, N * sizeof(double)));
            HIPCHK(c, dalloc(&ps.hx, (size_t)g.nnr * sizeof(double)));
            HIPCHK(c, dalloc(&ps.hz, (size_t)g.nnr * g.nnx * sizeof(double)));
            HIPCHK(c, hipHostMalloc((void**)&ps.h_chg, 4096 * sizeof(int), hipHostMallocDefault));
            HIPCHK(c, hipHostMalloc((void**)&ps.h_err, 4096 * sizeof(unsigned long long), hipHostMallocDefault));
        }
        std::vector<double> hx, hz;
        part_metrics(g, p, hx, hz);
        HIPCHK(c, hipMemcpy(ps.hx, hx.data(), hx.size() * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(ps.hz, hz.data(), hz.size() * sizeof(double), hipMemcpyHostToDevice));
        const long total = (long)N;
        k_pt_gridder<<<(int)std::min<long>((total + 255) / 256, 65536), 256, 0, c->stream>>>(g, p, c->d_velv, ps.slo);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipStreamSynchronize(c->stream));
        ps.model = true;
    }
    if (ps.cap < nsrc) {
        (void)hipFree(ps.Tprev); (void)hipFree(ps.T0); (void)hipFree(ps.hbits); (void)hipFree(ps.fnew);
        (void)hipFree(ps.on); (void)hipFree(ps.dsend); (void)hipFree(ps.march); (void)hipFree(ps.mini);
        (void)hipFree(ps.stamp); (void)hipFree(ps.chg); (void)hipFree(ps.pstart); (void)hipFree(ps.err);
        ps.Tprev = ps.T0 = nullptr; ps.hbits = nullptr; ps.fnew = nullptr; ps.on = ps.dsend = ps.march = nullptr;
        ps.mini = nullptr; ps.stamp = ps.chg = ps.pstart = nullptr; ps.err = nullptr;
        ps.cap = 0;
        size_t nf = 0;
        for (int k = 0; k < 6; ++k) {
            int outer, inner;
            pt_face_axes(k, outer, inner);
            nf = std::max(nf, (size_t)p.len(outer, 0) * p.len(inner, 0));
        }
        ps.nfmax = nf;
        if (2 * nf > 65536) return fail(c, FMM3D_E_ARG, "partitioned solve: block faces of %zu nodes exceed 32768", nf);
        const int npx = (p.len(1, 0) + PT_P - 1) / PT_P, npz = (p.len(2, 0) + PT_P - 1) / PT_P;
        HIPCHK(c, dalloc(&ps.Tprev, (size_t)nsrc * N * sizeof(double)));
        HIPCHK(c, dalloc(&ps.T0, (size_t)nsrc * N * sizeof(double)));
        HIPCHK(c, dalloc(&ps.hbits, (size_t)nsrc * words * sizeof(uint32_t)));
        HIPCHK(c, dalloc(&ps.fnew, (size_t)nsrc * 8 * nf * sizeof(double)));
        HIPCHK(c, dalloc(&ps.on, (size_t)nsrc * sizeof(int)));
        HIPCHK(c, dalloc(&ps.dsend, (size_t)nsrc * 8 * sizeof(int)));
        HIPCHK(c, dalloc(&ps.march, (size_t)nsrc * 8 * sizeof(int)));
        HIPCHK(c, dalloc(&ps.mini, (size_t)nsrc * 8 * sizeof(double)));
        HIPCHK(c, dalloc(&ps.stamp, (size_t)nsrc * 8 * npx * npz * sizeof(int)));
        HIPCHK(c, dalloc(&ps.chg, (size_t)nsrc * sizeof(int)));
        HIPCHK(c, dalloc(&ps.pstart, (size_t)nsrc * sizeof(int)));
        HIPCHK(c, dalloc(&ps.err, (size_t)nsrc * sizeof(unsigned long long)));
        ps.cap = nsrc;
    }
    return FMM3D_OK;
}

// The partitioned solve of nsrc bottom faces (host copies in `faces`, device
// copies in c->d_faces) into slots 0 .. nsrc-1 (float64 fields).
static int part_run(fmm3d_ctx* c, int nsrc, const double* faces, int order)
{
    const Grid& g = c->g;
    const PartGeo p = part_geo(g);
    if (nsrc > 4096) return fail(c, FMM3D_E_ARG, "partitioned solve: at most 4096 sources per call");
    for (int a_ = 0; a_ < 3; ++a_)
        if (p.len(a_, 1) < 3)
            return fail(c, FMM3D_E_ARG, "partitioned solve: each of the 2x2x2 blocks needs >= 3 nodes per axis");
    int rc = part_prepare(c, nsrc);
    if (rc) return rc;
    PartState& ps = *c->part;
    const size_t N = c->N, words = (N + 31) / 32, NL = (size_t)g.nnx * g.nnz;
    const int npx = (p.len(1, 0) + PT_P - 1) / PT_P, npz = (p.len(2, 0) + PT_P - 1) / PT_P;
    double* T = (double*)c->sl.T;
    PartArgs a{};
    a.T = T; a.Tprev = ps.Tprev; a.T0 = ps.T0; a.hbits = ps.hbits; a.slo = ps.slo; a.hx = ps.hx; a.hz = ps.hz;
    a.hr = g.dnr; a.p = p; a.nslots = nsrc; a.fom = order == FMM3D_ORDER_MIXED ? 1 : 0; a.N = N; a.words = words;
    a.on = ps.on; a.dsend = ps.dsend; a.march = ps.march; a.mini = ps.mini; a.stamp = ps.stamp; a.chg = ps.chg;
    a.pstart = ps.pstart; a.err = ps.err; a.fnew = ps.fnew; a.nfmax = ps.nfmax; a.npx = npx; a.npz = npz;
    a.maxpass = 16; a.updates = c->d_visits + 1;
    hipStream_t st = c->stream;
    const int gsz = (int)std::min<size_t>(((size_t)nsrc * N + 255) / 256, 65536);
    // pass 0: injection; the bottom block(s) holding the smallest face minimum march from it
    k_pt_inject<<<gsz, 256, 0, st>>>(g.nnr, g.nnx, g.nnz, nsrc, c->d_faces, T);
    k_fill_real<double><<<gsz, 256, 0, st>>>(ps.Tprev, (size_t)nsrc * N, __builtin_huge_val());
    k_pt_fill_u32<<<gsz, 256, 0, st>>>(ps.hbits, (size_t)nsrc * words);
    k_pt_fill_int<<<(nsrc * 8 * npx * npz + 255) / 256, 256, 0, st>>>(ps.stamp, (size_t)nsrc * 8 * npx * npz, -1000);
    std::vector<int> march((size_t)nsrc * 8, 0), seeds, on(nsrc, 1), zeros((size_t)nsrc * 8, 0);
    std::vector<double> mini((size_t)nsrc * 8, __builtin_huge_val());
    for (int s = 0; s < nsrc; ++s) {
        const double* f = faces + (size_t)s * NL;
        int have[4] = { 0, 0, 0, 0 }, seed[4] = { -1, -1, -1, -1 };
        double tmin[4] = { 0, 0, 0, 0 };
        for (int q = 0; q < 4; ++q) {
            const int bx = q >> 1, bz = q & 1;
            for (int ix = p.lo(1, bx); ix < p.lo(1, bx) + p.len(1, bx); ++ix)
                for (int iz = p.lo(2, bz); iz < p.lo(2, bz) + p.len(2, bz); ++iz) {
                    const double v = f[(size_t)ix * g.nnz + iz];
                    if (v > 0.0 && (!have[q] || v < tmin[q])) {
                        tmin[q] = v;
                        have[q] = 1;
                        seed[q] = (int)((size_t)(g.nnr - 1) * NL + (size_t)ix * g.nnz + iz);
                    }
                }
        }
        for (int q = 0; q < 4; ++q) {
            if (!have[q]) continue;
            bool smaller = false;
            for (int o = 0; o < 4; ++o)
                if (have[o] && tmin[q] > tmin[o]) smaller = true;
            if (smaller) continue;
            march[(size_t)s * 8 + 4 + q] = 1;
            seeds.push_back(s);
            seeds.push_back(seed[q]);
        }
    }
    if ((int)seeds.size() / 2 > ps.nseed_cap) {
        (void)hipFree(ps.seed);
        ps.seed = nullptr;
        ps.nseed_cap = 0;
        HIPCHK(c, dalloc(&ps.seed, seeds.size() * sizeof(int)));
        ps.nseed_cap = (int)seeds.size() / 2;
    }
    const int nseed = (int)seeds.size() / 2;
    if (nseed) HIPCHK(c, hipMemcpyAsync(ps.seed, seeds.data(), seeds.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(ps.march, march.data(), march.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(ps.mini, mini.data(), mini.size() * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(ps.on, on.data(), on.size() * sizeof(int), hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(ps.dsend, zeros.data(), zeros.size() * sizeof(int), hipMemcpyHostToDevice, st));
    k_pt_fill_int<<<(nsrc + 255) / 256, 256, 0, st>>>(ps.pstart, (size_t)nsrc, 1);
    HIPCHK(c, hipMemsetAsync(ps.err, 0, (size_t)nsrc * sizeof(unsigned long long), st));
    if (nseed) k_pt_setbits<<<(nseed + 255) / 256, 256, 0, st>>>(nseed, ps.seed, words, ps.hbits);
    HIPCHK(c, hipGetLastError());
    // the host vectors above are pageable: the copies must land before they go
    HIPCHK(c, hipStreamSynchronize(st));

    const double tol = (double)1.0e-5f;   // F:6403, a default-real literal
    const int walk_blocks = nsrc * 8 * npx * npz;
    const size_t face_lds = 2 * ps.nfmax;
    int iter = 1;
    long total_sweeps = 0;
    ps.passes.assign(nsrc, 0);
    const int max_pass = 1000;
    for (int pass = 0; pass <= max_pass; ++pass) {
        if (pass > 0) {
            k_pt_pass_begin<<<gsz, 256, 0, st>>>(a);
            for (int k = 0; k < 6; ++k) k_pt_face<<<nsrc * 8, PT_FACE_THREADS, face_lds, st>>>(a, k);
            k_pt_mark<<<nsrc * 8, 256, 0, st>>>(a);
        }
        k_pt_start<<<gsz, 256, 0, st>>>(a);
        HIPCHK(c, hipGetLastError());
        // the march: sweeps until one changes nothing in every slot
        for (int sw = 0;; ++sw) {
            if (sw >= 100000) return fail(c, FMM3D_E_NOCONV, "partitioned solve: pass %d did not settle", pass);
            k_pt_fill_int<<<(nsrc + 255) / 256, 256, 0, st>>>(ps.chg, (size_t)nsrc, 0);
            a.iter = iter;
            k_pt_walk<<<walk_blocks, 64, 0, st>>>(a);
            HIPCHK(c, hipGetLastError());
            HIPCHK(c, hipMemcpyAsync(ps.h_chg, ps.chg, nsrc * sizeof(int), hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipStreamSynchronize(st));
            ++iter;
            ++total_sweeps;
            bool any = false;
            for (int s = 0; s < nsrc; ++s) any |= ps.h_chg[s] != 0;
            if (!any) break;
        }
        k_pt_err<<<dim3(64, nsrc), 256, 0, st>>>(a);
        HIPCHK(c, hipMemcpyAsync(ps.h_err, ps.err, nsrc * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        int live = 0;
        for (int s = 0; s < nsrc; ++s) {
            if (!on[s]) continue;
            ps.passes[s] = pass;
            double e;
            memcpy(&e, &ps.h_err[s], sizeof e);
            if (!(e > tol)) on[s] = 0;
            live += on[s];
        }
        // blocks that marched send next pass; slots whose change is within the tolerance stop
        k_pt_pass_end<<<(nsrc * 8 + 255) / 256, 256, 0, st>>>(a, iter);
        HIPCHK(c, hipMemcpyAsync(ps.on, on.data(), on.size() * sizeof(int), hipMemcpyHostToDevice, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (!live) break;
        if (pass == max_pass) return fail(c, FMM3D_E_NOCONV, "partitioned solve: %d passes without settling", max_pass);
    }
    c->stats.sweeps = (int32_t)total_sweeps;
    int mp = 0;
    for (int s = 0; s < nsrc; ++s) mp = std::max(mp, ps.passes[s]);
    c->stats.phase_sweeps[0] = mp;
    return FMM3D_OK;
}
