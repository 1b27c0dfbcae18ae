/*
 * fmm3d_oracle.c -- CPU restatement of the fm3dt_parallel forward problem.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP product
 * path: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, and never as the thing measured or shipped.
 *
 * It restates, for ONE domain holding the whole grid (the reference splits
 * the grid over a 2x2x2 MPI layout), the reference algorithm of
 * fm3dt_parallel.f90 ("F" below):
 *
 *   or_gridder      cubic B-spline vertex -> node velocity   F:6989-7038
 *   or_fmm_tele     narrow-band FMM, binary heap, bottom-face
 *                   injection, seeded at the face minimum    F:78-155, 1006-1098,
 *                                                            6894-6961
 *     local update  mixed-order (fouds2_all) / first-order
 *                   (fouds1_all) octant quadratic            F:535-1004
 *   or_fim_sweep    one value-based (FIM-form) sweep          F:1100-1400
 *   or_fmm_fixed, or_relax1_fixed, or_point_region
 *                   general fixed sets and point sources with a
 *                   refined neighbourhood (build-only, no reference)
 *   or_srtimes      trilinear receiver times, float32 sum    F:7040-7248
 *   or_rpaths       steepest-descent rays + Frechet rows     F:7250-7777
 *   or_slice        velocity-perturbation slices of gmtplot/gmtslicet.f90:4-540
 *   or_gmat_gtx/gx  subinv's sparse products G^T x, (G x)/cd  subinv.f90:457-475,
 *                   612-616, 759-767
 *
 * Layouts.  Fields are (nnr, nnx, nnz) with phi fastest -- the reference
 * array order ttn(iz,ix,ir) -- and carry no halo at the interface; veln is
 * returned on the node range -1..n+1 per axis (the reference gridder's
 * range) as (nnr+3, nnx+3, nnz+3).  Vertex arrays are (nvz+2, nvx+2, nvr+2)
 * with r fastest, the .vtx record order.  Frechet columns are the 1-based
 * r-fastest vertex index.
 *
 * Arithmetic follows the reference: float64 throughout except the float32
 * receiver accumulator and the float32 Frechet accumulator; pi is the
 * reference's float32-rounded literal (applied by the caller when it builds
 * the grid descriptor).  Build with -O2 and no FMA contraction
 * (-ffp-contract=off) so the results do not depend on the host ISA.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    int nvr, nvx, nvz;              /* vertex counts (arrays carry +2)     */
    double gor, gox, goz;           /* vertex origin: height km, colat rad, lon rad */
    double dvr, dvx, dvz;           /* vertex spacing: km, rad, rad        */
    int nnr, nnx, nnz;              /* propagation grid                    */
    double rgor, rgox, rgoz;        /* node origin: height km, colat rad, lon rad */
    double dnr, dnx, dnz;           /* node spacing                        */
    double earth;                   /* km                                  */
} or_grid;

/* ------------------------------------------------------------------ gridder */

static void bspline4(double t, double w[4])
{
    /* uniform cubic B-spline basis at fraction t in [0,1) (F:7000-7003) */
    double s = 1.0 - t, t2 = t * t, t3 = t2 * t;
    w[0] = s * s * s / 6.0;
    w[1] = (4.0 - 6.0 * t2 + 3.0 * t3) / 6.0;
    w[2] = (1.0 + 3.0 * t + 3.0 * t2 - 3.0 * t3) / 6.0;
    w[3] = t3 / 6.0;
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* veln on nodes -1..n+1 per axis: out[(kr+1)][(kx+1)][(kz+1)], kz fastest.
 * Taps outside the vertex array are clamped into it (the reference reads
 * past its array there; such halo values are only ever used with weight 0). */
void or_gridder(const or_grid *g, const double *velv, double *veln)
{
    const int NR = g->nvr + 2, NX = g->nvx + 2, NZ = g->nvz + 2;
    const int MX = g->nnx + 3, MZ = g->nnz + 3;
    for (int kr = -1; kr <= g->nnr + 1; ++kr) {
        double h = g->rgor - (kr - 1) * g->dnr;
        int ri = (int)floor((g->gor - g->rgor + (kr - 1) * g->dnr) / g->dvr) + 1;
        double distr = (g->gor - (ri - 1) * g->dvr) - h;
        double wr[4];
        bspline4(distr / g->dvr, wr);
        for (int kx = -1; kx <= g->nnx + 1; ++kx) {
            double th = g->rgox + (kx - 1) * g->dnx;
            int xi = (int)floor((th - g->gox) / g->dvx) + 1;
            double wx[4];
            bspline4((th - g->gox - (xi - 1) * g->dvx) / g->dvx, wx);
            for (int kz = -1; kz <= g->nnz + 1; ++kz) {
                double ph = g->rgoz + (kz - 1) * g->dnz;
                int zi = (int)floor((ph - g->goz) / g->dvz) + 1;
                double wz[4];
                bspline4((ph - g->goz - (zi - 1) * g->dvz) / g->dvz, wz);
                double acc_z = 0.0;
                for (int a = 0; a < 4; ++a) {
                    int iz = clampi(zi - 1 + a, 0, NZ - 1);
                    double acc_x = 0.0;
                    for (int b = 0; b < 4; ++b) {
                        int ix = clampi(xi - 1 + b, 0, NX - 1);
                        const double *col = velv + ((size_t)iz * NX + ix) * NR;
                        double acc_r = 0.0;
                        for (int c = 0; c < 4; ++c)
                            acc_r += wr[c] * col[clampi(ri - 1 + c, 0, NR - 1)];
                        acc_x += wx[b] * acc_r;
                    }
                    acc_z += wz[a] * acc_x;
                }
                veln[((size_t)(kr + 1) * MX + (kx + 1)) * MZ + (kz + 1)] = acc_z;
            }
        }
    }
}

/* ------------------------------------------------------ local octant update */

/* One octant: per axis (r, theta, phi) an order 0 (absent), 1 or 2, the
 * upwind neighbour t1, the second neighbour t2 and the node spacing h.
 * Same discrete equation as the reference case table (F:712-982, F:548-610):
 * first order (T - t1)/h, second order (3T - 4t1 + t2)/(2h), sum of squares
 * = slowness^2, largest root, discriminant clamped at zero.  The unknown is
 * shifted to the "leading" axis -- the first first-order axis, else the
 * first second-order axis -- and the equation is multiplied through by the
 * squared spacings, as the reference does. */
double or_octant_solve(const int ord[3], const double t1[3], const double t2[3],
                       const double h[3], double slown)
{
    int L = -1;
    for (int a = 0; a < 3; ++a) if (ord[a] == 1) { L = a; break; }
    if (L < 0) for (int a = 0; a < 3; ++a) if (ord[a] == 2) { L = a; break; }
    if (L < 0) return INFINITY;
    /* the reference pairs the remaining axes as (v, w) = (theta, r) when
     * phi leads and in natural order otherwise */
    int o[2], no = 0;
    if (L == 2) { if (ord[1]) o[no++] = 1; if (ord[0]) o[no++] = 0; }
    else for (int a = 0; a < 3; ++a) if (a != L && ord[a]) o[no++] = a;
    double cL = ord[L] == 2 ? 3.0 : 1.0;
    double DL = ord[L] == 2 ? 4.0 * t1[L] - t2[L] : t1[L];
    double HL = ord[L] == 2 ? 2.0 * h[L] : h[L];
    double u2 = HL * HL, s2 = slown * slown;
    double a, b, c;
    if (no == 0) {
        a = 1.0; b = 0.0; c = -(u2 * s2);
    } else {
        double m[2], e[2], H2[2];
        for (int i = 0; i < no; ++i) {
            int ax = o[i];
            double ca = ord[ax] == 2 ? 3.0 : 1.0;
            double Da = ord[ax] == 2 ? 4.0 * t1[ax] - t2[ax] : t1[ax];
            double Ha = ord[ax] == 2 ? 2.0 * h[ax] : h[ax];
            m[i] = ca / cL;                 /* 1 or 3 */
            e[i] = m[i] * DL - Da;
            H2[i] = Ha * Ha;
        }
        if (no == 1) {
            a = H2[0] + m[0] * m[0] * u2;
            b = 2.0 * m[0] * e[0] * u2;
            c = u2 * (e[0] * e[0] - s2 * H2[0]);
        } else {
            a = H2[0] * H2[1] + m[0] * m[0] * u2 * H2[1] + m[1] * m[1] * u2 * H2[0];
            b = 2.0 * u2 * (m[0] * e[0] * H2[1] + m[1] * e[1] * H2[0]);
            c = u2 * (e[0] * e[0] * H2[1] + e[1] * e[1] * H2[0] - s2 * H2[0] * H2[1]);
        }
    }
    double disc = b * b - 4.0 * a * c;
    if (disc < 0.0) disc = 0.0;
    double X = (-b + sqrt(disc)) / (2.0 * a);
    return (DL + X) / cL;
}

/* ---------------------------------------------------------------- domains */

/* One subdomain of the propagation grid: the whole grid in global mode, one
 * of the 2x2x2 blocks in partitioned mode.  Arrays are padded by one node on
 * every side (local indices 0..n+1), phi fastest.  The metric tables and the
 * node velocities use the reference's per-rank offset arithmetic
 * (F:540-543, 6989-7038) so both modes see the numbers the reference does. */
typedef struct {
    const or_grid *g;
    int n[3];              /* local sizes r, theta, phi                 */
    int off[3];            /* global 1-based index of local index 1     */
    int NX, NZ;            /* padded dims (x, z); r padded = n[0]+2     */
    double *t;             /* travel times (0 = unknown, as the reference) */
    double *prev;          /* value before the current relaxation pass  */
    int *st;               /* >0 heap slot, 0 alive, -1 far             */
    double *slo;           /* slowness                                  */
    double *ri;            /* earth + rgor - (roff-1) dnr - (ir-1) dnr  */
    double *sth;           /* sin(theta) per local ix                   */
    int fom;
    int *heap, nh;
    /* partitioned mode only */
    int coord[3];          /* block coordinates                         */
    double *gh[6][2];      /* ghost faces: [face][0 near, 1 far]        */
    int sw_fmm, dsend, srecv[6];
} dom_t;

#define DIDX(D, r, x, z) ((((size_t)(r)) * (D)->NX + (x)) * (D)->NZ + (z))

static double node_velocity(const or_grid *g, const double *velv, double h, double th, double ph,
                            double hr_num)
{
    /* B-spline at one node; hr_num is the reference's (gor - h) expression */
    const int NR = g->nvr + 2, NX = g->nvx + 2, NZ = g->nvz + 2;
    int ri = (int)floor(hr_num / g->dvr) + 1;
    double wr[4], wx[4], wz[4];
    bspline4(((g->gor - (ri - 1) * g->dvr) - h) / g->dvr, wr);
    int xi = (int)floor((th - g->gox) / g->dvx) + 1;
    bspline4((th - g->gox - (xi - 1) * g->dvx) / g->dvx, wx);
    int zi = (int)floor((ph - g->goz) / g->dvz) + 1;
    bspline4((ph - g->goz - (zi - 1) * g->dvz) / g->dvz, wz);
    double acc_z = 0.0;
    for (int a = 0; a < 4; ++a) {
        int iz = clampi(zi - 1 + a, 0, NZ - 1);
        double acc_x = 0.0;
        for (int b = 0; b < 4; ++b) {
            int ix = clampi(xi - 1 + b, 0, NX - 1);
            const double *col = velv + ((size_t)iz * NX + ix) * NR;
            double acc_r = 0.0;
            for (int c = 0; c < 4; ++c) acc_r += wr[c] * col[clampi(ri - 1 + c, 0, NR - 1)];
            acc_x += wx[b] * acc_r;
        }
        acc_z += wz[a] * acc_x;
    }
    return acc_z;
}

static void dom_init(dom_t *D, const or_grid *g, const int n[3], const int off[3], int fom,
                     const double *velv, const double *veln_global)
{
    memset(D, 0, sizeof *D);
    D->g = g; D->fom = fom;
    for (int a = 0; a < 3; ++a) { D->n[a] = n[a]; D->off[a] = off[a]; }
    D->NX = n[1] + 2; D->NZ = n[2] + 2;
    size_t N = (size_t)(n[0] + 2) * D->NX * D->NZ;
    D->t = calloc(N, sizeof(double));
    D->prev = calloc(N, sizeof(double));
    D->st = malloc(N * sizeof(int));
    D->slo = calloc(N, sizeof(double));
    D->heap = malloc((N + 1) * sizeof(int));
    D->ri = malloc((n[0] + 2) * sizeof(double));
    D->sth = malloc((n[1] + 2) * sizeof(double));
    for (int ir = 0; ir <= n[0] + 1; ++ir)
        D->ri[ir] = g->rgor - (off[0] - 1) * g->dnr - (ir - 1) * g->dnr + g->earth;
    for (int ix = 0; ix <= n[1] + 1; ++ix)
        D->sth[ix] = sin(g->rgox + (off[1] - 1) * g->dnx + (ix - 1) * g->dnx);
    const int MX = g->nnx + 3, MZ = g->nnz + 3;
    for (int ir = 1; ir <= n[0]; ++ir)
        for (int ix = 1; ix <= n[1]; ++ix)
            for (int iz = 1; iz <= n[2]; ++iz) {
                double v;
                if (veln_global) {
                    int gr = off[0] + ir - 1, gx = off[1] + ix - 1, gz = off[2] + iz - 1;
                    v = veln_global[((size_t)(gr + 1) * MX + (gx + 1)) * MZ + (gz + 1)];
                } else {
                    double h = g->rgor - (off[0] - 1) * g->dnr - (ir - 1) * g->dnr;
                    double hr = g->gor - g->rgor + (off[0] - 1) * g->dnr + (ir - 1) * g->dnr;
                    double th = g->rgox + (off[1] - 1) * g->dnx + (ix - 1) * g->dnx;
                    double ph = g->rgoz + (off[2] - 1) * g->dnz + (iz - 1) * g->dnz;
                    v = node_velocity(g, velv, h, th, ph, hr);
                }
                D->slo[DIDX(D, ir, ix, iz)] = 1.0 / v;
            }
}

static void dom_reset(dom_t *D)
{
    size_t N = (size_t)(D->n[0] + 2) * D->NX * D->NZ;
    for (size_t i = 0; i < N; ++i) { D->t[i] = 0.0; D->prev[i] = 0.0; D->st[i] = -1; }
    D->nh = 0;
    D->sw_fmm = 0; D->dsend = 0;
    for (int k = 0; k < 6; ++k) {
        D->srecv[k] = 0;
        if (D->gh[k][0]) {
            int a = k / 2;   /* face axis: 0 r (up/down), 1 x (left/right), 2 z (in/out) */
            (void)a;
        }
    }
}

static void dom_free(dom_t *D)
{
    free(D->t); free(D->prev); free(D->st); free(D->slo); free(D->heap); free(D->ri); free(D->sth);
    for (int k = 0; k < 6; ++k) for (int l = 0; l < 2; ++l) free(D->gh[k][l]);
}

/* Node update in FMM form: alive neighbours only, local bounds for the
 * octant test and the second neighbour (F:635-1004, F:535-633). */
static double fmm_update(dom_t *D, int ir, int ix, int iz)
{
    const or_grid *g = D->g;
    const int c[3] = { ir, ix, iz };
    double ri = D->ri[ir];
    double h[3] = { g->dnr, ri * g->dnx, ri * D->sth[ix] * g->dnz };
    int ord[3][2];
    double t1[3][2], t2[3][2];
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd) {
            int d = sd ? 1 : -1;
            int p1[3] = { c[0], c[1], c[2] }, p2[3] = { c[0], c[1], c[2] };
            p1[a] += d; p2[a] += 2 * d;
            ord[a][sd] = 0;
            if (p1[a] < 1 || p1[a] > D->n[a]) { ord[a][sd] = -1; continue; }
            size_t i1 = DIDX(D, p1[0], p1[1], p1[2]);
            if (D->st[i1] != 0) continue;
            ord[a][sd] = 1;
            t1[a][sd] = D->t[i1];
            if (D->fom && p2[a] >= 1 && p2[a] <= D->n[a]) {
                size_t i2 = DIDX(D, p2[0], p2[1], p2[2]);
                if (D->st[i2] == 0 && D->t[i1] > D->t[i2]) { ord[a][sd] = 2; t2[a][sd] = D->t[i2]; }
            }
        }
    double best = INFINITY, s = D->slo[DIDX(D, ir, ix, iz)];
    for (int sr = 0; sr < 2; ++sr)
        for (int sx = 0; sx < 2; ++sx)
            for (int sz = 0; sz < 2; ++sz) {
                if (ord[0][sr] < 0 || ord[1][sx] < 0 || ord[2][sz] < 0) continue;
                int o[3] = { ord[0][sr], ord[1][sx], ord[2][sz] };
                if (!o[0] && !o[1] && !o[2]) continue;
                double a1[3] = { t1[0][sr], t1[1][sx], t1[2][sz] };
                double a2[3] = { t2[0][sr], t2[1][sx], t2[2][sz] };
                double T = or_octant_solve(o, a1, a2, h, s);
                if (T < best) best = T;
            }
    return best;
}

/* binary heap on t; sift rules and tie handling as F:1006-1098 */
static void heap_up(dom_t *D, int pos)
{
    int node = D->heap[pos];
    double key = D->t[node];
    while (pos > 1) {
        int par = pos / 2, pn = D->heap[par];
        if (!(key < D->t[pn])) break;
        D->heap[pos] = pn; D->st[pn] = pos;
        pos = par;
    }
    D->heap[pos] = node; D->st[node] = pos;
}

static void heap_push(dom_t *D, int node)
{
    D->heap[++D->nh] = node;
    D->st[node] = D->nh;
    heap_up(D, D->nh);
}

static int heap_pop(dom_t *D)
{
    int top = D->heap[1];
    if (D->nh == 1) { D->nh = 0; return top; }
    int last = D->heap[D->nh--];
    int pos = 1;
    double key = D->t[last];
    for (;;) {
        int ch = 2 * pos;
        if (ch > D->nh) break;
        if (ch < D->nh && D->t[D->heap[ch]] > D->t[D->heap[ch + 1]]) ch++;
        int cn = D->heap[ch];
        if (!(D->t[cn] < key)) break;
        D->heap[pos] = cn; D->st[cn] = pos;
        pos = ch;
    }
    D->heap[pos] = last; D->st[last] = pos;
    return top;
}

/* the narrow-band march (F:78-155) */
static void march(dom_t *D)
{
    while (D->nh > 0) {
        int node = heap_pop(D);
        D->st[node] = 0;
        int iz = node % D->NZ, ix = (node / D->NZ) % D->NX, ir = node / (D->NZ * D->NX);
        int c[3] = { ir, ix, iz };
        for (int a = 0; a < 3; ++a)
            for (int d = -1; d <= 1; d += 2) {
                int p[3] = { c[0], c[1], c[2] };
                p[a] += d;
                if (p[a] < 1 || p[a] > D->n[a]) continue;
                int nb = (int)DIDX(D, p[0], p[1], p[2]);
                if (D->st[nb] == 0) continue;
                double T = fmm_update(D, p[0], p[1], p[2]);
                if (D->t[nb] == 0.0 || D->t[nb] > T) D->t[nb] = T;
                if (D->st[nb] < 0) heap_push(D, nb);
                else heap_up(D, D->st[nb]);
            }
    }
}

/* Load this block's part of a bottom face: positive values are known, the
 * rest unknown (F:6914-6939).  Returns 1 and the first strict minimum (in
 * the reference's loop order) if any value is known. */
static int dom_inject_bottom(dom_t *D, const double *face, int *mn, double *tmin)
{
    const or_grid *g = D->g;
    int have = 0;
    int ir = D->n[0];
    for (int ix = 1; ix <= D->n[1]; ++ix)
        for (int iz = 1; iz <= D->n[2]; ++iz) {
            int gx = D->off[1] + ix - 1, gz = D->off[2] + iz - 1;
            double v = face[(size_t)(gx - 1) * g->nnz + (gz - 1)];
            size_t k = DIDX(D, ir, ix, iz);
            if (v > 0.0) {
                D->t[k] = v; D->st[k] = 0;
                if (!have || v < *tmin) { *tmin = v; *mn = (int)k; have = 1; }
            } else {
                D->t[k] = 0.0;
            }
        }
    return have;
}

static void dom_extract(const dom_t *D, double *ttn)
{
    const or_grid *g = D->g;
    for (int ir = 1; ir <= D->n[0]; ++ir)
        for (int ix = 1; ix <= D->n[1]; ++ix)
            for (int iz = 1; iz <= D->n[2]; ++iz) {
                int gr = D->off[0] + ir - 1, gx = D->off[1] + ix - 1, gz = D->off[2] + iz - 1;
                ttn[((size_t)(gr - 1) * g->nnx + (gx - 1)) * g->nnz + (gz - 1)] = D->t[DIDX(D, ir, ix, iz)];
            }
}

/* Teleseismic solve on ONE domain holding the whole grid: bottom-face
 * injection, heap seeded at the face minimum (F:6894-6961 + F:78-155).
 * face: (nnx, nnz) phi fastest; veln from or_gridder; ttn out (nnr, nnx,
 * nnz) phi fastest.  Returns the number of alive nodes. */
long or_fmm_tele(const or_grid *g, const double *veln, const double *face, int fom, double *ttn)
{
    dom_t D;
    int n[3] = { g->nnr, g->nnx, g->nnz }, off[3] = { 1, 1, 1 };
    dom_init(&D, g, n, off, fom, NULL, veln);
    dom_reset(&D);
    int mn = -1;
    double tmin = 0.0;
    if (dom_inject_bottom(&D, face, &mn, &tmin)) {
        heap_push(&D, mn);
        march(&D);
    }
    dom_extract(&D, ttn);
    long alive = 0;
    for (int ir = 1; ir <= n[0]; ++ir)
        for (int ix = 1; ix <= n[1]; ++ix)
            for (int iz = 1; iz <= n[2]; ++iz) alive += D.st[DIDX(&D, ir, ix, iz)] == 0;
    dom_free(&D);
    return alive;
}

/* ----------------------------------------------- partitioned (2x2x2) mode */

/* Faces in the reference's numbering: 0 left (x-), 1 right (x+), 2 in (z-),
 * 3 out (z+), 4 up (r-), 5 down (r+). */
static const int FACE_AXIS[6] = { 1, 1, 2, 2, 0, 0 };
static const int FACE_SIDE[6] = { -1, 1, -1, 1, -1, 1 };

/* the two in-plane axes of a face, in the reference's loop order (outer, inner) */
static void face_axes(int k, int *outer, int *inner)
{
    int a = FACE_AXIS[k];
    if (a == 1) { *outer = 0; *inner = 2; }        /* DO r; DO z */
    else if (a == 2) { *outer = 0; *inner = 1; }   /* DO r; DO x */
    else { *outer = 1; *inner = 2; }               /* DO x; DO z */
}

static size_t face_index(const dom_t *D, int k, int io, int ii)
{
    int outer, inner;
    face_axes(k, &outer, &inner);
    return (size_t)(io - 1) * D->n[inner] + (ii - 1);
    (void)outer;
}

/* Value-based update of face node c through the ghost face k (F:1100-2825):
 * a neighbour counts iff known and below the node (or the node unknown);
 * the ghost side always counts, second order iff near ghost > far ghost;
 * only octants on the ghost side of the face axis.  Acceptance as F:1378
 * (F:2817 for the down face). */
static double face_value(const dom_t *D, int k, int ir, int ix, int iz)
{
    const or_grid *g = D->g;
    const int c[3] = { ir, ix, iz };
    const int A = FACE_AXIS[k];
    double tc = D->t[DIDX(D, ir, ix, iz)];
    double ri = D->ri[ir];
    double h[3] = { g->dnr, ri * g->dnx, ri * D->sth[ix] * g->dnz };
    int ord[3][2];
    double t1[3][2], t2[3][2];
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd) {
            int d = sd ? 1 : -1;
            ord[a][sd] = 0;
            if (a == A) {
                if (d != FACE_SIDE[k]) { ord[a][sd] = -1; continue; }
                int outer, inner;
                face_axes(k, &outer, &inner);
                size_t fi = face_index(D, k, c[outer], c[inner]);
                double nr_ = D->gh[k][0][fi], fr_ = D->gh[k][1][fi];
                t1[a][sd] = nr_;
                if (nr_ > fr_) { ord[a][sd] = 2; t2[a][sd] = fr_; }
                else ord[a][sd] = 1;
                continue;
            }
            int p1[3] = { c[0], c[1], c[2] }, p2[3] = { c[0], c[1], c[2] };
            p1[a] += d; p2[a] += 2 * d;
            if (p1[a] < 1 || p1[a] > D->n[a]) { ord[a][sd] = -1; continue; }
            double v1 = D->t[DIDX(D, p1[0], p1[1], p1[2])];
            if (!((v1 < tc || tc == 0.0) && v1 != 0.0)) continue;
            ord[a][sd] = 1; t1[a][sd] = v1;
            /* the face updates are mixed order whatever .in line 9 says (F:197-383) */
            if (p2[a] >= 1 && p2[a] <= D->n[a]) {
                double v2 = D->t[DIDX(D, p2[0], p2[1], p2[2])];
                if (v2 != 0.0 && v1 > v2) { ord[a][sd] = 2; t2[a][sd] = v2; }
            }
        }
    double best = INFINITY, s = D->slo[DIDX(D, ir, ix, iz)];
    for (int sr = 0; sr < 2; ++sr)
        for (int sx = 0; sx < 2; ++sx)
            for (int sz = 0; sz < 2; ++sz) {
                if (ord[0][sr] < 0 || ord[1][sx] < 0 || ord[2][sz] < 0) continue;
                int o[3] = { ord[0][sr], ord[1][sx], ord[2][sz] };
                if (!o[0] && !o[1] && !o[2]) continue;
                double a1[3] = { t1[0][sr], t1[1][sx], t1[2][sz] };
                double a2[3] = { t2[0][sr], t2[1][sx], t2[2][sz] };
                double T = or_octant_solve(o, a1, a2, h, s);
                if (T < best) best = T;
            }
    return best;
}

/* acceptance as F:1378 (F:2817 for the down face): the new value goes into
 * the block's heap */
static void face_accept(dom_t *D, int k, int ir, int ix, int iz, double best)
{
    size_t node = DIDX(D, ir, ix, iz);
    double old = D->t[node];
    int accept = (k == 5) ? (old > best || old == 0.0) : (old >= best || old == 0.0);
    if (accept) {
        D->t[node] = best;
        if (D->st[node] <= 0) heap_push(D, (int)node);
        else heap_up(D, D->st[node]);
    }
}

static void face_update(dom_t *D, int k, int ir, int ix, int iz)
{
    face_accept(D, k, ir, ix, iz, face_value(D, k, ir, ix, iz));
}

/* Reclassification after the face relaxation (F:387-415): if any face node
 * was accepted (the heap is not empty) the block marches this pass; every
 * other known node at or below the heap minimum is alive, at or above the
 * heap maximum far (its value kept), otherwise back in the band. */
static void reclassify(dom_t *D)
{
    if (D->nh >= 1) {
        D->sw_fmm = 1;
        double mini = D->t[D->heap[1]], maxi = 0.0;
        for (int i = 1; i <= D->nh; ++i) if (D->t[D->heap[i]] > maxi) maxi = D->t[D->heap[i]];
        for (int ir = 1; ir <= D->n[0]; ++ir)
            for (int ix = 1; ix <= D->n[1]; ++ix)
                for (int iz = 1; iz <= D->n[2]; ++iz) {
                    size_t q = DIDX(D, ir, ix, iz);
                    if (D->st[q] > 0) continue;
                    if (D->t[q] > 0.0) {
                        if (D->t[q] <= mini) D->st[q] = 0;
                        else if (D->t[q] >= maxi) D->st[q] = -1;
                        else heap_push(D, (int)q);
                    } else {
                        D->st[q] = -1;
                    }
                }
    }
}

/* face relaxation of one block in the reference's order: per face, a heap of
 * the face nodes keyed by the near ghost, popped in key order (F:157-386) */
static void face_relax_heap(dom_t *D)
{
    for (int k = 0; k < 6; ++k) {
        if (!D->srecv[k]) continue;
        int outer, inner;
        face_axes(k, &outer, &inner);
        int A = FACE_AXIS[k];
        int fixed = FACE_SIDE[k] < 0 ? 1 : D->n[A];
        int cap = D->n[outer] * D->n[inner];
        int *fh = malloc((cap + 1) * sizeof(int));   /* face-heap of face_index */
        const double *key = D->gh[k][0];
        int nf = 0;
        for (int io = 1; io <= D->n[outer]; ++io)
            for (int ii = 1; ii <= D->n[inner]; ++ii) {
                int p[3];
                p[A] = fixed; p[outer] = io; p[inner] = ii;
                double tf = D->t[DIDX(D, p[0], p[1], p[2])];
                int fi = (int)face_index(D, k, io, ii);
                if (key[fi] < tf || tf == 0.0) {
                    int pos = ++nf;
                    fh[pos] = fi;
                    while (pos > 1 && key[fi] < key[fh[pos / 2]]) {
                        fh[pos] = fh[pos / 2]; pos /= 2;
                    }
                    fh[pos] = fi;
                }
            }
        while (nf > 0) {
            int fi = fh[1];
            int p[3];
            p[A] = fixed; p[outer] = fi / D->n[inner] + 1; p[inner] = fi % D->n[inner] + 1;
            face_update(D, k, p[0], p[1], p[2]);
            /* pop the face heap */
            if (nf == 1) { nf = 0; continue; }
            int last = fh[nf--];
            int pos = 1;
            for (;;) {
                int ch = 2 * pos;
                if (ch > nf) break;
                if (ch < nf && key[fh[ch]] > key[fh[ch + 1]]) ch++;
                if (!(key[fh[ch]] < key[last])) break;
                fh[pos] = fh[ch]; pos = ch;
            }
            fh[pos] = last;
        }
        free(fh);
    }
}

/* ------------------------- partitioned mode as block relaxation (the GPU's) */

/* The face relaxation of face_relax_heap as dataflow rounds, the form the GPU
 * runs.  A face node in the heap reads, through face_update, its in-plane
 * first and second neighbours; in the heap order it sees the new value of
 * every such neighbour with a smaller key and the old value of the others.
 * So a node is ready once its heap neighbours of smaller key are done, and
 * each round updates every ready node.  Ties in the key are taken in face
 * index order (the reference's heap breaks them by its sift order). */
static void face_relax_rounds(dom_t *D)
{
    static const int NB[8][2] = { {-1, 0}, {1, 0}, {-2, 0}, {2, 0}, {0, -1}, {0, 1}, {0, -2}, {0, 2} };
    for (int k = 0; k < 6; ++k) {
        if (!D->srecv[k]) continue;
        int outer, inner;
        face_axes(k, &outer, &inner);
        const int A = FACE_AXIS[k];
        const int fixed = FACE_SIDE[k] < 0 ? 1 : D->n[A];
        const int no = D->n[outer], ni = D->n[inner], nf = no * ni;
        const double *key = D->gh[k][0];
        unsigned char *inh = calloc(nf, 1), *done = calloc(nf, 1), *now = calloc(nf, 1);
        double *val = malloc(nf * sizeof(double));
        for (int fi = 0; fi < nf; ++fi) {
            int p[3];
            p[A] = fixed; p[outer] = fi / ni + 1; p[inner] = fi % ni + 1;
            double tf = D->t[DIDX(D, p[0], p[1], p[2])];
            inh[fi] = key[fi] < tf || tf == 0.0;
        }
        for (;;) {
            int any = 0;
            for (int fi = 0; fi < nf; ++fi) {
                now[fi] = 0;
                if (!inh[fi] || done[fi]) continue;
                const int io = fi / ni, ii = fi % ni;
                int ready = 1;
                for (int q = 0; q < 8 && ready; ++q) {
                    const int jo = io + NB[q][0], ji = ii + NB[q][1];
                    if (jo < 0 || jo >= no || ji < 0 || ji >= ni) continue;
                    const int fj = jo * ni + ji;
                    if (inh[fj] && !done[fj] && (key[fj] < key[fi] || (key[fj] == key[fi] && fj < fi))) ready = 0;
                }
                now[fi] = ready;
                any |= ready;
            }
            if (!any) break;
            /* all of a round's values from the field before it (a ready node
             * reads its larger-key heap neighbours' old values), then applied */
            for (int fi = 0; fi < nf; ++fi) {
                if (!now[fi]) continue;
                int p[3];
                p[A] = fixed; p[outer] = fi / ni + 1; p[inner] = fi % ni + 1;
                val[fi] = face_value(D, k, p[0], p[1], p[2]);
            }
            for (int fi = 0; fi < nf; ++fi) {
                if (!now[fi]) continue;
                int p[3];
                p[A] = fixed; p[outer] = fi / ni + 1; p[inner] = fi % ni + 1;
                face_accept(D, k, p[0], p[1], p[2], val[fi]);
                done[fi] = 1;
            }
        }
        free(inh); free(done); free(now); free(val);
    }
}

/* The heap march of one block (local_fmm, F:78-155) as a relaxation: the
 * pop-order update.  The march freezes every node alive at its start (st 0)
 * and gives every other node the minimum of its starting value and the
 * octant candidates computed each time a neighbour pops, from the
 * neighbours alive at that time.  A neighbour pops before the node iff its
 * value is below the node's value so far, so the update walks the known
 * neighbours in increasing value, and a neighbour joins the stencil while
 * it is below the running minimum; after each join the octants through it
 * give a candidate.  Second order on an axis iff the second neighbour is
 * known and below the first (F:645-707); neighbours outside the block are
 * out of bounds (F:708-710).  Unlike the value-based form (a neighbour
 * counts iff below the node's own value), the result depends on the
 * neighbours only, so the relaxation cannot cycle; it settles where the
 * march does.  T: current values (+inf unknown). */
static double pop_update(const dom_t *D, const double *T, int ir, int ix, int iz, double t0)
{
    const or_grid *g = D->g;
    const int c[3] = { ir, ix, iz };
    const double ri = D->ri[ir];
    const double h[3] = { g->dnr, ri * g->dnx, ri * D->sth[ix] * g->dnz };
    const double s = D->slo[DIDX(D, ir, ix, iz)];
    double v1[6] = { 0 }, v2[6] = { 0 };
    int side[6], sec[6], ord6[6], m = 0;
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd) {
            const int j = 2 * a + sd, d = sd ? 1 : -1;
            int p1[3] = { c[0], c[1], c[2] }, p2[3] = { c[0], c[1], c[2] };
            p1[a] += d; p2[a] += 2 * d;
            side[j] = -1; sec[j] = 0;                       /* -1 out of bounds */
            if (p1[a] < 1 || p1[a] > D->n[a]) continue;
            side[j] = 0;                                    /* 0 not (yet) in the stencil */
            v1[j] = T[DIDX(D, p1[0], p1[1], p1[2])];
            if (!(v1[j] < INFINITY)) continue;
            if (D->fom && p2[a] >= 1 && p2[a] <= D->n[a]) {
                v2[j] = T[DIDX(D, p2[0], p2[1], p2[2])];
                sec[j] = v2[j] < INFINITY && v1[j] > v2[j];
            }
            ord6[m++] = j;
        }
    for (int i = 1; i < m; ++i) {                           /* ascending value, ties by side */
        int j = ord6[i], q = i;
        while (q > 0 && v1[ord6[q - 1]] > v1[j]) { ord6[q] = ord6[q - 1]; --q; }
        ord6[q] = j;
    }
    double cur = t0;
    for (int i = 0; i < m; ++i) {
        const int j = ord6[i];
        if (!(v1[j] < cur)) break;
        side[j] = sec[j] ? 2 : 1;
        const int aj = j / 2;
        for (int u = 0; u < 2; ++u)
            for (int w = 0; w < 2; ++w) {
                int pick[3];
                pick[aj] = j;
                pick[(aj + 1) % 3] = 2 * ((aj + 1) % 3) + u;
                pick[(aj + 2) % 3] = 2 * ((aj + 2) % 3) + w;
                int o[3];
                double a1[3], a2[3];
                int bad = 0;
                for (int a = 0; a < 3; ++a) {
                    const int q = pick[a];
                    if (side[q] < 0) { bad = 1; break; }
                    o[a] = side[q]; a1[a] = v1[q]; a2[a] = v2[q];
                }
                if (bad) continue;
                const double Tv = or_octant_solve(o, a1, a2, h, s);
                if (Tv < cur) cur = Tv;
            }
    }
    return cur;
}

/* Relax one block from the state the march would start from (alive nodes
 * fixed, every other node bounded by its starting value, unknown = +inf),
 * Gauss-Seidel in column order with the walk direction alternating, until a
 * sweep changes nothing; then every reached node is alive, as after the
 * march.  Returns the sweeps made. */
static long block_relax(dom_t *D)
{
    const size_t N = (size_t)(D->n[0] + 2) * D->NX * D->NZ;
    double *T = malloc(N * sizeof(double)), *T0 = malloc(N * sizeof(double));
    unsigned char *fx = calloc(N, 1);
    for (size_t q = 0; q < N; ++q) { T[q] = INFINITY; T0[q] = INFINITY; }
    for (int ir = 1; ir <= D->n[0]; ++ir)
        for (int ix = 1; ix <= D->n[1]; ++ix)
            for (int iz = 1; iz <= D->n[2]; ++iz) {
                const size_t q = DIDX(D, ir, ix, iz);
                T0[q] = D->t[q] > 0.0 ? D->t[q] : INFINITY;
                T[q] = T0[q];
                fx[q] = D->st[q] == 0;
            }
    long it;
    for (it = 1; it <= 100000; ++it) {
        int moved = 0;
        const int down = !(it & 1);
        for (int qq = 0; qq < D->n[0]; ++qq) {
            const int ir = down ? 1 + qq : D->n[0] - qq;
            for (int ix = 1; ix <= D->n[1]; ++ix)
                for (int iz = 1; iz <= D->n[2]; ++iz) {
                    const size_t k = DIDX(D, ir, ix, iz);
                    if (fx[k]) continue;
                    const double nv = pop_update(D, T, ir, ix, iz, T0[k]);
                    if (nv != T[k]) { T[k] = nv; moved = 1; }
                }
        }
        if (!moved) break;
    }
    for (int ir = 1; ir <= D->n[0]; ++ir)
        for (int ix = 1; ix <= D->n[1]; ++ix)
            for (int iz = 1; iz <= D->n[2]; ++iz) {
                const size_t q = DIDX(D, ir, ix, iz);
                D->t[q] = T[q] < INFINITY ? T[q] : 0.0;
                D->st[q] = T[q] < INFINITY ? 0 : -1;
            }
    D->nh = 0;
    free(T); free(T0); free(fx);
    return it;
}

/* copy the two planes of block B next to face k of block A into A's ghosts */
static void exchange_face(dom_t *A, const dom_t *B, int k)
{
    int outer, inner;
    face_axes(k, &outer, &inner);
    int ax = FACE_AXIS[k];
    /* B lies on A's side FACE_SIDE[k]; its near plane is its far edge */
    int nearB = FACE_SIDE[k] < 0 ? B->n[ax] : 1;
    int farB = FACE_SIDE[k] < 0 ? B->n[ax] - 1 : 2;
    for (int io = 1; io <= A->n[outer]; ++io)
        for (int ii = 1; ii <= A->n[inner]; ++ii) {
            int p[3], q[3];
            p[ax] = nearB; p[outer] = io; p[inner] = ii;
            q[ax] = farB; q[outer] = io; q[inner] = ii;
            size_t fi = face_index(A, k, io, ii);
            A->gh[k][0][fi] = B->t[DIDX(B, p[0], p[1], p[2])];
            A->gh[k][1][fi] = B->t[DIDX(B, q[0], q[1], q[2])];
        }
}

/* Teleseismic solve with the reference's own 2x2x2 decomposition and outer
 * relaxation (F:6127-6171, 6371-6448, 6894-6987): eight blocks, each with
 * its own heap march, two-deep ghost faces exchanged from blocks that
 * marched in the previous pass, repeated until the largest change between
 * passes is at most the float32 literal 1e-5.  Blocks are emulated in one
 * process in lockstep; the MPI exchanges are deterministic, so this is the
 * reference's result.  velv: vertex array (node velocities are evaluated
 * per block with the block's offset arithmetic).  Returns passes made. */
static int part_solve(const or_grid *g, const double *velv, const double *face, int fom,
                      double *ttn, int max_pass, int relax)
{
    const int P = 2;
    int nn[3] = { g->nnr, g->nnx, g->nnz };
    int sz[3][2], of[3][2];
    for (int a = 0; a < 3; ++a) {
        for (int r = 0; r < P; ++r) {
            int q = nn[a] / P, m = nn[a] % P;
            sz[a][r] = q + (r < m ? 1 : 0);
        }
        of[a][0] = 1; of[a][1] = 1 + sz[a][0];
    }
    dom_t Dm[8];
    for (int br = 0; br < P; ++br)
        for (int bx = 0; bx < P; ++bx)
            for (int bz = 0; bz < P; ++bz) {
                dom_t *D = &Dm[(br * P + bx) * P + bz];
                int n[3] = { sz[0][br], sz[1][bx], sz[2][bz] };
                int off[3] = { of[0][br], of[1][bx], of[2][bz] };
                dom_init(D, g, n, off, fom, velv, NULL);
                D->coord[0] = br; D->coord[1] = bx; D->coord[2] = bz;
                for (int k = 0; k < 6; ++k) {
                    int outer, inner;
                    face_axes(k, &outer, &inner);
                    size_t fn = (size_t)n[outer] * n[inner];
                    D->gh[k][0] = calloc(fn, sizeof(double));
                    D->gh[k][1] = calloc(fn, sizeof(double));
                }
            }
    for (int b = 0; b < 8; ++b) dom_reset(&Dm[b]);
    /* neighbour lookup */
    int nbr[8][6];
    for (int b = 0; b < 8; ++b) {
        int c[3] = { Dm[b].coord[0], Dm[b].coord[1], Dm[b].coord[2] };
        for (int k = 0; k < 6; ++k) {
            int q[3] = { c[0], c[1], c[2] };
            q[FACE_AXIS[k]] += FACE_SIDE[k];
            nbr[b][k] = (q[FACE_AXIS[k]] < 0 || q[FACE_AXIS[k]] >= P) ? -1 : (q[0] * P + q[1]) * P + q[2];
        }
    }
    /* injection on the bottom blocks; the block(s) holding the smallest
     * minimum seed and march (F:6914-6959) */
    int have[8] = { 0 }, mn[8];
    double tmin[8];
    for (int b = 0; b < 8; ++b) {
        mn[b] = -1; tmin[b] = -100.0;
        if (Dm[b].coord[0] == P - 1) {
            double v = 0.0;
            int m = -1;
            if (dom_inject_bottom(&Dm[b], face, &m, &v)) { have[b] = 1; mn[b] = m; tmin[b] = v; }
        }
    }
    for (int b = 0; b < 8; ++b) {
        if (Dm[b].coord[0] != P - 1 || !have[b]) continue;
        int sw = 0;
        for (int o = 0; o < 8; ++o)
            if (Dm[o].coord[0] == P - 1 && have[o] && tmin[b] > tmin[o]) sw = 1;
        if (!sw) {
            heap_push(&Dm[b], mn[b]);
            if (relax) block_relax(&Dm[b]); else march(&Dm[b]);
            Dm[b].sw_fmm = 1;
        }
    }
    const double tol = (double)1.0e-5f;
    double err = 0.0;
    for (int b = 0; b < 8; ++b) {
        dom_t *D = &Dm[b];
        for (int ir = 1; ir <= D->n[0]; ++ir)
            for (int ix = 1; ix <= D->n[1]; ++ix)
                for (int iz = 1; iz <= D->n[2]; ++iz) {
                    size_t q = DIDX(D, ir, ix, iz);
                    double d = fabs(D->t[q] - D->prev[q]);
                    if (d > err) err = d;
                }
    }
    int pass = 0;
    while (err > tol && pass < max_pass) {
        for (int b = 0; b < 8; ++b) {
            dom_t *D = &Dm[b];
            size_t N = (size_t)(D->n[0] + 2) * D->NX * D->NZ;
            memcpy(D->prev, D->t, N * sizeof(double));
            D->dsend = D->sw_fmm;
            D->sw_fmm = 0;
        }
        for (int b = 0; b < 8; ++b)
            for (int k = 0; k < 6; ++k)
                if (nbr[b][k] >= 0) Dm[b].srecv[k] = Dm[nbr[b][k]].dsend;
        for (int b = 0; b < 8; ++b)
            for (int k = 0; k < 6; ++k)
                if (nbr[b][k] >= 0 && Dm[nbr[b][k]].dsend) exchange_face(&Dm[b], &Dm[nbr[b][k]], k);
        for (int b = 0; b < 8; ++b) {
            if (relax) face_relax_rounds(&Dm[b]); else face_relax_heap(&Dm[b]);
            reclassify(&Dm[b]);
            if (Dm[b].nh > 0) { if (relax) block_relax(&Dm[b]); else march(&Dm[b]); }
        }
        err = 0.0;
        for (int b = 0; b < 8; ++b) {
            dom_t *D = &Dm[b];
            for (int ir = 1; ir <= D->n[0]; ++ir)
                for (int ix = 1; ix <= D->n[1]; ++ix)
                    for (int iz = 1; iz <= D->n[2]; ++iz) {
                        size_t q = DIDX(D, ir, ix, iz);
                        double d = fabs(D->t[q] - D->prev[q]);
                        if (d > err) err = d;
                    }
        }
        pass++;
    }
    for (int b = 0; b < 8; ++b) { dom_extract(&Dm[b], ttn); dom_free(&Dm[b]); }
    return pass;
}

int or_fmm_tele_part(const or_grid *g, const double *velv, const double *face, int fom,
                     double *ttn, int max_pass)
{
    return part_solve(g, velv, face, fom, ttn, max_pass, 0);
}

/* The same partitioned solve with the two sequential parts in the parallel
 * form the GPU runs: each block's heap march as the pop-order relaxation
 * (block_relax) and each face heap as dataflow rounds (face_relax_rounds).
 * Returns the passes made. */
int or_relax_tele_part(const or_grid *g, const double *velv, const double *face, int fom,
                       double *ttn, int max_pass)
{
    return part_solve(g, velv, face, fom, ttn, max_pass, 1);
}

static void metric_tables(const or_grid *g, double *ri, double *sth)
{
    for (int ir = 0; ir <= g->nnr + 1; ++ir) ri[ir] = g->rgor - (ir - 1) * g->dnr + g->earth;
    for (int ix = 0; ix <= g->nnx + 1; ++ix) sth[ix] = sin(g->rgox + (ix - 1) * g->dnx);
}

/* ----------------------------------------------------- value-based update */

double or_fim_sweep_mode(const or_grid *g, const double *veln, const unsigned char *fixed,
                         int fom, const double *t_in, double *t_out, int mode);

/* One Jacobi sweep of the value-based (FIM-form) update over every node that
 * is not fixed: a neighbour counts iff it is known and smaller than the
 * node's current value (or the node is unknown); the second-order flag needs
 * the second neighbour known and below the first (F:1112-1150); the update is
 * accepted iff it does not increase the value (F:1378).  Unknown = +inf.
 * t_in/t_out: (nnr, nnx, nnz) phi fastest; fixed: same shape, nonzero = keep.
 * Returns max |t_out - t_in| over nodes whose value was finite before. */
double or_fim_sweep(const or_grid *g, const double *veln, const unsigned char *fixed,
                    int fom, const double *t_in, double *t_out)
{
    return or_fim_sweep_mode(g, veln, fixed, fom, t_in, t_out, 0);
}

/* mode 0: accept only non-increasing updates (F:1378); mode 1: plain Jacobi
 * (the update replaces the value whenever one exists) */
double or_fim_sweep_pred(const or_grid *g, const double *veln, const unsigned char *fixed,
                         int fom, const double *t_in, const double *t_pred, double *t_out, int mode);

double or_fim_sweep_mode(const or_grid *g, const double *veln, const unsigned char *fixed,
                         int fom, const double *t_in, double *t_out, int mode)
{
    return or_fim_sweep_pred(g, veln, fixed, fom, t_in, t_in, t_out, mode);
}

/* The same sweep with the stencil choice taken from a separate ordering
 * field t_pred (neighbour counts iff t_pred(nb) < t_pred(c); second order
 * iff t_pred(nb2) < t_pred(nb)); values come from t_in. */
/* The value-based node update shared by the relaxations: neighbours count
 * by the ordering field t_pred (nb counts iff t_pred(nb) < t_pred(c) and
 * its value is known; second order iff t_pred(nb2) < t_pred(nb)), values
 * come from t_in; minimum over the valid octants (F:535-1004, F:1112-1150).
 * Indices 1-based. */
static double relax_node(const or_grid *g, const double *veln, const double *ri, const double *sth,
                         int fom, const double *t_in, const double *t_pred, int ir, int ix, int iz)
{
    const int nr = g->nnr, nx = g->nnx, nz = g->nnz;
    const int MX = nx + 3, MZ = nz + 3;
#define TI(r, x, z) t_in[((size_t)((r) - 1) * nx + ((x) - 1)) * nz + ((z) - 1)]
#define TP(r, x, z) t_pred[((size_t)((r) - 1) * nx + ((x) - 1)) * nz + ((z) - 1)]
    const double pc = TP(ir, ix, iz);
    const int n[3] = { nr, nx, nz };
    const int c[3] = { ir, ix, iz };
    double rr = ri[ir];
    double h[3] = { g->dnr, rr * g->dnx, rr * sth[ix] * g->dnz };
    int ord[3][2];
    double t1[3][2], t2[3][2];
    for (int a = 0; a < 3; ++a)
        for (int sd = 0; sd < 2; ++sd) {
            int d = sd ? 1 : -1;
            int p1[3] = { c[0], c[1], c[2] }, p2[3] = { c[0], c[1], c[2] };
            p1[a] += d; p2[a] += 2 * d;
            ord[a][sd] = 0;
            if (p1[a] < 1 || p1[a] > n[a]) { ord[a][sd] = -1; continue; }
            double v1 = TI(p1[0], p1[1], p1[2]), q1 = TP(p1[0], p1[1], p1[2]);
            if (!(q1 < pc) || !(v1 < INFINITY)) continue;   /* inf never counts */
            ord[a][sd] = 1; t1[a][sd] = v1;
            if (fom && p2[a] >= 1 && p2[a] <= n[a]) {
                double v2 = TI(p2[0], p2[1], p2[2]), q2 = TP(p2[0], p2[1], p2[2]);
                if (q2 < INFINITY && v2 < INFINITY && q1 > q2) { ord[a][sd] = 2; t2[a][sd] = v2; }
            }
        }
#undef TI
#undef TP
    double best = INFINITY;
    for (int sr = 0; sr < 2; ++sr)
        for (int sx = 0; sx < 2; ++sx)
            for (int sz = 0; sz < 2; ++sz) {
                if (ord[0][sr] < 0 || ord[1][sx] < 0 || ord[2][sz] < 0) continue;
                int o[3] = { ord[0][sr], ord[1][sx], ord[2][sz] };
                if (!o[0] && !o[1] && !o[2]) continue;
                double a1[3] = { t1[0][sr], t1[1][sx], t1[2][sz] };
                double a2[3] = { t2[0][sr], t2[1][sx], t2[2][sz] };
                double T = or_octant_solve(o, a1, a2, h,
                    1.0 / veln[((size_t)(ir + 1) * MX + (ix + 1)) * MZ + (iz + 1)]);
                if (T < best) best = T;
            }
    return best;
}

double or_fim_sweep_pred(const or_grid *g, const double *veln, const unsigned char *fixed,
                         int fom, const double *t_in, const double *t_pred, double *t_out, int mode)
{
    const int nr = g->nnr, nx = g->nnx, nz = g->nnz;
    double *sth = malloc((nx + 2) * sizeof(double));
    double *ri = malloc((nr + 2) * sizeof(double));
    metric_tables(g, ri, sth);
    double maxd = 0.0;
    for (int ir = 1; ir <= nr; ++ir)
        for (int ix = 1; ix <= nx; ++ix)
            for (int iz = 1; iz <= nz; ++iz) {
                size_t k = ((size_t)(ir - 1) * nx + (ix - 1)) * nz + (iz - 1);
                double tc = t_in[k];
                if (fixed[k]) { t_out[k] = tc; continue; }
                double best = relax_node(g, veln, ri, sth, fom, t_in, t_pred, ir, ix, iz);
                double nv;
                if (mode == 1 && best == INFINITY && tc < INFINITY) nv = INFINITY;  /* unsupported value: drop */
                else nv = mode == 1 ? (best < INFINITY ? best : tc) : ((best <= tc) ? best : tc);
                t_out[k] = nv;
                if (tc < INFINITY) { double dd = fabs(nv - tc); if (dd > maxd) maxd = dd; }
            }
    free(sth); free(ri);
    return maxd;
}

/* Three-phase relaxation for a teleseismic bottom face: phase A relaxes the
 * value-based first-order update (Gauss-Seidel, unsupported values dropped)
 * to its fixed point; phases B, C, ... freeze the ordering from the previous
 * phase and relax the mixed-order update with it, until a phase changes
 * nothing.  Each phase iterates to exact convergence.  With a fully known
 * bottom face this equals or_fmm_tele bit for bit; where the face has
 * unknown nodes it can be earlier, because the heap march reaches such
 * nodes only from above (the face nodes other than the seed never leave the
 * alive set through the heap, F:6930-6958).  Returns the total number of
 * sweeps or -1. */
long or_relax_tele(const or_grid *g, const double *veln, const double *face, int fom, double *ttn)
{
    const size_t NL = (size_t)g->nnx * g->nnz, N = NL * g->nnr;
    unsigned char *fixed = calloc(N, 1);
    double *P = malloc(N * sizeof(double)), *prev = malloc(N * sizeof(double));
    for (size_t i = 0; i < N; ++i) ttn[i] = INFINITY;
    for (size_t f = 0; f < NL; ++f)
        if (face[f] > 0.0) { ttn[(size_t)(g->nnr - 1) * NL + f] = face[f]; fixed[(size_t)(g->nnr - 1) * NL + f] = 1; }
    long sweeps = 0;
    int phases = fom ? 4 : 1;
    for (int ph = 0; ph < phases; ++ph) {
        memcpy(P, ttn, N * sizeof(double));
        int moved_any = 0;
        for (int it = 0; it < 100000; ++it) {
            memcpy(prev, ttn, N * sizeof(double));
            or_fim_sweep_pred(g, veln, fixed, ph == 0 ? 0 : fom, ttn, ph == 0 ? ttn : P, ttn, 1);
            sweeps++;
            if (!memcmp(prev, ttn, N * sizeof(double))) break;
            moved_any = 1;
            if (it == 99999) { sweeps = -1; goto out; }
        }
        if (ph >= 2 && !moved_any) break;
    }
out:
    free(fixed); free(P); free(prev);
    return sweeps;
}

/* Single-phase relaxation in column-walk order for a teleseismic bottom
 * face: the value-based mixed-order update (the node's own value orders its
 * neighbours) with unsupported values dropped, applied in place layer by
 * layer from the bottom to the top; each layer gets four line passes
 * (phi ascending, phi descending, theta ascending, theta descending).
 * Sweeps repeat until one changes nothing.  This is the discrete system the
 * GPU column walk relaxes; its fixed point is the heap march's field where
 * the bottom face is fully known.  Returns the number of sweeps or -1. */
long or_relax1_tele(const or_grid *g, const double *veln, const double *face, int fom, double *ttn)
{
    const int nr = g->nnr, nx = g->nnx, nz = g->nnz;
    const size_t NL = (size_t)nx * nz, N = NL * nr;
    unsigned char *fixed = calloc(N, 1);
    double *sth = malloc((nx + 2) * sizeof(double));
    double *ri = malloc((nr + 2) * sizeof(double));
    metric_tables(g, ri, sth);
    for (size_t i = 0; i < N; ++i) ttn[i] = INFINITY;
    for (size_t f = 0; f < NL; ++f)
        if (face[f] > 0.0) { ttn[(size_t)(nr - 1) * NL + f] = face[f]; fixed[(size_t)(nr - 1) * NL + f] = 1; }
    long sweeps = -1;
    for (int it = 1; it <= 100000; ++it) {
        int moved = 0;
        for (int ir = nr; ir >= 1; --ir)
            for (int pass = 0; pass < 4; ++pass)
                for (int a = 0; a < (pass < 2 ? nz : nx); ++a)
                    for (int b = 0; b < (pass < 2 ? nx : nz); ++b) {
                        int ix, iz;
                        if (pass < 2) { iz = 1 + (pass == 0 ? a : nz - 1 - a); ix = 1 + b; }
                        else { ix = 1 + (pass == 2 ? a : nx - 1 - a); iz = 1 + b; }
                        size_t k = ((size_t)(ir - 1) * nx + (ix - 1)) * nz + (iz - 1);
                        if (fixed[k]) continue;
                        double nv = relax_node(g, veln, ri, sth, fom, ttn, ttn, ir, ix, iz);
                        if (nv != ttn[k]) { ttn[k] = nv; moved = 1; }
                    }
        if (!moved) { sweeps = it; break; }
    }
    free(fixed); free(sth); free(ri);
    return sweeps;
}


/* ------------------------------------------- general Dirichlet sets, point sources */

/* Textbook narrow-band march from a fixed (Dirichlet) node set: every
 * fixed node is alive with its given time and seeds the band through its
 * six neighbours; then the march of F:78-155 with the update of
 * F:535-1004.  (The reference injects only bottom faces and seeds at the
 * face minimum, F:6894-6961; with a fully known face both give the same
 * field.)  fixed/tfix: (nnr, nnx, nnz) phi fastest.  Returns alive nodes. */
long or_fmm_fixed(const or_grid *g, const double *veln, const unsigned char *fixed, const double *tfix,
                  int fom, double *ttn)
{
    dom_t D;
    int n[3] = { g->nnr, g->nnx, g->nnz }, off[3] = { 1, 1, 1 };
    dom_init(&D, g, n, off, fom, NULL, veln);
    dom_reset(&D);
    for (int ir = 1; ir <= n[0]; ++ir)
        for (int ix = 1; ix <= n[1]; ++ix)
            for (int iz = 1; iz <= n[2]; ++iz) {
                size_t k = ((size_t)(ir - 1) * n[1] + (ix - 1)) * n[2] + (iz - 1);
                if (!fixed[k]) continue;
                size_t d = DIDX(&D, ir, ix, iz);
                D.t[d] = tfix[k];
                D.st[d] = 0;
            }
    for (int ir = 1; ir <= n[0]; ++ir)
        for (int ix = 1; ix <= n[1]; ++ix)
            for (int iz = 1; iz <= n[2]; ++iz) {
                size_t k = ((size_t)(ir - 1) * n[1] + (ix - 1)) * n[2] + (iz - 1);
                if (!fixed[k]) continue;
                int c[3] = { ir, ix, iz };
                for (int a = 0; a < 3; ++a)
                    for (int dd = -1; dd <= 1; dd += 2) {
                        int p[3] = { c[0], c[1], c[2] };
                        p[a] += dd;
                        if (p[a] < 1 || p[a] > n[a]) continue;
                        int nb = (int)DIDX(&D, p[0], p[1], p[2]);
                        if (D.st[nb] == 0) continue;
                        double T = fmm_update(&D, p[0], p[1], p[2]);
                        if (D.t[nb] == 0.0 || D.t[nb] > T) D.t[nb] = T;
                        if (D.st[nb] < 0) heap_push(&D, nb);
                        else heap_up(&D, D.st[nb]);
                    }
            }
    march(&D);
    dom_extract(&D, ttn);
    long alive = 0;
    for (int ir = 1; ir <= n[0]; ++ir)
        for (int ix = 1; ix <= n[1]; ++ix)
            for (int iz = 1; iz <= n[2]; ++iz) alive += D.st[DIDX(&D, ir, ix, iz)] == 0;
    dom_free(&D);
    return alive;
}

/* The single-phase column relaxation (see or_relax1_tele) for a general
 * fixed set; with `alternate` the layer walk goes bottom to top on odd
 * sweeps and top to bottom on even ones.  Unknown nodes start at +inf.
 * Returns the number of sweeps or -1. */
long or_relax1_fixed(const or_grid *g, const double *veln, const unsigned char *fixed, const double *tfix,
                     int fom, int alternate, double *ttn)
{
    const int nr = g->nnr, nx = g->nnx, nz = g->nnz;
    const size_t N = (size_t)nr * nx * nz;
    double *sth = malloc((nx + 2) * sizeof(double));
    double *ri = malloc((nr + 2) * sizeof(double));
    metric_tables(g, ri, sth);
    for (size_t i = 0; i < N; ++i) ttn[i] = fixed[i] ? tfix[i] : INFINITY;
    long sweeps = -1;
    for (int it = 1; it <= 100000; ++it) {
        int moved = 0;
        const int down = alternate && !(it & 1);
        for (int q = 0; q < nr; ++q) {
            const int ir = down ? 1 + q : nr - q;
            for (int pass = 0; pass < 4; ++pass)
                for (int a = 0; a < (pass < 2 ? nz : nx); ++a)
                    for (int b = 0; b < (pass < 2 ? nx : nz); ++b) {
                        int ix, iz;
                        if (pass < 2) { iz = 1 + (pass == 0 ? a : nz - 1 - a); ix = 1 + b; }
                        else { ix = 1 + (pass == 2 ? a : nx - 1 - a); iz = 1 + b; }
                        size_t k = ((size_t)(ir - 1) * nx + (ix - 1)) * nz + (iz - 1);
                        if (fixed[k]) continue;
                        double nv = relax_node(g, veln, ri, sth, fom, ttn, ttn, ir, ix, iz);
                        if (nv != ttn[k]) { ttn[k] = nv; moved = 1; }
                    }
        }
        if (!moved) { sweeps = it; break; }
    }
    free(sth); free(ri);
    return sweeps;
}

/* Gauss-Seidel sweeps of the value-based mixed-order update (relax_node, the
 * node's own value orders its neighbours, unsupported values dropped) in
 * flow order: walk layers bottom to top (top to bottom when `down`), inside
 * a layer theta in the direction of fx (+1 ascending, -1 descending), then
 * phi in the direction of fz; every free node is updated once per sweep.
 * Starts from ttn as given (fixed nodes keep their values).  With maxsweeps
 * > 0 it stops after that many sweeps.  Returns the sweep that changed
 * nothing (converged), or -(sweeps run) when it stopped before that. */
long or_gs_flow(const or_grid *g, const double *veln, const unsigned char *fixed, int fom,
                int fx, int fz, int down, int maxsweeps, double *ttn)
{
    const int nr = g->nnr, nx = g->nnx, nz = g->nnz;
    double *sth = malloc((nx + 2) * sizeof(double));
    double *ri = malloc((nr + 2) * sizeof(double));
    metric_tables(g, ri, sth);
    long res = -1;
    const int lim = maxsweeps > 0 ? maxsweeps : 100000;
    int it;
    for (it = 1; it <= lim; ++it) {
        int moved = 0;
        for (int q = 0; q < nr; ++q) {
            const int ir = down ? 1 + q : nr - q;
            for (int a = 0; a < nx; ++a) {
                const int ix = fx > 0 ? 1 + a : nx - a;
                for (int b = 0; b < nz; ++b) {
                    const int iz = fz > 0 ? 1 + b : nz - b;
                    size_t k = ((size_t)(ir - 1) * nx + (ix - 1)) * nz + (iz - 1);
                    if (fixed[k]) continue;
                    double nv = relax_node(g, veln, ri, sth, fom, ttn, ttn, ir, ix, iz);
                    if (nv != ttn[k]) { ttn[k] = nv; moved = 1; }
                }
            }
        }
        if (!moved) { res = it; break; }
    }
    if (res < 0) res = -(long)(it - 1);
    free(sth); free(ri);
    return res;
}

/* Cartesian position of (height km, colat, lon) */
static void cart(const or_grid *g, double h, double th, double ph, double x[3])
{
    double r = g->earth + h;
    x[0] = r * sin(th) * cos(ph);
    x[1] = r * sin(th) * sin(ph);
    x[2] = r * cos(th);
}

/* Refined source neighbourhood for a point source (build-only: the
 * reference solves teleseismic injection only).  The coarse cell holding
 * the source, widened by `radius` nodes per side (clipped to the grid), is
 * refined by `refine` per axis.  On that sub-grid: node velocities by the
 * B-spline gridder; the 8 corners of the fine cell holding the source get
 * the straight-ray time |x - x_s| times the mean of the source and node
 * slowness; a heap march (`solve` 0) or the column relaxation (`solve` 1,
 * alternating walks) fills the sub-grid.  The coarse nodes it covers take
 * the fine times and become the fixed set of the coarse solve.
 * src: height km, colat rad, lon rad.  Outputs: lo/hi coarse node range of
 * the refined block (0-based, inclusive) and the coarse fixed set/times;
 * returns the number of fine nodes or -1 when the source lies outside. */
long or_point_region(const or_grid *g, const double *velv, const double src[3], int refine, int radius,
                     int fom, int solve, int lo[3], int hi[3], unsigned char *cfixed, double *ctfix)
{
    const int nn[3] = { g->nnr, g->nnx, g->nnz };
    double u[3] = { (g->rgor - src[0]) / g->dnr, (src[1] - g->rgox) / g->dnx, (src[2] - g->rgoz) / g->dnz };
    int c[3];
    for (int a = 0; a < 3; ++a) {
        if (!(u[a] >= 0.0 && u[a] <= nn[a] - 1)) return -1;
        c[a] = (int)floor(u[a]);
        if (c[a] > nn[a] - 2) c[a] = nn[a] - 2;
        lo[a] = c[a] - radius < 0 ? 0 : c[a] - radius;
        hi[a] = c[a] + 1 + radius > nn[a] - 1 ? nn[a] - 1 : c[a] + 1 + radius;
    }
    or_grid f = *g;
    f.nnr = (hi[0] - lo[0]) * refine + 1;
    f.nnx = (hi[1] - lo[1]) * refine + 1;
    f.nnz = (hi[2] - lo[2]) * refine + 1;
    f.rgor = g->rgor - lo[0] * g->dnr;
    f.rgox = g->rgox + lo[1] * g->dnx;
    f.rgoz = g->rgoz + lo[2] * g->dnz;
    f.dnr = g->dnr / refine; f.dnx = g->dnx / refine; f.dnz = g->dnz / refine;
    const size_t FN = (size_t)f.nnr * f.nnx * f.nnz;
    double *fveln = malloc((size_t)(f.nnr + 3) * (f.nnx + 3) * (f.nnz + 3) * sizeof(double));
    or_gridder(&f, velv, fveln);
    unsigned char *ffix = calloc(FN, 1);
    double *ftf = calloc(FN, sizeof(double)), *ft = malloc(FN * sizeof(double));
    double fu[3] = { (f.rgor - src[0]) / f.dnr, (src[1] - f.rgox) / f.dnx, (src[2] - f.rgoz) / f.dnz };
    const int fn[3] = { f.nnr, f.nnx, f.nnz };
    int fc[3];
    for (int a = 0; a < 3; ++a) {
        fc[a] = (int)floor(fu[a]);
        if (fc[a] < 0) fc[a] = 0;
        if (fc[a] > fn[a] - 2) fc[a] = fn[a] - 2;
    }
    double xs[3];
    cart(g, src[0], src[1], src[2], xs);
    const double ss = 1.0 / node_velocity(g, velv, src[0], src[1], src[2], g->gor - src[0]);
    const int MX = f.nnx + 3, MZ = f.nnz + 3;
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                int r = fc[0] + i, x = fc[1] + j, z = fc[2] + k;
                double xn[3];
                cart(g, f.rgor - r * f.dnr, f.rgox + x * f.dnx, f.rgoz + z * f.dnz, xn);
                double d = sqrt((xn[0] - xs[0]) * (xn[0] - xs[0]) + (xn[1] - xs[1]) * (xn[1] - xs[1]) +
                                (xn[2] - xs[2]) * (xn[2] - xs[2]));
                double sn = 1.0 / fveln[((size_t)(r + 2) * MX + (x + 2)) * MZ + (z + 2)];   /* node r: kr = r+1 */
                size_t q = ((size_t)r * f.nnx + x) * f.nnz + z;
                ffix[q] = 1;
                ftf[q] = d * 0.5 * (ss + sn);
            }
    if (solve == 0) or_fmm_fixed(&f, fveln, ffix, ftf, fom, ft);
    else (void)or_relax1_fixed(&f, fveln, ffix, ftf, fom, 1, ft);
    const size_t N = (size_t)g->nnr * g->nnx * g->nnz;
    memset(cfixed, 0, N);
    for (int r = lo[0]; r <= hi[0]; ++r)
        for (int x = lo[1]; x <= hi[1]; ++x)
            for (int z = lo[2]; z <= hi[2]; ++z) {
                size_t k = ((size_t)r * g->nnx + x) * g->nnz + z;
                size_t q = ((size_t)(r - lo[0]) * refine * f.nnx + (size_t)(x - lo[1]) * refine) * f.nnz + (size_t)(z - lo[2]) * refine;
                cfixed[k] = 1;
                ctfix[k] = ft[q];
            }
    free(fveln); free(ffix); free(ftf); free(ft);
    return (long)FN;
}

/* ------------------------------------------------------------- receivers */

/* Trilinear receiver times (F:7063-7106).  rcv: (nrc, 3) = height km,
 * colat rad, lon rad.  tout float32; receivers outside the grid get NaN and
 * are reported through the return value (count inside). */
int or_srtimes(const or_grid *g, const double *ttn, int nrc, const double *rcv, float *tout)
{
    int inside = 0;
    const int nx = g->nnx, nz = g->nnz;
    for (int i = 0; i < nrc; ++i) {
        double rr = rcv[3 * i], rx = rcv[3 * i + 1], rz = rcv[3 * i + 2];
        int irr = (int)floor((g->rgor - rr) / g->dnr) + 1;
        int irx = (int)floor((rx - g->rgox) / g->dnx) + 1;
        int irz = (int)floor((rz - g->rgoz) / g->dnz) + 1;
        if (irx == g->nnx) irx--;
        if (irz == g->nnz) irz--;
        if (irr < 1 || irr >= g->nnr || irx < 1 || irx >= g->nnx || irz < 1 || irz >= g->nnz) {
            tout[i] = NAN;
            continue;
        }
        inside++;
        double drx = (rx - g->rgox) - (irx - 1) * g->dnx;
        double drz = (rz - g->rgoz) - (irz - 1) * g->dnz;
        double drr = (g->rgor - rr) - (irr - 1) * g->dnr;
        float acc = 0.0f;
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k)
                for (int l = 0; l < 2; ++l) {
                    double w = (1.0 - fabs((l * g->dnz - drz) / g->dnz)) * (1.0 - fabs((k * g->dnx - drx) / g->dnx));
                    w = w * (1.0 - fabs((j * g->dnr - drr) / g->dnr));
                    double tv = ttn[((size_t)(irr - 1 + j) * nx + (irx - 1 + k)) * nz + (irz - 1 + l)];
                    acc = (float)((double)acc + tv * w);
                }
        tout[i] = acc;
    }
    return inside;
}

/* --------------------------------------------------------- rays + Frechet */

static double tri_vel(const or_grid *g, const double *veln, int ipr, int ipx, int ipz,
                      double drr, double drx, double drz)
{
    /* trilinear velocity in propagation cell (ipr,ipx,ipz) (F:7629-7641) */
    const int MX = g->nnx + 3, MZ = g->nnz + 3;
    double vel = 0.0;
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 2; ++l)
            for (int m = 0; m < 2; ++m) {
                double w = 1.0 - fabs((m * g->dnz - drz) / g->dnz);
                w = w * (1.0 - fabs((l * g->dnx - drx) / g->dnx));
                w = w * (1.0 - fabs((k * g->dnr - drr) / g->dnr));
                int zr = ipr + k, zx = ipx + l, zz = ipz + m;   /* ipX-1+{1,2} */
                if (zz <= g->nnz + 1 && zx <= g->nnx + 1 && zr <= g->nnr + 1)
                    vel += veln[((size_t)(zr + 1) * MX + (zx + 1)) * MZ + (zz + 1)] * w;
            }
    return vel;
}

/* For each receiver, trace a ray down the traveltime gradient and build its
 * Frechet row (F:7291-7766).  Output rows are appended in receiver order:
 * row_ptr[nrc+1] (entries), cols (1-based r-fastest vertex index, ascending),
 * vals; a receiver that starts outside the grid gets an empty row and
 * row_ok[i]=0.  cap bounds cols/vals.  Returns total entries, or -1 if cap
 * is too small.  steps_out (optional) receives the step count per ray. */
static long rpaths(const or_grid *g, const double *veln, const double *ttn, int nrc,
                   const double *rcv, long *row_ptr, int *cols, float *vals, long cap,
                   int *row_ok, int *steps_out, const double *src);

long or_rpaths(const or_grid *g, const double *veln, const double *ttn, int nrc,
               const double *rcv, long *row_ptr, int *cols, float *vals, long cap,
               int *row_ok, int *steps_out)
{
    return rpaths(g, veln, ttn, nrc, rcv, row_ptr, cols, vals, cap, row_ok, steps_out, NULL);
}

/* The same rays for a point source (build-only): a ray within one step of
 * the source (height km, colat, lon) ends with a step to the source, whose
 * Frechet increments use that step's true length. */
long or_rpaths_point(const or_grid *g, const double *veln, const double *ttn, int nrc,
                     const double *rcv, long *row_ptr, int *cols, float *vals, long cap,
                     int *row_ok, int *steps_out, const double *src)
{
    return rpaths(g, veln, ttn, nrc, rcv, row_ptr, cols, vals, cap, row_ok, steps_out, src);
}

static long rpaths(const or_grid *g, const double *veln, const double *ttn, int nrc,
                   const double *rcv, long *row_ptr, int *cols, float *vals, long cap,
                   int *row_ok, int *steps_out, const double *src)
{
    const int nx = g->nnx, nz = g->nnz;
    const int NVR = g->nvr + 2, NVX = g->nvx + 2, NVZ = g->nvz + 2;
    const size_t nv = (size_t)NVR * NVX * NVZ;
    float *fdm = malloc(nv * sizeof(float));
    const int maxrp = 100000;
    const float ftol = 1.0e-6f;
    const double dpl = g->dnr;
    long nent = 0;
#define TT(r, x, z) ttn[((size_t)((r) - 1) * nx + ((x) - 1)) * nz + ((z) - 1)]
#define FDM(iz, ix, ir) fdm[((size_t)(iz) * NVX + (ix)) * NVR + (ir)]
    row_ptr[0] = 0;
    for (int i = 0; i < nrc; ++i) {
        memset(fdm, 0, nv * sizeof(float));
        if (steps_out) steps_out[i] = 0;
        double r0 = rcv[3 * i], x0 = rcv[3 * i + 1], z0 = rcv[3 * i + 2];
        int ipr = (int)floor((g->rgor - r0) / g->dnr) + 1;
        int ipx = (int)floor((x0 - g->rgox) / g->dnx) + 1;
        int ipz = (int)floor((z0 - g->rgoz) / g->dnz) + 1;
        int sw = 0;
        if (ipr < 1 || ipr >= g->nnr) sw = 1;
        if (ipx < 1 || ipx >= g->nnx) sw = 1;
        if (ipz < 1 || ipz >= g->nnz) sw = 1;
        /* a start on the last node uses the cell below it (F:7324-7337) */
        if (ipr == g->nnr) { ipr--; sw = 0; }
        if (ipx == g->nnx) { ipx--; sw = 0; }
        if (ipz == g->nnz) { ipz--; sw = 0; }
        row_ok[i] = !sw;
        if (sw) { row_ptr[i + 1] = nent; continue; }
        double pr = r0, px = x0, pz = z0;
        double xsrc[3] = { 0.0, 0.0, 0.0 };
        if (src) {
            const double rr = g->earth + src[0];
            xsrc[0] = rr * sin(src[1]) * cos(src[2]); xsrc[1] = rr * sin(src[1]) * sin(src[2]); xsrc[2] = rr * cos(src[1]);
        }
        int step;
        for (step = 1; step <= maxrp; ++step) {
            /* gradient on the cell (F:7348-7378) */
            double dtr = 0.0;
            int cnt = 0;
            static const int cz[4] = { 0, 1, 1, 0 }, cx[4] = { 0, 0, 1, 1 };
            for (int q = 0; q < 4; ++q) {
                double lo = TT(ipr + 1, ipx + cx[q], ipz + cz[q]);
                double hi = TT(ipr, ipx + cx[q], ipz + cz[q]);
                /* teleseismic: only edges where t decreases downward (the
                 * reference's rule); a point source: every edge, since rays
                 * climb towards a source above them */
                if (src || lo < hi) { dtr = dtr + lo - hi; cnt++; }
            }
            if (cnt) dtr = dtr / (cnt * g->dnr);
            double dtx = TT(ipr, ipx + 1, ipz) - TT(ipr, ipx, ipz);
            dtx = dtx + TT(ipr, ipx + 1, ipz + 1) - TT(ipr, ipx, ipz + 1);
            dtx = dtx + TT(ipr + 1, ipx + 1, ipz + 1) - TT(ipr + 1, ipx, ipz + 1);
            dtx = dtx + TT(ipr + 1, ipx + 1, ipz) - TT(ipr + 1, ipx, ipz);
            double rcell = g->earth + g->rgor - (ipr - 1) * g->dnr;
            dtx = dtx / (4.0 * rcell * g->dnx);
            double dtz = TT(ipr, ipx, ipz + 1) - TT(ipr, ipx, ipz);
            dtz = dtz + TT(ipr + 1, ipx + 1, ipz + 1) - TT(ipr + 1, ipx + 1, ipz);
            dtz = dtz + TT(ipr, ipx + 1, ipz + 1) - TT(ipr, ipx + 1, ipz);
            dtz = dtz + TT(ipr + 1, ipx, ipz + 1) - TT(ipr + 1, ipx, ipz);
            dtz = dtz / (4.0 * rcell * sin(px) * g->dnz);
            double crad = g->earth + pr;
            double gn = sqrt(dtr * dtr + dtx * dtx + dtz * dtz);
            double nr_ = pr + dpl * dtr / gn;
            double nx_ = px - dpl * dtx / (crad * gn);
            double nz_ = pz - dpl * dtz / (crad * sin(px) * gn);
            double seg = dpl;
            int done = 0;
            if (src) {
                const double sp = sin(px), cp = cos(px);
                const double d0 = crad * sp * cos(pz) - xsrc[0], d1 = crad * sp * sin(pz) - xsrc[1], d2 = crad * cp - xsrc[2];
                const double d = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                if (d <= dpl || !(gn > 0.0)) { nr_ = src[0]; nx_ = src[1]; nz_ = src[2]; seg = d; done = 1; }
            }
            int ipro = ipr, ipxo = ipx, ipzo = ipz;
            ipr = (int)floor((g->rgor - nr_) / g->dnr) + 1;
            ipx = (int)floor((nx_ - g->rgox) / g->dnx) + 1;
            ipz = (int)floor((nz_ - g->rgoz) / g->dnz) + 1;
            if (done) {
                ipr = clampi(ipr, 1, g->nnr - 1);
                ipx = clampi(ipx, 1, g->nnx - 1);
                ipz = clampi(ipz, 1, g->nnz - 1);
            }
            int exr = !done && (ipr < 1 || ipr >= g->nnr);
            int exx = !done && (ipx < 1 || ipx >= g->nnx);
            int exz = !done && (ipz < 1 || ipz >= g->nnz);
            if (exr || exx || exz) {
                /* clip to the first face crossed (F:7463-7538) */
                int which = 0;
                double atio = 0.0;
                if (exr) {
                    double rb;
                    if (ipr < 1) { rb = g->rgor; ipr = 1; }
                    else { rb = g->rgor - (g->nnr - 1) * g->dnr; ipr = g->nnr - 1; }
                    atio = (rb - pr) / (nr_ - pr);
                    which = 1;
                }
                if (exx) {
                    double xb;
                    if (ipx < 1) { xb = g->rgox; ipx = 1; }
                    else { xb = g->rgox + (g->nnx - 1) * g->dnx; ipx = g->nnx - 1; }
                    double q = (xb - px) / (nx_ - px);
                    if (which == 1) { if (q < atio) { atio = q; which = 2; } }
                    else { atio = q; which = 2; }
                }
                if (exz) {
                    double zb;
                    if (ipz < 1) { zb = g->rgoz; ipz = 1; }
                    else { zb = g->rgoz + (g->nnz - 1) * g->dnz; ipz = g->nnz - 1; }
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
            /* Frechet contribution of this step (F:7542-7737) */
            int ivr = (int)floor((g->gor - nr_) / g->dvr) + 1;
            int ivx = (int)floor((nx_ - g->gox) / g->dvx) + 1;
            int ivz = (int)floor((nz_ - g->goz) / g->dvz) + 1;
            int ivro = (int)floor((g->gor - pr) / g->dvr) + 1;
            int ivxo = (int)floor((px - g->gox) / g->dvx) + 1;
            int ivzo = (int)floor((pz - g->goz) / g->dvz) + 1;
            int nhp = 0;
            double vrat[4];
            int chp[4];
            if (ivr != ivro) {
                double rb = (ivr > ivro) ? g->gor - (ivr - 1) * g->dvr : g->gor - ivr * g->dvr;
                vrat[nhp] = (rb - pr) / (nr_ - pr); chp[nhp] = 1; nhp++;
            }
            if (ivx != ivxo) {
                double xb = (ivx > ivxo) ? g->gox + (ivx - 1) * g->dvx : g->gox + ivx * g->dvx;
                double q = (xb - px) / (nx_ - px);
                /* insertion keeps vrat ascending */
                int pos = nhp;
                while (pos > 0 && q < vrat[pos - 1]) { vrat[pos] = vrat[pos - 1]; chp[pos] = chp[pos - 1]; pos--; }
                vrat[pos] = q; chp[pos] = 2; nhp++;
            }
            if (ivz != ivzo) {
                double zb = (ivz > ivzo) ? g->goz + (ivz - 1) * g->dvz : g->goz + ivz * g->dvz;
                double q = (zb - pz) / (nz_ - pz);
                int pos = nhp;
                while (pos > 0 && q < vrat[pos - 1]) { vrat[pos] = vrat[pos - 1]; chp[pos] = chp[pos - 1]; pos--; }
                vrat[pos] = q; chp[pos] = 3; nhp++;
            }
            vrat[nhp] = 1.0; chp[nhp] = 0; nhp++;
            double vel = tri_vel(g, veln, ipro, ipxo, ipzo,
                                 (g->rgor - pr) - (ipro - 1) * g->dnr,
                                 (px - g->rgox) - (ipxo - 1) * g->dnx,
                                 (pz - g->rgoz) - (ipzo - 1) * g->dnz);
            double ui[4], vi[4], wi[4];
            bspline4(((g->gor - pr) - (ivro - 1) * g->dvr) / g->dvr, ui);
            bspline4(((px - g->gox) - (ivxo - 1) * g->dvx) / g->dvx, vi);
            bspline4(((pz - g->goz) - (ivzo - 1) * g->dvz) / g->dvz, wi);
            int ivrt = ivro, ivxt = ivxo, ivzt = ivzo;
            for (int k = 0; k < nhp; ++k) {
                double velo = vel, uio[4], vio[4], wio[4];
                memcpy(uio, ui, sizeof ui); memcpy(vio, vi, sizeof vi); memcpy(wio, wi, sizeof wi);
                if (k > 0) {
                    if (chp[k - 1] == 1) ivrt = ivr;
                    else if (chp[k - 1] == 2) ivxt = ivx;
                    else if (chp[k - 1] == 3) ivzt = ivz;
                }
                double sz_ = pz + vrat[k] * (nz_ - pz);
                double sx_ = px + vrat[k] * (nx_ - px);
                double sr_ = pr + vrat[k] * (nr_ - pr);
                int iprt = (int)floor((g->rgor - sr_) / g->dnr) + 1;
                int ipxt = (int)floor((sx_ - g->rgox) / g->dnx) + 1;
                int ipzt = (int)floor((sz_ - g->rgoz) / g->dnz) + 1;
                vel = tri_vel(g, veln, iprt, ipxt, ipzt,
                              (g->rgor - sr_) - (iprt - 1) * g->dnr,
                              (sx_ - g->rgox) - (ipxt - 1) * g->dnx,
                              (sz_ - g->rgoz) - (ipzt - 1) * g->dnz);
                bspline4(((g->gor - sr_) - (ivrt - 1) * g->dvr) / g->dvr, ui);
                bspline4(((sx_ - g->gox) - (ivxt - 1) * g->dvx) / g->dvx, vi);
                bspline4(((sz_ - g->goz) - (ivzt - 1) * g->dvz) / g->dvz, wi);
                double dinc = (k == 0) ? vrat[0] * seg : (vrat[k] - vrat[k - 1]) * seg;
                /* trapezoid of w/v^2 at both sub-segment ends, added to the
                 * 64 taps of the current vertex cell (F:7717-7732) */
                for (int n = 0; n < 4; ++n)
                    for (int m = 0; m < 4; ++m)
                        for (int l = 0; l < 4; ++l) {
                            double e1 = ui[n] * vi[m] * wi[l] / (vel * vel);
                            double e0 = uio[n] * vio[m] * wio[l] / (velo * velo);
                            double inc = -(e1 + e0) * dinc / 2.0;
                            /* taps iv-1 .. iv+2 of a 0-based vertex array */
                            int zz = ivzt - 1 + l, xx = ivxt - 1 + m, rr = ivrt - 1 + n;
                            if (zz < 0 || zz >= NVZ || xx < 0 || xx >= NVX || rr < 0 || rr >= NVR) continue;
                            FDM(zz, xx, rr) = (float)(inc + (double)FDM(zz, xx, rr));
                        }
            }
            pr = nr_; px = nx_; pz = nz_;
            if (done) break;
        }
        if (steps_out) steps_out[i] = step;
        /* sparsify and emit in r-fastest column order (F:7739-7764) */
        for (int iz = 0; iz < NVZ; ++iz)
            for (int ix = 0; ix < NVX; ++ix)
                for (int ir = 0; ir < NVR; ++ir) {
                    float v = FDM(iz, ix, ir);
                    if (fabsf(v) >= ftol) {
                        if (nent >= cap) { free(fdm); return -1; }
                        cols[nent] = (int)(((size_t)iz * NVX + ix) * NVR + ir) + 1;
                        vals[nent] = v;
                        nent++;
                    }
                }
        row_ptr[i + 1] = nent;
    }
#undef TT
#undef FDM
    free(fdm);
    return nent;
}

/* ------------------------------------------------ model slices (gmtslicet) */

/* B-spline basis as gmtplot/gmtslicet.f90 writes it (lines 283-300, 346-375):
 * (1-u)^3/6, (4-6u^2+3u^3)/6, (1+3u+3u^2-3u^3)/6, u^3/6 */
static void slice_basis(double u, double w[4])
{
    w[0] = (1.0 - u) * (1.0 - u) * (1.0 - u) / 6.0;
    w[1] = (4.0 - 6.0 * (u * u) + 3.0 * (u * u * u)) / 6.0;
    w[2] = (1.0 + 3.0 * u + 3.0 * (u * u) - 3.0 * (u * u * u)) / 6.0;
    w[3] = u * u * u / 6.0;
}

/* Velocity (perturbation) slices of gmtslicet (gmtplot/gmtslicet.f90:4-540):
 * the vertex model minus the reference vertex model (vref, or none), the
 * cubic B-spline on the propagation grid with a one-node border in theta
 * and phi (velv, lines 273-306), then a second B-spline diced d1 x d2 per
 * cell on a plane of it:
 *   kind 0  depth slice at height `where` km, dicing (theta, phi)  (327-386)
 *   kind 1  N-S slice at longitude `where` deg, dicing (r, theta)  (398-460)
 *   kind 2  E-W slice at latitude `where` deg, dicing (r, phi)     (470-531)
 * out receives the values in the order gmtslicet writes them (outer loop
 * over the second slice axis, inner over the first from its far end);
 * n1, n2 = the first and second axis lengths.  Vertex arrays are
 * (nvz+2, nvx+2, nvr+2) with r fastest.  Returns 0, or -1 when a tap falls
 * outside the arrays (gmtslicet reads out of bounds there). */
int or_slice(const or_grid *g, const double deg[4], const double *vtrue, const double *vref, int kind,
             double where, int d1, int d2, double *out, int *n1, int *n2)
{
    /* deg: the .vtx header's propagation origin and spacing in degrees,
     * (lat of the N edge, lat spacing, lon of the W edge, lon spacing):
     * the N-S and E-W slices are placed with them (lines 398, 470) */
    const int NR = g->nvr + 2, NX = g->nvx + 2, NZ = g->nvz + 2;
    const int rr = g->nnr, rt = g->nnx, rp = g->nnz;
    const size_t nv = (size_t)NR * NX * NZ;
    double *vv = malloc(nv * sizeof(double));
    for (size_t i = 0; i < nv; ++i) vv[i] = vtrue[i] - (vref ? vref[i] : 0.0);
    /* velv(k = 1..rnnr, j = 0..rnnt+1, i = 0..rnnp+1) */
    const int VJ = rt + 2, VI = rp + 2;
#define VELV(k, j, i) velv[((size_t)((i)) * VJ + (j)) * rr + ((k) - 1)]
    double *velv = malloc((size_t)rr * VJ * VI * sizeof(double));
    int bad = 0;
    for (int k = 1; k <= rr; ++k) {
        const int ri = (int)((g->gor - g->rgor + (k - 1) * g->dnr) / g->dvr) + 1;
        const double distr = -(g->rgor - (k - 1) * g->dnr) + (g->gor - (ri - 1) * g->dvr);
        double ai[4], bi[4], ci[4];
        slice_basis(distr / g->dvr, ai);
        for (int j = 0; j <= rt + 1; ++j) {
            const int xi = (int)((g->rgox + (j - 1) * g->dnx - g->gox) / g->dvx) + 1;
            const double distx = g->rgox + (j - 1) * g->dnx - g->gox - (xi - 1) * g->dvx;
            slice_basis(distx / g->dvx, bi);
            for (int i = 0; i <= rp + 1; ++i) {
                const int zi = (int)((g->rgoz + (i - 1) * g->dnz - g->goz) / g->dvz) + 1;
                const double distz = g->rgoz + (i - 1) * g->dnz - g->goz - (zi - 1) * g->dvz;
                slice_basis(distz / g->dvz, ci);
                if (ri - 1 < 0 || ri + 2 >= NR || xi - 1 < 0 || xi + 2 >= NX || zi - 1 < 0 || zi + 2 >= NZ) {
                    bad = 1;
                    VELV(k, j, i) = 0.0;
                    continue;
                }
                double sumi = 0.0;
                for (int i1 = 0; i1 < 4; ++i1) {
                    double sumj = 0.0;
                    for (int j1 = 0; j1 < 4; ++j1) {
                        double sumk = 0.0;
                        const double *col = vv + ((size_t)(zi - 1 + i1) * NX + (xi - 1 + j1)) * NR;
                        for (int k1 = 0; k1 < 4; ++k1) sumk = sumk + ai[k1] * col[ri - 1 + k1];
                        sumj = sumj + bi[j1] * sumk;
                    }
                    sumi = sumi + ci[i1] * sumj;
                }
                VELV(k, j, i) = sumi;
            }
        }
    }
    /* dicing weights: ui for the first slice axis, vi for the second, wi across the plane */
    double *ui = malloc((size_t)(d1 + 1) * 4 * sizeof(double)), *vi = malloc((size_t)(d2 + 1) * 4 * sizeof(double));
    for (int i = 1; i <= d1 + 1; ++i) slice_basis((double)(i - 1) / (double)d1, ui + 4 * (i - 1));
    for (int i = 1; i <= d2 + 1; ++i) slice_basis((double)(i - 1) / (double)d2, vi + 4 * (i - 1));
    double wi[4];
    int rnode;
    if (kind == 0) {
        rnode = (int)((g->rgor - where) / g->dnr) + 1;
        slice_basis(fabs(g->rgor - where - (rnode - 1) * g->dnr) / g->dnr, wi);
        *n1 = (rt - 1) * d1 + 1; *n2 = (rp - 1) * d2 + 1;
    } else if (kind == 1) {
        rnode = (int)((where - deg[2]) / deg[3]) + 1;
        slice_basis(fabs(where - deg[2] - (rnode - 1) * deg[3]) / deg[3], wi);
        *n1 = (rr - 3) * d1 + 1; *n2 = (rt - 1) * d2 + 1;
    } else {
        rnode = (int)((deg[0] - where) / deg[1]) + 1;
        slice_basis(fabs(deg[0] - where - (rnode - 1) * deg[1]) / deg[1], wi);
        *n1 = (rr - 3) * d1 + 1; *n2 = (rp - 1) * d2 + 1;
    }
    const int nx1 = *n1;
    for (int p = 0; p < nx1 * *n2; ++p) out[p] = 0.0;
#define OUT(a, b) out[(size_t)((b) - 1) * nx1 + (nx1 - (a))]
    if (kind == 0) {
        if (rnode - 1 < 1 || rnode + 2 > rr) bad = 1;
        else
            for (int i = 1; i <= rp - 1; ++i) {
                const int conp = i == rp - 1 ? d2 + 1 : d2;
                for (int j = 1; j <= rt - 1; ++j) {
                    const int cont = j == rt - 1 ? d1 + 1 : d1;
                    for (int l = 1; l <= conp; ++l)
                        for (int m = 1; m <= cont; ++m) {
                            double sumi = 0.0;
                            for (int i1 = 1; i1 <= 4; ++i1) {
                                double sumj = 0.0;
                                for (int j1 = 1; j1 <= 4; ++j1) {
                                    double sumk = 0.0;
                                    for (int k1 = 1; k1 <= 4; ++k1)
                                        sumk = sumk + wi[k1 - 1] * VELV(rnode - 2 + k1, j - 2 + j1, i - 2 + i1);
                                    sumj = sumj + ui[4 * (m - 1) + j1 - 1] * sumk;
                                }
                                sumi = sumi + vi[4 * (l - 1) + i1 - 1] * sumj;
                            }
                            OUT(d1 * (j - 1) + m, d2 * (i - 1) + l) = sumi;
                        }
                }
            }
    } else if (kind == 1) {
        if (rnode - 1 < 0 || rnode + 2 > rp + 1) bad = 1;
        else
            for (int j = 1; j <= rt - 1; ++j) {
                const int cont = j == rt - 1 ? d2 + 1 : d2;
                for (int k = 2; k <= rr - 2; ++k) {
                    const int conr = k == rr - 2 ? d1 + 1 : d1;
                    for (int m = 1; m <= cont; ++m)
                        for (int n = 1; n <= conr; ++n) {
                            double sumi = 0.0;
                            for (int i1 = 1; i1 <= 4; ++i1) {
                                double sumj = 0.0;
                                for (int j1 = 1; j1 <= 4; ++j1) {
                                    double sumk = 0.0;
                                    for (int k1 = 1; k1 <= 4; ++k1)
                                        sumk = sumk + ui[4 * (n - 1) + k1 - 1] * VELV(k - 2 + k1, j - 2 + j1, rnode - 2 + i1);
                                    sumj = sumj + vi[4 * (m - 1) + j1 - 1] * sumk;
                                }
                                sumi = sumi + wi[i1 - 1] * sumj;
                            }
                            OUT(d1 * (k - 2) + n, d2 * (j - 1) + m) = sumi;
                        }
                }
            }
    } else {
        if (rnode - 1 < 0 || rnode + 2 > rt + 1) bad = 1;
        else
            for (int i = 1; i <= rp - 1; ++i) {
                const int conp = i == rp - 1 ? d2 + 1 : d2;
                for (int k = 2; k <= rr - 2; ++k) {
                    const int conr = k == rr - 2 ? d1 + 1 : d1;
                    for (int l = 1; l <= conp; ++l)
                        for (int n = 1; n <= conr; ++n) {
                            double sumi = 0.0;
                            for (int i1 = 1; i1 <= 4; ++i1) {
                                double sumj = 0.0;
                                for (int j1 = 1; j1 <= 4; ++j1) {
                                    double sumk = 0.0;
                                    for (int k1 = 1; k1 <= 4; ++k1)
                                        sumk = sumk + ui[4 * (n - 1) + k1 - 1] * VELV(k - 2 + k1, rnode - 2 + j1, i - 2 + i1);
                                    sumj = sumj + wi[j1 - 1] * sumk;
                                }
                                sumi = sumi + vi[4 * (l - 1) + i1 - 1] * sumj;
                            }
                            OUT(d1 * (k - 2) + n, d2 * (i - 1) + l) = sumi;
                        }
                }
            }
    }
#undef OUT
#undef VELV
    free(vv); free(velv); free(ui); free(vi);
    return bad ? -1 : 0;
}

/* ----------------------------------------- subinv's sparse products (float32) */

/* y = G^T x as subinv computes it: the transpose built column by column
 * with a per-row cursor (subinv.f90:457-475: for each column, every row
 * whose next entry is in that column gives its entry, rows ascending),
 * then per column a single-precision sum in that order (subinv.f90:612-616,
 * 773-776).  G: CSR with 1-based columns.  Entries of a row that are not in
 * ascending column order are never reached, as in subinv. */
void or_gmat_gtx(long nrows, long ncols, const long *row_ptr, const int *cols, const float *vals,
                 const float *x, float *y)
{
    long *stpv = malloc((nrows + 1) * sizeof(long));
    for (long j = 0; j < nrows; ++j) stpv[j] = 1;
    for (long i = 1; i <= ncols; ++i) {
        float rsum1 = 0.0f;
        for (long j = 1; j <= nrows; ++j) {
            const long n = row_ptr[j] - row_ptr[j - 1];
            if (stpv[j - 1] <= n) {
                const long k = row_ptr[j - 1] + stpv[j - 1] - 1;
                if (cols[k] == i) {
                    rsum1 = rsum1 + vals[k] * x[j - 1];
                    stpv[j - 1]++;
                }
            }
        }
        y[i - 1] = rsum1;
    }
    free(stpv);
}

/* y = (G x) / cd as subinv computes it: per row a single-precision sum in
 * the row's order, divided by cd (subinv.f90:759-767); cd NULL: 1. */
void or_gmat_gx(long nrows, const long *row_ptr, const int *cols, const float *vals, const float *x,
                const float *cd, float *y)
{
    for (long j = 0; j < nrows; ++j) {
        float rsum1 = 0.0f;
        for (long k = row_ptr[j]; k < row_ptr[j + 1]; ++k) rsum1 = rsum1 + vals[k] * x[cols[k] - 1];
        y[j] = cd ? rsum1 / cd[j] : rsum1;
    }
}
