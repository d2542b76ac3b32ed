// mhd_magneto.h -- the step in magnetosphere mode (included by mhd_gpu.hip).
//
// The reference's own default case (t3dsdsc.f with mhd03.for's settings): a
// dipole magnetosphere with an inner-sphere mask rr1 <= rcr at the start of
// every boundary routine, a driven inflow at x-left (t3dsdsc.f:1711-1772), a
// radial Hall coefficient (1398-1399) and, with totalE = .false., an electric
// field of the field perturbation plus the curl(v x B0) dipole terms
// (1370-1386, 1460-1548).  These kernels follow the reference's routine
// structure one sweep per launch: a producer writes the interior, then a fill
// runs the x rules (one thread per row) and the y and z rules (one launch) as
// separate launches, so the rules see each other's results in the reference's
// x, y, z order.  The inner-sphere mask at the start of every fill is applied
// by the producers as they store (the masked points are interior points; the
// host falls back to a mask launch per fill where they are not).
// Non-periodic y only.
// Every non-interior point of a buffer is rewritten by the fill that follows
// its producer, and with the auxiliary and state fills every value formed from
// a non-interior point is rewritten again by a later rule of the same fill, so
// producers write interior points only; before a bndrytt fill, whose y and z
// rules skip the x-left columns, sweeps carry the other points of their input
// along, as the reference's in-place sweep leaves them.
#pragma once

#include "mhd_device.h"

namespace mhd {

// single-precision Fortran literal as the compiler stores it
#define MF32(v) ((double)(float)(v))

struct Mag {
    const double *rr;          // |(x, y, z)| per point (t3dsdsc.f:773-774)
    const double *cht;         // Hall coefficient per point, or null (uniform di)
    double rcr;                // mask radius (mhd03.for rcr)
    int mask;                  // the inner-sphere mask at the start of every fill
    int inflow;                // driven x-left inflow in the state fill
    int partial;               // totalE = .false.
    const double *zd, *yd;     // inflow profiles tanh((z+0.6)/0.15) - tanh((z-0.6)/0.15), y alike (0.4)
    const double *zd2, *yd2;   // the same with 0.2 (z) and 0.3 (y) for the velocity
    const double *fb;          // face data bxinfc, byinfc, bzinfc: 3 x NZ*NY, [k*NY + j]
    const int *mrow;           // masked columns [lo, hi) of row (j, k) at [2 (k*NY + j)], or null
};

// the producers' store of the mask: the point is inside the sphere
__device__ inline bool mg_masked(const Geo &g, const int *mrow, int i, int j, int k)
{
    if (!mrow) return false;
    const int *r = mrow + 2 * ((long)k * g.NY + j);
    return i >= r[0] && i < r[1];
}

// arguments t, ddt, km of the running right(): the time factors of the drive
// (computed on the host), the reference state of the quotient and the stage
// state whose driven columns bndry overwrites
struct Stage {
    double tt;                 // tanh((t-9)/1) - tanh((t-39)/1.5)
    double sq;                 // sqrt(t^3 7.25^3)
    double de;                 // exp(7.25 t 0.0069f) + 1
    double ddt;
    double rde;                // 1 / de, and ddt + 1e-9 with its reciprocal: the
    double den, rden;          // drive's quotients through divc
    const double *ref;         // xfold (km = 1) or xm (km = 2)
    double *x;
};

__device__ inline bool mg_point(const Geo &g, int &i, int &j, int &k, long &q)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NY * g.NZ) return false;
    i = (int)(t % g.NX);
    const long r = t / g.NX;
    j = (int)(r % g.NY);
    k = (int)(r / g.NY);
    q = IDX(g, i, j, k);
    return true;
}

__device__ inline bool mg_inner(const Geo &g, int i, int j, int k)
{
    return i >= 2 && i <= g.NX - 3 && j >= 2 && j <= g.NY - 3 && k >= 2 && k <= g.NZ - 3;
}

// 5-point first derivatives of a point function F along x, y, z at (i, j, k)
template <class F>
__device__ inline double mg_dx(const Geo &g, int i, long q, F &&f)
{
    return fd5(f(q - 2), f(q - 1), f(q), f(q + 1), f(q + 2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
}
template <class F>
__device__ inline double mg_dy(const Geo &g, int j, long q, F &&f)
{
    const long s = g.sy;
    return fd5(f(q - 2 * s), f(q - s), f(q), f(q + s), f(q + 2 * s), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
}
template <class F>
__device__ inline double mg_dz(const Geo &g, int k, long q, F &&f)
{
    const long s = g.sz;
    return fd5(f(q - 2 * s), f(q - s), f(q), f(q + s), f(q + 2 * s), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
}

// ---------------------------------------------------------------------------
// fills: mask, then x, y, z rules
// ---------------------------------------------------------------------------
// mask of bndry/bndrytt/bndryc/bndryh/bndryE (t3dsdsc.f:1692-1707,
// 2003-2017, 2549-2559, 2709-2719, 2881-2891): zero where rr1 <= rcr
__global__ void km_mask(Geo g, Mag a, double *__restrict__ A, int nf)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q)) return;
    if (a.rr[q] <= a.rcr)
        for (int m = 0; m < nf; m++) A[q + m * g.S] = 0.0;
}

// driven inflow of bndry at column i (0: jx = 3, 1: jx = 4), row j, plane k,
// t3dsdsc.f:1714-1772 (density base rhoi at the point itself)
// the driven value of field m at (i, j, k)
__device__ inline double mg_inflow_v(const Geo &g, const Mag &a, const Stage &s, const double *__restrict__ eq,
                                     int m, int j, int k, long q)
{
    const long S = g.S, f = (long)k * g.NY + j, F = (long)g.NZ * g.NY;
    if (m == 0 || m == 1) {
        const double v0 = eq[q] + 1. * 0.5 * s.tt * 0.5 * a.zd[k] * 0.5 * a.yd[j];
        return m == 0 ? v0 : MF32(0.8) * v0 * 0.5 * s.tt * 0.5 * a.zd2[k] * 0.5 * a.yd2[j];
    }
    if (m <= 3) return 0.;
    if (m <= 6) return eq[q + (m - 3) * S] + divc(a.fb[f + (m - 4) * F] * MF32(0.0017) * s.sq, s.de, s.rde);
    return eq[q + 4 * S] + MF32(0.2) * 0.5 * s.tt * 0.5 * a.zd[k] * 0.5 * a.yd[j];
}

// its xdif: (x_bc - x_ref) / (ddt + 1e-9)
__device__ inline double mg_inflow_xd(const Geo &g, const Mag &a, const Stage &s, const double *__restrict__ eq,
                                      int m, int j, int k, long q)
{
    return divc(mg_inflow_v(g, a, s, eq, m, j, k, q) - s.ref[q + m * g.S], s.den, s.rden);
}

__device__ inline void mg_inflow(const Geo &g, const Mag &a, const Stage &s, const double *__restrict__ eq,
                                 double *__restrict__ XD, int i, int j, int k)
{
    const long q = IDX(g, i, j, k), S = g.S;
    for (int m = 0; m < 8; m++) {
        const double v = mg_inflow_v(g, a, s, eq, m, j, k, q);
        s.x[q + m * S] = v;
        XD[q + m * S] = divc(v - s.ref[q + m * S], s.den, s.rden);
    }
}

// x rules over every row and plane: BC_STATE (bndry: 0 / half at x-right,
// the mirrored rule or the driven inflow at x-left), BC_AUX and BC_TT
// (bndryc/h/E, bndrytt: one-sided extrapolation at x-left).  One thread per
// row and field; with the inflow, one more per row and driven column (the
// driven points read only eq, the reference state and the drive, and no
// rule of this launch writes columns 0 and 1 then)
template <int TYPE>
__global__ void km_fill_x(Geo g, Mag a, Stage s, const double *__restrict__ eq, double *__restrict__ A, int nf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long rows = (long)g.NY * g.NZ, nr = rows * nf;
    if (t >= nr) {
        if (TYPE == BC_STATE && a.inflow && t < nr + 2 * rows) {
            const long u = t - nr, row = u % rows;
            mg_inflow(g, a, s, eq, A, u < rows ? 1 : 0, (int)(row % g.NY), (int)(row / g.NY));
        }
        return;
    }
    const int m = (int)(t / rows);
    const long row = t % rows;
    const int j = (int)(row % g.NY), k = (int)(row / g.NY);
    const int L = g.NX - 1;
    double *f = A + IDX(g, 0, j, k) + m * g.S;
    if (TYPE == BC_STATE) {
        if (!a.inflow) {
            f[0] = 0.;
            f[1] = 0.5 * f[2];
        }
    } else {
        const double v1 = (g.ax4 * f[2] + g.bx4 * f[3] + g.cx4 * f[4]) / g.dx4;
        f[1] = v1;
        f[0] = (g.ax3 * v1 + g.bx3 * f[2] + g.cx3 * f[3]) / g.dx3;
    }
    f[L] = 0.;
    f[L - 1] = 0.5 * f[L - 2];
}

// y rules (non-periodic, 0 / half) over every plane; x range all columns,
// bndrytt from jx = x_first (t3dsdsc.f:1800-1819, 2108-2128, 2592-2612)
template <int TYPE>
__global__ void km_fill_y(Geo g, double *__restrict__ A, int nf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NZ) return;
    const int i = (int)(t % g.NX), k = (int)(t / g.NX);
    if (TYPE == BC_TT && i < 2) return;
    const long q0 = IDX(g, i, 0, k), sy = g.sy;
    const int L = g.NY - 1;
    for (int m = 0; m < nf; m++) {
        double *f = A + q0 + m * g.S;
        f[0] = 0.;
        f[sy] = 0.5 * f[2 * sy];
        f[L * sy] = 0.;
        f[(L - 1) * sy] = 0.5 * f[(L - 2) * sy];
    }
}

// z rules (0 / half) over every row (t3dsdsc.f:1821-1861, 2130-2168, 2614-2644)
template <int TYPE>
__global__ void km_fill_z(Geo g, double *__restrict__ A, int nf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)g.NX * g.NY) return;
    const int i = (int)(t % g.NX), j = (int)(t / g.NX);
    if (TYPE == BC_TT && i < 2) return;
    const long q0 = IDX(g, i, j, 0), sz = g.sz;
    const int L = g.NZ - 1;
    for (int m = 0; m < nf; m++) {
        double *f = A + q0 + m * g.S;
        f[0] = 0.;
        f[sz] = 0.5 * f[2 * sz];
        f[L * sz] = 0.;
        f[(L - 1) * sz] = 0.5 * f[(L - 2) * sz];
    }
}

// y and z rules in one launch.  Threads [0, NX (NZ-4)) apply the y rules on
// the interior planes; the rest, one per x-y column, the z rules, taking
// plane 2 (NZ-3) of an image row as the y rule makes it from row 2 (NY-3)
// (0 on the outer rows).  The values are those of the y launch followed by
// the z launch: the y rule's values on the z-face planes of the image rows
// are rewritten by the z rules there, and no thread reads what another
// writes (t3dsdsc.f:1800-1861, 2108-2168, 2592-2644)
template <int TYPE>
__global__ void km_fill_yz(Geo g, double *__restrict__ A, int nf)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long ny = (long)g.NX * (g.NZ - 4), nz = (long)g.NX * g.NY;
    const long sy = g.sy, sz = g.sz;
    if (t < ny * nf) {
        const int m = (int)(t / ny);
        const long u = t % ny;
        const int i = (int)(u % g.NX), k = 2 + (int)(u / g.NX);
        if (TYPE == BC_TT && i < 2) return;
        const int L = g.NY - 1;
        double *f = A + IDX(g, i, 0, k) + m * g.S;
        f[0] = 0.;
        f[sy] = 0.5 * f[2 * sy];
        f[L * sy] = 0.;
        f[(L - 1) * sy] = 0.5 * f[(L - 2) * sy];
        return;
    }
    const long v = t - ny * nf;
    if (v >= nz * nf) return;
    const int m = (int)(v / nz);
    const long u = v % nz;
    const int i = (int)(u % g.NX), j = (int)(u / g.NX);
    if (TYPE == BC_TT && i < 2) return;
    const int L = g.NZ - 1, M = g.NY - 1;
    double *f = A + IDX(g, i, j, 0) + m * g.S;
    double p2, pl;
    if (j == 0 || j == M) {
        p2 = pl = 0.;
    } else if (j == 1 || j == M - 1) {
        const double *s = A + IDX(g, i, j == 1 ? 2 : M - 2, 0) + m * g.S;
        p2 = 0.5 * s[2 * sz];
        pl = 0.5 * s[(L - 2) * sz];
    } else {
        p2 = f[2 * sz];
        pl = f[(L - 2) * sz];
    }
    f[0] = 0.;
    f[sz] = 0.5 * p2;
    f[L * sz] = 0.;
    f[(L - 1) * sz] = 0.5 * pl;
}

// the whole auxiliary or state fill (x, then y, then z rules; not bndrytt) in
// one launch.  Threads over the rows take the x rules of the interior rows
// and planes (and, with the inflow, the driven state of every row), threads
// over the x-z columns the y rules on the interior planes, threads over the
// x-y columns the z rules.  A thread that needs a value another rule of this
// fill forms -- an x-face column of row 2 or plane 2, say -- forms it itself
// with the same arithmetic, so no thread reads what another writes and the
// values are those of the x, y and z launches in turn.
template <int TYPE>
__device__ inline double mg_xrule(const Geo &g, const Mag &a, const Stage &s, const double *__restrict__ eq,
                                  const double *__restrict__ f, int m, int i, int j, int k)
{
    // f: field m at (0, j, k); i an x-face column (0, 1, NX-2, NX-1)
    const int L = g.NX - 1;
    if (i == L) return 0.;
    if (i == L - 1) return 0.5 * f[L - 2];
    if (TYPE == BC_STATE) {
        if (a.inflow) return mg_inflow_xd(g, a, s, eq, m, j, k, IDX(g, i, j, k));
        return i == 0 ? 0. : 0.5 * f[2];
    }
    const double v1 = (g.ax4 * f[2] + g.bx4 * f[3] + g.cx4 * f[4]) / g.dx4;
    return i == 1 ? v1 : (g.ax3 * v1 + g.bx3 * f[2] + g.cx3 * f[3]) / g.dx3;
}

template <int TYPE>
__global__ void km_fill_xyz(Geo g, Mag a, Stage s, const double *__restrict__ eq, double *__restrict__ A, int nf)
{
    static_assert(TYPE == BC_AUX || TYPE == BC_STATE, "bndrytt keeps its launches");
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const int L = g.NX - 1, M = g.NY - 1, N = g.NZ - 1;
    const long rows = (long)g.NY * g.NZ;
    auto xface = [&](int i) { return i < 2 || i > L - 2; };
    // value of field m at (i, j, k) once the x rules have run (j, k interior)
    auto after_x = [&](int m, int i, int j, int k) {
        const double *f = A + IDX(g, 0, j, k) + m * S;
        return xface(i) ? mg_xrule<TYPE>(g, a, s, eq, f, m, i, j, k) : f[i];
    };
    // ... and the y rules (k interior)
    auto after_y = [&](int m, int i, int j, int k) {
        if (j == 0 || j == M) return 0.;
        if (j == 1 || j == M - 1) return 0.5 * after_x(m, i, j == 1 ? 2 : M - 2, k);
        return after_x(m, i, j, k);
    };
    // x rules: interior rows and planes, one thread per row and field
    if (t < rows * nf) {
        const int m = (int)(t / rows);
        const long row = t % rows;
        const int j = (int)(row % g.NY), k = (int)(row / g.NY);
        if (j < 2 || j > M - 2 || k < 2 || k > N - 2) return;
        double *f = A + IDX(g, 0, j, k) + m * S;
        const bool driven = TYPE == BC_STATE && a.inflow;
        double v[4];
        #pragma unroll
        for (int c = 0; c < 4; c++) {
            const int i = c < 2 ? c : L - 3 + c;
            v[c] = driven && i < 2 ? 0. : mg_xrule<TYPE>(g, a, s, eq, f, m, i, j, k);
        }
        if (!driven) { f[0] = v[0]; f[1] = v[1]; }
        f[L - 1] = v[2];
        f[L] = v[3];
        return;
    }
    t -= rows * nf;
    // the driven columns: the state on every row, xdif on the interior ones
    if (TYPE == BC_STATE && a.inflow) {
        if (t < 2 * rows) {
            const long row = t % rows;
            const int i = t < rows ? 1 : 0, j = (int)(row % g.NY), k = (int)(row / g.NY);
            const long q = IDX(g, i, j, k);
            const bool in = j >= 2 && j <= M - 2 && k >= 2 && k <= N - 2;
            for (int m = 0; m < 8; m++) {
                const double v = mg_inflow_v(g, a, s, eq, m, j, k, q);
                s.x[q + m * S] = v;
                if (in) A[q + m * S] = divc(v - s.ref[q + m * S], s.den, s.rden);
            }
            return;
        }
        t -= 2 * rows;
    }
    // y rules on the interior planes, one thread per x-z column and field
    const long ny = (long)g.NX * (g.NZ - 4);
    if (t < ny * nf) {
        const int m = (int)(t / ny);
        const long u = t % ny;
        const int i = (int)(u % g.NX), k = 2 + (int)(u / g.NX);
        double *f = A + IDX(g, i, 0, k) + m * S;
        f[0] = 0.;
        f[sy] = 0.5 * after_x(m, i, 2, k);
        f[M * sy] = 0.;
        f[(M - 1) * sy] = 0.5 * after_x(m, i, M - 2, k);
        return;
    }
    t -= ny * nf;
    // z rules, one thread per x-y column and field
    const long nz = (long)g.NX * g.NY;
    if (t >= nz * nf) return;
    const int m = (int)(t / nz);
    const long u = t % nz;
    const int i = (int)(u % g.NX), j = (int)(u / g.NX);
    double *f = A + IDX(g, i, j, 0) + m * S;
    f[0] = 0.;
    f[sz] = 0.5 * after_y(m, i, j, 2);
    f[N * sz] = 0.;
    f[(N - 1) * sz] = 0.5 * after_y(m, i, j, N - 2);
}

// ---------------------------------------------------------------------------
// one smoothing sweep of nf fields A -> B on the interior (t3dsdsc.f:5630-5680):
// KIND_RADIAL, KIND_EIGHTH, or KIND_STATE (B components 4..6 take cbx).
// The reference sweeps in place; before a bndrytt fill (carry) B takes A's
// other points, since its x rules read face rows and planes that its y and z
// rules, which skip jx < 5, do not all rewrite.
// ---------------------------------------------------------------------------
template <int KIND>
__global__ void km_sweep(Geo g, const double *__restrict__ A, double *__restrict__ B, int nf, const int *mrow,
                         int carry)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q)) return;
    if (!mg_inner(g, i, j, k)) {
        // only a bndrytt fill keeps values formed from non-interior points
        if (carry)
            for (int m = 0; m < nf; m++) B[q + m * g.S] = A[q + m * g.S];
        return;
    }
    const long sy = g.sy, sz = g.sz;
    if (mg_masked(g, mrow, i, j, k)) {
        for (int m = 0; m < nf; m++) B[q + m * g.S] = 0.0;
        return;
    }
    for (int m = 0; m < nf; m++) {
        const double *f = A + m * g.S;
        double c;
        if (KIND == KIND_EIGHTH) c = 1.0 / 8.0;
        else if (KIND == KIND_STATE && m >= 4 && m <= 6) c = g.cbx[i];
        else c = g.crad[q];
        const double f0 = f[q];
        const double s = difw(f[q - 1], f0, f[q + 1], g.xwp[i], g.xwm[i]) +
                         difw(f[q - sy], f0, f[q + sy], g.ywp[j], g.ywm[j]) +
                         difw(f[q - sz], f0, f[q + sz], g.zwp[k], g.zwm[k]);
        B[q + m * g.S] = f0 + c * s;
    }
}

// ---------------------------------------------------------------------------
// current (t3dsdsc.f:2465-2491): J = curl(B - B_i) on the interior
// ---------------------------------------------------------------------------
__global__ void km_curl(Geo g, const double *__restrict__ X, const double *__restrict__ eq, double *__restrict__ cur,
                        const int *mrow)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q) || !mg_inner(g, i, j, k)) return;
    const long S = g.S;
    if (mg_masked(g, mrow, i, j, k)) {
        cur[q] = cur[q + S] = cur[q + 2 * S] = 0.0;
        return;
    }
    auto u = [&](long p) { return X[p + 4 * S] - eq[p + S]; };
    auto v = [&](long p) { return X[p + 5 * S] - eq[p + 2 * S]; };
    auto w = [&](long p) { return X[p + 6 * S] - eq[p + 3 * S]; };
    cur[q] = mg_dy(g, j, q, w) - mg_dz(g, k, q, v);
    cur[q + S] = mg_dz(g, k, q, u) - mg_dx(g, i, q, w);
    cur[q + 2 * S] = mg_dx(g, i, q, v) - mg_dy(g, j, q, u);
}

// ---------------------------------------------------------------------------
// right, first part (t3dsdsc.f:1334-1387, flux 2304-2347): on the interior
// xdif(1..4) = -div flux(m), then - xresd + J x B on the momentum; on every
// point v = m / rho and E (totalE, or the field perturbation only)
// ---------------------------------------------------------------------------
__device__ inline void mg_flux(const double *__restrict__ X, const double *__restrict__ pri, long S, long p, int m,
                               double &f, double &gg, double &h)
{
    const double x1 = X[p], x2 = X[p + S], x3 = X[p + 2 * S], x4 = X[p + 3 * S], x8 = X[p + 7 * S];
    const double pi0 = pri[p];
    if (m == 0) { f = x2; gg = x3; h = x4; }
    else if (m == 1) { f = x2 * x2 / x1 + (x8 - pi0); gg = x2 * x3 / x1; h = x2 * x4 / x1; }
    else if (m == 2) { f = x2 * x3 / x1; gg = x3 * x3 / x1 + (x8 - pi0); h = x4 * x3 / x1; }
    else { f = x2 * x4 / x1; gg = x3 * x4 / x1; h = x4 * x4 / x1 + (x8 - pi0); }
}

// the momentum-flux tensor of mg_flux at every point, formed once for the
// divergences of km_rhs_a: T22, T23, T24, T33, T34, T44 (T34 = x4 x3 / x1,
// equal to mg_flux's x3 x4 / x1)
__global__ void km_flux(Geo g, const double *__restrict__ X, const double *__restrict__ eq, double *__restrict__ T)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q)) return;
    const long S = g.S;
    double f, gg, h, t33, t44;
    mg_flux(X, eq + 4 * S, S, q, 1, f, gg, h);
    T[q] = f; T[q + S] = gg; T[q + 2 * S] = h;
    mg_flux(X, eq + 4 * S, S, q, 2, f, t33, h);
    T[q + 3 * S] = t33; T[q + 4 * S] = h;
    mg_flux(X, eq + 4 * S, S, q, 3, f, gg, t44);
    T[q + 5 * S] = t44;
}

__global__ void km_rhs_a(Geo g, Mag a, const double *__restrict__ X, const double *__restrict__ eq,
                         const double *__restrict__ cur, const double *__restrict__ xresd, double eta,
                         const double *__restrict__ T, double *__restrict__ XD, double *__restrict__ E,
                         double *__restrict__ vel)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q)) return;
    const long S = g.S;
    const double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
    const double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
    const double c1 = cur[q], c2 = cur[q + S], c3 = cur[q + 2 * S];
    if (mg_inner(g, i, j, k)) {
        // flux(m) of t3dsdsc.f:2304-2347: (x2, x3, x4) for the mass, the
        // tensor rows (T22, T23, T24), (T23, T33, T34), (T24, T34, T44)
        const double *fb[4][3] = {{X + S, X + 2 * S, X + 3 * S},
                                  {T, T + S, T + 2 * S},
                                  {T + S, T + 3 * S, T + 4 * S},
                                  {T + 2 * S, T + 4 * S, T + 5 * S}};
        double d[4];
        for (int m = 0; m < 4; m++) {
            const double *fx = fb[m][0], *fy = fb[m][1], *fz = fb[m][2];
            d[m] = -mg_dx(g, i, q, [&](long p) { return fx[p]; }) - mg_dy(g, j, q, [&](long p) { return fy[p]; }) -
                   mg_dz(g, k, q, [&](long p) { return fz[p]; });
        }
        // (masked points: km_rhs_b, which follows, zeroes all eight fields)
        XD[q] = d[0];
        XD[q + S] = d[1] - xresd[q] + c2 * x7 - c3 * x6;
        XD[q + 2 * S] = d[2] - xresd[q + S] + c3 * x5 - c1 * x7;
        XD[q + 3 * S] = d[3] - xresd[q + 2 * S] + c1 * x6 - c2 * x5;
    }
    vel[q] = x2 / x1;
    vel[q + S] = x3 / x1;
    vel[q + 2 * S] = x4 / x1;
    if (!a.partial) {
        E[q] = (-x3 * x7 + x4 * x6) / x1 + eta * c1;
        E[q + S] = (-x4 * x5 + x2 * x7) / x1 + eta * c2;
        E[q + 2 * S] = (-x2 * x6 + x3 * x5) / x1 + eta * c3;
    } else {
        const double b1 = eq[q + S], b2 = eq[q + 2 * S], b3 = eq[q + 3 * S];
        E[q] = (-x3 * (x7 - b3) + x4 * (x6 - b2)) / x1 + eta * c1;
        E[q + S] = (-x4 * (x5 - b1) + x2 * (x7 - b3)) / x1 + eta * c2;
        E[q + 2 * S] = (-x2 * (x6 - b2) + x3 * (x5 - b1)) / x1 + eta * c3;
    }
}

// Hall term on the interior (t3dsdsc.f:1404-1421): cht (J x B - grad p) / rho
__global__ void km_hall(Geo g, Mag a, const double *__restrict__ X, const double *__restrict__ cur, double di,
                        double *__restrict__ H)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q) || !mg_inner(g, i, j, k)) return;
    const long S = g.S;
    if (mg_masked(g, a.mrow, i, j, k)) {
        H[q] = H[q + S] = H[q + 2 * S] = 0.0;
        return;
    }
    const double cht = a.cht ? a.cht[q] : di;
    const double x1 = X[q], x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
    const double c1 = cur[q], c2 = cur[q + S], c3 = cur[q + 2 * S];
    auto p = [&](long r) { return X[r + 7 * S]; };
    H[q] = cht * (c2 * x7 - c3 * x6 - mg_dx(g, i, q, p)) / x1;
    H[q + S] = cht * (c3 * x5 - c1 * x7 - mg_dy(g, j, q, p)) / x1;
    H[q + 2 * S] = cht * (c1 * x6 - c2 * x5 - mg_dz(g, k, q, p)) / x1;
}

// E += xdh on every point (t3dsdsc.f:1426-1432)
__global__ void km_add3(long S, long n, double *__restrict__ E, const double *__restrict__ H)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    for (int m = 0; m < 3; m++) E[q + m * S] = E[q + m * S] + H[q + m * S];
}

// r^7 by binary exponentiation (the expansion of a constant integer power)
__device__ inline double mg_pow7(double x)
{
    const double x2 = x * x;
    return (x * x2) * (x2 * x2);
}

// right, second part (t3dsdsc.f:1435-1566): -curl E (+ the totalE = .false.
// dipole terms) and the pressure equation on the interior
__global__ void km_rhs_b(Geo g, Mag a, const double *__restrict__ X, const double *__restrict__ eq,
                         const double *__restrict__ cur, const double *__restrict__ E,
                         const double *__restrict__ vel, double eta, double gam, double *__restrict__ XD)
{
    int i, j, k;
    long q;
    if (!mg_point(g, i, j, k, q) || !mg_inner(g, i, j, k)) return;
    const long S = g.S;
    if (mg_masked(g, a.mrow, i, j, k)) {   // the mask of the xdif fill, all eight fields
        for (int m = 0; m < 8; m++) XD[q + m * S] = 0.0;
        return;
    }
    auto e1 = [&](long p) { return E[p]; };
    auto e2 = [&](long p) { return E[p + S]; };
    auto e3 = [&](long p) { return E[p + 2 * S]; };
    auto vx = [&](long p) { return vel[p]; };
    auto vy = [&](long p) { return vel[p + S]; };
    auto vz = [&](long p) { return vel[p + 2 * S]; };
    auto pr = [&](long p) { return X[p + 7 * S]; };
    double d5 = -mg_dy(g, j, q, e3) + mg_dz(g, k, q, e2);
    double d6 = -mg_dz(g, k, q, e1) + mg_dx(g, i, q, e3);
    double d7 = -mg_dx(g, i, q, e2) + mg_dy(g, j, q, e1);
    if (a.partial) {
        const double xg = g.xc[i], yg = g.yc[j], zg = g.zc[k];
        const double r0 = a.rr[q], r2 = r0 * r0, r7 = mg_pow7(r0);
        const double b1 = eq[q + S], b2 = eq[q + 2 * S], b3 = eq[q + 3 * S];
        const double u = vel[q], v = vel[q + S], w = vel[q + 2 * S];
        const double dv = mg_dx(g, i, q, vx) + mg_dy(g, j, q, vy) + mg_dz(g, k, q, vz);
        const double q2 = xg * xg + yg * yg - 2. * (zg * zg);
        const double a1 = -3. * zg * (r2 - 5. * (xg * xg)) / r7;
        const double a2 = 3. * xg * zg * 5. * yg / r7;
        const double a3 = -3. * xg * (r2 - 5. * (zg * zg)) / r7;
        const double f1 = 3. * yg * zg * 5. * xg / r7;
        const double f2 = -3. * zg * (r2 - 5. * (yg * yg)) / r7;
        const double f3 = -3. * yg * (r2 - 5. * (zg * zg)) / r7;
        const double h1 = (2. * xg * r2 - 5. * xg * q2) / r7;
        const double h2 = (2. * yg * r2 - 5. * yg * q2) / r7;
        const double h3 = (-4. * zg * r2 - 5. * zg * q2) / r7;
        d5 = d5 + (-b1 * dv + (b1 * mg_dx(g, i, q, vx) + b2 * mg_dy(g, j, q, vx) + b3 * mg_dz(g, k, q, vx)) -
                   (u * a1 + v * a2 + w * a3));
        d6 = d6 + (-b2 * dv + (b1 * mg_dx(g, i, q, vy) + b2 * mg_dy(g, j, q, vy) + b3 * mg_dz(g, k, q, vy)) -
                   (u * f1 + v * f2 + w * f3));
        d7 = d7 + (-b3 * dv + (b1 * mg_dx(g, i, q, vz) + b2 * mg_dy(g, j, q, vz) + b3 * mg_dz(g, k, q, vz)) -
                   (u * h1 + v * h2 + w * h3));
    }
    XD[q + 4 * S] = d5;
    XD[q + 5 * S] = d6;
    XD[q + 6 * S] = d7;
    const double c1 = cur[q], c2 = cur[q + S], c3 = cur[q + 2 * S];
    XD[q + 7 * S] = -mg_dx(g, i, q, pr) * vel[q] - mg_dy(g, j, q, pr) * vel[q + S] - mg_dz(g, k, q, pr) * vel[q + 2 * S] +
                    (-mg_dx(g, i, q, vx) - mg_dy(g, j, q, vy) - mg_dz(g, k, q, vz)) * gam * X[q + 7 * S] +
                    (gam - 1.) * eta * (c1 * c1 + c2 * c2 + c3 * c3);
}

// xm = x - eq (sign -1) or x = xm + eq (sign +1) on every point of the 8 fields
// (t3dsdsc.f:1135-1165; the momentum has no equilibrium part)
__global__ void km_eq(long S, long n, const double *__restrict__ A, const double *__restrict__ eq, int sign,
                      double *__restrict__ B)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    const int em[8] = {0, -1, -1, -1, 1, 2, 3, 4};
    for (int m = 0; m < 8; m++) {
        const double v = A[q + m * S];
        if (em[m] < 0) B[q + m * S] = v;
        else B[q + m * S] = sign < 0 ? v - eq[q + em[m] * S] : v + eq[q + em[m] * S];
    }
}

#undef MF32

} // namespace mhd
