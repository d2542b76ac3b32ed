// mhd_adjoint.h -- kernels of the discrete adjoint (reverse mode) of the
// RK4 step (included by mhd_gpu.hip).
//
// Each kernel is the transpose of one forward operation of the step,
// launched in the reverse order of the forward ones:
//   - linear maps (smoothing sweeps, boundary rules, 5-point derivatives,
//     RK register combinations) are transposed exactly;
//   - pointwise nonlinear terms (fluxes, E = -v x B + eta J, the Hall term,
//     the pressure equation) contribute the transpose of their Jacobian at
//     the forward state, which the caller supplies as the stage state X and
//     its (filtered) current J.
// Transposed derivatives and sweeps are gathers: the thread of point q sums
// w(p -> q) * a(p) over the interior points p whose stencil reads q, with
// the coefficients of p.  One thread per point of the unique range
// (periodic y: rows 2..NY-3, the image rows are implicit, as in the forward
// kernels; non-periodic: every row).  Reference routines: right t3dsdsc.f:1297-1673, current
// 2452-2500, bndry 1677-1985, bndrytt 1989-2284, bndryc/h/E 2539-3027,
// avrg* 5611-6054, stepon 1063-1185.
#pragma once

#include "mhd_device.h"

namespace mhd {

struct AdjPt {
    int i, j, k;
    bool ok;
};

// point of a (64, 4) block over (x, unique rows) and blockIdx.z = plane
__device__ inline AdjPt adj_point(const Geo &g)
{
    AdjPt p;
    p.i = blockIdx.x * blockDim.x + threadIdx.x;
    p.j = g.ylo + blockIdx.y * blockDim.y + threadIdx.y;
    p.k = blockIdx.z;
    p.ok = p.i < g.NX && p.j <= g.yhi;
    return p;
}

__device__ inline bool ax_in(int v, int n) { return v >= 2 && v <= n - 3; }

// interior point (y rows of the unique range are interior when periodic)
__device__ inline bool in3(const Geo &g, int i, int j, int k)
{
    return ax_in(i, g.NX) && ax_in(k, g.NZ) && (g.periodic || ax_in(j, g.NY));
}

// weight of the input at offset d = q - p in fd5 at output point p
__device__ inline double adj_dw(int d, double a, double b, double c, double dd)
{
    return d == 1 ? a : d == -1 ? -b : d == 2 ? c : d == -2 ? -dd : (-a + b - c + dd);
}

// (D^T v)(q) along x / y / z of the first derivative, v(i, j, k) the value
// at an interior output point (callers pass values supported on the interior)
template <class V>
__device__ inline double dTx(const Geo &g, int i, int j, int k, V &&v)
{
    if (!ax_in(k, g.NZ)) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int d = -2; d <= 2; d++) {
        int p = i - d;
        if (p >= 2 && p <= g.NX - 3) s += adj_dw(d, g.x1a[p], g.x1b[p], g.x1c[p], g.x1d[p]) * v(p, j, k);
    }
    return s;
}
template <class V>
__device__ inline double dTy(const Geo &g, int i, int j, int k, V &&v)
{
    if (!ax_in(i, g.NX) || !ax_in(k, g.NZ)) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int d = -2; d <= 2; d++) {
        int p = rmap(g, j - d);
        if (g.periodic || ax_in(p, g.NY)) s += adj_dw(d, g.y1a[p], g.y1b[p], g.y1c[p], g.y1d[p]) * v(i, p, k);
    }
    return s;
}
template <class V>
__device__ inline double dTz(const Geo &g, int i, int j, int k, V &&v)
{
    if (!ax_in(i, g.NX)) return 0.0;
    double s = 0.0;
#pragma unroll
    for (int d = -2; d <= 2; d++) {
        int p = k - d;
        if (p >= 2 && p <= g.NZ - 3) s += adj_dw(d, g.z1a[p], g.z1b[p], g.z1c[p], g.z1d[p]) * v(i, j, p);
    }
    return s;
}

// (L^T v)(q) of the 5-point Laplacian D2X + D2Y + D2Z (second-derivative
// coefficients), v supported on the interior
template <class V>
__device__ inline double lapT(const Geo &g, int i, int j, int k, V &&v)
{
    double s = 0.0;
    const bool xin = ax_in(i, g.NX), yin = g.periodic || ax_in(j, g.NY), zin = ax_in(k, g.NZ);
#pragma unroll
    for (int d = -2; d <= 2; d++) {
        const int px = i - d, py = rmap(g, j - d), pz = k - d;
        if (yin && zin && px >= 2 && px <= g.NX - 3)
            s += adj_dw(d, g.x2a[px], g.x2b[px], g.x2c[px], g.x2d[px]) * v(px, j, k);
        if (xin && zin && (g.periodic || ax_in(py, g.NY)))
            s += adj_dw(d, g.y2a[py], g.y2b[py], g.y2c[py], g.y2d[py]) * v(i, py, k);
        if (xin && yin && pz >= 2 && pz <= g.NZ - 3)
            s += adj_dw(d, g.z2a[pz], g.z2b[pz], g.z2c[pz], g.z2d[pz]) * v(i, j, pz);
    }
    return s;
}

// forward first derivatives of a field at an interior point
__device__ inline double fdx(const Geo &g, const double *f, int i, int j, int k)
{
    long q = IDX(g, i, j, k);
    return fd5(f[q - 2], f[q - 1], f[q], f[q + 1], f[q + 2], g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
}
__device__ inline double fdy(const Geo &g, const double *f, int i, int j, int k)
{
    long b = i + g.sz * k;
    return fd5(f[b + g.sy * rmap(g, j - 2)], f[b + g.sy * rmap(g, j - 1)], f[b + g.sy * j],
               f[b + g.sy * rmap(g, j + 1)], f[b + g.sy * rmap(g, j + 2)], g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
}
__device__ inline double fdz(const Geo &g, const double *f, int i, int j, int k)
{
    long q = IDX(g, i, j, k), s = g.sz;
    return fd5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s], g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
}

// ---------------------------------------------------------------------------
// boundary rules, transposed.  Forward order x, then y, then z: the
// transpose is z, then (non-periodic) y, then x; periodic y images are
// implicit.
// ---------------------------------------------------------------------------
// z rules (f0 = 0, f1 = f2/2; f(NZ-1) = 0, f(NZ-2) = f(NZ-3)/2) on columns
// i >= i0 (0 for bndry/bndryc/h/E, 2 for bndrytt: t3dsdsc.f:2130-2168)
__global__ void k_adj_bc_z(Geo g, double *__restrict__ a, int nf, int i0)
{
    const int nr = g.yhi - g.ylo + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * nr * nf;
    if (t >= n) return;
    int i = (int)(t % g.NX);
    long r = t / g.NX;
    int j = g.ylo + (int)(r % nr);
    int m = (int)(r / nr);
    if (i < i0) return;
    double *f = a + m * g.S + i + g.sy * j;
    const long s = g.sz;
    f[(g.NZ - 3) * s] += 0.5 * f[(g.NZ - 2) * s];
    f[(g.NZ - 2) * s] = 0.0;
    f[(g.NZ - 1) * s] = 0.0;
    f[2 * s] += 0.5 * f[s];
    f[s] = 0.0;
    f[0] = 0.0;
}

// non-periodic y rules (f0 = 0, f1 = f2/2; f(NY-1) = 0, f(NY-2) = f(NY-3)/2)
// on columns i >= i0 (t3dsdsc.f:1800-1819, 2592-2612; bndrytt 2108-2128)
__global__ void k_adj_bc_y(Geo g, double *__restrict__ a, int nf, int i0)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)g.NX * g.NZ * nf;
    if (t >= n) return;
    int i = (int)(t % g.NX);
    long r = t / g.NX;
    int k = (int)(r % g.NZ);
    int m = (int)(r / g.NZ);
    if (i < i0) return;
    double *f = a + m * g.S + i + g.sz * k;
    const long s = g.sy;
    f[(g.NY - 3) * s] += 0.5 * f[(g.NY - 2) * s];
    f[(g.NY - 2) * s] = 0.0;
    f[(g.NY - 1) * s] = 0.0;
    f[2 * s] += 0.5 * f[s];
    f[s] = 0.0;
    f[0] = 0.0;
}

// x rules: right f(NX-1) = 0, f(NX-2) = f(NX-3)/2; left the one-sided
// extrapolation (extrap, bndryc/h/E and bndrytt) or 0 / half (bndry)
__global__ void k_adj_bc_x(Geo g, double *__restrict__ a, int nf, int extrap)
{
    const int nr = g.yhi - g.ylo + 1;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long n = (long)nr * g.NZ * nf;
    if (t >= n) return;
    int j = g.ylo + (int)(t % nr);
    long r = t / nr;
    int k = (int)(r % g.NZ);
    int m = (int)(r / g.NZ);
    double *f = a + m * g.S + g.sy * j + g.sz * k;
    const int L = g.NX;
    f[L - 3] += 0.5 * f[L - 2];
    f[L - 2] = 0.0;
    f[L - 1] = 0.0;
    if (extrap) {
        double a0 = f[0];
        f[1] += g.ax3 / g.dx3 * a0;
        f[2] += g.bx3 / g.dx3 * a0;
        f[3] += g.cx3 / g.dx3 * a0;
        f[0] = 0.0;
        double a1 = f[1];
        f[2] += g.ax4 / g.dx4 * a1;
        f[3] += g.bx4 / g.dx4 * a1;
        f[4] += g.cx4 / g.dx4 * a1;
        f[1] = 0.0;
    } else {
        f[2] += 0.5 * f[1];
        f[1] = 0.0;
        f[0] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// smoothing sweep, transposed (out of place): f_new = f + c * sum(difc) on
// the interior.  KIND_STATE: fields 4..6 use cbx, the others crad.
// ---------------------------------------------------------------------------
template <int KIND>
__device__ inline double adj_coef(const Geo &g, int m, int i, long q)
{
    if (KIND == KIND_EIGHTH) return 1.0 / 8.0;
    if (KIND == KIND_STATE && m >= 4 && m <= 6) return g.cbx[i];
    return g.crad[q];
}

template <int KIND>
__global__ __launch_bounds__(256) void k_adj_filter(Geo g, const double *__restrict__ A, double *__restrict__ B,
                                                    int nf)
{
    AdjPt p = adj_point(g);
    if (!p.ok) return;
    const int i = p.i, j = p.j, k = p.k;
    const long q = IDX(g, i, j, k);
    const bool qi = in3(g, i, j, k);
    const bool xin = ax_in(i, g.NX), zin = ax_in(k, g.NZ);
    const int jm = rmap(g, j - 1), jp = rmap(g, j + 1);
    const long qjm = IDX(g, i, jm, k), qjp = IDX(g, i, jp, k);
    for (int m = 0; m < nf; m++) {
        const double *f = A + m * g.S;
        double s = f[q];
        if (qi)
            s -= adj_coef<KIND>(g, m, i, q) *
                 (g.xwp[i] + g.xwm[i] + g.ywp[j] + g.ywm[j] + g.zwp[k] + g.zwm[k]) * f[q];
        if (zin && (g.periodic || ax_in(j, g.NY))) {
            if (i - 1 >= 2 && i - 1 <= g.NX - 3) s += adj_coef<KIND>(g, m, i - 1, q - 1) * g.xwp[i - 1] * f[q - 1];
            if (i + 1 >= 2 && i + 1 <= g.NX - 3) s += adj_coef<KIND>(g, m, i + 1, q + 1) * g.xwm[i + 1] * f[q + 1];
        }
        if (xin && zin) {
            if (in3(g, i, jm, k)) s += adj_coef<KIND>(g, m, i, qjm) * g.ywp[jm] * f[qjm];
            if (in3(g, i, jp, k)) s += adj_coef<KIND>(g, m, i, qjp) * g.ywm[jp] * f[qjp];
        }
        if (xin && (g.periodic || ax_in(j, g.NY))) {
            if (k - 1 >= 2 && k - 1 <= g.NZ - 3) s += adj_coef<KIND>(g, m, i, q - g.sz) * g.zwp[k - 1] * f[q - g.sz];
            if (k + 1 >= 2 && k + 1 <= g.NZ - 3) s += adj_coef<KIND>(g, m, i, q + g.sz) * g.zwm[k + 1] * f[q + g.sz];
        }
        B[q + m * g.S] = s;
    }
}

// block sum of v added to the block's own slot of acc (launches on one
// stream are ordered, so the slots accumulate deterministically; the host
// sums them in a fixed order)
__device__ inline void block_add(double v, double *acc)
{
    __shared__ double red[256];
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    red[t] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) red[t] += red[t + s];
        __syncthreads();
    }
    if (t == 0) acc[blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z)] += red[0];
}

// ---------------------------------------------------------------------------
// -curl E and the pressure equation (t3dsdsc.f:1435-1566), transposed.
// In: ab (adjoint of xdif, 8 fields, zero off the interior), X, J.
// Out: eb = adjoint of E (set), vb = adjoint of v = m/rho (set), jb (set),
// xb(7) += adjoint of p, *etab += adjoint of eta.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_curle_p(Geo g, const double *__restrict__ ab,
                                                     const double *__restrict__ X, const double *__restrict__ J,
                                                     double gam, double eta, double *__restrict__ eb,
                                                     double *__restrict__ vb, double *__restrict__ jb,
                                                     double *__restrict__ xb, double *__restrict__ etab)
{
    AdjPt pt = adj_point(g);
    double es = 0.0;
    if (pt.ok) {
        const int i = pt.i, j = pt.j, k = pt.k;
        const long S = g.S, q = IDX(g, i, j, k);
        const double *a5 = ab + 4 * S, *a6 = ab + 5 * S, *a7 = ab + 6 * S, *a8 = ab + 7 * S;
        auto F = [&](const double *f) { return [=, &g](int x, int y, int z) { return f[IDX(g, x, y, z)]; }; };
        double e1 = -dTz(g, i, j, k, F(a6)) + dTy(g, i, j, k, F(a7));
        double e2 = dTz(g, i, j, k, F(a5)) - dTx(g, i, j, k, F(a7));
        double e3 = -dTy(g, i, j, k, F(a5)) + dTx(g, i, j, k, F(a6));
        // a8 * v and a8 * gam * p at the output points
        auto av = [&](int c) {
            return [=, &g](int x, int y, int z) {
                long r = IDX(g, x, y, z);
                return a8[r] * (X[r + c * S] / X[r]);
            };
        };
        auto agp = [=, &g](int x, int y, int z) {
            long r = IDX(g, x, y, z);
            return a8[r] * gam * X[r + 7 * S];
        };
        double p8 = -dTx(g, i, j, k, av(1)) - dTy(g, i, j, k, av(2)) - dTz(g, i, j, k, av(3));
        double v1 = -dTx(g, i, j, k, agp), v2 = -dTy(g, i, j, k, agp), v3 = -dTz(g, i, j, k, agp);
        double j1 = 0.0, j2 = 0.0, j3 = 0.0;
        if (in3(g, i, j, k)) {
            const double a = a8[q];
            const double *P = X + 7 * S;
            v1 -= a * fdx(g, P, i, j, k);
            v2 -= a * fdy(g, P, i, j, k);
            v3 -= a * fdz(g, P, i, j, k);
            // div v at q from v = m / rho at the neighbours
            auto vx = [&](long r) { return X[r + S] / X[r]; };
            auto vy = [&](long r) { return X[r + 2 * S] / X[r]; };
            auto vz = [&](long r) { return X[r + 3 * S] / X[r]; };
            const long sz = g.sz, b = i + sz * k;
            double dvx = fd5(vx(q - 2), vx(q - 1), vx(q), vx(q + 1), vx(q + 2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
            double dvy = fd5(vy(b + g.sy * rmap(g, j - 2)), vy(b + g.sy * rmap(g, j - 1)), vy(q),
                             vy(b + g.sy * rmap(g, j + 1)), vy(b + g.sy * rmap(g, j + 2)), g.y1a[j], g.y1b[j],
                             g.y1c[j], g.y1d[j]);
            double dvz = fd5(vz(q - 2 * sz), vz(q - sz), vz(q), vz(q + sz), vz(q + 2 * sz), g.z1a[k], g.z1b[k],
                             g.z1c[k], g.z1d[k]);
            p8 += a * gam * (-dvx - dvy - dvz);
            const double c1 = J[q], c2 = J[q + S], c3 = J[q + 2 * S];
            j1 = a * (gam - 1.) * eta * 2. * c1;
            j2 = a * (gam - 1.) * eta * 2. * c2;
            j3 = a * (gam - 1.) * eta * 2. * c3;
            es = a * (gam - 1.) * (c1 * c1 + c2 * c2 + c3 * c3);
        }
        eb[q] = e1; eb[q + S] = e2; eb[q + 2 * S] = e3;
        vb[q] = v1; vb[q + S] = v2; vb[q + 2 * S] = v3;
        jb[q] = j1; jb[q + S] = j2; jb[q + 2 * S] = j3;
        xb[q + 7 * S] += p8;
    }
    block_add(es, etab);
}

// ---------------------------------------------------------------------------
// viscosity (t3dsdsc.f:1585-1621): xdif(2..4) += fmu L(v), xdif8 +=
// fmu v.L(v); diffusion (1623-1663): xdif1 += frho L(rho), xdif8 += fpr
// L(p); transposed.  vb += adjoint of v, xb(0), xb(7) += .
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_visc(Geo g, const double *__restrict__ ab,
                                                  const double *__restrict__ X, double fmu, double frho,
                                                  double fpr, double *__restrict__ vb, double *__restrict__ xb)
{
    AdjPt pt = adj_point(g);
    if (!pt.ok) return;
    const int i = pt.i, j = pt.j, k = pt.k;
    const long S = g.S, q = IDX(g, i, j, k);
    const double *a8 = ab + 7 * S;
    auto vel = [&](long r, int m) { return X[r + (m + 1) * S] / X[r]; };
    if (fmu != 0.0) {
        const bool qi = in3(g, i, j, k);
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const double *am = ab + (m + 1) * S;
            double t = lapT(g, i, j, k, [&](int x, int y, int z) {
                const long r = IDX(g, x, y, z);
                return fmu * (am[r] + a8[r] * vel(r, m));
            });
            if (qi) {
                const long b = i + g.sz * k, sz = g.sz;
                const double lx = fd5(vel(q - 2, m), vel(q - 1, m), vel(q, m), vel(q + 1, m), vel(q + 2, m),
                                      g.x2a[i], g.x2b[i], g.x2c[i], g.x2d[i]);
                const double ly = fd5(vel(b + g.sy * rmap(g, j - 2), m), vel(b + g.sy * rmap(g, j - 1), m), vel(q, m),
                                      vel(b + g.sy * rmap(g, j + 1), m), vel(b + g.sy * rmap(g, j + 2), m),
                                      g.y2a[j], g.y2b[j], g.y2c[j], g.y2d[j]);
                const double lz = fd5(vel(q - 2 * sz, m), vel(q - sz, m), vel(q, m), vel(q + sz, m), vel(q + 2 * sz, m),
                                      g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]);
                t += fmu * a8[q] * (lx + ly + lz);
            }
            vb[q + m * S] += t;
        }
    }
    if (frho != 0.0 || fpr != 0.0) {
        xb[q] += lapT(g, i, j, k, [&](int x, int y, int z) { return frho * ab[IDX(g, x, y, z)]; });
        xb[q + 7 * S] += lapT(g, i, j, k, [&](int x, int y, int z) { return fpr * a8[IDX(g, x, y, z)]; });
    }
}

// ---------------------------------------------------------------------------
// Hall term xdh = cht (J x B - grad p) / rho on the interior
// (t3dsdsc.f:1391-1422), transposed.  In: hb = adjoint of xdh (after the
// transposed avrgh and bndryh).  Out: jb, xb += .
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_hall(Geo g, const double *__restrict__ hb, const double *__restrict__ X,
                                                  const double *__restrict__ J, double cht, double *__restrict__ jb,
                                                  double *__restrict__ xb)
{
    AdjPt pt = adj_point(g);
    if (!pt.ok) return;
    const int i = pt.i, j = pt.j, k = pt.k;
    const long S = g.S, q = IDX(g, i, j, k);
    auto sv = [&](int c) {
        return [=, &g](int x, int y, int z) {
            long r = IDX(g, x, y, z);
            return cht * hb[r + c * S] / X[r];
        };
    };
    double p8 = -dTx(g, i, j, k, sv(0)) - dTy(g, i, j, k, sv(1)) - dTz(g, i, j, k, sv(2));
    if (in3(g, i, j, k)) {
        const double x1 = X[q], x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
        const double c1 = J[q], c2 = J[q + S], c3 = J[q + 2 * S];
        const double b1 = hb[q], b2 = hb[q + S], b3 = hb[q + 2 * S];
        const double *P = X + 7 * S;
        const double h1 = cht * (c2 * x7 - c3 * x6 - fdx(g, P, i, j, k)) / x1;
        const double h2 = cht * (c3 * x5 - c1 * x7 - fdy(g, P, i, j, k)) / x1;
        const double h3 = cht * (c1 * x6 - c2 * x5 - fdz(g, P, i, j, k)) / x1;
        const double s1 = cht * b1 / x1, s2 = cht * b2 / x1, s3 = cht * b3 / x1;
        jb[q] += -s2 * x7 + s3 * x6;
        jb[q + S] += s1 * x7 - s3 * x5;
        jb[q + 2 * S] += -s1 * x6 + s2 * x5;
        xb[q + 4 * S] += s2 * c3 - s3 * c2;
        xb[q + 5 * S] += -s1 * c3 + s3 * c1;
        xb[q + 6 * S] += s1 * c2 - s2 * c1;
        xb[q] -= (b1 * h1 + b2 * h2 + b3 * h3) / x1;
    }
    xb[q + 7 * S] += p8;
}

// ---------------------------------------------------------------------------
// Extended-range block and fluxes (t3dsdsc.f:1334-1387, flux 2304-2347),
// transposed: E = (-m x B)/rho + eta J; xdif(2..4) += -xresd + J x B;
// v = m / rho; xdif(1..4) = -div flux(m).  In: ab (a1..a4), eb (after the
// transposed avrgE), vb.  Out: xb += , jb += , rb -= (when rb), *etab += .
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_adj_ext_flux(Geo g, const double *__restrict__ ab,
                                                      const double *__restrict__ eb, const double *__restrict__ vb,
                                                      const double *__restrict__ X, const double *__restrict__ J,
                                                      double eta,
                                                      double *__restrict__ xb, double *__restrict__ jb,
                                                      double *__restrict__ rb, double *__restrict__ etab)
{
    AdjPt pt = adj_point(g);
    double es = 0.0;
    if (pt.ok) {
        const int i = pt.i, j = pt.j, k = pt.k;
        const long S = g.S, q = IDX(g, i, j, k);
        const double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
        const double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
        const double c1 = J[q], c2 = J[q + S], c3 = J[q + 2 * S];
        const double e1 = eb[q], e2 = eb[q + S], e3 = eb[q + 2 * S];
        const double u1 = -x3 * x7 + x4 * x6, u2 = -x4 * x5 + x2 * x7, u3 = -x2 * x6 + x3 * x5;
        const double p2 = ab[q + S], p3 = ab[q + 2 * S], p4 = ab[q + 3 * S];
        const double v1 = vb[q], v2 = vb[q + S], v3 = vb[q + 2 * S];
        double g1 = -(e1 * u1 + e2 * u2 + e3 * u3) / (x1 * x1) - (v1 * x2 + v2 * x3 + v3 * x4) / (x1 * x1);
        double g2 = (e2 * x7 - e3 * x6) / x1 + v1 / x1;
        double g3 = (-e1 * x7 + e3 * x5) / x1 + v2 / x1;
        double g4 = (e1 * x6 - e2 * x5) / x1 + v3 / x1;
        double g5 = (-e2 * x4 + e3 * x3) / x1 + p3 * c3 - p4 * c2;
        double g6 = (e1 * x4 - e3 * x2) / x1 - p2 * c3 + p4 * c1;
        double g7 = (-e1 * x3 + e2 * x2) / x1 + p2 * c2 - p3 * c1;
        double g8 = 0.0;
        jb[q] += eta * e1 - p3 * x7 + p4 * x6;
        jb[q + S] += eta * e2 + p2 * x7 - p4 * x5;
        jb[q + 2 * S] += eta * e3 - p2 * x6 + p3 * x5;
        es = c1 * e1 + c2 * e2 + c3 * e3;
        if (rb) {
            rb[q] -= p2;
            rb[q + S] -= p3;
            rb[q + 2 * S] -= p4;
        }
        // fluxes: adjoints fb, gb, hb of (f, g, h)(m) = -D^T a(m)
        auto F = [&](const double *f) { return [=, &g](int x, int y, int z) { return f[IDX(g, x, y, z)]; }; };
        const double *a1 = ab, *a2 = ab + S, *a3 = ab + 2 * S, *a4 = ab + 3 * S;
        double fb = -dTx(g, i, j, k, F(a1)), gb = -dTy(g, i, j, k, F(a1)), hb = -dTz(g, i, j, k, F(a1));
        g2 += fb; g3 += gb; g4 += hb;
        const double x11 = x1 * x1;
        fb = -dTx(g, i, j, k, F(a2)); gb = -dTy(g, i, j, k, F(a2)); hb = -dTz(g, i, j, k, F(a2));
        g2 += (fb * 2. * x2 + gb * x3 + hb * x4) / x1;
        g3 += gb * x2 / x1;
        g4 += hb * x2 / x1;
        g8 += fb;
        g1 -= (fb * x2 * x2 + gb * x2 * x3 + hb * x2 * x4) / x11;
        fb = -dTx(g, i, j, k, F(a3)); gb = -dTy(g, i, j, k, F(a3)); hb = -dTz(g, i, j, k, F(a3));
        g2 += fb * x3 / x1;
        g3 += (fb * x2 + gb * 2. * x3 + hb * x4) / x1;
        g4 += hb * x3 / x1;
        g8 += gb;
        g1 -= (fb * x2 * x3 + gb * x3 * x3 + hb * x4 * x3) / x11;
        fb = -dTx(g, i, j, k, F(a4)); gb = -dTy(g, i, j, k, F(a4)); hb = -dTz(g, i, j, k, F(a4));
        g2 += fb * x4 / x1;
        g3 += gb * x4 / x1;
        g4 += (fb * x2 + gb * x3 + hb * 2. * x4) / x1;
        g8 += hb;
        g1 -= (fb * x2 * x4 + gb * x3 * x4 + hb * x4 * x4) / x11;
        xb[q] += g1; xb[q + S] += g2; xb[q + 2 * S] += g3; xb[q + 3 * S] += g4;
        xb[q + 4 * S] += g5; xb[q + 5 * S] += g6; xb[q + 6 * S] += g7; xb[q + 7 * S] += g8;
    }
    block_add(es, etab);
}

// curl(B - B_i) (t3dsdsc.f:2465-2491), transposed: xb(4..6) += curl^T jb
__global__ __launch_bounds__(256) void k_adj_curl(Geo g, const double *__restrict__ jb, double *__restrict__ xb)
{
    AdjPt pt = adj_point(g);
    if (!pt.ok) return;
    const int i = pt.i, j = pt.j, k = pt.k;
    const long S = g.S, q = IDX(g, i, j, k);
    auto F = [&](const double *f) { return [=, &g](int x, int y, int z) { return f[IDX(g, x, y, z)]; }; };
    const double *j1 = jb, *j2 = jb + S, *j3 = jb + 2 * S;
    xb[q + 4 * S] += dTz(g, i, j, k, F(j2)) - dTy(g, i, j, k, F(j3));
    xb[q + 5 * S] += dTx(g, i, j, k, F(j3)) - dTz(g, i, j, k, F(j1));
    xb[q + 6 * S] += dTy(g, i, j, k, F(j1)) - dTx(g, i, j, k, F(j2));
}

// RK4 register scheme (t3dsdsc.f:1090-1132), transposed, one stage at a
// time from the last: xb = adjoint of the stage output, m = of xm,
// f = of xfold, d = adjoint of that stage's xdif (set), t = adjoint of the
// stage input returned by the previous transposed right (read)
__global__ void k_adj_rk(long n, int stage, double dt, const double *__restrict__ xb, double *__restrict__ xbm,
                         double *__restrict__ xbf, double *__restrict__ d, double *__restrict__ t)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    if (stage == 3) {
        xbm[q] = xb[q];
        d[q] = div6(xb[q] * dt);
    } else if (stage == 2) {
        xbf[q] = t[q];
        d[q] = t[q] * dt + div3(xbm[q] * dt);
    } else if (stage == 1) {
        xbf[q] += t[q];
        d[q] = t[q] * dt / 2. + div3(xbm[q] * dt);
    } else {
        xbf[q] += t[q] + xbm[q];
        d[q] = t[q] * dt / 2. + div6(xbm[q] * dt);
    }
    t[q] = 0.0;
}

// objective at the final state and its seed (adjoint of x_N)
//   obj 0: magnetic energy 1/2 sum |B|^2 wx wy wz (trapezoid weights)
//   obj 1: reconnected flux 1/2 sum_i wx |Bz~| averaged over y, with Bz~ the
//          mean of planes k0, k0+1 (the two planes nearest z = 0)
__global__ __launch_bounds__(256) void k_adj_objective(Geo g, const double *__restrict__ X, int obj, int k0,
                                                       const double *__restrict__ wx, const double *__restrict__ wy,
                                                       const double *__restrict__ wz, double *__restrict__ seed,
                                                       double *__restrict__ val)
{
    AdjPt pt = adj_point(g);
    double v = 0.0;
    if (pt.ok) {
        const int i = pt.i, j = pt.j, k = pt.k;
        const long S = g.S, q = IDX(g, i, j, k);
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (obj == 0) {
            const double w = wx[i] * wy[j] * wz[k];
            const double b1 = X[q + 4 * S], b2 = X[q + 5 * S], b3 = X[q + 6 * S];
            v = 0.5 * w * (b1 * b1 + b2 * b2 + b3 * b3);
            s[4] = w * b1; s[5] = w * b2; s[6] = w * b3;
        } else if (k == k0 || k == k0 + 1) {
            const long q0 = IDX(g, i, j, k0);
            const double bt = 0.5 * (X[q0 + 6 * S] + X[q0 + g.sz + 6 * S]);
            const double cw = 0.5 * wx[i] / (g.yhi - g.ylo + 1);
            if (k == k0) v = cw * fabs(bt);
            s[6] = cw * (bt > 0.0 ? 0.5 : bt < 0.0 ? -0.5 : 0.0);
        }
#pragma unroll
        for (int m = 0; m < 8; m++) seed[q + m * S] = s[m];
    }
    block_add(v, val);
}

} // namespace mhd
