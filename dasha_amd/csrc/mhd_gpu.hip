// mhd_gpu.hip -- MI355X (gfx950) RK4 Hall-MHD step behind the C ABI of
// include/mhd_gpu.h.
//
// The step is the reference's stepon (t3dsdsc.f:1063-1185): four calls of
// the right-hand side `right` (t3dsdsc.f:1297-1673), each a chain of
// dependent sweeps with boundary fills between them, RK4 register updates,
// then a smoothing pass on the perturbation x - eq.  The sweeps are the
// z-marching tile kernels of mhd_kernels.h; each leaves its output buffer
// with the faces filled as the reference's boundary routine after that sweep
// would fill them, so only the post-step fill (bndrytt) runs separately.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "mhd_gpu.h"
#include "mhd_kernels.h"
#include "mhd_adjoint.h"
#include "mhd_magneto.h"

using namespace mhd;

namespace {

// ---------------------------------------------------------------------------
// RK4 register updates on the extended range, t3dsdsc.f:1090-1132
// ---------------------------------------------------------------------------
__global__ void k_rk(long n8, int stage, const double *__restrict__ dtp, const double *__restrict__ XD,
                     const double *__restrict__ XF, double *__restrict__ XM, double *__restrict__ X)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n8) return;
    double dt = *dtp, d = XD[q];
    if (stage == 0) {
        st(XM + q, XF[q] + div6(d * dt));
        st(X + q, XF[q] + d * dt / 2.);
    } else if (stage == 1) {
        st(XM + q, XM[q] + div3(d * dt));
        st(X + q, XF[q] + d * dt / 2.);
    } else if (stage == 2) {
        st(XM + q, XM[q] + div3(d * dt));
        st(X + q, XF[q] + d * dt);
    } else {
        st(X + q, XM[q] + div6(d * dt));
    }
}

// x = xm + eq, t3dsdsc.f:1157-1164
__global__ void k_add_eq(long S, long n, const double *__restrict__ XM, Eq eq, double *__restrict__ X)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
#pragma unroll
    for (int m = 0; m < 8; m++) {
        double v = XM[q + m * S];
        st(X + q + m * S, eq.f[m] ? v + eq.f[m][q] : v);
    }
}

__global__ void k_set(double *p, double v) { *p = v; }

// radial smoothing coefficient of every interior point (geometry only), the
// same expression the sweeps evaluate in place
__global__ void k_crad(Geo g, double *__restrict__ c)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= g.S) return;
    int i = (int)(q % g.NX);
    long r = q / g.NX;
    int j = (int)(r % g.NY), k = (int)(r / g.NY);
    bool in = i >= 2 && i <= g.NX - 3 && j >= 2 && j <= g.NY - 3 && k >= 2 && k <= g.NZ - 3;
    c[q] = in ? radial_coef(g.xc[i] * g.xc[i] + g.yc[j] * g.yc[j], g.zc[k]) : 0.0;
}

// plain streaming copy (reference rate for the roofline figures)
__global__ void k_copy(long n, const double *__restrict__ a, double *__restrict__ b)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) st(b + t, a[t]);
}

// U(-1, 1) from the splitmix64 finaliser of (seed, index): the top 53 bits
// of the mixed word scaled to [0, 1), then 2u - 1 (exact in FP64)
__device__ inline double splitmix_u(unsigned long long seed, unsigned long long index)
{
    unsigned long long z = seed * 0x9E3779B97F4A7C15ull + (index + 1ull) * 0xD1B54A32D192ED03ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z = z ^ (z >> 31);
    return (double)(z >> 11) * 0x1.0p-53 * 2.0 - 1.0;
}

// Harris sheet (mhd_init_harris): one thread per point (i, j, k) of the
// device volume, all 8 state fields and the 5 equilibrium fields
__global__ void k_harris(Geo g, mhd_harris hp, int nyu, double *__restrict__ X, double *__restrict__ EQ)
{
    const long n = (long)g.NX * g.NY * g.NZ;
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int i = (int)(t % g.NX);
    const long r = t / g.NX;
    const int j = (int)(r % g.NY), k = (int)(r / g.NY);
    const long q = IDX(g, i, j, k), S = g.S;
    const double Z = g.zc[k], xr = g.xc[i] - hp.x_offset;
    const double ch = cosh(Z / hp.lam);
    const double sech2 = 1.0 / (ch * ch);
    const double pri = 0.5 * hp.beta_inf + 0.5 * sech2;
    const double rhoi = pri / hp.temp;
    const double bxi = tanh(Z / hp.lam);
    const double w2 = hp.width * hp.width;
    const double gp = hp.amp * exp(-(xr * xr + Z * Z) / w2);
    const double dbx = 2.0 * Z / w2 * gp, dbz = -2.0 * xr / w2 * gp;
    EQ[q] = rhoi; EQ[q + S] = bxi; EQ[q + 2 * S] = hp.guide; EQ[q + 3 * S] = 0.0; EQ[q + 4 * S] = pri;
    X[q] = rhoi;
    X[q + 4 * S] = bxi + dbx;
    X[q + 5 * S] = hp.guide;
    X[q + 6 * S] = 0.0 + dbz;
    X[q + 7 * S] = pri;
    const int ju = g.periodic ? ((j - 2) % nyu + nyu) % nyu : j;
    for (int c = 0; c < 3; c++) {
        const unsigned long long idx =
            (((unsigned long long)c * g.NZ + k) * (unsigned long long)nyu + ju) * g.NX + i;
        X[q + (1 + c) * S] = rhoi * hp.noise * splitmix_u(hp.seed, idx);
    }
}

// ---------------------------------------------------------------------------
// setdt, t3dsdsc.f:4709-4808: minimum over jz = 3..mz, jy = 1..my-1,
// jx = 3..mx (k = 0..NZ-3, j = 0..NY-2, i = 0..NX-3); with periodic y the
// image rows repeat unique rows, so the unique rows give the same minimum
// ---------------------------------------------------------------------------
// rr (magnetosphere mode, else null): only points with rr1 >= 1 count (t3dsdsc.f:4712);
// block b of nb writes part[b]
__device__ __forceinline__ void dt_partial_body(const Geo &g, const double *__restrict__ X, double eta, double fmu,
                                                double cfl, double gam, double wcoef, const double *__restrict__ rr,
                                                double *__restrict__ part, unsigned b, unsigned nb)
{
    __shared__ double sm[256];
    const int j0 = g.periodic ? 2 : 0, j1 = g.periodic ? g.NY - 3 : g.NY - 2;
    const int ni = g.NX - 2, nj = j1 - j0 + 1, nk = g.NZ - 2;
    const unsigned nrow = (unsigned)nj * (unsigned)nk;
    double best = 100.;
    // block b takes the rows (j, k) b, b + nb, ...; its threads run along x
    for (unsigned row = b; row < nrow; row += nb)
    for (int i = threadIdx.x; i < ni; i += blockDim.x) {
        const int j = j0 + (int)(row % (unsigned)nj), k = (int)(row / (unsigned)nj);
        long q = IDX(g, i, j, k), S = g.S;
        if (rr && !(rr[q] >= 1.)) continue;
        // quotients with a shared or precomputed rounded reciprocal (divc:
        // equal to the IEEE quotients); (eta + fmu) / dd / 2 is a per-index
        // table computed the same way on the host
        const double x1 = X[q], y1 = 1.0 / x1;
        double vx = divc(X[q + S], x1, y1), vy = divc(X[q + 2 * S], x1, y1), vz = divc(X[q + 3 * S], x1, y1);
        double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
        double b2 = x5 * x5 + x6 * x6 + x7 * x7;
        double va2 = divc(b2, x1, y1);
        double cs2 = divc(gam * X[q + 7 * S], x1, y1);
        double vf = sqrt(fabs(va2 + cs2));
        double vpx = fabs(vx) + vf, vpy = fabs(vy) + vf, vpz = fabs(vz) + vf;
        double ddx = g.dxs[i], ddy = g.dys[j], ddz = g.dzs[k];
        if (wcoef != 0.0) {
            double wb = divc(wcoef * sqrt(b2), x1, y1);
            vpx = vpx + divc(wb, fabs(ddx), g.rdx[i]);
            vpy = vpy + divc(wb, fabs(ddy), g.rdy[j]);
            vpz = vpz + divc(wb, fabs(ddz), g.rdz[k]);
        }
        const double rcfl = 1.0 / cfl;
        double dtx = fabs(ddx) / (divc(vpx, cfl, rcfl) + g.hdx[i]);
        double dty = fabs(ddy) / (divc(vpy, cfl, rcfl) + g.hdy[j]);
        double dtz = fabs(ddz) / (divc(vpz, cfl, rcfl) + g.hdz[k]);
        double dt1 = dtx < dty ? dtx : dty;
        double dt2 = dtz < dt1 ? dtz : dt1;
        if (dt2 < best) best = dt2;
    }
    sm[threadIdx.x] = best;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && sm[threadIdx.x + s] < sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[b] = sm[0];
}

__global__ void k_dt_partial(Geo g, const double *__restrict__ X, double eta, double fmu, double cfl, double gam,
                             double wcoef, const double *__restrict__ rr, double *__restrict__ part)
{
    dt_partial_body(g, X, eta, fmu, cfl, gam, wcoef, rr, part, blockIdx.x, gridDim.x);
}

// final MIN; keeps the running time and the collapse flag on the device:
// dtv[0] = dt, dtv[1] = time (advanced by the previous dt), dtv[2] = dt0,
// dtv[3] = collapse flag, dtv[4] = previous dt, dtv[5] = the collapsing dt.
// A collapse (dt < 0.1 dt0, t3dsdsc.f:142-154) freezes dt and time: the
// step after it runs with the last good dt and k_restore undoes it, so the
// state and time stay those before the collapsing step, as the reference
// stops there.
__device__ __forceinline__ void dt_final_body(const double *__restrict__ part, int n, double *dtv, int advance_time)
{
    __shared__ double sm[256];
    double best = 100.;
    for (int t = threadIdx.x; t < n; t += blockDim.x)
        if (part[t] < best) best = part[t];
    sm[threadIdx.x] = best;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s && sm[threadIdx.x + s] < sm[threadIdx.x]) sm[threadIdx.x] = sm[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (dtv[3] != 0.0) return;
        double dt = sm[0];
        if (advance_time) dtv[1] = dtv[1] + dtv[4];
        if (dtv[2] > 0.0 && dt < 0.1 * dtv[2]) {
            dtv[3] = 1.0;
            dtv[5] = dt;
            return;
        }
        dtv[0] = dt;
        dtv[4] = dt;
    }
}

__global__ void k_dt_final(const double *__restrict__ part, int n, double *dtv, int advance_time)
{
    dt_final_body(part, n, dtv, advance_time);
}

// after a step of a guarded loop: once the collapse flag is up, the step
// that followed the collapsing setdt is undone (x = xfold, the state before it)
__device__ __forceinline__ void restore_body(const double *__restrict__ dtv, const double *__restrict__ XF,
                                             double *__restrict__ X, long n, unsigned b, unsigned nb)
{
    if (dtv[3] == 0.0) return;
    for (long q = (long)b * blockDim.x + threadIdx.x; q < n; q += (long)nb * blockDim.x) X[q] = XF[q];
}

__global__ void k_restore(const double *__restrict__ dtv, const double *__restrict__ XF, double *__restrict__ X, long n)
{
    restore_body(dtv, XF, X, n, blockIdx.x, gridDim.x);
}

// batched forms of setdt and the collapse undo (one launch over every run of
// an ensemble; {run, block} entries, arguments per run)
struct BDtArgs {
    Geo g;
    const double *X, *rr;
    double eta, fmu, cfl, gam, wcoef;
    double *part, *dtv;
    int nb, advance_time;
};

struct BRestoreArgs {
    const double *dtv, *XF;
    double *X;
    long n;
    int nb;
};

__global__ void kb_dt_partial(const BDtArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BDtArgs &a = run_args(args, run);
    dt_partial_body(a.g, a.X, a.eta, a.fmu, a.cfl, a.gam, a.wcoef, a.rr, a.part, tab[blockIdx.x].w, a.nb);
}

__global__ void kb_dt_final(const BDtArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const BDtArgs &a = run_args(args, tab[blockIdx.x].run);
    dt_final_body(a.part, a.nb, a.dtv, a.advance_time);
}

__global__ void kb_restore(const BRestoreArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BRestoreArgs &a = run_args(args, run);
    restore_body(a.dtv, a.XF, a.X, a.n, tab[blockIdx.x].w, a.nb);
}

// disarm the collapse guard (a stand-alone setdt reports every dt)
__global__ void k_disarm(double *dtv)
{
    dtv[2] = 0.0;
    dtv[3] = 0.0;
}

// ---------------------------------------------------------------------------
// diagnostics
// ---------------------------------------------------------------------------
// current extrema on the plane j0 over the extended x-z range (the
// reference's range adds ghost columns holding 0; the host folds 0 in)
__global__ void k_plane_extrema(Geo g, const double *__restrict__ cur, int j0, double *__restrict__ out)
{
    __shared__ double smax[256], smin[256];
    int m = blockIdx.y;
    double mx = -1000000., mn = 1000000.;
    long n = (long)g.NX * g.NZ;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        int i = (int)(t % g.NX), k = (int)(t / g.NX);
        double v = cur[IDX(g, i, j0, k) + m * g.S];
        if (v > mx) mx = v;
        if (v < mn) mn = v;
    }
    smax[threadIdx.x] = mx; smin[threadIdx.x] = mn;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (smax[threadIdx.x + s] > smax[threadIdx.x]) smax[threadIdx.x] = smax[threadIdx.x + s];
            if (smin[threadIdx.x + s] < smin[threadIdx.x]) smin[threadIdx.x] = smin[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[(m * gridDim.x + blockIdx.x) * 2] = smax[0];
        out[(m * gridDim.x + blockIdx.x) * 2 + 1] = smin[0];
    }
}

// per-(j,k) row sums of the energy densities times the volume weights;
// rows are summed on the host in (k, j) order
__global__ void k_energy_rows(Geo g, const double *__restrict__ X, Eq eq, const double *__restrict__ wx,
                              const double *__restrict__ wy, const double *__restrict__ wz, int j0,
                              int nyr, double gm1, double *__restrict__ rows)
{
    int jr = blockIdx.x, k = blockIdx.y;
    int j = j0 + jr;
    __shared__ double sm[4][64];
    double a0 = 0, a1 = 0, a2 = 0, a3 = 0;
    long S = g.S;
    for (int i = threadIdx.x; i < g.NX; i += 64) {
        long q = IDX(g, i, j, k);
        double wv = wx[i] * wy[j] * wz[k];
        double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
        double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S], x8 = X[q + 7 * S];
        double d5 = x5 - eq.f[4][q], d6 = x6 - eq.f[5][q], d7 = x7 - eq.f[6][q];
        a0 += 0.5 * (x2 * x2 + x3 * x3 + x4 * x4) / x1 * wv;
        a1 += 0.5 * (x5 * x5 + x6 * x6 + x7 * x7) * wv;
        a2 += x8 / gm1 * wv;
        a3 += 0.5 * (d5 * d5 + d6 * d6 + d7 * d7) * wv;
    }
    sm[0][threadIdx.x] = a0; sm[1][threadIdx.x] = a1; sm[2][threadIdx.x] = a2; sm[3][threadIdx.x] = a3;
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0;
        for (int t = 0; t < 64; t++) s += sm[threadIdx.x][t];
        rows[((long)k * nyr + jr) * 4 + threadIdx.x] = s;
    }
}

// div B interior (t3dsdsc.f:2520-2528), y neighbours through the image map
__global__ void k_divb(Geo g, const double *__restrict__ X, double *__restrict__ dv)
{
    int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    int k = 2 + blockIdx.z;
    if (i > g.NX - 3 || j > g.NY - 3) return;
    long S = g.S, q = IDX(g, i, j, k), sz = g.sz;
    const double *b5 = X + 4 * S, *b6 = X + 5 * S, *b7 = X + 6 * S;
    long r0 = (long)i + g.sz * k;
    double dbx = fd5(b5[q - 2], b5[q - 1], b5[q], b5[q + 1], b5[q + 2], g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
    double dby = fd5(b6[r0 + g.sy * rmap(g, j - 2)], b6[r0 + g.sy * rmap(g, j - 1)], b6[q],
                     b6[r0 + g.sy * rmap(g, j + 1)], b6[r0 + g.sy * rmap(g, j + 2)], g.y1a[j], g.y1b[j], g.y1c[j],
                     g.y1d[j]);
    double dbz = fd5(b7[q - 2 * sz], b7[q - sz], b7[q], b7[q + sz], b7[q + 2 * sz], g.z1a[k], g.z1b[k], g.z1c[k],
                     g.z1d[k]);
    dv[q] = dbx + dby + dbz;
}

// bndrydiv (t3dsdsc.f:3031-3712) as sequential face passes (x, y, z);
// pass 0: x faces for all j,k; pass 1: y rows (non-periodic); pass 2: z
__global__ void k_divb_bc(Geo g, double *__restrict__ dv, int pass)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (pass == 0) {
        long n = (long)g.NY * g.NZ;
        if (t >= n) return;
        int j = (int)(t % g.NY), k = (int)(t / g.NY);
        long b = IDX(g, 0, j, k);
        double *f = dv + b;
        int L = g.NX - 1;
        f[1] = (g.ax4 * f[2] + g.bx4 * f[3] + g.cx4 * f[4]) / (g.ax4 + g.bx4 + g.cx4);
        f[0] = (g.ax3 * f[1] + g.bx3 * f[2] + g.cx3 * f[3]) / (g.ax3 + g.bx3 + g.cx3);
        f[L - 1] = (g.axmL1 * f[L - 2] + g.bxmL1 * f[L - 3] + g.cxmL1 * f[L - 4]) / (g.axmL1 + g.bxmL1 + g.cxmL1);
        f[L] = (g.axmL * f[L - 1] + g.bxmL * f[L - 2] + g.cxmL * f[L - 3]) / (g.axmL + g.bxmL + g.cxmL);
    } else if (pass == 1) {
        long n = (long)g.NX * g.NZ;
        if (t >= n || g.periodic) return;
        int i = (int)(t % g.NX), k = (int)(t / g.NX);
        long sy = g.sy, b = IDX(g, i, 0, k);
        int M = g.NY - 1;
        double *f = dv + b;
        double aa1 = 5. / 2., bb1 = -2., cc1 = 1. / 2.;
        f[sy] = aa1 * f[2 * sy] + bb1 * f[3 * sy] + cc1 * f[4 * sy];
        f[(M - 1) * sy] = aa1 * f[(M - 2) * sy] + bb1 * f[(M - 3) * sy] + cc1 * f[(M - 4) * sy];
        f[0] = aa1 * f[sy] + bb1 * f[2 * sy] + cc1 * f[3 * sy];
        f[M * sy] = aa1 * f[(M - 1) * sy] + bb1 * f[(M - 2) * sy] + cc1 * f[(M - 3) * sy];
    } else {
        long n = (long)g.NX * g.NY;
        if (t >= n) return;
        int i = (int)(t % g.NX), j = (int)(t / g.NX);
        long sz = g.sz, b = IDX(g, i, j, 0);
        int M = g.NZ - 1;
        double *f = dv + b;
        double aa1 = 5. / 2., bb1 = -2., cc1 = 1. / 2.;
        f[sz] = aa1 * f[2 * sz] + bb1 * f[3 * sz] + cc1 * f[4 * sz];
        f[0] = aa1 * f[sz] + bb1 * f[2 * sz] + cc1 * f[3 * sz];
        f[(M - 1) * sz] = aa1 * f[(M - 2) * sz] + bb1 * f[(M - 3) * sz] + cc1 * f[(M - 4) * sz];
        f[M * sz] = aa1 * f[(M - 1) * sz] + bb1 * f[(M - 2) * sz] + cc1 * f[(M - 3) * sz];
    }
}

__global__ void k_absmax_sq(long n, const double *__restrict__ a, double *__restrict__ out)
{
    __shared__ double smx[256], ssq[256];
    double mx = 0, sq = 0;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        double v = a[t];
        if (fabs(v) > mx) mx = fabs(v);
        sq += v * v;
    }
    smx[threadIdx.x] = mx; ssq[threadIdx.x] = sq;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            if (smx[threadIdx.x + s] > smx[threadIdx.x]) smx[threadIdx.x] = smx[threadIdx.x + s];
            ssq[threadIdx.x] += ssq[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) { out[2 * blockIdx.x] = smx[0]; out[2 * blockIdx.x + 1] = ssq[0]; }
}

} // namespace

// ===========================================================================
// host side
// ===========================================================================
struct Prof {
    const char *name;
    double ms;
    int calls;
};

// block table of a temporal-blocked launch (device words in dispatch
// order, one per block; the same words in enumeration order on the host)
struct ZTab {
    unsigned *d = nullptr;
    unsigned n = 0;
    std::vector<unsigned> nat;
};

// One launch of a recorded step sequence, for the batched kernels of an
// ensemble: the launcher of the batched kernel, the run's argument block
// and the run's block words (enumeration order).
using BLaunch = void (*)(const void *args, const BEntry *tab, unsigned nblk, hipStream_t s);
struct RecLaunch {
    BLaunch fn;
    std::string arg;
    std::vector<unsigned> words;
};

struct mhd_handle {
    mhd_config cfg;
    int NX, NY, NZ;
    long S, N;
    hipStream_t stream = nullptr;
    // second stream for launches independent of the main stream's current
    // one (the face-window launch of a smoothing sweep beside its plain
    // launch), forked from and joined back into the main stream by events
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    // and a second such pair for the first Hall smoothing passes, which run
    // beside the smoothing of E
    hipStream_t aux = nullptr, aux_side = nullptr;
    hipEvent_t aux_ev = nullptr;
    int overlap = 1;
    Geo g;
    double *geo = nullptr;                 // all 1-D device arrays
    double *x = nullptr, *xfold = nullptr, *xm = nullptr, *xdA = nullptr, *xdB = nullptr;
    double *eq = nullptr;                  // rhoi, bxi, byi, bzi, pri
    double *xresd = nullptr;
    double *cA = nullptr, *cB = nullptr, *eA = nullptr, *eB = nullptr, *hA = nullptr, *hB = nullptr;
    double *xd = nullptr;                  // buffer holding the last xdif
    double *cur = nullptr;                 // buffer holding the last current
    double *red = nullptr;                 // reduction scratch
    double *crad = nullptr;                // radial smoothing coefficient table
    double *stage = nullptr;               // pinned staging buffer of host transfers
    double *dtv = nullptr;                 // device dt slots
    int nred = 0;
    std::vector<double> wx, wy, wz;        // energy weights (host copy)
    double *dwx = nullptr, *dwy = nullptr, *dwz = nullptr;
    std::string err;
    int dev = -1;                          // HIP device the handle was created on
    // magnetosphere mode: |r| per point, the Hall coefficient per point (radial
    // profile), v = m / rho, the inflow profiles and face data
    bool mag = false;
    double *rr = nullptr, *chtab = nullptr, *vel = nullptr, *mgeo = nullptr;
    int *mrow = nullptr;                   // masked column range per row, when the producers mask
    int fill_split = 0;                    // y and z rules of the fills as two launches
    const double *zd = nullptr, *yd = nullptr, *zd2 = nullptr, *yd2 = nullptr, *face = nullptr;
    bool have_eq = false;
    int kc = 16;                           // z planes per tile chunk
    int fuse_rk = 1;                       // RK update fused into the last avrg pass
    int zs = 1;                            // pairs of smoothing passes in one temporal-blocked kernel
    int rb2 = 1;                           // windowed curl/pressure kernels for right, part 2
    int zkc = 32;                          // interior planes per block of the temporal-blocked kernels
    int zord = 1;                          // block order of the temporal-blocked kernels (zs_tables)
    ZTab zt[2][2];                         // their block tables: [3 or 8 fields][without / with face rules]
    ZTab zt3[2];                           // the same for three passes on 3 fields
    int zs3 = 1;                           // last three avrgh passes (+ E) in one launch
    int face12 = 1;                        // 12-row blocks for the face windows (periodic y)
    int adj_overlap = 1;                   // adjoint: the Hall chain beside the E chain
    // stage recording of the adjoint's forward replay: when rec_x is set,
    // right() copies its state and current into slot rec_n and advances it
    double *rec_x = nullptr, *rec_c = nullptr;
    int rec_n = 0;
    // when set, the launches that have a batched form are appended here
    // instead of issued (ensemble batching); rec_err names a launch shape
    // the batched forms cannot take
    std::vector<RecLaunch> *rec = nullptr;
    std::string rec_err;
    // profiling
    bool prof = false;
    std::vector<Prof> pt;
    std::vector<hipEvent_t> pool;
    size_t nev = 0;
    std::vector<int> ev_tag;
};

#define HCHK(call)                                                                 \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess) {                                                    \
            h->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return MHD_ERR_HIP;                                                    \
        }                                                                          \
    } while (0)

// launch-time errors of void helpers surface at the next synchronising call
#define HCHK_V(call) (void)(call)

// ---- allocation-free setdt (mhd_compute_dt_noalloc) ----
// The block partial minima of setdt go to this fixed device array instead of
// the handle's reduction buffer.  Its host-side name is not a device address:
// kernels get the address hipGetSymbolAddress returns for it, looked up once
// per device when the first handle is created there (mhd_create).
static constexpr int MHD_SCRATCH_DOUBLES = 4096;
static constexpr int MHD_MAX_DEVICES = 64;
static __device__ __align__(256) double scratch[MHD_SCRATCH_DOUBLES];
static double *g_scratch_ptr[MHD_MAX_DEVICES];
static std::mutex g_scratch_mu;

// with device `dev` current: record the device address of `scratch`
static int scratch_setup(mhd_handle *h, int dev)
{
    if (dev < 0 || dev >= MHD_MAX_DEVICES) {
        h->err = "device index beyond the scratch table";
        return MHD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev]) return MHD_OK;
    void *p = nullptr;
    HCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) {
        h->err = "hipGetSymbolAddress returned no address for the scratch array";
        return MHD_ERR_HIP;
    }
    g_scratch_ptr[dev] = static_cast<double *>(p);
    return MHD_OK;
}

namespace {

// ---- host restatement of the FD coefficient algebra, t3dsdsc.f:336-606 ----
void coef_d2(const double *c, int j, double *A, double *B, double *Cc, double *D)
{
    double p1 = c[j + 1] - c[j], m1 = c[j] - c[j - 1], p2 = c[j + 2] - c[j], m2 = c[j] - c[j - 2];
    double p1s = p1 * p1, m2s = m2 * m2, m1s = m1 * m1, p2s = p2 * p2;
    double f1 = p1s + p1 * m2, f2 = p1s * p1 - p1 * m2s, f3 = p1s * p1s + p1 * (m2s * m2);
    double g1 = -m1s + m1 * m2, g2 = m1s * m1 - m1 * m2s, g3 = -(m1s * m1s) + m1 * (m2s * m2);
    double h1 = p2s + p2 * m2, h2 = p2s * p2 - p2 * m2s, h3 = p2s * p2s + p2 * (m2s * m2);
    double ca = (f1 * h3 - f3 * h1) * (g2 * h3 - g3 * h2) - (g1 * h3 - g3 * h1) * (f2 * h3 - f3 * h2);
    *A = 2. * h3 * (g2 * h3 - g3 * h2) / ca;
    *B = 2. * h3 * (h2 * f3 - h3 * f2) / ca;
    *Cc = 2. * h3 * (f2 * g3 - f3 * g2) / ca;
    *D = -(p1 * *A + m1 * *B + p2 * *Cc) / m2;
}

void coef_d1(const double *c, int j, double *A, double *B, double *Cc, double *D)
{
    double p1 = c[j + 1] - c[j], m1 = c[j] - c[j - 1], p2 = c[j + 2] - c[j], m2 = c[j] - c[j - 2];
    double p1s = p1 * p1, m1s = m1 * m1, p2s = p2 * p2, m2s = m2 * m2;
    double p14 = p1s * p1s, m14 = m1s * m1s, p24 = p2s * p2s, m23 = m2s * m2;
    double f1 = p1 + p14 / m23, f2 = p1s - p14 / m2s, f3 = p1s * p1 + p14 / m2;
    double g1 = m1 - m14 / m23, g2 = -m1s + m14 / m2s, g3 = m1s * m1 - m14 / m2;
    double h1 = p2 + p24 / m23, h2 = p2s - p24 / m2s, h3 = p2s * p2 + p24 / m2;
    double ca = (f1 * h3 - f3 * h1) * (g2 * h3 - g3 * h2) - (g1 * h3 - g3 * h1) * (f2 * h3 - f3 * h2);
    *A = h3 * (g2 * h3 - g3 * h2) / ca;
    *B = h3 * (h2 * f3 - h3 * f2) / ca;
    *Cc = h3 * (f2 * g3 - f3 * g2) / ca;
    *D = (p1s * *A - m1s * *B + p2s * *Cc) / m2s;
}

void coef_side(const double *c, int j, int dir, double *A, double *B, double *Cc)
{
    double d1 = dir * (c[j + dir] - c[j]), d2 = dir * (c[j + 2 * dir] - c[j]), d3 = dir * (c[j + 3 * dir] - c[j]);
    double d1s = d1 * d1, d2s = d2 * d2, d3s = d3 * d3;
    double f1 = d1 - d1s * d1 / d3s, f2 = d1s - d1s * d1 / d3;
    double g1 = d2 - d2s * d2 / d3s, g2 = d2s - d2s * d2 / d3;
    double ca = f1 * g2 - f2 * g1;
    *A = g2 / ca;
    *B = -f2 / ca;
    *Cc = (1 - *A * d1 - *B * d2) / d3;
}

int blocks(long n, int b) { return (int)((n + b - 1) / b); }

} // namespace

static void set_err(mhd_handle *h, const std::string &s) { if (h) h->err = s; }

// profiling helpers: events bracket launches of one family; events come
// from a pool that grows once and is reused
static void prof_mark(mhd_handle *h, int tag)
{
    if (!h->prof) return;
    if (h->nev == h->pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return;
        h->pool.push_back(e);
    }
    hipEvent_t e = h->pool[h->nev++];
    (void)hipEventRecord(e, h->stream);
    h->ev_tag.push_back(tag);
}

// kernel families (one name per kernel shape; _x2 = two passes in one launch)
enum {
    PT_CURL, PT_RAD2, PT_RAD1, PT_EIG2, PT_EIG1, PT_ST2, PT_ST1, PT_ST2RK, PT_ST1RK, PT_RK, PT_FILTTT, PT_BC,
    PT_RHSA, PT_RHSB, PT_POST, PT_DT, PT_TT2, PT_EIG3, PT_N
};
static const char *PT_NAMES[PT_N] = {"curl", "avrg_radial_x2", "avrg_radial_x1", "avrgh_x2", "avrgh_x1",
                                     "avrg_state_x2", "avrg_state_x1", "avrg_state_x2_rk", "avrg_state_x1_rk",
                                     "rk_update", "avrgtt", "bndrytt", "rhs_flux_E_hall", "rhs_curlE_pressure",
                                     "post_add_eq", "setdt", "avrgtt_x2_eq", "avrgh_x3_addE"};

static int prof_collect(mhd_handle *h)
{
    if (h->nev == 0) return 0;
    HCHK(hipEventSynchronize(h->pool[h->nev - 1]));
    for (size_t t = 0; t + 1 < h->nev; t += 2) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, h->pool[t], h->pool[t + 1]);
        int tag = h->ev_tag[t];
        h->pt[tag].ms += ms;
        h->pt[tag].calls++;
    }
    h->nev = 0;
    h->ev_tag.clear();
    return 0;
}

#define PROF_BEGIN(tag) prof_mark(h, tag)
#define PROF_END(tag) prof_mark(h, tag)

static dim3 int_grid(const mhd_handle *h, dim3 b)
{
    return dim3((unsigned)blocks(h->NX - 4, b.x), (unsigned)blocks(h->NY - 4, b.y), (unsigned)(h->NZ - 4));
}

static Eq eq_state(const mhd_handle *h, bool sub)
{
    Eq e;
    for (int m = 0; m < 8; m++) e.f[m] = nullptr;
    if (sub) {
        e.f[0] = h->eq; e.f[4] = h->eq + h->S; e.f[5] = h->eq + 2 * h->S;
        e.f[6] = h->eq + 3 * h->S; e.f[7] = h->eq + 4 * h->S;
    }
    return e;
}

static const dim3 BLK(64, 4, 1);

static dim3 tile_grid(const mhd_handle *h)
{
    int nyt = h->g.yhi - h->g.ylo + 1;
    return dim3((unsigned)blocks(h->NX, TX), (unsigned)blocks(nyt, TY), (unsigned)blocks(h->NZ - 4, h->kc));
}

static dim3 z_grid(const mhd_handle *h, int nfblocks, int uz)
{
    int nyt = h->g.yhi - h->g.ylo + 1;
    return dim3((unsigned)blocks(h->NX, TX), (unsigned)blocks(nyt, BY), (unsigned)(blocks(h->NZ, uz) * nfblocks));
}

// one smoothing sweep a -> b (RKS < 0) or the last avrg pass of RK stage
// RKS fused with its register update (b unused).  All NF fields in one
// block and 4 planes per thread measured fastest on MI355X at 256x128x256
// (against one field per block, 8 planes, XCD-contiguous block order).
template <int NF, int KIND, int TYPE, int RKS>
static void launch_sweep(mhd_handle *h, const double *a, double *b, const double *addf = nullptr)
{
    RKArgs rk{h->dtv, h->xfold, h->xm, h->x, h->eq};
    hipLaunchKernelGGL((k_sweepz<NF, KIND, TYPE, NF, SWEEP_UZ, RKS>), z_grid(h, 1, SWEEP_UZ), dim3(TX * BY), 0,
                       h->stream, h->g, a, b, rk, 0, addf);
}

// NP passes a -> b in one temporal-blocked kernel (one field per block);
// with RKS >= 0 the last pass ends in RK stage RKS (b unused)
// Block tables of the two-pass temporal-blocked kernels over nf fields: the
// windows that hold no x face and no image row (the march without face
// rules) and the rest.  The hardware hands dispatch position L
// to XCD L % 8.  Order 0 enumerates x tile, y tile, chunk, field (x fastest)
// by position.  Order 1 gives each XCD a contiguous range of an enumeration
// with the field fastest, then x tile, y tile, chunk: one tile's fields read
// the same radial coefficients together, and neighbouring tiles share halo
// columns and rows, in one L2.
// rows per block of the face-window launch: 12 with periodic y (the face
// windows are whole x-tile columns then, tiled in rows of their own), else
// the rows of the other windows
static int face_rows(const mhd_handle *h) { return h->g.periodic && h->face12 ? 12 : ZB; }

static int zs_tables(mhd_handle *h, int nf, int np, ZTab *out)
{
    const int W = TX - 2 * np, H = ZB - 2 * np, HF = face_rows(h) - 2 * np;
    const Geo &g = h->g;
    const unsigned gx = (unsigned)blocks(h->NX, W);
    const unsigned gy = (unsigned)blocks(g.yhi - g.ylo + 1, H);
    const unsigned gyf = (unsigned)blocks(g.yhi - g.ylo + 1, HF);
    const unsigned nch = (unsigned)blocks(h->NZ - 4, h->zkc);
    if (gx >= (1u << ZT_BY) || std::max(gy, gyf) >= (1u << (ZT_CH - ZT_BY)) || nch >= (1u << (ZT_F - ZT_CH)) ||
        nf > (1 << (32 - ZT_F))) {
        set_err(h, "grid too large for the smoothing kernels' block table");
        return MHD_ERR_ARG;
    }
    // the window of tile (bx, by) holds an x face or an image row (the same
    // window placement as the kernel's)
    auto faces = [&](unsigned bx, unsigned by) {
        const int i0 = std::max(0, std::min((int)bx * W - np, h->NX - TX));
        bool f = i0 < 2 || i0 + TX > h->NX - 2;
        if (!g.periodic) {
            const int j0 = std::max(0, std::min(g.ylo + (int)by * H - np, h->NY - ZB));
            f = f || j0 < 2 || j0 + ZB > h->NY - 2;
        }
        return f;
    };
    std::vector<unsigned> list[2];
    // class c of the windows: with periodic y a window holds a face exactly
    // when its x tile does, and the face class is tiled in rows of HF
    auto add = [&](int c, unsigned bx, unsigned by, unsigned ch, unsigned m) {
        const bool f = g.periodic ? faces(bx, 0) : faces(bx, by);
        if ((f ? 1 : 0) == c) list[c].push_back(bx | by << ZT_BY | ch << ZT_CH | m << ZT_F);
    };
    for (int c = 0; c < 2; c++) {
        const unsigned ny = (c == 1 && g.periodic) ? gyf : gy;
        if (h->zord == 0) {
            for (unsigned m = 0; m < (unsigned)nf; m++)
                for (unsigned ch = 0; ch < nch; ch++)
                    for (unsigned by = 0; by < ny; by++)
                        for (unsigned bx = 0; bx < gx; bx++) add(c, bx, by, ch, m);
        } else {
            for (unsigned ch = 0; ch < nch; ch++)
                for (unsigned by = 0; by < ny; by++)
                    for (unsigned bx = 0; bx < gx; bx++)
                        for (unsigned m = 0; m < (unsigned)nf; m++) add(c, bx, by, ch, m);
        }
    }
    for (int c = 0; c < 2; c++) {
        const unsigned T = (unsigned)list[c].size();
        out[c].n = T;
        if (T == 0) continue;
        std::vector<unsigned> w(T);
        const unsigned per = T / NXCD, rem = T % NXCD;
        for (unsigned L = 0; L < T; L++) {
            const unsigned xcd = L % NXCD, n = L / NXCD;
            w[L] = h->zord == 0 ? list[c][L] : list[c][xcd * per + (xcd < rem ? xcd : rem) + n];
        }
        HCHK(hipMalloc(&out[c].d, T * sizeof(unsigned)));
        HCHK(hipMemcpy(out[c].d, w.data(), T * sizeof(unsigned), hipMemcpyHostToDevice));
        out[c].nat = std::move(list[c]);
    }
    return MHD_OK;
}

// ---- recording for the batched ensemble kernels ----
template <class A>
static void rec_add(mhd_handle *h, BLaunch fn, const A &a, std::vector<unsigned> words)
{
    static_assert(std::is_trivially_copyable<A>::value, "argument blocks travel as bytes");
    h->rec->push_back(RecLaunch{fn, std::string((const char *)&a, sizeof(A)), std::move(words)});
}

// argument blocks start zeroed, padding included (they are compared as bytes)
template <class A>
static A zeroed()
{
    A a;
    std::memset((void *)&a, 0, sizeof(A));
    return a;
}

// words 0..n-1 (one-dimensional launches)
static std::vector<unsigned> lin_words(unsigned n)
{
    std::vector<unsigned> w(n);
    for (unsigned i = 0; i < n; i++) w[i] = i;
    return w;
}

// words of a gx x gy x gz grid, x fastest: bx | by << ZT_BY | bz << ZT_CH
static std::vector<unsigned> grid_words(mhd_handle *h, dim3 gr)
{
    if (gr.x >= (1u << ZT_BY) || gr.y >= (1u << (ZT_CH - ZT_BY)) || gr.z >= (1u << (32 - ZT_CH))) {
        h->rec_err = "grid too large for the batched kernels' block words";
        return {};
    }
    std::vector<unsigned> w;
    w.reserve((size_t)gr.x * gr.y * gr.z);
    for (unsigned z = 0; z < gr.z; z++)
        for (unsigned y = 0; y < gr.y; y++)
            for (unsigned x = 0; x < gr.x; x++) w.push_back(x | y << ZT_BY | z << ZT_CH);
    return w;
}

template <int NP, int KIND, int TYPE, int RKS, int MODE, int RB>
static void bl_zs(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL((kb_zsmooth<NP, KIND, TYPE, RKS, MODE, RB>), dim3(n), dim3(TX * RB), 0, s, (const BZArgs *)a, t);
}

static void bl_curl(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_curl, dim3(n), dim3(NT), 0, s, (const BCurlArgs *)a, t);
}

static void bl_rhs_a(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_rhs_a, dim3(n), dim3(NT), 0, s, (const BRhsAArgs *)a, t);
}

static void bl_rhs_b(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_rhs_b, dim3(n), dim3(NT), 0, s, (const BRhsBArgs *)a, t);
}

template <int PART, int RB>
static void bl_rhs_b2(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL((kb_rhs_b2<PART, RB>), dim3(n), dim3(TX * RB), 0, s, (const BRhsB2Args *)a, t);
}

static void bl_fill(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_fill_state, dim3(n), dim3(256), 0, s, (const BFillArgs *)a, t);
}

static void bl_dt_partial(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_dt_partial, dim3(n), dim3(256), 0, s, (const BDtArgs *)a, t);
}

static void bl_dt_final(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_dt_final, dim3(n), dim3(256), 0, s, (const BDtArgs *)a, t);
}

static void bl_restore(const void *a, const BEntry *t, unsigned n, hipStream_t s)
{
    hipLaunchKernelGGL(kb_restore, dim3(n), dim3(256), 0, s, (const BRestoreArgs *)a, t);
}

// blocks per run of the batched collapse undo (it copies only after a collapse)
static constexpr unsigned RESTORE_BLOCKS = 128;

// the side stream starts after the main stream's work so far (fork) and the
// main stream continues after the side stream's (join)
static void fork_side(mhd_handle *h)
{
    HCHK_V(hipEventRecord(h->fork_ev, h->stream));
    HCHK_V(hipStreamWaitEvent(h->side, h->fork_ev, 0));
}

static void join_side(mhd_handle *h)
{
    HCHK_V(hipEventRecord(h->join_ev, h->side));
    HCHK_V(hipStreamWaitEvent(h->stream, h->join_ev, 0));
}

template <int NP, int KIND, int TYPE, int RKS, int MODE>
static void launch_zs(mhd_handle *h, dim3 grid, const ZTab *zt, int nch, const double *a, double *b,
                      const RKArgs &rk, hipStream_t st = nullptr)
{
    if (!st) st = h->stream;
    const unsigned *tab = zt ? zt->d : nullptr;
    if (h->rec) {
        BZArgs z = zeroed<BZArgs>();
        z.g = h->g; z.A = a; z.B = b; z.rk = rk; z.kc = h->zkc;
        std::vector<unsigned> w;
        if (zt) {
            w = zt->nat;
        } else {
            // the full march's enumeration (k_zsmooth ZM_ALL): x tile fastest,
            // then y tile, chunk, field
            const unsigned nf = grid.z / (unsigned)nch;
            for (unsigned m = 0; m < nf; m++)
                for (unsigned ch = 0; ch < (unsigned)nch; ch++)
                    for (unsigned y = 0; y < grid.y; y++)
                        for (unsigned x = 0; x < grid.x; x++) w.push_back(x | y << ZT_BY | ch << ZT_CH | m << ZT_F);
        }
        BLaunch fn = bl_zs<NP, KIND, TYPE, RKS, MODE, ZB>;
        if constexpr (MODE == ZM_FACE) {
            if (face_rows(h) == 12) fn = bl_zs<NP, KIND, TYPE, RKS, MODE, 12>;
        }
        rec_add(h, fn, z, std::move(w));
        return;
    }
    if constexpr (MODE == ZM_FACE) {
        if (face_rows(h) == 12) {
            hipLaunchKernelGGL((k_zsmooth<NP, KIND, TYPE, RKS, MODE, 12>), grid, dim3(TX * 12), 0, st, h->g, a, b, rk,
                               tab, nch, h->zkc);
            return;
        }
    }
    hipLaunchKernelGGL((k_zsmooth<NP, KIND, TYPE, RKS, MODE>), grid, dim3(TX * ZB), 0, st, h->g, a, b, rk, tab, nch,
                       h->zkc);
}

// NP smoothing passes of NF fields a -> b.  The RK stage sweeps (two more
// operand streams per point) run as one launch over all windows with the
// full march; the other sweeps in two launches, the windows without face
// rules, then the windows with them (measured faster for each).
template <int NF, int NP, int KIND, int TYPE, int RKS>
static void launch_zsmooth(mhd_handle *h, const double *a, double *b, const double *addf = nullptr)
{
    static_assert(NF == 3 || NF == 8, "block tables exist for 3 and 8 fields");
    static_assert(NP == 2 || (NP == 3 && NF == 3), "three-pass tables exist for 3 fields");
    const RKArgs rk{h->dtv, h->xfold, h->xm, h->x, RKS == RKS_ADD ? addf : h->eq};
    if constexpr (RKS >= 0) {
        const int W = TX - 2 * NP, H = ZB - 2 * NP;
        const int nch = blocks(h->NZ - 4, h->zkc);
        const dim3 grid((unsigned)blocks(h->NX, W), (unsigned)blocks(h->g.yhi - h->g.ylo + 1, H), (unsigned)(nch * NF));
        launch_zs<NP, KIND, TYPE, RKS, ZM_ALL>(h, grid, nullptr, nch, a, b, rk);
    } else {
        const ZTab *zt = NP == 3 ? h->zt3 : h->zt[NF == 8 ? 1 : 0];
        if (zt[0].n && zt[1].n && h->overlap && !h->rec) {
            // the face windows on the side stream beside the plain ones
            fork_side(h);
            launch_zs<NP, KIND, TYPE, RKS, ZM_FACE>(h, dim3(zt[1].n), &zt[1], 0, a, b, rk, h->side);
            launch_zs<NP, KIND, TYPE, RKS, ZM_PLAIN>(h, dim3(zt[0].n), &zt[0], 0, a, b, rk);
            join_side(h);
            return;
        }
        if (zt[0].n) launch_zs<NP, KIND, TYPE, RKS, ZM_PLAIN>(h, dim3(zt[0].n), &zt[0], 0, a, b, rk);
        if (zt[1].n) launch_zs<NP, KIND, TYPE, RKS, ZM_FACE>(h, dim3(zt[1].n), &zt[1], 0, a, b, rk);
    }
}

// kk smoothing sweeps (+ fills) on a 3-field array; returns the result
// buffer.  With addf, the result is the smoothed field plus addf, added by
// the launch that runs the last pass (three passes for avrgh, else a
// single-pass sweep)
static double *smooth_aux(mhd_handle *h, double *a, double *b, int kind, int kk, const double *addf = nullptr)
{
    while (kk > 0) {
        // the last three avrgh passes and the added field in one launch
        if (h->zs && h->zs3 && kind == KIND_EIGHTH && addf && kk == 3) {
            PROF_BEGIN(PT_EIG3);
            launch_zsmooth<3, 3, KIND_EIGHTH, BC_AUX, RKS_ADD>(h, a, b, addf);
            PROF_END(PT_EIG3);
            std::swap(a, b);
            break;
        }
        const bool two = h->zs && kk >= 2 && !(addf && kk == 2);
        const int tag = kind == KIND_RADIAL ? (two ? PT_RAD2 : PT_RAD1) : (two ? PT_EIG2 : PT_EIG1);
        PROF_BEGIN(tag);
        if (two) {
            if (kind == KIND_RADIAL) launch_zsmooth<3, 2, KIND_RADIAL, BC_AUX, -1>(h, a, b);
            else launch_zsmooth<3, 2, KIND_EIGHTH, BC_AUX, -1>(h, a, b);
            kk -= 2;
        } else {
            const double *add = kk == 1 ? addf : nullptr;
            if (kind == KIND_RADIAL) launch_sweep<3, KIND_RADIAL, BC_AUX, -1>(h, a, b, add);
            else launch_sweep<3, KIND_EIGHTH, BC_AUX, -1>(h, a, b, add);
            kk -= 1;
        }
        PROF_END(tag);
        std::swap(a, b);
    }
    return a;
}

// current (t3dsdsc.f:2452-2500) of state X into h->cur
static void current(mhd_handle *h, const double *X)
{
    if (h->rec) {
        BCurlArgs a = zeroed<BCurlArgs>();
        a.g = h->g; a.X = X; a.bxi = h->eq + h->S; a.byi = h->eq + 2 * h->S; a.bzi = h->eq + 3 * h->S;
        a.cur = h->cA; a.kc = h->kc;
        rec_add(h, bl_curl, a, grid_words(h, tile_grid(h)));
    } else {
        PROF_BEGIN(PT_CURL);
        hipLaunchKernelGGL(k_curl, tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->eq + h->S, h->eq + 2 * h->S,
                           h->eq + 3 * h->S, h->cA, h->kc);
        PROF_END(PT_CURL);
    }
    h->cur = smooth_aux(h, h->cA, h->cB, KIND_RADIAL, 4);
}

// right (t3dsdsc.f:1297-1673) of state X; result in h->xd
// with rks >= 0 the last sweep performs RK stage rks (nothing is left in h->xd)
static void rhs(mhd_handle *h, const double *X, bool smooth_xdif, bool init, int rks)
{
    const mhd_config &c = h->cfg;
    double eta = init ? 0.0 : c.eta0, fmu = init ? 0.0 : c.fmu0;
    double frho = init ? 0.0 : c.frho0, fpr = init ? 0.0 : c.fpr0;
    current(h, X);
    if (h->rec_x) {
        (void)hipMemcpyAsync(h->rec_x + (long)h->rec_n * 8 * h->S, X, 8 * h->S * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream);
        (void)hipMemcpyAsync(h->rec_c + (long)h->rec_n * 3 * h->S, h->cur, 3 * h->S * sizeof(double),
                             hipMemcpyDeviceToDevice, h->stream);
        h->rec_n++;
    }
    if (h->rec) {
        BRhsAArgs a = zeroed<BRhsAArgs>();
        a.g = h->g; a.X = X; a.pri = h->eq + 4 * h->S; a.cur = h->cur; a.xresd = h->xresd;
        a.eta = eta; a.cht = c.di; a.hall = c.hall; a.kc = h->kc;
        a.XD = h->xdA; a.E = h->eA; a.XH = h->hA;
        rec_add(h, bl_rhs_a, a, grid_words(h, tile_grid(h)));
    } else {
        PROF_BEGIN(PT_RHSA);
        hipLaunchKernelGGL(k_rhs_a, tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->eq + 4 * h->S, h->cur,
                           h->xresd, eta, c.hall, c.di, h->xdA, h->eA, h->hA, h->kc);
        PROF_END(PT_RHSA);
    }
    // with the Hall term the last avrgh sweep leaves E + xdh, the field whose
    // curl enters the induction equation (t3dsdsc.f:1426-1433)
    double *E;
    if (c.hall && h->overlap && h->zs && h->zs3 && !h->rec && !h->prof) {
        // the first two avrgh passes need no E: they run on the second
        // stream pair (swapped in as the handle's streams for the launch)
        // while E is smoothed, and the last three add E on the main stream
        HCHK_V(hipEventRecord(h->aux_ev, h->stream));
        HCHK_V(hipStreamWaitEvent(h->aux, h->aux_ev, 0));
        std::swap(h->stream, h->aux);
        std::swap(h->side, h->aux_side);
        launch_zsmooth<3, 2, KIND_EIGHTH, BC_AUX, -1>(h, h->hA, h->hB);
        std::swap(h->stream, h->aux);
        std::swap(h->side, h->aux_side);
        E = smooth_aux(h, h->eA, h->eB, KIND_RADIAL, 4);
        HCHK_V(hipEventRecord(h->aux_ev, h->aux));
        HCHK_V(hipStreamWaitEvent(h->stream, h->aux_ev, 0));
        E = smooth_aux(h, h->hB, h->hA, KIND_EIGHTH, 3, E);
    } else {
        E = smooth_aux(h, h->eA, h->eB, KIND_RADIAL, 4);
        if (c.hall) E = smooth_aux(h, h->hA, h->hB, KIND_EIGHTH, 5, E);
    }
    const double *XH = nullptr;
    bool vd = fmu != 0.0 || frho != 0.0 || fpr != 0.0;
    PROF_BEGIN(PT_RHSB);
    if (h->rec && !vd && h->rb2) {
        const int nyt = h->g.yhi - h->g.ylo + 1;
        const int rb = h->rb2 == 2 ? 8 : 16;
        dim3 grid((unsigned)blocks(h->NX, TX - 4), (unsigned)blocks(nyt, rb - 4), (unsigned)blocks(h->NZ - 4, RB_KC));
        BRhsB2Args a = zeroed<BRhsB2Args>();
        a.g = h->g; a.X = X; a.cur = h->cur; a.E = E; a.eta = eta; a.gam = c.gamma; a.XD = h->xdA;
        rec_add(h, rb == 8 ? bl_rhs_b2<0, 8> : bl_rhs_b2<0, 16>, a, grid_words(h, grid));
        rec_add(h, rb == 8 ? bl_rhs_b2<1, 8> : bl_rhs_b2<1, 16>, a, grid_words(h, grid));
        BFillArgs f = zeroed<BFillArgs>();
        f.g = h->g; f.XD = h->xdA; f.nf = 4;
        rec_add(h, bl_fill, f, lin_words((unsigned)blocks(face_count(h->g), 256)));
    } else if (h->rec && !vd) {
        BRhsBArgs a = zeroed<BRhsBArgs>();
        a.g = h->g; a.X = X; a.cur = h->cur; a.E = E; a.XH = XH; a.eta = eta; a.gam = c.gamma;
        a.fmu = fmu; a.frho = frho; a.fpr = fpr; a.XD = h->xdA; a.kc = h->kc;
        rec_add(h, bl_rhs_b, a, grid_words(h, tile_grid(h)));
    } else if (vd) {
        hipLaunchKernelGGL((k_rhs_b<true>), tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->cur, E, XH, eta,
                           c.gamma, fmu, frho, fpr, h->xdA, h->kc);
    } else if (h->rb2) {
        // -curl(E + xdh), the pressure equation, and the fill of xdif(1..4)
        const int nyt = h->g.yhi - h->g.ylo + 1;
        const int rb = h->rb2 == 2 ? 8 : 16;
        dim3 grid((unsigned)blocks(h->NX, TX - 4), (unsigned)blocks(nyt, rb - 4), (unsigned)blocks(h->NZ - 4, RB_KC));
        // the two parts write different fields: part 0 on the side stream
        hipStream_t s0 = h->stream;
        if (h->overlap) {
            fork_side(h);
            s0 = h->side;
        }
        if (rb == 8) {
            hipLaunchKernelGGL((k_rhs_b2<0, 8>), grid, dim3(TX * 8), 0, s0, h->g, X, h->cur, E, eta, c.gamma, h->xdA);
            hipLaunchKernelGGL((k_rhs_b2<1, 8>), grid, dim3(TX * 8), 0, h->stream, h->g, X, h->cur, E, eta, c.gamma, h->xdA);
        } else {
            hipLaunchKernelGGL((k_rhs_b2<0, 16>), grid, dim3(TX * 16), 0, s0, h->g, X, h->cur, E, eta, c.gamma, h->xdA);
            hipLaunchKernelGGL((k_rhs_b2<1, 16>), grid, dim3(TX * 16), 0, h->stream, h->g, X, h->cur, E, eta, c.gamma, h->xdA);
        }
        hipLaunchKernelGGL(k_fill_state, dim3(blocks(face_count(h->g), 256)), dim3(256), 0, h->stream, h->g, h->xdA, 4);
        if (h->overlap) join_side(h);
    } else {
        hipLaunchKernelGGL((k_rhs_b<false>), tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->cur, E, XH, eta,
                           c.gamma, fmu, frho, fpr, h->xdA, h->kc);
    }
    PROF_END(PT_RHSB);
    double *a = h->xdA, *b = h->xdB;
    if (smooth_xdif && h->zs) {
        // avrg twice (t3dsdsc.f:1668-1671) in one temporal-blocked kernel;
        // with rks >= 0 it also performs the RK stage update
        PROF_BEGIN(rks >= 0 ? PT_ST2RK : PT_ST2);
        switch (rks) {
        case 0: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, 0>(h, a, nullptr); break;
        case 1: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, 1>(h, a, nullptr); break;
        case 2: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, 2>(h, a, nullptr); break;
        case 3: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, 3>(h, a, nullptr); break;
        case 4: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, 4>(h, a, nullptr); break;
        default: launch_zsmooth<8, 2, KIND_STATE, BC_STATE, -1>(h, a, b); std::swap(a, b); break;
        }
        PROF_END(rks >= 0 ? PT_ST2RK : PT_ST2);
        if (rks >= 0) a = nullptr;
    } else if (smooth_xdif) {
        // avrg twice (t3dsdsc.f:1668-1671); with rks >= 0 the second pass
        // also performs the RK stage update
        PROF_BEGIN(PT_ST1);
        launch_sweep<8, KIND_STATE, BC_STATE, -1>(h, a, b);
        PROF_END(PT_ST1);
        std::swap(a, b);
        PROF_BEGIN(rks >= 0 ? PT_ST1RK : PT_ST1);
        switch (rks) {
        case 0: launch_sweep<8, KIND_STATE, BC_STATE, 0>(h, a, nullptr); break;
        case 1: launch_sweep<8, KIND_STATE, BC_STATE, 1>(h, a, nullptr); break;
        case 2: launch_sweep<8, KIND_STATE, BC_STATE, 2>(h, a, nullptr); break;
        case 3: launch_sweep<8, KIND_STATE, BC_STATE, 3>(h, a, nullptr); break;
        default: launch_sweep<8, KIND_STATE, BC_STATE, -1>(h, a, b); std::swap(a, b); break;
        }
        PROF_END(rks >= 0 ? PT_ST1RK : PT_ST1);
        if (rks >= 0) a = nullptr;
    }
    h->xd = a;
}

static void rk(mhd_handle *h, int stage)
{
    long n8 = 8 * h->S;
    PROF_BEGIN(PT_RK);
    hipLaunchKernelGGL(k_rk, dim3(blocks(n8, 256)), dim3(256), 0, h->stream, n8, stage, h->dtv, h->xd, h->xfold,
                       h->xm, h->x);
    PROF_END(PT_RK);
}

static void launch_bc_tt(mhd_handle *h, double *B, const double *A, const double *X, Eq eq)
{
    long n = face_count(h->g);
    PROF_BEGIN(PT_BC);
    hipLaunchKernelGGL(k_bc_tt, dim3(blocks(n, 256)), dim3(256), 0, h->stream, h->g, B, A, X, eq);
    PROF_END(PT_BC);
}

// stepon body (t3dsdsc.f:1079-1165) with dt in h->dtv[0]
static void step_body(mhd_handle *h)
{
    std::swap(h->x, h->xfold);   // xfold = x (1079-1084) by renaming
    // temporal-blocked path: the last stage leaves xm = x - eq and one kernel
    // runs both avrgtt passes with their bndrytt fills and adds eq back
    const bool tb = h->zs && h->fuse_rk;
    for (int s = 0; s < 4; s++) {
        const double *X = s == 0 ? h->xfold : h->x;
        if (h->fuse_rk) {
            rhs(h, X, true, false, (tb && s == 3) ? 4 : s);
        } else {
            rhs(h, X, true, false, -1);
            rk(h, s);
        }
    }
    if (tb) {
        PROF_BEGIN(PT_TT2);
        launch_zsmooth<8, 2, KIND_STATE, BC_TT, -1>(h, h->xm, h->x);
        PROF_END(PT_TT2);
        return;
    }
    // perturbation smoothing (1135-1165): xm = x - eq; avrgtt x2; x = xm + eq
    Eq sub = eq_state(h, true);
    Eq none = eq_state(h, false);
    PROF_BEGIN(PT_FILTTT);
    hipLaunchKernelGGL(k_smooth_tt<SRC_XEQ>, tile_grid(h), dim3(NT), 0, h->stream, h->g, h->x, sub, h->xm, h->kc);
    PROF_END(PT_FILTTT);
    launch_bc_tt(h, h->xm, nullptr, h->x, sub);
    PROF_BEGIN(PT_FILTTT);
    hipLaunchKernelGGL(k_smooth_tt<SRC_BUF>, tile_grid(h), dim3(NT), 0, h->stream, h->g, h->xm, none, h->xdB, h->kc);
    PROF_END(PT_FILTTT);
    launch_bc_tt(h, h->xdB, h->xm, nullptr, none);
    PROF_BEGIN(PT_POST);
    hipLaunchKernelGGL(k_add_eq, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, h->S, h->S, h->xdB, sub, h->x);
    PROF_END(PT_POST);
}

// part: h->nred partial minima (the handle's reduction buffer by default)
static void dt_kernels(mhd_handle *h, int advance_time, double *part = nullptr)
{
    const mhd_config &c = h->cfg;
    double wcoef = (c.whistler_dt && c.hall && c.di != 0.0) ? 3.141592654 * c.di : 0.0;
    if (!part) part = h->red;
    const double *rr = h->mag && c.magnetosphere ? h->rr : nullptr;
    if (h->rec) {
        BDtArgs a = zeroed<BDtArgs>();
        a.g = h->g; a.X = h->x; a.rr = rr; a.eta = c.eta0; a.fmu = c.fmu0; a.cfl = c.cfl; a.gam = c.gamma;
        a.wcoef = wcoef; a.part = part; a.dtv = h->dtv; a.nb = h->nred; a.advance_time = advance_time;
        rec_add(h, bl_dt_partial, a, lin_words((unsigned)h->nred));
        rec_add(h, bl_dt_final, a, lin_words(1));
        return;
    }
    PROF_BEGIN(PT_DT);
    hipLaunchKernelGGL(k_dt_partial, dim3(h->nred), dim3(256), 0, h->stream, h->g, h->x, c.eta0, c.fmu0, c.cfl,
                       c.gamma, wcoef, rr, part);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, h->stream, part, h->nred, h->dtv, advance_time);
    PROF_END(PT_DT);
}

// the guarded loop's undo of a step after a collapse (see k_dt_final)
static void restore_after_collapse(mhd_handle *h)
{
    if (h->rec) {
        BRestoreArgs a = zeroed<BRestoreArgs>();
        a.dtv = h->dtv; a.XF = h->xfold; a.X = h->x; a.n = 8 * h->S; a.nb = RESTORE_BLOCKS;
        rec_add(h, bl_restore, a, lin_words(RESTORE_BLOCKS));
        return;
    }
    hipLaunchKernelGGL(k_restore, dim3(1024), dim3(256), 0, h->stream, h->dtv, h->xfold, h->x, 8 * h->S);
}

// ---- magnetosphere mode (kernels in mhd_magneto.h) ----
static Mag mag_args(const mhd_handle *h)
{
    Mag a;
    a.rr = h->rr;
    a.cht = h->chtab;
    a.rcr = h->cfg.rcr;
    a.mask = h->cfg.magnetosphere != 0;
    a.inflow = h->cfg.inflow != 0;
    a.partial = h->cfg.partial_e != 0;
    a.zd = h->zd; a.yd = h->yd; a.zd2 = h->zd2; a.yd2 = h->yd2;
    a.fb = h->face;
    a.mrow = h->mrow;
    return a;
}

// time factors of the drive at stage time t (t3dsdsc.f:1714-1768)
static Stage mag_stage(const mhd_handle *h, double t, double ddt, int km)
{
    Stage s;
    s.tt = tanh((t - 9.) / 1.) - tanh((t - 39.) / 1.5);
    s.sq = sqrt(t * t * t * 7.25 * 7.25 * 7.25);
    s.de = exp(t * 7.25 * (double)0.0069f) + 1.;
    s.ddt = ddt;
    s.rde = 1.0 / s.de;
    s.den = ddt + 1.e-9;
    s.rden = 1.0 / s.den;
    s.ref = km == 1 ? h->xfold : h->xm;
    s.x = h->x;
    return s;
}

static dim3 vol_grid(const mhd_handle *h) { return dim3((unsigned)blocks((long)h->NX * h->NY * h->NZ, 256)); }

// a boundary routine: mask (by the producer of A as it stores, else a
// launch over the volume), then the x rules, then the y and z rules
template <int TYPE>
static void mag_fill(mhd_handle *h, double *A, int nf, const Stage &s)
{
    const Mag a = mag_args(h);
    const Geo &g = h->g;
    if constexpr (TYPE != BC_TT) {
        // the whole fill in one launch (the mask is the producer's)
        if (!h->fill_split && (h->mrow || !a.mask)) {
            const long rows = (long)g.NY * g.NZ;
            const long n = rows * nf + (TYPE == BC_STATE && a.inflow ? 2 * rows : 0) +
                           ((long)g.NX * (g.NZ - 4) + (long)g.NX * g.NY) * nf;
            hipLaunchKernelGGL((km_fill_xyz<TYPE>), dim3(blocks(n, 256)), dim3(256), 0, h->stream, g, a, s, h->eq,
                               A, nf);
            return;
        }
    }
    if (a.mask && !h->mrow) hipLaunchKernelGGL(km_mask, vol_grid(h), dim3(256), 0, h->stream, g, a, A, nf);
    const long rows = (long)g.NY * g.NZ;
    const long nx = rows * nf + (TYPE == BC_STATE && a.inflow ? 2 * rows : 0);
    hipLaunchKernelGGL((km_fill_x<TYPE>), dim3(blocks(nx, 256)), dim3(256), 0, h->stream, g, a, s, h->eq, A, nf);
    if (h->fill_split) {
        hipLaunchKernelGGL((km_fill_y<TYPE>), dim3(blocks((long)g.NX * g.NZ, 256)), dim3(256), 0, h->stream, g, A, nf);
        hipLaunchKernelGGL((km_fill_z<TYPE>), dim3(blocks((long)g.NX * g.NY, 256)), dim3(256), 0, h->stream, g, A, nf);
    } else {
        const long n = ((long)g.NX * (g.NZ - 4) + (long)g.NX * g.NY) * nf;
        hipLaunchKernelGGL((km_fill_yz<TYPE>), dim3(blocks(n, 256)), dim3(256), 0, h->stream, g, A, nf);
    }
}

// kk sweeps of nf fields, each followed by the boundary routine TYPE;
// a and b alternate, the result is returned
template <int KIND, int TYPE>
static double *mag_smooth(mhd_handle *h, double *a, double *b, int nf, int kk, const Stage &s)
{
    for (int p = 0; p < kk; p++) {
        hipLaunchKernelGGL((km_sweep<KIND>), vol_grid(h), dim3(256), 0, h->stream, h->g, a, b, nf, h->mrow,
                           TYPE == BC_TT ? 1 : 0);
        mag_fill<TYPE>(h, b, nf, s);
        std::swap(a, b);
    }
    return a;
}

// current (t3dsdsc.f:2452-2500) of state X into h->cur
static void mag_current(mhd_handle *h, const double *X)
{
    const Stage none = mag_stage(h, 0.0, 1.0, 1);
    PROF_BEGIN(PT_CURL);
    hipLaunchKernelGGL(km_curl, vol_grid(h), dim3(256), 0, h->stream, h->g, X, h->eq, h->cA, h->mrow);
    mag_fill<BC_AUX>(h, h->cA, 3, none);
    h->cur = mag_smooth<KIND_RADIAL, BC_AUX>(h, h->cA, h->cB, 3, 4, none);
    PROF_END(PT_CURL);
}

// right(t, ddt, km) (t3dsdsc.f:1297-1673) of the state h->x; result in h->xd
static void mag_rhs(mhd_handle *h, double t, double ddt, int km, bool smooth_xdif, bool init)
{
    const mhd_config &c = h->cfg;
    const double eta = init ? 0.0 : c.eta0;
    const Mag a = mag_args(h);
    const Stage st = mag_stage(h, t, ddt, km);
    const double *X = h->x;
    mag_current(h, X);
    PROF_BEGIN(PT_RHSA);
    // the flux tensor once per point into xdB (free until the xdif smoothing)
    hipLaunchKernelGGL(km_flux, vol_grid(h), dim3(256), 0, h->stream, h->g, X, h->eq, h->xdB);
    hipLaunchKernelGGL(km_rhs_a, vol_grid(h), dim3(256), 0, h->stream, h->g, a, X, h->eq, h->cur, h->xresd, eta,
                       h->xdB, h->xdA, h->eA, h->vel);
    double *E = mag_smooth<KIND_RADIAL, BC_AUX>(h, h->eA, h->eB, 3, 4, st);
    PROF_END(PT_RHSA);
    if (c.hall) {
        PROF_BEGIN(PT_EIG1);
        hipLaunchKernelGGL(km_hall, vol_grid(h), dim3(256), 0, h->stream, h->g, a, X, h->cur, c.di, h->hA);
        mag_fill<BC_AUX>(h, h->hA, 3, st);
        const double *H = mag_smooth<KIND_EIGHTH, BC_AUX>(h, h->hA, h->hB, 3, 5, st);
        hipLaunchKernelGGL(km_add3, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, h->S, h->S, E, H);
        PROF_END(PT_EIG1);
    }
    PROF_BEGIN(PT_RHSB);
    hipLaunchKernelGGL(km_rhs_b, vol_grid(h), dim3(256), 0, h->stream, h->g, a, X, h->eq, h->cur, E, h->vel, eta,
                       c.gamma, h->xdA);
    mag_fill<BC_STATE>(h, h->xdA, 8, st);
    PROF_END(PT_RHSB);
    double *r = h->xdA;
    if (smooth_xdif) {
        PROF_BEGIN(PT_ST1);
        r = mag_smooth<KIND_STATE, BC_STATE>(h, h->xdA, h->xdB, 8, 2, st);
        PROF_END(PT_ST1);
    }
    h->xd = r;
}

// stepon (t3dsdsc.f:1063-1171) with dt in h->dtv[0]; the state is copied to
// xfold (the driven columns of x are overwritten inside right())
static void mag_step(mhd_handle *h, double time, double dt)
{
    HCHK_V(hipMemcpyAsync(h->xfold, h->x, 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    const double ts[4] = {time + dt / 2, time + dt / 2, time + dt, time + dt};
    const double dd[4] = {dt / 2., dt / 2., dt, dt / 6.};
    const int kms[4] = {1, 1, 1, 2};
    for (int s = 0; s < 4; s++) {
        mag_rhs(h, ts[s], dd[s], kms[s], true, false);
        rk(h, s);
    }
    // perturbation smoothing (1135-1165): xm = x - eq; avrgtt x2; x = xm + eq
    const Stage none = mag_stage(h, 0.0, 1.0, 1);
    PROF_BEGIN(PT_POST);
    hipLaunchKernelGGL(km_eq, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, h->S, h->S, h->x, h->eq, -1, h->xm);
    double *r = mag_smooth<KIND_STATE, BC_TT>(h, h->xm, h->xdB, 8, 2, none);
    hipLaunchKernelGGL(km_eq, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, h->S, h->S, r, h->eq, +1, h->x);
    PROF_END(PT_POST);
    h->xd = nullptr;
    mag_current(h, h->x);
}

// current of either path
static void any_current(mhd_handle *h, const double *X)
{
    if (h->mag) mag_current(h, X);
    else current(h, X);
}

// ---- host <-> device layout conversion ----
// Host arrays are the reference's padded layout (mx+4, my, mz+4); the device
// keeps the physical range only.  Transfers go through a pinned staging buffer
// in slabs of whole planes (one contiguous DMA per slab) and the rows are
// scattered / gathered on the host.
static constexpr size_t STAGE_BYTES = 64u << 20;

static int ensure_stage(mhd_handle *h)
{
    if (h->stage) return 0;
    HCHK(hipHostMalloc((void **)&h->stage, STAGE_BYTES, hipHostMallocDefault));
    return 0;
}

static int upload_fields(mhd_handle *h, double *dst, const double *src, int nf)
{
    const long px = h->cfg.mx + 4, plane = px * h->NY, total = plane * (h->cfg.mz + 4);
    const long dplane = (long)h->NX * h->NY;
    const int nzc = (int)std::max<long>(1, (long)(STAGE_BYTES / sizeof(double)) / dplane);
    if (ensure_stage(h)) return MHD_ERR_HIP;
    for (int m = 0; m < nf; m++) {
        const double *s = src + (long)m * total;
        for (int z0 = 0; z0 < h->NZ; z0 += nzc) {
            const int z1 = std::min(h->NZ, z0 + nzc);
            for (int k = z0; k < z1; k++)
                for (int j = 0; j < h->NY; j++)
                    memcpy(h->stage + ((long)(k - z0) * h->NY + j) * h->NX, s + (k + 2) * plane + j * px + 2,
                           sizeof(double) * h->NX);
            HCHK(hipMemcpyAsync(dst + m * h->S + (long)z0 * dplane, h->stage, sizeof(double) * (z1 - z0) * dplane,
                                hipMemcpyHostToDevice, h->stream));
            HCHK(hipStreamSynchronize(h->stream));
        }
    }
    return 0;
}

static int download_fields(mhd_handle *h, double *dst, const double *src, int nf, bool images = true)
{
    const long px = h->cfg.mx + 4, plane = px * h->NY;
    const long total = plane * (h->cfg.mz + 4);
    const long dplane = (long)h->NX * h->NY;
    const int nzc = (int)std::max<long>(1, (long)(STAGE_BYTES / sizeof(double)) / dplane);
    if (ensure_stage(h)) return MHD_ERR_HIP;
    for (int m = 0; m < nf; m++) {
        double *d = dst + (long)m * total;
        // the ghost planes and columns the device does not hold read 0
        memset(d, 0, sizeof(double) * 2 * plane);
        memset(d + (h->cfg.mz + 2) * plane, 0, sizeof(double) * 2 * plane);
        for (int z0 = 0; z0 < h->NZ; z0 += nzc) {
            const int z1 = std::min(h->NZ, z0 + nzc);
            HCHK(hipMemcpyAsync(h->stage, src + m * h->S + (long)z0 * dplane, sizeof(double) * (z1 - z0) * dplane,
                                hipMemcpyDeviceToHost, h->stream));
            HCHK(hipStreamSynchronize(h->stream));
            for (int k = z0; k < z1; k++)
                for (int j = 0; j < h->NY; j++) {
                    double *row = d + (k + 2) * plane + j * px;
                    row[0] = row[1] = row[px - 2] = row[px - 1] = 0.0;
                    memcpy(row + 2, h->stage + ((long)(k - z0) * h->NY + j) * h->NX, sizeof(double) * h->NX);
                }
        }
    }
    if (h->cfg.periodic_y && images) {
        // image rows (never computed on the device): 0,1,NY-2,NY-1 <- NY-4,NY-3,2,3
        int NY = h->NY;
        const int dstr[4] = {0, 1, NY - 2, NY - 1}, srcr[4] = {NY - 4, NY - 3, 2, 3};
        for (int m = 0; m < nf; m++)
            for (int k = 2; k < h->cfg.mz + 2; k++)
                for (int r = 0; r < 4; r++) {
                    double *d = dst + (long)m * total + k * plane;
                    memcpy(d + dstr[r] * px + 2, d + srcr[r] * px + 2, sizeof(double) * h->NX);
                }
    }
    return 0;
}

// ===========================================================================
// reverse mode: discrete adjoint of the step (kernels in mhd_adjoint.h)
// ===========================================================================
struct AdjBufs {
    double *ab = nullptr, *t8 = nullptr, *xb = nullptr, *xbm = nullptr, *xbf = nullptr, *tt = nullptr;  // 8 fields
    double *eb = nullptr, *hb = nullptr, *jb = nullptr, *vb = nullptr, *t3 = nullptr, *rb = nullptr;    // 3 fields
    double *xfin = nullptr;                // final state (8 fields)
    double *acc = nullptr;                 // per-block slots: adjoint of eta, then objective
    long nblk = 0;
    double *ck = nullptr;                  // checkpoints, 8 fields each
    double *tx = nullptr, *tc = nullptr;   // tape: stage states (8) and currents (3), 4 per step
};

static void adj_free(AdjBufs &A)
{
    double *p[] = {A.ab, A.t8, A.xb, A.xbm, A.xbf, A.tt, A.eb, A.hb, A.jb, A.vb, A.t3, A.rb, A.xfin, A.acc,
                   A.ck, A.tx, A.tc};
    for (double *q : p)
        if (q) (void)hipFree(q);
    A = AdjBufs();
}

static dim3 adj_grid(const mhd_handle *h)
{
    return dim3((unsigned)blocks(h->NX, 64), (unsigned)blocks(h->g.yhi - h->g.ylo + 1, 4), (unsigned)h->NZ);
}

static int adj_alloc(mhd_handle *h, AdjBufs &A, int nck, int K)
{
    const size_t fb = (size_t)h->S * sizeof(double);
    const dim3 gr = adj_grid(h);
    A.nblk = (long)gr.x * gr.y * gr.z;
    struct Want {
        double **p;
        size_t bytes;
    } w[] = {{&A.ab, 8 * fb}, {&A.t8, 8 * fb}, {&A.xb, 8 * fb}, {&A.xbm, 8 * fb}, {&A.xbf, 8 * fb},
             {&A.tt, 8 * fb}, {&A.xfin, 8 * fb}, {&A.eb, 3 * fb}, {&A.hb, 3 * fb}, {&A.jb, 3 * fb},
             {&A.vb, 3 * fb}, {&A.t3, 3 * fb}, {&A.rb, 3 * fb}, {&A.acc, 2 * A.nblk * sizeof(double)},
             {&A.ck, (size_t)nck * 8 * fb}, {&A.tx, (size_t)K * 4 * 8 * fb}, {&A.tc, (size_t)K * 4 * 3 * fb}};
    for (auto &x : w) {
        if (x.bytes == 0) continue;
        if (hipMalloc(x.p, x.bytes) != hipSuccess) {
            set_err(h, "adjoint: device allocation failed");
            return MHD_ERR_NOMEM;
        }
        // checkpoints and tape are written before they are read
        if (x.p != &A.ck && x.p != &A.tx && x.p != &A.tc) HCHK(hipMemsetAsync(*x.p, 0, x.bytes, h->stream));
    }
    return 0;
}

// the two accumulated sums (adjoint of eta, objective), in a fixed order
static int adj_sums(mhd_handle *h, const AdjBufs &A, double *eta_bar, double *obj)
{
    std::vector<double> v(2 * A.nblk);
    HCHK(hipMemcpyAsync(v.data(), A.acc, v.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    double e = 0.0, o = 0.0;
    for (long b = 0; b < A.nblk; b++) { e += v[b]; o += v[A.nblk + b]; }
    if (eta_bar) *eta_bar = e;
    if (obj) *obj = o;
    return 0;
}

// transposed fill: z rules, then x rules (the y images are implicit)
static void adj_bc(mhd_handle *h, double *a, int nf, int type)
{
    const Geo &g = h->g;
    const long nr = g.yhi - g.ylo + 1;
    const long nz = (long)g.NX * nr * nf, nx = nr * g.NZ * nf, ny = (long)g.NX * g.NZ * nf;
    const int i0 = type == BC_TT ? 2 : 0;
    hipLaunchKernelGGL(k_adj_bc_z, dim3(blocks(nz, 256)), dim3(256), 0, h->stream, g, a, nf, i0);
    if (!g.periodic)
        hipLaunchKernelGGL(k_adj_bc_y, dim3(blocks(ny, 256)), dim3(256), 0, h->stream, g, a, nf, i0);
    hipLaunchKernelGGL(k_adj_bc_x, dim3(blocks(nx, 256)), dim3(256), 0, h->stream, g, a, nf, type == BC_STATE ? 0 : 1);
}

// kk transposed passes of a smoothing routine (each: fill^T, then sweep^T),
// ping-ponging a and tmp; on return a holds the result, tmp is free
template <int KIND>
static void adj_smooth(mhd_handle *h, double *&a, double *&tmp, int nf, int kk, int bctype)
{
    for (int p = 0; p < kk; p++) {
        adj_bc(h, a, nf, bctype);
        hipLaunchKernelGGL(k_adj_filter<KIND>, adj_grid(h), dim3(64, 4), 0, h->stream, h->g, a, tmp, nf);
        std::swap(a, tmp);
    }
}

// reverse mode of right() at state X with current C: ab = adjoint of xdif
// (destroyed), xb += (dright/dX)^T ab, rb -= part on xresd (if rb),
// acc[0] += part on eta.  init: the residual capture (eta = 0, no avrg).
static void adj_rhs(mhd_handle *h, AdjBufs &A, const double *X, const double *C, double *ab, double *xb,
                    double *rb, bool init, bool smooth)
{
    const mhd_config &c = h->cfg;
    const Geo &g = h->g;
    const double eta = init ? 0.0 : c.eta0;
    const dim3 grid = adj_grid(h), blk(64, 4);
    double *a = ab, *t8 = A.t8;
    if (smooth) adj_smooth<KIND_STATE>(h, a, t8, 8, 2, BC_STATE);   // avrg x2 (1669)
    adj_bc(h, a, 8, BC_STATE);                                         // bndry (1665)
    hipLaunchKernelGGL(k_adj_curle_p, grid, blk, 0, h->stream, g, a, X, C, c.gamma, eta, A.eb, A.vb, A.jb, xb,
                       A.acc);
    const double fmu = init ? 0.0 : c.fmu0, frho = init ? 0.0 : c.frho0, fpr = init ? 0.0 : c.fpr0;
    if (fmu != 0.0 || frho != 0.0 || fpr != 0.0)   // viscosity, diffusion (1585-1663)
        hipLaunchKernelGGL(k_adj_visc, grid, blk, 0, h->stream, g, a, X, fmu, frho, fpr, A.vb, xb);
    const bool side = c.hall && h->adj_overlap && !h->prof;
    if (c.hall) {
        HCHK_V(hipMemcpyAsync(A.hb, A.eb, 3 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        // the Hall chain reads only its copy of E's adjoint: beside the E
        // chain on the second stream (its temporary the free half of the
        // state smoothing's pair), joined before the flux adjoint, which
        // accumulates into the same J and x adjoints after it
        double *hb = A.hb, *ht = side ? t8 : A.t3;
        if (side) {
            HCHK_V(hipEventRecord(h->aux_ev, h->stream));
            HCHK_V(hipStreamWaitEvent(h->aux, h->aux_ev, 0));
            std::swap(h->stream, h->aux);
        }
        adj_smooth<KIND_EIGHTH>(h, hb, ht, 3, 5, BC_AUX);              // avrgh x5 (1424)
        adj_bc(h, hb, 3, BC_AUX);                                      // bndryh (1423)
        hipLaunchKernelGGL(k_adj_hall, grid, blk, 0, h->stream, g, hb, X, C, c.di, A.jb, xb);
        if (side) {
            std::swap(h->stream, h->aux);
            HCHK_V(hipEventRecord(h->aux_ev, h->aux));
        }
    }
    adj_smooth<KIND_RADIAL>(h, A.eb, A.t3, 3, 4, BC_AUX);              // avrgE x4 (1389)
    if (side) HCHK_V(hipStreamWaitEvent(h->stream, h->aux_ev, 0));
    hipLaunchKernelGGL(k_adj_ext_flux, grid, blk, 0, h->stream, g, a, A.eb, A.vb, X, C, eta, xb, A.jb, rb, A.acc);
    adj_smooth<KIND_RADIAL>(h, A.jb, A.t3, 3, 4, BC_AUX);              // avrgc x4 (2495)
    adj_bc(h, A.jb, 3, BC_AUX);                                        // bndryc (2493)
    hipLaunchKernelGGL(k_adj_curl, grid, blk, 0, h->stream, g, A.jb, xb);
}

// reverse mode of one step from tape slot `slot` (stage states and
// currents of that step): A.xb = adjoint of x_{n+1} on entry, of x_n on exit
static void adj_step(mhd_handle *h, AdjBufs &A, int slot, double dt)
{
    const long n8 = 8 * h->S, s8 = 8 * h->S, s3 = 3 * h->S;
    const double *TX = A.tx + (long)slot * 4 * s8, *TC = A.tc + (long)slot * 4 * s3;
    adj_smooth<KIND_STATE>(h, A.xb, A.t8, 8, 2, BC_TT);                // avrgtt x2 + bndrytt (1150)
    const dim3 eg(blocks(n8, 256)), eb(256);
    for (int s = 3; s >= 0; s--) {
        hipLaunchKernelGGL(k_adj_rk, eg, eb, 0, h->stream, n8, s, dt, A.xb, A.xbm, A.xbf, A.ab, A.tt);
        adj_rhs(h, A, TX + s * s8, TC + s * s3, A.ab, s == 0 ? A.xbf : A.tt, A.rb, false, true);
    }
    std::swap(A.xb, A.xbf);
}

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

int mhd_create(const mhd_config *cfg, const double *xx, const double *yy, const double *zz, const double *dx,
               const double *dy, const double *dz, mhd_handle **out)
{
    if (!cfg || !xx || !yy || !zz || !dx || !dy || !dz || !out) return MHD_ERR_ARG;
    *out = nullptr;
    if (cfg->mx < 8 || cfg->mz < 8 || cfg->my < 8) return MHD_ERR_ARG;
    // one field must stay below 2^29 points (32-bit byte offsets inside a field)
    if ((long)cfg->mx * cfg->my * cfg->mz >= (1L << 29)) return MHD_ERR_ARG;
    if (!(cfg->gamma > 1.0) || !(cfg->cfl > 0.0)) return MHD_ERR_ARG;
    if (cfg->periodic_y && cfg->my < 8) return MHD_ERR_ARG;
    {
        // the in-place viscosity/diffusion update reads values another tile
        // may own when a shifted last tile overlaps its neighbour in the
        // boundary rows the rules read; these sizes are refused with it
        bool vd = cfg->fmu0 != 0.0 || cfg->frho0 != 0.0 || cfg->fpr0 != 0.0;
        int nyt = cfg->periodic_y ? cfg->my - 4 : cfg->my;
        bool xbad = cfg->mx > TX && (cfg->mx % TX == 1 || cfg->mx % TX == 2);
        bool ybad = !cfg->periodic_y && nyt > TY && (nyt % TY == 1 || nyt % TY == 2);
        if (vd && (xbad || ybad)) return MHD_ERR_ARG;
        // magnetosphere mode: non-periodic y, no viscosity or diffusion
        bool mag = cfg->magnetosphere || cfg->inflow || cfg->radial_hall || cfg->partial_e;
        if (mag && (cfg->periodic_y || vd)) return MHD_ERR_ARG;
    }
    mhd_handle *h = new mhd_handle();
    h->cfg = *cfg;
    h->NX = cfg->mx; h->NY = cfg->my; h->NZ = cfg->mz;
    h->N = (long)h->NX * h->NY * h->NZ;
    // field stride: the volume rounded up to 64 points plus 576 (measured
    // slightly faster than an unpadded stride at 512x256x512)
    long pad = 576;
    if (const char *e = getenv("MHD_PAD")) pad = 64L * std::max(0, atoi(e));
    h->S = (h->N + 63) / 64 * 64 + pad;
    HCHK(hipGetDevice(&h->dev));
    {
        int r = scratch_setup(h, h->dev);
        if (r) { *out = h; return r; }
    }
    HCHK(hipStreamCreate(&h->stream));
    HCHK(hipStreamCreate(&h->side));
    HCHK(hipStreamCreate(&h->aux));
    HCHK(hipStreamCreate(&h->aux_side));
    HCHK(hipEventCreateWithFlags(&h->aux_ev, hipEventDisableTiming));
    HCHK(hipEventCreateWithFlags(&h->fork_ev, hipEventDisableTiming));
    HCHK(hipEventCreateWithFlags(&h->join_ev, hipEventDisableTiming));

    // ---- 1-D geometry on the host (ext-indexed) ----
    int NX = h->NX, NY = h->NY, NZ = h->NZ;
    std::vector<double> X(NX), Y(NY), Z(NZ), DXs(NX), DYs(NY), DZs(NZ);
    for (int i = 0; i < NX; i++) { X[i] = xx[i + 2]; DXs[i] = dx[i + 2]; }
    for (int j = 0; j < NY; j++) { Y[j] = yy[j]; DYs[j] = dy[j]; }
    for (int k = 0; k < NZ; k++) { Z[k] = zz[k + 2]; DZs[k] = dz[k + 2]; }
    // coefficient arrays: 8 (d1,d2) + 3 (spacings) per axis
    std::vector<double> cx(13 * NX, 0.0), cy(13 * NY, 0.0), cz(13 * NZ, 0.0);
    for (int i = 2; i <= NX - 3; i++) {
        coef_d1(X.data(), i, &cx[0 * NX + i], &cx[1 * NX + i], &cx[2 * NX + i], &cx[3 * NX + i]);
        coef_d2(X.data(), i, &cx[4 * NX + i], &cx[5 * NX + i], &cx[6 * NX + i], &cx[7 * NX + i]);
        cx[8 * NX + i] = X[i + 1] - X[i]; cx[9 * NX + i] = X[i] - X[i - 1]; cx[10 * NX + i] = X[i + 1] - X[i - 1];
        cx[11 * NX + i] = 2. * cx[8 * NX + i] / cx[10 * NX + i]; cx[12 * NX + i] = 2. * cx[9 * NX + i] / cx[10 * NX + i];
    }
    for (int j = 2; j <= NY - 3; j++) {
        coef_d1(Y.data(), j, &cy[0 * NY + j], &cy[1 * NY + j], &cy[2 * NY + j], &cy[3 * NY + j]);
        coef_d2(Y.data(), j, &cy[4 * NY + j], &cy[5 * NY + j], &cy[6 * NY + j], &cy[7 * NY + j]);
        cy[8 * NY + j] = Y[j + 1] - Y[j]; cy[9 * NY + j] = Y[j] - Y[j - 1]; cy[10 * NY + j] = Y[j + 1] - Y[j - 1];
        cy[11 * NY + j] = 2. * cy[8 * NY + j] / cy[10 * NY + j]; cy[12 * NY + j] = 2. * cy[9 * NY + j] / cy[10 * NY + j];
    }
    for (int k = 2; k <= NZ - 3; k++) {
        coef_d1(Z.data(), k, &cz[0 * NZ + k], &cz[1 * NZ + k], &cz[2 * NZ + k], &cz[3 * NZ + k]);
        coef_d2(Z.data(), k, &cz[4 * NZ + k], &cz[5 * NZ + k], &cz[6 * NZ + k], &cz[7 * NZ + k]);
        cz[8 * NZ + k] = Z[k + 1] - Z[k]; cz[9 * NZ + k] = Z[k] - Z[k - 1]; cz[10 * NZ + k] = Z[k + 1] - Z[k - 1];
        cz[11 * NZ + k] = 2. * cz[8 * NZ + k] / cz[10 * NZ + k]; cz[12 * NZ + k] = 2. * cz[9 * NZ + k] / cz[10 * NZ + k];
    }
    std::vector<double> cbx(NX);
    for (int i = 0; i < NX; i++) cbx[i] = (1 - 0.0) * (1.0 / 32.0) * (1 - tanh((X[i] + 3) / 8.));
    // energy weights (trapezoid over unique physical points)
    h->wx.assign(NX, 0.0); h->wz.assign(NZ, 0.0); h->wy.assign(NY, 0.0);
    for (int i = 0; i < NX; i++)
        h->wx[i] = i == 0 ? 0.5 * (X[1] - X[0]) : i == NX - 1 ? 0.5 * (X[i] - X[i - 1]) : 0.5 * (X[i + 1] - X[i - 1]);
    for (int k = 0; k < NZ; k++)
        h->wz[k] = k == 0 ? 0.5 * (Z[1] - Z[0]) : k == NZ - 1 ? 0.5 * (Z[k] - Z[k - 1]) : 0.5 * (Z[k + 1] - Z[k - 1]);
    for (int j = 0; j < NY; j++) {
        if (cfg->periodic_y) h->wy[j] = (j >= 2 && j <= NY - 3) ? Y[j + 1] - Y[j] : 0.0;
        else h->wy[j] = j == 0 ? 0.5 * (Y[1] - Y[0]) : j == NY - 1 ? 0.5 * (Y[j] - Y[j - 1]) : 0.5 * (Y[j + 1] - Y[j - 1]);
    }
    // setdt tables: 1/|d| and (eta + fmu) / d / 2 per index
    std::vector<double> RDX(NX), RDY(NY), RDZ(NZ), HDX(NX), HDY(NY), HDZ(NZ);
    const double em = cfg->eta0 + cfg->fmu0;
    for (int i = 0; i < NX; i++) { RDX[i] = 1.0 / std::fabs(DXs[i]); HDX[i] = em / DXs[i] / 2.; }
    for (int j = 0; j < NY; j++) { RDY[j] = 1.0 / std::fabs(DYs[j]); HDY[j] = em / DYs[j] / 2.; }
    for (int k = 0; k < NZ; k++) { RDZ[k] = 1.0 / std::fabs(DZs[k]); HDZ[k] = em / DZs[k] / 2.; }
    size_t ngeo = cx.size() + cy.size() + cz.size() + 2 * (NX + NY + NZ) + NX + (NX + NY + NZ) + 2 * (NX + NY + NZ);
    HCHK(hipMalloc(&h->geo, ngeo * sizeof(double)));
    std::vector<double> hg;
    hg.reserve(ngeo);
    auto push = [&](const std::vector<double> &v) { size_t o = hg.size(); hg.insert(hg.end(), v.begin(), v.end()); return o; };
    size_t ocx = push(cx), ocy = push(cy), ocz = push(cz);
    size_t oX = push(X), oY = push(Y), oZ = push(Z), oDX = push(DXs), oDY = push(DYs), oDZ = push(DZs);
    size_t ocb = push(cbx);
    size_t owx = push(h->wx), owy = push(h->wy), owz = push(h->wz);
    size_t orx = push(RDX), ory = push(RDY), orz = push(RDZ), ohx = push(HDX), ohy = push(HDY), ohz = push(HDZ);
    HCHK(hipMemcpy(h->geo, hg.data(), hg.size() * sizeof(double), hipMemcpyHostToDevice));
    Geo &g = h->g;
    g.NX = NX; g.NY = NY; g.NZ = NZ; g.sy = NX; g.sz = (long)NX * NY; g.S = h->S; g.periodic = cfg->periodic_y;
    g.ylo = cfg->periodic_y ? 2 : 0;
    g.yhi = cfg->periodic_y ? NY - 3 : NY - 1;
    g.nyu = NY - 4;
    const double *G0 = h->geo;
    g.x1a = G0 + ocx; g.x1b = g.x1a + NX; g.x1c = g.x1b + NX; g.x1d = g.x1c + NX;
    g.x2a = g.x1d + NX; g.x2b = g.x2a + NX; g.x2c = g.x2b + NX; g.x2d = g.x2c + NX;
    g.xwp = g.x2d + 4 * NX; g.xwm = g.xwp + NX;
    g.y1a = G0 + ocy; g.y1b = g.y1a + NY; g.y1c = g.y1b + NY; g.y1d = g.y1c + NY;
    g.y2a = g.y1d + NY; g.y2b = g.y2a + NY; g.y2c = g.y2b + NY; g.y2d = g.y2c + NY;
    g.ywp = g.y2d + 4 * NY; g.ywm = g.ywp + NY;
    g.z1a = G0 + ocz; g.z1b = g.z1a + NZ; g.z1c = g.z1b + NZ; g.z1d = g.z1c + NZ;
    g.z2a = g.z1d + NZ; g.z2b = g.z2a + NZ; g.z2c = g.z2b + NZ; g.z2d = g.z2c + NZ;
    g.zwp = g.z2d + 4 * NZ; g.zwm = g.zwp + NZ;
    g.xc = G0 + oX; g.yc = G0 + oY; g.zc = G0 + oZ;
    g.dxs = G0 + oDX; g.dys = G0 + oDY; g.dzs = G0 + oDZ;
    g.rdx = G0 + orx; g.rdy = G0 + ory; g.rdz = G0 + orz;
    g.hdx = G0 + ohx; g.hdy = G0 + ohy; g.hdz = G0 + ohz;
    g.cbx = G0 + ocb;
    h->dwx = h->geo + owx; h->dwy = h->geo + owy; h->dwz = h->geo + owz;
    coef_side(X.data(), 0, +1, &g.ax3, &g.bx3, &g.cx3);
    coef_side(X.data(), 1, +1, &g.ax4, &g.bx4, &g.cx4);
    g.dx3 = g.ax3 + g.bx3 + g.cx3; g.rx3 = 1.0 / g.dx3;
    g.dx4 = g.ax4 + g.bx4 + g.cx4; g.rx4 = 1.0 / g.dx4;
    coef_side(X.data(), NX - 2, -1, &g.axmL1, &g.bxmL1, &g.cxmL1);
    coef_side(X.data(), NX - 1, -1, &g.axmL, &g.bxmL, &g.cxmL);

    // ---- volumes ----
    size_t fb = (size_t)h->S * sizeof(double);
    double **p8[] = {&h->x, &h->xfold, &h->xm, &h->xdA, &h->xdB};
    for (auto p : p8) {
        if (hipMalloc(p, 8 * fb) != hipSuccess) { set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM; }
        HCHK(hipMemset(*p, 0, 8 * fb));
    }
    if (hipMalloc(&h->eq, 5 * fb) != hipSuccess) { set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM; }
    HCHK(hipMemset(h->eq, 0, 5 * fb));
    double **p3[] = {&h->xresd, &h->cA, &h->cB, &h->eA, &h->eB, &h->hA, &h->hB};
    for (auto p : p3) {
        if (hipMalloc(p, 3 * fb) != hipSuccess) { set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM; }
        HCHK(hipMemset(*p, 0, 3 * fb));
    }
    if (hipMalloc(&h->crad, fb) != hipSuccess) { set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM; }
    g.crad = h->crad;
    hipLaunchKernelGGL(k_crad, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, g, h->crad);
    HCHK(hipGetLastError());
    h->nred = 1024;
    long nrows = (long)NY * NZ * 4;
    HCHK(hipMalloc(&h->red, std::max<long>(4096, nrows) * sizeof(double)));
    HCHK(hipMalloc(&h->dtv, 8 * sizeof(double)));
    HCHK(hipMemset(h->dtv, 0, 8 * sizeof(double)));
    h->mag = cfg->magnetosphere || cfg->inflow || cfg->radial_hall || cfg->partial_e;
    if (h->mag) {
        // |r| (t3dsdsc.f:773-774) and the radial Hall coefficient (1398-1399)
        // per point, the inflow profiles per plane / row (1714-1768), the
        // face data (zero until mhd_set_inflow_face), v = m / rho
        std::vector<double> rr(h->S, 0.0), ch;
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++)
                for (int i = 0; i < NX; i++)
                    rr[IDX(g, i, j, k)] = sqrt((X[i] - 0.0) * (X[i] - 0.0) + Y[j] * Y[j] + Z[k] * Z[k]);
        if (hipMalloc(&h->rr, fb) != hipSuccess || hipMalloc(&h->vel, 3 * fb) != hipSuccess) {
            set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM;
        }
        HCHK(hipMemcpy(h->rr, rr.data(), fb, hipMemcpyHostToDevice));
        HCHK(hipMemset(h->vel, 0, 3 * fb));
        if (cfg->magnetosphere) {
            // the masked points of each row (j, k) as one column range: rr is
            // monotone in |x| along a row; the producers apply the mask when
            // every masked point is interior and each row's set is one range
            std::vector<int> mr(2 * (size_t)NY * NZ, 0);
            bool ok = true;
            for (int k = 0; k < NZ && ok; k++)
                for (int j = 0; j < NY && ok; j++) {
                    int lo = -1, hi = -1;
                    for (int i = 0; i < NX; i++) {
                        if (!(rr[IDX(g, i, j, k)] <= cfg->rcr)) continue;
                        if (lo < 0) lo = i;
                        else if (hi != i) ok = false;
                        hi = i + 1;
                    }
                    if (lo < 0) continue;
                    if (lo < 2 || hi > NX - 2 || j < 2 || j > NY - 3 || k < 2 || k > NZ - 3) ok = false;
                    mr[2 * ((size_t)k * NY + j)] = lo;
                    mr[2 * ((size_t)k * NY + j) + 1] = hi;
                }
            if (const char *e = getenv("MHD_MAGMASK")) ok = ok && atoi(e) != 0;
            if (ok) {
                HCHK(hipMalloc(&h->mrow, mr.size() * sizeof(int)));
                HCHK(hipMemcpy(h->mrow, mr.data(), mr.size() * sizeof(int), hipMemcpyHostToDevice));
            }
        }
        if (const char *e = getenv("MHD_FILLSPLIT")) h->fill_split = atoi(e);
        if (cfg->radial_hall) {
            ch.assign(h->S, 0.0);
            for (size_t q = 0; q < ch.size(); q++)
                ch[q] = 0.5 * cfg->di * (tanh((rr[q] - (double)1.1f) / 8.) - tanh((rr[q] - 2.) / 10.));
            if (hipMalloc(&h->chtab, fb) != hipSuccess) { set_err(h, "device allocation failed"); *out = h; return MHD_ERR_NOMEM; }
            HCHK(hipMemcpy(h->chtab, ch.data(), fb, hipMemcpyHostToDevice));
        }
        std::vector<double> mg(2 * (size_t)NZ + 2 * (size_t)NY + 3 * (size_t)NZ * NY, 0.0);
        const double c06 = (double)0.6f, c04 = (double)0.4f, c02 = (double)0.2f, c03 = (double)0.3f,
                     c15 = (double)0.15f;
        for (int k = 0; k < NZ; k++) {
            mg[k] = tanh((Z[k] + c06) / c15) - tanh((Z[k] - c06) / c15);
            mg[NZ + NY + k] = tanh((Z[k] + c02) / c15) - tanh((Z[k] - c02) / c15);
        }
        for (int j = 0; j < NY; j++) {
            mg[NZ + j] = tanh((Y[j] + c04) / c15) - tanh((Y[j] - c04) / c15);
            mg[2 * NZ + NY + j] = tanh((Y[j] + c03) / c15) - tanh((Y[j] - c03) / c15);
        }
        HCHK(hipMalloc(&h->mgeo, mg.size() * sizeof(double)));
        HCHK(hipMemcpy(h->mgeo, mg.data(), mg.size() * sizeof(double), hipMemcpyHostToDevice));
        h->zd = h->mgeo; h->yd = h->mgeo + NZ; h->zd2 = h->mgeo + NZ + NY; h->yd2 = h->mgeo + 2 * NZ + NY;
        h->face = h->mgeo + 2 * NZ + 2 * NY;
    }
    h->xd = h->xdA;
    h->cur = h->cA;
    // z planes per chunk of the tile kernels: 16 on volumes of 2^25 points
    // and more; below, the longest of 32, 16, 8, 4 that still gives 400
    // work-groups (measured at 64x32x64 to 512x256x512: short chunks pay
    // for their z halo, too few work-groups leave CUs idle)
    {
        const long tiles = (long)blocks(NX, TX) * blocks(h->g.yhi - h->g.ylo + 1, TY);
        h->kc = 32;
        while (h->kc > 4 && tiles * blocks(NZ - 4, h->kc) < 400) h->kc /= 2;
        if (h->S >= (1L << 25)) h->kc = 16;
    }
    if (const char *e = getenv("MHD_KC")) h->kc = std::max(1, atoi(e));
    if (const char *e = getenv("MHD_FUSE_RK")) h->fuse_rk = atoi(e);
    if (const char *e = getenv("MHD_ZS")) h->zs = atoi(e);
    // windowed right-part-2 kernels pay off on large volumes (measured at
    // 512x256x512 against 256x128x256)
    h->rb2 = h->S >= (1L << 25) ? 1 : 0;
    // interior planes per chunk of the smoothing kernels, the chunks of a
    // column of equal length: about 72 planes on volumes of 2^25 points and
    // more (73 at 512x256x512 measured 0.7% faster than 64 and 85);
    // below, about the longest of 32, 16, 8 that gives 600 work-groups to a
    // 3-field two-pass launch (same measurements)
    {
        const long wins = (long)blocks(NX, TX - 4) * blocks(h->g.yhi - h->g.ylo + 1, ZB - 4) * 3;
        int z = 32;
        while (z > 8 && wins * blocks(NZ - 4, z) < 600) z /= 2;
        int nch = blocks(NZ - 4, z);
        if (h->S >= (1L << 25)) nch = std::max(1, (int)std::lround((NZ - 4) / 72.0));
        h->zkc = std::min(ZKC_MAX, blocks(NZ - 4, nch));
    }
    if (const char *e = getenv("MHD_ZKC")) h->zkc = std::min(ZKC_MAX, std::max(4, atoi(e)));
    if (const char *e = getenv("MHD_RB2")) h->rb2 = atoi(e);
    if (const char *e = getenv("MHD_ZORD")) h->zord = std::min(1, std::max(0, atoi(e)));
    if (const char *e = getenv("MHD_ZS3")) h->zs3 = atoi(e);
    if (const char *e = getenv("MHD_FACE12")) h->face12 = atoi(e);
    // side-stream launches beside the main stream's (face windows beside the
    // face-free ones, the parts of right part 2, the first Hall passes beside
    // the smoothing of E): 1.4% faster at 512x256x512 and 256x128x256, 4.7%
    // slower at 128x64x128, where the launches are short
    h->overlap = h->S >= (1L << 22) ? 1 : 0;
    if (const char *e = getenv("MHD_OVERLAP")) h->overlap = atoi(e);
    if (const char *e = getenv("MHD_ADJOVL")) h->adj_overlap = atoi(e);
    for (int c = 0; c < 2; c++) {
        int r = zs_tables(h, c ? 8 : 3, 2, h->zt[c]);
        if (r) { *out = h; return r; }
    }
    {
        int r = zs_tables(h, 3, 3, h->zt3);
        if (r) { *out = h; return r; }
    }
    h->pt.resize(PT_N);
    for (int t = 0; t < PT_N; t++) h->pt[t] = Prof{PT_NAMES[t], 0.0, 0};
    *out = h;
    return MHD_OK;
}

int mhd_set_equilibrium(mhd_handle *h, const double *rhoi, const double *bxi, const double *byi, const double *bzi,
                        const double *pri)
{
    if (!h || !rhoi || !bxi || !byi || !bzi || !pri) return MHD_ERR_ARG;
    const double *src[5] = {rhoi, bxi, byi, bzi, pri};
    for (int m = 0; m < 5; m++) {
        int r = upload_fields(h, h->eq + m * h->S, src[m], 1);
        if (r) return r;
    }
    h->have_eq = true;
    return MHD_OK;
}

int mhd_set_state(mhd_handle *h, const double *x)
{
    if (!h || !x) return MHD_ERR_ARG;
    return upload_fields(h, h->x, x, 8);
}

int mhd_init_harris(mhd_handle *h, const mhd_harris *p)
{
    if (!h || !p) return MHD_ERR_ARG;
    if (h->mag) { set_err(h, "mhd_init_harris builds the Harris sheet, not the magnetosphere"); return MHD_ERR_ARG; }
    if (!(p->lam != 0.0) || !(p->temp != 0.0) || !(p->width != 0.0)) {
        set_err(h, "mhd_init_harris: lam, temp and width must be nonzero");
        return MHD_ERR_ARG;
    }
    const int nyu = h->g.periodic ? h->NY - 4 : h->NY;
    const long n = (long)h->NX * h->NY * h->NZ;
    hipLaunchKernelGGL(k_harris, dim3((unsigned)blocks(n, 256)), dim3(256), 0, h->stream, h->g, *p, nyu, h->x, h->eq);
    HCHK(hipGetLastError());
    HCHK(hipStreamSynchronize(h->stream));
    h->have_eq = true;
    return MHD_OK;
}

int mhd_get_state(mhd_handle *h, double *x)
{
    if (!h || !x) return MHD_ERR_ARG;
    return download_fields(h, x, h->x, 8);
}

int mhd_init_residual(mhd_handle *h)
{
    if (!h) return MHD_ERR_ARG;
    if (!h->have_eq) { set_err(h, "equilibrium not set"); return MHD_ERR_ARG; }
    HCHK(hipMemsetAsync(h->xresd, 0, 3 * h->S * sizeof(double), h->stream));
    if (h->mag) {
        // xfold = x, right(0d0, 1.d-9, 1) (t3dsdsc.f:821-829)
        HCHK(hipMemcpyAsync(h->xfold, h->x, 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        mag_rhs(h, 0.0, 1.e-9, 1, false, true);
    } else {
        rhs(h, h->x, false, true, -1);
    }
    HCHK(hipMemcpyAsync(h->xresd, h->xd + h->S, 3 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    return MHD_OK;
}

int mhd_rhs(mhd_handle *h, double t)
{
    return mhd_rhs_stage(h, t, 1.e-9, 1);
}

int mhd_rhs_stage(mhd_handle *h, double t, double ddt, int km)
{
    if (!h || (km != 1 && km != 2)) return MHD_ERR_ARG;
    if (h->mag) mag_rhs(h, t, ddt, km, t != 0.0, false);
    else rhs(h, h->x, t != 0.0, false, -1);
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    return MHD_OK;
}

int mhd_set_inflow_face(mhd_handle *h, const double *bx, const double *by, const double *bz)
{
    if (!h || !bx || !by || !bz) return MHD_ERR_ARG;
    if (!h->mag) { set_err(h, "inflow face data need the magnetosphere mode"); return MHD_ERR_ARG; }
    // host (my, mz+4) Fortran order = C [jz-1][jy-1]; device [k * NY + j], jz = k + 3
    const int NY = h->NY, NZ = h->NZ;
    std::vector<double> f(3 * (size_t)NZ * NY);
    const double *src[3] = {bx, by, bz};
    for (int c = 0; c < 3; c++)
        for (int k = 0; k < NZ; k++)
            for (int j = 0; j < NY; j++) f[((size_t)c * NZ + k) * NY + j] = src[c][(size_t)(k + 2) * NY + j];
    HCHK(hipMemcpy(const_cast<double *>(h->face), f.data(), f.size() * sizeof(double), hipMemcpyHostToDevice));
    return MHD_OK;
}

int mhd_get_xdif(mhd_handle *h, double *xdif)
{
    if (!h || !xdif) return MHD_ERR_ARG;
    if (!h->xd) { set_err(h, "no xdif: the last call was a step (its final sweep updates the state in place)"); return MHD_ERR_ARG; }
    return download_fields(h, xdif, h->xd, 8);
}

int mhd_get_residual(mhd_handle *h, double *r)
{
    if (!h || !r) return MHD_ERR_ARG;
    return download_fields(h, r, h->xresd, 3);
}

int mhd_set_residual(mhd_handle *h, const double *r)
{
    if (!h || !r) return MHD_ERR_ARG;
    return upload_fields(h, h->xresd, r, 3);
}

int mhd_compute_dt(mhd_handle *h, double *dt)
{
    if (!h || !dt) return MHD_ERR_ARG;
    hipLaunchKernelGGL(k_disarm, dim3(1), dim3(1), 0, h->stream, h->dtv);
    dt_kernels(h, 0);
    double v[8];
    HCHK(hipMemcpyAsync(v, h->dtv, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    *dt = v[0];
    return MHD_OK;
}

int mhd_compute_dt_noalloc(mhd_handle *h, double *dt)
{
    if (!h || !dt) return MHD_ERR_ARG;
    double *part = (h->dev >= 0 && h->dev < MHD_MAX_DEVICES) ? g_scratch_ptr[h->dev] : nullptr;
    if (!part) { set_err(h, "scratch array not set up on the handle's device"); return MHD_ERR_HIP; }
    if (h->nred > MHD_SCRATCH_DOUBLES) {
        set_err(h, "setdt needs " + std::to_string(h->nred) + " partial minima; the scratch array holds " +
                       std::to_string(MHD_SCRATCH_DOUBLES));
        return MHD_ERR_ARG;
    }
    hipLaunchKernelGGL(k_disarm, dim3(1), dim3(1), 0, h->stream, h->dtv);
    dt_kernels(h, 0, part);
    double v[8];
    HCHK(hipMemcpyAsync(v, h->dtv, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    *dt = v[0];
    return MHD_OK;
}

int mhd_step(mhd_handle *h, double time, double dt)
{
    (void)time;   // no time-dependent boundary data in this configuration
    if (!h) return MHD_ERR_ARG;
    if (!h->have_eq) { set_err(h, "equilibrium not set"); return MHD_ERR_ARG; }
    if (!(dt > 0.0)) { set_err(h, "dt must be positive"); return MHD_ERR_ARG; }
    hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, h->stream, h->dtv, dt);
    if (h->mag) mag_step(h, time, dt);
    else step_body(h);
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    if (h->prof) return prof_collect(h);
    return MHD_OK;
}

// magnetosphere mode: the drive's time factors are evaluated on the host, so
// the loop reads each step's dt; it stops before the step whose setdt
// collapses (t3dsdsc.f:142-154)
static int mag_advance(mhd_handle *h, int nsteps, double dt0, double *time, double *dt_last, double *elapsed_ms)
{
    hipEvent_t e0, e1;
    hipLaunchKernelGGL(k_disarm, dim3(1), dim3(1), 0, h->stream, h->dtv);
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    HCHK(hipEventRecord(e0, h->stream));
    double t = *time, dl = 0.0;
    int rc = MHD_OK;
    for (int s = 0; s < nsteps; s++) {
        dt_kernels(h, 0);
        double v[8];
        HCHK(hipMemcpyAsync(v, h->dtv, 8 * sizeof(double), hipMemcpyDeviceToHost, h->stream));
        HCHK(hipStreamSynchronize(h->stream));
        const double dt = v[0];
        if (dt0 > 0.0 && dt < 0.1 * dt0) { rc = MHD_ERR_DT_COLLAPSE; dl = dt; break; }
        mag_step(h, t, dt);
        t += dt;
        dl = dt;
    }
    HCHK(hipEventRecord(e1, h->stream));
    HCHK(hipEventSynchronize(e1));
    float ms = 0;
    HCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    HCHK(hipGetLastError());
    if (h->prof) { int r = prof_collect(h); if (r) return r; }
    *time = t;
    if (dt_last) *dt_last = dl;
    if (elapsed_ms) *elapsed_ms = ms;
    if (rc) set_err(h, "dt collapsed below 0.1*dt0");
    return rc;
}

int mhd_advance(mhd_handle *h, int nsteps, double dt0, double *time, double *dt_last, double *elapsed_ms)
{
    if (!h || nsteps < 0 || !time) return MHD_ERR_ARG;
    if (!h->have_eq) { set_err(h, "equilibrium not set"); return MHD_ERR_ARG; }
    if (h->mag) return mag_advance(h, nsteps, dt0, time, dt_last, elapsed_ms);
    double init[8] = {0, *time, dt0, 0, 0, 0, 0, 0};
    HCHK(hipMemcpyAsync(h->dtv, init, 8 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipEvent_t e0, e1;
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    HCHK(hipEventRecord(e0, h->stream));
    for (int s = 0; s < nsteps; s++) {
        dt_kernels(h, s > 0);
        step_body(h);
        restore_after_collapse(h);
    }
    HCHK(hipEventRecord(e1, h->stream));
    HCHK(hipEventSynchronize(e1));
    float ms = 0;
    HCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    double v[8];
    HCHK(hipMemcpy(v, h->dtv, 8 * sizeof(double), hipMemcpyDeviceToHost));
    HCHK(hipGetLastError());
    if (h->prof) { int r = prof_collect(h); if (r) return r; }
    const bool collapsed = v[3] != 0.0;
    *time = v[1] + (nsteps > 0 && !collapsed ? v[4] : 0.0);
    if (dt_last) *dt_last = collapsed ? v[5] : v[4];
    if (elapsed_ms) *elapsed_ms = ms;
    if (collapsed) { set_err(h, "dt collapsed below 0.1*dt0"); return MHD_ERR_DT_COLLAPSE; }
    return MHD_OK;
}

int mhd_get_current(mhd_handle *h, double *cur)
{
    if (!h || !cur) return MHD_ERR_ARG;
    any_current(h, h->x);
    return download_fields(h, cur, h->cur, 3);
}

int mhd_adjoint_run(mhd_handle *h, int nsteps, const double *dts, int objective, int every, int flags,
                    double *value, double *grad_x, double *grad_eta, double *ms)
{
    if (!h || nsteps < 1 || !dts || objective < 0 || objective > 1) return MHD_ERR_ARG;
    if (!h->have_eq) { set_err(h, "equilibrium not set"); return MHD_ERR_ARG; }
    if (h->mag) { set_err(h, "the adjoint covers the Harris configuration, not the magnetosphere mode"); return MHD_ERR_ARG; }
    const mhd_config &c = h->cfg;
    for (int n = 0; n < nsteps; n++)
        if (!(dts[n] > 0.0)) { set_err(h, "adjoint: dt must be positive"); return MHD_ERR_ARG; }
    // every < 0: the forward pass records every step (no recomputation)
    const bool store_all = every < 0;
    const int K = store_all ? nsteps
                            : every > 0 ? std::min(every, nsteps) : (int)std::ceil(std::sqrt((double)nsteps));
    const int nck = (nsteps + K - 1) / K;
    // the two planes nearest z = 0 (the reconnection line of the sheet)
    int k0 = -1;
    {
        std::vector<double> Z(h->NZ);
        HCHK(hipMemcpy(Z.data(), h->g.zc, h->NZ * sizeof(double), hipMemcpyDeviceToHost));
        for (int k = 0; k + 1 < h->NZ; k++)
            if (Z[k] < 0.0 && Z[k + 1] >= 0.0) k0 = k;
        if (objective == 1 && k0 < 0) { set_err(h, "adjoint: z = 0 is not inside the mesh"); return MHD_ERR_ARG; }
    }
    // buffers, events and the recording hook are released on every return
    struct Scope {
        mhd_handle *h;
        AdjBufs A;
        hipEvent_t e[3] = {nullptr, nullptr, nullptr};
        ~Scope()
        {
            h->rec_x = h->rec_c = nullptr;
            if (h->stream) (void)hipStreamSynchronize(h->stream);
            adj_free(A);
            for (auto x : e)
                if (x) (void)hipEventDestroy(x);
        }
    } sc{h};
    AdjBufs &A = sc.A;
    hipEvent_t *e = sc.e;
    int r = adj_alloc(h, A, nck, K);
    if (r) return r;
    const size_t b8 = 8 * h->S * sizeof(double);
    for (int i = 0; i < 3; i++) HCHK(hipEventCreate(&e[i]));
    HCHK(hipEventRecord(e[0], h->stream));
    // forward pass, keeping every K-th state
    for (int n = 0; n < nsteps; n++) {
        if (n % K == 0)
            HCHK(hipMemcpyAsync(A.ck + (long)(n / K) * 8 * h->S, h->x, b8, hipMemcpyDeviceToDevice, h->stream));
        if (store_all) {
            h->rec_x = A.tx + (long)n * 4 * 8 * h->S;
            h->rec_c = A.tc + (long)n * 4 * 3 * h->S;
            h->rec_n = 0;
        }
        hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, h->stream, h->dtv, dts[n]);
        step_body(h);
    }
    h->rec_x = h->rec_c = nullptr;
    HCHK(hipEventRecord(e[1], h->stream));
    HCHK(hipMemcpyAsync(A.xfin, h->x, b8, hipMemcpyDeviceToDevice, h->stream));
    hipLaunchKernelGGL(k_adj_objective, adj_grid(h), dim3(64, 4), 0, h->stream, h->g, h->x, objective, k0,
                       h->dwx, h->dwy, h->dwz, A.xb, A.acc + A.nblk);
    // reverse sweep, one checkpoint segment at a time: replay the segment
    // recording stage states and currents, then run its steps backwards
    for (int ci = nck - 1; ci >= 0; ci--) {
        const int n0 = ci * K, n1 = std::min(nsteps, n0 + K);
        HCHK(hipMemcpyAsync(h->x, A.ck + (long)ci * 8 * h->S, b8, hipMemcpyDeviceToDevice, h->stream));
        for (int n = n0; n < n1 && !store_all; n++) {
            h->rec_x = A.tx + (long)(n - n0) * 4 * 8 * h->S;
            h->rec_c = A.tc + (long)(n - n0) * 4 * 3 * h->S;
            h->rec_n = 0;
            hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, h->stream, h->dtv, dts[n]);
            step_body(h);
        }
        h->rec_x = h->rec_c = nullptr;
        for (int n = n1 - 1; n >= n0; n--) adj_step(h, A, n - n0, dts[n]);
    }
    // the residual captured from x_0 (initia, t3dsdsc.f:816-834)
    if (flags & MHD_ADJ_RESIDUAL) {
        HCHK(hipMemcpyAsync(h->x, A.ck, b8, hipMemcpyDeviceToDevice, h->stream));
        current(h, h->x);
        HCHK(hipMemsetAsync(A.ab, 0, b8, h->stream));
        HCHK(hipMemcpyAsync(A.ab + h->S, A.rb, 3 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        adj_rhs(h, A, h->x, h->cur, A.ab, A.xb, nullptr, true, false);
    }
    HCHK(hipEventRecord(e[2], h->stream));
    HCHK(hipMemcpyAsync(h->x, A.xfin, b8, hipMemcpyDeviceToDevice, h->stream));
    r = adj_sums(h, A, grad_eta, value);
    if (r) return r;
    float f0 = 0, f1 = 0;
    HCHK(hipEventElapsedTime(&f0, e[0], e[1]));
    HCHK(hipEventElapsedTime(&f1, e[1], e[2]));
    if (ms) { ms[0] = f0; ms[1] = f1; }
    return grad_x ? download_fields(h, grad_x, A.xb, 8, false) : 0;
}

int mhd_adjoint_rhs(mhd_handle *h, const double *xdif_bar, double *x_bar, double *xresd_bar, double *eta_bar,
                    int init, int smooth)
{
    if (!h || !xdif_bar || !x_bar) return MHD_ERR_ARG;
    if (h->mag) { set_err(h, "the adjoint covers the Harris configuration, not the magnetosphere mode"); return MHD_ERR_ARG; }
    AdjBufs A;
    int r = adj_alloc(h, A, 0, 0);
    if (!r) r = upload_fields(h, A.ab, xdif_bar, 8);
    if (!r) {
        current(h, h->x);
        adj_rhs(h, A, h->x, h->cur, A.ab, A.xb, A.rb, init != 0, smooth != 0);
        r = adj_sums(h, A, eta_bar, nullptr);
    }
    if (!r) r = download_fields(h, x_bar, A.xb, 8, false);
    if (!r && xresd_bar) r = download_fields(h, xresd_bar, A.rb, 3, false);
    adj_free(A);
    return r;
}

int mhd_get_divb(mhd_handle *h, double *out)
{
    if (!h || !out) return MHD_ERR_ARG;
    double *dv = h->eB;   // scratch
    hipLaunchKernelGGL(k_divb, int_grid(h, BLK), BLK, 0, h->stream, h->g, h->x, dv);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NY * h->NZ, 256)), dim3(256), 0, h->stream, h->g, dv, 0);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NX * h->NZ, 256)), dim3(256), 0, h->stream, h->g, dv, 1);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NX * h->NY, 256)), dim3(256), 0, h->stream, h->g, dv, 2);
    return download_fields(h, out, dv, 1);
}

int mhd_diagnostics(mhd_handle *h, mhd_diag *d)
{
    if (!h || !d) return MHD_ERR_ARG;
    // diagn: current then extrema on jy = my/2+1 (j = my/2)
    any_current(h, h->x);
    int nb = 64;
    hipLaunchKernelGGL(k_plane_extrema, dim3(nb, 3), dim3(256), 0, h->stream, h->g, h->cur, h->cfg.my / 2, h->red);
    std::vector<double> ex(3 * nb * 2);
    HCHK(hipMemcpyAsync(ex.data(), h->red, ex.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    double vals[6];
    for (int m = 0; m < 3; m++) {
        double mx = 0.0, mn = 0.0;   // the reference's range includes two zero ghost columns
        for (int b = 0; b < nb; b++) {
            mx = std::max(mx, ex[(m * nb + b) * 2]);
            mn = std::min(mn, ex[(m * nb + b) * 2 + 1]);
        }
        vals[2 * m] = mx; vals[2 * m + 1] = mn;
    }
    d->cxmax = vals[0]; d->cxmin = vals[1]; d->cymax = vals[2]; d->cymin = vals[3];
    d->czmax = vals[4]; d->czmin = vals[5];
    // energy
    int j0 = h->cfg.periodic_y ? 2 : 0, nyr = h->cfg.periodic_y ? h->NY - 4 : h->NY;
    Eq sub = eq_state(h, true);
    hipLaunchKernelGGL(k_energy_rows, dim3(nyr, h->NZ), dim3(64), 0, h->stream, h->g, h->x, sub, h->dwx, h->dwy,
                       h->dwz, j0, nyr, h->cfg.gamma - 1.0, h->red);
    std::vector<double> rows((size_t)nyr * h->NZ * 4);
    HCHK(hipMemcpyAsync(rows.data(), h->red, rows.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    double e[4] = {0, 0, 0, 0};
    for (size_t r = 0; r < rows.size() / 4; r++)
        for (int t = 0; t < 4; t++) e[t] += rows[r * 4 + t];
    d->e_kin = e[0]; d->e_mag = e[1]; d->e_th = e[2]; d->e_mag_pert = e[3];
    // div B over the unique physical points
    double *dv = h->eB;
    hipLaunchKernelGGL(k_divb, int_grid(h, BLK), BLK, 0, h->stream, h->g, h->x, dv);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NY * h->NZ, 256)), dim3(256), 0, h->stream, h->g, dv, 0);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NX * h->NZ, 256)), dim3(256), 0, h->stream, h->g, dv, 1);
    hipLaunchKernelGGL(k_divb_bc, dim3(blocks((long)h->NX * h->NY, 256)), dim3(256), 0, h->stream, h->g, dv, 2);
    long plane = (long)h->NX * h->NY;
    double mx = 0, sq = 0;
    long cnt = 0;
    std::vector<double> part(2 * 64 * 8);
    // reduce over y rows j0..j0+nyr-1 of every z plane, plane by plane
    for (int k = 0; k < h->NZ; k++) {
        const double *base = dv + k * plane + (long)j0 * h->NX;
        long n = (long)nyr * h->NX;
        hipLaunchKernelGGL(k_absmax_sq, dim3(64), dim3(256), 0, h->stream, n, base, h->red + 2 * 64 * (k % 8));
        if (k % 8 == 7 || k == h->NZ - 1) {
            int nk = k % 8 + 1;
            HCHK(hipMemcpyAsync(part.data(), h->red, 2 * 64 * nk * sizeof(double), hipMemcpyDeviceToHost, h->stream));
            HCHK(hipStreamSynchronize(h->stream));
            for (int b = 0; b < 64 * nk; b++) { mx = std::max(mx, part[2 * b]); sq += part[2 * b + 1]; }
        }
        cnt += n;
    }
    d->divb_max = mx;
    d->divb_rms = std::sqrt(sq / (double)cnt);
    HCHK(hipGetLastError());
    if (std::fabs(d->cymax) >= 1.e7) { set_err(h, "current blow-up |Jy| >= 1e7"); return MHD_ERR_BLOWUP; }
    return MHD_OK;
}

int mhd_set_profiling(mhd_handle *h, int on)
{
    if (!h) return MHD_ERR_ARG;
    h->prof = on != 0;
    for (auto &p : h->pt) { p.ms = 0; p.calls = 0; }
    return MHD_OK;
}

int mhd_kernel_times(mhd_handle *h, int max, const char **names, double *ms, int *calls)
{
    if (!h) return MHD_ERR_ARG;
    int n = std::min<int>(max, (int)h->pt.size());
    for (int t = 0; t < n; t++) {
        if (names) names[t] = h->pt[t].name;
        if (ms) ms[t] = h->pt[t].ms;
        if (calls) calls[t] = h->pt[t].calls;
    }
    return n;
}

int mhd_copy_rate(long bytes, int iters, double *bytes_per_s)
{
    if (bytes < (1 << 20) || iters < 1 || !bytes_per_s) return MHD_ERR_ARG;
    long n = bytes / 8;
    double *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n * 8) != hipSuccess) return MHD_ERR_NOMEM;
    if (hipMalloc(&b, n * 8) != hipSuccess) { (void)hipFree(a); return MHD_ERR_NOMEM; }
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return MHD_ERR_HIP; }
    (void)hipMemsetAsync(a, 0, n * 8, st);
    dim3 gr(blocks(n, 256));
    hipLaunchKernelGGL(k_copy, gr, dim3(256), 0, st, n, a, b);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
    for (int t = 0; t < iters; t++) hipLaunchKernelGGL(k_copy, gr, dim3(256), 0, st, n, a, b);
    (void)hipEventRecord(e1, st);
    hipError_t e = hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipStreamDestroy(st);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess || ms <= 0) return MHD_ERR_HIP;
    *bytes_per_s = 2.0 * 8.0 * n * iters / (ms * 1e-3);
    return MHD_OK;
}

const char *mhd_last_error(const mhd_handle *h) { return h ? h->err.c_str() : "null handle"; }

} // extern "C"

// ---- ensemble: independent runs advanced concurrently ----
// Every run keeps its own stream.  A step pair (setdt; stepon; setdt;
// stepon) of a run is captured once as a HIP graph: stepon renames x and
// xfold on the host, so a pair leaves the pointers where it found them and
// the graph can be replayed; a graph is kept for each of the two pointer
// states a run can be in.  The pairs of all runs are launched round-robin,
// so the device works on as many runs at once as it has hardware queues.
struct RunGraphs {
    double *x = nullptr;                   // the run's x buffer when captured
    hipGraphExec_t first = nullptr;        // pair opening an advance (no time += dt first)
    hipGraphExec_t next = nullptr;
};

// The whole batch's step pair as one graph: every run's pair on its own
// stream, forked from and joined back to the lead stream during capture.
struct BatchGraphs {
    std::vector<double *> xs;              // the runs' x buffers when captured
    hipGraphExec_t first = nullptr;
    hipGraphExec_t next = nullptr;
};

// The batched-kernel form: each sweep of a step pair is one launch over
// the runs of a wave (a merged {run, block word} table and a table of the
// runs' arguments), the wave's launches captured once into a graph on the
// stream of its first run.  Runs whose launch sequences differ (Hall on /
// off, face-window rows) go to different waves; a wave holds at most
// wave_cells physical cells unless it is one run (0: no limit).  The waves
// replay concurrently.  The argument and block tables live in device memory
// held by the graphs.
struct KWave {
    std::vector<int> runs;
    hipStream_t s = nullptr;
    hipGraphExec_t first = nullptr;
    hipGraphExec_t next = nullptr;
};

struct KernGraphs {
    std::vector<double *> xs;              // the runs' x buffers when recorded
    std::vector<KWave> waves;
    std::vector<void *> mem;
    size_t bytes = 0;                      // device bytes of the tables
    int groups = 0, launches = 0;          // run groups, launches per step pair and wave
};

enum { FORM_RUNS = MHD_FORM_RUNS, FORM_BATCH_GRAPH = MHD_FORM_BATCH_GRAPH, FORM_KERNELS = MHD_FORM_KERNELS };

struct mhd_ensemble {
    std::vector<mhd_handle *> runs;
    std::vector<std::pair<int, int>> chunks;   // each run's own (kc, zkc), restored at destroy
    std::vector<int> face12s;                  // and its face-window rows choice
    std::vector<int> rb2s;                     // and its right-part-2 kernels
    int policy = 0;                            // z-chunk policy in force (ENS_POLICIES)
    int form = FORM_RUNS;                      // replay form (FORM_*)
    long wave_cells = 0;                       // batched kernels: cells per wave (0: no limit)
    std::vector<RunGraphs> graphs;         // two per run
    BatchGraphs bg[2];
    KernGraphs kg[2];
    std::vector<int> status;
    std::vector<hipEvent_t> done;          // end of each run's work
    hipStream_t lead = nullptr;            // timing stream
    hipEvent_t e0 = nullptr, e1 = nullptr, fork = nullptr;
    std::string err;
};

// tile shape of a handle: z chunk lengths of its tile and smoothing
// kernels, the rows of its face windows (the smoothing block tables are
// rebuilt) and the form of right part 2 (rb2 < 0 keeps the handle's).
// Results do not depend on them.
static int set_shape(mhd_handle *h, int kc, int zkc, int face12, int rb2 = -1)
{
    if (rb2 >= 0) h->rb2 = rb2;
    if (h->kc == kc && h->zkc == zkc && h->face12 == face12) return 0;
    HCHK(hipStreamSynchronize(h->stream));
    h->kc = kc;
    h->zkc = zkc;
    h->face12 = face12;
    for (auto &zc : h->zt)
        for (auto &t : zc) {
            if (t.d) (void)hipFree(t.d);
            t = ZTab();
        }
    for (auto &t : h->zt3) {
        if (t.d) (void)hipFree(t.d);
        t = ZTab();
    }
    for (int c = 0; c < 2; c++) {
        int r = zs_tables(h, c ? 8 : 3, 2, h->zt[c]);
        if (r) return r;
    }
    return zs_tables(h, 3, 3, h->zt3);
}

static int set_chunks(mhd_handle *h, int kc, int zkc) { return set_shape(h, kc, zkc, h->face12); }

// The ensemble keeps every launch of a run on the run's one stream: the
// side-stream launches (overlap) stay with the single-run loop (a 48-run
// ensemble of per-run graphs with forked branches ended in a host-side
// fault inside the runtime).
struct NoOverlap {
    mhd_handle *h;
    int saved;
    explicit NoOverlap(mhd_handle *hh) : h(hh), saved(hh->overlap) { h->overlap = 0; }
    ~NoOverlap() { h->overlap = saved; }
};

static int capture_pair(mhd_handle *h, int opening, hipGraphExec_t *out)
{
    hipGraph_t gr = nullptr;
    NoOverlap single(h);
    HCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    dt_kernels(h, opening ? 0 : 1);
    step_body(h);
    restore_after_collapse(h);
    dt_kernels(h, 1);
    step_body(h);
    restore_after_collapse(h);
    HCHK(hipStreamEndCapture(h->stream, &gr));
    const hipError_t e = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
        h->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
        return MHD_ERR_HIP;
    }
    return MHD_OK;
}

static void drop_graphs(RunGraphs &g)
{
    if (g.first) (void)hipGraphExecDestroy(g.first);
    if (g.next) (void)hipGraphExecDestroy(g.next);
    g = RunGraphs{};
}

static void drop_batch(BatchGraphs &g)
{
    if (g.first) (void)hipGraphExecDestroy(g.first);
    if (g.next) (void)hipGraphExecDestroy(g.next);
    g = BatchGraphs{};
}

// one step pair of every run, captured from the lead stream
static int capture_batch(mhd_ensemble *e, int opening, hipGraphExec_t *out)
{
    hipGraph_t gr = nullptr;
    auto fail = [&](hipError_t st, const char *what) {
        e->err = std::string(what) + ": " + hipGetErrorString(st);
        return MHD_ERR_HIP;
    };
    hipError_t st = hipStreamBeginCapture(e->lead, hipStreamCaptureModeThreadLocal);
    if (st != hipSuccess) return fail(st, "hipStreamBeginCapture");
    st = hipEventRecord(e->fork, e->lead);
    for (size_t r = 0; r < e->runs.size() && st == hipSuccess; r++) {
        mhd_handle *h = e->runs[r];
        st = hipStreamWaitEvent(h->stream, e->fork, 0);
        if (st != hipSuccess) break;
        NoOverlap single(h);
        dt_kernels(h, opening ? 0 : 1);
        step_body(h);
        restore_after_collapse(h);
        dt_kernels(h, 1);
        step_body(h);
        restore_after_collapse(h);
        st = hipEventRecord(e->done[r], h->stream);
        if (st == hipSuccess) st = hipStreamWaitEvent(e->lead, e->done[r], 0);
    }
    const hipError_t se = hipStreamEndCapture(e->lead, &gr);
    if (st != hipSuccess) { if (gr) (void)hipGraphDestroy(gr); return fail(st, "batch capture"); }
    if (se != hipSuccess) return fail(se, "hipStreamEndCapture");
    st = hipGraphInstantiate(out, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (st != hipSuccess) return fail(st, "hipGraphInstantiate");
    return MHD_OK;
}

// the batch graphs for the runs' current pointer states (captured when missing)
static int batch_graphs(mhd_ensemble *e, BatchGraphs **out)
{
    std::vector<double *> xs;
    for (auto *h : e->runs) xs.push_back(h->x);
    BatchGraphs *slot = nullptr;
    for (auto &g : e->bg) {
        if (g.first && g.xs == xs) { *out = &g; return MHD_OK; }
        if (!slot && !g.first) slot = &g;
    }
    if (!slot) { slot = &e->bg[0]; drop_batch(*slot); }
    int rc = capture_batch(e, 1, &slot->first);
    if (!rc) rc = capture_batch(e, 0, &slot->next);
    if (rc) { drop_batch(*slot); return rc; }
    slot->xs = xs;
    *out = slot;
    return MHD_OK;
}

static void drop_kern(mhd_ensemble *e, KernGraphs &g)
{
    (void)e;
    for (auto &w : g.waves) {
        if (w.s && (w.first || w.next)) (void)hipStreamSynchronize(w.s);
        if (w.first) (void)hipGraphExecDestroy(w.first);
        if (w.next) (void)hipGraphExecDestroy(w.next);
    }
    for (void *p : g.mem) (void)hipFree(p);
    g = KernGraphs{};
}

static void drop_all(mhd_ensemble *e)
{
    for (auto &g : e->graphs) drop_graphs(g);
    for (auto &g : e->bg) drop_batch(g);
    for (auto &g : e->kg) drop_kern(e, g);
}

// the launches of one step pair of run h, recorded instead of issued.  The
// run's stream is put in capture mode meanwhile, so that a launch without a
// batched form (viscosity, a non-default kernel switch) lands in a graph
// that is never run; such a run is refused.
static int record_pair(mhd_handle *h, int opening, std::vector<RecLaunch> &out)
{
    out.clear();
    h->rec_err.clear();
    HCHK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    h->rec = &out;
    dt_kernels(h, opening ? 0 : 1);
    step_body(h);
    restore_after_collapse(h);
    dt_kernels(h, 1);
    step_body(h);
    restore_after_collapse(h);
    h->rec = nullptr;
    hipGraph_t gr = nullptr;
    HCHK(hipStreamEndCapture(h->stream, &gr));
    size_t nodes = 0;
    const hipError_t st = hipGraphGetNodes(gr, nullptr, &nodes);
    (void)hipGraphDestroy(gr);
    HCHK(st);
    if (nodes != 0) {
        h->err = "a launch of this run's step has no batched form (viscosity or a non-default kernel switch)";
        return MHD_ERR_ARG;
    }
    if (!h->rec_err.empty()) {
        h->err = h->rec_err;
        return MHD_ERR_ARG;
    }
    return MHD_OK;
}

// the batched-kernel graphs for the runs' current pointer states
static int kern_graphs(mhd_ensemble *e, KernGraphs **out)
{
    const int n = (int)e->runs.size();
    std::vector<double *> xs;
    for (auto *h : e->runs) xs.push_back(h->x);
    KernGraphs *slot = nullptr;
    for (auto &g : e->kg) {
        if (!g.waves.empty() && g.xs == xs) { *out = &g; return MHD_OK; }
        if (!slot && g.waves.empty()) slot = &g;
    }
    if (!slot) { slot = &e->kg[0]; drop_kern(e, *slot); }
    // record both pair kinds of every run
    std::vector<std::vector<RecLaunch>> seq[2];
    for (int o = 0; o < 2; o++) {
        seq[o].resize(n);
        for (int r = 0; r < n; r++) {
            const int rc = record_pair(e->runs[r], o == 0 ? 1 : 0, seq[o][r]);
            if (rc) { e->err = "run " + std::to_string(r) + ": " + e->runs[r]->err; return rc; }
        }
    }
    // group the runs by launch sequence
    std::vector<std::vector<int>> groups;
    {
        std::map<std::vector<BLaunch>, int> gid;
        for (int r = 0; r < n; r++) {
            std::vector<BLaunch> sig;
            for (int o = 0; o < 2; o++)
                for (const auto &l : seq[o][r]) sig.push_back(l.fn);
            auto it = gid.find(sig);
            if (it == gid.end()) {
                gid.emplace(sig, (int)groups.size());
                groups.push_back({r});
            } else {
                groups[it->second].push_back(r);
            }
        }
    }
    // waves: consecutive runs of a group up to the cell budget
    KernGraphs &K = *slot;
    for (const auto &grp : groups) {
        long cells = 0;
        bool open = false;
        for (int r : grp) {
            const mhd_handle *h = e->runs[r];
            const long c = (long)h->cfg.mx * h->cfg.my * h->cfg.mz;
            if (!open || (e->wave_cells > 0 && cells + c > e->wave_cells)) {
                K.waves.emplace_back();
                cells = 0;
                open = true;
            }
            K.waves.back().runs.push_back(r);
            cells += c;
        }
    }
    for (auto &w : K.waves) w.s = e->runs[w.runs[0]]->stream;
    // merged tables, uploaded once (identical tables shared)
    std::map<std::string, void *> blobs;
    hipError_t st = hipSuccess;
    auto upload = [&](std::string &&bytes) -> void * {
        auto it = blobs.find(bytes);
        if (it != blobs.end()) return it->second;
        void *d = nullptr;
        if (st == hipSuccess) st = hipMalloc(&d, bytes.size());
        if (st == hipSuccess) st = hipMemcpy(d, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
        if (st != hipSuccess) { if (d) (void)hipFree(d); return nullptr; }
        K.mem.push_back(d);
        K.bytes += bytes.size();
        blobs.emplace(std::move(bytes), d);
        return d;
    };
    struct Plan {
        BLaunch fn;
        const void *args;
        const BEntry *tab;
        unsigned nblk;
    };
    std::vector<std::vector<Plan>> plan[2];
    for (int o = 0; o < 2 && st == hipSuccess; o++)
        for (const auto &wv : K.waves) {
            const std::vector<int> &grp = wv.runs;
            plan[o].emplace_back();
            const size_t nl = seq[o][grp[0]].size();
            for (size_t i = 0; i < nl && st == hipSuccess; i++) {
                std::string args;
                std::vector<BEntry> nat;
                for (size_t k = 0; k < grp.size(); k++) {
                    const RecLaunch &l = seq[o][grp[k]][i];
                    args += l.arg;
                    for (unsigned w : l.words) nat.push_back(BEntry{(unsigned)k, w});
                }
                if (nat.empty()) continue;
                // dispatch position L runs on XCD L % 8: each XCD takes a
                // contiguous range of the merged enumeration
                const unsigned T = (unsigned)nat.size(), per = T / NXCD, rem = T % NXCD;
                std::vector<BEntry> tab(T);
                for (unsigned L = 0; L < T; L++) {
                    const unsigned xcd = L % NXCD;
                    tab[L] = nat[xcd * per + (xcd < rem ? xcd : rem) + L / NXCD];
                }
                void *da = upload(std::move(args));
                void *dt = upload(std::string((const char *)tab.data(), T * sizeof(BEntry)));
                if (!da || !dt) break;
                plan[o].back().push_back(Plan{seq[o][grp[0]][i].fn, da, (const BEntry *)dt, T});
            }
        }
    if (st != hipSuccess) {
        e->err = std::string("batched tables: ") + hipGetErrorString(st);
        drop_kern(e, K);
        return MHD_ERR_NOMEM;
    }
    // capture each wave's pair on its stream
    for (size_t w = 0; w < K.waves.size(); w++)
        for (int o = 0; o < 2; o++) {
            KWave &wv = K.waves[w];
            hipGraph_t gr = nullptr;
            st = hipStreamBeginCapture(wv.s, hipStreamCaptureModeThreadLocal);
            if (st == hipSuccess) {
                for (const Plan &p : plan[o][w]) p.fn(p.args, p.tab, p.nblk, wv.s);
                st = hipStreamEndCapture(wv.s, &gr);
            }
            if (st == hipSuccess) st = hipGraphInstantiate(o == 0 ? &wv.first : &wv.next, gr, nullptr, nullptr, 0);
            if (gr) (void)hipGraphDestroy(gr);
            if (st != hipSuccess) {
                e->err = std::string("batched capture: ") + hipGetErrorString(st);
                drop_kern(e, K);
                return MHD_ERR_HIP;
            }
        }
    K.xs = xs;
    K.groups = (int)groups.size();
    K.launches = K.waves.empty() ? 0 : (int)plan[1][0].size();
    *out = &K;
    return MHD_OK;
}

// the graphs of run r for its current pointer state (captured when missing)
static int run_graphs(mhd_ensemble *e, int r, RunGraphs **out)
{
    mhd_handle *h = e->runs[r];
    RunGraphs *slot = nullptr;
    for (int k = 0; k < 2; k++) {
        RunGraphs &g = e->graphs[2 * r + k];
        if (g.x == h->x) { *out = &g; return MHD_OK; }
        if (!slot && !g.x) slot = &g;
    }
    if (!slot) { slot = &e->graphs[2 * r]; drop_graphs(*slot); }
    int rc = capture_pair(h, 1, &slot->first);
    if (!rc) rc = capture_pair(h, 0, &slot->next);
    if (rc) { drop_graphs(*slot); return rc; }
    slot->x = h->x;
    *out = slot;
    return MHD_OK;
}

// z-chunk policies the ensemble tuner times: 0 long chunks, 1 each run's
// own (sized to fill the device alone), 2 16/16, 3 8/8 (tile / smoothing)
static constexpr int ENS_POLICIES = 4;

static std::pair<int, int> policy_chunks(const mhd_ensemble *e, int r, int p)
{
    const bool big = e->runs[r]->S >= (1L << 25);
    switch (p) {
    case 0: return {big ? 16 : 32, big ? 64 : 32};
    case 1: return e->chunks[r];
    case 2: return {16, 16};
    default: return {8, 8};
    }
}

static int apply_policy(mhd_ensemble *e, int p)
{
    drop_all(e);   // captured with the old launch arguments
    for (size_t r = 0; r < e->runs.size(); r++) {
        const auto c = policy_chunks(e, (int)r, p);
        const int rc = set_chunks(e->runs[r], c.first, c.second);
        if (rc) { e->err = "run " + std::to_string(r) + ": " + e->runs[r]->err; return rc; }
    }
    e->policy = p;
    return MHD_OK;
}

extern "C" {

int mhd_ensemble_create(mhd_handle *const *runs, int n, mhd_ensemble **out)
{
    if (!runs || n <= 0 || !out) return MHD_ERR_ARG;
    *out = nullptr;
    for (int r = 0; r < n; r++)
        if (!runs[r] || runs[r]->mag) return MHD_ERR_ARG;   // Harris runs only
    for (int r = 0; r < n; r++)
        for (int q = 0; q < r; q++)
            if (runs[q] == runs[r]) return MHD_ERR_ARG;   // a run appears once
    mhd_ensemble *e = new mhd_ensemble();
    e->runs.assign(runs, runs + n);
    e->graphs.resize(2 * n);
    e->status.assign(n, MHD_OK);
    e->done.assign(n, nullptr);
    *out = e;
    hipError_t st = hipStreamCreate(&e->lead);
    if (st == hipSuccess) st = hipEventCreate(&e->e0);
    if (st == hipSuccess) st = hipEventCreate(&e->e1);
    if (st == hipSuccess) st = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    for (int r = 0; r < n && st == hipSuccess; r++) st = hipEventCreateWithFlags(&e->done[r], hipEventDisableTiming);
    if (st != hipSuccess) { e->err = std::string("ensemble set-up: ") + hipGetErrorString(st); return MHD_ERR_HIP; }
    // runs that share the device do not need short chunks to fill it: the
    // long ones (less z halo per plane) measured faster for the C4 batch;
    // mhd_ensemble_tune picks by measurement
    for (int r = 0; r < n; r++) {
        e->chunks.emplace_back(e->runs[r]->kc, e->runs[r]->zkc);
        e->face12s.push_back(e->runs[r]->face12);
        e->rb2s.push_back(e->runs[r]->rb2);
    }
    return apply_policy(e, 0);
}

int mhd_ensemble_advance(mhd_ensemble *e, int nsteps, const double *dt0, double *time, double *elapsed_ms)
{
    if (!e || nsteps < 0 || !dt0 || !time) return MHD_ERR_ARG;
    const int n = (int)e->runs.size();
    std::vector<RunGraphs *> gs(n, nullptr);
    for (int r = 0; r < n; r++) {
        mhd_handle *h = e->runs[r];
        e->status[r] = MHD_OK;
        if (!h->have_eq || h->prof) {
            e->err = "run " + std::to_string(r) + (h->prof ? ": profiling is on" : ": equilibrium not set");
            return MHD_ERR_ARG;
        }
        if (nsteps >= 2 && e->form == FORM_RUNS) {
            const int rc = run_graphs(e, r, &gs[r]);
            if (rc) { e->err = "run " + std::to_string(r) + ": " + h->err; return rc; }
        }
    }
    // the whole-batch graph: one graph per step pair on the lead stream;
    // the batched kernels: one graph per wave and step pair
    hipGraphExec_t gfirst = nullptr, gnext = nullptr;
    KernGraphs *kgs = nullptr;
    if (nsteps >= 2 && e->form == FORM_BATCH_GRAPH) {
        BatchGraphs *bgs = nullptr;
        const int rc = batch_graphs(e, &bgs);
        if (rc) return rc;
        gfirst = bgs->first;
        gnext = bgs->next;
    } else if (nsteps >= 2 && e->form == FORM_KERNELS) {
        const int rc = kern_graphs(e, &kgs);
        if (rc) return rc;
    }
    auto hip_fail = [&](hipError_t st, const char *what) {
        e->err = std::string(what) + ": " + hipGetErrorString(st);
        return MHD_ERR_HIP;
    };
    hipError_t st;
    for (int r = 0; r < n; r++) {
        mhd_handle *h = e->runs[r];
        const double init[8] = {0, time[r], dt0[r], 0, 0, 0, 0, 0};
        if ((st = hipMemcpyAsync(h->dtv, init, sizeof(init), hipMemcpyHostToDevice, h->stream)) != hipSuccess)
            return hip_fail(st, "dt upload");
        if ((st = hipStreamSynchronize(h->stream)) != hipSuccess) return hip_fail(st, "dt upload");
    }
    if ((st = hipEventRecord(e->e0, e->lead)) != hipSuccess) return hip_fail(st, "hipEventRecord");
    for (int r = 0; r < n; r++)
        if ((st = hipStreamWaitEvent(e->runs[r]->stream, e->e0, 0)) != hipSuccess) return hip_fail(st, "hipStreamWaitEvent");
    // step pairs round-robin over the runs (or the batch graph on the lead
    // stream), then an odd last step eagerly
    if (gfirst) {
        for (int p = 0; p < nsteps / 2; p++)
            if ((st = hipGraphLaunch(p == 0 ? gfirst : gnext, e->lead)) != hipSuccess)
                return hip_fail(st, "hipGraphLaunch");
        if ((st = hipEventRecord(e->fork, e->lead)) != hipSuccess) return hip_fail(st, "hipEventRecord");
        for (int r = 0; r < n; r++)
            if ((st = hipStreamWaitEvent(e->runs[r]->stream, e->fork, 0)) != hipSuccess)
                return hip_fail(st, "hipStreamWaitEvent");
    } else if (kgs) {
        for (int p = 0; p < nsteps / 2; p++)
            for (const KWave &w : kgs->waves)
                if ((st = hipGraphLaunch(p == 0 ? w.first : w.next, w.s)) != hipSuccess)
                    return hip_fail(st, "hipGraphLaunch");
        // the other runs of a wave continue after it (odd step, join)
        for (const KWave &w : kgs->waves) {
            const int r0 = w.runs[0];
            if ((st = hipEventRecord(e->done[r0], w.s)) != hipSuccess) return hip_fail(st, "hipEventRecord");
            for (size_t k = 1; k < w.runs.size(); k++)
                if ((st = hipStreamWaitEvent(e->runs[w.runs[k]]->stream, e->done[r0], 0)) != hipSuccess)
                    return hip_fail(st, "hipStreamWaitEvent");
        }
    } else {
        for (int p = 0; p < nsteps / 2; p++)
            for (int r = 0; r < n; r++)
                if ((st = hipGraphLaunch(p == 0 ? gs[r]->first : gs[r]->next, e->runs[r]->stream)) != hipSuccess)
                    return hip_fail(st, "hipGraphLaunch");
    }
    if (nsteps % 2)
        for (int r = 0; r < n; r++) {
            NoOverlap single(e->runs[r]);
            dt_kernels(e->runs[r], nsteps > 1 ? 1 : 0);
            step_body(e->runs[r]);
            restore_after_collapse(e->runs[r]);
        }
    for (int r = 0; r < n; r++) {
        if ((st = hipEventRecord(e->done[r], e->runs[r]->stream)) != hipSuccess) return hip_fail(st, "hipEventRecord");
        if ((st = hipStreamWaitEvent(e->lead, e->done[r], 0)) != hipSuccess) return hip_fail(st, "hipStreamWaitEvent");
    }
    if ((st = hipEventRecord(e->e1, e->lead)) != hipSuccess) return hip_fail(st, "hipEventRecord");
    if ((st = hipEventSynchronize(e->e1)) != hipSuccess) return hip_fail(st, "ensemble run");
    float ms = 0;
    if ((st = hipEventElapsedTime(&ms, e->e0, e->e1)) != hipSuccess) return hip_fail(st, "hipEventElapsedTime");
    if (elapsed_ms) *elapsed_ms = ms;
    if ((st = hipGetLastError()) != hipSuccess) return hip_fail(st, "ensemble run");
    int first = MHD_OK;
    for (int r = 0; r < n; r++) {
        mhd_handle *h = e->runs[r];
        double v[8];
        if ((st = hipMemcpy(v, h->dtv, sizeof(v), hipMemcpyDeviceToHost)) != hipSuccess) return hip_fail(st, "dt download");
        time[r] = v[1] + (nsteps > 0 && v[3] == 0.0 ? v[4] : 0.0);
        if (v[3] != 0.0) {
            set_err(h, "dt collapsed below 0.1*dt0");
            e->status[r] = MHD_ERR_DT_COLLAPSE;
            if (first == MHD_OK) { first = MHD_ERR_DT_COLLAPSE; e->err = "run " + std::to_string(r) + ": " + h->err; }
        }
    }
    return first;
}

int mhd_ensemble_tune(mhd_ensemble *e, int trial, const double *dt0, const double *time, int *best, double *ms)
{
    if (!e || trial < 1 || !dt0 || !time) return MHD_ERR_ARG;
    const int n = (int)e->runs.size();
    std::vector<double *> keep(n, nullptr);
    auto release = [&]() {
        for (auto p : keep)
            if (p) (void)hipFree(p);
    };
    for (int r = 0; r < n; r++) {
        mhd_handle *h = e->runs[r];
        if (hipMalloc(&keep[r], 8 * h->S * sizeof(double)) != hipSuccess ||
            hipMemcpyAsync(keep[r], h->x, 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            release();
            e->err = "tune: saving the states failed";
            return MHD_ERR_NOMEM;
        }
    }
    double tms[ENS_POLICIES];
    int pick = 0;
    for (int p = 0; p < ENS_POLICIES; p++) {
        int rc = apply_policy(e, p);
        std::vector<double> t(time, time + n);
        double el = 0;
        // two steps capture the graphs; the trial that follows is timed
        if (!rc) rc = mhd_ensemble_advance(e, 2, dt0, t.data(), nullptr);
        if (!rc) rc = mhd_ensemble_advance(e, trial, dt0, t.data(), &el);
        tms[p] = (rc == MHD_OK) ? el : 1e30;
        for (int r = 0; r < n; r++) {
            mhd_handle *h = e->runs[r];
            if (hipMemcpyAsync(h->x, keep[r], 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                release();
                e->err = "tune: restoring the states failed";
                return MHD_ERR_HIP;
            }
        }
        if (rc != MHD_OK && rc != MHD_ERR_DT_COLLAPSE) { release(); return rc; }
        if (tms[p] < tms[pick]) pick = p;
        if (ms) ms[p] = tms[p];
    }
    release();
    if (best) *best = pick;
    return apply_policy(e, pick);
}

int mhd_ensemble_set_shape(mhd_ensemble *e, const mhd_tile_shape *sh)
{
    if (!e || !sh || sh->kc < 1 || sh->zkc < 4 || sh->zkc > ZKC_MAX || (sh->face_rows != 12 && sh->face_rows != 16) ||
        sh->form < FORM_RUNS || sh->form > FORM_KERNELS || sh->rhs_rows < -1 || sh->rhs_rows > 2 ||
        sh->wave_cells < 0)
        return MHD_ERR_ARG;
    drop_all(e);
    for (size_t r = 0; r < e->runs.size(); r++) {
        const int rc = set_shape(e->runs[r], sh->kc, sh->zkc, sh->face_rows == 12 ? 1 : 0,
                                 sh->rhs_rows < 0 ? e->rb2s[r] : sh->rhs_rows);
        if (rc) { e->err = "run " + std::to_string(r) + ": " + e->runs[r]->err; return rc; }
    }
    e->form = sh->form;
    e->wave_cells = sh->wave_cells;
    return MHD_OK;
}

int mhd_ensemble_tune_shapes(mhd_ensemble *e, int trial, const double *dt0, const double *time, int max_trials,
                             mhd_tile_shape *trials, int *ntrials, mhd_tile_shape *best)
{
    if (!e || trial < 1 || !dt0 || !time || max_trials < 0 || (max_trials > 0 && !trials)) return MHD_ERR_ARG;
    const int n = (int)e->runs.size();
    std::vector<double *> keep(n, nullptr);
    auto release = [&]() {
        for (auto p : keep)
            if (p) (void)hipFree(p);
    };
    for (int r = 0; r < n; r++) {
        mhd_handle *h = e->runs[r];
        if (hipMalloc(&keep[r], 8 * h->S * sizeof(double)) != hipSuccess ||
            hipMemcpyAsync(keep[r], h->x, 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
            hipStreamSynchronize(h->stream) != hipSuccess) {
            release();
            e->err = "tune: saving the states failed";
            return MHD_ERR_NOMEM;
        }
    }
    auto apply = [&](const mhd_tile_shape &sh) {
        drop_all(e);
        for (int r = 0; r < n; r++) {
            const int rc = set_shape(e->runs[r], sh.kc, sh.zkc, sh.face_rows == 12 ? 1 : 0,
                                     sh.rhs_rows < 0 ? e->rb2s[r] : sh.rhs_rows);
            if (rc) { e->err = "run " + std::to_string(r) + ": " + e->runs[r]->err; return rc; }
        }
        e->form = sh.form;
        e->wave_cells = sh.wave_cells;
        return (int)MHD_OK;
    };
    int count = 0, rc = MHD_OK;
    // one trial: two untimed steps (they capture the graphs), then `trial`
    // timed steps of the whole batch; the states are restored afterwards
    auto measure = [&](mhd_tile_shape &sh) {
        int r0 = apply(sh);
        std::vector<double> t(time, time + n);
        double el = 0;
        if (!r0) r0 = mhd_ensemble_advance(e, 2, dt0, t.data(), nullptr);
        if (!r0) r0 = mhd_ensemble_advance(e, trial, dt0, t.data(), &el);
        sh.ms = (r0 == MHD_OK) ? el : 1e30;
        for (int r = 0; r < n; r++) {
            mhd_handle *h = e->runs[r];
            if (hipMemcpyAsync(h->x, keep[r], 8 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream) != hipSuccess ||
                hipStreamSynchronize(h->stream) != hipSuccess) {
                e->err = "tune: restoring the states failed";
                return (int)MHD_ERR_HIP;
            }
        }
        if (count < max_trials) trials[count] = sh;
        count++;
        // runs the batched kernels cannot take (viscosity) rule that form out
        if (r0 == MHD_ERR_ARG && sh.form == FORM_KERNELS) return (int)MHD_OK;
        return (r0 == MHD_OK || r0 == MHD_ERR_DT_COLLAPSE) ? (int)MHD_OK : r0;
    };
    // coordinate search from the long-chunk shape: the replay form (per-run
    // graphs, or batched kernels over all runs at once or in waves of 1/4,
    // 1/8, 1/16 of the cells), the smoothing chunk, the tile-kernel chunk,
    // the face-window rows, the right-part-2 kernels, then the other forms
    long total = 0;
    for (auto *h : e->runs) total += (long)h->cfg.mx * h->cfg.my * h->cfg.mz;
    mhd_tile_shape cur{32, 32, 12, FORM_RUNS, -1, 0, 0.0};
    rc = measure(cur);
    mhd_tile_shape kb = cur;
    kb.ms = 1e30;
    for (int d : {1, 4, 8, 16}) {
        if (rc) break;
        mhd_tile_shape s = cur;
        s.form = FORM_KERNELS;
        s.wave_cells = d == 1 ? 0 : (int)std::min<long>(INT_MAX, (total + d - 1) / d);
        rc = measure(s);
        if (!rc && s.ms < kb.ms) kb = s;
    }
    if (!rc && kb.ms < cur.ms) cur = kb;
    const int zkcs[] = {16, 24, 48, 64};
    const int kcs[] = {8, 16, 64};
    for (int v : zkcs) {
        if (rc) break;
        mhd_tile_shape s = cur;
        s.zkc = v;
        rc = measure(s);
        if (!rc && s.ms < cur.ms) cur = s;
    }
    for (int v : kcs) {
        if (rc) break;
        mhd_tile_shape s = cur;
        s.kc = v;
        rc = measure(s);
        if (!rc && s.ms < cur.ms) cur = s;
    }
    if (!rc) {
        mhd_tile_shape s = cur;
        s.face_rows = 16;
        rc = measure(s);
        if (!rc && s.ms < cur.ms) cur = s;
    }
    // right part 2 in one kind for every run: the tile kernel, the 16-row
    // and the 8-row windowed kernels
    for (int v = 0; v <= 2 && !rc; v++) {
        mhd_tile_shape s = cur;
        s.rhs_rows = v;
        rc = measure(s);
        if (!rc && s.ms < cur.ms) cur = s;
    }
    const mhd_tile_shape at = cur;
    for (int f = FORM_RUNS; f <= FORM_KERNELS && !rc; f++) {
        if (f == at.form) continue;
        mhd_tile_shape s = at;
        s.form = f;
        if (f == FORM_KERNELS) s.wave_cells = kb.wave_cells;
        rc = measure(s);
        if (!rc && s.ms < cur.ms) cur = s;
    }
    release();
    if (ntrials) *ntrials = count;
    if (rc) return rc;
    if (best) *best = cur;
    return apply(cur);
}

int mhd_ensemble_status(const mhd_ensemble *e, int *status)
{
    if (!e || !status) return MHD_ERR_ARG;
    for (size_t r = 0; r < e->status.size(); r++) status[r] = e->status[r];
    return MHD_OK;
}

const char *mhd_ensemble_error(const mhd_ensemble *e) { return e ? e->err.c_str() : "null ensemble"; }

void mhd_ensemble_destroy(mhd_ensemble *e)
{
    if (!e) return;
    if (e->lead) (void)hipStreamSynchronize(e->lead);
    for (auto *h : e->runs) (void)hipStreamSynchronize(h->stream);
    drop_all(e);
    for (size_t r = 0; r < e->chunks.size(); r++)
        (void)set_shape(e->runs[r], e->chunks[r].first, e->chunks[r].second, e->face12s[r], e->rb2s[r]);
    if (e->fork) (void)hipEventDestroy(e->fork);
    for (auto ev : e->done)
        if (ev) (void)hipEventDestroy(ev);
    if (e->e0) (void)hipEventDestroy(e->e0);
    if (e->e1) (void)hipEventDestroy(e->e1);
    if (e->lead) (void)hipStreamDestroy(e->lead);
    delete e;
}

void mhd_destroy(mhd_handle *h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    if (h->side) hipStreamSynchronize(h->side);
    if (h->aux) hipStreamSynchronize(h->aux);
    if (h->aux_side) hipStreamSynchronize(h->aux_side);
    double *ps[] = {h->geo, h->x, h->xfold, h->xm, h->xdA, h->xdB, h->eq, h->xresd, h->cA, h->cB,
                    h->eA, h->eB, h->hA, h->hB, h->red, h->dtv, h->crad, h->rr, h->chtab, h->vel, h->mgeo};
    for (auto p : ps)
        if (p) hipFree(p);
    if (h->mrow) hipFree(h->mrow);
    for (auto &zc : h->zt)
        for (auto &t : zc)
            if (t.d) hipFree(t.d);
    for (auto &t : h->zt3)
        if (t.d) hipFree(t.d);
    for (auto e : h->pool) (void)hipEventDestroy(e);
    if (h->stage) (void)hipHostFree(h->stage);
    if (h->fork_ev) (void)hipEventDestroy(h->fork_ev);
    if (h->join_ev) (void)hipEventDestroy(h->join_ev);
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->aux_ev) (void)hipEventDestroy(h->aux_ev);
    if (h->aux) (void)hipStreamDestroy(h->aux);
    if (h->aux_side) (void)hipStreamDestroy(h->aux_side);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

} // extern "C"
