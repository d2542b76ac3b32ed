// mhd_gpu.hip -- MI355X (gfx950) RK4 Hall-MHD step behind the C ABI of
// include/mhd_gpu.h.
//
// The step is the reference's stepon (t3dsdsc.f:1063-1185): four calls of
// the right-hand side `right` (t3dsdsc.f:1297-1673), each a chain of
// dependent sweeps with boundary fills between them, RK4 register updates,
// then a smoothing pass on the perturbation x - eq.
//
// Device layout: one structure-of-arrays FP64 volume per field covering the
// reference's extended range only (x: jx = 3..mx+2, y: 1..my, z: 3..mz+2),
// x contiguous.  The reference's outer ghost columns are never read on a
// single domain and are not stored.  Index (i,j,k) = (jx-3, jy-1, jz-3);
// the interior (stencil) points are i in [2, NX-3], j in [2, NY-3],
// k in [2, NZ-3].
//
// Sweeps write interior points only.  Every boundary routine of the
// reference (bndry, bndryc/E/h, bndrytt: sequential x, then y, then z
// rules) runs as ONE launch over the face points: each face thread
// evaluates the composed rule as a pure function of the freshly swept
// interior and, where the reference's rule reads them, of the face values
// the sweep started from.  Ping-pong buffers keep those available, so no
// in-place races and no copy of the face layer are needed.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mhd_gpu.h"

namespace {

// ---------------------------------------------------------------------------
// geometry passed by value to kernels
// ---------------------------------------------------------------------------
struct Geo {
    int NX, NY, NZ;
    long sy, sz, S;           // y stride, z stride, field stride (elements)
    int periodic;
    // 1-D first-derivative coefficients a,b,c,d (t3dsdsc.f:439-504)
    const double *x1a, *x1b, *x1c, *x1d, *y1a, *y1b, *y1c, *y1d, *z1a, *z1b, *z1c, *z1d;
    // 1-D second-derivative coefficients (t3dsdsc.f:336-437)
    const double *x2a, *x2b, *x2c, *x2d, *y2a, *y2b, *y2c, *y2d, *z2a, *z2b, *z2c, *z2d;
    // smoothing spacings (xp1-x0, x0-xm1, xp1-xm1) per axis (t3dsdsc.f:5618)
    const double *xhp, *xhm, *xhs, *yhp, *yhm, *yhs, *zhp, *zhm, *zhs;
    // difc weights 2*hp/hs, 2*hm/hs per axis
    const double *xwp, *xwm, *ywp, *ywm, *zwp, *zwm;
    const double *xc, *yc, *zc;          // coordinates
    const double *dxs, *dys, *dzs;       // setdt spacings
    const double *cbx;                   // B-component smoothing coefficient per x
    double ax3, bx3, cx3, ax4, bx4, cx4; // axp at jx=3,4 (x-left extrapolation)
    double axmL1, bxmL1, cxmL1, axmL, bxmL, cxmL; // axm at xlast-1, xlast
};

__host__ __device__ inline long IDX(const Geo &g, int i, int j, int k)
{
    return (long)i + g.sy * j + g.sz * k;
}

__device__ inline bool interior(const Geo &g, int i, int j, int k)
{
    return i >= 2 && i <= g.NX - 3 && j >= 2 && j <= g.NY - 3 && k >= 2 && k <= g.NZ - 3;
}

// 5-point derivative, t3dsdsc.f:1308-1313 (d1fc and d2fc share the form)
__device__ inline double fd5(double fm2, double fm, double f0, double fp, double fp2, double a,
                             double b, double c, double d)
{
    return a * (fp - f0) + b * (f0 - fm) + c * (fp2 - f0) + d * (f0 - fm2);
}

__device__ inline double dx1(const Geo &g, const double *__restrict__ f, long q, int i)
{
    return fd5(f[q - 2], f[q - 1], f[q], f[q + 1], f[q + 2], g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
}
__device__ inline double dy1(const Geo &g, const double *__restrict__ f, long q, int j)
{
    long s = g.sy;
    return fd5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s], g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
}
__device__ inline double dz1(const Geo &g, const double *__restrict__ f, long q, int k)
{
    long s = g.sz;
    return fd5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s], g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
}
__device__ inline double dx2(const Geo &g, const double *__restrict__ f, long q, int i)
{
    return fd5(f[q - 2], f[q - 1], f[q], f[q + 1], f[q + 2], g.x2a[i], g.x2b[i], g.x2c[i], g.x2d[i]);
}
__device__ inline double dy2(const Geo &g, const double *__restrict__ f, long q, int j)
{
    long s = g.sy;
    return fd5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s], g.y2a[j], g.y2b[j], g.y2c[j], g.y2d[j]);
}
__device__ inline double dz2(const Geo &g, const double *__restrict__ f, long q, int k)
{
    long s = g.sz;
    return fd5(f[q - 2 * s], f[q - s], f[q], f[q + s], f[q + 2 * s], g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]);
}

// three-point smoothing term difc, t3dsdsc.f:5618-5619
__device__ inline double difc(double fm, double f0, double fp, double hp, double hm, double hs)
{
    return 2. * ((fp - f0) * hp - (f0 - fm) * hm) / hs;
}

// the same term with the per-axis weights wp = 2*hp/hs, wm = 2*hm/hs
// precomputed (no division per cell; differs from difc by rounding only)
__device__ inline double difw(double fm, double f0, double fp, double wp, double wm)
{
    return (fp - f0) * wp - (f0 - fm) * wm;
}

// radial smoothing coefficient, t3dsdsc.f:5659-5661 / 5895-5897 / 5981-5983
__device__ inline double radial_coef(const Geo &g, int i, int j, int k)
{
    double x = g.xc[i], y = g.yc[j], z = g.zc[k];
    double r = sqrt(x * x + y * y + z * z);
    return 1.0 / 32.0 + (1.0 / 8.0 - 1.0 / 32.0) * 0.5 * (1 + tanh((r - 3.) / 3.));
}

// the same coefficient through 1 + tanh(u) = 2 / (1 + exp(-2u)):
// 1/32 + (3/32) / (1 + exp(-2(r-3)/3)); one exp and one division instead of
// tanh (equal to rounding)
__device__ inline double radial_coef_fast(double x2y2, double z)
{
    double r = sqrt(x2y2 + z * z);
    return 1.0 / 32.0 + (3.0 / 32.0) / (1.0 + exp(-2.0 * (r - 3.) / 3.));
}

enum { KIND_RADIAL = 0, KIND_EIGHTH = 1, KIND_STATE = 2 };
enum { BC_AUX = 0, BC_STATE = 1, BC_TT = 2 };
enum { SRC_BUF = 0, SRC_XEQ = 1 };

struct Eq {
    const double *f[8];   // equilibrium per state field or nullptr (x - eq)
};

// ---------------------------------------------------------------------------
// J = curl(B - B_i) on the interior, t3dsdsc.f:2465-2491
// ---------------------------------------------------------------------------
__global__ void k_curl(Geo g, const double *__restrict__ X, const double *__restrict__ bxi,
                       const double *__restrict__ byi, const double *__restrict__ bzi,
                       double *__restrict__ cur)
{
    int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    int k = 2 + blockIdx.z;
    if (i > g.NX - 3 || j > g.NY - 3) return;
    long q = IDX(g, i, j, k);
    const double *b5 = X + 4 * g.S, *b6 = X + 5 * g.S, *b7 = X + 6 * g.S;
    long sy = g.sy, sz = g.sz;
#define U(o) (b5[q + (o)] - bxi[q + (o)])
#define V(o) (b6[q + (o)] - byi[q + (o)])
#define W(o) (b7[q + (o)] - bzi[q + (o)])
    double wy = fd5(W(-2 * sy), W(-sy), W(0), W(sy), W(2 * sy), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
    double vz = fd5(V(-2 * sz), V(-sz), V(0), V(sz), V(2 * sz), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
    double uz = fd5(U(-2 * sz), U(-sz), U(0), U(sz), U(2 * sz), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
    double wx = fd5(W(-2), W(-1), W(0), W(1), W(2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
    double vx = fd5(V(-2), V(-1), V(0), V(1), V(2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
    double uy = fd5(U(-2 * sy), U(-sy), U(0), U(sy), U(2 * sy), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
#undef U
#undef V
#undef W
    cur[q] = wy - vz;
    cur[q + g.S] = uz - wx;
    cur[q + 2 * g.S] = vx - uy;
}

// ---------------------------------------------------------------------------
// one smoothing sweep on NF fields (interior), t3dsdsc.f:5630-5680 and the
// avrgtt/avrgE/avrgc/avrgh copies of it
// ---------------------------------------------------------------------------
template <int NF, int KIND, int SRC>
__global__ void k_filter(Geo g, const double *__restrict__ A, Eq eq, double *__restrict__ B)
{
    int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    int k = 2 + blockIdx.z;
    if (i > g.NX - 3 || j > g.NY - 3) return;
    long q = IDX(g, i, j, k);
    long sy = g.sy, sz = g.sz;
    double hpx = g.xhp[i], hmx = g.xhm[i], hsx = g.xhs[i];
    double hpy = g.yhp[j], hmy = g.yhm[j], hsy = g.yhs[j];
    double hpz = g.zhp[k], hmz = g.zhm[k], hsz = g.zhs[k];
    double crad = (KIND == KIND_EIGHTH) ? 1.0 / 8.0 : radial_coef(g, i, j, k);
#pragma unroll
    for (int m = 0; m < NF; m++) {
        const double *a = A + m * g.S;
        double f0, fxm, fxp, fym, fyp, fzm, fzp;
        if (SRC == SRC_XEQ && eq.f[m]) {
            const double *e = eq.f[m];
            f0 = a[q] - e[q];
            fxm = a[q - 1] - e[q - 1]; fxp = a[q + 1] - e[q + 1];
            fym = a[q - sy] - e[q - sy]; fyp = a[q + sy] - e[q + sy];
            fzm = a[q - sz] - e[q - sz]; fzp = a[q + sz] - e[q + sz];
        } else {
            f0 = a[q];
            fxm = a[q - 1]; fxp = a[q + 1];
            fym = a[q - sy]; fyp = a[q + sy];
            fzm = a[q - sz]; fzp = a[q + sz];
        }
        double s = difc(fxm, f0, fxp, hpx, hmx, hsx) + difc(fym, f0, fyp, hpy, hmy, hsy) +
                   difc(fzm, f0, fzp, hpz, hmz, hsz);
        double c = (KIND == KIND_STATE && m >= 4 && m <= 6) ? g.cbx[i] : crad;
        B[q + m * g.S] = f0 + c * s;
    }
}

// ---------------------------------------------------------------------------
// z-marching (2.5-D) tiles: a block owns a TX x TY column tile of the
// extended x-y range and walks a chunk of interior z planes.  Values needed
// at x/y neighbours sit in an LDS plane with a halo of R; values needed at
// z neighbours sit in a per-thread register queue.  Each input value is
// loaded once per plane (plus the tile halo).
// ---------------------------------------------------------------------------
constexpr int TX = 64;
constexpr int TY = 8;
constexpr int NT = TX * TY;

struct Chunk {
    int i, j, tx, ty, kb, ke;
    bool valid, inner;
};

__device__ inline Chunk chunk_of(const Geo &g, int kc)
{
    Chunk c;
    c.tx = threadIdx.x % TX;
    c.ty = threadIdx.x / TX;
    c.i = blockIdx.x * TX + c.tx;
    c.j = blockIdx.y * TY + c.ty;
    c.kb = 2 + blockIdx.z * kc;
    c.ke = min(c.kb + kc, g.NZ - 2);
    c.valid = c.i < g.NX && c.j < g.NY;
    c.inner = c.i >= 2 && c.i <= g.NX - 3 && c.j >= 2 && c.j <= g.NY - 3;
    return c;
}

// load helper for a smoothing source: buffer value or x - eq on the fly
template <int SRC>
__device__ inline double src_at(const double *__restrict__ a, const double *__restrict__ e, long q)
{
    if (SRC == SRC_XEQ && e) return a[q] - e[q];
    return a[q];
}

template <int NF, int KIND, int SRC>
__global__ __launch_bounds__(NT) void k_filter25(Geo g, const double *__restrict__ A, Eq eq,
                                                  double *__restrict__ B, int kc)
{
    __shared__ double sp[NF][TY + 2][TX + 2];
    Chunk c = chunk_of(g, kc);
    if (c.kb >= c.ke) return;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const long q0 = (long)c.i + sy * c.j;
    double zm[NF], z0[NF], zp[NF];
    if (c.valid) {
#pragma unroll
        for (int m = 0; m < NF; m++) {
            zm[m] = src_at<SRC>(A + m * S, eq.f[m], q0 + sz * (c.kb - 1));
            z0[m] = src_at<SRC>(A + m * S, eq.f[m], q0 + sz * c.kb);
        }
    }
    const double wpx = c.inner ? g.xwp[c.i] : 0, wmx = c.inner ? g.xwm[c.i] : 0;
    const double wpy = c.inner ? g.ywp[c.j] : 0, wmy = c.inner ? g.ywm[c.j] : 0;
    const double cb = (KIND == KIND_STATE && c.inner) ? g.cbx[c.i] : 0.0;
    const double xy2 = c.inner ? g.xc[c.i] * g.xc[c.i] + g.yc[c.j] * g.yc[c.j] : 0.0;
    // halo role of this thread (x edges and y edges of the tile)
    const int hx = c.tx == 0 ? -1 : (c.tx == TX - 1 ? 1 : 0);
    const int hy = c.ty == 0 ? -1 : (c.ty == TY - 1 ? 1 : 0);
    const bool hxok = hx != 0 && c.j < g.NY && c.i + hx >= 0 && c.i + hx < g.NX;
    const bool hyok = hy != 0 && c.i < g.NX && c.j + hy >= 0 && c.j + hy < g.NY;
    for (int k = c.kb; k < c.ke; k++) {
        const long qk = q0 + sz * k;
        if (c.valid) {
#pragma unroll
            for (int m = 0; m < NF; m++) {
                zp[m] = src_at<SRC>(A + m * S, eq.f[m], qk + sz);
                sp[m][c.ty + 1][c.tx + 1] = z0[m];
            }
        }
        if (hxok) {
#pragma unroll
            for (int m = 0; m < NF; m++) sp[m][c.ty + 1][c.tx + 1 + hx] = src_at<SRC>(A + m * S, eq.f[m], qk + hx);
        }
        if (hyok) {
#pragma unroll
            for (int m = 0; m < NF; m++) sp[m][c.ty + 1 + hy][c.tx + 1] = src_at<SRC>(A + m * S, eq.f[m], qk + hy * sy);
        }
        __syncthreads();
        if (c.inner) {
            const double wpz = g.zwp[k], wmz = g.zwm[k];
            const double crad = (KIND == KIND_EIGHTH) ? 1.0 / 8.0 : radial_coef_fast(xy2, g.zc[k]);
#pragma unroll
            for (int m = 0; m < NF; m++) {
                double f0 = z0[m];
                double s = difw(sp[m][c.ty + 1][c.tx], f0, sp[m][c.ty + 1][c.tx + 2], wpx, wmx) +
                           difw(sp[m][c.ty][c.tx + 1], f0, sp[m][c.ty + 2][c.tx + 1], wpy, wmy) +
                           difw(zm[m], f0, zp[m], wpz, wmz);
                double cc = (KIND == KIND_STATE && m >= 4 && m <= 6) ? cb : crad;
                B[qk + m * S] = f0 + cc * s;
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NF; m++) { zm[m] = z0[m]; z0[m] = zp[m]; }
    }
}

// curl(B - B_i) on the interior (t3dsdsc.f:2465-2491), z-marching
__global__ __launch_bounds__(NT) void k_curl25(Geo g, const double *__restrict__ X, const double *__restrict__ bxi,
                                               const double *__restrict__ byi, const double *__restrict__ bzi,
                                               double *__restrict__ cur, int kc)
{
    __shared__ double sp[3][TY + 4][TX + 4];
    Chunk c = chunk_of(g, kc);
    if (c.kb >= c.ke) return;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const long q0 = (long)c.i + sy * c.j;
    const double *b5 = X + 4 * S, *b6 = X + 5 * S, *b7 = X + 6 * S;
    double uq[5], vq[5];   // z queue of u = Bx - bxi, v = By - byi (planes k-2..k+2)
    if (c.valid) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            long q = q0 + sz * (c.kb - 2 + t);
            uq[t] = b5[q] - bxi[q];
            vq[t] = b6[q] - byi[q];
        }
    }
    const int hx = c.tx < 2 ? -2 : (c.tx >= TX - 2 ? 2 : 0);
    const int hy = c.ty < 2 ? -2 : (c.ty >= TY - 2 ? 2 : 0);
    const bool hxok = hx != 0 && c.j < g.NY && c.i + hx >= 0 && c.i + hx < g.NX;
    const bool hyok = hy != 0 && c.i < g.NX && c.j + hy >= 0 && c.j + hy < g.NY;
    double ax[4], ay[4];
    if (c.inner) {
        ax[0] = g.x1a[c.i]; ax[1] = g.x1b[c.i]; ax[2] = g.x1c[c.i]; ax[3] = g.x1d[c.i];
        ay[0] = g.y1a[c.j]; ay[1] = g.y1b[c.j]; ay[2] = g.y1c[c.j]; ay[3] = g.y1d[c.j];
    }
    for (int k = c.kb; k < c.ke; k++) {
        const long qk = q0 + sz * k;
        if (c.valid) {
            long q = qk + 2 * sz;
            uq[4] = b5[q] - bxi[q];
            vq[4] = b6[q] - byi[q];
            sp[0][c.ty + 2][c.tx + 2] = uq[2];
            sp[1][c.ty + 2][c.tx + 2] = vq[2];
            sp[2][c.ty + 2][c.tx + 2] = b7[qk] - bzi[qk];
        }
        if (hxok) {
            long q = qk + hx;
            sp[0][c.ty + 2][c.tx + 2 + hx] = b5[q] - bxi[q];
            sp[1][c.ty + 2][c.tx + 2 + hx] = b6[q] - byi[q];
            sp[2][c.ty + 2][c.tx + 2 + hx] = b7[q] - bzi[q];
        }
        if (hyok) {
            long q = qk + hy * sy;
            sp[0][c.ty + 2 + hy][c.tx + 2] = b5[q] - bxi[q];
            sp[1][c.ty + 2 + hy][c.tx + 2] = b6[q] - byi[q];
            sp[2][c.ty + 2 + hy][c.tx + 2] = b7[q] - bzi[q];
        }
        __syncthreads();
        if (c.inner) {
            const int X0 = c.tx + 2, Y0 = c.ty + 2;
            double az[4] = {g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]};
#define DXS(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], ax[0], ax[1], ax[2], ax[3])
#define DYS(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], ay[0], ay[1], ay[2], ay[3])
            double wy = DYS(2), wx = DXS(2), vx = DXS(1), uy = DYS(0);
            double vz = fd5(vq[0], vq[1], vq[2], vq[3], vq[4], az[0], az[1], az[2], az[3]);
            double uz = fd5(uq[0], uq[1], uq[2], uq[3], uq[4], az[0], az[1], az[2], az[3]);
#undef DXS
#undef DYS
            cur[qk] = wy - vz;
            cur[qk + S] = uz - wx;
            cur[qk + 2 * S] = vx - uy;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; t++) { uq[t] = uq[t + 1]; vq[t] = vq[t + 1]; }
    }
}

// ---------------------------------------------------------------------------
// right, part 1 as a z-marching tile.  Per point the momentum flux tensor
// (t3dsdsc.f:2311-2346) is formed once:
//   T22 = x2*x2/x1 + (p - p_i), T23 = x2*x3/x1, T24 = x2*x4/x1,
//   T33 = x3*x3/x1 + (p - p_i), T34 = x4*x3/x1, T44 = x4*x4/x1 + (p - p_i)
// and xdif(1..4) = -(d/dx, d/dy, d/dz) of (x2,x3,x4), (T22,T23,T24),
// (T23,T33,T34), (T24,T34,T44): the same values the reference's flux(m)
// arrays hold.  LDS holds {x2,x3,T22,T23,T24,T33,T34,p} of plane k with a
// halo of 2; the register queue holds {x4,T24,T34,T44,p} of planes k-2..k+2.
// ---------------------------------------------------------------------------
struct Rec {
    double x2, x3, x4, t22, t23, t24, t33, t34, t44, p;
};

__device__ inline Rec rec_at(const double *__restrict__ X, const double *__restrict__ pri, long S, long q)
{
    Rec r;
    double x1 = X[q];
    r.x2 = X[q + S]; r.x3 = X[q + 2 * S]; r.x4 = X[q + 3 * S]; r.p = X[q + 7 * S];
    double dp = r.p - pri[q];
    r.t22 = r.x2 * r.x2 / x1 + dp;
    r.t23 = r.x2 * r.x3 / x1;
    r.t24 = r.x2 * r.x4 / x1;
    r.t33 = r.x3 * r.x3 / x1 + dp;
    r.t34 = r.x4 * r.x3 / x1;
    r.t44 = r.x4 * r.x4 / x1 + dp;
    return r;
}

enum { LX2 = 0, LX3, LT22, LT23, LT24, LT33, LT34, LP, NLA };

__device__ inline void rec_to_lds(double (*sp)[TY + 4][TX + 4], int y, int x, const Rec &r)
{
    sp[LX2][y][x] = r.x2; sp[LX3][y][x] = r.x3; sp[LT22][y][x] = r.t22; sp[LT23][y][x] = r.t23;
    sp[LT24][y][x] = r.t24; sp[LT33][y][x] = r.t33; sp[LT34][y][x] = r.t34; sp[LP][y][x] = r.p;
}

__device__ inline void e_point(const double *__restrict__ X, const double *__restrict__ cur, double eta, long S,
                               long q, double *__restrict__ E)
{
    double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
    double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
    E[q] = (-x3 * x7 + x4 * x6) / x1 + eta * cur[q];
    E[q + S] = (-x4 * x5 + x2 * x7) / x1 + eta * cur[q + S];
    E[q + 2 * S] = (-x2 * x6 + x3 * x5) / x1 + eta * cur[q + 2 * S];
}

__global__ __launch_bounds__(NT) void k_rhs_a25(Geo g, const double *__restrict__ X, const double *__restrict__ pri,
                                                const double *__restrict__ cur, const double *__restrict__ xresd,
                                                double eta, int hall, double cht, double *__restrict__ XD,
                                                double *__restrict__ E, double *__restrict__ XH, int kc)
{
    __shared__ double sp[NLA][TY + 4][TX + 4];
    Chunk c = chunk_of(g, kc);
    if (c.kb >= c.ke) return;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const long q0 = (long)c.i + sy * c.j;
    // E on the planes outside the interior z range (first / last chunk)
    if (c.valid) {
        if (c.kb == 2) { e_point(X, cur, eta, S, q0, E); e_point(X, cur, eta, S, q0 + sz, E); }
        if (c.ke == g.NZ - 2) {
            e_point(X, cur, eta, S, q0 + sz * (g.NZ - 2), E);
            e_point(X, cur, eta, S, q0 + sz * (g.NZ - 1), E);
        }
    }
    // queues: z fields for planes k-2..k+2, x-y fields for planes k..k+2
    double qx4[5], qt24[5], qt34[5], qt44[5], qp[5];
    double yx2[3], yx3[3], yt22[3], yt23[3], yt33[3];
    if (c.valid) {
#pragma unroll
        for (int t = 0; t < 4; t++) {
            Rec r = rec_at(X, pri, S, q0 + sz * (c.kb - 2 + t));
            qx4[t] = r.x4; qt24[t] = r.t24; qt34[t] = r.t34; qt44[t] = r.t44; qp[t] = r.p;
            if (t >= 2) {
                yx2[t - 2] = r.x2; yx3[t - 2] = r.x3; yt22[t - 2] = r.t22; yt23[t - 2] = r.t23; yt33[t - 2] = r.t33;
            }
        }
    }
    const int hx = c.tx < 2 ? -2 : (c.tx >= TX - 2 ? 2 : 0);
    const int hy = c.ty < 2 ? -2 : (c.ty >= TY - 2 ? 2 : 0);
    const bool hxok = hx != 0 && c.j < g.NY && c.i + hx >= 0 && c.i + hx < g.NX;
    const bool hyok = hy != 0 && c.i < g.NX && c.j + hy >= 0 && c.j + hy < g.NY;
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
    if (c.inner) {
        ax[0] = g.x1a[c.i]; ax[1] = g.x1b[c.i]; ax[2] = g.x1c[c.i]; ax[3] = g.x1d[c.i];
        ay[0] = g.y1a[c.j]; ay[1] = g.y1b[c.j]; ay[2] = g.y1c[c.j]; ay[3] = g.y1d[c.j];
    }
    for (int k = c.kb; k < c.ke; k++) {
        const long qk = q0 + sz * k;
        if (c.valid) {
            Rec r = rec_at(X, pri, S, qk + 2 * sz);
            qx4[4] = r.x4; qt24[4] = r.t24; qt34[4] = r.t34; qt44[4] = r.t44; qp[4] = r.p;
            yx2[2] = r.x2; yx3[2] = r.x3; yt22[2] = r.t22; yt23[2] = r.t23; yt33[2] = r.t33;
            const int y = c.ty + 2, x = c.tx + 2;
            sp[LX2][y][x] = yx2[0]; sp[LX3][y][x] = yx3[0]; sp[LT22][y][x] = yt22[0]; sp[LT23][y][x] = yt23[0];
            sp[LT24][y][x] = qt24[2]; sp[LT33][y][x] = yt33[0]; sp[LT34][y][x] = qt34[2]; sp[LP][y][x] = qp[2];
            e_point(X, cur, eta, S, qk, E);
        }
        if (hxok) rec_to_lds(sp, c.ty + 2, c.tx + 2 + hx, rec_at(X, pri, S, qk + hx));
        if (hyok) rec_to_lds(sp, c.ty + 2 + hy, c.tx + 2, rec_at(X, pri, S, qk + hy * sy));
        __syncthreads();
        if (c.inner) {
            const int X0 = c.tx + 2, Y0 = c.ty + 2;
            double az[4] = {g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]};
#define DXL(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], ax[0], ax[1], ax[2], ax[3])
#define DYL(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], ay[0], ay[1], ay[2], ay[3])
#define DZQ(a) fd5(a[0], a[1], a[2], a[3], a[4], az[0], az[1], az[2], az[3])
            double xd0 = -DXL(LX2) - DYL(LX3) - DZQ(qx4);
            double xd1 = -DXL(LT22) - DYL(LT23) - DZQ(qt24);
            double xd2 = -DXL(LT23) - DYL(LT33) - DZQ(qt34);
            double xd3 = -DXL(LT24) - DYL(LT34) - DZQ(qt44);
            double x1 = X[qk], x5 = X[qk + 4 * S], x6 = X[qk + 5 * S], x7 = X[qk + 6 * S];
            double c1 = cur[qk], c2 = cur[qk + S], c3 = cur[qk + 2 * S];
            xd1 = xd1 - xresd[qk] + c2 * x7 - c3 * x6;
            xd2 = xd2 - xresd[qk + S] + c3 * x5 - c1 * x7;
            xd3 = xd3 - xresd[qk + 2 * S] + c1 * x6 - c2 * x5;
            XD[qk] = xd0; XD[qk + S] = xd1; XD[qk + 2 * S] = xd2; XD[qk + 3 * S] = xd3;
            if (hall) {
                XH[qk] = cht * (c2 * x7 - c3 * x6 - DXL(LP)) / x1;
                XH[qk + S] = cht * (c3 * x5 - c1 * x7 - DYL(LP)) / x1;
                XH[qk + 2 * S] = cht * (c1 * x6 - c2 * x5 - DZQ(qp)) / x1;
            }
#undef DZQ
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; t++) {
            qx4[t] = qx4[t + 1]; qt24[t] = qt24[t + 1]; qt34[t] = qt34[t + 1]; qt44[t] = qt44[t + 1]; qp[t] = qp[t + 1];
        }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            yx2[t] = yx2[t + 1]; yx3[t] = yx3[t + 1]; yt22[t] = yt22[t + 1]; yt23[t] = yt23[t + 1]; yt33[t] = yt33[t + 1];
        }
    }
}

// ---------------------------------------------------------------------------
// right, part 2 as a z-marching tile: -curl(E + xdh) and the pressure
// equation (t3dsdsc.f:1426-1566), optional viscosity/diffusion (1585-1663).
// Per point e = E + xdh and v = m / rho are formed once.
// LDS: {e1,e2,e3,p,vx,vy[,vz,rho]}; queue: {e1,e2,p,vz[,vx,vy,rho]}.
// ---------------------------------------------------------------------------
enum { LE1 = 0, LE2, LE3, LPP, LVX, LVY, LVZ, LRH };

template <bool VD>
__global__ __launch_bounds__(NT) void k_rhs_b25(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                                                const double *__restrict__ E, const double *__restrict__ XH,
                                                double eta, double gam, double fmu, double frho, double fpr,
                                                double *__restrict__ XD, int kc)
{
    constexpr int NL = VD ? 8 : 6;
    __shared__ double sp[NL][TY + 4][TX + 4];
    Chunk c = chunk_of(g, kc);
    if (c.kb >= c.ke) return;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const long q0 = (long)c.i + sy * c.j;
    auto ept = [&](long q, int m) { return XH ? E[q + m * S] + XH[q + m * S] : E[q + m * S]; };
    double qe1[5], qe2[5], qp[5], qvz[5], qvx[5], qvy[5], qrh[5];
    auto load_q = [&](int t, long q) {
        double r = X[q];
        qe1[t] = ept(q, 0); qe2[t] = ept(q, 1); qp[t] = X[q + 7 * S]; qvz[t] = X[q + 3 * S] / r;
        if (VD) { qvx[t] = X[q + S] / r; qvy[t] = X[q + 2 * S] / r; qrh[t] = r; }
    };
    auto to_lds = [&](int y, int x, long q) {
        double r = X[q];
        sp[LE1][y][x] = ept(q, 0); sp[LE2][y][x] = ept(q, 1); sp[LE3][y][x] = ept(q, 2);
        sp[LPP][y][x] = X[q + 7 * S]; sp[LVX][y][x] = X[q + S] / r; sp[LVY][y][x] = X[q + 2 * S] / r;
        if (VD) { sp[LVZ][y][x] = X[q + 3 * S] / r; sp[LRH][y][x] = r; }
    };
    if (c.valid) {
#pragma unroll
        for (int t = 0; t < 4; t++) load_q(t, q0 + sz * (c.kb - 2 + t));
    }
    const int hx = c.tx < 2 ? -2 : (c.tx >= TX - 2 ? 2 : 0);
    const int hy = c.ty < 2 ? -2 : (c.ty >= TY - 2 ? 2 : 0);
    const bool hxok = hx != 0 && c.j < g.NY && c.i + hx >= 0 && c.i + hx < g.NX;
    const bool hyok = hy != 0 && c.i < g.NX && c.j + hy >= 0 && c.j + hy < g.NY;
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0}, bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
    if (c.inner) {
        ax[0] = g.x1a[c.i]; ax[1] = g.x1b[c.i]; ax[2] = g.x1c[c.i]; ax[3] = g.x1d[c.i];
        ay[0] = g.y1a[c.j]; ay[1] = g.y1b[c.j]; ay[2] = g.y1c[c.j]; ay[3] = g.y1d[c.j];
        if (VD) {
            bx[0] = g.x2a[c.i]; bx[1] = g.x2b[c.i]; bx[2] = g.x2c[c.i]; bx[3] = g.x2d[c.i];
            by[0] = g.y2a[c.j]; by[1] = g.y2b[c.j]; by[2] = g.y2c[c.j]; by[3] = g.y2d[c.j];
        }
    }
    for (int k = c.kb; k < c.ke; k++) {
        const long qk = q0 + sz * k;
        if (c.valid) {
            load_q(4, qk + 2 * sz);
            to_lds(c.ty + 2, c.tx + 2, qk);
        }
        if (hxok) to_lds(c.ty + 2, c.tx + 2 + hx, qk + hx);
        if (hyok) to_lds(c.ty + 2 + hy, c.tx + 2, qk + hy * sy);
        __syncthreads();
        if (c.inner) {
            const int X0 = c.tx + 2, Y0 = c.ty + 2;
            double az[4] = {g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]};
#define DZQ(a) fd5(a[0], a[1], a[2], a[3], a[4], az[0], az[1], az[2], az[3])
            double xd5 = -DYL(LE3) + DZQ(qe2);
            double xd6 = -DZQ(qe1) + DXL(LE3);
            double xd7 = -DXL(LE2) + DYL(LE1);
            double fs0 = sp[LVX][Y0][X0], gs0 = sp[LVY][Y0][X0], hs0 = qvz[2], p0 = qp[2];
            double dvx = DXL(LVX), dvy = DYL(LVY), dvz = DZQ(qvz);
            double c1 = cur[qk], c2 = cur[qk + S], c3 = cur[qk + 2 * S];
            double xd8 = -DXL(LPP) * fs0 - DYL(LPP) * gs0 - DZQ(qp) * hs0 + (-dvx - dvy - dvz) * gam * p0 +
                         (gam - 1.) * eta * (c1 * c1 + c2 * c2 + c3 * c3);
#undef DZQ
            if (VD) {
                double bz[4] = {g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]};
#define D2X(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], bx[0], bx[1], bx[2], bx[3])
#define D2Y(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], by[0], by[1], by[2], by[3])
#define D2Z(a) fd5(a[0], a[1], a[2], a[3], a[4], bz[0], bz[1], bz[2], bz[3])
                if (fmu != 0.0) {
                    double af = D2X(LVX) + D2Y(LVX) + D2Z(qvx);
                    double ag = D2X(LVY) + D2Y(LVY) + D2Z(qvy);
                    double ah = D2X(LVZ) + D2Y(LVZ) + D2Z(qvz);
                    XD[qk + S] = XD[qk + S] + fmu * af;
                    XD[qk + 2 * S] = XD[qk + 2 * S] + fmu * ag;
                    XD[qk + 3 * S] = XD[qk + 3 * S] + fmu * ah;
                    xd8 = xd8 + fmu * (fs0 * af + gs0 * ag + hs0 * ah);
                }
                if (frho != 0.0 || fpr != 0.0) {
                    XD[qk] = XD[qk] + frho * (D2X(LRH) + D2Y(LRH) + D2Z(qrh));
                    xd8 = xd8 + fpr * (D2X(LPP) + D2Y(LPP) + D2Z(qp));
                }
#undef D2X
#undef D2Y
#undef D2Z
            }
            XD[qk + 4 * S] = xd5;
            XD[qk + 5 * S] = xd6;
            XD[qk + 6 * S] = xd7;
            XD[qk + 7 * S] = xd8;
        }
        __syncthreads();
#pragma unroll
        for (int t = 0; t < 4; t++) {
            qe1[t] = qe1[t + 1]; qe2[t] = qe2[t + 1]; qp[t] = qp[t + 1]; qvz[t] = qvz[t + 1];
            if (VD) { qvx[t] = qvx[t + 1]; qvy[t] = qvy[t + 1]; qrh[t] = qrh[t + 1]; }
        }
    }
}
#undef DXL
#undef DYL

// ---------------------------------------------------------------------------
// boundary fills: bndryc/bndryE/bndryh (BC_AUX, t3dsdsc.f:2539-3027),
// bndry (BC_STATE, t3dsdsc.f:1677-1985 with the x-right rule mirrored on the
// left), bndrytt (BC_TT, t3dsdsc.f:1989-2284).  The reference applies the x
// rule, then the y rule, then the z rule, each reading the results of the
// previous; here each face point evaluates that composition directly.
// ---------------------------------------------------------------------------
template <int TYPE>
struct BCEval {
    const Geo &g;
    const double *G;     // swept values (interior valid)
    const double *A;     // face values before this fill (BC_TT), or nullptr
    const double *X;     // state field when A is x - eq on the fly
    const double *E;     // its equilibrium (or nullptr)

    __device__ double P(int i, int j, int k) const
    {
        long q = IDX(g, i, j, k);
        if (interior(g, i, j, k)) return G[q];
        if (A) return A[q];
        if (X) return E ? X[q] - E[q] : X[q];
        return G[q];
    }
    __device__ double Xv(int i, int j, int k) const
    {
        if (i >= 2 && i <= g.NX - 3) return P(i, j, k);
        if (i == g.NX - 1) return 0.0;
        if (i == g.NX - 2) return 0.5 * P(g.NX - 3, j, k);
        if (TYPE == BC_STATE) return i == 0 ? 0.0 : 0.5 * P(2, j, k);
        double x1 = (g.ax4 * P(2, j, k) + g.bx4 * P(3, j, k) + g.cx4 * P(4, j, k)) / (g.ax4 + g.bx4 + g.cx4);
        if (i == 1) return x1;
        return (g.ax3 * x1 + g.bx3 * P(2, j, k) + g.cx3 * P(3, j, k)) / (g.ax3 + g.bx3 + g.cx3);
    }
    __device__ double Yv(int i, int j, int k) const
    {
        if (j >= 2 && j <= g.NY - 3) return Xv(i, j, k);
        if (TYPE == BC_TT && !g.periodic && i < 2) return Xv(i, j, k);
        if (g.periodic) {
            int jj = j == 0 ? g.NY - 4 : j == 1 ? g.NY - 3 : j == g.NY - 2 ? 2 : 3;
            return Xv(i, jj, k);
        }
        if (j == 0 || j == g.NY - 1) return 0.0;
        if (j == 1) return 0.5 * Xv(i, 2, k);
        return 0.5 * Xv(i, g.NY - 3, k);
    }
    __device__ double Zv(int i, int j, int k) const
    {
        if (k >= 2 && k <= g.NZ - 3) return Yv(i, j, k);
        if (TYPE == BC_TT && i < 2) return Yv(i, j, k);
        if (k == 0 || k == g.NZ - 1) return 0.0;
        if (k == 1) return 0.5 * Yv(i, j, 2);
        return 0.5 * Yv(i, j, g.NZ - 3);
    }
};

// face point enumeration: x faces (4*NY*NZ), then y faces over interior i
// ((NX-4)*4*NZ), then z faces over interior i,j ((NX-4)*(NY-4)*4)
__device__ inline bool face_point(const Geo &g, long t, int &i, int &j, int &k)
{
    long nxf = 4L * g.NY * g.NZ;
    long nyf = (long)(g.NX - 4) * 4 * g.NZ;
    long nzf = (long)(g.NX - 4) * (g.NY - 4) * 4;
    if (t < nxf) {
        int f = (int)(t % 4); long r = t / 4;
        j = (int)(r % g.NY); k = (int)(r / g.NY);
        i = f < 2 ? f : g.NX - 4 + f;
        return true;
    }
    t -= nxf;
    if (t < nyf) {
        i = 2 + (int)(t % (g.NX - 4)); long r = t / (g.NX - 4);
        int f = (int)(r % 4); k = (int)(r / 4);
        j = f < 2 ? f : g.NY - 4 + f;
        return true;
    }
    t -= nyf;
    if (t < nzf) {
        i = 2 + (int)(t % (g.NX - 4)); long r = t / (g.NX - 4);
        j = 2 + (int)(r % (g.NY - 4)); int f = (int)(r / (g.NY - 4));
        k = f < 2 ? f : g.NZ - 4 + f;
        return true;
    }
    return false;
}

template <int TYPE, int NF>
__global__ void k_bc(Geo g, double *B, const double *A, const double *X, Eq eq)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!face_point(g, t, i, j, k)) return;
    long q = IDX(g, i, j, k);
#pragma unroll
    for (int m = 0; m < NF; m++) {
        BCEval<TYPE> ev{g, B + m * g.S, A ? A + m * g.S : nullptr, X ? X + m * g.S : nullptr,
                        X ? eq.f[m] : nullptr};
        B[q + m * g.S] = ev.Zv(i, j, k);
    }
}

// ---------------------------------------------------------------------------
// right, part 1 (t3dsdsc.f:1332-1422): divergence of the mass and momentum
// fluxes (interior), + J x B - xresd, E = -v x B + eta J (whole extended
// range), Hall term xdh (interior)
// ---------------------------------------------------------------------------
__device__ inline void flux_at(int m, const double *X, const double *pri, long S, long p, double &f,
                               double &gg, double &h)
{
    double x1 = X[p], x2 = X[p + S], x3 = X[p + 2 * S], x4 = X[p + 3 * S];
    if (m == 1) { f = x2; gg = x3; h = x4; return; }
    double dp = X[p + 7 * S] - pri[p];
    if (m == 2) { f = x2 * x2 / x1 + dp; gg = x2 * x3 / x1; h = x2 * x4 / x1; }
    else if (m == 3) { f = x2 * x3 / x1; gg = x3 * x3 / x1 + dp; h = x4 * x3 / x1; }
    else { f = x2 * x4 / x1; gg = x3 * x4 / x1; h = x4 * x4 / x1 + dp; }
}

__device__ inline double flux_f(int m, const double *X, const double *pri, long S, long p)
{
    double f, gg, h; flux_at(m, X, pri, S, p, f, gg, h); return f;
}
__device__ inline double flux_g(int m, const double *X, const double *pri, long S, long p)
{
    double f, gg, h; flux_at(m, X, pri, S, p, f, gg, h); return gg;
}
__device__ inline double flux_h(int m, const double *X, const double *pri, long S, long p)
{
    double f, gg, h; flux_at(m, X, pri, S, p, f, gg, h); return h;
}

__global__ void k_rhs_a(Geo g, const double *__restrict__ X, const double *__restrict__ pri,
                        const double *__restrict__ cur, const double *__restrict__ xresd,
                        double eta, int hall, double cht, double *__restrict__ XD,
                        double *__restrict__ E, double *__restrict__ XH)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int j = blockIdx.y * blockDim.y + threadIdx.y;
    int k = blockIdx.z;
    if (i >= g.NX || j >= g.NY) return;
    long S = g.S, q = IDX(g, i, j, k), sy = g.sy, sz = g.sz;
    double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
    double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
    double c1 = cur[q], c2 = cur[q + S], c3 = cur[q + 2 * S];
    // E over the extended range (t3dsdsc.f:1373-1378)
    E[q] = (-x3 * x7 + x4 * x6) / x1 + eta * c1;
    E[q + S] = (-x4 * x5 + x2 * x7) / x1 + eta * c2;
    E[q + 2 * S] = (-x2 * x6 + x3 * x5) / x1 + eta * c3;
    if (!interior(g, i, j, k)) return;
    double xd[4];
#pragma unroll
    for (int m = 1; m <= 4; m++) {
        double fx = fd5(flux_f(m, X, pri, S, q - 2), flux_f(m, X, pri, S, q - 1), flux_f(m, X, pri, S, q),
                        flux_f(m, X, pri, S, q + 1), flux_f(m, X, pri, S, q + 2), g.x1a[i], g.x1b[i],
                        g.x1c[i], g.x1d[i]);
        double gy = fd5(flux_g(m, X, pri, S, q - 2 * sy), flux_g(m, X, pri, S, q - sy),
                        flux_g(m, X, pri, S, q), flux_g(m, X, pri, S, q + sy),
                        flux_g(m, X, pri, S, q + 2 * sy), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
        double hz = fd5(flux_h(m, X, pri, S, q - 2 * sz), flux_h(m, X, pri, S, q - sz),
                        flux_h(m, X, pri, S, q), flux_h(m, X, pri, S, q + sz),
                        flux_h(m, X, pri, S, q + 2 * sz), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
        xd[m - 1] = -fx - gy - hz;
    }
    // t3dsdsc.f:1357-1362
    xd[1] = xd[1] - xresd[q] + c2 * x7 - c3 * x6;
    xd[2] = xd[2] - xresd[q + S] + c3 * x5 - c1 * x7;
    xd[3] = xd[3] - xresd[q + 2 * S] + c1 * x6 - c2 * x5;
    XD[q] = xd[0]; XD[q + S] = xd[1]; XD[q + 2 * S] = xd[2]; XD[q + 3 * S] = xd[3];
    if (hall) {   // t3dsdsc.f:1404-1422
        const double *p8 = X + 7 * S;
        XH[q] = cht * (c2 * x7 - c3 * x6 - dx1(g, p8, q, i)) / x1;
        XH[q + S] = cht * (c3 * x5 - c1 * x7 - dy1(g, p8, q, j)) / x1;
        XH[q + 2 * S] = cht * (c1 * x6 - c2 * x5 - dz1(g, p8, q, k)) / x1;
    }
}

// ---------------------------------------------------------------------------
// right, part 2 (t3dsdsc.f:1426-1663): -curl(E + xdh), pressure equation,
// optional viscosity / diffusion (interior)
// ---------------------------------------------------------------------------
__global__ void k_rhs_b(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                        const double *__restrict__ E, const double *__restrict__ XH, double eta,
                        double gam, double fmu, double frho, double fpr, double *__restrict__ XD)
{
    int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    int k = 2 + blockIdx.z;
    if (i > g.NX - 3 || j > g.NY - 3) return;
    long S = g.S, q = IDX(g, i, j, k), sy = g.sy, sz = g.sz;
    const double *e1 = E, *e2 = E + S, *e3 = E + 2 * S;
    const double *h1 = XH, *h2 = XH ? XH + S : nullptr, *h3 = XH ? XH + 2 * S : nullptr;
#define ET(e, h, o) ((h) ? (e)[q + (o)] + (h)[q + (o)] : (e)[q + (o)])
#define DXE(e, h) fd5(ET(e, h, -2), ET(e, h, -1), ET(e, h, 0), ET(e, h, 1), ET(e, h, 2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i])
#define DYE(e, h) fd5(ET(e, h, -2 * sy), ET(e, h, -sy), ET(e, h, 0), ET(e, h, sy), ET(e, h, 2 * sy), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j])
#define DZE(e, h) fd5(ET(e, h, -2 * sz), ET(e, h, -sz), ET(e, h, 0), ET(e, h, sz), ET(e, h, 2 * sz), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k])
    double xd5 = -DYE(e3, h3) + DZE(e2, h2);
    double xd6 = -DZE(e1, h1) + DXE(e3, h3);
    double xd7 = -DXE(e2, h2) + DYE(e1, h1);
#undef DXE
#undef DYE
#undef DZE
#undef ET
    const double *r = X, *m2 = X + S, *m3 = X + 2 * S, *m4 = X + 3 * S, *p8 = X + 7 * S;
#define VX(o) (m2[q + (o)] / r[q + (o)])
#define VY(o) (m3[q + (o)] / r[q + (o)])
#define VZ(o) (m4[q + (o)] / r[q + (o)])
    double fs0 = VX(0), gs0 = VY(0), hs0 = VZ(0);
    double dvx = fd5(VX(-2), VX(-1), fs0, VX(1), VX(2), g.x1a[i], g.x1b[i], g.x1c[i], g.x1d[i]);
    double dvy = fd5(VY(-2 * sy), VY(-sy), gs0, VY(sy), VY(2 * sy), g.y1a[j], g.y1b[j], g.y1c[j], g.y1d[j]);
    double dvz = fd5(VZ(-2 * sz), VZ(-sz), hs0, VZ(sz), VZ(2 * sz), g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]);
    double c1 = cur[q], c2 = cur[q + S], c3 = cur[q + 2 * S];
    double xd8 = -dx1(g, p8, q, i) * fs0 - dy1(g, p8, q, j) * gs0 - dz1(g, p8, q, k) * hs0 +
                 (-dvx - dvy - dvz) * gam * p8[q] + (gam - 1.) * eta * (c1 * c1 + c2 * c2 + c3 * c3);
    if (fmu != 0.0) {   // t3dsdsc.f:1585-1621
        double af = fd5(VX(-2), VX(-1), fs0, VX(1), VX(2), g.x2a[i], g.x2b[i], g.x2c[i], g.x2d[i]) +
                    fd5(VX(-2 * sy), VX(-sy), fs0, VX(sy), VX(2 * sy), g.y2a[j], g.y2b[j], g.y2c[j], g.y2d[j]) +
                    fd5(VX(-2 * sz), VX(-sz), fs0, VX(sz), VX(2 * sz), g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]);
        double ag = fd5(VY(-2), VY(-1), gs0, VY(1), VY(2), g.x2a[i], g.x2b[i], g.x2c[i], g.x2d[i]) +
                    fd5(VY(-2 * sy), VY(-sy), gs0, VY(sy), VY(2 * sy), g.y2a[j], g.y2b[j], g.y2c[j], g.y2d[j]) +
                    fd5(VY(-2 * sz), VY(-sz), gs0, VY(sz), VY(2 * sz), g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]);
        double ah = fd5(VZ(-2), VZ(-1), hs0, VZ(1), VZ(2), g.x2a[i], g.x2b[i], g.x2c[i], g.x2d[i]) +
                    fd5(VZ(-2 * sy), VZ(-sy), hs0, VZ(sy), VZ(2 * sy), g.y2a[j], g.y2b[j], g.y2c[j], g.y2d[j]) +
                    fd5(VZ(-2 * sz), VZ(-sz), hs0, VZ(sz), VZ(2 * sz), g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]);
        XD[q + S] = XD[q + S] + fmu * af;
        XD[q + 2 * S] = XD[q + 2 * S] + fmu * ag;
        XD[q + 3 * S] = XD[q + 3 * S] + fmu * ah;
        xd8 = xd8 + fmu * (fs0 * af + gs0 * ag + hs0 * ah);
    }
#undef VX
#undef VY
#undef VZ
    if (frho != 0.0 || fpr != 0.0) {   // t3dsdsc.f:1623-1663
        XD[q] = XD[q] + frho * (dx2(g, r, q, i) + dy2(g, r, q, j) + dz2(g, r, q, k));
        xd8 = xd8 + fpr * (dx2(g, p8, q, i) + dy2(g, p8, q, j) + dz2(g, p8, q, k));
    }
    XD[q + 4 * S] = xd5;
    XD[q + 5 * S] = xd6;
    XD[q + 6 * S] = xd7;
    XD[q + 7 * S] = xd8;
}

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
        XM[q] = XF[q] + d * dt / 6.;
        X[q] = XF[q] + d * dt / 2.;
    } else if (stage == 1) {
        XM[q] = XM[q] + d * dt / 3.;
        X[q] = XF[q] + d * dt / 2.;
    } else if (stage == 2) {
        XM[q] = XM[q] + d * dt / 3.;
        X[q] = XF[q] + d * dt;
    } else {
        X[q] = XM[q] + d * dt / 6.;
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
        X[q + m * S] = eq.f[m] ? v + eq.f[m][q] : v;
    }
}

__global__ void k_copy3(long S, long n, const double *__restrict__ A, double *__restrict__ B)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= n) return;
    B[q] = A[q]; B[q + S] = A[q + S]; B[q + 2 * S] = A[q + 2 * S];
}

__global__ void k_set(double *p, double v) { *p = v; }

// plain streaming copy (reference rate for the roofline figures)
__global__ void k_copy(long n, const double2 *__restrict__ a, double2 *__restrict__ b)
{
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) b[t] = a[t];
}

// ---------------------------------------------------------------------------
// setdt, t3dsdsc.f:4709-4808: minimum over jz = 3..mz, jy = 1..my-1,
// jx = 3..mx  (k = 0..NZ-3, j = 0..NY-2, i = 0..NX-3)
// ---------------------------------------------------------------------------
__global__ void k_dt_partial(Geo g, const double *__restrict__ X, double eta, double fmu, double cfl,
                             double gam, double wcoef, double *__restrict__ part)
{
    __shared__ double sm[256];
    long ni = g.NX - 2, nj = g.NY - 1, nk = g.NZ - 2;
    long n = ni * nj * nk;
    double best = 100.;
    for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
        int i = (int)(t % ni); long r = t / ni;
        int j = (int)(r % nj); int k = (int)(r / nj);
        long q = IDX(g, i, j, k), S = g.S;
        double x1 = X[q];
        double vx = X[q + S] / x1, vy = X[q + 2 * S] / x1, vz = X[q + 3 * S] / x1;
        double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
        double b2 = x5 * x5 + x6 * x6 + x7 * x7;
        double va2 = b2 / x1;
        double cs2 = gam * X[q + 7 * S] / x1;
        double vf = sqrt(fabs(va2 + cs2));
        double vpx = fabs(vx) + vf, vpy = fabs(vy) + vf, vpz = fabs(vz) + vf;
        double ddx = g.dxs[i], ddy = g.dys[j], ddz = g.dzs[k];
        if (wcoef != 0.0) {
            double wb = wcoef * sqrt(b2) / x1;
            vpx = vpx + wb / fabs(ddx);
            vpy = vpy + wb / fabs(ddy);
            vpz = vpz + wb / fabs(ddz);
        }
        double dtx = fabs(ddx) / (vpx / cfl + (eta + fmu) / ddx / 2.);
        double dty = fabs(ddy) / (vpy / cfl + (eta + fmu) / ddy / 2.);
        double dtz = fabs(ddz) / (vpz / cfl + (eta + fmu) / ddz / 2.);
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
    if (threadIdx.x == 0) part[blockIdx.x] = sm[0];
}

// final MIN; also keeps the running time and the collapse flag on device:
// dtv[0] = dt, dtv[1] = time (advanced by the previous dt), dtv[2] = dt0,
// dtv[3] = collapse flag, dtv[4] = previous dt
__global__ void k_dt_final(const double *__restrict__ part, int n, double *dtv, int advance_time)
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
        double dt = sm[0];
        if (advance_time) dtv[1] = dtv[1] + dtv[4];
        dtv[0] = dt;
        dtv[4] = dt;
        if (dtv[2] > 0.0 && dt < 0.1 * dtv[2]) dtv[3] = 1.0;
    }
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
    double wyz = wy[j] * wz[k];
    (void)wyz;
    // lane t accumulates i = t, t+64, ... in order (fixed per lane)
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

// div B interior (t3dsdsc.f:2520-2528)
__global__ void k_divb(Geo g, const double *__restrict__ X, double *__restrict__ dv)
{
    int i = 2 + blockIdx.x * blockDim.x + threadIdx.x;
    int j = 2 + blockIdx.y * blockDim.y + threadIdx.y;
    int k = 2 + blockIdx.z;
    if (i > g.NX - 3 || j > g.NY - 3) return;
    long q = IDX(g, i, j, k), S = g.S;
    dv[q] = dx1(g, X + 4 * S, q, i) + dy1(g, X + 5 * S, q, j) + dz1(g, X + 6 * S, q, k);
}

// bndrydiv (t3dsdsc.f:3031-3712) as sequential face passes (x, y, z);
// pass 0: x faces for all j,k; pass 1: y rows for the x range; pass 2: z
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
        if (t >= n) return;
        int i = (int)(t % g.NX), k = (int)(t / g.NX);
        long sy = g.sy, b = IDX(g, i, 0, k);
        int M = g.NY - 1;
        double *f = dv + b;
        if (g.periodic) {
            f[0] = f[(M - 3) * sy]; f[sy] = f[(M - 2) * sy];
            f[(M - 1) * sy] = f[2 * sy]; f[M * sy] = f[3 * sy];
        } else {
            double aa1 = 5. / 2., bb1 = -2., cc1 = 1. / 2.;
            f[sy] = aa1 * f[2 * sy] + bb1 * f[3 * sy] + cc1 * f[4 * sy];
            f[(M - 1) * sy] = aa1 * f[(M - 2) * sy] + bb1 * f[(M - 3) * sy] + cc1 * f[(M - 4) * sy];
            f[0] = aa1 * f[sy] + bb1 * f[2 * sy] + cc1 * f[3 * sy];
            f[M * sy] = aa1 * f[(M - 1) * sy] + bb1 * f[(M - 2) * sy] + cc1 * f[(M - 3) * sy];
        }
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

struct mhd_handle {
    mhd_config cfg;
    int NX, NY, NZ;
    long S, N;
    hipStream_t stream = nullptr;
    Geo g;
    double *geo = nullptr;                 // all 1-D device arrays
    double *x = nullptr, *xfold = nullptr, *xm = nullptr, *xdA = nullptr, *xdB = nullptr;
    double *eq = nullptr;                  // rhoi, bxi, byi, bzi, pri
    double *xresd = nullptr;
    double *cA = nullptr, *cB = nullptr, *eA = nullptr, *eB = nullptr, *hA = nullptr, *hB = nullptr;
    double *xd = nullptr;                  // buffer holding the last xdif
    double *cur = nullptr;                 // buffer holding the last current
    double *red = nullptr;                 // reduction scratch
    double *dtv = nullptr;                 // device dt slots
    int nred = 0;
    std::vector<double> wx, wy, wz;        // energy weights (host copy)
    double *dwx = nullptr, *dwy = nullptr, *dwz = nullptr;
    std::string err;
    bool have_eq = false;
    int kc = 16;                           // z planes per tile
    bool naive = false;                    // one-point-per-thread kernels (A/B checks)
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

enum { PT_CURL, PT_FILT3, PT_FILT8, PT_BC, PT_RHSA, PT_RHSB, PT_RK, PT_POST, PT_DT, PT_N };
static const char *PT_NAMES[PT_N] = {"curl", "smooth3", "smooth8", "boundary", "rhs_flux_E_hall",
                                     "rhs_curlE_pressure", "rk_update", "post_add_eq", "setdt"};

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

static long n_faces(const mhd_handle *h)
{
    return 4L * h->NY * h->NZ + (long)(h->NX - 4) * 4 * h->NZ + (long)(h->NX - 4) * (h->NY - 4) * 4;
}

static void launch_bc(mhd_handle *h, int type, int nf, double *B, const double *A, const double *X, Eq eq)
{
    long nf_pts = n_faces(h);
    int bs = 256;
    dim3 gr(blocks(nf_pts, bs));
    PROF_BEGIN(PT_BC);
    if (type == BC_AUX)
        hipLaunchKernelGGL((k_bc<BC_AUX, 3>), gr, dim3(bs), 0, h->stream, h->g, B, A, X, eq);
    else if (type == BC_STATE)
        hipLaunchKernelGGL((k_bc<BC_STATE, 8>), gr, dim3(bs), 0, h->stream, h->g, B, A, X, eq);
    else
        hipLaunchKernelGGL((k_bc<BC_TT, 8>), gr, dim3(bs), 0, h->stream, h->g, B, A, X, eq);
    PROF_END(PT_BC);
}

static const dim3 BLK(64, 4, 1);

static dim3 tile_grid(const mhd_handle *h)
{
    return dim3((unsigned)blocks(h->NX, TX), (unsigned)blocks(h->NY, TY), (unsigned)blocks(h->NZ - 4, h->kc));
}

// kk smoothing passes on a 3-field array with its aux boundary fill;
// returns the buffer holding the result
static double *smooth_aux(mhd_handle *h, double *a, double *b, int kind, int kk)
{
    Eq none = eq_state(h, false);
    for (int p = 0; p < kk; p++) {
        PROF_BEGIN(PT_FILT3);
        if (h->naive) {
            if (kind == KIND_RADIAL)
                hipLaunchKernelGGL((k_filter<3, KIND_RADIAL, SRC_BUF>), int_grid(h, BLK), BLK, 0, h->stream, h->g, a, none, b);
            else
                hipLaunchKernelGGL((k_filter<3, KIND_EIGHTH, SRC_BUF>), int_grid(h, BLK), BLK, 0, h->stream, h->g, a, none, b);
        } else if (kind == KIND_RADIAL) {
            hipLaunchKernelGGL((k_filter25<3, KIND_RADIAL, SRC_BUF>), tile_grid(h), dim3(NT), 0, h->stream, h->g, a, none, b,
                               h->kc);
        } else {
            hipLaunchKernelGGL((k_filter25<3, KIND_EIGHTH, SRC_BUF>), tile_grid(h), dim3(NT), 0, h->stream, h->g, a, none, b,
                               h->kc);
        }
        PROF_END(PT_FILT3);
        launch_bc(h, BC_AUX, 3, b, nullptr, nullptr, none);
        std::swap(a, b);
    }
    return a;
}

// current (t3dsdsc.f:2452-2500) of state X into h->cur
static void current(mhd_handle *h, const double *X)
{
    Eq none = eq_state(h, false);
    PROF_BEGIN(PT_CURL);
    if (h->naive)
        hipLaunchKernelGGL(k_curl, int_grid(h, BLK), BLK, 0, h->stream, h->g, X, h->eq + h->S, h->eq + 2 * h->S,
                           h->eq + 3 * h->S, h->cA);
    else
        hipLaunchKernelGGL(k_curl25, tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->eq + h->S, h->eq + 2 * h->S,
                           h->eq + 3 * h->S, h->cA, h->kc);
    PROF_END(PT_CURL);
    launch_bc(h, BC_AUX, 3, h->cA, nullptr, nullptr, none);
    h->cur = smooth_aux(h, h->cA, h->cB, KIND_RADIAL, 4);
}

// right (t3dsdsc.f:1297-1673) of state X; result in h->xd
static void rhs(mhd_handle *h, const double *X, bool smooth_xdif, bool init)
{
    const mhd_config &c = h->cfg;
    double eta = init ? 0.0 : c.eta0, fmu = init ? 0.0 : c.fmu0;
    double frho = init ? 0.0 : c.frho0, fpr = init ? 0.0 : c.fpr0;
    Eq none = eq_state(h, false);
    current(h, X);
    dim3 gext((unsigned)blocks(h->NX, BLK.x), (unsigned)blocks(h->NY, BLK.y), (unsigned)h->NZ);
    PROF_BEGIN(PT_RHSA);
    if (h->naive)
        hipLaunchKernelGGL(k_rhs_a, gext, BLK, 0, h->stream, h->g, X, h->eq + 4 * h->S, h->cur, h->xresd, eta,
                           c.hall, c.di, h->xdA, h->eA, h->hA);
    else
        hipLaunchKernelGGL(k_rhs_a25, tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->eq + 4 * h->S, h->cur, h->xresd,
                           eta, c.hall, c.di, h->xdA, h->eA, h->hA, h->kc);
    PROF_END(PT_RHSA);
    double *E = smooth_aux(h, h->eA, h->eB, KIND_RADIAL, 4);
    double *XH = nullptr;
    if (c.hall) {
        launch_bc(h, BC_AUX, 3, h->hA, nullptr, nullptr, none);
        XH = smooth_aux(h, h->hA, h->hB, KIND_EIGHTH, 5);
    }
    PROF_BEGIN(PT_RHSB);
    bool vd = fmu != 0.0 || frho != 0.0 || fpr != 0.0;
    if (h->naive)
        hipLaunchKernelGGL(k_rhs_b, int_grid(h, BLK), BLK, 0, h->stream, h->g, X, h->cur, E, XH, eta, c.gamma, fmu,
                           frho, fpr, h->xdA);
    else if (vd)
        hipLaunchKernelGGL((k_rhs_b25<true>), tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->cur, E, XH, eta, c.gamma,
                           fmu, frho, fpr, h->xdA, h->kc);
    else
        hipLaunchKernelGGL((k_rhs_b25<false>), tile_grid(h), dim3(NT), 0, h->stream, h->g, X, h->cur, E, XH, eta,
                           c.gamma, fmu, frho, fpr, h->xdA, h->kc);
    PROF_END(PT_RHSB);
    launch_bc(h, BC_STATE, 8, h->xdA, nullptr, nullptr, none);
    double *a = h->xdA, *b = h->xdB;
    if (smooth_xdif) {
        for (int p = 0; p < 2; p++) {
            PROF_BEGIN(PT_FILT8);
            if (h->naive)
                hipLaunchKernelGGL((k_filter<8, KIND_STATE, SRC_BUF>), int_grid(h, BLK), BLK, 0, h->stream, h->g, a,
                                   none, b);
            else
                hipLaunchKernelGGL((k_filter25<8, KIND_STATE, SRC_BUF>), tile_grid(h), dim3(NT), 0, h->stream, h->g, a,
                                   none, b, h->kc);
            PROF_END(PT_FILT8);
            launch_bc(h, BC_STATE, 8, b, nullptr, nullptr, none);
            std::swap(a, b);
        }
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

// stepon body (t3dsdsc.f:1079-1165) with dt in h->dtv[0]
static void step_body(mhd_handle *h)
{
    std::swap(h->x, h->xfold);   // xfold = x (1079-1084) by renaming
    rhs(h, h->xfold, true, false);
    rk(h, 0);
    for (int s = 1; s < 4; s++) {
        rhs(h, h->x, true, false);
        rk(h, s);
    }
    // perturbation smoothing (1135-1165): xm = x - eq; avrgtt x2; x = xm + eq
    Eq sub = eq_state(h, true);
    PROF_BEGIN(PT_FILT8);
    if (h->naive)
        hipLaunchKernelGGL((k_filter<8, KIND_STATE, SRC_XEQ>), int_grid(h, BLK), BLK, 0, h->stream, h->g, h->x, sub,
                           h->xm);
    else
        hipLaunchKernelGGL((k_filter25<8, KIND_STATE, SRC_XEQ>), tile_grid(h), dim3(NT), 0, h->stream, h->g, h->x, sub,
                           h->xm, h->kc);
    PROF_END(PT_FILT8);
    launch_bc(h, BC_TT, 8, h->xm, nullptr, h->x, sub);
    Eq none = eq_state(h, false);
    PROF_BEGIN(PT_FILT8);
    if (h->naive)
        hipLaunchKernelGGL((k_filter<8, KIND_STATE, SRC_BUF>), int_grid(h, BLK), BLK, 0, h->stream, h->g, h->xm, none,
                           h->xdB);
    else
        hipLaunchKernelGGL((k_filter25<8, KIND_STATE, SRC_BUF>), tile_grid(h), dim3(NT), 0, h->stream, h->g, h->xm, none,
                           h->xdB, h->kc);
    PROF_END(PT_FILT8);
    launch_bc(h, BC_TT, 8, h->xdB, h->xm, nullptr, none);
    PROF_BEGIN(PT_POST);
    hipLaunchKernelGGL(k_add_eq, dim3(blocks(h->S, 256)), dim3(256), 0, h->stream, h->S, h->S, h->xdB, sub, h->x);
    PROF_END(PT_POST);
}

static void dt_kernels(mhd_handle *h, int advance_time)
{
    const mhd_config &c = h->cfg;
    double wcoef = (c.whistler_dt && c.hall && c.di != 0.0) ? 3.141592654 * c.di : 0.0;
    PROF_BEGIN(PT_DT);
    hipLaunchKernelGGL(k_dt_partial, dim3(h->nred), dim3(256), 0, h->stream, h->g, h->x, c.eta0, c.fmu0, c.cfl,
                       c.gamma, wcoef, h->red);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(256), 0, h->stream, h->red, h->nred, h->dtv, advance_time);
    PROF_END(PT_DT);
}

// ---- host <-> device layout conversion ----
static int upload_fields(mhd_handle *h, double *dst, const double *src, int nf)
{
    long px = h->cfg.mx + 4, plane = px * h->NY;
    for (int m = 0; m < nf; m++) {
        const double *s = src + (long)m * plane * (h->cfg.mz + 4) + 2 * plane + 2;
        HCHK(hipMemcpy2DAsync(dst + m * h->S, h->NX * sizeof(double), s, px * sizeof(double), h->NX * sizeof(double),
                              (size_t)h->NY * h->NZ, hipMemcpyHostToDevice, h->stream));
    }
    HCHK(hipStreamSynchronize(h->stream));
    return 0;
}

static int download_fields(mhd_handle *h, double *dst, const double *src, int nf)
{
    long px = h->cfg.mx + 4, plane = px * h->NY;
    long total = plane * (h->cfg.mz + 4);
    for (int m = 0; m < nf; m++) {
        double *d = dst + (long)m * total;
        memset(d, 0, sizeof(double) * total);
        HCHK(hipMemcpy2DAsync(d + 2 * plane + 2, px * sizeof(double), src + m * h->S, h->NX * sizeof(double),
                              h->NX * sizeof(double), (size_t)h->NY * h->NZ, hipMemcpyDeviceToHost, h->stream));
    }
    HCHK(hipStreamSynchronize(h->stream));
    return 0;
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
    if (!(cfg->gamma > 1.0) || !(cfg->cfl > 0.0)) return MHD_ERR_ARG;
    mhd_handle *h = new mhd_handle();
    h->cfg = *cfg;
    h->NX = cfg->mx; h->NY = cfg->my; h->NZ = cfg->mz;
    h->N = (long)h->NX * h->NY * h->NZ;
    h->S = (h->N + 63) / 64 * 64;
    HCHK(hipStreamCreate(&h->stream));

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
    size_t ngeo = cx.size() + cy.size() + cz.size() + 2 * (NX + NY + NZ) + NX + (NX + NY + NZ);
    HCHK(hipMalloc(&h->geo, ngeo * sizeof(double)));
    std::vector<double> hg;
    hg.reserve(ngeo);
    auto push = [&](const std::vector<double> &v) { size_t o = hg.size(); hg.insert(hg.end(), v.begin(), v.end()); return o; };
    size_t ocx = push(cx), ocy = push(cy), ocz = push(cz);
    size_t oX = push(X), oY = push(Y), oZ = push(Z), oDX = push(DXs), oDY = push(DYs), oDZ = push(DZs);
    size_t ocb = push(cbx);
    size_t owx = push(h->wx), owy = push(h->wy), owz = push(h->wz);
    HCHK(hipMemcpy(h->geo, hg.data(), hg.size() * sizeof(double), hipMemcpyHostToDevice));
    Geo &g = h->g;
    g.NX = NX; g.NY = NY; g.NZ = NZ; g.sy = NX; g.sz = (long)NX * NY; g.S = h->S; g.periodic = cfg->periodic_y;
    const double *G0 = h->geo;
    g.x1a = G0 + ocx; g.x1b = g.x1a + NX; g.x1c = g.x1b + NX; g.x1d = g.x1c + NX;
    g.x2a = g.x1d + NX; g.x2b = g.x2a + NX; g.x2c = g.x2b + NX; g.x2d = g.x2c + NX;
    g.xhp = g.x2d + NX; g.xhm = g.xhp + NX; g.xhs = g.xhm + NX; g.xwp = g.xhs + NX; g.xwm = g.xwp + NX;
    g.y1a = G0 + ocy; g.y1b = g.y1a + NY; g.y1c = g.y1b + NY; g.y1d = g.y1c + NY;
    g.y2a = g.y1d + NY; g.y2b = g.y2a + NY; g.y2c = g.y2b + NY; g.y2d = g.y2c + NY;
    g.yhp = g.y2d + NY; g.yhm = g.yhp + NY; g.yhs = g.yhm + NY; g.ywp = g.yhs + NY; g.ywm = g.ywp + NY;
    g.z1a = G0 + ocz; g.z1b = g.z1a + NZ; g.z1c = g.z1b + NZ; g.z1d = g.z1c + NZ;
    g.z2a = g.z1d + NZ; g.z2b = g.z2a + NZ; g.z2c = g.z2b + NZ; g.z2d = g.z2c + NZ;
    g.zhp = g.z2d + NZ; g.zhm = g.zhp + NZ; g.zhs = g.zhm + NZ; g.zwp = g.zhs + NZ; g.zwm = g.zwp + NZ;
    g.xc = G0 + oX; g.yc = G0 + oY; g.zc = G0 + oZ;
    g.dxs = G0 + oDX; g.dys = G0 + oDY; g.dzs = G0 + oDZ;
    g.cbx = G0 + ocb;
    h->dwx = h->geo + owx; h->dwy = h->geo + owy; h->dwz = h->geo + owz;
    coef_side(X.data(), 0, +1, &g.ax3, &g.bx3, &g.cx3);
    coef_side(X.data(), 1, +1, &g.ax4, &g.bx4, &g.cx4);
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
    h->nred = 1024;
    long nrows = (long)NY * NZ * 4;
    HCHK(hipMalloc(&h->red, std::max<long>(4096, nrows) * sizeof(double)));
    HCHK(hipMalloc(&h->dtv, 8 * sizeof(double)));
    HCHK(hipMemset(h->dtv, 0, 8 * sizeof(double)));
    h->xd = h->xdA;
    h->cur = h->cA;
    if (const char *e = getenv("MHD_KC")) h->kc = std::max(1, atoi(e));
    if (const char *e = getenv("MHD_NAIVE")) h->naive = atoi(e) != 0;
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
    rhs(h, h->x, false, true);
    HCHK(hipMemcpyAsync(h->xresd, h->xd + h->S, 3 * h->S * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    return MHD_OK;
}

int mhd_rhs(mhd_handle *h, double t)
{
    if (!h) return MHD_ERR_ARG;
    rhs(h, h->x, t != 0.0, false);
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    return MHD_OK;
}

int mhd_get_xdif(mhd_handle *h, double *xdif)
{
    if (!h || !xdif) return MHD_ERR_ARG;
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
    dt_kernels(h, 0);
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
    step_body(h);
    HCHK(hipStreamSynchronize(h->stream));
    HCHK(hipGetLastError());
    if (h->prof) return prof_collect(h);
    return MHD_OK;
}

int mhd_advance(mhd_handle *h, int nsteps, double dt0, double *time, double *dt_last, double *elapsed_ms)
{
    if (!h || nsteps < 0 || !time) return MHD_ERR_ARG;
    if (!h->have_eq) { set_err(h, "equilibrium not set"); return MHD_ERR_ARG; }
    double init[8] = {0, *time, dt0, 0, 0, 0, 0, 0};
    HCHK(hipMemcpyAsync(h->dtv, init, 8 * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipEvent_t e0, e1;
    HCHK(hipEventCreate(&e0));
    HCHK(hipEventCreate(&e1));
    HCHK(hipEventRecord(e0, h->stream));
    for (int s = 0; s < nsteps; s++) {
        dt_kernels(h, s > 0);
        step_body(h);
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
    *time = v[1] + (nsteps > 0 ? v[4] : 0.0);
    if (dt_last) *dt_last = v[4];
    if (elapsed_ms) *elapsed_ms = ms;
    if (v[3] != 0.0) { set_err(h, "dt collapsed below 0.1*dt0"); return MHD_ERR_DT_COLLAPSE; }
    return MHD_OK;
}

int mhd_get_current(mhd_handle *h, double *cur)
{
    if (!h || !cur) return MHD_ERR_ARG;
    current(h, h->x);
    return download_fields(h, cur, h->cur, 3);
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
    current(h, h->x);
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
    long n2 = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n2 * 16) != hipSuccess) return MHD_ERR_NOMEM;
    if (hipMalloc(&b, n2 * 16) != hipSuccess) { (void)hipFree(a); return MHD_ERR_NOMEM; }
    hipStream_t st;
    if (hipStreamCreate(&st) != hipSuccess) { (void)hipFree(a); (void)hipFree(b); return MHD_ERR_HIP; }
    (void)hipMemsetAsync(a, 0, n2 * 16, st);
    int nb = 256 * 8 * 4;
    hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, st, n2, a, b);
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
    for (int t = 0; t < iters; t++) hipLaunchKernelGGL(k_copy, dim3(nb), dim3(256), 0, st, n2, a, b);
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
    *bytes_per_s = 2.0 * 16.0 * n2 * iters / (ms * 1e-3);
    return MHD_OK;
}

const char *mhd_last_error(const mhd_handle *h) { return h ? h->err.c_str() : "null handle"; }

void mhd_destroy(mhd_handle *h)
{
    if (!h) return;
    if (h->stream) hipStreamSynchronize(h->stream);
    double *ps[] = {h->geo, h->x, h->xfold, h->xm, h->xdA, h->xdB, h->eq, h->xresd, h->cA, h->cB,
                    h->eA, h->eB, h->hA, h->hB, h->red, h->dtv};
    for (auto p : ps)
        if (p) hipFree(p);
    for (auto e : h->pool) (void)hipEventDestroy(e);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

} // extern "C"
