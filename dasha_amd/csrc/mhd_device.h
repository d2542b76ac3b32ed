// mhd_device.h -- geometry, tiling and boundary-rule helpers shared by the
// kernels of mhd_gpu.hip (single translation unit).
//
// Device layout: one FP64 volume per field covering the reference's extended
// range (x: jx = 3..mx+2, y: 1..my, z: jz = 3..mz+2), x contiguous, field
// stride S.  Index (i, j, k) = (jx-3, jy-1, jz-3).  Interior (stencil)
// points: i in [2, NX-3], j in [2, NY-3], k in [2, NZ-3].
//
// Periodic y: rows 0, 1, NY-2, NY-1 are images of NY-4, NY-3, 2, 3.  The
// device never computes them: every y-neighbour read goes through rmap()
// and the tiles cover the unique rows 2..NY-3 only (the host fills the images
// when it downloads).  Non-periodic y: all NY rows are real and carry the
// reference's "f2 = 0" rules.
#pragma once

#include <hip/hip_runtime.h>

namespace mhd {

struct Geo {
    int NX, NY, NZ;
    long sy, sz, S;           // y stride, z stride, field stride (elements)
    int periodic;             // periodic y (image rows)
    int ylo, yhi, nyu;        // rows a tile may own; unique row count when periodic
    // 1-D first-derivative coefficients a,b,c,d (t3dsdsc.f:439-504)
    const double *x1a, *x1b, *x1c, *x1d, *y1a, *y1b, *y1c, *y1d, *z1a, *z1b, *z1c, *z1d;
    // 1-D second-derivative coefficients (t3dsdsc.f:336-437)
    const double *x2a, *x2b, *x2c, *x2d, *y2a, *y2b, *y2c, *y2d, *z2a, *z2b, *z2c, *z2d;
    // smoothing weights 2*hp/hs, 2*hm/hs per axis (t3dsdsc.f:5618-5619)
    const double *xwp, *xwm, *ywp, *ywm, *zwp, *zwm;
    const double *xc, *yc, *zc;          // coordinates
    const double *dxs, *dys, *dzs;       // setdt spacings
    const double *rdx, *rdy, *rdz;       // 1/|spacing| (rounded), for divc
    const double *hdx, *hdy, *hdz;       // (eta + fmu) / spacing / 2 of setdt
    const double *cbx;                   // B-component smoothing coefficient per x
    const double *crad;                  // radial smoothing coefficient per point (0 off the interior)
    double ax3, bx3, cx3, ax4, bx4, cx4; // axp at jx = 3, 4 (x-left extrapolation)
    double dx3, dx4, rx3, rx4;           // their sums (ax+bx)+cx and the rounded reciprocals
    double axmL1, bxmL1, cxmL1, axmL, bxmL, cxmL; // axm at xlast-1, xlast
};

__host__ __device__ inline long IDX(const Geo &g, int i, int j, int k)
{
    return (long)i + g.sy * j + g.sz * k;
}

// element e of a field with a wave-uniform base: the byte offset fits 32
// bits (fields hold fewer than 2^29 points, checked at create), so the access
// can use a scalar base and a 32-bit vector offset
__device__ inline const double &fld(const double *base, unsigned e)
{
    return *(const double *)((const char *)base + (e << 3));
}
__device__ inline double *fldp(double *base, unsigned e) { return (double *)((char *)base + (e << 3)); }

// physical row of logical row j (periodic image map; identity otherwise)
__device__ inline int rmap(const Geo &g, int j)
{
    if (!g.periodic) return j;
    int r = (j - 2) % g.nyu;
    if (r < 0) r += g.nyu;
    return 2 + r;
}

// 5-point derivative, t3dsdsc.f:1308-1313 (d1fc and d2fc share the form)
__device__ inline double fd5(double fm2, double fm, double f0, double fp, double fp2, double a,
                             double b, double c, double d)
{
    return a * (fp - f0) + b * (f0 - fm) + c * (fp2 - f0) + d * (f0 - fm2);
}

// three-point smoothing term difc (t3dsdsc.f:5618-5619) with the per-axis
// weights wp = 2*hp/hs, wm = 2*hm/hs precomputed (equal to rounding)
__device__ inline double difw(double fm, double f0, double fp, double wp, double wm)
{
    return (fp - f0) * wp - (f0 - fm) * wm;
}

// radial smoothing coefficient 1/32 + (3/32)(1 + tanh((r-3)/3))/2
// (t3dsdsc.f:5659-5661, 5895-5897, 5981-5983) through
// 1 + tanh(u) = 2 / (1 + exp(-2u)); equal to rounding
__device__ inline double radial_coef(double x2y2, double z)
{
    double r = sqrt(x2y2 + z * z);
    return 1.0 / 32.0 + (3.0 / 32.0) / (1.0 + exp(-2.0 * (r - 3.) / 3.));
}

enum { KIND_RADIAL = 0, KIND_EIGHTH = 1, KIND_STATE = 2 };
enum { BC_AUX = 0, BC_STATE = 1, BC_TT = 2 };
enum { SRC_BUF = 0, SRC_XEQ = 1 };

struct Eq {
    const double *f[8];   // equilibrium per state field or nullptr
};

// ---------------------------------------------------------------------------
// z-marching tiles: a block of TX x TY threads owns a column tile of the x-y
// range and walks a chunk of interior z planes.  The last tile in x (and in
// y) is shifted back so that it is full: it overlaps its neighbour and owns
// only the columns (rows) past the neighbour's end.  A full tile at each x
// (and non-periodic y) end contains every column (row) the boundary rules of
// that end read.
// ---------------------------------------------------------------------------
constexpr int TX = 64;
constexpr int TY = 8;
constexpr int NT = TX * TY;
constexpr int NXCD = 8;          // accelerator complex dies, each with its own L2

// XCD-aware block order.  Work-groups are dispatched round-robin over the
// XCDs (linear id % 8).  Remap the linear id so that every XCD receives a
// contiguous range of the tile enumeration (x tile fastest, then y tile,
// then z chunk, then field): tiles that share halo rows or planes then run
// on the same XCD close together in time and meet in its L2.
struct BlockPos {
    int bx, by, bz;
};

__device__ inline BlockPos block_pos()
{
    long T = (long)gridDim.x * gridDim.y * gridDim.z;
    long L = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    long per = T / NXCD, rem = T % NXCD;
    long xcd = L % NXCD, n = L / NXCD;
    long V = xcd * per + (xcd < rem ? xcd : rem) + n;
    BlockPos b;
    b.bx = (int)(V % gridDim.x);
    long r = V / gridDim.x;
    b.by = (int)(r % gridDim.y);
    b.bz = (int)(r / gridDim.y);
    return b;
}


struct Tile {
    int tx, ty, i, j;
    bool valid;   // (i, j) is a computed column of the range
    bool own;     // this tile writes column (i, j)
    bool inner;   // interior column in x and y
    int kb, ke;   // interior z planes of the chunk
};

__device__ inline Tile tile_of(const Geo &g, int kc, const BlockPos &bp)
{
    Tile t;
    int chunk = bp.bz;
    t.tx = threadIdx.x % TX;
    t.ty = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);   // one row per wavefront
    int nomx = bp.bx * TX;
    int i0 = g.NX >= TX ? min(nomx, g.NX - TX) : 0;
    int nyt = g.yhi - g.ylo + 1;
    int nomy = g.ylo + bp.by * TY;
    int j0 = nyt >= TY ? min(nomy, g.yhi + 1 - TY) : g.ylo;
    t.i = i0 + t.tx;
    t.j = j0 + t.ty;
    t.valid = t.i < g.NX && t.j <= g.yhi;
    t.own = t.valid && t.i >= nomx && t.j >= nomy;
    t.inner = t.valid && t.i >= 2 && t.i <= g.NX - 3 && t.j >= 2 && t.j <= g.NY - 3;
    t.kb = 2 + chunk * kc;
    t.ke = min(t.kb + kc, g.NZ - 2);
    return t;
}

// can the value of column (i, row j of the tile) be loaded (periodic rows
// beyond the range map back onto unique rows)
__device__ inline bool loadable(const Geo &g, int i, int j)
{
    if (i < 0 || i >= g.NX) return false;
    if (g.periodic) return true;
    return j >= 0 && j < g.NY;
}

// Load plane k of NFL values into sp (halo R) through v(m, q) -> value.
template <int R, int NFL, class F>
__device__ inline void load_plane(const Geo &g, const Tile &t, double (*sp)[TY + 2 * R][TX + 2 * R], long kz,
                                  F &&v)
{
    if (loadable(g, t.i, t.j)) {
        long q = t.i + g.sy * rmap(g, t.j) + kz;
#pragma unroll
        for (int m = 0; m < NFL; m++) sp[m][t.ty + R][t.tx + R] = v(m, q);
    }
    if (t.tx < R || t.tx >= TX - R) {
        int hx = t.tx < R ? -R : R;
        if (loadable(g, t.i + hx, t.j)) {
            long q = t.i + hx + g.sy * rmap(g, t.j) + kz;
#pragma unroll
            for (int m = 0; m < NFL; m++) sp[m][t.ty + R][t.tx + R + hx] = v(m, q);
        }
    }
    if (t.ty < R || t.ty >= TY - R) {
        int hy = t.ty < R ? -R : R;
        if (loadable(g, t.i, t.j + hy)) {
            long q = t.i + g.sy * rmap(g, t.j + hy) + kz;
#pragma unroll
            for (int m = 0; m < NFL; m++) sp[m][t.ty + R + hy][t.tx + R] = v(m, q);
        }
    }
}

// ---------------------------------------------------------------------------
// In-tile boundary rules (bndryc/bndryE/bndryh = BC_AUX, t3dsdsc.f:2539-3027;
// bndry = BC_STATE, t3dsdsc.f:1677-1985 with the x-right rule mirrored on the
// left).  The reference applies x, then y, then z rules over the whole
// extended range; the final face values depend only on freshly computed
// interior values of the same row (x), column (y) or column (z), which the
// tile holds.  Usage per plane, after the interior values v[] are known:
//   1. columns 2, 3, 4 and NX-3 publish v into sx (LDS), sync
//   2. x-face columns evaluate their rule from sx
//   3. (non-periodic y) rows 2 and NY-3 publish into sy2, sync; image rows
//      evaluate 0 / half rules from sy2
// The z rules are applied by the first / last chunk from plane 2 / NZ-3.
// ---------------------------------------------------------------------------
__device__ inline int xslot(const Geo &g, int i)
{
    return i == 2 ? 0 : i == 3 ? 1 : i == 4 ? 2 : i == g.NX - 3 ? 3 : -1;
}

// a / b correctly rounded from the rounded reciprocal y = 1/b: one Newton
// correction of a*y with an exact FMA residual (Markstein); equal to the
// IEEE quotient for every finite a, b without overflow or underflow
__device__ inline double divc(double a, double b, double y)
{
    const double q0 = a * y;
    return fma(fma(-q0, b, a), y, q0);
}

// a / 3 and a / 6 of the RK4 register updates the same way, from the rounded
// reciprocals 1/3 and 1/6
__device__ inline double div3(double a) { return divc(a, 3.0, 0x1.5555555555555p-2); }
__device__ inline double div6(double a) { return divc(a, 6.0, 0x1.5555555555555p-3); }

// x-left extrapolation of the auxiliary fills (bndryc/E/h, t3dsdsc.f:2569-2572
// and the like): f(4) = (ax f5 + bx f6 + cx f7) / (ax + bx + cx), then f(3)
__device__ inline double xleft1(const Geo &g, double p2, double p3, double p4)
{
    return divc(g.ax4 * p2 + g.bx4 * p3 + g.cx4 * p4, g.dx4, g.rx4);
}
__device__ inline double xleft0(const Geo &g, double x1, double p2, double p3)
{
    return divc(g.ax3 * x1 + g.bx3 * p2 + g.cx3 * p3, g.dx3, g.rx3);
}

template <int TYPE>
__device__ inline double xrule(const Geo &g, int i, double p2, double p3, double p4, double pL)
{
    if (i == g.NX - 1) return 0.0;
    if (i == g.NX - 2) return 0.5 * pL;
    if (TYPE == BC_STATE) return i == 0 ? 0.0 : 0.5 * p2;
    double x1 = xleft1(g, p2, p3, p4);
    if (i == 1) return x1;
    return xleft0(g, x1, p2, p3);
}

__device__ inline bool xface(const Geo &g, int i) { return i < 2 || i > g.NX - 3; }
__device__ inline bool yimage(const Geo &g, int j) { return !g.periodic && (j < 2 || j > g.NY - 3); }

// LDS scratch of the in-tile rules for NF fields
template <int NF>
struct BCScratch {
    double sx[NF][TY][4];
    double sy2[NF][2][TX];
};

// Complete the boundary values of one plane: v[m] holds the interior value
// on inner columns (anything elsewhere); on return v[m] holds the final
// value on every valid column.  Contains the block-wide syncs.
template <int NF, int TYPE>
__device__ inline void plane_rules(const Geo &g, const Tile &t, BCScratch<NF> &bs, double *v)
{
    int s = t.valid ? xslot(g, t.i) : -1;
    if (s >= 0 && t.j >= 2 && t.j <= g.NY - 3) {
#pragma unroll
        for (int m = 0; m < NF; m++) bs.sx[m][t.ty][s] = v[m];
    }
    __syncthreads();
    if (t.valid && xface(g, t.i) && t.j >= 2 && t.j <= g.NY - 3) {
#pragma unroll
        for (int m = 0; m < NF; m++)
            v[m] = xrule<TYPE>(g, t.i, bs.sx[m][t.ty][0], bs.sx[m][t.ty][1], bs.sx[m][t.ty][2], bs.sx[m][t.ty][3]);
    }
    if (!g.periodic) {
        if (t.valid && (t.j == 2 || t.j == g.NY - 3)) {
#pragma unroll
            for (int m = 0; m < NF; m++) bs.sy2[m][t.j == 2 ? 0 : 1][t.tx] = v[m];
        }
        __syncthreads();
        if (t.valid && yimage(g, t.j)) {
#pragma unroll
            for (int m = 0; m < NF; m++)
                v[m] = (t.j == 0 || t.j == g.NY - 1) ? 0.0 : 0.5 * bs.sy2[m][t.j == 1 ? 0 : 1][t.tx];
        }
    }
}

} // namespace mhd
