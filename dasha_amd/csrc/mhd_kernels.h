// mhd_kernels.h -- the sweeps of one RK4 step (included by mhd_gpu.hip).
//
// Each kernel is a z-marching tile (see mhd_device.h) that produces its
// output on the interior and, through plane_rules() and the chunk-end z
// rules, on the boundary faces as well, so every buffer leaves a kernel with
// its faces filled the way the reference's boundary routine after that sweep
// would fill them.
#pragma once

#include "mhd_device.h"

namespace mhd {

// non-temporal store: outputs are read by the next kernel, not by this one
__device__ inline void st(double *p, double v) { __builtin_nontemporal_store(v, p); }

// z rules of AUX/STATE fills on the first / last chunk: plane 1 (NZ-2) is
// half of plane 2 (NZ-3), plane 0 (NZ-1) is zero (t3dsdsc.f:2614-2644,
// 1821-1861)
template <int NF>
__device__ inline void zrules(const Geo &g, const Tile &t, double *__restrict__ B, const double *v2,
                              const double *vL)
{
    if (!t.own) return;
    long q = t.i + g.sy * t.j;
    if (t.kb == 2) {
#pragma unroll
        for (int m = 0; m < NF; m++) {
            st(B + q + m * g.S, 0.0);
            st(B + q + g.sz + m * g.S, 0.5 * v2[m]);
        }
    }
    if (t.ke == g.NZ - 2) {
#pragma unroll
        for (int m = 0; m < NF; m++) {
            st(B + q + g.sz * (g.NZ - 1) + m * g.S, 0.0);
            st(B + q + g.sz * (g.NZ - 2) + m * g.S, 0.5 * vL[m]);
        }
    }
}

// ---------------------------------------------------------------------------
// current (t3dsdsc.f:2465-2491) + bndryc: J = curl(B - B_i)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void k_curl(Geo g, const double *__restrict__ X, const double *__restrict__ bxi,
                                             const double *__restrict__ byi, const double *__restrict__ bzi,
                                             double *__restrict__ cur, int kc)
{
    __shared__ double sp[3][TY + 4][TX + 4];
    __shared__ BCScratch<3> bs;
    BlockPos bp = block_pos();
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    const double *b5 = X + 4 * S, *b6 = X + 5 * S, *b7 = X + 6 * S;
    double uq[5], vq[5];
    if (ld) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            long q = q0 + sz * (t.kb - 2 + s);
            uq[s] = b5[q] - bxi[q];
            vq[s] = b6[q] - byi[q];
        }
    }
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
    if (t.inner) {
        ax[0] = g.x1a[t.i]; ax[1] = g.x1b[t.i]; ax[2] = g.x1c[t.i]; ax[3] = g.x1d[t.i];
        ay[0] = g.y1a[t.j]; ay[1] = g.y1b[t.j]; ay[2] = g.y1c[t.j]; ay[3] = g.y1d[t.j];
    }
    double v2[3] = {0, 0, 0}, vL[3] = {0, 0, 0};
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        if (ld) {
            long q = q0 + kz + 2 * sz;
            uq[4] = b5[q] - bxi[q];
            vq[4] = b6[q] - byi[q];
        }
        load_plane<2, 3>(g, t, sp, kz, [&](int m, long q) {
            return m == 0 ? b5[q] - bxi[q] : m == 1 ? b6[q] - byi[q] : b7[q] - bzi[q];
        });
        __syncthreads();
        double v[3] = {0, 0, 0};
        if (t.inner) {
            const int X0 = t.tx + 2, Y0 = t.ty + 2;
            double az[4] = {g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]};
#define DXS(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], ax[0], ax[1], ax[2], ax[3])
#define DYS(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], ay[0], ay[1], ay[2], ay[3])
            double vz = fd5(vq[0], vq[1], vq[2], vq[3], vq[4], az[0], az[1], az[2], az[3]);
            double uz = fd5(uq[0], uq[1], uq[2], uq[3], uq[4], az[0], az[1], az[2], az[3]);
            v[0] = DYS(2) - vz;
            v[1] = uz - DXS(2);
            v[2] = DXS(1) - DYS(0);
#undef DXS
#undef DYS
        }
        plane_rules<3, BC_AUX>(g, t, bs, v);
        if (t.own) {
            long q = q0 + kz;
            st(cur + q, v[0]); st(cur + q + S, v[1]); st(cur + q + 2 * S, v[2]);
        }
        if (k == 2) { v2[0] = v[0]; v2[1] = v[1]; v2[2] = v[2]; }
        if (k == g.NZ - 3) { vL[0] = v[0]; vL[1] = v[1]; vL[2] = v[2]; }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) { uq[s] = uq[s + 1]; vq[s] = vq[s + 1]; }
    }
    zrules<3>(g, t, cur, v2, vL);
}

// ---------------------------------------------------------------------------
// one smoothing sweep + its boundary fill on NF fields:
// avrgc/avrgE/avrgh (3 fields, BC_AUX) and avrg on xdif (8 fields,
// BC_STATE; the B components use the x-profile coefficient)
// t3dsdsc.f:5611-5720, 5866-5918, 5952-6004, 6008-6054
// ---------------------------------------------------------------------------
template <int NF, int KIND, int TYPE>
__global__ __launch_bounds__(NT) void k_smooth(Geo g, const double *__restrict__ A, double *__restrict__ B, int kc)
{
    __shared__ double sp[NF][TY + 2][TX + 2];
    __shared__ BCScratch<NF> bs;
    BlockPos bp = block_pos();
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    double zm[NF], z0[NF], zp[NF];
    if (ld) {
#pragma unroll
        for (int m = 0; m < NF; m++) {
            zm[m] = A[m * S + q0 + sz * (t.kb - 1)];
            z0[m] = A[m * S + q0 + sz * t.kb];
        }
    }
    const double wpx = t.inner ? g.xwp[t.i] : 0, wmx = t.inner ? g.xwm[t.i] : 0;
    const double wpy = t.inner ? g.ywp[t.j] : 0, wmy = t.inner ? g.ywm[t.j] : 0;
    const double cb = (KIND == KIND_STATE && t.inner) ? g.cbx[t.i] : 0.0;
    const double xy2 = t.inner ? g.xc[t.i] * g.xc[t.i] + g.yc[t.j] * g.yc[t.j] : 0.0;
    double v2[NF], vL[NF];
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        if (ld) {
#pragma unroll
            for (int m = 0; m < NF; m++) zp[m] = A[m * S + q0 + kz + sz];
        }
        load_plane<1, NF>(g, t, sp, kz, [&](int m, long q) { return A[m * S + q]; });
        __syncthreads();
        double v[NF];
        if (t.inner) {
            const double wpz = g.zwp[k], wmz = g.zwm[k];
            const double crad = (KIND == KIND_EIGHTH) ? 1.0 / 8.0 : radial_coef(xy2, g.zc[k]);
#pragma unroll
            for (int m = 0; m < NF; m++) {
                double f0 = z0[m];
                double s = difw(sp[m][t.ty + 1][t.tx], f0, sp[m][t.ty + 1][t.tx + 2], wpx, wmx) +
                           difw(sp[m][t.ty][t.tx + 1], f0, sp[m][t.ty + 2][t.tx + 1], wpy, wmy) +
                           difw(zm[m], f0, zp[m], wpz, wmz);
                double cc = (KIND == KIND_STATE && m >= 4 && m <= 6) ? cb : crad;
                v[m] = f0 + cc * s;
            }
        } else {
#pragma unroll
            for (int m = 0; m < NF; m++) v[m] = 0.0;
        }
        plane_rules<NF, TYPE>(g, t, bs, v);
        if (t.own) {
#pragma unroll
            for (int m = 0; m < NF; m++) B[m * S + q0 + kz] = v[m];
        }
        if (k == 2) {
#pragma unroll
            for (int m = 0; m < NF; m++) v2[m] = v[m];
        }
        if (k == g.NZ - 3) {
#pragma unroll
            for (int m = 0; m < NF; m++) vL[m] = v[m];
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NF; m++) { zm[m] = z0[m]; z0[m] = zp[m]; }
    }
    zrules<NF>(g, t, B, v2, vL);
}

// ---------------------------------------------------------------------------
// Sync-free smoothing sweep: one field per block, one tile row per 64-lane
// wavefront.  x neighbours come from the neighbouring lanes (the two tile
// edge lanes load their outer neighbour), y neighbours are read from L1/L2,
// z neighbours from the register queue.  The x-face rules use lane
// shuffles within the row; the non-periodic image rows evaluate the sweep at
// row 2 (NY-3) themselves.  No LDS and no block barrier.
// ---------------------------------------------------------------------------
constexpr int SY = 8;            // rows (wavefronts) per block

struct RowTile {
    int lane, i, j, nomx, nomy, kb, ke;
    bool valid, own;
};

__device__ inline RowTile rowtile_of(const Geo &g, int kc, const BlockPos &bp, int chunk)
{
    RowTile t;
    t.lane = threadIdx.x % TX;
    int row = threadIdx.x / TX;
    t.nomx = bp.bx * TX;
    int i0 = g.NX >= TX ? min(t.nomx, g.NX - TX) : 0;
    int nyt = g.yhi - g.ylo + 1;
    t.nomy = g.ylo + bp.by * SY;
    int j0 = nyt >= SY ? min(t.nomy, g.yhi + 1 - SY) : g.ylo;
    t.i = i0 + t.lane;
    t.j = j0 + row;
    t.valid = t.i < g.NX && t.j <= g.yhi;
    t.own = t.valid && t.i >= t.nomx && t.j >= t.nomy;
    t.kb = 2 + chunk * kc;
    t.ke = min(t.kb + kc, g.NZ - 2);
    return t;
}

__device__ inline double shfl(double v, int src) { return __shfl(v, src, TX); }

// x-face values of a row from the lanes holding columns 2, 3, 4, NX-3
// x-face values of a row (lane shuffles; every lane of the row must call)
template <int TYPE>
__device__ inline double row_xrule_lanes(const Geo &g, int i, int i0, bool valid, double v)
{
    int l2 = 2 - i0, lL = g.NX - 3 - i0;
    double p2 = shfl(v, l2 >= 0 && l2 < TX ? l2 : 0);
    double p3 = shfl(v, l2 + 1 >= 0 && l2 + 1 < TX ? l2 + 1 : 0);
    double p4 = shfl(v, l2 + 2 >= 0 && l2 + 2 < TX ? l2 + 2 : 0);
    double pL = shfl(v, lL >= 0 && lL < TX ? lL : 0);
    if (valid && xface(g, i)) return xrule<TYPE>(g, i, p2, p3, p4, pL);
    return v;
}

template <int TYPE>
__device__ inline double row_xrule(const Geo &g, const RowTile &t, int i0, double v)
{
    int l2 = 2 - i0, lL = g.NX - 3 - i0;
    double p2 = shfl(v, l2 >= 0 && l2 < TX ? l2 : 0);
    double p3 = shfl(v, l2 + 1 >= 0 && l2 + 1 < TX ? l2 + 1 : 0);
    double p4 = shfl(v, l2 + 2 >= 0 && l2 + 2 < TX ? l2 + 2 : 0);
    double pL = shfl(v, lL >= 0 && lL < TX ? lL : 0);
    if (t.valid && xface(g, t.i)) return xrule<TYPE>(g, t.i, p2, p3, p4, pL);
    return v;
}

template <int KIND, int TYPE>
__global__ __launch_bounds__(TX * SY) void k_sweep(Geo g, const double *__restrict__ A, double *__restrict__ B,
                                                   int kc, int nchunk)
{
    BlockPos bp = block_pos();
    const int m = bp.bz / nchunk;
    RowTile t = rowtile_of(g, kc, bp, bp.bz % nchunk);
    if (t.kb >= t.ke) return;
    const int i0 = t.i - t.lane;
    const long S = g.S, sz = g.sz, sy = g.sy;
    const double *a = A + m * S;
    double *b = B + m * S;
    const bool inner = t.valid && t.i >= 2 && t.i <= g.NX - 3 && t.j >= 2 && t.j <= g.NY - 3;
    const bool image = t.valid && yimage(g, t.j);
    // the column this thread evaluates the sweep at: its own, or row 2 /
    // NY-3 for a non-periodic image row
    const int je = image ? (t.j < 2 ? 2 : g.NY - 3) : t.j;
    const bool ecol = t.i < g.NX && (t.valid || !g.periodic ? true : true);
    const bool einner = t.valid && t.i >= 2 && t.i <= g.NX - 3;   // x-interior column, row je interior
    const long qc = (long)min(t.i, g.NX - 1) + sy * rmap(g, je);  // column of evaluation
    const long qm = (long)min(t.i, g.NX - 1) + sy * rmap(g, je - 1);
    const long qp = (long)min(t.i, g.NX - 1) + sy * rmap(g, je + 1);
    const bool ld = t.i < g.NX && (g.periodic || (je >= 0 && je < g.NY));
    (void)ecol;
    double cb = 0.0, wpx = 0, wmx = 0, wpy = 0, wmy = 0, xy2 = 0;
    if (einner) {
        wpx = g.xwp[t.i]; wmx = g.xwm[t.i]; wpy = g.ywp[je]; wmy = g.ywm[je];
        xy2 = g.xc[t.i] * g.xc[t.i] + g.yc[je] * g.yc[je];
        if (KIND == KIND_STATE) cb = g.cbx[t.i];
    }
    const bool bcoef = KIND == KIND_STATE && m >= 4 && m <= 6;
    // edge lanes' outer x neighbour
    const int ex = t.lane == 0 ? -1 : (t.lane == TX - 1 ? 1 : 0);
    const bool exok = ex != 0 && t.i + ex >= 0 && t.i + ex < g.NX;
    double zm = 0, z0 = 0, zp = 0;
    if (ld) { zm = a[qc + sz * (t.kb - 1)]; z0 = a[qc + sz * t.kb]; }
    double v2 = 0, vL = 0;
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        double ym = 0, yp = 0, xo = 0;
        if (ld) {
            zp = a[qc + kz + sz];
            ym = a[qm + kz];
            yp = a[qp + kz];
        }
        if (exok) xo = a[qc + ex + kz];
        double xm = shfl(z0, t.lane > 0 ? t.lane - 1 : 0);
        double xp = shfl(z0, t.lane < TX - 1 ? t.lane + 1 : TX - 1);
        if (ex < 0) xm = xo;
        if (ex > 0) xp = xo;
        double v = 0.0;
        if (einner) {
            double c;
            if (KIND == KIND_EIGHTH) c = 1.0 / 8.0;
            else if (bcoef) c = cb;
            else c = radial_coef(xy2, g.zc[k]);
            double s = difw(xm, z0, xp, wpx, wmx) + difw(ym, z0, yp, wpy, wmy) +
                       difw(zm, z0, zp, g.zwp[k], g.zwm[k]);
            v = z0 + c * s;
        }
        // x rule on the evaluated row, then the image-row rule
        v = row_xrule<TYPE>(g, t, i0, v);
        if (image) v = (t.j == 0 || t.j == g.NY - 1) ? 0.0 : 0.5 * v;
        if (t.own) b[(long)t.i + sy * t.j + kz] = v;
        if (k == 2) v2 = v;
        if (k == g.NZ - 3) vL = v;
        zm = z0;
        z0 = zp;
    }
    if (t.own) {
        long q = (long)t.i + sy * t.j;
        if (t.kb == 2) { b[q] = 0.0; b[q + sz] = 0.5 * v2; }
        if (t.ke == g.NZ - 2) { b[q + sz * (g.NZ - 1)] = 0.0; b[q + sz * (g.NZ - 2)] = 0.5 * vL; }
    }
    (void)inner;
}

// Multi-field variant of k_sweep: NF fields per thread (coefficient once per
// cell), loads of the next plane issued one iteration ahead.
template <int NF, int KIND, int TYPE>
__global__ __launch_bounds__(TX * SY) void k_sweepn(Geo g, const double *__restrict__ A, double *__restrict__ B,
                                                    int kc)
{
    BlockPos bp = block_pos();
    RowTile t = rowtile_of(g, kc, bp, bp.bz);
    if (t.kb >= t.ke) return;
    const int i0 = t.i - t.lane;
    const long S = g.S, sz = g.sz, sy = g.sy;
    const bool image = t.valid && yimage(g, t.j);
    const int je = image ? (t.j < 2 ? 2 : g.NY - 3) : t.j;
    const bool einner = t.valid && t.i >= 2 && t.i <= g.NX - 3;
    const int ic = min(t.i, g.NX - 1);
    const long qc = (long)ic + sy * rmap(g, je);
    const long qm = (long)ic + sy * rmap(g, je - 1);
    const long qp = (long)ic + sy * rmap(g, je + 1);
    const bool ld = t.i < g.NX && (g.periodic || (je >= 0 && je < g.NY));
    double cb = 0.0, wpx = 0, wmx = 0, wpy = 0, wmy = 0, xy2 = 0;
    if (einner) {
        wpx = g.xwp[t.i]; wmx = g.xwm[t.i]; wpy = g.ywp[je]; wmy = g.ywm[je];
        xy2 = g.xc[t.i] * g.xc[t.i] + g.yc[je] * g.yc[je];
        if (KIND == KIND_STATE) cb = g.cbx[t.i];
    }
    const int ex = t.lane == 0 ? -1 : (t.lane == TX - 1 ? 1 : 0);
    const bool exok = ex != 0 && t.i + ex >= 0 && t.i + ex < g.NX;
    // registers: plane k-1, k, k+1 of the own column; rows j-1, j+1 and the
    // outer x neighbour of plane k; the prefetched plane-k+1 rows
    double zm[NF], z0[NF], zp[NF], ym[NF], yp[NF], xo[NF];
    double nym[NF], nyp[NF], nxo[NF];
#pragma unroll
    for (int m = 0; m < NF; m++) {
        zm[m] = z0[m] = zp[m] = ym[m] = yp[m] = xo[m] = 0.0;
        nym[m] = nyp[m] = nxo[m] = 0.0;
    }
    {
        const long kz = sz * t.kb;
        if (ld) {
#pragma unroll
            for (int m = 0; m < NF; m++) {
                const double *a = A + m * S;
                zm[m] = a[qc + kz - sz]; z0[m] = a[qc + kz]; zp[m] = a[qc + kz + sz];
                ym[m] = a[qm + kz]; yp[m] = a[qp + kz];
            }
        }
        if (exok) {
#pragma unroll
            for (int m = 0; m < NF; m++) xo[m] = A[m * S + qc + ex + kz];
        }
    }
    double v2[NF], vL[NF];
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        // prefetch plane k+1 (rows) and plane k+2 (own column)
        double nzp[NF];
        const bool more = k + 1 < t.ke;
#pragma unroll
        for (int m = 0; m < NF; m++) nzp[m] = 0.0;
        if (ld && more) {
#pragma unroll
            for (int m = 0; m < NF; m++) {
                const double *a = A + m * S;
                nzp[m] = a[qc + kz + 2 * sz];
                nym[m] = a[qm + kz + sz];
                nyp[m] = a[qp + kz + sz];
            }
        }
        if (exok && more) {
#pragma unroll
            for (int m = 0; m < NF; m++) nxo[m] = A[m * S + qc + ex + kz + sz];
        }
        double crad = 0.0;
        if (KIND == KIND_EIGHTH) crad = 1.0 / 8.0;
        else if (einner) crad = radial_coef(xy2, g.zc[k]);
        const double wpz = g.zwp[k], wmz = g.zwm[k];
        double v[NF];
#pragma unroll
        for (int m = 0; m < NF; m++) {
            double xm = shfl(z0[m], t.lane > 0 ? t.lane - 1 : 0);
            double xp = shfl(z0[m], t.lane < TX - 1 ? t.lane + 1 : TX - 1);
            if (ex < 0) xm = xo[m];
            if (ex > 0) xp = xo[m];
            v[m] = 0.0;
            if (einner) {
                double c = (KIND == KIND_STATE && m >= 4 && m <= 6) ? cb : crad;
                double s = difw(xm, z0[m], xp, wpx, wmx) + difw(ym[m], z0[m], yp[m], wpy, wmy) +
                           difw(zm[m], z0[m], zp[m], wpz, wmz);
                v[m] = z0[m] + c * s;
            }
            v[m] = row_xrule<TYPE>(g, t, i0, v[m]);
            if (image) v[m] = (t.j == 0 || t.j == g.NY - 1) ? 0.0 : 0.5 * v[m];
        }
        if (t.own) {
#pragma unroll
            for (int m = 0; m < NF; m++) B[m * S + (long)t.i + sy * t.j + kz] = v[m];
        }
        if (k == 2) {
#pragma unroll
            for (int m = 0; m < NF; m++) v2[m] = v[m];
        }
        if (k == g.NZ - 3) {
#pragma unroll
            for (int m = 0; m < NF; m++) vL[m] = v[m];
        }
#pragma unroll
        for (int m = 0; m < NF; m++) {
            zm[m] = z0[m]; z0[m] = zp[m]; zp[m] = nzp[m];
            ym[m] = nym[m]; yp[m] = nyp[m]; xo[m] = nxo[m];
        }
    }
    if (t.own) {
        long q = (long)t.i + sy * t.j;
#pragma unroll
        for (int m = 0; m < NF; m++) {
            double *b = B + m * S;
            if (t.kb == 2) { b[q] = 0.0; b[q + sz] = 0.5 * v2[m]; }
            if (t.ke == g.NZ - 2) { b[q + sz * (g.NZ - 1)] = 0.0; b[q + sz * (g.NZ - 2)] = 0.5 * vL[m]; }
        }
    }
}

// ---------------------------------------------------------------------------
// Point-parallel smoothing sweep: one thread per (i, j, k) of the extended
// range and NF fields, no z march, so every thread issues its loads at
// once (the 7-point reuse is served by L1/L2).  A block is 64 x PY x 1:
// each wavefront is one x row, so the x-face rules run on lane shuffles.
// Points on an image row or a z face evaluate the sweep at row 2 (NY-3) /
// plane 2 (NZ-3) themselves and apply the rules to that value.
// ---------------------------------------------------------------------------
constexpr int PY = 4;

template <int NF, int KIND, int TYPE>
__global__ __launch_bounds__(TX * PY) void k_sweep_pt(Geo g, const double *__restrict__ A, double *__restrict__ B)
{
    const int lane = threadIdx.x % TX, row = threadIdx.x / TX;
    const int nomx = blockIdx.x * TX;
    const int i0 = g.NX >= TX ? min(nomx, g.NX - TX) : 0;
    const int nyt = g.yhi - g.ylo + 1;
    const int nomy = g.ylo + blockIdx.y * PY;
    const int j0 = nyt >= PY ? min(nomy, g.yhi + 1 - PY) : g.ylo;
    const int i = i0 + lane, j = j0 + row, k = blockIdx.z;
    const bool valid = i < g.NX && j <= g.yhi;
    const bool own = valid && i >= nomx && j >= nomy;
    const long S = g.S, sy = g.sy, sz = g.sz;
    // point of evaluation
    const bool image = valid && yimage(g, j);
    const bool zface = k < 2 || k > g.NZ - 3;
    const int je = image ? (j < 2 ? 2 : g.NY - 3) : j;
    const int ke = zface ? (k < 2 ? 2 : g.NZ - 3) : k;
    const bool einner = valid && i >= 2 && i <= g.NX - 3;
    const int ic = min(i, g.NX - 1);
    const bool ld = i < g.NX && (g.periodic || je < g.NY);
    double v[NF];
#pragma unroll
    for (int m = 0; m < NF; m++) v[m] = 0.0;
    if (einner && ld) {
        const long qc = (long)ic + sy * rmap(g, je) + sz * ke;
        const long qm = (long)ic + sy * rmap(g, je - 1) + sz * ke;
        const long qp = (long)ic + sy * rmap(g, je + 1) + sz * ke;
        double c = 1.0 / 8.0;
        if (KIND != KIND_EIGHTH) c = radial_coef(g.xc[i] * g.xc[i] + g.yc[je] * g.yc[je], g.zc[ke]);
        const double cb = KIND == KIND_STATE ? g.cbx[i] : 0.0;
        const double wpx = g.xwp[i], wmx = g.xwm[i], wpy = g.ywp[je], wmy = g.ywm[je];
        const double wpz = g.zwp[ke], wmz = g.zwm[ke];
#pragma unroll
        for (int m = 0; m < NF; m++) {
            const double *a = A + m * S;
            double f0 = a[qc];
            double s = difw(a[qc - 1], f0, a[qc + 1], wpx, wmx) + difw(a[qm], f0, a[qp], wpy, wmy) +
                       difw(a[qc - sz], f0, a[qc + sz], wpz, wmz);
            double cc = (KIND == KIND_STATE && m >= 4 && m <= 6) ? cb : c;
            v[m] = f0 + cc * s;
        }
    }
    // every lane takes part in the shuffles of its row
#pragma unroll
    for (int m = 0; m < NF; m++) {
        v[m] = row_xrule_lanes<TYPE>(g, i, i0, valid, v[m]);
        if (image) v[m] = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * v[m];
        if (zface) v[m] = (k == 0 || k == g.NZ - 1) ? 0.0 : 0.5 * v[m];
    }
    if (own) {
        const long q = (long)i + sy * j + sz * k;
#pragma unroll
        for (int m = 0; m < NF; m++) B[q + m * S] = v[m];
    }
}

// ---------------------------------------------------------------------------
// z-unrolled point-parallel smoothing sweep (the production sweep).  Thread
// (i, j) computes UZ consecutive planes [k0, k0+UZ) of all NF fields (one
// field at a time); the column values of planes k0-1..k0+UZ are loaded once
// and shared between the planes.  The plane groups tile [0, NZ) from plane
// 0; the last group is shifted back to end at NZ-1 (it writes only its own
// planes), so the z-face planes 0, 1 (NZ-2, NZ-1) always sit in the group
// that also computes plane 2 (NZ-3).  Stores are non-temporal: the output
// is not re-read by this sweep, and keeping it out of L2 and the Infinity
// Cache leaves room for the neighbour planes the other threads still read.
// ---------------------------------------------------------------------------
constexpr int UZ = 4;
constexpr int BY = 8;

template <int NF, int KIND, int TYPE, int FPB>
__global__ __launch_bounds__(TX * BY) void k_sweepz(Geo g, const double *__restrict__ A, double *__restrict__ B)
{
    // FPB fields per block: blockIdx.z = plane group + ngroups * field block
    const int ngz = (g.NZ + UZ - 1) / UZ;
    const int mb = (blockIdx.z / ngz) * FPB;
    const int lane = threadIdx.x % TX, row = threadIdx.x / TX;
    const int nomx = blockIdx.x * TX;
    const int i0 = g.NX >= TX ? min(nomx, g.NX - TX) : 0;
    const int nyt = g.yhi - g.ylo + 1;
    const int nomy = g.ylo + blockIdx.y * BY;
    const int j0 = nyt >= BY ? min(nomy, g.yhi + 1 - BY) : g.ylo;
    const int nomz = (blockIdx.z % ngz) * UZ;
    const int k0 = min(nomz, g.NZ - UZ);
    const int i = i0 + lane, j = j0 + row;
    const bool valid = i < g.NX && j <= g.yhi;
    const bool own = valid && i >= nomx && j >= nomy;
    // does this row tile hold an x face (uniform over the wavefront)
    const bool xft = i0 < 2 || i0 + TX > g.NX - 2;
    const long S = g.S, sy = g.sy, sz = g.sz;
    const bool image = valid && yimage(g, j);
    const int je = image ? (j < 2 ? 2 : g.NY - 3) : j;
    const bool einner = valid && i >= 2 && i <= g.NX - 3;
    const int ic = min(i, g.NX - 1);
    const bool ld = i < g.NX && (g.periodic || je < g.NY);
    const long qc = (long)ic + sy * rmap(g, je), qm = (long)ic + sy * rmap(g, je - 1),
               qp = (long)ic + sy * rmap(g, je + 1);
    // per-plane coefficients and weights (interior planes of the group)
    double cr[UZ], wpz[UZ], wmz[UZ];
    const bool needr = KIND == KIND_RADIAL || (KIND == KIND_STATE && !(FPB == 1 && mb >= 4 && mb <= 6));
    const double xy2 = einner ? g.xc[i] * g.xc[i] + g.yc[je] * g.yc[je] : 0.0;
#pragma unroll
    for (int u = 0; u < UZ; u++) {
        int k = k0 + u;
        bool kin = k >= 2 && k <= g.NZ - 3;
        cr[u] = (KIND == KIND_EIGHTH) ? 1.0 / 8.0 : ((needr && einner && kin) ? radial_coef(xy2, g.zc[k]) : 0.0);
        wpz[u] = kin ? g.zwp[k] : 0.0;
        wmz[u] = kin ? g.zwm[k] : 0.0;
    }
    const double wpx = einner ? g.xwp[i] : 0, wmx = einner ? g.xwm[i] : 0;
    const double wpy = einner ? g.ywp[je] : 0, wmy = einner ? g.ywm[je] : 0;
    const double cb = (KIND == KIND_STATE && einner) ? g.cbx[i] : 0.0;
#pragma unroll
    for (int mm = 0; mm < FPB; mm++) {
        const int m = mb + mm;
        if (m >= NF) break;
        const double *a = A + m * S;
        double col[UZ + 2], xm[UZ], xp[UZ], ym[UZ], yp[UZ];
#pragma unroll
        for (int u = 0; u < UZ + 2; u++) {
            int k = k0 - 1 + u;
            col[u] = (ld && k >= 0 && k < g.NZ) ? a[qc + sz * k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UZ; u++) {
            int k = k0 + u;
            bool kin = ld && einner && k >= 2 && k <= g.NZ - 3;
            long qk = sz * k;
            xm[u] = kin ? a[qc + qk - 1] : 0.0;
            xp[u] = kin ? a[qc + qk + 1] : 0.0;
            ym[u] = kin ? a[qm + qk] : 0.0;
            yp[u] = kin ? a[qp + qk] : 0.0;
        }
        const bool bcoef = KIND == KIND_STATE && m >= 4 && m <= 6;
        double v[UZ];
#pragma unroll
        for (int u = 0; u < UZ; u++) {
            int k = k0 + u;
            v[u] = 0.0;
            if (einner && k >= 2 && k <= g.NZ - 3) {
                double f0 = col[u + 1];
                double s = difw(xm[u], f0, xp[u], wpx, wmx) + difw(ym[u], f0, yp[u], wpy, wmy) +
                           difw(col[u], f0, col[u + 2], wpz[u], wmz[u]);
                v[u] = f0 + (bcoef ? cb : cr[u]) * s;
            }
            // x rule on the row, then the image-row rule
            if (xft) v[u] = row_xrule_lanes<TYPE>(g, i, i0, valid, v[u]);
            if (image) v[u] = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * v[u];
        }
        // z-face planes from plane 2 / NZ-3 of the same group
#pragma unroll
        for (int u = 0; u < UZ; u++) {
            int k = k0 + u;
            if (k == 0 || k == g.NZ - 1) v[u] = 0.0;
            else if (k == 1) v[u] = 0.5 * v[u + 1];
            else if (k == g.NZ - 2) v[u] = 0.5 * v[u - 1];
        }
        if (own) {
            double *b = B + m * S + (long)i + sy * j;
#pragma unroll
            for (int u = 0; u < UZ; u++)
                if (k0 + u >= nomz) st(b + sz * (k0 + u), v[u]);
        }
    }
}

// ---------------------------------------------------------------------------
// right, part 1 (t3dsdsc.f:1332-1433): per point the momentum flux tensor
// (t3dsdsc.f:2311-2346) is formed once,
//   T22 = x2*x2/x1 + (p - p_i), T23 = x2*x3/x1, T24 = x2*x4/x1,
//   T33 = x3*x3/x1 + (p - p_i), T34 = x4*x3/x1, T44 = x4*x4/x1 + (p - p_i),
// and xdif(1..4) = -div of (x2,x3,x4), (T22,T23,T24), (T23,T33,T34),
// (T24,T34,T44) -- the values of the reference's flux(m) arrays -- then
// + J x B - xresd on the interior; E = -v x B + eta J on every column; the
// Hall term xdh on the interior plus its bndryh fill.
// LDS: {x2,x3,T22,T23,T24,T33,T34,p} of plane k (halo 2); queues:
// {x4,T24,T34,T44,p} of planes k-2..k+2 and the x-y values of k..k+2.
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

__device__ inline double rec_field(const Rec &r, int m)
{
    switch (m) {
    case LX2: return r.x2;
    case LX3: return r.x3;
    case LT22: return r.t22;
    case LT23: return r.t23;
    case LT24: return r.t24;
    case LT33: return r.t33;
    case LT34: return r.t34;
    default: return r.p;
    }
}

__device__ inline void e_point(const double *__restrict__ X, const double *__restrict__ cur, double eta, long S,
                               long q, double *__restrict__ E)
{
    double x1 = X[q], x2 = X[q + S], x3 = X[q + 2 * S], x4 = X[q + 3 * S];
    double x5 = X[q + 4 * S], x6 = X[q + 5 * S], x7 = X[q + 6 * S];
    st(E + q, (-x3 * x7 + x4 * x6) / x1 + eta * cur[q]);
    st(E + q + S, (-x4 * x5 + x2 * x7) / x1 + eta * cur[q + S]);
    st(E + q + 2 * S, (-x2 * x6 + x3 * x5) / x1 + eta * cur[q + 2 * S]);
}

__global__ __launch_bounds__(NT) void k_rhs_a(Geo g, const double *__restrict__ X, const double *__restrict__ pri,
                                              const double *__restrict__ cur, const double *__restrict__ xresd,
                                              double eta, int hall, double cht, double *__restrict__ XD,
                                              double *__restrict__ E, double *__restrict__ XH, int kc)
{
    __shared__ double sp[NLA][TY + 4][TX + 4];
    __shared__ BCScratch<3> bs;
    BlockPos bp = block_pos();
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    if (t.own) {   // E on the planes outside the interior z range
        if (t.kb == 2) { e_point(X, cur, eta, S, q0, E); e_point(X, cur, eta, S, q0 + sz, E); }
        if (t.ke == g.NZ - 2) {
            e_point(X, cur, eta, S, q0 + sz * (g.NZ - 2), E);
            e_point(X, cur, eta, S, q0 + sz * (g.NZ - 1), E);
        }
    }
    double qx4[5], qt24[5], qt34[5], qt44[5], qp[5];
    double yx2[3], yx3[3], yt22[3], yt23[3], yt33[3];
    if (ld) {
#pragma unroll
        for (int s = 0; s < 4; s++) {
            Rec r = rec_at(X, pri, S, q0 + sz * (t.kb - 2 + s));
            qx4[s] = r.x4; qt24[s] = r.t24; qt34[s] = r.t34; qt44[s] = r.t44; qp[s] = r.p;
            if (s >= 2) {
                yx2[s - 2] = r.x2; yx3[s - 2] = r.x3; yt22[s - 2] = r.t22; yt23[s - 2] = r.t23; yt33[s - 2] = r.t33;
            }
        }
    }
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
    if (t.inner) {
        ax[0] = g.x1a[t.i]; ax[1] = g.x1b[t.i]; ax[2] = g.x1c[t.i]; ax[3] = g.x1d[t.i];
        ay[0] = g.y1a[t.j]; ay[1] = g.y1b[t.j]; ay[2] = g.y1c[t.j]; ay[3] = g.y1d[t.j];
    }
    double h2[3] = {0, 0, 0}, hL[3] = {0, 0, 0};
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        const long qk = q0 + kz;
        if (ld) {
            Rec r = rec_at(X, pri, S, qk + 2 * sz);
            qx4[4] = r.x4; qt24[4] = r.t24; qt34[4] = r.t34; qt44[4] = r.t44; qp[4] = r.p;
            yx2[2] = r.x2; yx3[2] = r.x3; yt22[2] = r.t22; yt23[2] = r.t23; yt33[2] = r.t33;
            const int y = t.ty + 2, x = t.tx + 2;
            sp[LX2][y][x] = yx2[0]; sp[LX3][y][x] = yx3[0]; sp[LT22][y][x] = yt22[0]; sp[LT23][y][x] = yt23[0];
            sp[LT24][y][x] = qt24[2]; sp[LT33][y][x] = yt33[0]; sp[LT34][y][x] = qt34[2]; sp[LP][y][x] = qp[2];
        }
        if (t.own) e_point(X, cur, eta, S, qk, E);
        // halo only (the own value is in place already)
        if (t.tx < 2 || t.tx >= TX - 2) {
            int hx = t.tx < 2 ? -2 : 2;
            if (loadable(g, t.i + hx, t.j)) {
                Rec r = rec_at(X, pri, S, t.i + hx + g.sy * rmap(g, t.j) + kz);
#pragma unroll
                for (int m = 0; m < NLA; m++) sp[m][t.ty + 2][t.tx + 2 + hx] = rec_field(r, m);
            }
        }
        if (t.ty < 2 || t.ty >= TY - 2) {
            int hy = t.ty < 2 ? -2 : 2;
            if (loadable(g, t.i, t.j + hy)) {
                Rec r = rec_at(X, pri, S, t.i + g.sy * rmap(g, t.j + hy) + kz);
#pragma unroll
                for (int m = 0; m < NLA; m++) sp[m][t.ty + 2 + hy][t.tx + 2] = rec_field(r, m);
            }
        }
        __syncthreads();
        double hv[3] = {0, 0, 0};
        if (t.inner) {
            const int X0 = t.tx + 2, Y0 = t.ty + 2;
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
            st(XD + qk, xd0); st(XD + qk + S, xd1); st(XD + qk + 2 * S, xd2); st(XD + qk + 3 * S, xd3);
            if (hall) {
                hv[0] = cht * (c2 * x7 - c3 * x6 - DXL(LP)) / x1;
                hv[1] = cht * (c3 * x5 - c1 * x7 - DYL(LP)) / x1;
                hv[2] = cht * (c1 * x6 - c2 * x5 - DZQ(qp)) / x1;
            }
#undef DZQ
#undef DXL
#undef DYL
        }
        if (hall) {
            plane_rules<3, BC_AUX>(g, t, bs, hv);
            if (t.own) { st(XH + qk, hv[0]); st(XH + qk + S, hv[1]); st(XH + qk + 2 * S, hv[2]); }
            if (k == 2) { h2[0] = hv[0]; h2[1] = hv[1]; h2[2] = hv[2]; }
            if (k == g.NZ - 3) { hL[0] = hv[0]; hL[1] = hv[1]; hL[2] = hv[2]; }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            qx4[s] = qx4[s + 1]; qt24[s] = qt24[s + 1]; qt34[s] = qt34[s + 1]; qt44[s] = qt44[s + 1]; qp[s] = qp[s + 1];
        }
#pragma unroll
        for (int s = 0; s < 2; s++) {
            yx2[s] = yx2[s + 1]; yx3[s] = yx3[s + 1]; yt22[s] = yt22[s + 1]; yt23[s] = yt23[s + 1]; yt33[s] = yt33[s + 1];
        }
    }
    if (hall) zrules<3>(g, t, XH, h2, hL);
}

// ---------------------------------------------------------------------------
// right, part 2 (t3dsdsc.f:1426-1566, 1585-1663, 1665): -curl(E + xdh), the
// pressure equation, optional viscosity / diffusion, then the bndry fill of
// all eight xdif fields.  Per point e = E + xdh and v = m / rho are formed
// once.  LDS: {e1,e2,e3,p,vx,vy[,vz,rho]}; queue: {e1,e2,p,vz[,vx,vy,rho]}.
// ---------------------------------------------------------------------------
enum { LE1 = 0, LE2, LE3, LPP, LVX, LVY, LVZ, LRH };

template <bool VD>
__global__ __launch_bounds__(NT) void k_rhs_b(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                                              const double *__restrict__ E, const double *__restrict__ XH,
                                              double eta, double gam, double fmu, double frho, double fpr,
                                              double *__restrict__ XD, int kc)
{
    constexpr int NL = VD ? 8 : 6;
    __shared__ double sp[NL][TY + 4][TX + 4];
    __shared__ BCScratch<8> bs;
    BlockPos bp = block_pos();
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    auto ept = [&](long q, int m) { return XH ? E[q + m * S] + XH[q + m * S] : E[q + m * S]; };
    double qe1[5], qe2[5], qp[5], qvz[5], qvx[5], qvy[5], qrh[5];
    auto load_q = [&](int s, long q) {
        double r = X[q];
        qe1[s] = ept(q, 0); qe2[s] = ept(q, 1); qp[s] = X[q + 7 * S]; qvz[s] = X[q + 3 * S] / r;
        if (VD) { qvx[s] = X[q + S] / r; qvy[s] = X[q + 2 * S] / r; qrh[s] = r; }
    };
    if (ld) {
#pragma unroll
        for (int s = 0; s < 4; s++) load_q(s, q0 + sz * (t.kb - 2 + s));
    }
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0}, bx[4] = {0, 0, 0, 0}, by[4] = {0, 0, 0, 0};
    if (t.inner) {
        ax[0] = g.x1a[t.i]; ax[1] = g.x1b[t.i]; ax[2] = g.x1c[t.i]; ax[3] = g.x1d[t.i];
        ay[0] = g.y1a[t.j]; ay[1] = g.y1b[t.j]; ay[2] = g.y1c[t.j]; ay[3] = g.y1d[t.j];
        if (VD) {
            bx[0] = g.x2a[t.i]; bx[1] = g.x2b[t.i]; bx[2] = g.x2c[t.i]; bx[3] = g.x2d[t.i];
            by[0] = g.y2a[t.j]; by[1] = g.y2b[t.j]; by[2] = g.y2c[t.j]; by[3] = g.y2d[t.j];
        }
    }
    double v2[8], vL[8];
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        const long qk = q0 + kz;
        if (ld) load_q(4, qk + 2 * sz);
        load_plane<2, NL>(g, t, sp, kz, [&](int m, long q) {
            switch (m) {
            case LE1: return ept(q, 0);
            case LE2: return ept(q, 1);
            case LE3: return ept(q, 2);
            case LPP: return X[q + 7 * S];
            case LVX: return X[q + S] / X[q];
            case LVY: return X[q + 2 * S] / X[q];
            case LVZ: return X[q + 3 * S] / X[q];
            default: return X[q];
            }
        });
        __syncthreads();
        double v[8];
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = 0.0;
        if (t.inner) {
            const int X0 = t.tx + 2, Y0 = t.ty + 2;
            double az[4] = {g.z1a[k], g.z1b[k], g.z1c[k], g.z1d[k]};
#define DXL(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], ax[0], ax[1], ax[2], ax[3])
#define DYL(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], ay[0], ay[1], ay[2], ay[3])
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
            v[0] = XD[qk]; v[1] = XD[qk + S]; v[2] = XD[qk + 2 * S]; v[3] = XD[qk + 3 * S];
            if (VD) {
                double bz[4] = {g.z2a[k], g.z2b[k], g.z2c[k], g.z2d[k]};
#define D2X(f) fd5(sp[f][Y0][X0 - 2], sp[f][Y0][X0 - 1], sp[f][Y0][X0], sp[f][Y0][X0 + 1], sp[f][Y0][X0 + 2], bx[0], bx[1], bx[2], bx[3])
#define D2Y(f) fd5(sp[f][Y0 - 2][X0], sp[f][Y0 - 1][X0], sp[f][Y0][X0], sp[f][Y0 + 1][X0], sp[f][Y0 + 2][X0], by[0], by[1], by[2], by[3])
#define D2Z(a) fd5(a[0], a[1], a[2], a[3], a[4], bz[0], bz[1], bz[2], bz[3])
                if (fmu != 0.0) {
                    double af = D2X(LVX) + D2Y(LVX) + D2Z(qvx);
                    double ag = D2X(LVY) + D2Y(LVY) + D2Z(qvy);
                    double ah = D2X(LVZ) + D2Y(LVZ) + D2Z(qvz);
                    v[1] = v[1] + fmu * af;
                    v[2] = v[2] + fmu * ag;
                    v[3] = v[3] + fmu * ah;
                    xd8 = xd8 + fmu * (fs0 * af + gs0 * ag + hs0 * ah);
                }
                if (frho != 0.0 || fpr != 0.0) {
                    v[0] = v[0] + frho * (D2X(LRH) + D2Y(LRH) + D2Z(qrh));
                    xd8 = xd8 + fpr * (D2X(LPP) + D2Y(LPP) + D2Z(qp));
                }
#undef D2X
#undef D2Y
#undef D2Z
            }
#undef DXL
#undef DYL
            v[4] = xd5; v[5] = xd6; v[6] = xd7; v[7] = xd8;
        }
        plane_rules<8, BC_STATE>(g, t, bs, v);
        if (t.own) {
#pragma unroll
            for (int m = 0; m < 8; m++) st(XD + qk + m * S, v[m]);
        }
        if (k == 2) {
#pragma unroll
            for (int m = 0; m < 8; m++) v2[m] = v[m];
        }
        if (k == g.NZ - 3) {
#pragma unroll
            for (int m = 0; m < 8; m++) vL[m] = v[m];
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 4; s++) {
            qe1[s] = qe1[s + 1]; qe2[s] = qe2[s + 1]; qp[s] = qp[s + 1]; qvz[s] = qvz[s + 1];
            if (VD) { qvx[s] = qvx[s + 1]; qvy[s] = qvy[s + 1]; qrh[s] = qrh[s + 1]; }
        }
    }
    zrules<8>(g, t, XD, v2, vL);
}

// ---------------------------------------------------------------------------
// post-step smoothing of the perturbation (t3dsdsc.f:1135-1165, avrgtt
// 5754-5862): one sweep on 8 fields of A (or of x - eq on the fly), interior
// only; its bndrytt fill runs in k_bc_tt.
// ---------------------------------------------------------------------------
template <int SRC>
__global__ __launch_bounds__(NT) void k_smooth_tt(Geo g, const double *__restrict__ A, Eq eq,
                                                  double *__restrict__ B, int kc)
{
    constexpr int NF = 8;
    __shared__ double sp[NF][TY + 2][TX + 2];
    BlockPos bp = block_pos();
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    auto val = [&](int m, long q) {
        double a = A[m * S + q];
        return (SRC == SRC_XEQ && eq.f[m]) ? a - eq.f[m][q] : a;
    };
    double zm[NF], z0[NF], zp[NF];
    if (ld) {
#pragma unroll
        for (int m = 0; m < NF; m++) { zm[m] = val(m, q0 + sz * (t.kb - 1)); z0[m] = val(m, q0 + sz * t.kb); }
    }
    const double wpx = t.inner ? g.xwp[t.i] : 0, wmx = t.inner ? g.xwm[t.i] : 0;
    const double wpy = t.inner ? g.ywp[t.j] : 0, wmy = t.inner ? g.ywm[t.j] : 0;
    const double cb = t.inner ? g.cbx[t.i] : 0.0;
    const double xy2 = t.inner ? g.xc[t.i] * g.xc[t.i] + g.yc[t.j] * g.yc[t.j] : 0.0;
    for (int k = t.kb; k < t.ke; k++) {
        const long kz = sz * k;
        if (ld) {
#pragma unroll
            for (int m = 0; m < NF; m++) zp[m] = val(m, q0 + kz + sz);
        }
        load_plane<1, NF>(g, t, sp, kz, val);
        __syncthreads();
        if (t.inner && t.own) {
            const double wpz = g.zwp[k], wmz = g.zwm[k];
            const double crad = radial_coef(xy2, g.zc[k]);
#pragma unroll
            for (int m = 0; m < NF; m++) {
                double f0 = z0[m];
                double s = difw(sp[m][t.ty + 1][t.tx], f0, sp[m][t.ty + 1][t.tx + 2], wpx, wmx) +
                           difw(sp[m][t.ty][t.tx + 1], f0, sp[m][t.ty + 2][t.tx + 1], wpy, wmy) +
                           difw(zm[m], f0, zp[m], wpz, wmz);
                double cc = (m >= 4 && m <= 6) ? cb : crad;
                st(B + m * S + q0 + kz, f0 + cc * s);
            }
        }
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NF; m++) { zm[m] = z0[m]; z0[m] = zp[m]; }
    }
}

// bndrytt (t3dsdsc.f:1989-2284) over the face points: x rule on every row
// and plane, y rule (non-periodic) and z rule for jx >= x_first only; the x
// rule reads the face values the sweep started from (A, or x - eq).
struct TTEval {
    const Geo &g;
    const double *G, *A, *X, *E;
    __device__ bool inner3(int i, int j, int k) const
    {
        return i >= 2 && i <= g.NX - 3 && j >= 2 && j <= g.NY - 3 && k >= 2 && k <= g.NZ - 3;
    }
    __device__ double P(int i, int j, int k) const
    {
        long q = IDX(g, i, rmap(g, j), k);
        if (inner3(i, j, k)) return G[q];
        if (A) return A[q];
        return E ? X[q] - E[q] : X[q];
    }
    __device__ double Xv(int i, int j, int k) const
    {
        if (i >= 2 && i <= g.NX - 3) return P(i, j, k);
        return xrule<BC_AUX>(g, i, P(2, j, k), P(3, j, k), P(4, j, k), P(g.NX - 3, j, k));
    }
    __device__ double Yv(int i, int j, int k) const
    {
        if (g.periodic || (j >= 2 && j <= g.NY - 3) || i < 2) return Xv(i, j, k);
        if (j == 0 || j == g.NY - 1) return 0.0;
        return 0.5 * Xv(i, j == 1 ? 2 : g.NY - 3, k);
    }
    __device__ double Zv(int i, int j, int k) const
    {
        if ((k >= 2 && k <= g.NZ - 3) || i < 2) return Yv(i, j, k);
        if (k == 0 || k == g.NZ - 1) return 0.0;
        return 0.5 * Yv(i, j, k == 1 ? 2 : g.NZ - 3);
    }
};

// face points over the row range [ylo, yhi]: x faces (4 per row and plane),
// non-periodic image rows over interior columns, z faces over interior
// columns and rows
__device__ inline bool face_point(const Geo &g, long t, int &i, int &j, int &k)
{
    int nrow = g.yhi - g.ylo + 1;
    long nxf = 4L * nrow * g.NZ;
    long nyf = g.periodic ? 0 : (long)(g.NX - 4) * 4 * g.NZ;
    long nzf = (long)(g.NX - 4) * (g.NY - 4) * 4;
    if (t < nxf) {
        int f = (int)(t % 4); long r = t / 4;
        j = g.ylo + (int)(r % nrow); k = (int)(r / nrow);
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

__host__ inline long face_count(const Geo &g)
{
    int nrow = g.yhi - g.ylo + 1;
    return 4L * nrow * g.NZ + (g.periodic ? 0 : (long)(g.NX - 4) * 4 * g.NZ) + (long)(g.NX - 4) * (g.NY - 4) * 4;
}

__global__ void k_bc_tt(Geo g, double *B, const double *A, const double *X, Eq eq)
{
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!face_point(g, t, i, j, k)) return;
    long q = IDX(g, i, j, k);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        TTEval ev{g, B + m * g.S, A ? A + m * g.S : nullptr, X ? X + m * g.S : nullptr, X ? eq.f[m] : nullptr};
        B[q + m * g.S] = ev.Zv(i, j, k);
    }
}

} // namespace mhd
