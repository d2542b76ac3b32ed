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

// z rules of AUX/STATE fills, applied as soon as plane 2 (NZ-3) is known:
// plane 1 (NZ-2) is half of it, plane 0 (NZ-1) is zero (t3dsdsc.f:2614-2644,
// 1821-1861).  q = offset of the column in field 0.
template <int NF>
__device__ inline void zface_store(const Geo &g, double *__restrict__ B, long q, int k, const double *v)
{
    if (k != 2 && k != g.NZ - 3) return;
    const long qh = q + g.sz * (k == 2 ? 1 : g.NZ - 2), q0 = q + g.sz * (k == 2 ? 0 : g.NZ - 1);
    #pragma unroll
    for (int m = 0; m < NF; m++) {
        st(B + qh + m * g.S, 0.5 * v[m]);
        st(B + q0 + m * g.S, 0.0);
    }
}

// ---------------------------------------------------------------------------
// current (t3dsdsc.f:2465-2491) + bndryc: J = curl(B - B_i)
// Iteration t loads plane t of B - B_i into LDS (halo 2), forms the x/y
// derivative parts of plane t and pushes the column's bx - bxi, by - byi
// into the z queues; plane t-2 is completed from the queues.  Every value of
// B and B_i is loaded once per plane (planes outside the chunk: the column
// only), and the sums keep the one-pass order of operations.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void curl_body(Geo g, const double *__restrict__ X, const double *__restrict__ bxi,
                                          const double *__restrict__ byi, const double *__restrict__ bzi,
                                          double *__restrict__ cur, int kc, BlockPos bp)
{
    __shared__ double sp[2][3][TY + 4][TX + 4];   // plane t by parity
    __shared__ BCScratch<3> bs;
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S;
    const unsigned sz = (unsigned)g.sz;
    const bool ld = loadable(g, t.i, t.j);
    // element offsets (fields hold fewer than 2^29 points, checked at create)
    const unsigned e0 = ld ? (unsigned)t.i + (unsigned)g.sy * (unsigned)rmap(g, t.j) : 0u;
    const double *b5 = X + 4 * S, *b6 = X + 5 * S, *b7 = X + 6 * S;
    double uq[5] = {0, 0, 0, 0, 0}, vq[5] = {0, 0, 0, 0, 0};   // planes t-4..t
    double pa[3] = {0, 0, 0}, pb[3] = {0, 0, 0}, pc[3] = {0, 0, 0};   // planes t-2..t
    double ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0};
    if (t.inner) {
        ax[0] = g.x1a[t.i]; ax[1] = g.x1b[t.i]; ax[2] = g.x1c[t.i]; ax[3] = g.x1d[t.i];
        ay[0] = g.y1a[t.j]; ay[1] = g.y1b[t.j]; ay[2] = g.y1c[t.j]; ay[3] = g.y1d[t.j];
    }
    for (int tp = t.kb - 2; tp < t.ke + 2; tp++) {
        const unsigned kz = sz * (unsigned)tp;
        double u = 0.0, w = 0.0, xa = 0.0, xb = 0.0, xc = 0.0;
        if (tp >= t.kb && tp < t.ke) {
            // block-uniform: one barrier per plane (the LDS planes alternate)
            double (*sq)[TY + 4][TX + 4] = sp[tp & 1];
            load_plane<2, 3>(g, t, sq, (long)kz, [&](int m, long q) {
                const unsigned e = (unsigned)q;
                return m == 0 ? fld(b5, e) - fld(bxi, e) : m == 1 ? fld(b6, e) - fld(byi, e) : fld(b7, e) - fld(bzi, e);
            });
            __syncthreads();
            const int X0 = t.tx + 2, Y0 = t.ty + 2;
            if (ld) { u = sq[0][Y0][X0]; w = sq[1][Y0][X0]; }
            if (t.inner) {
#define DXS(f) fd5(sq[f][Y0][X0 - 2], sq[f][Y0][X0 - 1], sq[f][Y0][X0], sq[f][Y0][X0 + 1], sq[f][Y0][X0 + 2], ax[0], ax[1], ax[2], ax[3])
#define DYS(f) fd5(sq[f][Y0 - 2][X0], sq[f][Y0 - 1][X0], sq[f][Y0][X0], sq[f][Y0 + 1][X0], sq[f][Y0 + 2][X0], ay[0], ay[1], ay[2], ay[3])
                xa = DYS(2);
                xb = DXS(2);
                xc = DXS(1) - DYS(0);
#undef DXS
#undef DYS
            }
        } else if (ld) {
            const unsigned e = e0 + kz;
            u = fld(b5, e) - fld(bxi, e);
            w = fld(b6, e) - fld(byi, e);
        }
        #pragma unroll
        for (int s = 0; s < 4; s++) { uq[s] = uq[s + 1]; vq[s] = vq[s + 1]; }
        uq[4] = u; vq[4] = w;
        pa[0] = pa[1]; pa[1] = pa[2]; pa[2] = xa;
        pb[0] = pb[1]; pb[1] = pb[2]; pb[2] = xb;
        pc[0] = pc[1]; pc[1] = pc[2]; pc[2] = xc;
        const int k = tp - 2;
        if (k < t.kb) continue;   // block-uniform
        double v[3] = {0, 0, 0};
        if (t.inner) {
            const double az0 = g.z1a[k], az1 = g.z1b[k], az2 = g.z1c[k], az3 = g.z1d[k];
            const double vz = fd5(vq[0], vq[1], vq[2], vq[3], vq[4], az0, az1, az2, az3);
            const double uz = fd5(uq[0], uq[1], uq[2], uq[3], uq[4], az0, az1, az2, az3);
            v[0] = pa[0] - vz;
            v[1] = uz - pb[0];
            v[2] = pc[0];
        }
        // past the chunk no plane load separates two calls of the rules,
        // whose LDS slots the previous call may still be reading
        if (tp >= t.ke) __syncthreads();
        plane_rules<3, BC_AUX>(g, t, bs, v);
        if (t.own) {
            const unsigned e = e0 + sz * (unsigned)k;
            st(fldp(cur, e), v[0]); st(fldp(cur + S, e), v[1]); st(fldp(cur + 2 * S, e), v[2]);
            zface_store<3>(g, cur, (long)e0, k, v);
        }
    }
}

// ---------------------------------------------------------------------------
// lane helpers of the row-per-wavefront sweeps
// ---------------------------------------------------------------------------
// value held by lane l of the wavefront (l uniform over the wavefront)
__device__ inline double lane_of(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned int)lo);
}

// x-face values of a row from the lanes holding columns 2, 3, 4 (left) and
// NX-3 (right); i0 = column of lane 0.  Every lane of the wavefront must
// call; only windows that hold a face pay for its rule.
template <int TYPE>
__device__ inline double row_xrule_rl(const Geo &g, int i, int i0, bool apply, double v)
{
    double out = v;
    if (i0 + TX > g.NX - 2) {       // right face in the window
        const double pL = lane_of(v, min(max(g.NX - 3 - i0, 0), TX - 1));
        if (apply && i == g.NX - 1) out = 0.0;
        else if (apply && i == g.NX - 2) out = 0.5 * pL;
    }
    if (i0 < 2) {                   // left face in the window
        const int l2 = min(max(2 - i0, 0), TX - 3);
        const double p2 = lane_of(v, l2);
        if (TYPE == BC_STATE) {
            if (apply && i == 0) out = 0.0;
            else if (apply && i == 1) out = 0.5 * p2;
        } else {
            const double x1 = xleft1(g, p2, lane_of(v, l2 + 1), lane_of(v, l2 + 2));
            if (apply && i == 1) out = x1;
            else if (apply && i == 0) out = xleft0(g, x1, p2, lane_of(v, l2 + 1));
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// z-unrolled point-parallel smoothing sweep (the production sweep).  Thread
// (i, j) computes UZ consecutive planes [k0, k0+UZ) of FPB fields (one field
// at a time); the column values of planes k0-1..k0+UZ are loaded once and
// shared between the planes.  The plane groups tile [0, NZ) from plane 0;
// the last group is shifted back to end at NZ-1 (it writes only its own
// planes), so the z-face planes 0, 1 (NZ-2, NZ-1) always sit in the group
// that also computes plane 2 (NZ-3).  Stores are non-temporal: the output is
// not re-read by this sweep.
// Block order (remap != 0): x tile fastest, then y tile, plane group and
// field block, handed to the XCDs in contiguous ranges (block_pos), so the
// tiles that share halo rows and planes meet in one L2.
// With RKS >= 0 the sweep is the second avrg pass of RK stage RKS and ends
// in that stage's register update (t3dsdsc.f:1090-1132) instead of storing
// xdif.
// ---------------------------------------------------------------------------
constexpr int BY = 8;
constexpr int SWEEP_UZ = 4;

// interior planes per chunk of k_zsmooth: a launch argument up to this
constexpr int ZKC_MAX = 96;

struct RKArgs {
    const double *dtp;     // device dt
    const double *XF;      // xfold
    double *XM, *X;        // xm, x
    const double *EQ;      // equilibrium rho, bx, by, bz, p (field stride S)
};

template <int RKS>
__device__ inline void rk_point(const RKArgs &r, long q, double d, double dt)
{
    if (RKS == 0) {
        double xf = r.XF[q];
        st(r.XM + q, xf + div6(d * dt));
        st(r.X + q, xf + d * dt / 2.);
    } else if (RKS == 1) {
        st(r.XM + q, r.XM[q] + div3(d * dt));
        st(r.X + q, r.XF[q] + d * dt / 2.);
    } else if (RKS == 2) {
        st(r.XM + q, r.XM[q] + div3(d * dt));
        st(r.X + q, r.XF[q] + d * dt);
    } else {
        st(r.X + q, r.XM[q] + div6(d * dt));
    }
}

template <int NF, int KIND, int TYPE, int FPB, int UZ, int RKS>
__global__ __launch_bounds__(TX * BY) void k_sweepz(Geo g, const double *__restrict__ A, double *__restrict__ B,
                                                    RKArgs rk, int remap, const double *__restrict__ addf)
{
    BlockPos bp;
    if (remap) bp = block_pos();
    else { bp.bx = blockIdx.x; bp.by = blockIdx.y; bp.bz = blockIdx.z; }
    // FPB fields per block: bz = plane group + ngz * field block
    const int ngz = (g.NZ + UZ - 1) / UZ;
    const int mb = (bp.bz / ngz) * FPB;
    const int lane = threadIdx.x % TX, row = threadIdx.x / TX;
    const int nomx = bp.bx * TX;
    const int i0 = g.NX >= TX ? min(nomx, g.NX - TX) : 0;
    const int nyt = g.yhi - g.ylo + 1;
    const int nomy = g.ylo + bp.by * BY;
    const int j0 = nyt >= BY ? min(nomy, g.yhi + 1 - BY) : g.ylo;
    const int nomz = (bp.bz % ngz) * UZ;
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
    const double dt = RKS >= 0 ? *rk.dtp : 0.0;
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
            if (xft) v[u] = row_xrule_rl<TYPE>(g, i, i0, valid, v[u]);
            if (image) v[u] = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * v[u];
        }
        // z-face planes from plane 2 / NZ-3 of the same group
        #pragma unroll
        for (int u = 0; u < UZ; u++) {
            int k = k0 + u;
            if (k == 0 || k == g.NZ - 1) v[u] = 0.0;
            else if (k == 1 && u + 1 < UZ) v[u] = 0.5 * v[u + 1];
            else if (k == g.NZ - 2 && u >= 1) v[u] = 0.5 * v[u - 1];
        }
        if (own) {
            const long qb = m * S + (long)i + sy * j;
            #pragma unroll
            for (int u = 0; u < UZ; u++) {
                if (k0 + u < nomz) continue;
                const long qq = qb + sz * (k0 + u);
                // addf: the sweep's output plus a field (E + xdh for the curl of right)
                if (RKS < 0) st(B + qq, addf ? addf[qq] + v[u] : v[u]);
                else rk_point<RKS>(rk, qq, v[u], dt);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Temporal blocking of the smoothing chains: NP consecutive passes of one
// smoothing kind (with the boundary fill after each pass) on one field, in
// one z-marching block.  A block of 64 x ZB threads covers a window of 64
// columns x ZB rows; "level p" is the field after p passes.  Iteration t
// brings plane t of level 0 (loaded ZPF planes ahead) and forms level p at
// plane t-p, p = 1..NP:
//   z neighbours  the thread's own register queue of level p-1,
//   x, y neighbours  level p-1's plane in LDS (written one iteration earlier,
//                 so one barrier per iteration suffices; rows padded by a
//                 column on either side).
// Each pass costs the window one column and row at every edge that is not a
// domain face, so a tile owns the inner (64-2NP) x (ZB-2NP) columns x rows.
// Every level gets the boundary rules of its pass: x faces from columns 2,
// 3, 4, NX-3 of the row (lane reads), non-periodic image rows from rows 2 /
// NY-3 (an extra LDS exchange), z faces from planes 2 / NZ-3 (the register
// queue of planes 0, 1 is rewritten once plane 2 is known).  The output of
// the last level is stored, or with RKS >= 0 the RK stage update is applied
// to it (the second avrg pass of the stage).
// The grid is one-dimensional: the host's block table gives the tile, chunk
// and field of each dispatch position, so the order in which the XCDs
// receive the blocks is the host's choice.
// ---------------------------------------------------------------------------
constexpr int ZB = 16;            // rows (wavefronts) per block
// block table word: x tile | y tile << ZT_BY | chunk << ZT_CH | field << ZT_F
constexpr int ZT_BY = 10, ZT_CH = 20, ZT_F = 29;
// launch modes: every window from a gx x gy x (nch * nf) grid with the full
// march; from a block table, the windows without x faces or image rows, or
// the windows with them.  The table modes use the fast plane groups.  The
// face windows' march needs more than 64 registers (6 waves per SIMD): with
// periodic y their blocks have RB = 12 rows, so two fit a CU.
enum { ZM_ALL = 0, ZM_PLAIN = 1, ZM_FACE = 2 };
// RKS value of a sweep that stores its output plus a field (E + xdh after the
// last avrgh pass, t3dsdsc.f:1426-1433); the field comes in rk.EQ
constexpr int RKS_ADD = -2;
constexpr int ZPF = 3;            // planes loaded ahead

template <int NP, int KIND, int TYPE, int RKS, int MODE, int RB = ZB>
__device__ __forceinline__ void zsmooth_body(const Geo &g, const double *__restrict__ A, double *__restrict__ B,
                                             const RKArgs &rk, int bx, int by, int chunk, int m, int kc)
{
    static_assert(NP >= 1 && (NP <= 2 || KIND == KIND_EIGHTH), "the radial coefficient ring holds planes t-2..t");
    static_assert(NP <= 3 && (RKS != RKS_ADD || NP == 3), "passes per launch");
    constexpr int W = TX - 2 * NP, H = RB - 2 * NP;
    constexpr int NZW = ZKC_MAX + 4 * NP;      // z-weight planes of the longest chunk
    constexpr bool RAD = KIND != KIND_EIGHTH;   // reads a smoothing coefficient stream
    constexpr bool TT = TYPE == BC_TT;          // post-step passes on x - eq (bndrytt fills)
    constexpr bool ADD = RKS == RKS_ADD;        // store the output plus a field
    constexpr int XT = TT ? BC_AUX : TYPE;      // x-face rule of the pass
    // the plane levels in LDS, column l of a row at index l + 1: the x and y
    // neighbours of a point are read there (the two pad columns feed only
    // the window's edge lanes, whose values no owned point uses)
    __shared__ double s0[3][RB][TX + 2];       // level 0 by plane mod 3
    __shared__ double sl[NP][2][RB][TX + 2];   // level p (1..NP) by plane parity
    __shared__ double szw[2][NZW];             // z weights of the chunk's planes
    const int lane = threadIdx.x % TX;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);   // uniform over the wavefront
    // columns: lane l holds column i0 + l
    const int nomx = bx * W;
    const int i0 = max(0, min(nomx - NP, g.NX - TX));
    const int i = i0 + lane;
    // the march, compiled for windows with face rules (FACES) and without
    auto march = [&](auto faces_tag) {
        constexpr bool FACES = decltype(faces_tag)::value;
        // FACES false: the window holds no x face and no image row (the
        // host's block table says so), every lane is an interior point
        const bool vx = !FACES || i < g.NX;
        const bool ownx = vx && i >= nomx && i < nomx + W;
        // rows: wavefront r holds logical row j0 + r (periodic rows map through rmap)
        const int nomy = g.ylo + by * H;
        const int j0 = g.periodic ? max(g.ylo - NP, min(nomy - NP, g.yhi + NP + 1 - RB))
                                  : max(0, min(nomy - NP, g.NY - RB));
        const int j = j0 + r;
        const bool vy = !FACES || g.periodic || j < g.NY;
        const bool own = ownx && vy && j >= nomy && j < nomy + H && j <= g.yhi;
        // the block's owned rows are r in [olo, ohi) (NP..RB-1-NP, narrower in a
        // clamped last row tile); level p is read only within NP-p rows of them
        const int olo = nomy - j0, ohi = min(nomy + H, g.yhi + 1) - j0;
        const bool ld = vx && vy;
        const int jp = vy ? rmap(g, j) : 0;
        const bool yin = !FACES || g.periodic || (j >= 2 && j <= g.NY - 3);
        const bool pin = !FACES || (ld && yin && i >= 2 && i <= g.NX - 3);   // the stencil applies here
        const bool yex = FACES && !g.periodic;         // image rows exchanged through LDS
        const bool img = yex && vy && (j < 2 || j > g.NY - 3);
        const int rsrc = img ? (j < 2 ? 2 - j0 : g.NY - 3 - j0) : 0;
        const bool xft = FACES && (i0 < 2 || i0 + TX > g.NX - 2);       // the window holds an x face
        const long S = g.S;
        const unsigned sz = (unsigned)g.sz;
        // every lane reads a valid address; lanes outside the range read column 0
        const unsigned qc = ld ? (unsigned)i + (unsigned)g.sy * (unsigned)jp : 0u;
        const double *a = A + m * S;                // wave-uniform field bases
        const double *crad = g.crad;
        const double wpx = pin ? g.xwp[i] : 0.0, wmx = pin ? g.xwm[i] : 0.0;
        const double wpy = pin ? g.ywp[jp] : 0.0, wmy = pin ? g.ywm[jp] : 0.0;
        const bool bcomp = KIND == KIND_STATE && m >= 4 && m <= 6;
        // smoothing coefficient stream of the field: the radial table, or
        // for the magnetic field of the state the x-only cbx (plane stride 0,
        // so those loads hit the cache)
        const double *cbase = bcomp ? g.cbx : crad;
        const unsigned cq0 = bcomp ? (unsigned)i : qc, csz = bcomp ? 0u : sz;

        const int kb = 2 + chunk * kc, ke = min(kb + kc, g.NZ - 2);
        const int t0 = kb - NP, t1 = ke - 1 + NP;
        const int tl = min(t1, g.NZ - 1);           // last plane loaded (three passes reach past NZ-1)
        const int kz0 = t0 - NP;                    // lowest plane a level touches
        for (int n = threadIdx.x; n < kc + 4 * NP; n += TX * RB) {
            const int k = kz0 + n;
            const bool kin = k >= 2 && k <= g.NZ - 3;
            szw[0][n] = kin ? g.zwp[k] : 0.0;
            szw[1][n] = kin ? g.zwm[k] : 0.0;
        }
        const double dt = RKS >= 0 ? *rk.dtp : 0.0;
        // owned points are unique rows, where rmap(j) == j: the output offset is qc
        double *bo = RKS < 0 ? B + m * S : nullptr;
        const double *xf = RKS >= 0 ? rk.XF + m * S : nullptr;
        double *xmo = RKS >= 0 ? rk.XM + m * S : nullptr;
        double *xo = RKS >= 0 ? rk.X + m * S : nullptr;
        // equilibrium of field m (rho, B, p; none for the momentum)
        const double *eqm = ((TT || RKS == 4) && (m == 0 || m >= 4)) ? rk.EQ + (m == 0 ? 0 : m - 3) * S : nullptr;
        // the field added to the output (ADD: rk.EQ holds a 3-field array)
        const double *addm = ADD ? rk.EQ + m * S : nullptr;
        // final value v of point e: store it (plus eq after the post-step
        // passes, t3dsdsc.f:1157-1164), or apply the RK stage update; RKS 4 is the
        // last stage followed by xm = x - eq (t3dsdsc.f:1131, 1138-1145)
        // the same in two parts: put_pre reads the operands (eq, xfold, xm) that
        // put_with combines with v
        auto put_pre = [&](unsigned e, double &o1, double &o2) {
            o1 = o2 = 0.0;
            if (RKS < 0) { if ((TT && eqm) || ADD) o1 = fld(ADD ? addm : eqm, e); }
            else if (RKS == 0) o1 = fld(xf, e);
            else if (RKS <= 2) { o1 = fld(xmo, e); o2 = fld(xf, e); }
            else if (RKS == 3) o1 = fld(xmo, e);
            else { o1 = fld(xmo, e); if (eqm) o2 = fld(eqm, e); }
        };
        auto put_with = [&](unsigned e, double v, double o1, double o2) {
            if (RKS < 0) st(fldp(bo, e), ((TT && eqm) || ADD) ? v + o1 : v);
            else if (RKS == 0) { st(fldp(xmo, e), o1 + div6(v * dt)); st(fldp(xo, e), o1 + v * dt / 2.); }
            else if (RKS == 1) { st(fldp(xmo, e), o1 + div3(v * dt)); st(fldp(xo, e), o2 + v * dt / 2.); }
            else if (RKS == 2) { st(fldp(xmo, e), o1 + div3(v * dt)); st(fldp(xo, e), o2 + v * dt); }
            else if (RKS == 3) st(fldp(xo, e), o1 + div6(v * dt));
            else { const double xn = o1 + div6(v * dt); st(fldp(xmo, e), eqm ? xn - o2 : xn); }
        };
        auto put = [&](unsigned e, double v) {
            if (RKS < 0) st(fldp(bo, e), ((TT && eqm) || ADD) ? v + fld(ADD ? addm : eqm, e) : v);
            else if (RKS == 0) { const double f = fld(xf, e); st(fldp(xmo, e), f + div6(v * dt)); st(fldp(xo, e), f + v * dt / 2.); }
            else if (RKS == 1) { st(fldp(xmo, e), fld(xmo, e) + div3(v * dt)); st(fldp(xo, e), fld(xf, e) + v * dt / 2.); }
            else if (RKS == 2) { st(fldp(xmo, e), fld(xmo, e) + div3(v * dt)); st(fldp(xo, e), fld(xf, e) + v * dt); }
            else if (RKS == 3) st(fldp(xo, e), fld(xmo, e) + div6(v * dt));
            else { const double xn = fld(xmo, e) + div6(v * dt); st(fldp(xmo, e), eqm ? xn - fld(eqm, e) : xn); }
        };
        // x-left columns of a row from the level held by lanes 2, 3, 4 (every
        // lane of the wavefront calls; the window holds the left face)
        auto tt_left = [&](double lv, double v) {
            const int l2 = 2 - i0;
            const double p2 = lane_of(lv, l2), p3 = lane_of(lv, l2 + 1);
            const double x1 = xleft1(g, p2, p3, lane_of(lv, l2 + 2));
            return i == 1 ? x1 : i == 0 ? xleft0(g, x1, p2, p3) : v;
        };

        double q[NP + 1][3];   // level p at its three latest planes, oldest first
        double cr[3];          // radial coefficient of planes t-2, t-1, t
        double pf[ZPF], pc[ZPF];
        #pragma unroll
        for (int p = 0; p <= NP; p++) q[p][0] = q[p][1] = q[p][2] = 0.0;
        cr[0] = cr[1] = cr[2] = 0.0;
        if (RAD && t0 >= 1) cr[2] = fld(cbase, cq0 + csz * (unsigned)(t0 - 1));
        #pragma unroll
        for (int d = 0; d < ZPF; d++) {
            // three passes start at plane t0 = -1 on the first chunk: planes
            // below 0 load plane 0 (they feed only z-face planes, which the
            // rules overwrite)
            const unsigned td = (unsigned)min(max(t0 + d, 0), tl);
            pf[d] = fld(a, qc + sz * td);
            pc[d] = RAD ? fld(cbase, cq0 + csz * td) : 0.0;
        }

        // one plane of the march with every rule: x faces, image rows, z faces
        auto step = [&](const int u, const int t) {
            const double v0 = ld ? pf[u] : 0.0;
            const double cv = pc[u];
            {
                const unsigned tz = (unsigned)min(t + ZPF, tl);
                pf[u] = fld(a, qc + sz * tz);
                if (RAD) pc[u] = fld(cbase, cq0 + csz * tz);
            }
            q[0][0] = q[0][1]; q[0][1] = q[0][2]; q[0][2] = v0;
            s0[(t + 3) % 3][r][lane + 1] = v0;
            if (RAD) { cr[0] = cr[1]; cr[1] = cr[2]; cr[2] = cv; }
            __syncthreads();
            #pragma unroll
            for (int p = 1; p <= NP; p++) {
                // a row farther than NP-p rows from the owned ones skips level p
                // (nothing reads it), unless image rows exchange through LDS with
                // a block-wide barrier.  The RK-stage sweeps skip only below the
                // last level (fewer spills there: 5.12 ms against 5.40 with every
                // row, 5.62 skipping both levels, at 512x256x512)
                if (!yex && (r < olo - (NP - p) || r >= ohi + (NP - p)) && (RKS < 0 || p < NP)) continue;
                const int k = t - p;                        // plane of level p
                const bool kin = k >= 2 && k <= g.NZ - 3;
                const int rm = r > 0 ? r - 1 : 0, rp = r < RB - 1 ? r + 1 : RB - 1;
                const double f0 = q[p - 1][1];              // level p-1 at plane k
                double ym, yp, xm, xp;
                if (p == 1) {
                    const int b = (k + 3) % 3;
                    ym = s0[b][rm][lane + 1];
                    yp = s0[b][rp][lane + 1];
                    xm = s0[b][r][lane];
                    xp = s0[b][r][lane + 2];
                } else {
                    ym = sl[p - 2][k & 1][rm][lane + 1];
                    yp = sl[p - 2][k & 1][rp][lane + 1];
                    xm = sl[p - 2][k & 1][r][lane];
                    xp = sl[p - 2][k & 1][r][lane + 2];
                }
                double v = 0.0;
                if (pin && kin) {
                    const double c = KIND == KIND_EIGHTH ? 0.125 : cr[RAD ? 2 - p : 0];
                    const double s = difw(xm, f0, xp, wpx, wmx) + difw(ym, f0, yp, wpy, wmy) +
                                     difw(q[p - 1][0], f0, q[p - 1][2], szw[0][k - kz0], szw[1][k - kz0]);
                    v = f0 + c * s;
                }
                if (xft) v = row_xrule_rl<XT>(g, i, i0, vx && yin, v);
                if (yex && kin) {
                    // image rows: 0 on the outer row, half of row 2 (NY-3) on the next
                    sl[p - 1][k & 1][r][lane + 1] = v;
                    __syncthreads();
                    if (img) {
                        v = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * sl[p - 1][k & 1][rsrc][lane + 1];
                        sl[p - 1][k & 1][r][lane + 1] = v;
                    }
                } else if (p < NP) {
                    sl[p - 1][k & 1][r][lane + 1] = v;
                }
                // z faces above the interior: plane NZ-2 is half of NZ-3, NZ-1 is 0
                if (k == g.NZ - 2) v = 0.5 * q[p][2];
                else if (k == g.NZ - 1) v = 0.0;
                // bndrytt skips x < x_first in its y and z rules: on image rows
                // and z-face planes the x-left columns extrapolate this pass's
                // input of that row and plane (t3dsdsc.f:2108-2168)
                if (TT && FACES && i0 < 2 && (img || !kin)) v = tt_left(f0, v);
                q[p][0] = q[p][1]; q[p][1] = q[p][2]; q[p][2] = v;
                // below: planes 1 and 0 follow from plane 2 (bndrytt: the x-left
                // columns of plane 1 from the input's plane 1; plane 0 stays 0
                // as the input is 0 there from the second pass on)
                if (k == 2 && p < NP) {
                    q[p][1] = 0.5 * v;
                    q[p][0] = 0.0;
                    if (TT && FACES && i0 < 2) q[p][1] = tt_left(q[p - 1][0], q[p][1]);
                }
                if (p == NP) {
                    // the last level's z-face planes: 1 (NZ-2) is half of plane
                    // 2 (NZ-3), 0 (NZ-1) is 0; under bndrytt the x-left columns
                    // of planes 1 and NZ-2 extrapolate the previous level there
                    const bool lo = k == 2 && kb == 2, hi = k == g.NZ - 3 && ke == g.NZ - 2;
                    double vf = 0.5 * v;
                    if (TT && FACES && i0 < 2 && (lo || hi)) vf = tt_left(lo ? q[p - 1][0] : q[p - 1][2], vf);
                    if (own) {
                        if (k >= kb && k < ke) put(qc + sz * (unsigned)k, v);
                        if (lo || hi) {
                            put(qc + sz * (unsigned)(lo ? 1 : g.NZ - 2), vf);
                            put(qc + sz * (unsigned)(lo ? 0 : g.NZ - 1), 0.0);
                        }
                    }
                }
            }
        };

        // plane u of a group of ZPF planes none of whose levels touches a z-face
        // rule (3 <= t-p <= NZ-4): the stencil with the x-face and image-row
        // rules of the window (none without FACES).  The queues are rings here:
        // after the group's u-th push the oldest plane of a level sits in slot
        // (u+1) % 3, the newest in u, so a whole group leaves them in the order
        // the full step keeps.
        auto fast_step = [&](const int u, const int t) {
            // the output plane's operands, read before the passes (every lane
            // reads an address inside the volume: plane t-NP >= 3) and before
            // the prefetch, so that waiting for them leaves the prefetch in flight
            const unsigned eo = qc + sz * (unsigned)(t - NP);
            double o1, o2;
            put_pre(eo, o1, o2);
            const double v0 = ld ? pf[u] : 0.0;
            const double cv = pc[u];
            {
                const unsigned tz = (unsigned)min(t + ZPF, tl);
                pf[u] = fld(a, qc + sz * tz);
                if (RAD) pc[u] = fld(cbase, cq0 + csz * tz);
            }
            q[0][u] = v0;
            s0[(t + 3) % 3][r][lane + 1] = v0;
            if (RAD) cr[u] = cv;
            __syncthreads();
            #pragma unroll
            for (int p = 1; p <= NP; p++) {
                if (!yex && (r < olo - (NP - p) || r >= ohi + (NP - p))) continue;   // as in step
                const int k = t - p;
                const int rm = r > 0 ? r - 1 : 0, rp = r < RB - 1 ? r + 1 : RB - 1;
                const double zm = q[p - 1][(u + 1) % 3], f0 = q[p - 1][(u + 2) % 3], zp = q[p - 1][u];
                double ym, yp, xm, xp;
                if (p == 1) {
                    const int b = (k + 3) % 3;
                    ym = s0[b][rm][lane + 1];
                    yp = s0[b][rp][lane + 1];
                    xm = s0[b][r][lane];
                    xp = s0[b][r][lane + 2];
                } else {
                    ym = sl[p - 2][k & 1][rm][lane + 1];
                    yp = sl[p - 2][k & 1][rp][lane + 1];
                    xm = sl[p - 2][k & 1][r][lane];
                    xp = sl[p - 2][k & 1][r][lane + 2];
                }
                const double c = KIND == KIND_EIGHTH ? 0.125 : cr[RAD ? (u + 3 - p) % 3 : 0];
                const double s = difw(xm, f0, xp, wpx, wmx) + difw(ym, f0, yp, wpy, wmy) +
                                 difw(zm, f0, zp, szw[0][k - kz0], szw[1][k - kz0]);
                double v = pin ? f0 + c * s : 0.0;
                if (FACES) {
                    if (xft) v = row_xrule_rl<XT>(g, i, i0, vx && yin, v);
                    if (yex) {
                        sl[p - 1][k & 1][r][lane + 1] = v;
                        __syncthreads();
                        if (img) {
                            v = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * sl[p - 1][k & 1][rsrc][lane + 1];
                            sl[p - 1][k & 1][r][lane + 1] = v;
                        }
                        if (TT && i0 < 2 && img) v = tt_left(f0, v);
                    } else if (p < NP) {
                        sl[p - 1][k & 1][r][lane + 1] = v;
                    }
                } else if (p < NP) {
                    sl[p - 1][k & 1][r][lane + 1] = v;
                }
                q[p][u] = v;
                if (p == NP && own && k >= kb && k < ke) put_with(eo, v, o1, o2);
            }
        };

        // the march runs in groups of ZPF planes so that every prefetch slot is a
        // fixed register (the loads stay in flight); planes past t1 are dummies
        // that load plane t1 again and store nothing
        static_assert(ZPF == 3, "the fast step's rings rotate once per group of 3 planes");
        const int tend = t0 + ((t1 - t0 + ZPF) / ZPF) * ZPF;
        // the plane groups that may use the fast step form one run
        auto fast = [&](int b) { return b >= NP + 3 && b + ZPF - 1 <= g.NZ - 3; };
        int tb = t0;
        if (MODE == ZM_PLAIN || MODE == ZM_FACE) {
        for (; tb < tend && !fast(tb); tb += ZPF) {
            #pragma unroll
            for (int u = 0; u < ZPF; u++) step(u, tb + u);
        }
        for (; tb < tend && fast(tb); tb += ZPF) {
            #pragma unroll
            for (int u = 0; u < ZPF; u++) fast_step(u, tb + u);
        }
        }
        for (; tb < tend; tb += ZPF) {
            #pragma unroll
            for (int u = 0; u < ZPF; u++) step(u, tb + u);
        }
    };
    march(std::integral_constant<bool, MODE != ZM_PLAIN>{});
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

// e: element offset of the point (wave-uniform field bases, 32-bit offsets)
__device__ inline Rec rec_at(const double *__restrict__ X, const double *__restrict__ pri, long S, unsigned e)
{
    Rec r;
    double x1 = fld(X, e);
    r.x2 = fld(X + S, e); r.x3 = fld(X + 2 * S, e); r.x4 = fld(X + 3 * S, e); r.p = fld(X + 7 * S, e);
    double dp = r.p - fld(pri, e);
    // the six quotients by rho from one rounded reciprocal (divc: equal to
    // the IEEE quotients)
    const double y = 1.0 / x1;
    r.t22 = divc(r.x2 * r.x2, x1, y) + dp;
    r.t23 = divc(r.x2 * r.x3, x1, y);
    r.t24 = divc(r.x2 * r.x4, x1, y);
    r.t33 = divc(r.x3 * r.x3, x1, y) + dp;
    r.t34 = divc(r.x4 * r.x3, x1, y);
    r.t44 = divc(r.x4 * r.x4, x1, y) + dp;
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

// E at point q from the momentum (x2, x3, x4) the caller already holds
__device__ inline void e_point_m(const double *__restrict__ X, const double *__restrict__ cur, double eta, long S,
                                 unsigned e, double x2, double x3, double x4, double *__restrict__ E)
{
    double x1 = fld(X, e);
    double x5 = fld(X + 4 * S, e), x6 = fld(X + 5 * S, e), x7 = fld(X + 6 * S, e);
    const double y = 1.0 / x1;
    st(fldp(E, e), divc(-x3 * x7 + x4 * x6, x1, y) + eta * fld(cur, e));
    st(fldp(E + S, e), divc(-x4 * x5 + x2 * x7, x1, y) + eta * fld(cur + S, e));
    st(fldp(E + 2 * S, e), divc(-x2 * x6 + x3 * x5, x1, y) + eta * fld(cur + 2 * S, e));
}

__device__ inline void e_point(const double *__restrict__ X, const double *__restrict__ cur, double eta, long S,
                               unsigned e, double *__restrict__ E)
{
    e_point_m(X, cur, eta, S, e, fld(X + S, e), fld(X + 2 * S, e), fld(X + 3 * S, e), E);
}

// GT: Geo (a kernel argument, copied) or const Geo & (a run's block of a batch)
template <class GT>
__device__ __forceinline__ void rhs_a_body(GT g, const double *__restrict__ X, const double *__restrict__ pri,
                                           const double *__restrict__ cur, const double *__restrict__ xresd,
                                           double eta, int hall, double cht, double *__restrict__ XD,
                                           double *__restrict__ E, double *__restrict__ XH, int kc, BlockPos bp)
{
    __shared__ double sp[NLA][TY + 4][TX + 4];
    __shared__ BCScratch<3> bs;
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S;
    const unsigned sz = (unsigned)g.sz;
    const bool ld = loadable(g, t.i, t.j);
    // element offsets (fields hold fewer than 2^29 points, checked at create)
    const unsigned e0 = ld ? (unsigned)t.i + (unsigned)g.sy * (unsigned)rmap(g, t.j) : 0u;
    if (t.own) {   // E on the planes outside the interior z range
        if (t.kb == 2) { e_point(X, cur, eta, S, e0, E); e_point(X, cur, eta, S, e0 + sz, E); }
        if (t.ke == g.NZ - 2) {
            e_point(X, cur, eta, S, e0 + sz * (unsigned)(g.NZ - 2), E);
            e_point(X, cur, eta, S, e0 + sz * (unsigned)(g.NZ - 1), E);
        }
    }
    double qx4[5], qt24[5], qt34[5], qt44[5], qp[5];
    double yx2[3], yx3[3], yt22[3], yt23[3], yt33[3];
    if (ld) {
        #pragma unroll
        for (int s = 0; s < 4; s++) {
            Rec r = rec_at(X, pri, S, e0 + sz * (unsigned)(t.kb - 2 + s));
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
    for (int k = t.kb; k < t.ke; k++) {
        const unsigned kz = sz * (unsigned)k;
        const unsigned ek = e0 + kz;
        if (ld) {
            Rec r = rec_at(X, pri, S, ek + 2 * sz);
            qx4[4] = r.x4; qt24[4] = r.t24; qt34[4] = r.t34; qt44[4] = r.t44; qp[4] = r.p;
            yx2[2] = r.x2; yx3[2] = r.x3; yt22[2] = r.t22; yt23[2] = r.t23; yt33[2] = r.t33;
            const int y = t.ty + 2, x = t.tx + 2;
            sp[LX2][y][x] = yx2[0]; sp[LX3][y][x] = yx3[0]; sp[LT22][y][x] = yt22[0]; sp[LT23][y][x] = yt23[0];
            sp[LT24][y][x] = qt24[2]; sp[LT33][y][x] = yt33[0]; sp[LT34][y][x] = qt34[2]; sp[LP][y][x] = qp[2];
        }
        // plane k's momentum is in the z queues (loaded two planes ago)
        if (t.own) e_point_m(X, cur, eta, S, ek, yx2[0], yx3[0], qx4[2], E);
        // halo only (the own value is in place already)
        if (t.tx < 2 || t.tx >= TX - 2) {
            int hx = t.tx < 2 ? -2 : 2;
            if (loadable(g, t.i + hx, t.j)) {
                Rec r = rec_at(X, pri, S, (unsigned)(t.i + hx) + (unsigned)g.sy * (unsigned)rmap(g, t.j) + kz);
                #pragma unroll
                for (int m = 0; m < NLA; m++) sp[m][t.ty + 2][t.tx + 2 + hx] = rec_field(r, m);
            }
        }
        if (t.ty < 2 || t.ty >= TY - 2) {
            int hy = t.ty < 2 ? -2 : 2;
            if (loadable(g, t.i, t.j + hy)) {
                Rec r = rec_at(X, pri, S, (unsigned)t.i + (unsigned)g.sy * (unsigned)rmap(g, t.j + hy) + kz);
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
            double x1 = fld(X, ek), x5 = fld(X + 4 * S, ek), x6 = fld(X + 5 * S, ek), x7 = fld(X + 6 * S, ek);
            double c1 = fld(cur, ek), c2 = fld(cur + S, ek), c3 = fld(cur + 2 * S, ek);
            xd1 = xd1 - fld(xresd, ek) + c2 * x7 - c3 * x6;
            xd2 = xd2 - fld(xresd + S, ek) + c3 * x5 - c1 * x7;
            xd3 = xd3 - fld(xresd + 2 * S, ek) + c1 * x6 - c2 * x5;
            st(fldp(XD, ek), xd0); st(fldp(XD + S, ek), xd1); st(fldp(XD + 2 * S, ek), xd2);
            st(fldp(XD + 3 * S, ek), xd3);
            if (hall) {
                const double y = 1.0 / x1;
                hv[0] = divc(cht * (c2 * x7 - c3 * x6 - DXL(LP)), x1, y);
                hv[1] = divc(cht * (c3 * x5 - c1 * x7 - DYL(LP)), x1, y);
                hv[2] = divc(cht * (c1 * x6 - c2 * x5 - DZQ(qp)), x1, y);
            }
#undef DZQ
#undef DXL
#undef DYL
        }
        if (hall) {
            plane_rules<3, BC_AUX>(g, t, bs, hv);
            if (t.own) {
                st(fldp(XH, ek), hv[0]); st(fldp(XH + S, ek), hv[1]); st(fldp(XH + 2 * S, ek), hv[2]);
                zface_store<3>(g, XH, (long)e0, k, hv);
            }
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
}

// ---------------------------------------------------------------------------
// right, part 2 (t3dsdsc.f:1426-1566, 1585-1663, 1665): -curl(E + xdh), the
// pressure equation, optional viscosity / diffusion, then the bndry fill of
// all eight xdif fields.  Per point e = E + xdh and v = m / rho are formed
// once.  LDS: {e1,e2,e3,p,vx,vy[,vz,rho]}; queue: {e1,e2,p,vz[,vx,vy,rho]}.
// ---------------------------------------------------------------------------
enum { LE1 = 0, LE2, LE3, LPP, LVX, LVY, LVZ, LRH };

template <bool VD, class GT = Geo>
__device__ __forceinline__ void rhs_b_body(GT g, const double *__restrict__ X, const double *__restrict__ cur,
                                           const double *__restrict__ E, const double *__restrict__ XH,
                                           double eta, double gam, double fmu, double frho, double fpr,
                                           double *__restrict__ XD, int kc, BlockPos bp)
{
    constexpr int NL = VD ? 8 : 6;
    __shared__ double sp[NL][TY + 4][TX + 4];
    __shared__ BCScratch<8> bs;
    Tile t = tile_of(g, kc, bp);
    if (t.kb >= t.ke) return;
    const long S = g.S, sz = g.sz;
    const bool ld = loadable(g, t.i, t.j);
    const long q0 = ld ? (long)t.i + g.sy * rmap(g, t.j) : 0;
    auto ept = [&](long q, int m) { return XH ? E[q + m * S] + XH[q + m * S] : E[q + m * S]; };
    double qe1[5], qe2[5], qp[5], qvz[5], qvx[5], qvy[5], qrh[5];
    auto load_q = [&](int s, long q) {
        double r = X[q];
        qe1[s] = ept(q, 0); qe2[s] = ept(q, 1); qp[s] = X[q + 7 * S];
        if (VD) {
            const double y = 1.0 / r;
            qvz[s] = divc(X[q + 3 * S], r, y);
            qvx[s] = divc(X[q + S], r, y); qvy[s] = divc(X[q + 2 * S], r, y); qrh[s] = r;
        } else {
            qvz[s] = X[q + 3 * S] / r;
        }
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
            zface_store<8>(g, XD, q0, k, v);
        }
        __syncthreads();
        #pragma unroll
        for (int s = 0; s < 4; s++) {
            qe1[s] = qe1[s + 1]; qe2[s] = qe2[s + 1]; qp[s] = qp[s + 1]; qvz[s] = qvz[s + 1];
            if (VD) { qvx[s] = qvx[s + 1]; qvy[s] = qvy[s + 1]; qrh[s] = qrh[s + 1]; }
        }
    }
}

// ---------------------------------------------------------------------------
// right, part 2 without viscosity and diffusion, in the windowed z-march form
// of the smoothing passes:
//   PART 0  xdif(5..7) = -curl(E + xdh)          (t3dsdsc.f:1435-1452)
//   PART 1  xdif(8), the pressure equation       (t3dsdsc.f:1550-1566)
// each with the bndry fill of its fields (t3dsdsc.f:1665; xdif(1..4) are
// filled by k_fill_state).  A block of 64 x ZB threads covers 64 columns x ZB
// rows and owns the inner 60 x ZB-4 (the 5-point stencils reach 2 points).
// Iteration t loads plane t, forms the x- and y-derivative parts of plane t
// from the plane in LDS, and completes plane t-2 with the z derivatives from
// the register queues.  Partial sums keep the reference's order of
// operations, so the results are those of the one-pass formulas.
// ---------------------------------------------------------------------------
constexpr int RB_KC = 32;        // interior planes per block

template <int PART, int RB>
__device__ __forceinline__ void rhs_b2_body(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                                            const double *__restrict__ E, double eta, double gam,
                                            double *__restrict__ XD, int bx, int by, int bz)
{
    constexpr int W = TX - 4, H = RB - 4;
    constexpr int NO = PART == 0 ? 3 : 1;        // output fields
    __shared__ double sp[2][3][RB][TX];         // the plane's fields, by plane parity
    __shared__ double sx[NO][2][TX];            // rows 2 / NY-3 of the outputs
    const int lane = threadIdx.x % TX;
    const int r = __builtin_amdgcn_readfirstlane(threadIdx.x / TX);
    const int nomx = bx * W;
    const int i0 = max(0, min(nomx - 2, g.NX - TX));
    const int i = i0 + lane;
    const bool vx = i < g.NX;
    const bool ownx = vx && i >= nomx && i < nomx + W;
    const int nomy = g.ylo + by * H;
    const int j0 = g.periodic ? max(g.ylo - 2, min(nomy - 2, g.yhi + 3 - RB)) : max(0, min(nomy - 2, g.NY - RB));
    const int j = j0 + r;
    const bool vy = g.periodic || j < g.NY;
    const bool own = ownx && vy && j >= nomy && j < nomy + H && j <= g.yhi;
    const bool ld = vx && vy;
    const int jp = vy ? rmap(g, j) : 0;
    const bool yin = g.periodic || (j >= 2 && j <= g.NY - 3);
    const bool pin = ld && yin && i >= 2 && i <= g.NX - 3;
    const bool img = !g.periodic && vy && (j < 2 || j > g.NY - 3);
    const bool xft = i0 < 2 || i0 + TX > g.NX - 2;
    // with periodic y a row outside the block's owned rows only loads the
    // plane into LDS for its neighbours' y derivatives: its x/y parts and its
    // output feed nothing (wave-uniform; non-periodic image rows exchange
    // through LDS with a block-wide barrier, so there every row works)
    const bool work = !g.periodic || (j >= nomy && j < min(nomy + H, g.yhi + 1));
    const int rm1 = max(r - 1, 0), rm2 = max(r - 2, 0), rp1 = min(r + 1, RB - 1), rp2 = min(r + 2, RB - 1);
    const int lm1 = max(lane - 1, 0), lm2 = max(lane - 2, 0), lp1 = min(lane + 1, TX - 1), lp2 = min(lane + 2, TX - 1);
    const long S = g.S;
    const unsigned sz = (unsigned)g.sz;
    const unsigned qc = ld ? (unsigned)i + (unsigned)g.sy * (unsigned)jp : 0u;
    const double ax0 = pin ? g.x1a[i] : 0.0, ax1 = pin ? g.x1b[i] : 0.0, ax2 = pin ? g.x1c[i] : 0.0,
                 ax3 = pin ? g.x1d[i] : 0.0;
    const double ay0 = g.y1a[jp], ay1 = g.y1b[jp], ay2 = g.y1c[jp], ay3 = g.y1d[jp];   // row-uniform
    const int kb = 2 + bz * RB_KC, ke = min(kb + RB_KC, g.NZ - 2);
    const int t0 = kb - 2, t1 = ke + 1;
#define DXS(f) fd5(sp[b][f][r][lm2], sp[b][f][r][lm1], sp[b][f][r][lane], sp[b][f][r][lp1], sp[b][f][r][lp2], ax0, ax1, ax2, ax3)
#define DYS(f) fd5(sp[b][f][rm2][lane], sp[b][f][rm1][lane], sp[b][f][r][lane], sp[b][f][rp1][lane], sp[b][f][rp2][lane], ay0, ay1, ay2, ay3)
#define DZQ(a) fd5(a[0], a[1], a[2], a[3], a[4], az0, az1, az2, az3)
    // z queues (planes t-4..t) and partial sums of planes t-1, t
    double qa[5], qb[5];
    double pa[2], pb[2], pc[2];
    #pragma unroll
    for (int s = 0; s < 5; s++) qa[s] = qb[s] = 0.0;
    pa[0] = pa[1] = pb[0] = pb[1] = pc[0] = pc[1] = 0.0;
    // raw values of the next plane
    double n0, n1, n2, n3 = 0.0, n4 = 0.0;
    auto fetch = [&](int t) {
        const unsigned e = qc + sz * (unsigned)t;
        if (PART == 0) {
            n0 = fld(E, e); n1 = fld(E + S, e); n2 = fld(E + 2 * S, e);
        } else {
            n0 = fld(X, e); n1 = fld(X + S, e); n2 = fld(X + 2 * S, e); n3 = fld(X + 3 * S, e); n4 = fld(X + 7 * S, e);
        }
    };
    fetch(t0);
    for (int t = t0; t <= t1; t++) {
        // plane t: LDS fields and queue entries of the own point
        double f0, f1, f2, za, zb;
        if (PART == 0) {
            f0 = n0; f1 = n1; f2 = n2;           // e1, e2, e3
            za = n0; zb = n1;                     // z derivatives of e1 and e2
        } else {
            const double y = 1.0 / n0;
            f0 = n4;                              // p
            f1 = divc(n1, n0, y);                 // vx
            f2 = divc(n2, n0, y);                 // vy
            za = n4; zb = divc(n3, n0, y);        // p, vz
        }
        if (!ld) { f0 = f1 = f2 = za = zb = 0.0; }
        if (t < t1) fetch(t + 1);
        const int b = t & 1;
        sp[b][0][r][lane] = f0; sp[b][1][r][lane] = f1; sp[b][2][r][lane] = f2;
        #pragma unroll
        for (int s = 0; s < 4; s++) { qa[s] = qa[s + 1]; qb[s] = qb[s + 1]; }
        qa[4] = za; qb[4] = zb;
        // J^2 of the output plane t-2 (pressure equation)
        double j2 = 0.0;
        if (PART == 1 && t - 2 >= kb && work) {
            const unsigned e = qc + sz * (unsigned)(t - 2);
            const double c1 = fld(cur, e), c2 = fld(cur + S, e), c3 = fld(cur + 2 * S, e);
            j2 = c1 * c1 + c2 * c2 + c3 * c3;
        }
        __syncthreads();
        // x/y parts of plane t
        double xa = 0.0, xb = 0.0, xc = 0.0;
        if (work && pin && t >= 2 && t <= g.NZ - 3) {
            if (PART == 0) {
                xa = -DYS(2);                     // xd5 = -d/dy e3 + ...
                xb = DXS(2);                      // xd6 = ... + d/dx e3
                xc = -DXS(1) + DYS(0);            // xd7 = -d/dx e2 + d/dy e1
            } else {
                xa = -DXS(0) * f1 - DYS(0) * f2;  // -dp/dx vx - dp/dy vy
                xb = -DXS(1) - DYS(2);            // -dvx/dx - dvy/dy
            }
        }
        // plane k = t-2: complete, fill, store
        const int k = t - 2;
        if (k >= kb && work) {
            double v[NO];
            if (pin) {
                const double az0 = g.z1a[k], az1 = g.z1b[k], az2 = g.z1c[k], az3 = g.z1d[k];
                if (PART == 0) {
                    v[0] = pa[0] + DZQ(qb);               // xd5
                    v[NO > 1 ? 1 : 0] = -DZQ(qa) + pb[0]; // xd6
                    v[NO > 2 ? 2 : 0] = pc[0];            // xd7
                } else {
                    const double hs0 = qb[2], p0 = qa[2];
                    v[0] = pa[0] - DZQ(qa) * hs0 + (pb[0] - DZQ(qb)) * gam * p0 + (gam - 1.) * eta * j2;
                }
            } else {
                #pragma unroll
                for (int o = 0; o < NO; o++) v[o] = 0.0;
            }
            if (xft) {
                #pragma unroll
                for (int o = 0; o < NO; o++) v[o] = row_xrule_rl<BC_STATE>(g, i, i0, vx && yin, v[o]);
            }
            if (!g.periodic) {
                if (j == 2 || j == g.NY - 3) {
                    #pragma unroll
                    for (int o = 0; o < NO; o++) sx[o][j == 2 ? 0 : 1][lane] = v[o];
                }
                __syncthreads();
                if (img) {
                    #pragma unroll
                    for (int o = 0; o < NO; o++)
                        v[o] = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * sx[o][j < 2 ? 0 : 1][lane];
                }
            }
            if (own) {
                const unsigned e = qc + sz * (unsigned)k;
                const int m0 = PART == 0 ? 4 : 7;
                #pragma unroll
                for (int o = 0; o < NO; o++) st(fldp(XD + (m0 + o) * S, e), v[o]);
                if (k == 2 || k == g.NZ - 3) {
                    const unsigned eh = qc + sz * (unsigned)(k == 2 ? 1 : g.NZ - 2);
                    const unsigned e0 = qc + sz * (unsigned)(k == 2 ? 0 : g.NZ - 1);
                    #pragma unroll
                    for (int o = 0; o < NO; o++) {
                        st(fldp(XD + (m0 + o) * S, eh), 0.5 * v[o]);
                        st(fldp(XD + (m0 + o) * S, e0), 0.0);
                    }
                }
            }
        }
        pa[0] = pa[1]; pb[0] = pb[1]; pc[0] = pc[1];
        pa[1] = xa; pb[1] = xb; pc[1] = xc;
    }
#undef DXS
#undef DYS
#undef DZQ
}

template <int PART, int RB>
__global__ __launch_bounds__(TX * RB) void k_rhs_b2(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                                                   const double *__restrict__ E, double eta, double gam,
                                                   double *__restrict__ XD)
{
    rhs_b2_body<PART, RB>(g, X, cur, E, eta, gam, XD, blockIdx.x, blockIdx.y, blockIdx.z);
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

// bndry fill of xdif(1..4) (t3dsdsc.f:1677-1985 with the x-left rule
// mirrored): the face points from the interior values the flux kernel left,
// in the order of the fills inside the sweeps (x rule on interior rows, the
// image-row rule from rows 2 / NY-3, then the z rule from planes 2 / NZ-3)
__device__ inline double state_face(const Geo &g, const double *G, int i, int j, int k)
{
    const int kk = k < 2 ? 2 : k > g.NZ - 3 ? g.NZ - 3 : k;
    const int jj = g.periodic ? j : (j < 2 ? 2 : j > g.NY - 3 ? g.NY - 3 : j);
    double v;
    const long r = g.sy * rmap(g, jj) + g.sz * kk;
    if (i >= 2 && i <= g.NX - 3) v = G[r + i];
    else if (i == 0 || i == g.NX - 1) v = 0.0;
    else v = 0.5 * G[r + (i == 1 ? 2 : g.NX - 3)];
    if (jj != j) v = (j == 0 || j == g.NY - 1) ? 0.0 : 0.5 * v;
    if (kk != k) v = (k == 0 || k == g.NZ - 1) ? 0.0 : 0.5 * v;
    return v;
}

// block b of the face-point fill (256 points per block)
__device__ __forceinline__ void fill_state_body(const Geo &g, double *__restrict__ XD, int nf, unsigned b)
{
    long t = (long)b * blockDim.x + threadIdx.x;
    int i, j, k;
    if (!face_point(g, t, i, j, k)) return;
    const long q = IDX(g, i, j, k);
    for (int m = 0; m < nf; m++) XD[q + m * g.S] = state_face(g, XD + m * g.S, i, j, k);
}

__global__ void k_fill_state(Geo g, double *__restrict__ XD, int nf)
{
    fill_state_body(g, XD, nf, blockIdx.x);
}


// ---------------------------------------------------------------------------
// launch forms of the tile and smoothing bodies.  One run: the block position
// comes from the grid (XCD-contiguous ranges, block_pos) or from the run's
// block table.  A batch of runs: each block reads {run, word} from a merged
// block table and the run's arguments from a device table, so one launch
// covers the same sweep of every run.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6, 8))) void k_curl(Geo g, const double *__restrict__ X, const double *__restrict__ bxi,
                                             const double *__restrict__ byi, const double *__restrict__ bzi,
                                             double *__restrict__ cur, int kc)
{
    curl_body(g, X, bxi, byi, bzi, cur, kc, block_pos());
}

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8))) void k_rhs_a(
    Geo g, const double *__restrict__ X, const double *__restrict__ pri, const double *__restrict__ cur,
    const double *__restrict__ xresd, double eta, int hall, double cht, double *__restrict__ XD,
    double *__restrict__ E, double *__restrict__ XH, int kc)
{
    rhs_a_body<Geo>(g, X, pri, cur, xresd, eta, hall, cht, XD, E, XH, kc, block_pos());
}

template <bool VD>
__global__ __launch_bounds__(NT) void k_rhs_b(Geo g, const double *__restrict__ X, const double *__restrict__ cur,
                                              const double *__restrict__ E, const double *__restrict__ XH,
                                              double eta, double gam, double fmu, double frho, double fpr,
                                              double *__restrict__ XD, int kc)
{
    rhs_b_body<VD>(g, X, cur, E, XH, eta, gam, fmu, frho, fpr, XD, kc, block_pos());
}

// table word of a smoothing block: x tile | y tile << ZT_BY | chunk << ZT_CH | field << ZT_F
__device__ inline void zs_decode(unsigned zw, int &bx, int &by, int &chunk, int &m)
{
    bx = __builtin_amdgcn_readfirstlane(zw & ((1u << ZT_BY) - 1));
    by = __builtin_amdgcn_readfirstlane((zw >> ZT_BY) & ((1u << (ZT_CH - ZT_BY)) - 1));
    chunk = __builtin_amdgcn_readfirstlane((zw >> ZT_CH) & ((1u << (ZT_F - ZT_CH)) - 1));
    m = __builtin_amdgcn_readfirstlane(zw >> ZT_F);
}

template <int NP, int KIND, int TYPE, int RKS, int MODE, int RB = ZB>
__global__ __launch_bounds__(TX * RB) __attribute__((amdgpu_waves_per_eu(MODE == ZM_FACE ? 4 : 8, 8))) void k_zsmooth(
    Geo g, const double *__restrict__ A, double *__restrict__ B, RKArgs rk, const unsigned *__restrict__ ztab, int nch,
    int kc)
{
    // the block's tile, chunk and field: from the grid position (ZM_ALL), or
    // from the host's block table by dispatch position (work-groups go to
    // XCD id % 8)
    int bx, by, chunk, m;
    if (MODE == ZM_ALL) {
        // the grid order (x tile fastest, then y tile, chunk, field) handed to
        // the XCDs in contiguous ranges: neighbouring tiles of a field share
        // halo rows in one L2
        const unsigned gx = gridDim.x, gy = gridDim.y;
        const unsigned T = gx * gy * gridDim.z;
        const unsigned L = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
        const unsigned per = T / NXCD, rem = T % NXCD, xcd = L % NXCD;
        unsigned V = xcd * per + (xcd < rem ? xcd : rem) + L / NXCD;
        bx = __builtin_amdgcn_readfirstlane(V % gx);
        V /= gx;
        by = __builtin_amdgcn_readfirstlane(V % gy);
        V /= gy;
        chunk = __builtin_amdgcn_readfirstlane(V % (unsigned)nch);
        m = __builtin_amdgcn_readfirstlane(V / (unsigned)nch);
    } else {
        zs_decode(ztab[blockIdx.x], bx, by, chunk, m);
    }
    zsmooth_body<NP, KIND, TYPE, RKS, MODE, RB>(g, A, B, rk, bx, by, chunk, m, kc);
}

// ---- batched forms ----
struct BEntry {
    unsigned run;      // index into the launch's argument table
    unsigned w;        // smoothing: table word; tile kernels: bx | by << ZT_BY | bz << ZT_CH
};

struct BZArgs {
    Geo g;
    const double *A;
    double *B;
    RKArgs rk;
    int kc;
};

struct BCurlArgs {
    Geo g;
    const double *X, *bxi, *byi, *bzi;
    double *cur;
    int kc;
};

struct BRhsAArgs {
    Geo g;
    const double *X, *pri, *cur, *xresd;
    double eta, cht;
    int hall, kc;
    double *XD, *E, *XH;
};

struct BRhsBArgs {
    Geo g;
    const double *X, *cur, *E, *XH;
    double eta, gam, fmu, frho, fpr;
    double *XD;
    int kc;
};

// a run's argument block, read through the constant address space: the
// blocks do not change during a launch, so the compiler may load their
// fields again where needed instead of holding them in registers (as it
// does for kernel arguments)
template <class A>
__device__ inline const A &run_args(const A *args, unsigned run)
{
    return *(const A *)((const __attribute__((address_space(4))) A *)args + run);
}

__device__ inline BlockPos tile_decode(unsigned w)
{
    BlockPos b;
    b.bx = __builtin_amdgcn_readfirstlane(w & ((1u << ZT_BY) - 1));
    b.by = __builtin_amdgcn_readfirstlane((w >> ZT_BY) & ((1u << (ZT_CH - ZT_BY)) - 1));
    b.bz = __builtin_amdgcn_readfirstlane(w >> ZT_CH);
    return b;
}

template <int NP, int KIND, int TYPE, int RKS, int MODE, int RB = ZB>
__global__ __launch_bounds__(TX * RB) __attribute__((amdgpu_waves_per_eu(MODE == ZM_FACE ? 4 : 8, 8))) void kb_zsmooth(
    const BZArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    int bx, by, chunk, m;
    zs_decode(tab[blockIdx.x].w, bx, by, chunk, m);
    const BZArgs &a = run_args(args, run);
    zsmooth_body<NP, KIND, TYPE, RKS, MODE, RB>(a.g, a.A, a.B, a.rk, bx, by, chunk, m, a.kc);
}

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6, 8))) void kb_curl(const BCurlArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BCurlArgs &a = run_args(args, run);
    curl_body(a.g, a.X, a.bxi, a.byi, a.bzi, a.cur, a.kc, tile_decode(tab[blockIdx.x].w));
}

__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4, 8))) void kb_rhs_a(
    const BRhsAArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BRhsAArgs &a = run_args(args, run);
    rhs_a_body<const Geo &>(a.g, a.X, a.pri, a.cur, a.xresd, a.eta, a.hall, a.cht, a.XD, a.E, a.XH, a.kc,
               tile_decode(tab[blockIdx.x].w));
}

__global__ __launch_bounds__(NT) void kb_rhs_b(const BRhsBArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BRhsBArgs &a = run_args(args, run);
    rhs_b_body<false, const Geo &>(a.g, a.X, a.cur, a.E, a.XH, a.eta, a.gam, a.fmu, a.frho, a.fpr, a.XD, a.kc,
                      tile_decode(tab[blockIdx.x].w));
}

struct BRhsB2Args {
    Geo g;
    const double *X, *cur, *E;
    double eta, gam;
    double *XD;
};

template <int PART, int RB>
__global__ __launch_bounds__(TX * RB) void kb_rhs_b2(const BRhsB2Args *__restrict__ args,
                                                    const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BRhsB2Args &a = run_args(args, run);
    const BlockPos b = tile_decode(tab[blockIdx.x].w);
    rhs_b2_body<PART, RB>(a.g, a.X, a.cur, a.E, a.eta, a.gam, a.XD, b.bx, b.by, b.bz);
}

struct BFillArgs {
    Geo g;
    double *XD;
    int nf;
};

__global__ void kb_fill_state(const BFillArgs *__restrict__ args, const BEntry *__restrict__ tab)
{
    const unsigned run = __builtin_amdgcn_readfirstlane(tab[blockIdx.x].run);
    const BFillArgs &a = run_args(args, run);
    fill_state_body(a.g, a.XD, a.nf, tab[blockIdx.x].w);
}

} // namespace mhd
