/*
 * mhd_oracle_adj.c -- CPU ORACLE of the discrete adjoint (reverse mode) of
 * right() and stepon() in mhd_oracle.c, which includes this file.
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP adjoint.  It follows the
 * forward restatement routine by routine, in reverse order:
 *   - linear maps (smoothing sweeps, boundary rules, 5-point derivatives,
 *     periodic images, RK register combinations) are transposed exactly;
 *   - the pointwise nonlinear terms (fluxes, E = -v x B + eta J, the Hall
 *     term, the pressure equation) contribute the transpose of their
 *     Jacobian at the forward state.
 * The time-step size is an input (the dt sequence is frozen), so setdt is
 * not differentiated.  Parity pins (tests/test_adjoint.py):
 * <A u, w> = <u, A^T w> for every linear primitive against the forward
 * primitive, and central differences of right(), stepon() and multi-step
 * objectives.
 */

static int adj_interior(const orc *o, int jx, int jy, int jz)
{
    return jx >= o->x_f && jx <= o->x_l && jy >= 3 && jy <= o->ny - 2 && jz >= o->z_f && jz <= o->z_l;
}

/* transpose of periodic_y on nf fields */
static void adj_periodic_y(orc *o, double *a, int nf)
{
    int m, jx, jz, my = o->ny;
    for (m = 1; m <= nf; m++)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                FM(a, jx, 4, jz, m) += FM(a, jx, my, jz, m); FM(a, jx, my, jz, m) = 0;
                FM(a, jx, 3, jz, m) += FM(a, jx, my - 1, jz, m); FM(a, jx, my - 1, jz, m) = 0;
                FM(a, jx, my - 2, jz, m) += FM(a, jx, 2, jz, m); FM(a, jx, 2, jz, m) = 0;
                FM(a, jx, my - 3, jz, m) += FM(a, jx, 1, jz, m); FM(a, jx, 1, jz, m) = 0;
            }
}

/* transpose of the z rules (f3 = 0, f4 = f5/2; f(zl) = 0, f(zl-1) = f(zl-2)/2)
 * over columns jx0..x_last+2 */
static void adj_zrules(orc *o, double *a, int nf, int jx0)
{
    int m, jx, jy, zl = o->zl;
    for (m = 1; m <= nf; m++)
        for (jy = 1; jy <= o->ny; jy++)
            for (jx = jx0; jx <= o->x_l + 2; jx++) {
                FM(a, jx, jy, zl - 2, m) += 0.5 * FM(a, jx, jy, zl - 1, m); FM(a, jx, jy, zl - 1, m) = 0;
                FM(a, jx, jy, zl, m) = 0;
                FM(a, jx, jy, 5, m) += 0.5 * FM(a, jx, jy, 4, m); FM(a, jx, jy, 4, m) = 0;
                FM(a, jx, jy, 3, m) = 0;
            }
}

/* transpose of the x rules: extrap = 1 the one-sided extrapolation on the
 * left (bndryc/h/E, bndrytt), 0 the mirrored 0 / half rule (bndry); the
 * right side is 0 / half in both */
static void adj_xrules(orc *o, double *a, int nf, int extrap)
{
    int m, jy, jz, xl = o->xl;
    double s3 = V1(o->axp, 3) + V1(o->bxp, 3) + V1(o->cxp, 3);
    double s4 = V1(o->axp, 4) + V1(o->bxp, 4) + V1(o->cxp, 4);
    for (m = 1; m <= nf; m++)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jy = 1; jy <= o->ny; jy++) {
                FM(a, xl - 2, jy, jz, m) += 0.5 * FM(a, xl - 1, jy, jz, m); FM(a, xl - 1, jy, jz, m) = 0;
                FM(a, xl, jy, jz, m) = 0;
                if (extrap) {
                    double a3 = FM(a, 3, jy, jz, m), a4;
                    FM(a, 4, jy, jz, m) += V1(o->axp, 3) / s3 * a3;
                    FM(a, 5, jy, jz, m) += V1(o->bxp, 3) / s3 * a3;
                    FM(a, 6, jy, jz, m) += V1(o->cxp, 3) / s3 * a3;
                    FM(a, 3, jy, jz, m) = 0;
                    a4 = FM(a, 4, jy, jz, m);
                    FM(a, 5, jy, jz, m) += V1(o->axp, 4) / s4 * a4;
                    FM(a, 6, jy, jz, m) += V1(o->bxp, 4) / s4 * a4;
                    FM(a, 7, jy, jz, m) += V1(o->cxp, 4) / s4 * a4;
                    FM(a, 4, jy, jz, m) = 0;
                } else {
                    FM(a, 5, jy, jz, m) += 0.5 * FM(a, 4, jy, jz, m); FM(a, 4, jy, jz, m) = 0;
                    FM(a, 3, jy, jz, m) = 0;
                }
            }
}

/* transpose of the non-periodic y rules (f1 = 0, f2 = f3/2; f(my) = 0,
 * f(my-1) = f(my-2)/2) over columns jx0..x_last+2 (bndryc/h/E, bndry:
 * t3dsdsc.f:1800-1819, 2592-2612; bndrytt from x_first, 2108-2128) */
static void adj_yrules(orc *o, double *a, int nf, int jx0)
{
    int m, jx, jz, my = o->ny;
    for (m = 1; m <= nf; m++)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jx = jx0; jx <= o->x_l + 2; jx++) {
                FM(a, jx, my - 2, jz, m) += 0.5 * FM(a, jx, my - 1, jz, m); FM(a, jx, my - 1, jz, m) = 0;
                FM(a, jx, my, jz, m) = 0;
                FM(a, jx, 3, jz, m) += 0.5 * FM(a, jx, 2, jz, m); FM(a, jx, 2, jz, m) = 0;
                FM(a, jx, 1, jz, m) = 0;
            }
}

static void adj_y(orc *o, double *a, int nf, int jx0)
{
    if (o->c.periodic_y) adj_periodic_y(o, a, nf);
    else adj_yrules(o, a, nf, jx0);
}

/* transposes of bndry_aux, bndry and bndrytt (x, then y, then z forward) */
static void adj_bndry_aux(orc *o, double *a)
{
    if (!o->c.bcs) return;
    adj_zrules(o, a, 3, o->x_f - 2);
    adj_y(o, a, 3, o->x_f - 2);
    adj_xrules(o, a, 3, 1);
}

static void adj_bndry(orc *o, double *a)
{
    if (!o->c.bcs) return;
    adj_zrules(o, a, 8, o->x_f - 2);
    adj_y(o, a, 8, o->x_f - 2);
    adj_xrules(o, a, 8, 0);
}

static void adj_bndrytt(orc *o, double *a)
{
    if (!o->c.bcs) return;
    adj_zrules(o, a, 8, o->x_f);
    adj_y(o, a, 8, o->x_f);
    adj_xrules(o, a, 8, 1);
}

/* difc weights of f(p+1) and f(p-1) at the interior point p of axis c */
static inline double adj_wp(const double *c, int p)
{
    return 2. * (V1(c, p + 1) - V1(c, p)) / (V1(c, p + 1) - V1(c, p - 1));
}
static inline double adj_wm(const double *c, int p)
{
    return 2. * (V1(c, p) - V1(c, p - 1)) / (V1(c, p + 1) - V1(c, p - 1));
}

/* transpose of filter_pass on field m of a (w: one-field scratch):
 * f_new = f + c * sum(difc) on the interior, f elsewhere */
static void adj_filter_pass(orc *o, double *a, int m, int kind, double *w)
{
    int jx, jy, jz;
    memset(w, 0, sizeof(double) * o->n3);
#pragma omp parallel for private(jy, jx)
    for (jz = o->z_f; jz <= o->z_l; jz++)
        for (jy = 3; jy <= o->ny - 2; jy++)
            for (jx = o->x_f; jx <= o->x_l; jx++)
                F(w, jx, jy, jz) = filt_coef(o, kind, jx, jy, jz) * FM(a, jx, jy, jz, m);
#pragma omp parallel for private(jy, jx)
    for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
        for (jy = 1; jy <= o->ny; jy++)
            for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                double s = 0;
                if (adj_interior(o, jx, jy, jz))
                    s -= (adj_wp(o->xx, jx) + adj_wm(o->xx, jx) + adj_wp(o->yy, jy) + adj_wm(o->yy, jy) +
                          adj_wp(o->zz, jz) + adj_wm(o->zz, jz)) * F(w, jx, jy, jz);
                if (adj_interior(o, jx - 1, jy, jz)) s += adj_wp(o->xx, jx - 1) * F(w, jx - 1, jy, jz);
                if (adj_interior(o, jx + 1, jy, jz)) s += adj_wm(o->xx, jx + 1) * F(w, jx + 1, jy, jz);
                if (adj_interior(o, jx, jy - 1, jz)) s += adj_wp(o->yy, jy - 1) * F(w, jx, jy - 1, jz);
                if (adj_interior(o, jx, jy + 1, jz)) s += adj_wm(o->yy, jy + 1) * F(w, jx, jy + 1, jz);
                if (adj_interior(o, jx, jy, jz - 1)) s += adj_wp(o->zz, jz - 1) * F(w, jx, jy, jz - 1);
                if (adj_interior(o, jx, jy, jz + 1)) s += adj_wm(o->zz, jz + 1) * F(w, jx, jy, jz + 1);
                FM(a, jx, jy, jz, m) += s;
            }
}

/* transposes of avrg_aux (kk passes, then the fill) and avrg_state */
static void adj_avrg_aux(orc *o, double *a, int kind, int kk, double *w)
{
    int k, m;
    if (!o->c.filters) return;
    for (k = kk; k >= 1; k--) {
        adj_bndry_aux(o, a);
        for (m = 1; m <= 3; m++) adj_filter_pass(o, a, m, kind, w);
    }
}

static void adj_avrg_state(orc *o, double *a, int kk, int post, double *w)
{
    int k, m;
    if (!o->c.filters) return;
    for (k = kk; k >= 1; k--) {
        if (post) adj_bndrytt(o, a); else adj_bndry(o, a);
        for (m = 1; m <= 8; m++) adj_filter_pass(o, a, m, (m >= 5 && m <= 7) ? COEF_BX : COEF_RADIAL, w);
    }
}

/* weight of the input at offset d = q - p in d1fc at output point p */
static inline double adj_dw(int d, double a, double b, double c, double dd)
{
    return d == 1 ? a : d == -1 ? -b : d == 2 ? c : d == -2 ? -dd : (-a + b - c + dd);
}

/* (D^T g)(q) along one axis for g supported on the interior */
static double adj_dx(const orc *o, const double *g, int jx, int jy, int jz)
{
    double s = 0;
    int p;
    for (p = jx - 2; p <= jx + 2; p++)
        if (p >= o->x_f && p <= o->x_l)
            s += adj_dw(jx - p, V1(o->ax1, p), V1(o->bx1, p), V1(o->cx1, p), V1(o->dx1, p)) * F(g, p, jy, jz);
    return s;
}
static double adj_dy(const orc *o, const double *g, int jx, int jy, int jz)
{
    double s = 0;
    int p;
    for (p = jy - 2; p <= jy + 2; p++)
        if (p >= 3 && p <= o->ny - 2)
            s += adj_dw(jy - p, V1(o->ay1, p), V1(o->by1, p), V1(o->cy1, p), V1(o->dy1, p)) * F(g, jx, p, jz);
    return s;
}
static double adj_dz(const orc *o, const double *g, int jx, int jy, int jz)
{
    double s = 0;
    int p;
    for (p = jz - 2; p <= jz + 2; p++)
        if (p >= o->z_f && p <= o->z_l)
            s += adj_dw(jz - p, V1(o->az1, p), V1(o->bz1, p), V1(o->cz1, p), V1(o->dz1, p)) * F(g, jx, jy, p);
    return s;
}

/* (L^T g)(q) of the 5-point Laplacian D2X + D2Y + D2Z (second-derivative
 * coefficients) for g supported on the interior */
static double adj_lap(const orc *o, const double *g, int jx, int jy, int jz)
{
    double s = 0;
    int p;
    for (p = jx - 2; p <= jx + 2; p++)
        if (p >= o->x_f && p <= o->x_l)
            s += adj_dw(jx - p, V1(o->ax2, p), V1(o->bx2, p), V1(o->cx2, p), V1(o->dx2, p)) * F(g, p, jy, jz);
    for (p = jy - 2; p <= jy + 2; p++)
        if (p >= 3 && p <= o->ny - 2)
            s += adj_dw(jy - p, V1(o->ay2, p), V1(o->by2, p), V1(o->cy2, p), V1(o->dy2, p)) * F(g, jx, p, jz);
    for (p = jz - 2; p <= jz + 2; p++)
        if (p >= o->z_f && p <= o->z_l)
            s += adj_dw(jz - p, V1(o->az2, p), V1(o->bz2, p), V1(o->cz2, p), V1(o->dz2, p)) * F(g, jx, jy, p);
    return s;
}

#define ADJ_INT_LOOP for (jz = o->z_f; jz <= o->z_l; jz++) for (jy = 3; jy <= o->ny - 2; jy++) \
                     for (jx = o->x_f; jx <= o->x_l; jx++)
#define ADJ_EXT_LOOP for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++) for (jy = 1; jy <= o->ny; jy++) \
                     for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++)

/* Reverse mode of right(): X the state, Cu = current(X) (filtered), ab the
 * adjoint of xdif (destroyed), xb += (dright/dX)^T ab, rb -= the part on
 * xresd (when rb), *etab += the part on eta.  init: the residual capture
 * (eta = 0, no final avrg); smooth: the final avrg (t != 0). */
static void adj_right(orc *o, const double *X, const double *Cu, double *ab, double *xb, double *rb,
                      double *etab, int init, int smooth)
{
    size_t n3 = o->n3;
    int jx, jy, jz, m;
    double eta = init ? 0.0 : o->etan, gam = o->c.gamma, cht = o->c.di, es = 0.0;
    double *eb = zalloc(3 * n3), *hb = zalloc(3 * n3), *jb = zalloc(3 * n3), *vb = zalloc(3 * n3);
    double *g = zalloc(3 * n3), *w = zalloc(n3), *vel = zalloc(3 * n3);
    double *g1 = g, *g2 = g + n3, *g3 = g + 2 * n3;
    double *fs = vel, *gs = vel + n3, *hs = vel + 2 * n3;
    const double *a1 = ab, *a2 = ab + n3, *a3 = ab + 2 * n3, *a4 = ab + 3 * n3;
    const double *a5 = ab + 4 * n3, *a6 = ab + 5 * n3, *a7 = ab + 6 * n3, *a8 = ab + 7 * n3;
    const double *x8 = X + 7 * n3;
    const double fmu = init ? 0.0 : o->fmu, frho = init ? 0.0 : o->frho, fpr = init ? 0.0 : o->fpr;
    /* final avrg (1669) and bndry (1665) */
    if (smooth) adj_avrg_state(o, ab, 2, 0, w);
    adj_bndry(o, ab);

    /* velocities of the extended range (1348-1351) */
    ADJ_EXT_LOOP {
        double x1 = FM(X, jx, jy, jz, 1);
        F(fs, jx, jy, jz) = FM(X, jx, jy, jz, 2) / x1;
        F(gs, jx, jy, jz) = FM(X, jx, jy, jz, 3) / x1;
        F(hs, jx, jy, jz) = FM(X, jx, jy, jz, 4) / x1;
    }

    /* -curl E (1435-1452): xdif5 = -Dy E3 + Dz E2, xdif6 = -Dz E1 + Dx E3,
     * xdif7 = -Dx E2 + Dy E1 */
#pragma omp parallel for private(jy, jx)
    for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
        for (jy = 1; jy <= o->ny; jy++)
            for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                FM(eb, jx, jy, jz, 1) = -adj_dz(o, a6, jx, jy, jz) + adj_dy(o, a7, jx, jy, jz);
                FM(eb, jx, jy, jz, 2) = adj_dz(o, a5, jx, jy, jz) - adj_dx(o, a7, jx, jy, jz);
                FM(eb, jx, jy, jz, 3) = -adj_dy(o, a5, jx, jy, jz) + adj_dx(o, a6, jx, jy, jz);
            }

    /* pressure equation (1550-1566):
     * xdif8 = -Dp.v + (-div v) gam p + (gam-1) eta |J|^2 */
    ADJ_INT_LOOP {
        double a = F(a8, jx, jy, jz);
        F(g1, jx, jy, jz) = a * F(fs, jx, jy, jz);
        F(g2, jx, jy, jz) = a * F(gs, jx, jy, jz);
        F(g3, jx, jy, jz) = a * F(hs, jx, jy, jz);
    }
#pragma omp parallel for private(jy, jx)
    for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
        for (jy = 1; jy <= o->ny; jy++)
            for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++)
                FM(xb, jx, jy, jz, 8) += -adj_dx(o, g1, jx, jy, jz) - adj_dy(o, g2, jx, jy, jz) -
                                         adj_dz(o, g3, jx, jy, jz);
    memset(g, 0, sizeof(double) * 3 * n3);
    ADJ_INT_LOOP F(g1, jx, jy, jz) = F(a8, jx, jy, jz) * gam * F(x8, jx, jy, jz);
#pragma omp parallel for private(jy, jx)
    for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
        for (jy = 1; jy <= o->ny; jy++)
            for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                FM(vb, jx, jy, jz, 1) = -adj_dx(o, g1, jx, jy, jz);
                FM(vb, jx, jy, jz, 2) = -adj_dy(o, g1, jx, jy, jz);
                FM(vb, jx, jy, jz, 3) = -adj_dz(o, g1, jx, jy, jz);
            }
    ADJ_INT_LOOP {
        double a = F(a8, jx, jy, jz);
        double c1 = FM(Cu, jx, jy, jz, 1), c2 = FM(Cu, jx, jy, jz, 2), c3 = FM(Cu, jx, jy, jz, 3);
        FM(vb, jx, jy, jz, 1) -= a * DXM(X, 8);
        FM(vb, jx, jy, jz, 2) -= a * DYM(X, 8);
        FM(vb, jx, jy, jz, 3) -= a * DZM(X, 8);
        FM(xb, jx, jy, jz, 8) += a * gam * (-DX(fs) - DY(gs) - DZ(hs));
        FM(jb, jx, jy, jz, 1) += a * (gam - 1.) * eta * 2. * c1;
        FM(jb, jx, jy, jz, 2) += a * (gam - 1.) * eta * 2. * c2;
        FM(jb, jx, jy, jz, 3) += a * (gam - 1.) * eta * 2. * c3;
        es += a * (gam - 1.) * (c1 * c1 + c2 * c2 + c3 * c3);
    }

    /* viscosity (1585-1621): xdif(2..4) += fmu L(v), xdif8 += fmu v.L(v);
     * diffusion (1623-1663): xdif1 += frho L(rho), xdif8 += fpr L(p) */
    if (fmu != 0.0) {
        const double *vv[3] = {fs, gs, hs}, *aa[3] = {a2, a3, a4};
        for (m = 0; m < 3; m++) {
            memset(g, 0, sizeof(double) * n3);
            ADJ_INT_LOOP F(g1, jx, jy, jz) = fmu * (F(aa[m], jx, jy, jz) + F(a8, jx, jy, jz) * F(vv[m], jx, jy, jz));
#pragma omp parallel for private(jy, jx)
            for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
                for (jy = 1; jy <= o->ny; jy++)
                    for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                        double t = adj_lap(o, g1, jx, jy, jz);
                        if (adj_interior(o, jx, jy, jz)) {
                            const double *f = vv[m];
                            t += fmu * F(a8, jx, jy, jz) * (D2X(f) + D2Y(f) + D2Z(f));
                        }
                        FM(vb, jx, jy, jz, m + 1) += t;
                    }
        }
    }
    if (frho != 0.0 || fpr != 0.0) {
        memset(g, 0, sizeof(double) * 2 * n3);
        ADJ_INT_LOOP {
            F(g1, jx, jy, jz) = frho * F(a1, jx, jy, jz);
            F(g2, jx, jy, jz) = fpr * F(a8, jx, jy, jz);
        }
#pragma omp parallel for private(jy, jx)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jy = 1; jy <= o->ny; jy++)
                for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                    FM(xb, jx, jy, jz, 1) += adj_lap(o, g1, jx, jy, jz);
                    FM(xb, jx, jy, jz, 8) += adj_lap(o, g2, jx, jy, jz);
                }
    }

    /* Hall term (1391-1433): E += avrg(bndryh(xdh)) with
     * xdh = cht (J x B - grad p) / rho on the interior */
    if (o->c.hall) {
        memcpy(hb, eb, sizeof(double) * 3 * n3);
        adj_avrg_aux(o, hb, COEF_EIGHTH, 5, w);
        adj_bndry_aux(o, hb);
        memset(g, 0, sizeof(double) * 3 * n3);
        ADJ_INT_LOOP {
            double x1 = FM(X, jx, jy, jz, 1), x5 = FM(X, jx, jy, jz, 5), x6 = FM(X, jx, jy, jz, 6);
            double x7 = FM(X, jx, jy, jz, 7);
            double c1 = FM(Cu, jx, jy, jz, 1), c2 = FM(Cu, jx, jy, jz, 2), c3 = FM(Cu, jx, jy, jz, 3);
            double b1 = FM(hb, jx, jy, jz, 1), b2 = FM(hb, jx, jy, jz, 2), b3 = FM(hb, jx, jy, jz, 3);
            double h1 = cht * (c2 * x7 - c3 * x6 - DXM(X, 8)) / x1;
            double h2 = cht * (c3 * x5 - c1 * x7 - DYM(X, 8)) / x1;
            double h3 = cht * (c1 * x6 - c2 * x5 - DZM(X, 8)) / x1;
            double s1 = cht * b1 / x1, s2 = cht * b2 / x1, s3 = cht * b3 / x1;
            FM(jb, jx, jy, jz, 1) += -s2 * x7 + s3 * x6;
            FM(jb, jx, jy, jz, 2) += s1 * x7 - s3 * x5;
            FM(jb, jx, jy, jz, 3) += -s1 * x6 + s2 * x5;
            FM(xb, jx, jy, jz, 5) += s2 * c3 - s3 * c2;
            FM(xb, jx, jy, jz, 6) += -s1 * c3 + s3 * c1;
            FM(xb, jx, jy, jz, 7) += s1 * c2 - s2 * c1;
            FM(xb, jx, jy, jz, 1) -= (b1 * h1 + b2 * h2 + b3 * h3) / x1;
            F(g1, jx, jy, jz) = s1; F(g2, jx, jy, jz) = s2; F(g3, jx, jy, jz) = s3;
        }
#pragma omp parallel for private(jy, jx)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jy = 1; jy <= o->ny; jy++)
                for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++)
                    FM(xb, jx, jy, jz, 8) += -adj_dx(o, g1, jx, jy, jz) - adj_dy(o, g2, jx, jy, jz) -
                                             adj_dz(o, g3, jx, jy, jz);
    }

    /* avrgE (1389) */
    adj_avrg_aux(o, eb, COEF_RADIAL, 4, w);

    /* extended-range block (1348-1387): E = (-m x B)/rho + eta J,
     * xdif(2..4) += -xresd + J x B, v = m / rho */
    ADJ_EXT_LOOP {
        double x1 = FM(X, jx, jy, jz, 1), x2 = FM(X, jx, jy, jz, 2), x3 = FM(X, jx, jy, jz, 3);
        double x4 = FM(X, jx, jy, jz, 4), x5 = FM(X, jx, jy, jz, 5), x6 = FM(X, jx, jy, jz, 6);
        double x7 = FM(X, jx, jy, jz, 7);
        double c1 = FM(Cu, jx, jy, jz, 1), c2 = FM(Cu, jx, jy, jz, 2), c3 = FM(Cu, jx, jy, jz, 3);
        double e1 = FM(eb, jx, jy, jz, 1), e2 = FM(eb, jx, jy, jz, 2), e3 = FM(eb, jx, jy, jz, 3);
        double u1 = -x3 * x7 + x4 * x6, u2 = -x4 * x5 + x2 * x7, u3 = -x2 * x6 + x3 * x5;
        double p2 = F(a2, jx, jy, jz), p3 = F(a3, jx, jy, jz), p4 = F(a4, jx, jy, jz);
        double v1 = FM(vb, jx, jy, jz, 1), v2 = FM(vb, jx, jy, jz, 2), v3 = FM(vb, jx, jy, jz, 3);
        FM(xb, jx, jy, jz, 1) += -(e1 * u1 + e2 * u2 + e3 * u3) / (x1 * x1) -
                                 (v1 * x2 + v2 * x3 + v3 * x4) / (x1 * x1);
        FM(xb, jx, jy, jz, 2) += (e2 * x7 - e3 * x6) / x1 + v1 / x1;
        FM(xb, jx, jy, jz, 3) += (-e1 * x7 + e3 * x5) / x1 + v2 / x1;
        FM(xb, jx, jy, jz, 4) += (e1 * x6 - e2 * x5) / x1 + v3 / x1;
        FM(xb, jx, jy, jz, 5) += (-e2 * x4 + e3 * x3) / x1 + p3 * c3 - p4 * c2;
        FM(xb, jx, jy, jz, 6) += (e1 * x4 - e3 * x2) / x1 - p2 * c3 + p4 * c1;
        FM(xb, jx, jy, jz, 7) += (-e1 * x3 + e2 * x2) / x1 + p2 * c2 - p3 * c1;
        FM(jb, jx, jy, jz, 1) += eta * e1 - p3 * x7 + p4 * x6;
        FM(jb, jx, jy, jz, 2) += eta * e2 + p2 * x7 - p4 * x5;
        FM(jb, jx, jy, jz, 3) += eta * e3 - p2 * x6 + p3 * x5;
        es += c1 * e1 + c2 * e2 + c3 * e3;
        if (rb) {
            FM(rb, jx, jy, jz, 1) -= p2;
            FM(rb, jx, jy, jz, 2) -= p3;
            FM(rb, jx, jy, jz, 3) -= p4;
        }
    }

    /* mass and momentum fluxes (1334-1346 with flux 2304-2347):
     * xdif(m) = -Dx f - Dy g - Dz h */
    for (m = 1; m <= 4; m++) {
        const double *am = m == 1 ? a1 : m == 2 ? a2 : m == 3 ? a3 : a4;
#pragma omp parallel for private(jy, jx)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jy = 1; jy <= o->ny; jy++)
                for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                    double fb = -adj_dx(o, am, jx, jy, jz), gb = -adj_dy(o, am, jx, jy, jz);
                    double hbv = -adj_dz(o, am, jx, jy, jz);
                    double x1 = FM(X, jx, jy, jz, 1), x2 = FM(X, jx, jy, jz, 2), x3 = FM(X, jx, jy, jz, 3);
                    double x4 = FM(X, jx, jy, jz, 4);
                    if (m == 1) {
                        FM(xb, jx, jy, jz, 2) += fb; FM(xb, jx, jy, jz, 3) += gb; FM(xb, jx, jy, jz, 4) += hbv;
                    } else if (m == 2) {
                        FM(xb, jx, jy, jz, 2) += (fb * 2. * x2 + gb * x3 + hbv * x4) / x1;
                        FM(xb, jx, jy, jz, 3) += gb * x2 / x1;
                        FM(xb, jx, jy, jz, 4) += hbv * x2 / x1;
                        FM(xb, jx, jy, jz, 8) += fb;
                        FM(xb, jx, jy, jz, 1) -= (fb * x2 * x2 + gb * x2 * x3 + hbv * x2 * x4) / (x1 * x1);
                    } else if (m == 3) {
                        FM(xb, jx, jy, jz, 2) += fb * x3 / x1;
                        FM(xb, jx, jy, jz, 3) += (fb * x2 + gb * 2. * x3 + hbv * x4) / x1;
                        FM(xb, jx, jy, jz, 4) += hbv * x3 / x1;
                        FM(xb, jx, jy, jz, 8) += gb;
                        FM(xb, jx, jy, jz, 1) -= (fb * x2 * x3 + gb * x3 * x3 + hbv * x4 * x3) / (x1 * x1);
                    } else {
                        FM(xb, jx, jy, jz, 2) += fb * x4 / x1;
                        FM(xb, jx, jy, jz, 3) += gb * x4 / x1;
                        FM(xb, jx, jy, jz, 4) += (fb * x2 + gb * x3 + hbv * 2. * x4) / x1;
                        FM(xb, jx, jy, jz, 8) += hbv;
                        FM(xb, jx, jy, jz, 1) -= (fb * x2 * x4 + gb * x3 * x4 + hbv * x4 * x4) / (x1 * x1);
                    }
                }
    }

    /* current (2452-2500): J = avrg(bndryc(curl(B - B_i))) */
    adj_avrg_aux(o, jb, COEF_RADIAL, 4, w);
    adj_bndry_aux(o, jb);
    {
        const double *j1 = jb, *j2 = jb + n3, *j3 = jb + 2 * n3;
#pragma omp parallel for private(jy, jx)
        for (jz = o->z_f - 2; jz <= o->z_l + 2; jz++)
            for (jy = 1; jy <= o->ny; jy++)
                for (jx = o->x_f - 2; jx <= o->x_l + 2; jx++) {
                    FM(xb, jx, jy, jz, 5) += adj_dz(o, j2, jx, jy, jz) - adj_dy(o, j3, jx, jy, jz);
                    FM(xb, jx, jy, jz, 6) += adj_dx(o, j3, jx, jy, jz) - adj_dz(o, j1, jx, jy, jz);
                    FM(xb, jx, jy, jz, 7) += adj_dy(o, j1, jx, jy, jz) - adj_dx(o, j2, jx, jy, jz);
                }
    }
    if (etab) *etab += es;
    free(eb); free(hb); free(jb); free(vb); free(g); free(w); free(vel);
}

/* ------------------------------------------------------------------ */
/* public adjoint API (tests)                                          */
/* ------------------------------------------------------------------ */

/* transposes of the primitives of orc_apply: kind 0 bndry, 1 bndrytt,
 * 2 bndry_aux (3 fields), 3 filter (nf fields, coefficient ck) */
void orc_adj_apply(orc *o, int kind, double *a, int nf, int ck)
{
    if (kind == 0) adj_bndry(o, a);
    else if (kind == 1) adj_bndrytt(o, a);
    else if (kind == 2) adj_bndry_aux(o, a);
    else {
        double *w = zalloc(o->n3);
        for (int m = 1; m <= nf; m++) adj_filter_pass(o, a, m, ck, w);
        free(w);
    }
}

/* vjp of right(t) (smooth = t != 0) or of the residual capture (init) at
 * the current state: xb (8 fields) += (dright/dx)^T ab, rb (3 fields, or
 * NULL) -= part on xresd, *etab += part on eta.  ab is not modified. */
void orc_adj_rhs(orc *o, const double *ab, double *xb, double *rb, double *etab, int init, int smooth)
{
    double *a = zalloc(8 * o->n3);
    memcpy(a, ab, sizeof(double) * 8 * o->n3);
    current(o);
    adj_right(o, o->x, o->cur, a, xb, rb, etab, init, smooth);
    free(a);
}

/* Reverse mode of one stepon(time, dt) from the current state x_n: replays
 * the step (the state becomes x_{n+1}, exactly as orc_step), then maps
 * xb = adjoint of x_{n+1} (8 fields, in/out) to the adjoint of x_n; rb
 * (or NULL) accumulates the adjoint of xresd, *etab that of eta. */
void orc_adj_step(orc *o, double time, double dt, double *xb, double *rb, double *etab)
{
    size_t n3 = o->n3, b8 = sizeof(double) * 8 * n3, b3 = sizeof(double) * 3 * n3, q;
    int m, s;
    double *X = o->x, *XM = o->xm, *XF = o->xfold, *XD = o->xdif;
    double *Xs = zalloc(4 * 8 * n3), *Cs = zalloc(4 * 3 * n3);
    double *xbm = zalloc(8 * n3), *xbf = zalloc(8 * n3), *d = zalloc(8 * n3), *t = zalloc(8 * n3);
    double *w = zalloc(n3);
    const double ts[4] = {time + dt / 2, time + dt / 2, time + dt, time + dt};
    /* forward replay of stepon (1079-1171), keeping stage states and currents */
    for (m = 0; m < 8; m++) EXT_LOOP(XF[m * n3 + q] = X[m * n3 + q];);
    for (s = 0; s < 4; s++) {
        memcpy(Xs + s * 8 * n3, X, b8);
        right(o, ts[s], 0);
        memcpy(Cs + s * 3 * n3, o->cur, b3);
        if (s == 0)
            for (m = 0; m < 8; m++) EXT_LOOP(
                XM[m * n3 + q] = XF[m * n3 + q] + XD[m * n3 + q] * dt / 6.;
                X[m * n3 + q] = XF[m * n3 + q] + XD[m * n3 + q] * dt / 2.;);
        else if (s == 1)
            for (m = 0; m < 8; m++) EXT_LOOP(
                XM[m * n3 + q] = XM[m * n3 + q] + XD[m * n3 + q] * dt / 3.;
                X[m * n3 + q] = XF[m * n3 + q] + XD[m * n3 + q] * dt / 2.;);
        else if (s == 2)
            for (m = 0; m < 8; m++) EXT_LOOP(
                XM[m * n3 + q] = XM[m * n3 + q] + XD[m * n3 + q] * dt / 3.;
                X[m * n3 + q] = XF[m * n3 + q] + XD[m * n3 + q] * dt;);
        else
            for (m = 0; m < 8; m++) EXT_LOOP(X[m * n3 + q] = XM[m * n3 + q] + XD[m * n3 + q] * dt / 6.;);
    }
    {
        const double *eqs[8] = {o->rhoi, NULL, NULL, NULL, o->bxi, o->byi, o->bzi, o->pri};
        for (m = 0; m < 8; m++) EXT_LOOP(XM[m * n3 + q] = X[m * n3 + q] - (eqs[m] ? eqs[m][q] : 0.0););
        avrg_state(o, XM, 2, 1);
        for (m = 0; m < 8; m++) EXT_LOOP(X[m * n3 + q] = XM[m * n3 + q] + (eqs[m] ? eqs[m][q] : 0.0););
    }
    current(o);

    /* reverse: perturbation filter (1135-1165) */
    adj_avrg_state(o, xb, 2, 1, w);
    /* stage 4: x = xm + d4 dt/6 */
    memcpy(xbm, xb, b8);
    for (q = 0; q < 8 * n3; q++) d[q] = xb[q] * dt / 6.;
    memset(t, 0, b8);
    adj_right(o, Xs + 3 * 8 * n3, Cs + 3 * 3 * n3, d, t, rb, etab, 0, 1);
    /* stage 3: x3 = xf + d3 dt, xm += d3 dt/3 */
    memcpy(xbf, t, b8);
    for (q = 0; q < 8 * n3; q++) d[q] = t[q] * dt + xbm[q] * dt / 3.;
    memset(t, 0, b8);
    adj_right(o, Xs + 2 * 8 * n3, Cs + 2 * 3 * n3, d, t, rb, etab, 0, 1);
    /* stage 2: x2 = xf + d2 dt/2, xm += d2 dt/3 */
    for (q = 0; q < 8 * n3; q++) { xbf[q] += t[q]; d[q] = t[q] * dt / 2. + xbm[q] * dt / 3.; }
    memset(t, 0, b8);
    adj_right(o, Xs + 1 * 8 * n3, Cs + 1 * 3 * n3, d, t, rb, etab, 0, 1);
    /* stage 1: x1 = xf + d1 dt/2, xm = xf + d1 dt/6 */
    for (q = 0; q < 8 * n3; q++) { xbf[q] += t[q] + xbm[q]; d[q] = t[q] * dt / 2. + xbm[q] * dt / 6.; }
    adj_right(o, Xs, Cs, d, xbf, rb, etab, 0, 1);
    memcpy(xb, xbf, b8);
    free(Xs); free(Cs); free(xbm); free(xbf); free(d); free(t); free(w);
}

/* vjp of the residual capture (initia, t3dsdsc.f:816-834) at the current
 * state: xb += (dxresd/dx)^T rb */
void orc_adj_residual(orc *o, const double *rb, double *xb)
{
    size_t n3 = o->n3;
    double *a = zalloc(8 * n3), *keep = zalloc(3 * n3);
    /* the capture runs right with xresd = 0 */
    memcpy(keep, o->xresd, sizeof(double) * 3 * n3);
    memset(o->xresd, 0, sizeof(double) * 3 * n3);
    memcpy(a + n3, rb, sizeof(double) * 3 * n3);
    current(o);
    adj_right(o, o->x, o->cur, a, xb, NULL, NULL, 1, 0);
    memcpy(o->xresd, keep, sizeof(double) * 3 * n3);
    free(a); free(keep);
}

/* transpose of periodic_y on nf fields (folds image-row adjoints into the
 * rows they copy) */
void orc_adj_fold_y(orc *o, double *a, int nf)
{
    if (o->c.periodic_y) adj_periodic_y(o, a, nf);
}
