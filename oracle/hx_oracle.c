/*
 * hx_oracle.c -- FP64 CPU restatement of HydroEXA's per-level shallow-water
 * advance.  TEST INFRASTRUCTURE ONLY: the parity tests, the package smoke
 * check and the CPU-baseline leg of the benchmark load this library as the
 * checker.  Nothing in the product path links or calls it.
 *
 * Every function restates the reference routine named in its header comment,
 * keeping the reference's operation order so that a build with
 * -ffp-contract=off produces the same doubles as the HIP kernels built with
 * the same policy.  Citations are path:line in the HydroEXA source tree.
 *
 * Pinning (what checks this file independently of a retyped formula) lives in
 * tests/test_oracle_*.py: upwind/consistency/symmetry identities of the Roe
 * fluxes, lake-at-rest (C-property), mass conservation, y-invariance of a 1-D
 * dam break, the SWASHES dam-break closed form, van Leer's harmonic-mean form,
 * the implicit Manning equation, linear-field gradients for tagging, and
 * brute-force numpy restatements for masked_average_down and pruning.
 *
 * Layout: dense row-major arrays, i fastest, element (i,j) at j*nx + i.
 * Boundary codes per side: 0 = reflective wall, 1 = transmissive.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define GRAV   9.81      /* include/utils/Constants.H:25 */
#define TOLDRY 1e-6      /* include/utils/Constants.H:63 */
#define TOL12  1e-12     /* include/utils/Constants.H:70 */
#define TOL10  1e-10     /* include/utils/Constants.H:78 */
#define NODATA (-9999.0)

#define HXO_OK 0
#define HXO_E_BADARG 2
#define HXO_E_DT_COLLAPSE 3
#define HXO_E_NOMEM 5

/* std::max / std::min semantics: max(a,b) = (a<b) ? b : a. */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* ------------------------------------------------------------------------ */
/* Per-face plain Roe flux with split bed/friction sources.                 */
/* Restates include/solvers/kernels/Roe.H:50-309.                           */
/* L and R are {h, hu, hv, z, n} of the left and right cells.               */
/* ------------------------------------------------------------------------ */
void hxo_roe_face(const double *L, const double *R, double dt, double ds,
                  double nx, double ny, double *F, double *Sm, double *Sp)
{
    (void)dt; /* carried but unused by the reference kernel */
    const double hl = (L[0] >= TOLDRY) ? L[0] : 0.0;            /* :68-69 */
    const double hr = (R[0] >= TOLDRY) ? R[0] : 0.0;
    const double sql = (hl > 0.0) ? sqrt(hl) : 0.0;             /* :71-72 */
    const double sqr = (hr > 0.0) ? sqrt(hr) : 0.0;
    const double ul = (hl > 0.0) ? L[1] / hl : 0.0;              /* :78-82 */
    const double vl = (hl > 0.0) ? L[2] / hl : 0.0;
    const double ur = (hr > 0.0) ? R[1] / hr : 0.0;
    const double vr = (hr > 0.0) ? R[2] / hr : 0.0;
    const double unl = ul * nx + vl * ny;                        /* :85-89 */
    const double utl = -ul * ny + vl * nx;
    const double unr = ur * nx + vr * ny;
    const double utr = -ur * ny + vr * nx;

    const double hmean = 0.5 * (hl + hr);                        /* :96-102 */
    const double ht = (hmean > TOLDRY) ? hmean : TOLDRY;
    const double ct = sqrt(GRAV * ht);
    const double dsq = sql + sqr + TOL12;
    const double ut = (unl * sql + unr * sqr) / dsq;
    const double vt = (utl * sql + utr * sqr) / dsq;

    double lam0 = ut - ct;                                       /* :104-106 */
    const double lam1 = ut;
    double lam2 = ut + ct;

    {   /* Harten-Hyman entropy fix, :113-127 */
        double el = unl - sqrt(GRAV * hl);
        double er = unr - sqrt(GRAV * hr);
        if (el < 0.0 && er > 0.0) lam0 = el * (er - lam0) / (er - el);
        el = unl + sqrt(GRAV * hl);
        er = unr + sqrt(GRAV * hr);
        if (el < 0.0 && er > 0.0) lam2 = er * (lam2 - el) / (er - el);
    }

    const double den = 2.0 * ct + TOL12;                         /* :129 */
    const double dh = hr - hl;                                   /* :132-134 */
    const double dhu = hr * unr - hl * unl;
    const double dhv = hr * utr - hl * utl;
    const double a0 = (lam2 * dh - dhu) / den;                   /* :140-142 */
    const double a1 = (-vt * dh + dhv) / ct;
    const double a2 = (-lam0 * dh + dhu) / den;

    /* bed thrust and Manning friction, :151-189 */
    const double dz = R[3] - L[3];
    const double etl = hl + L[3];
    const double etr = hr + R[3];
    const double dd = etr - etl;
    const double thrust_a = -GRAV * ht * dz;
    double dzt;
    if ((dz >= 0.0) && (etl < R[3]))      dzt = hl;
    else if ((dz < 0.0) && (etr < L[3]))  dzt = hr;
    else                                  dzt = dz;
    const double hup = (dz >= 0.0) ? hl : hr;
    const double thrust_b = -GRAV * (hup - 0.5 * fabs(dzt)) * dzt;
    const double thrust = ((dz * dd >= 0.0) && (ut * dz > 0.0)) ? smax(thrust_a, thrust_b) : thrust_b;
    const double nt = 0.5 * (L[4] + R[4]);
    const double sf = (nt * nt * sqrt(ut * ut + vt * vt) * ut) / smax(TOLDRY, ht * cbrt(ht));
    const double tau = GRAV * ht * sf * ds;

    double b0 = (thrust - tau) / den;                            /* :196-198 */
    const double b1 = 0.0;
    double b2 = -b0;

    {   /* positivity guard, :200-247 */
        const double l0s = (fabs(lam0) > TOL12) ? lam0 : (lam0 >= 0.0 ? TOL12 : -TOL12);
        const double l2s = (fabs(lam2) > TOL12) ? lam2 : (lam2 >= 0.0 ? TOL12 : -TOL12);
        const double h1s = hl + a0 - (b0 / l0s);
        const double h3s = hr - a2 + (b2 / l2s);
        double mlam = fabs(lam0);
        mlam = smax(mlam, fabs(lam1));
        mlam = smax(mlam, fabs(lam2));
        const double loc_dt = ds / (mlam + TOL12);
        const double d1 = hl - h1s;
        const double d1s = (fabs(d1) > TOL12) ? d1 : (d1 >= 0.0 ? TOL12 : -TOL12);
        const double d3 = hr - h3s;
        const double d3s = (fabs(d3) > TOL12) ? d3 : (d3 >= 0.0 ? TOL12 : -TOL12);
        const double dt1 = (ds * 0.5 / l0s) * (hl / d1s);
        const double dt3 = (ds * 0.5 / l2s) * (hr / d3s);
        if ((h1s < -TOLDRY) && (h3s > TOLDRY) && (dt1 < loc_dt)) {
            const double m1 = -(hl + a0) * fabs(lam0);
            const double m3 = -(hr - a2) * lam2;
            if (-m1 >= m3) b0 = m1;
            b2 = -b0;
        }
        if ((h1s > TOLDRY) && (h3s < -TOLDRY) && (dt3 < loc_dt)) {
            const double m1 = -(hl + a0) * fabs(lam0);
            const double m3 = -(hr - a2) * lam2;
            if (-m3 >= m1) b2 = m3;
            b0 = -b2;
        }
    }

    for (int c = 0; c < 3; ++c) { Sm[c] = 0.0; Sp[c] = 0.0; }   /* :249-252 */
    F[0] = 0.5 * (hl * unl + hr * unr);                          /* :255-258 */
    F[1] = 0.5 * (hl * unl * unl + 0.5 * GRAV * hl * hl +
                  hr * unr * unr + 0.5 * GRAV * hr * hr);
    F[2] = 0.5 * (hl * unl * utl + hr * unr * utr);

    /* wave sum and upwind source distribution, :265-290 */
    const double lam[3] = {lam0, lam1, lam2};
    const double alp[3] = {a0, a1, a2};
    const double bet[3] = {b0, b1, b2};
    for (int k = 0; k < 3; ++k) {
        const double e0 = (k == 1) ? 0.0 : 1.0;
        const double e1 = (k == 0) ? lam0 : (k == 1 ? 0.0 : lam2);
        const double e2 = (k == 0) ? vt : (k == 1 ? ct : vt);
        const double w = 0.5 * fabs(lam[k]) * alp[k];
        F[0] -= w * e0;
        F[1] -= w * e1;
        F[2] -= w * e2;
        if (lam[k] > 0.0) {
            Sm[0] += bet[k] * e0; Sm[1] += bet[k] * e1; Sm[2] += bet[k] * e2;
        } else if (lam[k] < 0.0) {
            Sp[0] += bet[k] * e0; Sp[1] += bet[k] * e1; Sp[2] += bet[k] * e2;
        }
    }

    /* rotate back, :293-308 */
    double t1 = F[1], t2 = F[2];
    F[1] = t1 * nx - t2 * ny;  F[2] = t1 * ny + t2 * nx;
    t1 = Sm[1]; t2 = Sm[2];
    Sm[1] = t1 * nx - t2 * ny; Sm[2] = t1 * ny + t2 * nx;
    t1 = Sp[1]; t2 = Sp[2];
    Sp[1] = t1 * nx - t2 * ny; Sp[2] = t1 * ny + t2 * nx;
}

/* ------------------------------------------------------------------------ */
/* Homogeneous Roe flux on reconstructed face states.                       */
/* Restates include/solvers/kernels/HydrostaticRoe.H:43-159.                */
/* ------------------------------------------------------------------------ */
void hxo_hroe_face(double hi, double hui, double hvi, double hj, double huj, double hvj,
                   double nx, double ny, double *F)
{
    hi = (hi >= TOLDRY) ? hi : 0.0;                              /* :59-60 */
    hj = (hj >= TOLDRY) ? hj : 0.0;
    const double sql = (hi > 0.0) ? sqrt(hi) : 0.0;
    const double sqr = (hj > 0.0) ? sqrt(hj) : 0.0;
    const double ul = (hi > 0.0) ? hui / hi : 0.0;
    const double vl = (hi > 0.0) ? hvi / hi : 0.0;
    const double ur = (hj > 0.0) ? huj / hj : 0.0;
    const double vr = (hj > 0.0) ? hvj / hj : 0.0;
    const double unl = ul * nx + vl * ny;
    const double utl = -ul * ny + vl * nx;
    const double unr = ur * nx + vr * ny;
    const double utr = -ur * ny + vr * nx;
    const double hmean = 0.5 * (hi + hj);                        /* :83-89 */
    const double ht = (hmean > TOLDRY) ? hmean : TOLDRY;
    const double ct = sqrt(GRAV * ht);
    const double dsq = sql + sqr + TOL12;
    const double ut = (unl * sql + unr * sqr) / dsq;
    const double vt = (utl * sql + utr * sqr) / dsq;
    double lam0 = ut - ct;
    const double lam1 = ut;
    double lam2 = ut + ct;
    {   /* :99-113 */
        double el = unl - sqrt(GRAV * hi);
        double er = unr - sqrt(GRAV * hj);
        if (el < 0.0 && er > 0.0) lam0 = el * (er - lam0) / (er - el);
        el = unl + sqrt(GRAV * hi);
        er = unr + sqrt(GRAV * hj);
        if (el < 0.0 && er > 0.0) lam2 = er * (lam2 - el) / (er - el);
    }
    const double den = 2.0 * ct + TOL12;
    const double dh = hj - hi;
    const double dhu = hj * unr - hi * unl;
    const double dhv = hj * utr - hi * utl;
    const double a0 = (lam2 * dh - dhu) / den;                   /* :123-125 */
    const double a1 = (-vt * dh + dhv) / ct;
    const double a2 = (-lam0 * dh + dhu) / den;
    F[0] = 0.5 * (hi * unl + hj * unr);                          /* :128-131 */
    F[1] = 0.5 * (hi * unl * unl + 0.5 * GRAV * hi * hi +
                  hj * unr * unr + 0.5 * GRAV * hj * hj);
    F[2] = 0.5 * (hi * unl * utl + hj * unr * utr);
    const double lam[3] = {lam0, lam1, lam2};
    const double alp[3] = {a0, a1, a2};
    for (int k = 0; k < 3; ++k) {                                /* :137-150 */
        const double e0 = (k == 1) ? 0.0 : 1.0;
        const double e1 = (k == 0) ? lam0 : (k == 1 ? 0.0 : lam2);
        const double e2 = (k == 0) ? vt : (k == 1 ? ct : vt);
        const double w = 0.5 * fabs(lam[k]) * alp[k];
        F[0] -= w * e0;
        F[1] -= w * e1;
        F[2] -= w * e2;
    }
    const double t1 = F[1], t2 = F[2];                           /* :153-158 */
    F[1] = t1 * nx - t2 * ny;
    F[2] = t1 * ny + t2 * nx;
}

/* van Leer limited slope; src/solvers/HydrostaticRoe.cpp:173-195.  The
 * limiter's abs() is taken on the floating value (the device binding). */
double hxo_vanleer_slope(double qm, double q, double qp)
{
    const double dm = q - qm;
    const double dp = qp - q;
    if (fabs(dp) <= TOL12) return 0.0;
    const double r = dm / dp;
    return ((r + fabs(r)) / (1.0 + fabs(r))) * dp;
}

/* One hydrostatic-reconstruction face.  c holds four cells along the face
 * normal, {Lm, L, R, Rp}, each {h, hu, hv, z} with h already rebuilt as
 * max(0, eta - z).  Restates src/solvers/HydrostaticRoe.cpp:218-396. */
void hxo_hydro_face(const double *c, double nx, double ny, double *FL, double *FR)
{
    const double *Lm = c, *L = c + 4, *R = c + 8, *Rp = c + 12;
    const double zL = L[3], zR = R[3];
    if (fabs(zL + 9999.0) < TOL10 || fabs(zR + 9999.0) < TOL10) {  /* :237-245 */
        for (int k = 0; k < 3; ++k) { FL[k] = 0.0; FR[k] = 0.0; }
        return;
    }
    const double zf = smax(zL, zR);                              /* :247 */
    /* eta, :256-267 */
    const double eLm = Lm[0] + Lm[3];
    const double eL = L[0] + zL;
    const double eR = R[0] + zR;
    const double eRp = Rp[0] + Rp[3];
    const double etaLf = eL + 0.5 * hxo_vanleer_slope(eLm, eL, eR);
    const double etaRf = eR - 0.5 * hxo_vanleer_slope(eL, eR, eRp);
    /* u and v, :276-341 */
    const double uLm = (Lm[0] > TOL12) ? Lm[1] / Lm[0] : 0.0;
    const double uL = (L[0] > TOL12) ? L[1] / L[0] : 0.0;
    const double uR = (R[0] > TOL12) ? R[1] / R[0] : 0.0;
    const double uRp = (Rp[0] > TOL12) ? Rp[1] / Rp[0] : 0.0;
    const double uLf = uL + 0.5 * hxo_vanleer_slope(uLm, uL, uR);
    const double uRf = uR - 0.5 * hxo_vanleer_slope(uL, uR, uRp);
    const double vLm = (Lm[0] > TOL12) ? Lm[2] / Lm[0] : 0.0;
    const double vL = (L[0] > TOL12) ? L[2] / L[0] : 0.0;
    const double vR = (R[0] > TOL12) ? R[2] / R[0] : 0.0;
    const double vRp = (Rp[0] > TOL12) ? Rp[2] / Rp[0] : 0.0;
    const double vLf = vL + 0.5 * hxo_vanleer_slope(vLm, vL, vR);
    const double vRf = vR - 0.5 * hxo_vanleer_slope(vL, vR, vRp);
    /* Audusse hydrostatic reconstruction, :351-361 */
    const double hLu = smax(0.0, etaLf - zL);
    const double hRu = smax(0.0, etaRf - zR);
    const double hLf = smax(0.0, etaLf - zf);
    const double hRf = smax(0.0, etaRf - zf);
    const double huLf = (hLf > 0.0) ? hLf * uLf : 0.0;
    const double hvLf = (hLf > 0.0) ? hLf * vLf : 0.0;
    const double huRf = (hRf > 0.0) ? hRf * uRf : 0.0;
    const double hvRf = (hRf > 0.0) ? hRf * vRf : 0.0;
    double f[3];
    hxo_hroe_face(hLf, huLf, hvLf, hRf, huRf, hvRf, nx, ny, f);   /* :369-373 */
    const double pL = 0.5 * GRAV * (hLu * hLu - hLf * hLf);      /* :385-396 */
    const double pR = 0.5 * GRAV * (hRu * hRu - hRf * hRf);
    FL[0] = f[0]; FL[1] = f[1] + pL * nx; FL[2] = f[2] + pL * ny;
    FR[0] = f[0]; FR[1] = f[1] + pR * nx; FR[2] = f[2] + pR * ny;
}

/* ------------------------------------------------------------------------ */
/* Ghost fill.  Padded arrays hold G=2 ghost layers.  Physical-boundary     */
/* semantics follow include/boundaries/BCFill.H:49-132 and the BCRec codes  */
/* set in src/state/AmrMeshState_Init.cpp:52-70: the free surface uses the  */
/* scalar rule (reflect_even / foextrap), the normal momentum reflect_odd / */
/* foextrap, the tangential momentum reflect_even / foextrap, and terrain   */
/* always foextrap.  x sides are filled first for valid rows, then y sides  */
/* for all columns, so corner ghosts take the y rule (they are never read   */
/* by the face stencils).                                                   */
/* ------------------------------------------------------------------------ */
#define G 2
typedef struct { int nx, ny, P; } pad_t;
static inline size_t pidx(const pad_t *p, int i, int j) { return (size_t)(j + G) * p->P + (i + G); }

/* kind: 0 scalar-even, 1 normal-x (odd in x, even in y), 2 normal-y (even in
 * x, odd in y), 3 always foextrap */
static void fill_ghosts(const pad_t *p, double *a, int kind, const int *bclo, const int *bchi)
{
    const int nx = p->nx, ny = p->ny;
    for (int j = 0; j < ny; ++j) {
        for (int g = 1; g <= G; ++g) {
            /* low x side: ghost i = -g */
            {
                const int refl = (kind != 3) && (bclo[0] == 0);
                const double sgn = (refl && kind == 1) ? -1.0 : 1.0;
                const int src = refl ? (g - 1) : 0;
                a[pidx(p, -g, j)] = refl ? sgn * a[pidx(p, src, j)] : a[pidx(p, src, j)];
            }
            {
                const int refl = (kind != 3) && (bchi[0] == 0);
                const double sgn = (refl && kind == 1) ? -1.0 : 1.0;
                const int src = refl ? (nx - g) : (nx - 1);
                a[pidx(p, nx - 1 + g, j)] = refl ? sgn * a[pidx(p, src, j)] : a[pidx(p, src, j)];
            }
        }
    }
    for (int i = -G; i < nx + G; ++i) {
        for (int g = 1; g <= G; ++g) {
            {
                const int refl = (kind != 3) && (bclo[1] == 0);
                const double sgn = (refl && kind == 2) ? -1.0 : 1.0;
                const int src = refl ? (g - 1) : 0;
                a[pidx(p, i, -g)] = refl ? sgn * a[pidx(p, i, src)] : a[pidx(p, i, src)];
            }
            {
                const int refl = (kind != 3) && (bchi[1] == 0);
                const double sgn = (refl && kind == 2) ? -1.0 : 1.0;
                const int src = refl ? (ny - g) : (ny - 1);
                a[pidx(p, i, ny - 1 + g)] = refl ? sgn * a[pidx(p, i, src)] : a[pidx(p, i, src)];
            }
        }
    }
}

static int check_bc(const int *lo, const int *hi)
{
    for (int d = 0; d < 2; ++d)
        if ((lo[d] != 0 && lo[d] != 1) || (hi[d] != 0 && hi[d] != 1)) return 0;
    return 1;
}

/* Build the padded, halo-filled working state shared by both solvers:
 * eta = h + z on valid cells (src/state/AmrMeshState_AMR.cpp:218-220), ghost
 * fill as above, then the one rebuild primitive h = max(0, eta - z) over the
 * grown region (src/solvers/Roe.cpp:244-247, HydrostaticRoe.cpp:495-499). */
typedef struct { double *he, *hu, *hv, *z, *n, *eta, *faces; } work_t;

/* Workspace of one level: padded halo arrays and the face arrays, allocated
 * once per loop (the reference's per-call MultiFabs come from AMReX's
 * caching arena, so they are not re-faulted every step either). */
static int work_alloc(const pad_t *p, work_t *w)
{
    const size_t np = (size_t)p->P * (p->ny + 2 * G);
    const size_t nf = (size_t)(p->nx + 1) * p->ny + (size_t)p->nx * (p->ny + 1);
    double *buf = (double *)malloc(6 * np * sizeof(double));
    double *fb = (double *)malloc(9 * nf * sizeof(double));
    if (!buf || !fb) { free(buf); free(fb); return HXO_E_NOMEM; }
    w->he = buf; w->hu = buf + np; w->hv = buf + 2 * np; w->z = buf + 3 * np; w->n = buf + 4 * np;
    w->eta = buf + 5 * np;
    w->faces = fb;
    return HXO_OK;
}

static void work_free(work_t *w) { free(w->he); free(w->faces); }

static void build_work(const pad_t *p, const double *h, const double *hu, const double *hv,
                       const double *z, const double *n, const int *bclo, const int *bchi,
                       work_t *w)
{
    const size_t np = (size_t)p->P * (p->ny + 2 * G);
    double *eta = w->eta;
    const int nx = p->nx, ny = p->ny;
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t s = (size_t)j * nx + i, d = pidx(p, i, j);
            eta[d] = h[s] + z[s];
            w->hu[d] = hu[s]; w->hv[d] = hv[s]; w->z[d] = z[s]; w->n[d] = n[s];
        }
    fill_ghosts(p, eta, 0, bclo, bchi);
    fill_ghosts(p, w->hu, 1, bclo, bchi);
    fill_ghosts(p, w->hv, 2, bclo, bchi);
    fill_ghosts(p, w->z, 3, bclo, bchi);
    fill_ghosts(p, w->n, 3, bclo, bchi);
    #pragma omp parallel for schedule(static)
    for (size_t k = 0; k < np; ++k) w->he[k] = smax(0.0, eta[k] - w->z[k]);
}

/* ------------------------------------------------------------------------ */
/* Plain Roe step on one box covering the level.                            */
/* Restates src/solvers/Roe.cpp:250-434 (swap, fill, rebuild, faces,        */
/* fluctuation-difference update).  U_o's depth is overwritten by the       */
/* rebuilt h before the update (:339), so the update starts from it.        */
/* ------------------------------------------------------------------------ */
static void roe_core(const pad_t *pp, const work_t *wp, double dt, double dx, double dy,
                     double *ho, double *huo, double *hvo)
{
    const pad_t p = *pp;
    const work_t w = *wp;
    const int nx = p.nx, ny = p.ny;
    const size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *fb = w.faces;
    double *Fx = fb, *Smx = Fx + 3 * nfx, *Spx = Smx + 3 * nfx;
    double *Fy = Spx + 3 * nfx, *Smy = Fy + 3 * nfy, *Spy = Smy + 3 * nfy;

    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {   /* x faces, Roe.cpp:201-234, dir 0 */
            const size_t a = pidx(&p, i - 1, j), b = pidx(&p, i, j);
            const double L[5] = {w.he[a], w.hu[a], w.hv[a], w.z[a], w.n[a]};
            const double R[5] = {w.he[b], w.hu[b], w.hv[b], w.z[b], w.n[b]};
            const size_t f = (size_t)j * (nx + 1) + i;
            hxo_roe_face(L, R, dt, dx, 1.0, 0.0, Fx + 3 * f, Smx + 3 * f, Spx + 3 * f);
        }
    #pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {    /* y faces, dir 1 */
            const size_t a = pidx(&p, i, j - 1), b = pidx(&p, i, j);
            const double L[5] = {w.he[a], w.hu[a], w.hv[a], w.z[a], w.n[a]};
            const double R[5] = {w.he[b], w.hu[b], w.hv[b], w.z[b], w.n[b]};
            const size_t f = (size_t)j * nx + i;
            hxo_roe_face(L, R, dt, dy, 0.0, 1.0, Fy + 3 * f, Smy + 3 * f, Spy + 3 * f);
        }
    /* update, Roe.cpp:416-432 */
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t c = pidx(&p, i, j), s = (size_t)j * nx + i;
            const size_t xl = (size_t)j * (nx + 1) + i, xr = xl + 1;
            const size_t yl = (size_t)j * nx + i, yr = yl + nx;
            const double uo[3] = {w.he[c], w.hu[c], w.hv[c]};
            double un[3];
            for (int k = 0; k < 3; ++k)
                un[k] = uo[k]
                      - (dt / dx) * (Fx[3 * xr + k] - Fx[3 * xl + k])
                      - (dt / dy) * (Fy[3 * yr + k] - Fy[3 * yl + k])
                      - (dt / dx) * (Spx[3 * xr + k] + Smx[3 * xl + k])
                      - (dt / dy) * (Spy[3 * yr + k] + Smy[3 * yl + k]);
            ho[s] = un[0]; huo[s] = un[1]; hvo[s] = un[2];
        }
}

static void roe_step_ws(int nx, int ny, const double *h, const double *hu, const double *hv,
                        const double *z, const double *n, double dt, double dx, double dy,
                        const int *bclo, const int *bchi, double *ho, double *huo, double *hvo,
                        work_t *wp)
{
    pad_t p = {nx, ny, nx + 2 * G};
    build_work(&p, h, hu, hv, z, n, bclo, bchi, wp);
    roe_core(&p, wp, dt, dx, dy, ho, huo, hvo);
}

int hxo_roe_step(int nx, int ny, const double *h, const double *hu, const double *hv,
                 const double *z, const double *n, double dt, double dx, double dy,
                 const int *bclo, const int *bchi, double *ho, double *huo, double *hvo)
{
    if (nx < 2 || ny < 2 || !check_bc(bclo, bchi)) return HXO_E_BADARG;
    pad_t p = {nx, ny, nx + 2 * G};
    work_t w;
    if (work_alloc(&p, &w)) return HXO_E_NOMEM;
    roe_step_ws(nx, ny, h, hu, hv, z, n, dt, dx, dy, bclo, bchi, ho, huo, hvo, &w);
    work_free(&w);
    return HXO_OK;
}

/* ------------------------------------------------------------------------ */
/* Hydrostatic-reconstruction (well-balanced) Roe step on one box.          */
/* Restates src/solvers/HydrostaticRoe.cpp:400-606.  The update starts from */
/* U_o as stored (its depth is not rebuilt; the rebuild goes to U_ext).     */
/* ------------------------------------------------------------------------ */
/* Stored state of the update: element (i, j) at so + j*sp + i. */
static void hydro_core(const pad_t *pp, const work_t *wp, const double *h, const double *hu,
                       const double *hv, const double *n, size_t sp, size_t so, double dt,
                       double dx, double dy, double *ho, double *huo, double *hvo)
{
    const pad_t p = *pp;
    const work_t w = *wp;
    const int nx = p.nx, ny = p.ny;
    const size_t nfx = (size_t)(nx + 1) * ny, nfy = (size_t)nx * (ny + 1);
    double *fb = w.faces;
    double *FLx = fb, *FRx = FLx + 3 * nfx, *FLy = FRx + 3 * nfx, *FRy = FLy + 3 * nfy;

    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i <= nx; ++i) {
            double c[16];
            for (int q = 0; q < 4; ++q) {
                const size_t a = pidx(&p, i - 2 + q, j);
                c[4 * q] = w.he[a]; c[4 * q + 1] = w.hu[a]; c[4 * q + 2] = w.hv[a]; c[4 * q + 3] = w.z[a];
            }
            const size_t f = (size_t)j * (nx + 1) + i;
            hxo_hydro_face(c, 1.0, 0.0, FLx + 3 * f, FRx + 3 * f);
        }
    #pragma omp parallel for schedule(static)
    for (int j = 0; j <= ny; ++j)
        for (int i = 0; i < nx; ++i) {
            double c[16];
            for (int q = 0; q < 4; ++q) {
                const size_t a = pidx(&p, i, j - 2 + q);
                c[4 * q] = w.he[a]; c[4 * q + 1] = w.hu[a]; c[4 * q + 2] = w.hv[a]; c[4 * q + 3] = w.z[a];
            }
            const size_t f = (size_t)j * nx + i;
            hxo_hydro_face(c, 0.0, 1.0, FLy + 3 * f, FRy + 3 * f);
        }
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t s = (size_t)j * nx + i, si = so + (size_t)j * sp + i;
            const size_t xl = (size_t)j * (nx + 1) + i, xr = xl + 1;
            const size_t yl = (size_t)j * nx + i, yr = yl + nx;
            const double uo[3] = {h[si], hu[si], hv[si]};
            double un[3];
            for (int k = 0; k < 3; ++k)                          /* :554-564 */
                un[k] = uo[k]
                      - (dt / dx) * (FLx[3 * xr + k] - FRx[3 * xl + k])
                      - (dt / dy) * (FLy[3 * yr + k] - FRy[3 * yl + k]);
            const double hn = un[0];
            if (hn >= TOLDRY) {                                  /* :566-595 */
                const double qx = un[1], qy = un[2];
                const double nsq = n[si] * n[si];
                const double inv_h = 1.0 / hn;
                const double u = qx * inv_h, v = qy * inv_h;
                const double modM = sqrt(u * u + v * v);
                if (nsq > TOL12 && modM > TOL12) {
                    const double h43 = hn * cbrt(hn);
                    const double tt = dt * GRAV * nsq * modM / h43;
                    const double ff = sqrt(1.0 + 4.0 * tt);
                    const double fac = 2.0 / (1.0 + ff);
                    un[1] = qx * fac;
                    un[2] = qy * fac;
                }
            } else {                                             /* :596-603 */
                un[0] = smax(hn, 0.0);
                un[1] = 0.0;
                un[2] = 0.0;
            }
            ho[s] = un[0]; huo[s] = un[1]; hvo[s] = un[2];
        }
}

static void hydro_step_ws(int nx, int ny, const double *h, const double *hu, const double *hv,
                          const double *z, const double *n, double dt, double dx, double dy,
                          const int *bclo, const int *bchi, double *ho, double *huo, double *hvo,
                          work_t *wp)
{
    pad_t p = {nx, ny, nx + 2 * G};
    build_work(&p, h, hu, hv, z, n, bclo, bchi, wp);
    hydro_core(&p, wp, h, hu, hv, n, (size_t)nx, 0, dt, dx, dy, ho, huo, hvo);
}

int hxo_hydro_step(int nx, int ny, const double *h, const double *hu, const double *hv,
                   const double *z, const double *n, double dt, double dx, double dy,
                   const int *bclo, const int *bchi, double *ho, double *huo, double *hvo)
{
    if (nx < 2 || ny < 2 || !check_bc(bclo, bchi)) return HXO_E_BADARG;
    pad_t p = {nx, ny, nx + 2 * G};
    work_t w;
    if (work_alloc(&p, &w)) return HXO_E_NOMEM;
    hydro_step_ws(nx, ny, h, hu, hv, z, n, dt, dx, dy, bclo, bchi, ho, huo, hvo, &w);
    work_free(&w);
    return HXO_OK;
}

/* ------------------------------------------------------------------------ */
/* Step of a box whose 2 ghost layers are given (a fine level inside a      */
/* coarse one: FillPatchTwoLevels wrote them).  Inputs are padded           */
/* (ny+4) x (nx+4), pitch nx+4, valid cell (0,0) at 2*(nx+4)+2; eta = h + z */
/* and h = max(0, eta - z) over the whole padded region, then the same      */
/* faces and update as above.  Outputs are the nx x ny valid cells.  eta   */
/* (optional, padded) replaces h + z as the free surface of each cell: a    */
/* ghost across a physical side carries its source cell's h + z on the     */
/* edge cell's bed.                                                         */
/* ------------------------------------------------------------------------ */
int hxo_step_padded(int solver, int nx, int ny, const double *h, const double *hu,
                    const double *hv, const double *z, const double *n, const double *eta,
                    double dt, double dx, double dy, double *ho, double *huo, double *hvo)
{
    if (nx < 2 || ny < 2 || (solver != 0 && solver != 1)) return HXO_E_BADARG;
    pad_t p = {nx, ny, nx + 2 * G};
    work_t w;
    if (work_alloc(&p, &w)) return HXO_E_NOMEM;
    const size_t np = (size_t)p.P * (ny + 2 * G);
    for (size_t k = 0; k < np; ++k) {
        w.eta[k] = eta ? eta[k] : h[k] + z[k];
        w.hu[k] = hu[k]; w.hv[k] = hv[k]; w.z[k] = z[k]; w.n[k] = n[k];
        w.he[k] = smax(0.0, w.eta[k] - w.z[k]);
    }
    if (solver == 0)
        roe_core(&p, &w, dt, dx, dy, ho, huo, hvo);
    else
        hydro_core(&p, &w, h, hu, hv, n, (size_t)p.P, (size_t)G * p.P + G, dt, dx, dy, ho, huo,
                   hvo);
    work_free(&w);
    return HXO_OK;
}

/* ------------------------------------------------------------------------ */
/* CFL estimate of one level before the cfl factor.                         */
/* Restates src/solvers/Roe.cpp:38-81 and HydrostaticRoe.cpp:37-70 (the two */
/* cell functions are identical: 1e-12 and TOL12 are the same value).       */
/* Empty levels return the ReduceOpMin/Max identities.                      */
/* ------------------------------------------------------------------------ */
void hxo_level_dt(int nx, int ny, const double *h, const double *hu, const double *hv,
                  const double *z, double dx, double dy, double *dt_est, double *max_vel)
{
    double dmin = 1.7976931348623157e308, vmax = -1.7976931348623157e308;
    #pragma omp parallel for schedule(static) reduction(min:dmin) reduction(max:vmax)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t s = (size_t)j * nx + i;
            double dtc, mu;
            if (fabs(z[s] - (-9999.0)) < TOL10) { dtc = 1.e30; mu = 0.0; }
            else {
                const double hh = h[s];
                dtc = 1.e6; mu = 0.0;
                if (hh > 1.e-12) {
                    const double u = hu[s] / hh, v = hv[s] / hh;
                    const double c = sqrt(GRAV * hh);
                    const double sx = fabs(u) + c, sy = fabs(v) + c;
                    dtc = 1.0 / (sx / dx + sy / dy);
                    mu = smax(fabs(u), fabs(v));
                }
            }
            dmin = smin(dmin, dtc);
            vmax = smax(vmax, mu);
        }
    *dt_est = dmin;
    *max_vel = vmax;
}

/* ------------------------------------------------------------------------ */
/* Gradient tagging on a halo-filled state (every ghost from the physical   */
/* BC of its component; terrain foextrap).  solver 0 restates               */
/* src/solvers/Roe.cpp:121-179, solver 1 HydrostaticRoe.cpp:109-169.        */
/* tags[s] = 1 where the reference sets TagBox::SET, else 0.                */
/* ------------------------------------------------------------------------ */
static inline double umag_at(const pad_t *p, const double *h, const double *hu,
                             const double *hv, int i, int j)
{
    const size_t a = pidx(p, i, j);
    const double hh = h[a];
    if (hh < 1.0e-6) return 0.0;
    const double u = hu[a] / hh, v = hv[a] / hh;
    return sqrt(u * u + v * v);
}

int hxo_tag(int solver, int nx, int ny, const double *h, const double *hu, const double *hv,
            const double *z, double dx, double dy, const int *bclo, const int *bchi,
            double time, double h_thresh, double u_thresh, double z_thresh, uint8_t *tags)
{
    if (nx < 2 || ny < 2 || !check_bc(bclo, bchi) || (solver != 0 && solver != 1)) return HXO_E_BADARG;
    pad_t p = {nx, ny, nx + 2 * G};
    const size_t np = (size_t)p.P * (ny + 2 * G);
    double *buf = (double *)calloc(4 * np, sizeof(double));
    if (!buf) return HXO_E_NOMEM;
    double *H = buf, *HU = buf + np, *HV = buf + 2 * np, *Z = buf + 3 * np;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t s = (size_t)j * nx + i, d = pidx(&p, i, j);
            H[d] = h[s]; HU[d] = hu[s]; HV[d] = hv[s]; Z[d] = z[s];
        }
    fill_ghosts(&p, H, 0, bclo, bchi);
    fill_ghosts(&p, HU, 1, bclo, bchi);
    fill_ghosts(&p, HV, 2, bclo, bchi);
    fill_ghosts(&p, Z, 3, bclo, bchi);
    #pragma omp parallel for schedule(static)
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const double zE = Z[pidx(&p, i + 1, j)], zW = Z[pidx(&p, i - 1, j)];
            const double zN = Z[pidx(&p, i, j + 1)], zS = Z[pidx(&p, i, j - 1)];
            uint8_t t = 0;
            int skip;
            if (solver == 0)
                skip = (zE == -9999.0 || zW == -9999.0 || zN == -9999.0 || zS == -9999.0);
            else
                skip = (fabs(zE + 9999.0) < TOL10 || fabs(zW + 9999.0) < TOL10 ||
                        fabs(zN + 9999.0) < TOL10 || fabs(zS + 9999.0) < TOL10);
            if (!skip) {
                const double dhdx = (H[pidx(&p, i + 1, j)] - H[pidx(&p, i - 1, j)]) / (2.0 * dx);
                const double dhdy = (H[pidx(&p, i, j + 1)] - H[pidx(&p, i, j - 1)]) / (2.0 * dy);
                if (sqrt(dhdx * dhdx + dhdy * dhdy) > h_thresh) t = 1;
                int zpass_done = 0;
                if (solver == 0 && time <= 1.0e-8) {
                    const double dzdx = (zE - zW) / (2.0 * dx), dzdy = (zN - zS) / (2.0 * dy);
                    if (sqrt(dzdx * dzdx + dzdy * dzdy) > z_thresh) t = 1;
                    zpass_done = 1;
                }
                if (!t) {
                    const double uE = umag_at(&p, H, HU, HV, i + 1, j);
                    const double uW = umag_at(&p, H, HU, HV, i - 1, j);
                    const double uN = umag_at(&p, H, HU, HV, i, j + 1);
                    const double uS = umag_at(&p, H, HU, HV, i, j - 1);
                    const double dudx = (uE - uW) / (2.0 * dx), dudy = (uN - uS) / (2.0 * dy);
                    if (sqrt(dudx * dudx + dudy * dudy) > u_thresh) t = 1;
                    if (solver == 1 && !zpass_done && time <= 1.0e-8) {
                        const double dzdx = (zE - zW) / (2.0 * dx), dzdy = (zN - zS) / (2.0 * dy);
                        if (sqrt(dzdx * dzdx + dzdy * dzdy) > z_thresh) t = 1;
                    }
                }
            }
            tags[(size_t)j * nx + i] = t;
        }
    free(buf);
    return HXO_OK;
}

/* ------------------------------------------------------------------------ */
/* NODATA-aware volume-weighted restriction.                                */
/* Restates include/utils/MaskedAverageDown.H:63-104 (cell loop) and        */
/* src/utils/MaskedAverageDown.cpp:30-59 (fine volume from the fine cell    */
/* size).  fine: ncomp planes of nxf*nyf; crse: ncomp planes of             */
/* (nxf/rx)*(nyf/ry); nvalid: one int per coarse cell and component.       */
/* ------------------------------------------------------------------------ */
int hxo_masked_average_down(int nxf, int nyf, int ncomp, const double *fine, int rx, int ry,
                            double dxf, double dyf, double nodata, double *crse, int32_t *nvalid)
{
    if (rx < 1 || ry < 1 || nxf % rx || nyf % ry || ncomp < 1) return HXO_E_BADARG;
    const int nxc = nxf / rx, nyc = nyf / ry;
    const double fv = dxf * dyf;
    for (int c = 0; c < ncomp; ++c) {
        const double *f = fine + (size_t)c * nxf * nyf;
        #pragma omp parallel for schedule(static)
        for (int j = 0; j < nyc; ++j)
            for (int i = 0; i < nxc; ++i) {
                double cd = 0.0, cv = 0.0;
                int nv = 0;
                const int total = rx * ry;
                for (int jr = 0; jr < ry; ++jr)
                    for (int ir = 0; ir < rx; ++ir) {
                        const double val = f[(size_t)(j * ry + jr) * nxf + (i * rx + ir)];
                        if (!isnan(val) && fabs(val - nodata) > 1e-3) {
                            ++nv;
                            cv += fv;
                            cd += val * fv;
                        }
                    }
                const size_t o = (size_t)c * nxc * nyc + (size_t)j * nxc + i;
                crse[o] = (nv * 2 >= total && cv > 0.0) ? cd / cv : nodata;
                nvalid[o] = nv;
            }
    }
    return HXO_OK;
}

/* ------------------------------------------------------------------------ */
/* Base-grid pruning predicate.                                             */
/* Restates src/state/AmrMeshState_AMR.cpp:318-405.  boxes: nbox x 4 ints  */
/* {ilo, jlo, ihi, jhi} in level-0 index space; z: static terrain bed of    */
/* nxs*nys cells at 2^scale_log2 times the level-0 resolution.              */
/* keep[b] = box survives; nodata[b] = box holds NODATA.  Returns the       */
/* number of kept boxes (all boxes are kept, pure fluid, when none has a    */
/* valid cell).                                                             */
/* ------------------------------------------------------------------------ */
int hxo_prune_boxes(int nbox, const int32_t *boxes, int nxs, int nys, const double *z,
                    int scale_log2, uint8_t *keep, uint8_t *nodata)
{
    const int sc = 1 << scale_log2;
    int nkeep = 0;
    for (int b = 0; b < nbox; ++b) {
        int ilo = boxes[4 * b] * sc, jlo = boxes[4 * b + 1] * sc;
        int ihi = (boxes[4 * b + 2] + 1) * sc - 1, jhi = (boxes[4 * b + 3] + 1) * sc - 1;
        if (ilo < 0) ilo = 0;
        if (jlo < 0) jlo = 0;
        if (ihi > nxs - 1) ihi = nxs - 1;
        if (jhi > nys - 1) jhi = nys - 1;
        int hv = 0, hn = 0;
        for (int j = jlo; j <= jhi; ++j)
            for (int i = ilo; i <= ihi; ++i) {
                const int isnd = (z[(size_t)j * nxs + i] <= -9998.0);
                if (isnd) hn = 1; else hv = 1;
            }
        keep[b] = (uint8_t)hv;
        nodata[b] = (uint8_t)hn;
        nkeep += hv;
    }
    if (nkeep == 0) {
        for (int b = 0; b < nbox; ++b) { keep[b] = 1; nodata[b] = 0; }
        nkeep = nbox;
    }
    return nkeep;
}

/* ------------------------------------------------------------------------ */
/* Step-driven single-level loop: ComputeDt then AdvanceLevel each step.    */
/* Restates src/HydroEXA.cpp:260-271 with a step count instead of max_time, */
/* src/state/AmrMeshState_Physics.cpp:13-19,104-108, BaseSolver.H:120-152   */
/* for one level, and the dt-collapse abort of Roe.cpp:88-95.               */
/* State is advanced in place; dt_hist (optional) receives each step's dt.  */
/* ------------------------------------------------------------------------ */
int hxo_run(int solver, int nx, int ny, double *h, double *hu, double *hv,
            const double *z, const double *n, double dx, double dy, const int *bclo,
            const int *bchi, double cfl, int nsteps, double *dt_hist, double *max_vel_hist)
{
    if (solver != 0 && solver != 1) return HXO_E_BADARG;
    if (nx < 2 || ny < 2 || !check_bc(bclo, bchi)) return HXO_E_BADARG;
    const size_t nc = (size_t)nx * ny;
    double *tmp = (double *)malloc(3 * nc * sizeof(double));
    if (!tmp) return HXO_E_NOMEM;
    pad_t p = {nx, ny, nx + 2 * G};
    work_t w;
    if (work_alloc(&p, &w)) { free(tmp); return HXO_E_NOMEM; }
    int rc = HXO_OK;
    for (int s = 0; s < nsteps; ++s) {
        double dte, mv;
        hxo_level_dt(nx, ny, h, hu, hv, z, dx, dy, &dte, &mv);
        if (max_vel_hist) max_vel_hist[s] = mv;
        if (dte < 1.e-12) { rc = HXO_E_DT_COLLAPSE; break; }
        const double dt = dte * cfl;
        if (dt_hist) dt_hist[s] = dt;
        if (solver == 0)
            roe_step_ws(nx, ny, h, hu, hv, z, n, dt, dx, dy, bclo, bchi, tmp, tmp + nc, tmp + 2 * nc, &w);
        else
            hydro_step_ws(nx, ny, h, hu, hv, z, n, dt, dx, dy, bclo, bchi, tmp, tmp + nc, tmp + 2 * nc, &w);
        memcpy(h, tmp, nc * sizeof(double));
        memcpy(hu, tmp + nc, nc * sizeof(double));
        memcpy(hv, tmp + 2 * nc, nc * sizeof(double));
    }
    work_free(&w);
    free(tmp);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Conservation diagnostics of one level.                                   */
/* Restates AmrMeshState::ComputeMassDiagnostics,                           */
/* src/state/AmrMeshState_Utils.cpp:213-343: over cells whose roughness is  */
/* not NODATA (== -9999) and h > 0, mass = sum h dA, px = sum hu dA,        */
/* py = sum hv dA, KE = sum 0.5 (hu^2 + hv^2) / h dA, and min h (DBL_MAX    */
/* when no cell qualifies).  The sums run in cell order (j outer, i inner), */
/* the order of a serial CPU build.  out = {mass, px, py, KE, min_h}.       */
/* ------------------------------------------------------------------------ */
void hxo_conservation(int nx, int ny, const double *h, const double *hu, const double *hv,
                      const double *n, double dx, double dy, double *out)
{
    const double dA = dx * dy;
    double m = 0.0, px = 0.0, py = 0.0, ke = 0.0, mn = 1.7976931348623157e308;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t s = (size_t)j * nx + i;
            const double hh = h[s], qx = hu[s], qy = hv[s];
            if (!(n[s] == NODATA) && hh > 0.0) {
                m += hh * dA;
                px += qx * dA;
                py += qy * dA;
                ke += 0.5 * (qx * qx + qy * qy) / hh * dA;
                mn = smin(mn, hh);
            }
        }
    out[0] = m; out[1] = px; out[2] = py; out[3] = ke; out[4] = mn;
}
