/*
 * imogen_oracle.c -- FP64 CPU restatement of the Imogen fluid step.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; never shipped,
 * never called by the product library). Built with -ffp-contract=off so every
 * expression is evaluated exactly as written, in the reference's operation
 * order.
 *
 * Each function cites the reference code it restates. Parity status of each
 * part (what pins it) is listed in the test files under tests/.
 */
#include "imogen_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_EINVAL (-100)
#define ORC_ENOIMPL (-105)
#define ORC_ENOMEM (-106)

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------ */
/* Limiters: gpuclass/limiter_functions.h:11-119                              */
/* ------------------------------------------------------------------------ */

/* fluxLimiter_minmod, limiter_functions.h:22-27 */
static double flux_minmod(double a, double b)
{
    if (a * b < 0) return 0.0;
    if (fabs(a) > fabs(b)) return b;
    return a;
}

/* fluxLimiter_Ospre, limiter_functions.h:37-43 (used by the Xin-Jin corrector) */
static double flux_ospre(double a, double b)
{
    double r = a * b;
    if (r <= 0.0) return 0.0;
    return 1.5 * r * (a + b) / (a * a + r + b * b);
}

/* half-slope limiters, limiter_functions.h:50-117 */
static double slope_ospre(double a, double b)
{
    double r = a * b;
    if (r > 0) {
        double s = a + b;
        return .75 * r * s / (s * s - r);
    }
    return 0.0;
}

static double slope_vanleer(double a, double b)
{
    double r = a * b;
    if (r > 0.0) return r / (a + b);
    return 0;
}

/* slopeLimiter_minmod returns float in the reference (limiter_functions.h:70) */
static double slope_minmod(double a, double b)
{
    if (a * b < 0) return (double)(float)0.0;
    if (fabs(a) > fabs(b)) return (double)(float)(.5 * b);
    return (double)(float)(.5 * a);
}

/* slopeLimiter_superbee, limiter_functions.h:80-86 (first branch has no 1/2) */
static double slope_superbee(double a, double b)
{
    if (a * b < 0) return 0.0;
    if (b < a) return flux_minmod(a, 2 * b);
    return .5 * flux_minmod(2 * a, b);
}

static double slope_vanalbada(double a, double b)
{
    double r = a * b;
    if (r < 0) return 0;
    return .5 * r * (a + b) / (a * a + b * b);
}

static double slope_mc(double a, double b)
{
    if (a * b <= 0) return 0;
    double r = b / a;
    double s = .25 + .25 * r;
    s = (s < r) ? s : r;
    s = (s < 1) ? s : 1.0;
    return s * a;
}

double orc_slope_limiter(int which, double a, double b)
{
    switch (which) {
    case ORC_LIM_OSPRE: return slope_ospre(a, b);
    case ORC_LIM_VANLEER: return slope_vanleer(a, b);
    case ORC_LIM_MINMOD: return slope_minmod(a, b);
    case ORC_LIM_SUPERBEE: return slope_superbee(a, b);
    case ORC_LIM_VANALBADA: return slope_vanalbada(a, b);
    case ORC_LIM_MC: return slope_mc(a, b);
    default: return 0.0;
    }
}

/* ------------------------------------------------------------------------ */
/* HLLC sweep: gpuclass/cudaFluidStep.cu:1124-1716, Einfeldt bound 1728-1754 */
/* ------------------------------------------------------------------------ */

typedef struct {
    double gamma, gm1, govergm1;
    int lim;
} thermo_t;

/* computeEinfeldtBounds, cudaFluidStep.cu:1728-1754 */
static void einfeldt(const thermo_t *t, double rhoL, double vL, double PL, double rhoR, double vR,
                     double PR, double dvsq, double *sl, double *sr)
{
    double csql = t->gamma * PL / rhoL;
    double csqr = t->gamma * PR / rhoR;
    double Fa = sqrt(rhoR / rhoL);
    double Fb = 1.0 / (1.0 + Fa);
    double Croe = ((csql + Fa * csqr) + .5 * t->gm1 * dvsq * Fa * Fb) * Fb;
    double Vroe = (vL + Fa * vR) * Fb;
    Croe = Vroe - sqrt(Croe);
    double vl = vL - sqrt(csql);
    double vr = vR + sqrt(csqr);
    vl = (Croe < vl) ? Croe : vl;
    Croe = 2 * Vroe - Croe;
    vr = (Croe > vr) ? Croe : vr;
    *sl = vl;
    *sr = vr;
}

/* HLLC flux at one interface. L = right face of cell i, R = left face of cell
 * i+1, each (rho, P, vN, vT1, vT2). out = (mass, energy, pN, pT1, pT2) flux.
 * cudaFluidStep.cu:1331-1431 */
static void hllc_flux(const thermo_t *t, const double *L, const double *R, double *out)
{
    double C = L[2] - R[2];
    double D = L[3] - R[3];
    double E = L[4] - R[4];
    C = C * C + D * D + E * E;
    double sl, sr;
    einfeldt(t, L[0], L[2], L[1], R[0], R[2], R[1], C, &sl, &sr);
    C = sl;
    D = sr;
    double A = L[0] * (C - L[2]);
    double B = R[0] * (D - R[2]);
    E = (B * R[2] - R[1] - A * L[2] + L[1]) / (B - A);

    int mode;
    if (E > 0) mode = (C > 0) ? 0 : 2;
    else mode = (D > 0) ? 3 : 1;

    double F, G, H;
    switch (mode) {
    case 0:
    case 1: {
        const double *S = (mode == 0) ? L : R;
        A = S[0];
        B = S[2];
        C = S[3];
        D = S[4];
        E = .5 * (B * B + C * C + D * D) * A;
        F = S[1];
        A *= B;
        C *= A;
        D *= A;
        G = A * B + F;
        B *= (E + t->govergm1 * F);
        out[0] = A;
        out[1] = B;
        out[2] = G;
        out[3] = C;
        out[4] = D;
    } break;
    case 2:
        A = L[0];
        B = L[2];
        D = A * (C - B) / (C - E);
        F = L[3];
        H = L[4];
        G = .5 * (B * B + F * F + H * H);
        A = D * E;
        F *= A;
        H *= A;
        C *= (E - B);
        B = L[1] + D * (C + E * B);
        G = D * E * (t->govergm1 * L[1] / L[0] + G + C);
        out[0] = A;
        out[1] = G;
        out[2] = B;
        out[3] = F;
        out[4] = H;
        break;
    default: /* 3 */
        A = R[0];
        B = R[2];
        C = A * (D - B) / (D - E);
        F = R[3];
        H = R[4];
        G = .5 * (B * B + F * F + H * H);
        A = C * E;
        F *= A;
        H *= A;
        D *= (E - B);
        B = R[1] + C * (D + E * B);
        G = C * E * (t->govergm1 * R[1] / R[0] + G + D);
        out[0] = A;
        out[1] = G;
        out[2] = B;
        out[3] = F;
        out[4] = H;
        break;
    }
}

/* Interface fluxes fl[v][i] (at i+1/2) of one periodic line of n cells.
 * in[v] are the 5 line arrays ordered (rho, E, pN, pT1, pT2). ws must hold
 * 15*n doubles. cudaFluidStep.cu:1185-1431 (MUSCL on rho, e_int, v with
 * RECONSTRUCT_VELOCITY, gpuclass/Makefile:5). */
static void hllc_line_fluxes(const thermo_t *t, int n, double *const in[5], double *fl[5], double *ws)
{
    double *prim = ws;           /* 5*n */
    double *lf = ws + 5 * n;     /* 5*n left faces  (rho, P, vN, vT1, vT2) */
    double *rf = ws + 10 * n;    /* 5*n right faces */
    int i, v;
    for (i = 0; i < n; i++) {
        double A = in[0][i], B = in[1][i], C = in[2][i], D = in[3][i], E = in[4][i];
        double F = 1.0 / A;
        B -= .5 * (C * C + D * D + E * E) * F;
        C *= F;
        D *= F;
        E *= F;
        prim[0 * n + i] = A;
        prim[1 * n + i] = B;
        prim[2 * n + i] = C;
        prim[3 * n + i] = D;
        prim[4 * n + i] = E;
    }
    for (v = 0; v < 5; v++) {
        const double *p = prim + v * n;
        for (i = 0; i < n; i++) {
            int im = (i == 0) ? n - 1 : i - 1;
            int ip = (i == n - 1) ? 0 : i + 1;
            double dl = p[i] - p[im];
            double dr = p[ip] - p[i];
            double s = orc_slope_limiter(t->lim, dl, dr);
            if (v == 1) { /* store pressure, not internal energy (cudaFluidStep.cu:1277-1278) */
                lf[v * n + i] = t->gm1 * (p[i] - s);
                rf[v * n + i] = t->gm1 * (p[i] + s);
            } else {
                lf[v * n + i] = p[i] - s;
                rf[v * n + i] = p[i] + s;
            }
        }
    }
    for (i = 0; i < n; i++) {
        int ip = (i == n - 1) ? 0 : i + 1;
        double L[5], R[5], o[5];
        for (v = 0; v < 5; v++) {
            L[v] = rf[v * n + i];
            R[v] = lf[v * n + ip];
        }
        hllc_flux(t, L, R, o);
        for (v = 0; v < 5; v++) fl[v][i] = o[v];
    }
}

/* ------------------------------------------------------------------------ */
/* HLL sweep: gpuclass/cudaFluidStep.cu:1801-2422, pressure solver 2813-2864 */
/* ------------------------------------------------------------------------ */

typedef struct {
    double gamma, gm1, govergm1, rhomin, emin, ciso, cisosq;
    int lim;
} hllconst_t;

/* cukern_PressureSolverHydro with CFD_ENFORCE_MINVALS (cudaFluidStep.cu:2813-2864):
 * floors rho and e in place on the physical-order cell c = (rho, E, px, py, pz)
 * and returns the gas pressure. */
static double hll_floor_pressure(const hllconst_t *t, double *c)
{
    double rho = c[0], E, z, momsq, P;
    if (rho < t->rhomin) {
        P = t->rhomin / rho;
        rho = c[0] = t->rhomin;
        E = c[1] = P * c[1];
        z = c[2] = P * c[2];
        momsq = z * z;
        z = c[3] = P * c[3];
        momsq += z * z;
        z = c[4] = P * c[4];
        momsq += z * z;
    } else {
        E = c[1];
        z = c[2];
        momsq = z * z;
        z = c[3];
        momsq += z * z;
        z = c[4];
        momsq += z * z;
    }
    P = E - .5 * momsq / rho;
    if (P < rho * t->emin) {
        P = t->emin * rho;
        c[1] = (.5 * momsq / rho) + P;
    }
    return P * t->gm1;
}

/* computeEinfeldtBounds_isothermal (cudaFluidStep.cu:1779-1799) */
static void einfeldt_iso(double rhoL, double vL, double rhoR, double vR, double cs, double *sl, double *sr)
{
    double Fa = sqrt(rhoR / rhoL);
    double Fb = 1.0 / (1.0 + Fa);
    double Vroe = (vL + Fa * vR) * Fb;
    cs = Vroe - cs;
    double vl = vL - cs;
    double vr = vR + cs;
    vl = (cs < vl) ? cs : vl;
    cs = 2 * Vroe - cs;
    vr = (cs > vr) ? cs : vr;
    *sl = vl;
    *sr = vr;
}

/* HLL flux between left state L and right state R, each (rho, vN, P, vT1, vT2)
 * as the reference stores them (A..E). out = (mass, energy, pN, pT1, pT2).
 * cudaFluidStep.cu:2124-2196 */
static void hll_flux(const hllconst_t *t, const double *L, const double *R, double *out)
{
    double Ale = L[0], Ble = L[1], Cle = L[2], Dle = L[3], Ele = L[4];
    double Are = R[0], Bre = R[1], Cre = R[2], Dre = R[3], Ere = R[4];
    double Utilde = (Bre - Ble);
    double Sleft = (Dre - Dle);
    double Sright = (Ere - Ele);
    double sl, sr;
    if (t->ciso == -1) {
        thermo_t th;
        th.gamma = t->gamma;
        th.gm1 = t->gm1;
        th.govergm1 = t->govergm1;
        th.lim = t->lim;
        einfeldt(&th, Ale, Ble, Cle, Are, Bre, Cre, Utilde * Utilde + Sleft * Sleft + Sright * Sright, &sl, &sr);
    } else {
        /* the reference passes C_iso^2 where the sound speed is expected */
        einfeldt_iso(Ale, Ble, Are, Bre, t->cisosq, &sl, &sr);
    }
    Sleft = sl;
    Sright = sr;
    double Atilde = 1 / (Sright - Sleft);
    double Fa = Ale * Ble;
    double Fb = Are * Bre;
    int mode = 1; /* HLL */
    if (Sleft >= 0) mode = 0;
    if (Sright <= 0) mode = 2;
    switch (mode) {
    case 0:
        out[0] = Fa;
        if (t->ciso == -1) {
            out[1] = Ble * (t->govergm1 * Cle + .5 * Ale * (Ble * Ble + Dle * Dle + Ele * Ele));
            out[2] = Fa * Ble + Cle;
        } else {
            out[1] = Ble * (Cle + .5 * Ale * (Ble * Ble + Dle * Dle + Ele * Ele));
            out[2] = Fa * Ble + t->cisosq * Ale;
        }
        out[3] = Fa * Dle;
        out[4] = Fa * Ele;
        break;
    case 1:
        out[0] = (Sright * (Ble - Sleft) * Ale - Sleft * (Bre - Sright) * Are) * Atilde;
        if (t->ciso == -1)
            out[2] = (Sright * (Ble - Sleft) * Fa - Sleft * (Bre - Sright) * Fb + Sright * Cle - Sleft * Cre) * Atilde;
        else
            out[2] = (Sright * (Ble - Sleft) * Fa - Sleft * (Bre - Sright) * Fb +
                      t->cisosq * (Sright * Ale - Sleft * Are)) * Atilde;
        out[3] = (Sright * (Ble - Sleft) * Ale * Dle - Sleft * (Bre - Sright) * Are * Dre) * Atilde;
        out[4] = (Sright * (Ble - Sleft) * Ale * Ele - Sleft * (Bre - Sright) * Are * Ere) * Atilde;
        if (t->ciso == -1) {
            Fa = .5 * Ale * (Ble * Ble + Dle * Dle + Ele * Ele);
            Fb = .5 * Are * (Bre * Bre + Dre * Dre + Ere * Ere);
            double e = (Sright * (Ble - Sleft) * Fa - Sleft * (Bre - Sright) * Fb);
            out[1] = (e + (t->govergm1 * (Sright * Cle * Ble - Sleft * Cre * Bre) +
                           Sleft * Sright * (Cre - Cle) / t->gm1)) * Atilde;
        } else {
            /* the isothermal energy flux lacks the 1/(Sr-Sl) factor in the reference (:2185-2189) */
            Fa = .5 * Ale * (Ble * Ble + Dle * Dle + Ele * Ele) + Cle;
            Fb = .5 * Are * (Bre * Bre + Dre * Dre + Ere * Ere) + Cre;
            out[1] = (Sright * (Ble - Sleft) * Fa - Sleft * (Bre - Sright) * Fb);
        }
        break;
    default:
        out[0] = Fb;
        if (t->ciso == -1) {
            out[1] = Bre * (t->govergm1 * Cre + .5 * Are * (Bre * Bre + Dre * Dre + Ere * Ere));
            out[2] = Fb * Bre + Cre;
        } else {
            out[1] = Bre * (Cre + .5 * Are * (Bre * Bre + Dre * Dre + Ere * Ere));
            out[2] = Fb * Bre + t->ciso * Are; /* C_iso, not C_iso^2, on this side (:2191) */
        }
        out[3] = Fb * Dre;
        out[4] = Fb * Ere;
        break;
    }
}

/* Fluxes at i+1/2 of a periodic line for HLL. in[v] holds the line in HLL's
 * component order (rho, E, pN, pT1, pT2) after the floors; P its pressure. */
static void hll_line_fluxes(const hllconst_t *t, int n, double *const in[5], const double *P, double *fl[5],
                            double *ws)
{
    double *prim = ws;        /* 5n: rho, vN, P, vT1, vT2 */
    double *lf = ws + 5 * n;
    double *rf = ws + 10 * n;
    int i, v;
    for (i = 0; i < n; i++) {
        double A = in[0][i];
        prim[0 * n + i] = A;
        prim[1 * n + i] = in[2][i] / A;
        prim[2 * n + i] = P[i];
        prim[3 * n + i] = in[3][i] / A;
        prim[4 * n + i] = in[4][i] / A;
    }
    for (v = 0; v < 5; v++) {
        const double *p = prim + v * n;
        for (i = 0; i < n; i++) {
            int im = (i == 0) ? n - 1 : i - 1;
            int ip = (i == n - 1) ? 0 : i + 1;
            double s = orc_slope_limiter(t->lim, p[i] - p[im], p[ip] - p[i]);
            lf[v * n + i] = p[i] - s;
            rf[v * n + i] = p[i] + s;
        }
    }
    for (i = 0; i < n; i++) {
        int ip = (i == n - 1) ? 0 : i + 1;
        double L[5], R[5], o[5];
        for (v = 0; v < 5; v++) {
            L[v] = rf[v * n + i];
            R[v] = lf[v * n + ip];
        }
        hll_flux(t, L, R, o);
        for (v = 0; v < 5; v++) fl[v][i] = o[v];
    }
}

/* ------------------------------------------------------------------------ */
/* Xin-Jin relaxing TVD: cudaFluidStep.cu:2427-2654, freeze 2868-2941        */
/* ------------------------------------------------------------------------ */

/* Per-line freeze speed max(|p_x|/rho + sqrt(gamma P/rho)), always with the
 * physical x momentum (cudaFluidStep.cu:2887-2896). c: physical-order line. */
static double xj_freeze(double gamma, double gm1, int n, double *const c[5])
{
    double cmax = 0.0;
    int i;
    for (i = 0; i < n; i++) {
        double invrho = 1.0 / c[0][i];
        double psq = c[3][i];
        double px = c[4][i];
        psq = psq * psq + px * px;
        px = c[2][i];
        psq += px * px;
        double P = (c[1][i] - .5 * psq * invrho) * gm1;
        double loc = fabs(px) * invrho + sqrt(gamma * P * invrho);
        cmax = (loc > cmax) ? loc : cmax;
    }
    return cmax;
}

/* One Xin-Jin stage on a periodic line (physical slab order). predict:
 * out = floor(Q - lam dF(Q)); correct: out = floor(base - lam dF_tvd(Q)).
 * nslab = slab index of the normal momentum (2,3,4). ws holds >= 4n doubles. */
static void xj_stage(const hllconst_t *t, int n, int nslab, double cf, double lam, int correct,
                     double *const Q[5], double *const base[5], double *out[5], double *ws)
{
    double *wl = ws, *wr = ws + n, *dl = ws + 2 * n, *dr = ws + 3 * n;
    int i, v;
    double *pr = (double *)malloc(sizeof(double) * 5 * (size_t)n);
    for (v = 0; v < 5; v++) {
        for (i = 0; i < n; i++) {
            double rho = Q[0][i];
            double P = t->gm1 * (Q[1][i] - .5 * (Q[4][i] * Q[4][i] + Q[3][i] * Q[3][i] + Q[2][i] * Q[2][i]) / rho);
            double vx = Q[nslab][i] / rho;
            double w;
            if (v == 0) w = Q[nslab][i];
            else if (v == 1) w = vx * (Q[1][i] + P);
            else if (v == nslab) w = vx * Q[v][i] + P;
            else w = vx * Q[v][i];
            wl[i] = (cf * Q[v][i] - w);
            wr[i] = (cf * Q[v][i] + w);
        }
        if (correct) {
            for (i = 0; i < n; i++) {
                int im = (i == 0) ? n - 1 : i - 1;
                dl[i] = 0.5 * (wl[i] - wl[im]);
                dr[i] = 0.5 * (wr[i] - wr[im]);
            }
            for (i = 0; i < n; i++) {
                int ip = (i == n - 1) ? 0 : i + 1;
                wl[i] -= flux_ospre(dl[i], dl[ip]);
                wr[i] += flux_ospre(dr[i], dr[ip]);
            }
        }
        for (i = 0; i < n; i++) {
            int im = (i == 0) ? n - 1 : i - 1;
            int ip = (i == n - 1) ? 0 : i + 1;
            double b = correct ? base[v][i] : Q[v][i];
            pr[v * n + i] = b - lam * (wr[i] - wr[im] - wl[ip] + wl[i]);
        }
    }
    for (i = 0; i < n; i++) {
        double p0 = pr[i], p1 = pr[n + i], p2 = pr[2 * n + i], p3 = pr[3 * n + i], p4 = pr[4 * n + i];
        p0 = (p0 < t->rhomin) ? t->rhomin : p0;
        double w = .5 * (p2 * p2 + p3 * p3 + p4 * p4) / p0;
        if ((p1 - w) < p0 * t->emin) p1 = w + p0 * t->emin;
        out[0][i] = p0;
        out[1][i] = p1;
        out[2][i] = p2;
        out[3][i] = p3;
        out[4][i] = p4;
    }
    free(pr);
}

/* ------------------------------------------------------------------------ */
/* Line gather/scatter over the 3D array                                     */
/* ------------------------------------------------------------------------ */

/* Slab used for (normal, tangent1, tangent2) momentum per direction, as the
 * reference kernels read them (cudaFluidStep.cu:1199-1216). */
static const int mom_slab[4][3] = {{0, 0, 0}, {2, 3, 4}, {3, 2, 4}, {4, 3, 2}};

typedef struct {
    long stride;   /* element stride along the line */
    int n;         /* line length */
    long nlines;
} lineinfo_t;

static void line_geometry(const orc_grid *g, int dir, lineinfo_t *li)
{
    switch (dir) {
    case 1: li->stride = 1; li->n = g->nx; li->nlines = (long)g->ny * g->nz; break;
    case 2: li->stride = g->nx; li->n = g->ny; li->nlines = (long)g->nx * g->nz; break;
    default: li->stride = (long)g->nx * g->ny; li->n = g->nz; li->nlines = (long)g->nx * g->ny; break;
    }
}

static long line_base(const orc_grid *g, int dir, long l)
{
    long nx = g->nx, ny = g->ny;
    switch (dir) {
    case 1: return l * nx;                              /* l = y + ny*z */
    case 2: return (l % nx) + (l / nx) * nx * ny;       /* l = x + nx*z */
    default: return l;                                  /* l = x + nx*y */
    }
}

static int slab_for(int dir, int v)
{
    if (v < 2) return v;
    return mom_slab[dir][v - 2];
}

/* ------------------------------------------------------------------------ */
/* Directional update                                                        */
/* ------------------------------------------------------------------------ */

static int update_1d(double *q, const orc_grid *g, int dir, int method, double lambda,
                     const orc_fluid_params *p, int predict_only)
{
    if (dir < 1 || dir > 3) return ORC_EINVAL;
    lineinfo_t li;
    line_geometry(g, dir, &li);
    if (li.n < 3) return ORC_OK; /* cudaFluidStep.cu:301 */
    if (method != ORC_HLLC && method != ORC_HLL && method != ORC_XINJIN) return ORC_EINVAL;

    thermo_t t;
    t.gamma = p->gamma;
    t.gm1 = p->gamma - 1.0;
    t.govergm1 = p->gamma / (p->gamma - 1.0);
    t.lim = p->limiter;
    hllconst_t hc;
    hc.gamma = p->gamma;
    hc.gm1 = p->gamma - 1.0;
    hc.govergm1 = p->gamma / (p->gamma - 1.0);
    hc.rhomin = p->rhomin;
    hc.emin = 1e-5 / (p->gamma - 1); /* cudaFluidStep.cu:298 */
    hc.ciso = p->ciso;
    hc.cisosq = p->ciso * p->ciso;
    hc.lim = p->limiter;

    int n = li.n;
    int failed = 0;
    if (method != ORC_HLLC) {
        /* HLL momentum order (cudaFluidStep.cu:1869-1885): X (px,py,pz), Y (py,px,pz), Z (pz,px,py) */
        static const int hll_slab[4][3] = {{0, 0, 0}, {2, 3, 4}, {3, 2, 4}, {4, 2, 3}};
#pragma omp parallel
        {
            double *buf = (double *)malloc(sizeof(double) * (size_t)n * 48);
            if (!buf) {
#pragma omp atomic write
                failed = 1;
            }
            long l;
#pragma omp for schedule(static)
            for (l = 0; l < li.nlines; l++) {
                if (!buf) continue;
                double *C[5], *U[5], *FL[5], *H[5], *O[5];
                double *P = buf + 25 * n, *ws = buf + 26 * n;
                int v, i;
                for (v = 0; v < 5; v++) {
                    C[v] = buf + v * n;
                    U[v] = buf + 5 * n + v * n;
                    FL[v] = buf + 10 * n + v * n;
                    H[v] = buf + 15 * n + v * n;
                    O[v] = buf + 20 * n + v * n;
                }
                long base = line_base(g, dir, l);
                for (v = 0; v < 5; v++) {
                    const double *src = q + (long)v * g->pitch + base;
                    for (i = 0; i < n; i++) C[v][i] = src[(long)i * li.stride];
                }
                if (method == ORC_XINJIN) {
                    /* freeze(Q); predict 1/4 lambda; freeze(U1); correct 1/2 lambda (cudaFluidStep.cu:357-475) */
                    double cf = xj_freeze(hc.gamma, hc.gm1, n, C);
                    xj_stage(&hc, n, dir + 1, cf, 0.25 * lambda, 0, C, C, predict_only ? O : U, ws);
                    if (!predict_only) {
                        cf = xj_freeze(hc.gamma, hc.gm1, n, U);
                        xj_stage(&hc, n, dir + 1, cf, 0.5 * lambda, 1, U, C, O, ws);
                    }
                } else {
                    const int *ms = hll_slab[dir];
                    double c[5];
                    /* floors + pressure of Q (in place), HLL component order */
                    for (i = 0; i < n; i++) {
                        for (v = 0; v < 5; v++) c[v] = C[v][i];
                        P[i] = hll_floor_pressure(&hc, c);
                        for (v = 0; v < 5; v++) C[v][i] = c[v];
                    }
                    double *Hq[5] = {C[0], C[1], C[ms[0]], C[ms[1]], C[ms[2]]};
                    hll_line_fluxes(&hc, n, Hq, P, FL, ws);
                    for (v = 0; v < 5; v++)
                        for (i = 0; i < n; i++) {
                            int im = (i == 0) ? n - 1 : i - 1;
                            H[v][i] = Hq[v][i] - lambda * (FL[v][i] - FL[v][im]);
                        }
                    /* U1 back to physical order, floors + pressure of U1 */
                    for (i = 0; i < n; i++) {
                        c[0] = H[0][i];
                        c[1] = H[1][i];
                        c[ms[0]] = H[2][i];
                        c[ms[1]] = H[3][i];
                        c[ms[2]] = H[4][i];
                        P[i] = hll_floor_pressure(&hc, c);
                        for (v = 0; v < 5; v++) U[v][i] = c[v];
                    }
                    double *Hu[5] = {U[0], U[1], U[ms[0]], U[ms[1]], U[ms[2]]};
                    hll_line_fluxes(&hc, n, Hu, P, FL, ws);
                    double *Ho[5] = {O[0], O[1], O[ms[0]], O[ms[1]], O[ms[2]]};
                    for (v = 0; v < 5; v++)
                        for (i = 0; i < n; i++) {
                            int im = (i == 0) ? n - 1 : i - 1;
                            Ho[v][i] = .5 * (Hq[v][i] + Hu[v][i] - lambda * (FL[v][i] - FL[v][im]));
                        }
                }
                for (v = 0; v < 5; v++) {
                    double *dst = q + (long)v * g->pitch + base;
                    for (i = 0; i < n; i++) dst[(long)i * li.stride] = O[v][i];
                }
            }
            free(buf);
        }
        return failed ? ORC_ENOMEM : ORC_OK;
    }
#pragma omp parallel
    {
        double *buf = (double *)malloc(sizeof(double) * (size_t)n * 35);
        if (!buf) {
#pragma omp atomic write
            failed = 1;
        }
        long l;
#pragma omp for schedule(static)
        for (l = 0; l < li.nlines; l++) {
            if (!buf) continue;
            double *Q[5], *U[5], *FL[5];
            double *ws = buf + 15 * n;
            int v, i;
            for (v = 0; v < 5; v++) {
                Q[v] = buf + v * n;
                U[v] = buf + 5 * n + v * n;
                FL[v] = buf + 10 * n + v * n;
            }
            /* ws occupies [15n, 30n); the remaining 5n hold the updated line */
            double *O[5];
            for (v = 0; v < 5; v++) O[v] = buf + 30 * n + v * n;
            long base = line_base(g, dir, l);
            for (v = 0; v < 5; v++) {
                const double *src = q + (long)slab_for(dir, v) * g->pitch + base;
                for (i = 0; i < n; i++) Q[v][i] = src[(long)i * li.stride];
            }
            /* SSP-RK2 stage A: U = Q - lambda dF(Q)   (TIMESCHEME_SSPRK_A, :1510-1516) */
            hllc_line_fluxes(&t, n, Q, FL, ws);
            for (v = 0; v < 5; v++)
                for (i = 0; i < n; i++) {
                    int im = (i == 0) ? n - 1 : i - 1;
                    U[v][i] = Q[v][i] - lambda * (FL[v][i] - FL[v][im]);
                }
            /* stage B: Q = .5 (U + Q - lambda dF(U))   (TIMESCHEME_SSPRK_B, :1555-1561) */
            hllc_line_fluxes(&t, n, U, FL, ws);
            for (v = 0; v < 5; v++)
                for (i = 0; i < n; i++) {
                    int im = (i == 0) ? n - 1 : i - 1;
                    O[v][i] = .5 * (U[v][i] + Q[v][i] - lambda * (FL[v][i] - FL[v][im]));
                }
            for (v = 0; v < 5; v++) {
                double *dst = q + (long)slab_for(dir, v) * g->pitch + base;
                for (i = 0; i < n; i++) dst[(long)i * li.stride] = O[v][i];
            }
        }
        free(buf);
    }
    return failed ? ORC_ENOMEM : ORC_OK;
}

int orc_fluid_update_1d(double *q, const orc_grid *g, int dir, int method, double lambda,
                        const orc_fluid_params *p)
{
    return update_1d(q, g, dir, method, lambda, p, 0);
}

int orc_xj_predict_1d(double *q, const orc_grid *g, int dir, double lambda, const orc_fluid_params *p)
{
    return update_1d(q, g, dir, ORC_XINJIN, lambda, p, 1);
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions: gpuclass/cudaStatics.cu:112-286, 1086-1532           */
/* ------------------------------------------------------------------------ */

/* Applies one array's BC on one face. memdir = 1,2,3; sas: 0 symmetric,
 * 1 antisymmetric, 2 constant, 3 linear (x only). Mirrors callBCKernel's
 * kernel table (cudaStatics.cu:1169-1240) including the per-direction index
 * conventions of each kernel. */
static int array_face_bc(double *phi, const orc_grid *g, int side, int memdir, int sas, int d)
{
    long nx = g->nx, ny = g->ny, nz = g->nz;
    long i, j, k;
    int q;
    long n = (memdir == 1) ? nx : (memdir == 2 ? ny : nz);
    if ((sas == 0 || sas == 1) && n < 2 * d + 2) return ORC_EINVAL;
    if (n < d + 2) return ORC_EINVAL;
    /* cukern_extrapolateLinearBdyXPlus (cudaStatics.cu:1406-1415) writes one
     * cell past the end of each row and is marked "not done" in the reference
     * (:1180); linear extrapolation exists only on the -x face here. */
    if (sas == 3 && (memdir != 1 || side == 1)) return ORC_ENOIMPL;
    if (memdir == 1) {
        for (k = 0; k < nz; k++)
            for (j = 0; j < ny; j++) {
                double *r = phi + j * nx + k * nx * ny;
                if (side == 0) {
                    if (sas == 0) for (q = 0; q < d; q++) r[d - 1 - q] = r[d + 1 + q];
                    else if (sas == 1) for (q = 0; q < d; q++) r[d - 1 - q] = -r[d + 1 + q];
                    else if (sas == 2) for (q = 0; q < d; q++) r[q] = r[d];
                    else {
                        /* cukern_extrapolateLinearBdyXMinus (cudaStatics.cu:1372-1380) */
                        double f0 = r[d], f1 = r[d + 1];
                        for (q = 0; q < d; q++) r[q] = r[d] + (d - q) * (f0 - f1);
                    }
                } else {
                    if (sas == 0) for (q = 0; q < d; q++) r[nx - d + q] = r[nx - d - 2 - q];
                    else if (sas == 1) for (q = 0; q < d; q++) r[nx - d + q] = -r[nx - d - 2 - q];
                    else for (q = 0; q < d; q++) r[nx - d + q] = r[nx - d - 1];
                }
            }
        return ORC_OK;
    }
    if (sas == 3) return ORC_ENOIMPL; /* callBCKernel cases 19/27/35/43 */
    if (memdir == 2) {
        for (k = 0; k < nz; k++)
            for (i = 0; i < nx; i++) {
                double *c = phi + i + k * nx * ny;
                if (side == 0) {
                    if (sas == 0) for (q = 0; q < d; q++) c[nx * q] = c[nx * (2 * d - q)];
                    else if (sas == 1) for (q = 0; q < d; q++) c[nx * q] = -c[nx * (2 * d - q)];
                    else { double f = c[d * nx]; for (q = 0; q < d; q++) c[q * nx] = f; }
                } else {
                    if (sas == 0) for (q = 0; q < d; q++) c[nx * (ny - 1 - q)] = c[nx * (ny - 1 - 2 * d + q)];
                    else if (sas == 1) for (q = 0; q < d; q++) c[nx * (ny - 1 - q)] = -c[nx * (ny - 1 - 2 * d + q)];
                    else { /* copies the innermost ghost, not the last interior cell (cudaStatics.cu:1461-1466) */
                        double f = c[(ny - d) * nx];
                        for (q = 0; q < d; q++) c[(ny - d + q) * nx] = f;
                    }
                }
            }
        return ORC_OK;
    }
    /* memdir 3 */
    {
        long st = nx * ny;
        for (j = 0; j < ny; j++)
            for (i = 0; i < nx; i++) {
                double *c = phi + i + nx * j;
                if (side == 0) {
                    /* z-minus mirror reflects about 4.5 (2*depth+MIRROR_OFFSET-q, :1496-1500) */
                    if (sas == 0) for (q = 0; q < d; q++) c[q * st] = c[(2 * d + 1 - q) * st];
                    else if (sas == 1) for (q = 0; q < d; q++) c[q * st] = -c[(2 * d + 1 - q) * st];
                    else for (q = 0; q < d; q++) c[st * q] = c[st * d];
                } else {
                    if (sas == 0) for (q = 0; q < d; q++) c[st * (nz - 1 - q)] = c[st * (nz - 1 - 2 * d + q)];
                    else if (sas == 1) for (q = 0; q < d; q++) c[st * (nz - 1 - q)] = -c[st * (nz - 1 - 2 * d + q)];
                    else for (q = 0; q < d; q++) c[st * (nz - 1 - q)] = c[st * (nz - d)];
                }
            }
    }
    return ORC_OK;
}

/* Outflow face (BCModeTypes outflow, cudaCommon.h:46: scalars and transverse
 * properties constant, normal momentum constant when outward, zero when
 * inward), with the per-slab rules of cukern_SetOutflowX's in-bounds +x branch
 * (cudaStatics.cu:486-531) on every face and both sides: the normal momentum
 * is read at the interior cell next to the face (cell depth or N-1-depth
 * along dir). Outward (sign of the face times that momentum > 0): all five
 * slabs of the ghosts copy that cell. Otherwise the ghosts get its density,
 * slab 2 (the reference's radial momentum) zero, slab 3 rho * v with v the
 * rest-frame speed on x faces (1.05 on -x, -1.198 - 0.0047 i for ghost i of
 * +x counted outward, cudaStatics.cu:382-385, 471-474) and the cell's own
 * velocity elsewhere, slab 4 zero, and the cell's total energy changed by
 * the kinetic energy of each replaced momentum, in that order (:510-518).
 * Ghost cells are 0..depth-1 and N-depth..N-1 of each line along dir. */
int orc_outflow_face(double *q, const orc_grid *g, int dir, int side, int depth)
{
    const long n3[3] = {g->nx, g->ny, g->nz};
    const long s3[3] = {1, (long)g->nx, (long)g->nx * g->ny};
    const int a1 = (dir == 1) ? 1 : 0, a2 = (dir == 3) ? 1 : 2;
    if (dir < 1 || dir > 3 || side < 0 || side > 1) return ORC_EINVAL;
    const long N = n3[dir - 1], st = s3[dir - 1], P = g->pitch;
    const int d = depth;
    if (N < d + 2) return ORC_EINVAL;
    const long c = side ? N - 1 - d : d; /* interior cell next to the face */
    const double dirsign = side ? 1.0 : -1.0;
    long i, j;
    int k;
    for (j = 0; j < n3[a2]; j++)
        for (i = 0; i < n3[a1]; i++) {
            double *r = q + i * s3[a1] + j * s3[a2];
            const double a = r[(1 + dir) * P + c * st];
            for (k = 0; k < d; k++) {
                const long gi = side ? N - d + k : k; /* ghost cell */
                if (a * dirsign > 0) {
                    int sl;
                    for (sl = 0; sl < 5; sl++) r[sl * P + gi * st] = r[sl * P + c * st];
                } else {
                    const double rho = r[c * st];
                    double E = r[P + c * st], p1, p2, v;
                    r[gi * st] = rho;
                    p1 = r[2 * P + c * st];
                    p2 = rho * 0.0;
                    r[2 * P + gi * st] = p2;
                    E += .5 * (p2 - p1) * (p2 + p1) / rho;
                    p1 = r[3 * P + c * st];
                    if (dir == 1) v = side ? -1.198 - .0047 * k : 1.05;
                    else v = p1 / rho;
                    p2 = rho * v;
                    r[3 * P + gi * st] = p2;
                    E += .5 * (p2 - p1) * (p2 + p1) / rho;
                    p1 = r[4 * P + c * st];
                    p2 = rho * 0.0;
                    r[4 * P + gi * st] = p2;
                    E += .5 * (p2 - p1) * (p2 + p1) / rho;
                    r[P + gi * st] = E;
                }
            }
        }
    return ORC_OK;
}

/* Freebalance face (BCModeTypes freebalance, cudaCommon.h:48; setFreeBalanceCondition and
 * cukern_SetFreeBalance{X,Y,Z}, gpuclass/cudaStatics.cu:700-1003). The interior cell next to
 * the face (cell depth, or N-1-depth, along dir) gives rho, v = p/rho, the internal energy
 * density e = E - rho|v|^2/2 and T = e/rho. Each ghost holds the same velocity and T with the
 * density of isothermal balance in the potential: rho_g = rho exp(-(phi_g - phi_c) / ((gamma-1) T))
 * (:916-922). For dir 1 (radial in the disk setups) the azimuthal velocity vy, taken to the
 * inertial frame (vy + Omega r_c), is extrapolated as a Kepler curve and its centripetal
 * potential enters the balance (:924-937): with q = r_g / r_c, the potential (q-1)/q w^2 and the
 * ghost's frame velocity w q^-1/2 - Omega r_c q (hostRadQ, :718-745; r = Rinner + i h0, the
 * reference cell at i = depth, or nx0 - 1 - depth on the plus face). A cell with rho < 1e-7
 * (the solar-wind side of those setups, :908-910) keeps phi_g = phi_c and loses a z velocity
 * that points into the domain. The ghost energy is rho_g T + rho_g |v|^2 / 2 (:945). */
int orc_freebalance_face(double *q, const orc_grid *g, const double *phi, int dir, int side, int depth,
                         double gamma, double Rinner, double h0, double frameOmega, int nx0)
{
    const long n3[3] = {g->nx, g->ny, g->nz};
    const long s3[3] = {1, (long)g->nx, (long)g->nx * g->ny};
    const int a1 = (dir == 1) ? 1 : 0, a2 = (dir == 3) ? 1 : 2;
    if (dir < 1 || dir > 3 || side < 0 || side > 1 || depth < 1 || depth > 8) return ORC_EINVAL;
    const long N = n3[dir - 1], st = s3[dir - 1], P = g->pitch;
    const int d = depth;
    if (N < d + 2) return ORC_EINVAL;
    /* radial constants of this face (hostRadQ) and the frame velocity at the reference cell */
    double x0[8], qh[8], qq[8];
    const double rc = side ? Rinner + (nx0 - 1 - d) * h0 : Rinner + d * h0;
    const double vFrame = frameOmega * rc;
    int k;
    for (k = 0; k < d; k++) {
        const double r = side ? Rinner + (nx0 - d + k) * h0 : Rinner + k * h0;
        const double qr = r / rc;
        x0[k] = (qr - 1) / qr;
        qh[k] = pow(qr, -0.5);
        qq[k] = qr;
    }
    const double invgm1 = 1 / (gamma - 1);
    const long c = side ? N - 1 - d : d;
    const double direct = side ? 1.0 : -1.0;
    long i, j;
    for (j = 0; j < n3[a2]; j++)
        for (i = 0; i < n3[a1]; i++) {
            double *r = q + i * s3[a1] + j * s3[a2];
            const double *ph = phi + i * s3[a1] + j * s3[a2];
            const double rho = r[c * st];
            const double vx = r[2 * P + c * st] / rho;
            const double vy = r[3 * P + c * st] / rho;
            double vz = r[4 * P + c * st] / rho;
            const double E = r[P + c * st] - .5 * rho * (vx * vx + vy * vy + vz * vz);
            const double T = E / rho;
            const double phi0 = ph[c * st];
            for (k = 0; k < d; k++) {
                const long gi = (side ? N - d + k : k) * st;
                double phi1 = ph[gi], rho2, th;
                if (rho < 1e-7) {
                    phi1 = phi0;
                    vz = (vz * direct < 0) ? 0 : vz;
                }
                if (dir != 1) {
                    rho2 = rho * exp(-invgm1 * (phi1 - phi0) / T);
                    th = vy;
                } else {
                    double cp;
                    th = vy + vFrame;
                    cp = th * th * x0[k];
                    th = th * qh[k] - vFrame * qq[k];
                    rho2 = rho * exp(-invgm1 * (phi1 - phi0 - cp) / T);
                }
                r[gi] = rho2;
                r[2 * P + gi] = rho2 * vx;
                r[3 * P + gi] = rho2 * th;
                r[4 * P + gi] = rho2 * vz;
                r[P + gi] = rho2 * T + .5 * rho2 * (vx * vx + th * th + vz * vz);
            }
        }
    return ORC_OK;
}

/* Static cells of one array (setArrayStaticCells and cukern_applySpecial_fade,
 * gpuclass/cudaStatics.cu:288-350, 1251-1281): entry s of the table is the
 * cell index table[s], the value f0 = table[s + blk] and the coefficient
 * c = table[s + 2 blk]; c = 1 sets the cell to f0, otherwise it fades to
 * f0 c + (1 - c) phi (f0 c where that is NaN). */
void orc_static_cells(double *phi, const double *table, int nspec, long blk)
{
    int s;
    for (s = 0; s < nspec; s++) {
        const long x = (long)table[s];
        const double f0 = table[s + blk], c = table[s + 2 * blk];
        if (c != 1.0) {
            double w = f0 * c + (1.0 - c) * phi[x];
            if (isnan(w)) w = f0 * c;
            phi[x] = w;
        } else {
            phi[x] = f0;
        }
    }
}

int orc_set_fluid_boundary(double *q, const orc_grid *g, int dir, const int *modes, int depth)
{
    if (dir < 1 || dir > 3) return ORC_EINVAL;
    int side, s;
    for (side = 0; side < 2; side++) {
        int bt = modes[2 * (dir - 1) + side];
        if (bt == ORC_BC_FREEBALANCE || bt == ORC_BC_WALL) return ORC_ENOIMPL;
        if (bt == ORC_BC_CIRCULAR || bt == ORC_BC_STATIONARY) continue;
        if (bt == ORC_BC_OUTFLOW) {
            int rc = orc_outflow_face(q, g, dir, side, depth);
            if (rc) return rc;
            continue;
        }
        for (s = 0; s < 5; s++) {
            int sas;
            int veccomp = (s < 2) ? 0 : s - 1; /* px,py,pz carry vectorComponent 1,2,3 */
            if (bt == ORC_BC_MIRROR) sas = (veccomp == dir);
            else if (bt == ORC_BC_CONST) sas = 2;
            else sas = 3;
            int rc = array_face_bc(q + (long)s * g->pitch, g, side, dir, sas, depth);
            if (rc) return rc;
        }
    }
    return ORC_OK;
}

/* The boundaries the fluid step sets after sweeping direction dir. flux.cu
 * rotates the arrays so that the swept direction is memory x
 * (cudaArrayRotateB before each sweep, the preperm / permcall tables of
 * flux.cu:46-48) and then calls setFluidBoundary (flux.cu:99-107), which picks
 * its kernels by memory axis (cudaStatics.cu:230): the memory-x kernels'
 * conventions (cudaStatics.cu:1308-1380) apply along dir, whatever dir is. */
static int array_face_bc_swept(double *phi, const orc_grid *g, int side, int dir, int sas, int d)
{
    const long n3[3] = {g->nx, g->ny, g->nz};
    const long s3[3] = {1, (long)g->nx, (long)g->nx * g->ny};
    const int a1 = (dir == 1) ? 1 : 0, a2 = (dir == 3) ? 1 : 2; /* the two other axes */
    const long N = n3[dir - 1], st = s3[dir - 1];
    long i, j;
    int q;
    if ((sas == 0 || sas == 1) && N < 2 * d + 2) return ORC_EINVAL;
    if (N < d + 2) return ORC_EINVAL;
    if (sas == 3 && side == 1) return ORC_ENOIMPL;
    for (j = 0; j < n3[a2]; j++)
        for (i = 0; i < n3[a1]; i++) {
            double *r = phi + i * s3[a1] + j * s3[a2];
            if (side == 0) {
                if (sas == 0) for (q = 0; q < d; q++) r[(d - 1 - q) * st] = r[(d + 1 + q) * st];
                else if (sas == 1) for (q = 0; q < d; q++) r[(d - 1 - q) * st] = -r[(d + 1 + q) * st];
                else if (sas == 2) for (q = 0; q < d; q++) r[q * st] = r[d * st];
                else {
                    double f0 = r[d * st], f1 = r[(d + 1) * st];
                    for (q = 0; q < d; q++) r[q * st] = f0 + (d - q) * (f0 - f1);
                }
            } else {
                if (sas == 0) for (q = 0; q < d; q++) r[(N - d + q) * st] = r[(N - d - 2 - q) * st];
                else if (sas == 1) for (q = 0; q < d; q++) r[(N - d + q) * st] = -r[(N - d - 2 - q) * st];
                else for (q = 0; q < d; q++) r[(N - d + q) * st] = r[(N - d - 1) * st];
            }
        }
    return ORC_OK;
}

int orc_fluid_boundary_swept(double *q, const orc_grid *g, int dir, const int *modes, int depth)
{
    if (dir < 1 || dir > 3) return ORC_EINVAL;
    int side, s;
    for (side = 0; side < 2; side++) {
        int bt = modes[2 * (dir - 1) + side];
        if (bt == ORC_BC_FREEBALANCE || bt == ORC_BC_WALL) return ORC_ENOIMPL;
        if (bt == ORC_BC_CIRCULAR || bt == ORC_BC_STATIONARY) continue;
        if (bt == ORC_BC_OUTFLOW) {
            int rc = orc_outflow_face(q, g, dir, side, depth);
            if (rc) return rc;
            continue;
        }
        for (s = 0; s < 5; s++) {
            int sas;
            int veccomp = (s < 2) ? 0 : s - 1;
            if (bt == ORC_BC_MIRROR) sas = (veccomp == dir);
            else if (bt == ORC_BC_CONST) sas = 2;
            else sas = 3;
            int rc = array_face_bc_swept(q + (long)s * g->pitch, g, side, dir, sas, depth);
            if (rc) return rc;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* fluidstep: gpuclass/flux.cu:27-155                                        */
/* ------------------------------------------------------------------------ */

/* fluxcall table, flux.cu:46 */
static const int fluxcall[3][6] = {{1, 2, 1, 2, 3, 3}, {3, 1, 2, 3, 1, 2}, {2, 3, 3, 1, 2, 1}};

void orc_sweep_order(int sweep, int *dirs3)
{
    int n;
    for (n = 0; n < 3; n++) dirs3[n] = fluxcall[n][((sweep % 6) + 6) % 6];
}

int orc_fluidstep(double *q, const orc_grid *g, const double *h, double dt, int stepDirection,
                  int stepNumber, int method, const orc_fluid_params *p, const int *modes, int depth)
{
    int rc;
    if (g->ny == 1 && g->nz == 1) {
        rc = orc_fluid_update_1d(q, g, 1, method, dt / h[0], p);
        if (rc) return rc;
        return orc_fluid_boundary_swept(q, g, 1, modes, depth);
    }
    int sweep = (stepNumber + 3 * (stepDirection > 0)) % 6;
    int dims[3] = {g->nx, g->ny, g->nz};
    int n;
    for (n = 0; n < 3; n++) {
        int d = fluxcall[n][sweep];
        if (dims[d - 1] > 3) {
            rc = orc_fluid_update_1d(q, g, d, method, dt / h[d - 1], p);
            if (rc) return rc;
            rc = orc_fluid_boundary_swept(q, g, d, modes, depth);
            if (rc) return rc;
        }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Sound speed and CFL                                                       */
/* ------------------------------------------------------------------------ */

void orc_soundspeed(const double *q, const orc_grid *g, double gamma, double *cs)
{
    long n = (long)g->nx * g->ny * g->nz, x;
    double gg1 = gamma * (gamma - 1);
    const double *rho = q, *E = q + g->pitch, *px = q + 2 * g->pitch, *py = q + 3 * g->pitch,
                 *pz = q + 4 * g->pitch;
#pragma omp parallel for schedule(static)
    for (x = 0; x < n; x++) {
        double rhoinv = 1 / rho[x];
        double csq = gg1 * (E[x] - .5 * (px[x] * px[x] + py[x] * py[x] + pz[x] * pz[x]) * rhoinv) * rhoinv;
        if (csq < 0.0) csq = 0.0;
        cs[x] = sqrt(csq);
    }
}

double orc_cfl_timestep(const double *q, const double *cs, const orc_grid *g, const double *h,
                        int method, const int *globalRez)
{
    long n = (long)g->nx * g->ny * g->nz, x;
    int spacedim = 0;
    if (globalRez[1] > 1) spacedim = 1;
    if (globalRez[2] > 1) spacedim = 2;
    int dim = spacedim + 1;
    const double *rho = q, *px = q + 2 * g->pitch, *py = q + 3 * g->pitch, *pz = q + 4 * g->pitch;
    double tmin = 1e38;
    /* the XJ 1D table entry is an HLLC-formula instance (cflTimestep.cu:183); same value */
    int xj = (method == ORC_XINJIN) && dim > 1;
    double hmin = h[0];
    if (dim > 1 && h[1] < hmin) hmin = h[1];
    if (dim > 2 && h[2] < hmin) hmin = h[2];
#pragma omp parallel for reduction(min : tmin) schedule(static)
    for (x = 0; x < n; x++) {
        double u;
        if (!xj) {
            double pm;
            if (dim == 1) pm = fabs(px[x]);
            else {
                double pp = px[x] * px[x];
                pp = pp + py[x] * py[x];
                if (dim > 2) pp = pp + pz[x] * pz[x];
                pm = sqrt(pp);
            }
            u = pm / rho[x] + cs[x];
            u = hmin / u;
        } else {
            double r = rho[x], v;
            u = h[0] / (fabs(px[x]) / r + cs[x]);
            v = h[1] / (fabs(py[x]) / r + cs[x]);
            u = (u < v) ? u : v;
            if (dim > 2) {
                v = h[2] / (fabs(pz[x]) / r + cs[x]);
                u = (u < v) ? u : v;
            }
        }
        if (u < tmin) tmin = u;
    }
    if (tmin > 1e37) tmin = 1e37; /* per-block initial value, cflTimestep.cu:290 */
    if (method == ORC_HLL || method == ORC_HLLC) tmin = tmin / 2;
    return tmin;
}
