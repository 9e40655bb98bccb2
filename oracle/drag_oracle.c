/*
 * drag_oracle.c -- FP64 CPU restatement of the gas-dust drag source.
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP path; never shipped).
 * Built with -ffp-contract=off, so expressions evaluate as written.
 *
 * Restates sourcefunction_2FluidDrag's constants (gpuclass/cudaSource2FluidDrag.cu:189-243),
 * the drag-rate evaluation cukern_GasDustDrag_GeneralLinearCore + allregimeCdrag
 * (:819-890), the initial-state preparation cukern_prepareForERK_h0 (:1874-1908),
 * the LogTrap2/3 integrator cukern_LogTrapSolve (:1532-1686), and the explicit
 * midpoint and classical RK4 integrators of |dv| (solveDragEMP :305-370,
 * solveDragRK4 :372-467: cukern_findInitialDeltaV :1014-1056,
 * cukern_GasDustDrag_GeneralAccel :898-946, cukern_SolveRK_single/final
 * :1062-1100, cukern_applyFinalDeltaV :1160-1237), and exponential Euler
 * (solveDragETDRK1 :477-577 with ACCOUNT_GRADP undefined (:63): the drag
 * rate frozen at t = 0, cukern_ETDRK1 :1426-1530).
 *
 * The "exact" integrator (method 6) has no counterpart in the reference's
 * compiled code; it integrates the same scalar relaxation problem that the
 * reference's analysis script solves with ode113 (experiment/Advection/
 * DustyBoxAnalyzer.m:91-104, 135-140): d|dv|/dt = -k(|dv|, u(|dv|)) |dv| with
 * u(|dv|) = u_0 + rho_d (|dv_0|^2 - |dv|^2) / (2 (rho_g + rho_d)), here as
 * y = ln|dv| with the Dormand-Prince 5(4) pair (J. Comput. Appl. Math. 6
 * (1980) 19) and an absolute tolerance on y (a relative tolerance on |dv|).
 */
#include "imogen_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>

/* hostDrag[0..8] of sourcefunction_2FluidDrag (:210-220) */
void orc_drag_params(const orc_thermo *tg, const orc_thermo *td, double *P)
{
    const double gam = tg->gamma;
    const double ddust = sqrt(td->sigma0 / 3.141592653589793);
    P[0] = tg->mu0;
    P[1] = -tg->muTindex;
    P[2] = tg->sigmaTindex;
    P[3] = tg->m * (gam - 1.0) / (298.15 * tg->kBolt);
    P[4] = sqrt(2.0) * tg->m / (tg->sigma0 * ddust);
    P[5] = ddust / tg->mu0;
    P[6] = td->sigma0 / (1.0 * td->m);
    P[7] = 0.0;
    P[8] = gam - 1.0;
}

/* allregimeCdrag (:819-833) */
static double cdrag(double Re, double Kn)
{
    double cunningham = 1 + Kn * (1.142 + 1 * 0.558 * exp(-0.999 / Kn));
    return (3 / Re + .45 * pow(Re, -.319) + .0509 * Re / (8710 + Re)) / cunningham;
}

/* cukern_GasDustDrag_GeneralLinearCore (:836-890): drag rate k at |dv| = magdv
 * and specific internal energy u; zero below |dv| = 1e-9. */
double orc_drag_k(const double *P, double rhog, double rhod, double magdv, double u)
{
    if (magdv < 1e-9) return 0.0;
    double Tn = P[3] * u;
    double Re = P[5] * rhog * magdv * pow(Tn, P[1]);
    double Kn = P[4] * pow(Tn, P[2]) / rhog;
    double Cd = cdrag(Re, Kn);
    return Cd * magdv * (rhog + rhod) * P[6];
}

/* cukern_prepareForERK_h0 (:1883-1908): |dv| and u_0 = (E - |p|^2/2rho)/rho */
static void prepare(const double *g, const double *d, long pitch, double *magdv, double *u0)
{
    double dv, dvsq;
    dv = g[2 * pitch] / g[0] - d[2 * pitch] / d[0];
    dvsq = dv * dv;
    dv = g[3 * pitch] / g[0] - d[3 * pitch] / d[0];
    dvsq += dv * dv;
    dv = g[4 * pitch] / g[0] - d[4 * pitch] / d[0];
    dvsq += dv * dv;
    double press = g[pitch] - .5 * (g[2 * pitch] * g[2 * pitch] + g[3 * pitch] * g[3 * pitch] +
                                   g[4 * pitch] * g[4 * pitch]) / g[0];
    *magdv = sqrt(dvsq);
    *u0 = press / g[0];
}

/* cukern_LogTrapSolve<order> (:1532-1630): the path-averaged drag rate */
static double logtrap_rate(const double *P, double rhog, double rhod, double dv0, double u0, double t, int order)
{
    double mu = rhod / (rhog + rhod);
    double tmp0 = dv0, tmp1 = u0;
    double ksum = orc_drag_k(P, rhog, rhod, tmp0, tmp1);
    double k = ksum;
    double dv_i = tmp0;
    double dv_t = dv_i * exp(-t * k);
    double dvsq = dv_t * dv_t;
    double duint = -.5 * dv_i * dv_i * mu * expm1(-2 * k * t);
    tmp0 = sqrt(dvsq);
    tmp1 += duint;
    /* accumulate k_1; the core skips the add below |dv| = 1e-9 (:853-859) */
    if (tmp0 >= 1e-9) ksum += orc_drag_k(P, rhog, rhod, tmp0, tmp1);
    if (order == 3) {
        tmp1 -= duint;
        k = .25 * ksum;
        dv_t = dv_i * exp(-t * k);
        dvsq = dv_t * dv_t;
        duint = -0.5 * dv_i * dv_i * mu * expm1(-2 * k * t);
        tmp0 = sqrt(dvsq);
        tmp1 += duint;
        double khalf = orc_drag_k(P, rhog, rhod, tmp0, tmp1);
        k = (0.16666666666666666667 * ksum + 0.66666666666666666667 * khalf);
    } else {
        k = .5 * ksum;
    }
    return k;
}

/* dy/dt with y = ln|dv| */
static double exact_rhs(const double *P, double rhog, double rhod, double s0sq, double u0, double c, double y)
{
    double s = exp(y);
    double u = u0 + c * (s0sq - s * s);
    return -orc_drag_k(P, rhog, rhod, s, u);
}

/* |dv(t)| / |dv(0)| by adaptive Dormand-Prince 5(4) on y = ln|dv|. Returns
 * the ratio; *steps receives the accepted + rejected step count (negative
 * when the step budget ran out). A cell the relaxation problem is not defined
 * on -- a density that is not positive and finite, a non-finite |dv|, or a
 * moving cell whose gas internal energy is not positive and finite, or whose
 * drag rate is not finite -- returns NaN at once with *steps =
 * ORC_DRAG_INVALID. A trial step whose error estimate is not finite (an
 * overflowing stage of a too-long step) is rejected and shrunk. */
double orc_drag_exact_ratio(const double *P, double rhog, double rhod, double s0, double u0, double t, double tol,
                            int *steps)
{
    int n = 0;
    if (steps) *steps = 0;
    if (!(rhog > 0 && rhog <= DBL_MAX) || !(rhod > 0 && rhod <= DBL_MAX) || !(s0 <= DBL_MAX)) {
        if (steps) *steps = ORC_DRAG_INVALID;
        return NAN;
    }
    if (s0 < 1e-9 || t <= 0) return 1.0;
    if (!(u0 > 0 && u0 <= DBL_MAX)) {
        if (steps) *steps = ORC_DRAG_INVALID;
        return NAN;
    }
    const double c = .5 * rhod / (rhog + rhod);
    const double s0sq = s0 * s0;
    const double y0 = log(s0);
    double y = y0, tt = 0.0, h = t;
    double k1 = exact_rhs(P, rhog, rhod, s0sq, u0, c, y);
    if (!(fabs(k1) <= DBL_MAX)) {
        if (steps) *steps = ORC_DRAG_INVALID;
        return NAN;
    }
    while (tt < t) {
        if (n >= 4096) {
            if (steps) *steps = -n;
            break;
        }
        if (h > t - tt) h = t - tt;
        double k2 = exact_rhs(P, rhog, rhod, s0sq, u0, c, y + h * (1.0 / 5) * k1);
        double k3 = exact_rhs(P, rhog, rhod, s0sq, u0, c, y + h * ((3.0 / 40) * k1 + (9.0 / 40) * k2));
        double k4 = exact_rhs(P, rhog, rhod, s0sq, u0, c,
                              y + h * ((44.0 / 45) * k1 - (56.0 / 15) * k2 + (32.0 / 9) * k3));
        double k5 = exact_rhs(P, rhog, rhod, s0sq, u0, c,
                              y + h * ((19372.0 / 6561) * k1 - (25360.0 / 2187) * k2 + (64448.0 / 6561) * k3 -
                                       (212.0 / 729) * k4));
        double k6 = exact_rhs(P, rhog, rhod, s0sq, u0, c,
                              y + h * ((9017.0 / 3168) * k1 - (355.0 / 33) * k2 + (46732.0 / 5247) * k3 +
                                       (49.0 / 176) * k4 - (5103.0 / 18656) * k5));
        double yn = y + h * ((35.0 / 384) * k1 + (500.0 / 1113) * k3 + (125.0 / 192) * k4 -
                             (2187.0 / 6784) * k5 + (11.0 / 84) * k6);
        double k7 = exact_rhs(P, rhog, rhod, s0sq, u0, c, yn);
        double err = h * ((71.0 / 57600) * k1 - (71.0 / 16695) * k3 + (71.0 / 1920) * k4 -
                          (17253.0 / 339200) * k5 + (22.0 / 525) * k6 - (1.0 / 40) * k7);
        err = fabs(err) / tol;
        n++;
        if (!(err <= DBL_MAX)) err = 1e300; /* reject, shrink by the minimum factor */
        if (err <= 1.0) {
            if (!(fabs(k7) <= DBL_MAX)) {
                if (steps) *steps = ORC_DRAG_INVALID;
                return NAN;
            }
            tt += h;
            y = yn;
            k1 = k7;
        }
        double f = (err > 0) ? 0.9 * pow(err, -0.2) : 5.0;
        f = (f < 0.2) ? 0.2 : ((f > 5.0) ? 5.0 : f);
        h *= f;
    }
    if (steps && *steps == 0) *steps = n;
    return exp(y - y0);
}

/* New momenta about the barycentric velocity and the gas energy change, given
 * the decay factor of every dv component (cukern_LogTrapSolve :1632-1676). */
static void apply_factor(double *g, double *d, long pitch, double f)
{
    double mu = d[0] / (g[0] + d[0]) * g[0];
    double rhoginv = 1.0 / g[0];
    double pdustsq = 0.0;
    for (int c = 2; c <= 4; c++) {
        pdustsq -= d[c * pitch] * d[c * pitch];
        double vstick = (g[c * pitch] + d[c * pitch]) / (g[0] + d[0]);
        double dv_i = (g[c * pitch]) * rhoginv - d[c * pitch] / d[0];
        double dv_t = dv_i * f;
        g[c * pitch] = g[0] * vstick + dv_t * mu;
        double q = d[0] * vstick - dv_t * mu;
        d[c * pitch] = q;
        pdustsq += q * q;
    }
    g[pitch] -= .5 * pdustsq / d[0];
}

/* cukern_GasDustDrag_GeneralAccel (:898-946): d|dv|/dt at |dv| = magdv and
 * specific internal energy u (no small-|dv| guard, unlike the linear core). */
static double rk_accel(const double *P, double rhog, double rhod, double magdv, double u)
{
    double Tn = P[3] * u;
    double Re = P[5] * rhog * magdv * pow(Tn, P[1]);
    double Kn = P[4] * pow(Tn, P[2]) / rhog;
    double Cd = cdrag(Re, Kn);
    return -(Cd * magdv * magdv * (rhog + rhod) * P[6]);
}

/* the gas heating of a stage at |dv| = s (GeneralAccel<false>, :921-924) */
static double rk_u(double u0, double s0, double s, double rhog, double rhod)
{
    return u0 + .5 * rhod * (s0 * s0 - s * s) / (rhog + rhod);
}

/* cukern_findInitialDeltaV (:1014-1056) */
static void rk_initial(const double *g, const double *d, long pitch, double *magdv, double *u0)
{
    double rhoginv = 1 / g[0];
    double rhodinv = 1 / d[0];
    double q = g[2 * pitch];
    double u = q * rhoginv - d[2 * pitch] * rhodinv;
    double momsq = q * q;
    double dvsq = u * u;
    q = g[3 * pitch];
    u = q * rhoginv - d[3 * pitch] * rhodinv;
    momsq += q * q;
    dvsq += u * u;
    q = g[4 * pitch];
    u = q * rhoginv - d[4 * pitch] * rhodinv;
    momsq += q * q;
    dvsq += u * u;
    *magdv = sqrt(dvsq);
    *u0 = (g[pitch] - .5 * momsq * rhoginv) * rhoginv;
}

/* cukern_applyFinalDeltaV (:1160-1237): rescale dv to |dv| = s1, rebuild the
 * momenta about the barycentric velocity, conserve gas E + dust KE, and reset
 * the dust energy to its kinetic energy plus 1e-4 rho_d (:1230-1233). */
static void rk_apply(double *g, double *d, long pitch, double s1)
{
    double rhog = g[0], rhod = d[0];
    double vstick[3], dvhat[3];
    for (int c = 0; c < 3; c++) {
        double p1 = g[(c + 2) * pitch], p2 = d[(c + 2) * pitch];
        vstick[c] = (p1 + p2) / (rhog + rhod);
        dvhat[c] = p1 / rhog - p2 / rhod;
    }
    double a = s1 / sqrt(dvhat[0] * dvhat[0] + dvhat[1] * dvhat[1] + dvhat[2] * dvhat[2]);
    dvhat[0] *= a;
    dvhat[1] *= a;
    dvhat[2] *= a;
    double b = rhog * rhod / (rhog + rhod);
    double dustmom = 0.0, dustmomfin = 0.0;
    for (int c = 0; c < 3; c++) {
        double *gp = g + (c + 2) * pitch, *dp = d + (c + 2) * pitch;
        if (c == 0) dustmom = dp[0] * dp[0];
        else dustmom += dp[0] * dp[0];
        gp[0] = rhog * vstick[c] + dvhat[c] * b;
        double q = rhod * vstick[c] - dvhat[c] * b;
        dp[0] = q;
        if (c == 0) dustmomfin = q * q;
        else dustmomfin += q * q;
    }
    g[pitch] += .5 * (dustmom - dustmomfin) / d[0];
    d[pitch] = .5 * dustmomfin / d[0] + 1e-4 * d[0];
}

/* |dv| after dt by the explicit midpoint rule (solveDragEMP, :305-370) or
 * Kutta's classical RK4 (solveDragRK4, :372-467), in the reference's order of
 * stage and accumulator updates. */
double orc_drag_rk_magdv(const double *P, double rhog, double rhod, double s0, double u0, double dt, int method)
{
    double k = rk_accel(P, rhog, rhod, s0, u0);
    if (method == ORC_DRAG_EMP) {
        double acc = 0.0 * k;                    /* SolveRK_single<true>, B = 0 */
        double s = s0 + (.5 * dt) * k;
        k = rk_accel(P, rhog, rhod, s, rk_u(u0, s0, s, rhog, rhod));
        return s0 + dt * (acc + 1.0 * k);         /* SolveRK_final, B = 1, W = dt */
    }
    double acc = 1.0 * k;
    double s = s0 + (.5 * dt) * k;
    k = rk_accel(P, rhog, rhod, s, rk_u(u0, s0, s, rhog, rhod));
    s = s0 + (.5 * dt) * k;
    acc += 2.0 * k;
    k = rk_accel(P, rhog, rhod, s, rk_u(u0, s0, s, rhog, rhod));
    s = s0 + (1.0 * dt) * k;
    acc += 2.0 * k;
    k = rk_accel(P, rhog, rhod, s, rk_u(u0, s0, s, rhog, rhod));
    return s0 + (dt / 6.0) * (acc + 1.0 * k);
}

/* cukern_ETDRK1 (:1426-1530) without the pressure-gradient drive: each dv
 * component decays by exp(-t k0) about the barycentric velocity; the
 * reference's driven form dpdt/k + (dv - dpdt/k) e^{-kt} is kept (it is NaN
 * when k0 = 0, i.e. below the |dv| cut-off). */
static void etdrk1_apply(double *g, double *d, long pitch, double t, double k)
{
    double rhoginv = 1.0 / g[0];
    double mu = g[0] * d[0] / (g[0] + d[0]);
    double pdustsq = 0.0;
    for (int c = 2; c <= 4; c++) {
        pdustsq -= d[c * pitch] * d[c * pitch];
        double vstick = (g[c * pitch] + d[c * pitch]) / (g[0] + d[0]);
        double dv_i = (g[c * pitch]) * rhoginv - d[c * pitch] / d[0];
        double dpdt = 0.0;
        dpdt *= mu * rhoginv;
        double dv_t = dpdt / k + (dv_i - dpdt / k) * exp(-t * k);
        g[c * pitch] = g[0] * vstick + dv_t * mu;
        double q = d[0] * vstick - dv_t * mu;
        d[c * pitch] = q;
        pdustsq += q * q;
    }
    g[pitch] -= .5 * pdustsq / d[0];
}

int orc_drag(double *gas, double *dust, const orc_grid *g, const orc_thermo *tg, const orc_thermo *td, double dt,
             int method, double tol, long *work)
{
    if (method != ORC_DRAG_LOGTRAP2 && method != ORC_DRAG_LOGTRAP3 && method != ORC_DRAG_EXACT &&
        method != ORC_DRAG_EMP && method != ORC_DRAG_RK4 && method != ORC_DRAG_ETDRK1)
        return -105;
    double P[9];
    orc_drag_params(tg, td, P);
    const long n = (long)g->nx * g->ny * g->nz;
    const long pitch = g->pitch;
    long total = 0, bad = 0;
    long i;
#pragma omp parallel for schedule(static) reduction(+ : total, bad)
    for (i = 0; i < n; i++) {
        double *gc = gas + i, *dc = dust + i;
        double s0, u0;
        if (method == ORC_DRAG_EMP || method == ORC_DRAG_RK4) {
            rk_initial(gc, dc, pitch, &s0, &u0);
            rk_apply(gc, dc, pitch, orc_drag_rk_magdv(P, gc[0], dc[0], s0, u0, dt, method));
            continue;
        }
        prepare(gc, dc, pitch, &s0, &u0);
        if (method == ORC_DRAG_ETDRK1) {
            etdrk1_apply(gc, dc, pitch, dt, orc_drag_k(P, gc[0], dc[0], s0, u0));
            continue;
        }
        double f;
        if (method == ORC_DRAG_EXACT) {
            int steps;
            f = orc_drag_exact_ratio(P, gc[0], dc[0], s0, u0, dt, tol, &steps);
            if (steps == ORC_DRAG_INVALID) bad++;
            else if (steps < 0) bad++, total += -steps;
            else total += steps;
        } else {
            double k = logtrap_rate(P, gc[0], dc[0], s0, u0, dt, method == ORC_DRAG_LOGTRAP3 ? 3 : 2);
            f = exp(-dt * k);
        }
        apply_factor(gc, dc, pitch, f);
    }
    if (work) *work = total;
    return bad ? -101 : 0; /* ERROR_CRASH: invalid cells or an exhausted step budget */
}
