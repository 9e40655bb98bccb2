// drag_device.h -- per-cell arithmetic of the gas-dust drag source (FP64).
//
// Drag rate: cukern_GasDustDrag_GeneralLinearCore + allregimeCdrag
// (gpuclass/cudaSource2FluidDrag.cu:819-890). Powers of the normalised
// temperature share one logarithm: x^a = exp(a ln x) (within a few ulp of pow).
#pragma once
#include <hip/hip_runtime.h>

#include "../../include/imogen_abi.h"

namespace imogen {

// absolute tolerance on ln|dv| of the adaptive reference integrator
constexpr double kDragExactTol = 1e-12;

struct DragConsts {
    double visc0, viscpow, lampow, alpha, beta, delta, epsilon, gm1;
    double tol;
};

struct DragCell {
    double rg, rd;   // gas and dust density
    double s0, u0;   // |dv| and gas specific internal energy
};

__device__ __forceinline__ double pow_from_log(double lx, double a)
{
    return (a == 0.0) ? 1.0 : exp(a * lx);
}

// drag rate k(|dv|, u); no drag below |dv| = 1e-9 (:853-859)
__device__ __forceinline__ double drag_k(const DragConsts &c, double rhog, double rhod, double magdv, double u)
{
    if (magdv < 1e-9) return 0.0;
    const double lT = log(c.alpha * u);
    const double Re = c.delta * rhog * magdv * pow_from_log(lT, c.viscpow);
    const double Kn = c.beta * pow_from_log(lT, c.lampow) / rhog;
    const double cunningham = 1 + Kn * (1.142 + 0.558 * exp(-0.999 / Kn));
    const double Cd = (3 / Re + .45 * exp(-.319 * log(Re)) + .0509 * Re / (8710 + Re)) / cunningham;
    return Cd * magdv * (rhog + rhod) * c.epsilon;
}

// |dv| and u_0 of a cell (cukern_prepareForERK_h0, :1883-1908)
__device__ __forceinline__ DragCell drag_load(const double *g, const double *d, int64_t p)
{
    DragCell s;
    s.rg = g[0];
    s.rd = d[0];
    const double E = g[p], gx = g[2 * p], gy = g[3 * p], gz = g[4 * p];
    const double dx = d[2 * p], dy = d[3 * p], dz = d[4 * p];
    double dv = gx / s.rg - dx / s.rd;
    double dvsq = dv * dv;
    dv = gy / s.rg - dy / s.rd;
    dvsq += dv * dv;
    dv = gz / s.rg - dz / s.rd;
    dvsq += dv * dv;
    s.s0 = sqrt(dvsq);
    s.u0 = (E - .5 * (gx * gx + gy * gy + gz * gz) / s.rg) / s.rg;
    return s;
}

// New momenta about the barycentric velocity, dv scaled by f, and the gas
// energy change conserving total energy (:1632-1676).
__device__ __forceinline__ void drag_apply(double *g, double *d, int64_t p, const DragCell &s, double f)
{
    const double mu = s.rd / (s.rg + s.rd) * s.rg;
    const double rhoginv = 1.0 / s.rg;
    double pdustsq = 0.0;
#pragma unroll
    for (int c = 2; c <= 4; c++) {
        const double pg = g[c * p], pd = d[c * p];
        pdustsq -= pd * pd;
        const double vstick = (pg + pd) / (s.rg + s.rd);
        const double dv_t = (pg * rhoginv - pd / s.rd) * f;
        g[c * p] = s.rg * vstick + dv_t * mu;
        const double q = s.rd * vstick - dv_t * mu;
        d[c * p] = q;
        pdustsq += q * q;
    }
    g[p] -= .5 * pdustsq / s.rd;
}

// Exponential Euler update (cukern_ETDRK1, :1426-1530, with ACCOUNT_GRADP
// undefined, :63): every dv component decays by exp(-t k0) about the
// barycentric velocity, in the reference's driven form dpdt/k + (dv - dpdt/k)
// e^{-kt} with dpdt = 0 (NaN when k0 = 0, as there)
__device__ __forceinline__ void drag_etd1_apply(double *g, double *d, int64_t p, double t, double k)
{
    const double rhoginv = 1.0 / g[0];
    const double mu = g[0] * d[0] / (g[0] + d[0]);
    double pdustsq = 0.0;
#pragma unroll
    for (int c = 2; c <= 4; c++) {
        const double pg = g[c * p], pd = d[c * p];
        pdustsq -= pd * pd;
        const double vstick = (pg + pd) / (g[0] + d[0]);
        const double dv_i = pg * rhoginv - pd / d[0];
        double dpdt = 0.0;
        dpdt *= mu * rhoginv;
        const double dv_t = dpdt / k + (dv_i - dpdt / k) * exp(-t * k);
        g[c * p] = g[0] * vstick + dv_t * mu;
        const double q = d[0] * vstick - dv_t * mu;
        d[c * p] = q;
        pdustsq += q * q;
    }
    g[p] -= .5 * pdustsq / d[0];
}

// path-averaged drag rate of LogTrap2 / LogTrap3 (cukern_LogTrapSolve, :1546-1630)
__device__ __forceinline__ double drag_logtrap_rate(const DragConsts &c, double rhog, double rhod, double dv0,
                                                    double u0, double t, int order)
{
    const double mu = rhod / (rhog + rhod);
    double tmp0 = dv0, tmp1 = u0;
    double ksum = drag_k(c, rhog, rhod, tmp0, tmp1);
    double k = ksum;
    double dv_t = dv0 * exp(-t * k);
    double duint = -.5 * dv0 * dv0 * mu * expm1(-2 * k * t);
    tmp0 = sqrt(dv_t * dv_t);
    tmp1 += duint;
    ksum += drag_k(c, rhog, rhod, tmp0, tmp1);
    if (order == 3) {
        tmp1 -= duint;
        k = .25 * ksum;
        dv_t = dv0 * exp(-t * k);
        duint = -0.5 * dv0 * dv0 * mu * expm1(-2 * k * t);
        tmp0 = sqrt(dv_t * dv_t);
        tmp1 += duint;
        const double khalf = drag_k(c, rhog, rhod, tmp0, tmp1);
        return 0.16666666666666666667 * ksum + 0.66666666666666666667 * khalf;
    }
    return .5 * ksum;
}

// d ln|dv| / dt along the relaxation path, with the gas heated by the
// dissipated relative kinetic energy
__device__ __forceinline__ double drag_exact_rhs(const DragConsts &c, double rhog, double rhod, double s0sq,
                                                 double u0, double h, double y)
{
    const double s = exp(y);
    return -drag_k(c, rhog, rhod, s, u0 + h * (s0sq - s * s));
}

// Whether the relaxation problem is defined on a cell: positive finite
// densities, a finite |dv| and, when the cell moves (|dv| >= 1e-9, the rate's
// cut-off), a positive finite gas internal energy.
__device__ __forceinline__ bool drag_cell_defined(double rhog, double rhod, double s0, double u0)
{
    const double big = 1.7976931348623157e308;
    return (rhog > 0 && rhog <= big) && (rhod > 0 && rhod <= big) && (s0 <= big) &&
           (s0 < 1e-9 || (u0 > 0 && u0 <= big));
}

// drag_exact_ratio's step count for a cell the problem is not defined on
constexpr int kDragInvalid = -(1 << 30);

// |dv(t)|/|dv(0)|: adaptive Dormand-Prince 5(4) (J. Comput. Appl. Math. 6
// (1980) 19) on y = ln|dv|, absolute tolerance c.tol on y, first trial step t.
// *steps receives the number of steps tried (negative if the budget ran out).
// A cell the problem is not defined on (drag_cell_defined), or whose rate is
// not finite, returns NaN at once with *steps = kDragInvalid; a trial step
// whose error estimate is not finite (an overflowing stage of a too-long
// step) is rejected and shrunk by the minimum factor.
__device__ __forceinline__ double drag_exact_ratio(const DragConsts &c, double rhog, double rhod, double s0,
                                                   double u0, double t, int *steps)
{
    const double big = 1.7976931348623157e308;
    *steps = 0;
    if (!drag_cell_defined(rhog, rhod, s0, u0)) {
        *steps = kDragInvalid;
        return __builtin_nan("");
    }
    if (s0 < 1e-9 || t <= 0) return 1.0;
    const double hc = .5 * rhod / (rhog + rhod);
    const double s0sq = s0 * s0;
    const double y0 = log(s0);
    double y = y0, tt = 0.0, h = t;
    double k1 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc, y);
    if (!(fabs(k1) <= big)) {
        *steps = kDragInvalid;
        return __builtin_nan("");
    }
    int n = 0;
    while (tt < t) {
        if (n >= 4096) {
            *steps = -n;
            return exp(y - y0);
        }
        if (h > t - tt) h = t - tt;
        const double k2 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc, y + h * (1.0 / 5) * k1);
        const double k3 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc, y + h * ((3.0 / 40) * k1 + (9.0 / 40) * k2));
        const double k4 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc,
                                         y + h * ((44.0 / 45) * k1 - (56.0 / 15) * k2 + (32.0 / 9) * k3));
        const double k5 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc,
                                         y + h * ((19372.0 / 6561) * k1 - (25360.0 / 2187) * k2 +
                                                  (64448.0 / 6561) * k3 - (212.0 / 729) * k4));
        const double k6 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc,
                                         y + h * ((9017.0 / 3168) * k1 - (355.0 / 33) * k2 + (46732.0 / 5247) * k3 +
                                                  (49.0 / 176) * k4 - (5103.0 / 18656) * k5));
        const double yn = y + h * ((35.0 / 384) * k1 + (500.0 / 1113) * k3 + (125.0 / 192) * k4 -
                                   (2187.0 / 6784) * k5 + (11.0 / 84) * k6);
        const double k7 = drag_exact_rhs(c, rhog, rhod, s0sq, u0, hc, yn);
        double err = h * ((71.0 / 57600) * k1 - (71.0 / 16695) * k3 + (71.0 / 1920) * k4 -
                          (17253.0 / 339200) * k5 + (22.0 / 525) * k6 - (1.0 / 40) * k7);
        err = fabs(err) / c.tol;
        n++;
        if (!(err <= big)) err = 1e300;  // reject, shrink by the minimum factor
        if (err <= 1.0) {
            if (!(fabs(k7) <= big)) {
                *steps = kDragInvalid;
                return __builtin_nan("");
            }
            tt += h;
            y = yn;
            k1 = k7;
        }
        double f = (err > 0) ? 0.9 * pow(err, -0.2) : 5.0;
        f = (f < 0.2) ? 0.2 : ((f > 5.0) ? 5.0 : f);
        h *= f;
    }
    *steps = n;
    return exp(y - y0);
}

// ---- explicit midpoint / RK4 on |dv| (solveDragEMP :305-370, solveDragRK4 :372-467) ----

// d|dv|/dt of cukern_GasDustDrag_GeneralAccel (:898-946); unlike the linear
// core it has no small-|dv| cut-off
__device__ __forceinline__ double drag_rk_accel(const DragConsts &c, double rhog, double rhod, double magdv,
                                                double u)
{
    const double lT = log(c.alpha * u);
    const double Re = c.delta * rhog * magdv * pow_from_log(lT, c.viscpow);
    const double Kn = c.beta * pow_from_log(lT, c.lampow) / rhog;
    const double cunningham = 1 + Kn * (1.142 + 0.558 * exp(-0.999 / Kn));
    const double Cd = (3 / Re + .45 * exp(-.319 * log(Re)) + .0509 * Re / (8710 + Re)) / cunningham;
    return -(Cd * magdv * magdv * (rhog + rhod) * c.epsilon);
}

// |dv| after dt; the stage and accumulator updates of cukern_SolveRK_single /
// _final (:1062-1100) in the reference's order, the gas heated by the
// dissipated relative kinetic energy at every later stage (:921-924)
template <int METHOD>
__device__ __forceinline__ double drag_rk_magdv(const DragConsts &c, double rhog, double rhod, double s0,
                                                double u0, double dt)
{
    const double hc = .5 * rhod;
    const double rs = rhog + rhod;
    auto u_at = [&](double s) { return u0 + hc * (s0 * s0 - s * s) / rs; };
    double k = drag_rk_accel(c, rhog, rhod, s0, u0);
    if (METHOD == DRAG_EMP) {
        const double acc = 0.0 * k;
        const double s = s0 + (.5 * dt) * k;
        k = drag_rk_accel(c, rhog, rhod, s, u_at(s));
        return s0 + dt * (acc + 1.0 * k);
    }
    double acc = 1.0 * k;
    double s = s0 + (.5 * dt) * k;
    k = drag_rk_accel(c, rhog, rhod, s, u_at(s));
    s = s0 + (.5 * dt) * k;
    acc += 2.0 * k;
    k = drag_rk_accel(c, rhog, rhod, s, u_at(s));
    s = s0 + (1.0 * dt) * k;
    acc += 2.0 * k;
    k = drag_rk_accel(c, rhog, rhod, s, u_at(s));
    return s0 + (dt / 6.0) * (acc + 1.0 * k);
}

// cukern_findInitialDeltaV (:1014-1056): |dv| and u_0 with reciprocal densities
__device__ __forceinline__ DragCell drag_rk_load(const double *g, const double *d, int64_t p)
{
    DragCell s;
    s.rg = g[0];
    s.rd = d[0];
    const double rgi = 1 / s.rg, rdi = 1 / s.rd;
    double q = g[2 * p];
    double u = q * rgi - d[2 * p] * rdi;
    double momsq = q * q;
    double dvsq = u * u;
    q = g[3 * p];
    u = q * rgi - d[3 * p] * rdi;
    momsq += q * q;
    dvsq += u * u;
    q = g[4 * p];
    u = q * rgi - d[4 * p] * rdi;
    momsq += q * q;
    dvsq += u * u;
    s.s0 = sqrt(dvsq);
    s.u0 = (g[p] - .5 * momsq * rgi) * rgi;
    return s;
}

// cukern_applyFinalDeltaV (:1160-1237): dv rescaled to |dv| = s1, momenta about
// the barycentric velocity, gas E + dust KE conserved, and the dust energy
// reset to its kinetic energy plus 1e-4 rho_d as the reference does
__device__ __forceinline__ void drag_rk_apply(double *g, double *d, int64_t p, double s1)
{
    const double rhog = g[0], rhod = d[0];
    double vs[3], dv[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double p1 = g[(k + 2) * p], p2 = d[(k + 2) * p];
        vs[k] = (p1 + p2) / (rhog + rhod);
        dv[k] = p1 / rhog - p2 / rhod;
    }
    const double a = s1 / sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    const double b = rhog * rhod / (rhog + rhod);
    double dustmom = 0.0, dustmomfin = 0.0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
        const double pd = d[(k + 2) * p];
        dustmom = (k == 0) ? pd * pd : dustmom + pd * pd;
        const double dvk = dv[k] * a;
        g[(k + 2) * p] = rhog * vs[k] + dvk * b;
        const double q = rhod * vs[k] - dvk * b;
        d[(k + 2) * p] = q;
        dustmomfin = (k == 0) ? q * q : dustmomfin + q * q;
    }
    g[p] += .5 * (dustmom - dustmomfin) / rhod;
    d[p] = .5 * dustmomfin / rhod + 1e-4 * rhod;
}

// The time interval of a drag call: a host value, or frac * dTime read on the
// device (loops that keep dTime in device memory). Scaling by frac = 1/4 or
// 1/2 is exact, so both forms give the same dt for the same dTime.
struct DragDt {
    double dt;
    const double *p;
    double frac;
};
__device__ __forceinline__ double drag_dt(const DragDt &t)
{
    return t.p ? t.frac * t.p[0] : t.dt;
}
inline DragDt host_dt(double dt)
{
    return DragDt{dt, nullptr, 0.0};
}

DragConsts drag_consts(const ThermoDetails *tg, const ThermoDetails *td);
// Every drag kernel counts the cells the method cannot be applied to (and the
// adaptive integrator's budget overruns) into the device error word. calls:
// 1, or 2 for two back-to-back drag calls of dt each applied in one pass.
int launch_drag(double *g, double *d, int64_t pitch, int64_t n, const DragConsts &c, DragDt dt, int method,
                hipStream_t s, int calls = 1);
// queue: one int64 per cell, or null for the library's scratch block; calls
// as launch_drag (two calls: the same result as two launches)
int launch_learned_drag(double *g, double *d, int64_t pitch, int64_t n, const DragConsts &c, DragDt dt,
                        int64_t *queue, hipStream_t s, int calls = 1);

// Drag over many gas/dust array pairs (the dusty problems of an ensemble):
// table entry t = (gas, dust, slab pitch, cells, problem index for dtp), items
// = 64-cell chunks (t, first cell). Each cell gets dt = frac * dtp[problem].
// Methods 4, 5, 6 and 7 (learned, with the queue of rejected cells integrated
// exactly). queue: one entry per cell of all tables per call (2 x cells for
// calls = 2); qcount: 3 counters.
struct EnsDragTable {
    double *g, *d;
    int64_t pitch, n;
    int p, pad;
};
struct EnsDragItem {
    int t, pad;
    int64_t base;
};
int launch_ens_drag(const EnsDragTable *tab, const EnsDragItem *items, int nitems, int64_t ncells,
                    const double *dtp, double frac, const DragConsts &c, int method, int64_t *queue,
                    unsigned long long *qcount, hipStream_t s, int calls = 1);
// changes whenever imogen_drag_set_model loads a model (captured graphs hold it by value)
uint64_t learned_model_generation();
// Allocates the learned drag's device counters (cells, fallbacks, queue
// length) if needed. Call at setup, outside any graph capture: the launchers
// of captured loops never allocate.
int learned_drag_prepare();

}  // namespace imogen
