// sweep_device.h -- per-cell arithmetic of the HLLC and HLL sweeps (gfx950, FP64).
//
// Same formulas, in the same operation order, as the reference kernels
// cukern_HLLC_2ndorder (gpuclass/cudaFluidStep.cu:1124-1716) and
// cukern_HLL_step + cukern_PressureSolverHydro (:1801-2422, 2813-2864):
// primitive recovery (RECONSTRUCT_VELOCITY, gpuclass/Makefile:5), half-slope
// limiters (gpuclass/limiter_functions.h:50-117), Einfeldt speed bounds
// (:1728-1799), the HLLC contact speed and four-way flux selection
// (:1357-1431), the HLL flux (:2124-2196) and the HLL density/energy floors.
// Divisions and square roots use refined hardware estimates (within 1-2 ulp).
// The data movement around these functions (which lane owns which cell, how
// neighbours are exchanged) is this library's own.
#pragma once
#include <hip/hip_runtime.h>

namespace imogen {

// Variables of one cell or face, ordered (rho, e|P, vN, vT1, vT2): normal
// velocity first, then the two tangential components in the order the
// reference kernels read them (X: y,z; Y: x,z; Z: y,x).
struct V5 {
    double a, b, c, d, e;
};

struct HydroConsts {
    double gamma;     // adiabatic index
    double gm1;       // gamma - 1
    double govergm1;  // gamma / (gamma - 1)
    double rhomin;    // density floor (HLL)
    double emin;      // specific internal energy floor 1e-5/(gamma-1) (cudaFluidStep.cu:298)
    double ciso;      // isothermal sound speed, -1 = adiabatic (HLL)
    double cisosq;    // ciso^2
};

__device__ __forceinline__ V5 operator-(const V5 &x, const V5 &y)
{
    return V5{x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d, x.e - y.e};
}

// ---- fast FP64 reciprocal, quotient and square root ------------------------
// The IEEE-exact sequences the compiler emits for '/' and sqrt() handle
// denormal and overflow scaling that the fluid variables never need; these
// use the hardware estimates (v_rcp_f64, v_rsq_f64) with Newton/Goldschmidt
// refinement to full double precision (within 1-2 ulp). Arguments are finite,
// non-zero normals (rcp, quotient) or >= 0 normals (sqrt).
// v_rcp_f64 / v_rsq_f64 are accurate to ~4.6e-8 and ~5.2e-8 relative
// (measured on gfx950 over 4M values spanning 2^-30..2^30).
__device__ __forceinline__ double rcp_nr(double x)
{
    // third-order refinement r0 (1 + e + e^2), e = 1 - x r0: error ~ e^3
    const double r = __builtin_amdgcn_rcp(x);
    const double e = fma(-x, r, 1.0);
    return fma(r, fma(e, e, e), r);
}

__device__ __forceinline__ double div_nr(double a, double b)
{
    return a * rcp_nr(b);
}

// Quotient with one residual correction: q + (a - q b) / b. With a correctly
// rounded reciprocal this is the correctly rounded quotient (Markstein,
// IBM J. Res. Dev. 34 (1990) 111), i.e. what the reference's `a / b` gives.
__device__ __forceinline__ double div_cr(double a, double b)
{
    const double r = rcp_nr(b);
    const double q = a * r;
    return fma(fma(-q, b, a), r, q);
}

__device__ __forceinline__ double sqrt_nr(double x)
{
    // One Goldschmidt step and one Newton correction reach full precision from
    // the hardware estimate. The 1e-300 offset keeps rsq finite at x = 0 (so
    // sqrt(0) = 0) without changing any argument >= 1e-284; x < 0 stays NaN.
    const double y = __builtin_amdgcn_rsq(x + 1e-300);
    double g = x * y;
    double h = 0.5 * y;
    const double r = fma(-g, h, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    const double d = fma(-g, g, x);
    return fma(d, h, g);
}

// Reduced-refinement variants for the sweep arithmetic: one quadratic Newton
// step for the reciprocal (relative error ~ e^2 + 1 ulp ~ 3e-15) and one
// Goldschmidt step for the square root (~5e-15). The sweep results stay far
// inside the 1e-12 parity bar; the correctly rounded quotient (div_cr) keeps
// the full-precision reciprocal.
#ifndef IMOGEN_SWEEP_FAST_MATH
#define IMOGEN_SWEEP_FAST_MATH 1
#endif
__device__ __forceinline__ double rcp_fast(double x)
{
#if IMOGEN_SWEEP_FAST_MATH
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return rcp_nr(x);
#endif
}

__device__ __forceinline__ double div_fast(double a, double b)
{
    return a * rcp_fast(b);
}

__device__ __forceinline__ double sqrt_fast(double x)
{
#if IMOGEN_SWEEP_FAST_MATH
    const double y = __builtin_amdgcn_rsq(x + 1e-300);
    const double g = x * y;
    const double r = fma(-g, 0.5 * y, 0.5);
    return fma(g, r, g);
#else
    return sqrt_nr(x);
#endif
}

// sqrt_fast for arguments known to be > 0 (ratios of densities)
__device__ __forceinline__ double sqrt_fast_pos(double x)
{
#if IMOGEN_SWEEP_FAST_MATH
    const double y = __builtin_amdgcn_rsq(x);
    const double g = x * y;
    const double r = fma(-g, 0.5 * y, 0.5);
    return fma(g, r, g);
#else
    return sqrt_nr(x);
#endif
}

// ---- half-slope limiters (limiter_functions.h) ----------------------------
__device__ __forceinline__ double flux_minmod(double a, double b)
{
    if (a * b < 0) return 0.0;
    return (fabs(a) > fabs(b)) ? b : a;
}

template <int LIM>
__device__ __forceinline__ double slope_limit(double a, double b, int rt)
{
    const int which = (LIM >= 0) ? LIM : rt;
    switch (which) {
    case 0: {  // Ospre: .75 A B (A+B) / (A^2 + AB + B^2)
        // With A B > 0 the denominator falls below the normal range only for
        // |A|, |B| < 1.5e-154, where the numerator has already underflowed to
        // zero and the quotient is +-0; the hardware reciprocal of such a
        // denominator is infinite, so it is raised to the smallest normal.
        double r = a * b;
        double s = a + b;
        double v = div_fast(.75 * r * s, fmax(s * s - r, 0x1p-1022));
        return (r > 0) ? v : 0.0;
    }
    case 1: {  // van Leer: A B / (A+B)
        double r = a * b;
        double v = div_fast(r, a + b);
        return (r > 0.0) ? v : 0.0;
    }
    case 2: {  // minmod, returned through float as the reference does (:70)
        if (a * b < 0) return 0.0;
        return (double)(float)((fabs(a) > fabs(b)) ? .5 * b : .5 * a);
    }
    case 3: {  // superbee (first branch without the 1/2, :80-86)
        if (a * b < 0) return 0.0;
        if (b < a) return flux_minmod(a, 2 * b);
        return .5 * flux_minmod(2 * a, b);
    }
    case 4: {  // van Albada
        double r = a * b;
        if (r < 0) return 0;
        return .5 * r * (a + b) / (a * a + b * b);
    }
    case 5: {  // monotonized central
        if (a * b <= 0) return 0;
        double r = b / a;
        double s = .25 + .25 * r;
        s = (s < r) ? s : r;
        s = (s < 1) ? s : 1.0;
        return s * a;
    }
    default:
        return 0.0;
    }
}

// Conserved (rho, E, pN, pT1, pT2) -> primitive (rho, e_int, vN, vT1, vT2),
// cudaFluidStep.cu:1218-1234.
__device__ __forceinline__ V5 to_prim(const V5 &q)
{
    double F = rcp_fast(q.a);
    double B = q.b - .5 * (q.c * q.c + q.d * q.d + q.e * q.e) * F;
    return V5{q.a, B, q.c * F, q.d * F, q.e * F};
}

// Left/right face states of a cell from its primitive w and the backward
// differences dl = w_i - w_{i-1}, dr = w_{i+1} - w_i. The energy slot holds
// pressure (gm1 * e) on the faces (cudaFluidStep.cu:1277-1278).
template <int LIM>
__device__ __forceinline__ void faces(const HydroConsts &k, const V5 &w, const V5 &dl, const V5 &dr,
                                      int rt, V5 &L, V5 &R)
{
    double s;
    s = slope_limit<LIM>(dl.a, dr.a, rt);
    L.a = w.a - s;
    R.a = w.a + s;
    s = slope_limit<LIM>(dl.b, dr.b, rt);
    L.b = k.gm1 * (w.b - s);
    R.b = k.gm1 * (w.b + s);
    s = slope_limit<LIM>(dl.c, dr.c, rt);
    L.c = w.c - s;
    R.c = w.c + s;
    s = slope_limit<LIM>(dl.d, dr.d, rt);
    L.d = w.d - s;
    R.d = w.d + s;
    s = slope_limit<LIM>(dl.e, dr.e, rt);
    L.e = w.e - s;
    R.e = w.e + s;
}

// HLLC flux through the interface between face state L (right face of cell i)
// and R (left face of cell i+1). Returns (mass, energy, pN, pT1, pT2) fluxes.
// The four reference cases are evaluated without divergence: the upwind side
// and its wave speed are selected first, then the plain and star fluxes of
// that side are both formed and one is kept.
// BRANCH: form the plain or the star flux under a branch (fewer live values;
// the X kernel's register budget needs it) instead of the select form below.
template <bool BRANCH>
__device__ __forceinline__ V5 hllc_flux(const HydroConsts &k, const V5 &L, const V5 &R)
{
    // velocity jump squared (cudaFluidStep.cu:1318-1322)
    double C = L.c - R.c;
    double D = L.d - R.d;
    double E = L.e - R.e;
    double dvsq = C * C + D * D + E * E;

    // Einfeldt bounds (cudaFluidStep.cu:1728-1754). The two face densities'
    // reciprocals serve both sound speeds, the Roe ratio and the star flux.
    const double iL = rcp_fast(L.a);
    const double iR = rcp_fast(R.a);
    double csql = k.gamma * L.b * iL;
    double csqr = k.gamma * R.b * iR;
    double Fa = sqrt_fast_pos(R.a * iL);
    double Fb = rcp_fast(1.0 + Fa);
    double Croe = ((csql + Fa * csqr) + .5 * k.gm1 * dvsq * Fa * Fb) * Fb;
    double Vroe = (L.c + Fa * R.c) * Fb;
    Croe = Vroe - sqrt_fast(Croe);
    double sl = L.c - sqrt_fast(csql);
    double sr = R.c + sqrt_fast(csqr);
    // the reference's selects, not fmin/fmax: a NaN bound (negative face
    // pressure) must propagate as it does there
    sl = (Croe < sl) ? Croe : sl;
    Croe = 2 * Vroe - Croe;
    sr = (Croe > sr) ? Croe : sr;

    // contact speed, Batten et al. eq. 34 (cudaFluidStep.cu:1352-1357)
    double A = L.a * (sl - L.c);
    double B = R.a * (sr - R.c);
    double sm = div_fast(B * R.c - R.b - A * L.c + L.b, B - A);

    // case selection (cudaFluidStep.cu:1363-1369)
    const bool left = sm > 0;
    const bool star = left ? !(sl > 0) : (sr > 0);
    const V5 S = left ? L : R;
    const double sk = left ? sl : sr;
    const double iS = left ? iL : iR;

    double ksum = S.c * S.c + S.d * S.d + S.e * S.e;  // |v|^2 of the upwind face
    if (BRANCH) {
        V5 out;
        if (!star) {
            // cases 0/1 (cudaFluidStep.cu:1372-1383, 1416-1430)
            const double Ek = .5 * ksum * S.a;
            const double m = S.a * S.c;
            out.a = m;
            out.b = S.c * (Ek + k.govergm1 * S.b);
            out.c = m * S.c + S.b;
            out.d = S.d * m;
            out.e = S.e * m;
        } else {
            // cases 2/3 (cudaFluidStep.cu:1384-1415)
            const double rb = div_fast(S.a * (sk - S.c), sk - sm);  // rho * beta
            const double G = .5 * ksum;
            const double m = rb * sm;
            const double Cc = sk * (sm - S.c);
            out.a = m;
            out.b = rb * sm * (k.govergm1 * S.b * iS + G + Cc);
            out.c = S.b + rb * (Cc + sm * S.c);
            out.d = S.d * m;
            out.e = S.e * m;
        }
        return out;
    }
    // Plain flux of the upwind side (cases 0/1, cudaFluidStep.cu:1372-1383,
    // 1416-1430) and its star flux (cases 2/3, :1384-1415) share their form:
    // with mass flux m (rho v of the face, or rho beta S_M), the energy flux
    // is m (gamma/(gamma-1) P/rho + |v|^2/2 [+ S_K (S_M - v)]) and the
    // tangential fluxes are v_t m, so one select per differing term replaces
    // forming both fluxes and selecting five results.
    const double rb = div_fast(S.a * (sk - S.c), sk - sm);  // rho * beta (star)
    const double G = .5 * ksum;
    const double mp = S.a * S.c;
    const double Cc = sk * (sm - S.c);
    const double m = star ? rb * sm : mp;
    const double cz = star ? Cc : 0.0;
    const double tn = star ? rb * (Cc + sm * S.c) : mp * S.c;
    V5 out;
    out.a = m;
    out.b = m * (k.govergm1 * S.b * iS + G + cz);
    out.c = S.b + tn;
    out.d = S.d * m;
    out.e = S.e * m;
    return out;
}

// ---------------------------------------------------------------------------
// Method policies used by the sweep kernels
// ---------------------------------------------------------------------------
// floor(q): applies the method's pointwise floors to the conserved cell q in
// place (the floored q is what the SSP-RK2 combination uses) and returns the
// internal energy density the floors settled on. prims(q, e): reconstruction
// variables of a floored cell. face(): limited left/right face states.
// flux(): interface flux in slab order (mass, energy, normal, t1, t2).

struct Hllc {
    template <int DIR>
    __device__ static __forceinline__ double floor(const HydroConsts &, V5 &)
    {
        return 0.0;
    }
    __device__ static __forceinline__ V5 prims(const HydroConsts &, const V5 &q, double)
    {
        return to_prim(q);
    }
    template <int LIM>
    __device__ static __forceinline__ void face(const HydroConsts &k, const V5 &w, const V5 &dl, const V5 &dr,
                                                int rt, V5 &L, V5 &R)
    {
        faces<LIM>(k, w, dl, dr, rt, L, R);
    }
    template <bool LOWREG = false>
    __device__ static __forceinline__ V5 flux(const HydroConsts &k, const V5 &L, const V5 &R)
    {
        return hllc_flux<LOWREG>(k, L, R);
    }
};

struct Hll {
    // q = (rho, E, pN, pT1, pT2) in HLL's component order for direction DIR:
    // X (px,py,pz), Y (py,px,pz), Z (pz,px,py) (cudaFluidStep.cu:1869-1885).
    // The floors accumulate |p|^2 in physical x,y,z order (:2821-2843).
    template <int DIR>
    __device__ static __forceinline__ double floor(const HydroConsts &k, V5 &q)
    {
        double rho = q.a;
        if (rho < k.rhomin) {
            const double f = k.rhomin / rho;
            rho = q.a = k.rhomin;
            q.b = f * q.b;
            q.c = f * q.c;
            q.d = f * q.d;
            q.e = f * q.e;
        }
        const double px = (DIR == 1) ? q.c : q.d;
        const double py = (DIR == 1) ? q.d : ((DIR == 2) ? q.c : q.e);
        const double pz = (DIR == 3) ? q.c : q.e;
        double momsq = px * px;
        momsq += py * py;
        momsq += pz * pz;
        const double ke = div_cr(.5 * momsq, rho);
        double e = q.b - ke;
        if (e < rho * k.emin) {
            e = k.emin * rho;
            q.b = ke + e;
        }
        return e;
    }
    // (rho, vN, P, vT1, vT2); velocities are quotients, as in the reference (:1887-1889)
    __device__ static __forceinline__ V5 prims(const HydroConsts &k, const V5 &q, double e)
    {
        return V5{q.a, div_cr(q.c, q.a), e * k.gm1, div_cr(q.d, q.a), div_cr(q.e, q.a)};
    }
    template <int LIM>
    __device__ static __forceinline__ void face(const HydroConsts &, const V5 &w, const V5 &dl, const V5 &dr,
                                                int rt, V5 &L, V5 &R)
    {
        double s;
        s = slope_limit<LIM>(dl.a, dr.a, rt);
        L.a = w.a - s;
        R.a = w.a + s;
        s = slope_limit<LIM>(dl.b, dr.b, rt);
        L.b = w.b - s;
        R.b = w.b + s;
        s = slope_limit<LIM>(dl.c, dr.c, rt);
        L.c = w.c - s;
        R.c = w.c + s;
        s = slope_limit<LIM>(dl.d, dr.d, rt);
        L.d = w.d - s;
        R.d = w.d + s;
        s = slope_limit<LIM>(dl.e, dr.e, rt);
        L.e = w.e - s;
        R.e = w.e + s;
    }
    // L, R = (rho, vN, P, vT1, vT2), the reference's A..E (cudaFluidStep.cu:2124-2196)
    template <bool LOWREG = false>
    __device__ static __forceinline__ V5 flux(const HydroConsts &k, const V5 &L, const V5 &R)
    {
        const double Ut = R.b - L.b, St1 = R.d - L.d, St2 = R.e - L.e;
        const double dvsq = Ut * Ut + St1 * St1 + St2 * St2;
        double sl, sr;
        const double iL = rcp_fast(L.a);
        const double Fa = sqrt_fast(R.a * iL);
        const double Fb = rcp_fast(1.0 + Fa);
        const double Vroe = (L.b + Fa * R.b) * Fb;
        if (k.ciso == -1) {
            const double csql = k.gamma * L.c * iL;
            const double csqr = k.gamma * R.c * rcp_fast(R.a);
            double Croe = ((csql + Fa * csqr) + .5 * k.gm1 * dvsq * Fa * Fb) * Fb;
            Croe = Vroe - sqrt_fast(Croe);
            sl = L.b - sqrt_fast(csql);
            sr = R.b + sqrt_fast(csqr);
            sl = (Croe < sl) ? Croe : sl;
            Croe = 2 * Vroe - Croe;
            sr = (Croe > sr) ? Croe : sr;
        } else {
            // computeEinfeldtBounds_isothermal, passed C_iso^2 for the sound speed (:2131)
            double cs = Vroe - k.cisosq;
            sl = L.b - cs;
            sr = R.b + cs;
            sl = (cs < sl) ? cs : sl;
            cs = 2 * Vroe - cs;
            sr = (cs > sr) ? cs : sr;
        }
        const double At = rcp_fast(sr - sl);
        const double Fl = L.a * L.b, Fr = R.a * R.b;
        V5 o;
        if (sl >= 0 || sr <= 0) {  // one-sided (left unless the right case also holds)
            const bool right = sr <= 0;
            const V5 S = right ? R : L;
            const double F = right ? Fr : Fl;
            const double ke = .5 * S.a * (S.b * S.b + S.d * S.d + S.e * S.e);
            o.a = F;
            if (k.ciso == -1) {
                o.b = S.b * (k.govergm1 * S.c + ke);
                o.c = F * S.b + S.c;
            } else {
                o.b = S.b * (S.c + ke);
                o.c = F * S.b + (right ? k.ciso : k.cisosq) * S.a;  // C_iso on the right side (:2191)
            }
            o.d = F * S.d;
            o.e = F * S.e;
        } else {
            const double wl = sr * (L.b - sl);
            const double wr = sl * (R.b - sr);
            o.a = (wl * L.a - wr * R.a) * At;
            if (k.ciso == -1) o.c = (wl * Fl - wr * Fr + sr * L.c - sl * R.c) * At;
            else o.c = (wl * Fl - wr * Fr + k.cisosq * (sr * L.a - sl * R.a)) * At;
            o.d = (wl * L.a * L.d - wr * R.a * R.d) * At;
            o.e = (wl * L.a * L.e - wr * R.a * R.e) * At;
            if (k.ciso == -1) {
                const double kl = .5 * L.a * (L.b * L.b + L.d * L.d + L.e * L.e);
                const double kr = .5 * R.a * (R.b * R.b + R.d * R.d + R.e * R.e);
                const double e = (wl * kl - wr * kr);
                o.b = (e + (k.govergm1 * (sr * L.c * L.b - sl * R.c * R.b) + sl * sr * (R.c - L.c) / k.gm1)) * At;
            } else {
                // no 1/(Sr-Sl) on the isothermal energy flux in the reference (:2185-2189)
                const double kl = .5 * L.a * (L.b * L.b + L.d * L.d + L.e * L.e) + L.c;
                const double kr = .5 * R.a * (R.b * R.b + R.d * R.d + R.e * R.e) + R.c;
                o.b = (wl * kl - wr * kr);
            }
        }
        return o;
    }
};

// Slabs holding (normal, t1, t2) momentum per method and direction.
template <class M, int DIR>
struct MomSlabs;
template <>
struct MomSlabs<Hllc, 1> {
    static constexpr int n = 2, t1 = 3, t2 = 4;
};
template <>
struct MomSlabs<Hllc, 2> {
    static constexpr int n = 3, t1 = 2, t2 = 4;
};
template <>
struct MomSlabs<Hllc, 3> {
    static constexpr int n = 4, t1 = 3, t2 = 2;
};
template <>
struct MomSlabs<Hll, 1> {
    static constexpr int n = 2, t1 = 3, t2 = 4;
};
template <>
struct MomSlabs<Hll, 2> {
    static constexpr int n = 3, t1 = 2, t2 = 4;
};
template <>
struct MomSlabs<Hll, 3> {
    static constexpr int n = 4, t1 = 2, t2 = 3;
};

}  // namespace imogen
