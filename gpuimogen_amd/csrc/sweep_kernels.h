// sweep_kernels.h -- the HLLC / HLL sweep kernels (templates) shared by the
// single-array sweeps (fluid_sweep.hip) and the batched ensemble sweeps
// (ensemble.hip). See fluid_sweep.hip for the data-movement design.
#pragma once
#include "cfl_device.h"
#include "sweep_device.h"

namespace imogen {

// ---------------------------------------------------------------------------
// cross-lane helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ double shfl_d(double v, int src)
{
    return __shfl(v, src, 64);
}

__device__ __forceinline__ V5 shfl5(const V5 &v, int src)
{
    return V5{shfl_d(v.a, src), shfl_d(v.b, src), shfl_d(v.c, src), shfl_d(v.d, src),
              shfl_d(v.e, src)};
}

__device__ __forceinline__ V5 sel5(bool c, const V5 &x, const V5 &y)
{
    return c ? x : y;
}

__device__ __forceinline__ V5 axpy_update(const V5 &q, double lam, const V5 &fr, const V5 &fl)
{
    // Qout = Qin - lambda*(F_{i+1/2} - F_{i-1/2})      (cudaFluidStep.cu:1510-1516)
    return V5{q.a - lam * (fr.a - fl.a), q.b - lam * (fr.b - fl.b), q.c - lam * (fr.c - fl.c),
              q.d - lam * (fr.d - fl.d), q.e - lam * (fr.e - fl.e)};
}

__device__ __forceinline__ V5 avg_update(const V5 &s, double lam, const V5 &fr, const V5 &fl)
{
    // Qout = .5*((Qin + Qout) - lambda*(F_{i+1/2} - F_{i-1/2}))   (cudaFluidStep.cu:1555-1561)
    // s already holds (Qin + Qout).
    return V5{.5 * (s.a - lam * (fr.a - fl.a)), .5 * (s.b - lam * (fr.b - fl.b)),
              .5 * (s.c - lam * (fr.c - fl.c)), .5 * (s.d - lam * (fr.d - fl.d)),
              .5 * (s.e - lam * (fr.e - fl.e))};
}

__device__ __forceinline__ V5 add5(const V5 &x, const V5 &y)
{
    return V5{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d, x.e + y.e};
}

// NT: loads and stores with the nontemporal hint. The single-grid Y/Z
// marches over rows of a multiple of 16 cells read and write whole 128-byte
// cache lines once per sweep and use it (Sedov 256^3 device loop 1 % faster,
// measured). The X sweep's lanes hold K consecutive cells, so one of its
// store instructions covers each cache line only in part: nontemporal stores
// there doubled its time. Rows that start at other 8-byte offsets (the
// ensemble's) are 10 % faster without the hint.
template <class Meth, int DIR, bool NT = false>
__device__ __forceinline__ V5 load_cell(const double *__restrict__ p, int64_t pitch)
{
    using M = MomSlabs<Meth, DIR>;
    if (NT)
        return V5{__builtin_nontemporal_load(p), __builtin_nontemporal_load(p + pitch),
                  __builtin_nontemporal_load(p + M::n * pitch), __builtin_nontemporal_load(p + M::t1 * pitch),
                  __builtin_nontemporal_load(p + M::t2 * pitch)};
    return V5{p[0], p[pitch], p[M::n * pitch], p[M::t1 * pitch], p[M::t2 * pitch]};
}

template <class Meth, int DIR, bool NT = false>
__device__ __forceinline__ void store_cell(double *__restrict__ p, int64_t pitch, const V5 &v)
{
    using M = MomSlabs<Meth, DIR>;
    if (NT) {
        __builtin_nontemporal_store(v.a, p);
        __builtin_nontemporal_store(v.b, p + pitch);
        __builtin_nontemporal_store(v.c, p + M::n * pitch);
        __builtin_nontemporal_store(v.d, p + M::t1 * pitch);
        __builtin_nontemporal_store(v.e, p + M::t2 * pitch);
        return;
    }
    p[0] = v.a;
    p[pitch] = v.b;
    p[M::n * pitch] = v.c;
    p[M::t1 * pitch] = v.d;
    p[M::t2 * pitch] = v.e;
}

// time-step limit of a stored cell (physical x, y, z momentum order, as the
// CFL kernel reads the slabs)
template <class Meth, int DIR>
__device__ __forceinline__ double cfl_of_cell(const CflArgs &a, const V5 &o)
{
    using M = MomSlabs<Meth, DIR>;
    const double px = (M::n == 2) ? o.c : ((M::t1 == 2) ? o.d : o.e);
    const double py = (M::n == 3) ? o.c : ((M::t1 == 3) ? o.d : o.e);
    const double pz = (M::n == 4) ? o.c : ((M::t1 == 4) ? o.d : o.e);
    return cfl_cell(a, o.a, o.b, px, py, pz, 0.0);
}

// ---------------------------------------------------------------------------
// X sweep: lanes own K consecutive cells of a line
// ---------------------------------------------------------------------------

// Interface fluxes of the cells a lane owns. F[k] is the flux at k+1/2 for
// k < cnt; Fm is the flux at -1/2 (the left lane's last interface). The
// method's floors are applied to U in place first.
template <class M, int K, int LIM>
__device__ __forceinline__ void x_fluxes(const HydroConsts &hc, int rt, V5 (&U)[K], int cnt, int srcL,
                                         int srcR, V5 (&F)[K], V5 &Fm)
{
    V5 w[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const double e = M::template floor<1>(hc, U[k]);
        w[k] = M::prims(hc, U[k], e);
    }

    V5 wlast = w[0];
#pragma unroll
    for (int k = 1; k < K; k++) wlast = sel5(k == cnt - 1, w[k], wlast);
    const V5 wl = shfl5(wlast, srcL);
    const V5 wr = shfl5(w[0], srcR);

    // backward differences d[k] = w[k] - w[k-1]; d[cnt] uses the right neighbour
    V5 d[K + 1];
    d[0] = w[0] - wl;
#pragma unroll
    for (int k = 1; k < K; k++) d[k] = sel5(k < cnt, w[k], wr) - w[k - 1];
    d[K] = wr - w[K - 1];

    V5 Lf[K], Rf[K];
#pragma unroll
    for (int k = 0; k < K; k++) M::template face<LIM>(hc, w[k], d[k], d[k + 1], rt, Lf[k], Rf[k]);

    const V5 LfR = shfl5(Lf[0], srcR);
#pragma unroll
    for (int k = 0; k < K; k++) {
        const V5 right = (k + 1 < K) ? sel5(k + 1 < cnt, Lf[(k + 1 < K) ? k + 1 : 0], LfR) : LfR;
        F[k] = M::template flux<(K >= 4)>(hc, Rf[k], right);
    }
    V5 Flast = F[0];
#pragma unroll
    for (int k = 1; k < K; k++) Flast = sel5(k == cnt - 1, F[k], Flast);
    Fm = shfl5(Flast, srcL);
}

// Each wave loops over groups of lines (lpw lines per group, groups strided by
// the number of waves in the grid). While one group is computed, the next
// group's cells are already being loaded. FULL: every line is exactly nl*K
// cells and nl divides 64, so every lane holds K cells and all memory
// operations are unconditional (the compiler can then count them exactly and
// wait only for the loads it needs); lanes past the last line read a clamped
// valid line and write into a private dummy slot.
template <class M, int K, int LIM, bool FULL, bool CFL>
__global__ __launch_bounds__(256, (K <= 4) ? 2 : 1) void sweep_x(double *__restrict__ q, int64_t pitch, int N, int64_t nlines,
                                               int nl, double lambda, const double *__restrict__ lamp,
                                               HydroConsts hc, int rt, double *__restrict__ dummy, CflFuse cf)
{
    if (lamp) lambda = lamp[0];  // time step kept on the device
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int64_t nwaves = (int64_t)gridDim.x * (blockDim.x >> 6);
    const int lpw = 64 / nl;
    const int sub = lane / nl;
    const int j = lane - sub * nl;
    const bool laneUsed = sub < lpw;
    const int lbase = sub * nl;
    const int srcL = laneUsed ? lbase + (j == 0 ? nl - 1 : j - 1) : lane;
    const int srcR = laneUsed ? lbase + (j == nl - 1 ? 0 : j + 1) : lane;
    const int x0 = laneUsed ? j * K : 0;
    int cntL = N - x0;
    cntL = cntL < K ? cntL : K;
    if (!laneUsed) cntL = 0;
    const int cnt = FULL ? K : cntL;
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    double *mydummy = dummy + gtid;
    const int64_t groupStride = nwaves * lpw;

    auto line_ptr = [&](int64_t ln) {
        const int64_t c = ln < nlines ? ln : nlines - 1;
        return q + c * (int64_t)N + x0;
    };
    auto load_group = [&](int64_t ln, V5 (&Qd)[K]) {
        double *b = line_ptr(ln);
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (k < cnt) Qd[k] = load_cell<M, 1>(b + k, pitch);
            else Qd[k] = V5{1.0, 1.0, 0.0, 0.0, 0.0};
        }
    };

    int64_t line = wave * lpw + (laneUsed ? sub : 0);
    int64_t gbase = wave * lpw;  // first line of this wave's group (uniform)
    if (gbase >= nlines) {
        if (CFL && lane == 0) cf.part[wave] = 1e38;
        return;
    }
    double cmin = 1e38;
    V5 Q[K], Qn[K];
    load_group(line, Q);

    for (;;) {
        const bool more = gbase + groupStride < nlines;  // uniform across the wave
        load_group(line + groupStride, Qn);  // unconditional: clamped to a valid line when !more

        // stage A: U1 = Q - lambda dF(Q)
        V5 F[K], Fm;
        x_fluxes<M, K, LIM>(hc, rt, Q, cnt, srcL, srcR, F, Fm);
        V5 U1[K];
#pragma unroll
        for (int k = 0; k < K; k++) U1[k] = axpy_update(Q[k], lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);

        // stage B: Q = .5((U1 + Q) - lambda dF(U1))
        x_fluxes<M, K, LIM>(hc, rt, U1, cnt, srcL, srcR, F, Fm);
        const bool own = laneUsed && line < nlines;
        double *dst = own ? line_ptr(line) : mydummy;
        const int64_t st = own ? pitch : nthr;
#pragma unroll
        for (int k = 0; k < K; k++) {
            if (k < cnt) {
                const V5 o = avg_update(add5(U1[k], Q[k]), lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
                store_cell<M, 1>(own ? dst + k : dst, st, o);
                if (CFL && own && x0 + k >= cf.lo && x0 + k < cf.hi) {
                    const double u = cfl_of_cell<M, 1>(cf.a, o);
                    cmin = (u < cmin) ? u : cmin;
                }
            }
        }
        if (!more) break;
#pragma unroll
        for (int k = 0; k < K; k++) Q[k] = Qn[k];
        line += groupStride;
        gbase += groupStride;
    }
    if (CFL) {
        cmin = wave_min(cmin);
        if (lane == 0) cf.part[wave] = cmin;
    }
}

// X sweep of lines of 257-512 cells split into halves, one wave per half. A
// wave holds its half and the 4 cells beyond each end (the two stages'
// stencil, periodic within the line, as the whole-line sweep reads them), so
// it forms its half's outputs alone; the window's outer cells only feed
// discarded values. The two waves of a line are in one workgroup (4 waves: 2
// lines), and every load of the workgroup precedes the barrier while every
// store follows it, so no wave overwrites a cell its partner has still to
// read. K = 3, 4 or 5 cells per lane keeps two waves per SIMD where a whole
// line needs 6 or 8 at one.
template <class M, int K, int LIM, bool CFL>
__global__ __launch_bounds__(256, 2) void sweep_xh(double *__restrict__ q, int64_t pitch, int N, int64_t nlines,
                                                  double lambda, const double *__restrict__ lamp, HydroConsts hc,
                                                  int rt, double *__restrict__ dummy, CflFuse cf)
{
    if (lamp) lambda = lamp[0];  // time step kept on the device
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int64_t wave = (int64_t)blockIdx.x * 4 + wib;
    int64_t line = (int64_t)blockIdx.x * 2 + (wib >> 1);
    const bool lineValid = line < nlines;
    line = lineValid ? line : nlines - 1;  // waves past the last line still meet the barrier
    const int part = wib & 1;
    const int h = (N + 1) >> 1;
    const int s0 = part ? h : 0;             // first output cell of this half
    const int Wn = (part ? N - h : h) + 8;   // window: the half and 4 cells beyond each end
    const int nl = (Wn + K - 1) / K;
    const bool laneUsed = lane < nl;
    const int srcL = laneUsed ? (lane == 0 ? nl - 1 : lane - 1) : lane;
    const int srcR = laneUsed ? (lane == nl - 1 ? 0 : lane + 1) : lane;
    const int w0 = lane * K;
    int cnt = Wn - w0;
    cnt = cnt < K ? cnt : K;
    if (!laneUsed) cnt = 0;
    double *row = q + line * (int64_t)N;
    V5 Q[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        int g = s0 - 4 + w0 + k;  // window cell -> line cell, periodic
        g = (g < 0) ? g + N : g;
        g = (g >= N) ? g - N : g;
        g = (g < N) ? g : N - 1;
        Q[k] = (k < cnt) ? load_cell<M, 1>(row + g, pitch) : V5{1.0, 1.0, 0.0, 0.0, 0.0};
    }
    __syncthreads();
    V5 F[K], Fm;
    x_fluxes<M, K, LIM>(hc, rt, Q, cnt, srcL, srcR, F, Fm);
    V5 U1[K];
#pragma unroll
    for (int k = 0; k < K; k++) U1[k] = axpy_update(Q[k], lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
    x_fluxes<M, K, LIM>(hc, rt, U1, cnt, srcL, srcR, F, Fm);
    const int64_t gtid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * blockDim.x;
    double cmin = 1e38;
#pragma unroll
    for (int k = 0; k < K; k++) {
        if (k < cnt) {
            const int w = w0 + k;
            const bool own = lineValid && w >= 4 && w < Wn - 4;
            const int x = s0 + w - 4;
            const V5 o = avg_update(add5(U1[k], Q[k]), lambda, F[k], k == 0 ? Fm : F[k > 0 ? k - 1 : 0]);
            store_cell<M, 1>(own ? row + x : dummy + gtid, own ? pitch : nthr, o);
            if (CFL && own && x >= cf.lo && x < cf.hi) {
                const double u = cfl_of_cell<M, 1>(cf.a, o);
                cmin = (u < cmin) ? u : cmin;
            }
        }
    }
    if (CFL) {
        cmin = wave_min(cmin);
        if (lane == 0) cf.part[wave] = cmin;
    }
}

// ---------------------------------------------------------------------------
// Y/Z sweep: register-window march along strided columns
// ---------------------------------------------------------------------------
// One wave-uniform march over the segments of W = 64/S columns: lane `lane`
// walks cells [j0, j0+L) of the column starting at `col` (stride sstride, N
// cells, periodic), the wave's stencil ring in LDS at `ring`. defer/gtid/nthr
// address this lane's deferred-output slots. seglen is uniform over the wave.
// CFL: cmin takes the minimum CFL limit of the outputs at column cells
// [cf.lo, cf.hi), nanc counts the NaN limits among them.
template <class M, int DIR, int LIM, bool CFL = false, bool NT = false>
__device__ __forceinline__ void march_body(double *__restrict__ col, int64_t pitch, int64_t sstride, int N,
                                           int seglen, int j0, int L, double lambda, const HydroConsts &hc,
                                           int rt, double *__restrict__ defer, int64_t gtid, int64_t nthr,
                                           double (*ring)[5][64], int lane, const CflFuse &cf, double &cmin,
                                           unsigned &nanc)
{
    // wrapped sweep index of the first load (m = j0 - 4), in [0, N)
    int mw = (j0 - 4) % N;
    if (mw < 0) mw += N;
    double *outp = col + (int64_t)(j0 - 9) * sstride;  // address of output t - 9
    double *const dfr = defer + gtid;
    const int64_t dstride = 5 * nthr;

    const V5 one = V5{1.0, 1.0, 0.0, 0.0, 0.0};
    const V5 zero = V5{0.0, 0.0, 0.0, 0.0, 0.0};
    // Every window variable lives in two slots that trade roles on even and
    // odd steps (prev <-> new), so the loop carries no register moves.
    V5 QA = one, QB = one;    // Q_{m-2} in, Q_m out
    V5 U1A = one, U1B = one;  // U1 (floored) of the previous step in, this step's out
    double eA = 1.0, eB = 1.0;  // internal energy the floors left in U1A / U1B
    V5 wA = one, wB = one, dA = zero, dB = zero, RA = one, RB = one, FAA = zero, FAB = zero;
    V5 w1A = one, w1B = one, d1A = zero, d1B = zero, R1A = one, R1B = one, FBA = zero, FBB = zero;

    // Stage B runs one step behind stage A (it consumes the previous step's
    // U1), so the two long dependency chains of a step are independent and
    // interleave. (U1+Q) of a cell is needed three steps after it is formed:
    // it waits in a 4-slot ring in LDS, private to the lane.

    auto step = [&](int t, V5 &Qcur, V5 &Qs, const V5 &U1p, V5 &U1n, const double &ep, double &en,
                    const V5 &wp, V5 &wn, const V5 &dp, V5 &dn, const V5 &Rp, V5 &Rn, const V5 &FAp, V5 &FAn,
                    const V5 &w1p, V5 &w1n, const V5 &d1p, V5 &d1n, const V5 &R1p, V5 &R1n, const V5 &FBp,
                    V5 &FBn) {
        // stage A at cell m-2
        {
            const double e = M::template floor<DIR>(hc, Qcur);
            wn = M::prims(hc, Qcur, e);
        }
        dn = wn - wp;
        V5 Lc;
        M::template face<LIM>(hc, wp, dp, dn, rt, Lc, Rn);  // faces of cell m-1
        FAn = M::flux(hc, Rp, Lc);                           // interface (m-2)+1/2
        U1n = axpy_update(Qs, lambda, FAn, FAp);
        en = M::template floor<DIR>(hc, U1n);
        {
            const V5 S2 = add5(U1n, Qs);  // (U1+Q) of cell m-2, used at step m+3
            double *r = &ring[t & 3][0][lane];
            r[0] = S2.a;
            r[64] = S2.b;
            r[128] = S2.c;
            r[192] = S2.d;
            r[256] = S2.e;
        }
        Qs = Qcur;

        // stage B at cell m-5, from the previous step's U1 (cell m-3)
        w1n = M::prims(hc, U1p, ep);
        d1n = w1n - w1p;
        V5 L1c;
        M::template face<LIM>(hc, w1p, d1p, d1n, rt, L1c, R1n);  // faces of cell m-4
        FBn = M::flux(hc, R1p, L1c);                              // interface (m-5)+1/2
        const double *r = &ring[(t + 1) & 3][0][lane];  // slot of step t-3
        const V5 S5{r[0], r[64], r[128], r[192], r[256]};
        const V5 o = avg_update(S5, lambda, FBn, FBp);  // cell m-5

        // Every step stores exactly 5 doubles: to the cell (outputs 4..L-1), to
        // a deferred slot (outputs 0..3) or to this lane's dummy slot (no
        // output yet / past the segment). With a fixed store count per step the
        // compiler can wait for a prefetched load without waiting for stores
        // (stores share the vector-memory counter on gfx9).
        // outp walks the column one cell per step (it points at output oi);
        // the deferred slots are wave-uniform offsets of this lane's block.
        const int oi = t - 9;  // output index within the segment
        if (CFL && oi >= 0 && oi < L && j0 + oi >= cf.lo && j0 + oi < cf.hi) {
            const double u = cfl_of_cell<M, DIR>(cf.a, o);
            cmin = (u < cmin) ? u : cmin;
            nanc += (u != u);
        }
        const bool direct = (oi >= 4) & (oi < L);
        const int slot = ((oi >= 0) & (oi < 4) & (oi < L)) ? oi : 4;
        double *dq = dfr + (int64_t)slot * dstride;
        double *dst = direct ? outp : dq;
        const int64_t st = direct ? pitch : nthr;
        outp += sstride;
        store_cell<M, DIR, NT>(dst, st, o);
    };

    const int maxsteps = seglen + 9;
    // Prefetch slots for even / odd steps. The loads are unconditional (a lane
    // past its segment, or of an inactive column, reads a valid cell it then
    // ignores: mw always wraps into [0, N) and `col` of an inactive column is
    // element 0), so the compiler neither copies slots nor waits for a load it
    // has just issued.
    V5 N0 = load_cell<M, DIR, NT>(col + (int64_t)mw * sstride, pitch);
    V5 N1;
    // Enter the loop with the same vector-memory counter pattern as the
    // back-edge (load group, then 5 stores) so the wait at the loop head can
    // leave stores outstanding: the stores write the loaded cell into this
    // lane's dummy slot, which also pins them after the load.
    store_cell<M, DIR, NT>(defer + gtid + (int64_t)4 * 5 * nthr, nthr, N0);

    // Steps go in pairs. When maxsteps is odd the last pair's second step is
    // one past the segment: it loads a valid (wrapped) cell and stores to the
    // dummy slot, so running it unconditionally changes no output and keeps
    // the loop to one path (no register copies at the back edge).
    for (int t = 0; t < maxsteps; t += 2) {
        mw = (mw + 1 == N) ? 0 : mw + 1;
        N1 = load_cell<M, DIR, NT>(col + (int64_t)mw * sstride, pitch);
        step(t, N0, QA, U1A, U1B, eA, eB, wA, wB, dA, dB, RA, RB, FAA, FAB, w1A, w1B, d1A, d1B, R1A, R1B, FBA,
             FBB);
        mw = (mw + 1 == N) ? 0 : mw + 1;
        N0 = load_cell<M, DIR, NT>(col + (int64_t)mw * sstride, pitch);
        step(t + 1, N1, QB, U1B, U1A, eB, eA, wB, wA, dB, dA, RB, RA, FAB, FAA, w1B, w1A, d1B, d1A, R1B, R1A, FBB,
             FBA);
    }

    // every lane of the wave has finished reading: release the deferred outputs
    const int nd = L < 4 ? L : 4;
    for (int oi = 0; oi < nd; oi++) {
        const V5 o = load_cell<M, DIR, NT>(defer + gtid + (int64_t)oi * 5 * nthr, nthr);
        store_cell<M, DIR, NT>(col + (int64_t)(j0 + oi) * sstride, pitch, o);
    }
}

template <class M, int DIR, int LIM, bool CFL, bool NT>
__global__ __launch_bounds__(256) void sweep_march(double *__restrict__ q, int64_t pitch, int nx, int64_t sstride,
                                                   int64_t ostride, int N, int S, int seglen, int wavesPerRow,
                                                   double lambda, const double *__restrict__ lamp,
                                                   HydroConsts hc, int rt, double *__restrict__ defer, CflFuse cf)
{
    if (lamp) lambda = lamp[0];  // time step kept on the device
    const int lane = threadIdx.x & 63;
    const int wib = threadIdx.x >> 6;
    const int64_t waveId = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * (blockDim.x >> 6) + wib;
    const int W = 64 / S;
    const int s = lane / W;
    const int c = lane - s * W;
    const int64_t wrow = (int64_t)blockIdx.x * (blockDim.x >> 6) + wib;  // wave index within this row
    if (wrow >= wavesPerRow) {
        if (CFL && lane == 0) cf.part[waveId] = 1e38;
        return;
    }
    const int64_t x = wrow * W + c;
    const bool colValid = (x < nx) && (s < S);
    // lanes of no column (x past the grid, or s >= S when S does not divide
    // 64) march a zero-length segment from row 0: their unconditional loads
    // then stay inside the column
    const int j0 = colValid ? s * seglen : 0;
    int L = N - j0;
    L = L < seglen ? L : seglen;
    if (!colValid) L = 0;

    double *col = q + (colValid ? x + (int64_t)blockIdx.y * ostride : 0);
    const int64_t gtid = ((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nthr = (int64_t)gridDim.x * gridDim.y * blockDim.x;
    __shared__ double sring[4][4][5][64];  // [wave in block][slot][variable][lane]
    double cmin = 1e38;
    unsigned nanc = 0;  // not reported by the single-array fold
    march_body<M, DIR, LIM, CFL, NT>(col, pitch, sstride, N, seglen, j0, L, lambda, hc, rt, defer, gtid, nthr,
                                     sring[wib], lane, cf, cmin, nanc);
    if (CFL) {
        cmin = wave_min(cmin);
        if (lane == 0) cf.part[waveId] = cmin;
    }
}

}  // namespace imogen
