// imogen_core -- standalone C++ driver of libimogen_hip.so (no MATLAB, no Python).
//
// The counterpart of the reference's imogenCore (gpuclass/core/gpuimogen_core.cu):
// it builds the fluid arrays through the same tag primitives the MATLAB side
// uses (GPU_upload, GPU_makeslab, GPU_setslab, GPU_getslab), then runs the
// iteration of performCompleteTimestep (gpuimogen_core.cu:954-1054):
//     tau = min over fluids of the CFL time step; dt = cfl * tau
//     source(dt/2); performFluidUpdate_3D(+1) of every fluid; source(dt);
//     performFluidUpdate_3D(-1) of every fluid; source(dt/2)
// where source is performSourceFunctions (sourceStep.cu:200-226): nothing for
// one fluid, the two-fluid drag (type 8, one sourcefunction_2FluidDrag call
// over the interval) for gas + dust.
//
//   imogen_core --problem sedov|sod|gasdust [--n N] [--steps K]
//               [--method hllc|hll|xinjin] [--limiter 0..6] [--cfl C]
//               [--drag M] [--thermo FILE] [--ic FILE] [--out FILE]
//
// --ic reads the initial state as raw little-endian doubles shaped
// (5, nz, ny, nx) with nx = ny = nz = N (sedov) or nx = N, ny = nz = 1 (sod);
// gasdust (required --ic) reads (2, 5, N, N, N): gas then dust, periodic
// boundaries, drag method M (default 4, LogTrap2), and --thermo the two
// ThermoDetails records (gas, dust) as 16 doubles in field order. --out writes
// the final state in the input layout. The last line printed is one JSON
// object with the timing and conserved totals.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "imogen_abi.h"

namespace {

struct Fluid {
    int64_t owner[GPU_TAG_MAXLEN] = {0};
    int64_t slab[5][GPU_TAG_MAXLEN] = {{0}};
    MGArray arr[5];
};

int check(int rc, const char *what)
{
    if (rc != SUCCESSFUL) {
        fprintf(stderr, "imogen_core: %s failed (%d): %s\n", what, rc, imogen_last_error());
        exit(1);
    }
    return rc;
}

// FluidManager's DataHolder: a 5-slab owner whose slab views are the fluid arrays
// (save/uploadDataArrays.m:93-111)
void upload_fluid(Fluid &f, const std::vector<double> &state, const int dims[3], const int bc[6])
{
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    int64_t first[GPU_TAG_MAXLEN] = {0};
    check(GPU_upload(state.data(), dims, 4, PARTITION_X, first), "GPU_upload");
    check(GPU_makeslab(first, 5, f.owner), "GPU_makeslab");
    check(GPU_getslab(f.owner, 0, f.slab[0]), "GPU_getslab");
    for (int k = 1; k < 5; k++) check(GPU_setslab(f.owner, k, state.data() + k * n, f.slab[k]), "GPU_setslab");
    for (int k = 0; k < 5; k++) {
        check(deserializeTagToMGArray(f.slab[k], &f.arr[k]), "deserializeTagToMGArray");
        f.arr[k].vectorComponent = (k < 2) ? 0 : k - 1;
        for (int s = 0; s < 6; s++) f.arr[k].boundaryConditions.mode[s] = (BCModeTypes)bc[s];
    }
}

std::vector<double> download_fluid(Fluid &f, const int dims[3])
{
    const size_t n = (size_t)dims[0] * dims[1] * dims[2];
    std::vector<double> out(5 * n);
    for (int k = 0; k < 5; k++) check(GPU_download(f.slab[k], out.data() + k * n), "GPU_download");
    return out;
}

// Sedov-Taylor blast on the unit cube: rho = 1, v = 0, P = 1e-5, E = 1 spread
// over the cells within r^2 <= 9 (integer offsets) of the centre cell
// floor(N/2 + 0.5) (1-based), gamma = 5/3
std::vector<double> sedov(int n, double gamma, int *ndep)
{
    const size_t N = (size_t)n * n * n;
    std::vector<double> q(5 * N, 0.0);
    const double dx = 1.0 / n;
    for (size_t i = 0; i < N; i++) {
        q[i] = 1.0;
        q[N + i] = 1e-5 / (gamma - 1);
    }
    const int c = (int)std::floor(n / 2.0 + 0.5) - 1;
    int cnt = 0;
    for (int k = -3; k <= 3; k++)
        for (int j = -3; j <= 3; j++)
            for (int i = -3; i <= 3; i++)
                if (i * i + j * j + k * k <= 9) cnt++;
    const double e = 1.0 / (cnt * dx * dx * dx);
    for (int k = -3; k <= 3; k++)
        for (int j = -3; j <= 3; j++)
            for (int i = -3; i <= 3; i++)
                if (i * i + j * j + k * k <= 9)
                    q[N + (size_t)(c + k) * n * n + (size_t)(c + j) * n + (c + i)] = e;
    *ndep = cnt;
    return q;
}

// Sod shock tube on [0, 1], gamma = 1.4
std::vector<double> sod(int n, double gamma)
{
    std::vector<double> q(5 * (size_t)n, 0.0);
    for (int i = 0; i < n; i++) {
        const double x = (i + 0.5) / n;
        q[i] = x < 0.5 ? 1.0 : 0.125;
        q[n + i] = (x < 0.5 ? 1.0 : 0.1) / (gamma - 1);
    }
    return q;
}

}  // namespace

int main(int argc, char **argv)
{
    std::string problem = "sedov", method = "hllc", icfile, outfile, thermofile;
    int n = 64, steps = 10, limiter = LIMITER_OSPRE, drag = DRAG_LOGTRAP2;
    double cfl = 0.4;
    for (int i = 1; i < argc; i++) {
        std::string a = argv[i];
        auto next = [&]() -> const char * {
            if (i + 1 >= argc) {
                fprintf(stderr, "missing value for %s\n", a.c_str());
                exit(2);
            }
            return argv[++i];
        };
        if (a == "--problem") problem = next();
        else if (a == "--n") n = atoi(next());
        else if (a == "--steps") steps = atoi(next());
        else if (a == "--method") method = next();
        else if (a == "--limiter") limiter = atoi(next());
        else if (a == "--cfl") cfl = atof(next());
        else if (a == "--ic") icfile = next();
        else if (a == "--out") outfile = next();
        else if (a == "--drag") drag = atoi(next());
        else if (a == "--thermo") thermofile = next();
        else {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            return 2;
        }
    }
    FluidMethods fm = METHOD_HLLC;
    if (method == "hll") fm = METHOD_HLL;
    else if (method == "xinjin") fm = METHOD_XINJIN;
    else if (method != "hllc") {
        fprintf(stderr, "unknown method %s\n", method.c_str());
        return 2;
    }

    int dims[3];
    double gamma;
    std::vector<double> q;
    int bc[6];
    int ndep = 0;
    int nfluids = 1;
    ThermoDetails thermo[2];
    memset(thermo, 0, sizeof(thermo));
    if (problem == "gasdust") {
        dims[0] = dims[1] = dims[2] = n;
        nfluids = 2;
        for (int s = 0; s < 6; s++) bc[s] = circular;
        FILE *fh = thermofile.empty() ? nullptr : fopen(thermofile.c_str(), "rb");
        if (!fh || fread(thermo, sizeof(ThermoDetails), 2, fh) != 2) {
            fprintf(stderr, "gasdust needs --thermo with 2 ThermoDetails records\n");
            return 1;
        }
        fclose(fh);
        gamma = thermo[0].gamma;
        q.assign(2 * 5 * (size_t)n * n * n, 0.0);
        if (icfile.empty()) {
            fprintf(stderr, "gasdust needs --ic\n");
            return 1;
        }
    } else if (problem == "sedov") {
        dims[0] = dims[1] = dims[2] = n;
        gamma = 5.0 / 3.0;
        q = sedov(n, gamma, &ndep);
        for (int s = 0; s < 6; s++) bc[s] = extrapConstant;
    } else if (problem == "sod") {
        dims[0] = n;
        dims[1] = dims[2] = 1;
        gamma = 1.4;
        q = sod(n, gamma);
        for (int s = 0; s < 6; s++) bc[s] = extrapConstant;
    } else {
        fprintf(stderr, "unknown problem %s\n", problem.c_str());
        return 2;
    }
    if (!icfile.empty()) {
        FILE *fh = fopen(icfile.c_str(), "rb");
        if (!fh || fread(q.data(), sizeof(double), q.size(), fh) != q.size()) {
            fprintf(stderr, "cannot read %zu doubles from %s\n", q.size(), icfile.c_str());
            return 1;
        }
        fclose(fh);
    }

    const size_t N = (size_t)dims[0] * dims[1] * dims[2];
    Fluid fl[2];
    for (int k = 0; k < nfluids; k++) {
        std::vector<double> part(q.begin() + k * 5 * N, q.begin() + (k + 1) * 5 * N);
        upload_fluid(fl[k], part, dims, bc);
    }
    const double gam[2] = {gamma, nfluids > 1 ? thermo[1].gamma : gamma};

    GeometryParams geo;
    memset(&geo, 0, sizeof(geo));
    geo.shape = SQUARE;
    for (int d = 0; d < 3; d++) {
        geo.h[d] = (d == 0 || dims[d] > 1) ? 1.0 / dims[0] : 1.0;
        geo.globalRez[d] = geo.localRez[d] = dims[d];
    }
    double rez[3] = {(double)dims[0], (double)dims[1], (double)dims[2]};
    ParallelTopology topo;
    memset(&topo, 0, sizeof(topo));
    topo.ndim = 3;
    topo.nproc[0] = topo.nproc[1] = topo.nproc[2] = 1;

    FluidStepParams p;
    memset(&p, 0, sizeof(p));
    p.onlyHydro = 1;
    p.thermoGamma = gamma;
    p.minimumRho = 1e-8;
    p.Cisothermal = -1;
    p.stepMethod = fm;
    p.geometry = geo;
    p.slopeLimiter = limiter;

    // the fluids as performSourceFunctions sees them (GridFluid, cudaCommon.h:224-234)
    GridFluid gf[2];
    memset(gf, 0, sizeof(gf));
    for (int k = 0; k < nfluids; k++) {
        gf[k].DataHolder = fl[k].arr[0];
        for (int s = 0; s < 5; s++) {
            gf[k].data[s] = fl[k].arr[s];
            gf[k].bcond[s] = fl[k].arr[s].boundaryConditions;
        }
        gf[k].thermo = thermo[k];
        gf[k].rhoMin = 1e-8;
    }
    ParametricRadiation rad;
    memset(&rad, 0, sizeof(rad));
    const int srcType = (nfluids > 1) ? 8 : 0;
    p.multifluidDragMethod = drag;
    auto source = [&](double dt) {
        FluidStepParams sp = p;
        sp.dt = dt;
        check(performSourceFunctions(srcType, gf, nfluids, sp, &topo, nullptr, &rad, nullptr), "performSourceFunctions");
    };
    auto fluid_step = [&](int direction) {
        for (int k = 0; k < nfluids; k++) {
            p.thermoGamma = gam[k];
            p.stepDirection = direction;
            check(performFluidUpdate_3D(fl[k].arr, &topo, p, nullptr), "performFluidUpdate_3D");
        }
    };

    double t = 0.0;
    check(imogen_synchronize(), "imogen_synchronize");
    const auto t0 = std::chrono::steady_clock::now();
    for (int it = 0; it < steps; it++) {
        double tau = 0.0;
        for (int k = 0; k < nfluids; k++) {
            double tk = 0.0;
            check(imogen_cflTimestepFused(fl[k].arr, &geo, fm, gam[k], rez, &tk), "imogen_cflTimestepFused");
            tau = (k == 0 || tk < tau) ? tk : tau;
        }
        p.dt = cfl * tau;
        p.stepNumber = it;
        source(0.5 * p.dt);
        fluid_step(1);
        source(p.dt);
        fluid_step(-1);
        source(0.5 * p.dt);
        t += 2 * p.dt;
    }
    check(imogen_synchronize(), "imogen_synchronize");
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

    std::vector<double> out;
    for (int k = 0; k < nfluids; k++) {
        std::vector<double> o = download_fluid(fl[k], dims);
        out.insert(out.end(), o.begin(), o.end());
    }
    double mass = 0, energy = 0;
    for (size_t i = 0; i < N; i++) {
        mass += out[i];
        energy += out[N + i];
    }
    if (!outfile.empty()) {
        FILE *fh = fopen(outfile.c_str(), "wb");
        if (!fh || fwrite(out.data(), sizeof(double), out.size(), fh) != out.size()) {
            fprintf(stderr, "cannot write %s\n", outfile.c_str());
            return 1;
        }
        fclose(fh);
    }
    for (int k = 0; k < nfluids; k++) check(GPU_free(fl[k].owner), "GPU_free");
    const double cu = (double)N * 2 * steps;
    printf("{\"problem\": \"%s\", \"method\": \"%s\", \"dims\": [%d, %d, %d], \"steps\": %d, \"seconds\": %.6f, "
           "\"cell_updates_per_s\": %.6e, \"sim_time\": %.9e, \"mass\": %.17g, \"energy\": %.17g, "
           "\"deposit_cells\": %d, \"fluids\": %d}\n",
           problem.c_str(), method.c_str(), dims[0], dims[1], dims[2], steps, secs, cu / secs, t, mass, energy,
           ndep, nfluids);
    return 0;
}
