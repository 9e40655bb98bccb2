// imogen_mex.cpp -- MATLAB object access for the MEX gateways (see imogen_mex.h).
//
// Property and field names are the MATLAB classes' own (gpuclass/GPU_Type.m,
// ImogenArray, fluid/FluidManager.m, GeometryManager.m); the lookup order of a
// tag follows gpuclass/cudaCommonML.cu:19-87 (int64 class, GPU_Type.GPU_MemPtr,
// ImogenArray.gputag) and the boundary-mode names :203-237.
#include "imogen_mex.h"

#include <cstdio>
#include <cstring>

static int set_fail(int status, const char *what)
{
    mexPrintf("imogen MEX: %s\n", what);
    return status;
}

// mxGetProperty returns a copy the caller destroys; a struct field is owned
// by its struct. Prop holds either and frees only what it owns.
struct Prop {
    mxArray *v = nullptr;
    bool owned = false;
    Prop(const mxArray *m, int idx, const char *name)
    {
        if (!m) return;
        if (mxIsStruct(m)) {
            v = mxGetField(m, (mwIndex)idx, name);
        } else {
            v = mxGetProperty(m, (mwIndex)idx, name);
            owned = true;
        }
    }
    ~Prop()
    {
        if (owned && v) mxDestroyArray(v);
    }
    Prop(const Prop &) = delete;
    Prop &operator=(const Prop &) = delete;
};

double imogen_mex_scalar(const mxArray *m, int idx, const char *name, int *ok)
{
    if (ok) *ok = 0;
    Prop p(m, idx, name);
    if (!p.v || !mxIsDouble(p.v) || mxGetNumberOfElements(p.v) < 1) return 0.0;
    if (ok) *ok = 1;
    return mxGetPr(p.v)[0];
}

static int tag_from_int64(const mxArray *t, int64_t *tag)
{
    const size_t n = mxGetNumberOfElements(t);
    if (n < GPU_TAG_LENGTH || n > GPU_TAG_MAXLEN)
        return set_fail(ERROR_GET_GPUTAG_FAILED, "a GPU tag holds 11 + 2 nGPUs int64 values");
    const int64_t *q = (const int64_t *)mxGetData(t);
    if (n != (size_t)(GPU_TAG_LENGTH + 2 * q[GPU_TAG_NGPUS]))
        return set_fail(ERROR_GET_GPUTAG_FAILED, "GPU tag length does not match its GPU count");
    memcpy(tag, q, n * sizeof(int64_t));
    return SUCCESSFUL;
}

int imogen_mex_tag(const mxArray *m, int idx, int64_t *tag)
{
    if (!m || !tag) return set_fail(ERROR_NULL_POINTER, "null argument");
    if (mxGetClassID(m) == mxINT64_CLASS) return tag_from_int64(m, tag);
    const char *cls = mxGetClassName(m);
    Prop t(m, idx, (cls && strcmp(cls, "GPU_Type") == 0) ? "GPU_MemPtr" : "gputag");
    if (!t.v || mxGetClassID(t.v) != mxINT64_CLASS)
        return set_fail(ERROR_GET_GPUTAG_FAILED, "argument is not a GPU tag, GPU_Type or ImogenArray");
    return tag_from_int64(t.v, tag);
}

static int bc_mode_of(const char *s, BCModeTypes *mode)
{
    static const struct {
        const char *name;
        BCModeTypes mode;
    } table[] = {{"circ", circular},   {"const", extrapConstant}, {"linear", extrapLinear},
                 {"mirror", mirror},   {"bcstatic", stationary},  {"wall", wall},
                 {"outflow", outflow}, {"freebalance", freebalance}};
    for (const auto &e : table)
        if (strcmp(s, e.name) == 0) {
            *mode = e.mode;
            return SUCCESSFUL;
        }
    return ERROR_INVALID_ARGS;
}

// ImogenArray.boundaryData.bcModes: a 2 x 3 cell of mode names, (side, direction)
static int read_bc_modes(const mxArray *m, int idx, BCSettings *b)
{
    if (mxGetClassID(m) == mxINT64_CLASS || mxIsStruct(m)) return SUCCESSFUL;
    Prop bd(m, idx, "boundaryData");
    if (!bd.v || !mxIsStruct(bd.v)) return SUCCESSFUL;  // a plain GPU_Type carries none
    const mxArray *modes = mxGetField(bd.v, 0, "bcModes");
    if (!modes || !mxIsCell(modes) || mxGetNumberOfElements(modes) < 6) return SUCCESSFUL;
    for (int k = 0; k < 6; k++) {
        const mxArray *s = mxGetCell(modes, (mwIndex)k);
        char name[32];
        if (!s || !mxIsChar(s) || mxGetString(s, name, sizeof(name)) != 0)
            return set_fail(ERROR_INVALID_ARGS, "boundaryData.bcModes holds a non-string entry");
        if (bc_mode_of(name, &b->mode[k]) != SUCCESSFUL) {
            mexPrintf("imogen MEX: unknown boundary mode '%s'\n", name);
            return ERROR_INVALID_ARGS;
        }
    }
    return SUCCESSFUL;
}

// The boundaryData fields the library reads through externalData (the
// freebalance potential and gamma, the static cells), for one array, kept in a
// ring of slots that outlives the gateway call that fills it (a gateway
// passes at most a few fluids' 5 arrays).
struct BoundaryDataSlot {
    ImogenBoundaryData bd;
    MGArray gravpot, statics;
};
static BoundaryDataSlot g_bdSlots[64];
static int g_bdNext = 0;

static int array_of(const mxArray *m, MGArray *out)
{
    int64_t tag[GPU_TAG_MAXLEN];
    int rc = imogen_mex_tag(m, 0, tag);
    return rc ? rc : deserializeTagToMGArray(tag, out);
}

// boundaryData.gravpot.field and .gamma (cudaStatics.cu:165-176), .staticsData
// and .compOffset (:302-321)
static int read_boundary_data(const mxArray *m, int idx, BCSettings *b)
{
    if (mxGetClassID(m) == mxINT64_CLASS || mxIsStruct(m)) return SUCCESSFUL;
    Prop bd(m, idx, "boundaryData");
    if (!bd.v || !mxIsStruct(bd.v)) return SUCCESSFUL;
    const mxArray *gp = mxGetField(bd.v, 0, "gravpot");
    const mxArray *st = mxGetField(bd.v, 0, "staticsData");
    const mxArray *co = mxGetField(bd.v, 0, "compOffset");
    const bool grav = gp && mxGetNumberOfElements(gp) > 0;
    const bool stat = st && mxGetNumberOfElements(st) > 0 && co && mxIsDouble(co) && mxGetNumberOfElements(co) >= 12;
    if (!grav && !stat) return SUCCESSFUL;
    BoundaryDataSlot &s = g_bdSlots[g_bdNext];
    g_bdNext = (g_bdNext + 1) % 64;
    memset(&s, 0, sizeof(s));
    if (grav) {
        Prop field(gp, 0, "field");
        int rc = array_of(field.v ? field.v : gp, &s.gravpot);
        if (rc) return rc;
        int ok;
        s.bd.gravpot = &s.gravpot;
        s.bd.gamma = imogen_mex_scalar(bd.v, 0, "gamma", &ok);
        if (!ok) return set_fail(ERROR_INVALID_ARGS, "boundaryData.gamma missing (freebalance)");
    }
    if (stat) {
        int rc = array_of(st, &s.statics);
        if (rc) return rc;
        s.bd.statics = &s.statics;
        size_t n = mxGetNumberOfElements(co);
        if (n > 12 * MAX_GPUS_USED) n = 12 * MAX_GPUS_USED;
        for (size_t k = 0; k < n; k++) s.bd.compOffset[k] = mxGetPr(co)[k];
    }
    b->externalData = &s.bd;
    return SUCCESSFUL;
}

int imogen_mex_array(const mxArray *m, int idx, MGArray *out)
{
    int64_t tag[GPU_TAG_MAXLEN];
    int rc = imogen_mex_tag(m, idx, tag);
    if (rc) return rc;
    rc = deserializeTagToMGArray(tag, out);
    if (rc) return rc;
    out->boundaryConditions.externalData = nullptr;
    out->boundaryConditions.extIndex = idx;
    rc = read_bc_modes(m, idx, &out->boundaryConditions);
    return rc ? rc : read_boundary_data(m, idx, &out->boundaryConditions);
}

int imogen_mex_fluid(const mxArray *fm, int fluidIdx, MGArray fluid[5])
{
    if (!fm) return set_fail(ERROR_NULL_POINTER, "null FluidManager");
    Prop mass(fm, fluidIdx, "mass"), ener(fm, fluidIdx, "ener"), mom(fm, fluidIdx, "mom");
    if (!mass.v || !ener.v || !mom.v)
        return set_fail(ERROR_INVALID_ARGS, "argument is not a FluidManager (mass / ener / mom)");
    if (mxGetNumberOfElements(mom.v) < 3) return set_fail(ERROR_INVALID_ARGS, "FluidManager.mom has fewer than 3 arrays");
    int rc = imogen_mex_array(mass.v, 0, &fluid[0]);
    if (!rc) rc = imogen_mex_array(ener.v, 0, &fluid[1]);
    for (int k = 0; k < 3 && !rc; k++) rc = imogen_mex_array(mom.v, k, &fluid[2 + k]);
    return rc;
}

int imogen_mex_thermo(const mxArray *fm, int fluidIdx, ThermoDetails *t)
{
    Prop tdp(fm, fluidIdx, "thermoDetails");
    const mxArray *td = tdp.v;
    if (!td || !mxIsStruct(td)) return set_fail(ERROR_INVALID_ARGS, "fluid has no thermoDetails struct");
    static const char *names[] = {"gamma", "mass", "Cisothermal", "kBolt", "dynViscosity", "viscTindex",
                                  "sigma", "sigmaTindex"};
    double v[8];
    for (int k = 0; k < 8; k++) {
        int ok;
        v[k] = imogen_mex_scalar(td, 0, names[k], &ok);
        if (!ok) {
            mexPrintf("imogen MEX: thermoDetails.%s missing\n", names[k]);
            return ERROR_INVALID_ARGS;
        }
    }
    t->gamma = v[0];
    t->m = v[1];
    t->Cisothermal = v[2];
    t->kBolt = v[3];
    t->mu0 = v[4];
    t->muTindex = v[5];
    t->sigma0 = v[6];
    t->sigmaTindex = v[7];
    return SUCCESSFUL;
}

static int vector3(const mxArray *m, const char *name, double *out)
{
    Prop p(m, 0, name);
    if (!p.v || !mxIsDouble(p.v) || mxGetNumberOfElements(p.v) < 1) return ERROR_INVALID_ARGS;
    const size_t n = mxGetNumberOfElements(p.v);
    for (size_t k = 0; k < 3; k++) out[k] = mxGetPr(p.v)[k < n ? k : n - 1];
    return SUCCESSFUL;
}

int imogen_mex_geometry(const mxArray *gm, GeometryParams *g)
{
    if (!gm) return set_fail(ERROR_NULL_POINTER, "null GeometryManager");
    memset(g, 0, sizeof(*g));
    int ok;
    const double shape = imogen_mex_scalar(gm, 0, "pGeometryType", &ok);
    if (!ok) return set_fail(ERROR_INVALID_ARGS, "GeometryManager.pGeometryType missing");
    g->shape = (shape == 2) ? CYLINDRICAL : SQUARE;  // GeometryManager: 1 square, 2 cylindrical
    if (vector3(gm, "d3h", g->h)) return set_fail(ERROR_INVALID_ARGS, "GeometryManager.d3h missing");
    double a[3] = {0, 0, 0}, rez[3];
    (void)vector3(gm, "affine", a);
    g->x0 = a[0];
    g->y0 = a[1];
    g->z0 = a[2];
    g->Rinner = imogen_mex_scalar(gm, 0, "pInnerRadius", &ok);
    (void)vector3(gm, "frameRotationCenter", g->frameRotateCenter);
    g->frameOmega = imogen_mex_scalar(gm, 0, "frameRotationOmega", &ok);
    if (vector3(gm, "globalDomainRez", rez)) return set_fail(ERROR_INVALID_ARGS, "GeometryManager.globalDomainRez missing");
    for (int k = 0; k < 3; k++) g->globalRez[k] = g->localRez[k] = (int)rez[k];
    return SUCCESSFUL;
}

void imogen_mex_topology(ParallelTopology *t)
{
    memset(t, 0, sizeof(*t));
    t->ndim = 3;
    for (int k = 0; k < 3; k++) t->nproc[k] = 1;
}

mxArray *imogen_mex_tag_array(const int64_t *tag)
{
    const mwSize n = (mwSize)(GPU_TAG_LENGTH + 2 * tag[GPU_TAG_NGPUS]);
    mxArray *r = mxCreateNumericMatrix(n, 1, mxINT64_CLASS, mxREAL);
    memcpy(mxGetData(r), tag, (size_t)n * sizeof(int64_t));
    return r;
}

void imogen_mex_fail(int status, const char *gateway)
{
    if (status == SUCCESSFUL) return;
    char msg[1200];
    snprintf(msg, sizeof(msg), "%s: status %d: %s", gateway, status, imogen_last_error());
    mexErrMsgIdAndTxt("imogen:compiledCode", "%s", msg);
}

void imogen_mex_usage(const char *gateway, const char *form)
{
    mexErrMsgIdAndTxt("imogen:usage", "%s: call as %s", gateway, form);
}
