/*
 * imogen_mex.h -- MATLAB-side helpers shared by the MEX gateways in mex/.
 *
 * The gateways keep the reference's MATLAB call forms (the mexFunction
 * bodies of gpuclass/ in the reference) and call libimogen_hip.so through include/imogen_abi.h.
 * These helpers read the MATLAB objects the reference's gateways read
 * (gpuclass/cudaCommonML.cu): GPU tags stored as int64 vectors, as
 * GPU_Type.GPU_MemPtr or as ImogenArray.gputag; a fluid as
 * FluidManager.mass / .ener / .mom; boundary modes from
 * ImogenArray.boundaryData.bcModes; ThermoDetails from the fluid's
 * thermoDetails struct; GeometryParams from a GeometryManager.
 *
 * Errors: every helper returns an Imogen status; imogen_mex_fail() turns a
 * non-zero status into mexErrMsgIdAndTxt with the library's last message,
 * the role of DROP_MEX_ERROR in the reference (gpuclass/cudaCommonML.cu:578).
 */
#ifndef IMOGEN_MEX_H
#define IMOGEN_MEX_H

#include "mex.h"
#include "imogen_abi.h"

/* Copies the tag held by m (an int64 vector, a GPU_Type or an ImogenArray;
 * element idx of an object array) into tag[GPU_TAG_MAXLEN]. */
int imogen_mex_tag(const mxArray *m, int idx, int64_t *tag);

/* MGArray of argument m (element idx), with its boundary modes when m carries
 * boundaryData.bcModes, and its ImogenBoundaryData (externalData) when
 * boundaryData holds a gravity potential or static cells. */
int imogen_mex_array(const mxArray *m, int idx, MGArray *out);

/* The five slab views of fluid fluidIdx of a FluidManager (array). */
int imogen_mex_fluid(const mxArray *fluidManager, int fluidIdx, MGArray fluid[5]);

/* ThermoDetails of fluid fluidIdx (FluidManager.thermoDetails). */
int imogen_mex_thermo(const mxArray *fluidManager, int fluidIdx, ThermoDetails *t);

/* GeometryParams of a GeometryManager, globalRez from globalDomainRez. */
int imogen_mex_geometry(const mxArray *geometryManager, GeometryParams *g);

/* One-rank ParallelTopology (the GeometryManager's topology is not read:
 * this library runs one GPU per process). */
void imogen_mex_topology(ParallelTopology *t);

/* Double-valued scalar property or struct field; *ok = 0 when absent. */
double imogen_mex_scalar(const mxArray *m, int idx, const char *name, int *ok);

/* A new int64 column holding tag (serialized MGArray of one GPU). */
mxArray *imogen_mex_tag_array(const int64_t *tag);

/* Aborts the MEX call (mexErrMsgIdAndTxt) when status is not SUCCESSFUL. */
void imogen_mex_fail(int status, const char *gateway);

/* Aborts with a usage message. */
void imogen_mex_usage(const char *gateway, const char *form);

#endif
