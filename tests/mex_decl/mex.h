/* Declarations of the MATLAB C Matrix / MEX API functions the gateways in
 * mex/ call, as MathWorks documents them (MATLAB External Interfaces, C
 * Matrix API). Declarations only: the CPU test suite compiles the gateways
 * against this header to check them (g++ -fsyntax-only); a real build uses
 * MATLAB's own mex.h through mex/Makefile. */
#ifndef IMOGEN_TEST_MEX_DECL_H
#define IMOGEN_TEST_MEX_DECL_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef enum {
    mxUNKNOWN_CLASS, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS, mxVOID_CLASS, mxDOUBLE_CLASS,
    mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS, mxUINT16_CLASS, mxINT32_CLASS, mxUINT32_CLASS,
    mxINT64_CLASS, mxUINT64_CLASS, mxFUNCTION_CLASS
} mxClassID;
typedef enum { mxREAL, mxCOMPLEX } mxComplexity;

double *mxGetPr(const mxArray *pm);
void *mxGetData(const mxArray *pm);
double mxGetScalar(const mxArray *pm);
size_t mxGetNumberOfElements(const mxArray *pm);
mwSize mxGetNumberOfDimensions(const mxArray *pm);
const mwSize *mxGetDimensions(const mxArray *pm);
mxClassID mxGetClassID(const mxArray *pm);
const char *mxGetClassName(const mxArray *pm);
bool mxIsDouble(const mxArray *pm);
bool mxIsComplex(const mxArray *pm);
bool mxIsStruct(const mxArray *pm);
bool mxIsCell(const mxArray *pm);
bool mxIsChar(const mxArray *pm);
mxArray *mxGetField(const mxArray *pm, mwIndex index, const char *fieldname);
mxArray *mxGetProperty(const mxArray *pa, mwIndex index, const char *propname);
void mxSetProperty(mxArray *pa, mwIndex index, const char *propname, const mxArray *value);
mxArray *mxGetCell(const mxArray *pm, mwIndex index);
int mxGetString(const mxArray *pm, char *str, mwSize strlen);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity complexFlag);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity complexFlag);
mxArray *mxCreateNumericArray(mwSize ndim, const mwSize *dims, mxClassID classid, mxComplexity complexFlag);
mxArray *mxCreateString(const char *str);
void mxDestroyArray(mxArray *pm);
int mexPrintf(const char *message, ...);
void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...);
#ifdef __cplusplus
}
#endif
#endif
