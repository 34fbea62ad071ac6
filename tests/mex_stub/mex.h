/*
 * mex.h -- the subset of MATLAB's C MEX and C Matrix API (interleaved-complex
 * API, "mex -R2018a") that gerjoii_amd/mex/gjw_mex.c uses, declared from the
 * names and signatures in MathWorks' published documentation.  It exists so
 * that the gateway compiles here, where MATLAB is absent, and is driven by the
 * test-side stand-in mx_stub.c.  On a MATLAB host the gateway is built with
 * MathWorks' own mex.h instead (`mex -R2018a gjw_mex.c ...`); nothing in this
 * file is needed there.
 *
 * Enumerator values are this stand-in's own: the gateway uses the names only.
 */
#ifndef GJW_MEX_SUBSET_H
#define GJW_MEX_SUBSET_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef size_t mwIndex;
typedef double mxDouble;
typedef float mxSingle;
typedef int32_t mxInt32;
typedef uint32_t mxUint32;

typedef enum {
  mxUNKNOWN_CLASS = 0,
  mxCELL_CLASS,
  mxSTRUCT_CLASS,
  mxLOGICAL_CLASS,
  mxCHAR_CLASS,
  mxVOID_CLASS,
  mxDOUBLE_CLASS,
  mxSINGLE_CLASS,
  mxINT8_CLASS,
  mxUINT8_CLASS,
  mxINT16_CLASS,
  mxUINT16_CLASS,
  mxINT32_CLASS,
  mxUINT32_CLASS,
  mxINT64_CLASS,
  mxUINT64_CLASS,
  mxFUNCTION_CLASS
} mxClassID;

typedef enum { mxREAL = 0, mxCOMPLEX } mxComplexity;

/* C Matrix API */
mxArray *mxGetField(const mxArray *pm, mwIndex index, const char *fieldname);
double mxGetScalar(const mxArray *pm);
size_t mxGetM(const mxArray *pm);
size_t mxGetN(const mxArray *pm);
size_t mxGetNumberOfElements(const mxArray *pm);
mxClassID mxGetClassID(const mxArray *pm);
bool mxIsStruct(const mxArray *pm);
bool mxIsChar(const mxArray *pm);
bool mxIsDouble(const mxArray *pm);
bool mxIsSingle(const mxArray *pm);
void *mxGetData(const mxArray *pm);
mxDouble *mxGetDoubles(const mxArray *pa);
mxSingle *mxGetSingles(const mxArray *pa);
mxInt32 *mxGetInt32s(const mxArray *pa);
int mxGetString(const mxArray *pm, char *str, mwSize strlen);
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray *mxDuplicateArray(const mxArray *in);
void mxDestroyArray(mxArray *pm);
void *mxMalloc(mwSize n);
void mxFree(void *ptr);

/* C MEX API */
void mexErrMsgIdAndTxt(const char *errorid, const char *errormsg, ...);
int mexAtExit(void (*ExitFcn)(void));
void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]);

#ifdef __cplusplus
}
#endif
#endif
