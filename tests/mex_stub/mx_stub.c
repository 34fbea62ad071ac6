/*
 * mx_stub.c -- TEST INFRASTRUCTURE: a stand-in for MATLAB's mxArray runtime,
 * enough of it to drive gerjoii_amd/mex/gjw_mex.c from Python (ctypes)
 * through the real libgjw C ABI.  Implements the functions tests/mex_stub/mex.h
 * declares with the behaviour MathWorks documents for them (column-major
 * numeric matrices, 1 x 1 structs, char arrays; mexErrMsgIdAndTxt does not
 * return; mxMalloc memory not freed by the MEX function is released when the
 * call ends), plus a small driver API (stub_*) for the tests.
 */
#define _POSIX_C_SOURCE 200809L  /* strdup */
#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"

#define MAXF 32

struct mxArray_tag {
  mxClassID cls;
  size_t m, n;
  void *data;
  int nf;
  char *fname[MAXF];
  mxArray *fval[MAXF];
};

static size_t esize(mxClassID c) {
  switch (c) {
    case mxDOUBLE_CLASS: case mxINT64_CLASS: case mxUINT64_CLASS: return 8;
    case mxSINGLE_CLASS: case mxINT32_CLASS: case mxUINT32_CLASS: return 4;
    case mxINT16_CLASS: case mxUINT16_CLASS: case mxCHAR_CLASS: return 2;
    default: return 1;
  }
}

/* ---- C Matrix API ---------------------------------------------------- */
mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag) {
  (void)flag;
  mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
  a->cls = classid;
  a->m = m;
  a->n = n;
  a->data = calloc(m * n > 0 ? m * n : 1, esize(classid));
  return a;
}
mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag) {
  return mxCreateNumericMatrix(m, n, mxDOUBLE_CLASS, flag);
}
mxArray *mxDuplicateArray(const mxArray *in) {
  mxArray *a = mxCreateNumericMatrix(in->m, in->n, in->cls, mxREAL);
  memcpy(a->data, in->data, in->m * in->n * esize(in->cls));
  for (int k = 0; k < in->nf; ++k) {
    a->fname[k] = strdup(in->fname[k]);
    a->fval[k] = mxDuplicateArray(in->fval[k]);
  }
  a->nf = in->nf;
  return a;
}
void mxDestroyArray(mxArray *a) {
  if (!a) return;
  for (int k = 0; k < a->nf; ++k) {
    free(a->fname[k]);
    mxDestroyArray(a->fval[k]);
  }
  free(a->data);
  free(a);
}
mxArray *mxGetField(const mxArray *pm, mwIndex index, const char *name) {
  if (!pm || pm->cls != mxSTRUCT_CLASS || index != 0) return NULL;
  for (int k = 0; k < pm->nf; ++k)
    if (!strcmp(pm->fname[k], name)) return pm->fval[k];
  return NULL;
}
size_t mxGetM(const mxArray *pm) { return pm->m; }
size_t mxGetN(const mxArray *pm) { return pm->n; }
size_t mxGetNumberOfElements(const mxArray *pm) { return pm->m * pm->n; }
mxClassID mxGetClassID(const mxArray *pm) { return pm->cls; }
bool mxIsStruct(const mxArray *pm) { return pm && pm->cls == mxSTRUCT_CLASS; }
bool mxIsChar(const mxArray *pm) { return pm && pm->cls == mxCHAR_CLASS; }
bool mxIsDouble(const mxArray *pm) { return pm && pm->cls == mxDOUBLE_CLASS; }
bool mxIsSingle(const mxArray *pm) { return pm && pm->cls == mxSINGLE_CLASS; }
void *mxGetData(const mxArray *pm) { return pm->data; }
mxDouble *mxGetDoubles(const mxArray *pa) {
  return pa->cls == mxDOUBLE_CLASS ? (mxDouble *)pa->data : NULL;
}
mxSingle *mxGetSingles(const mxArray *pa) {
  return pa->cls == mxSINGLE_CLASS ? (mxSingle *)pa->data : NULL;
}
mxInt32 *mxGetInt32s(const mxArray *pa) {
  return pa->cls == mxINT32_CLASS ? (mxInt32 *)pa->data : NULL;
}
double mxGetScalar(const mxArray *pm) {
  if (!pm || pm->m * pm->n == 0) return 0;
  switch (pm->cls) {
    case mxDOUBLE_CLASS: return ((double *)pm->data)[0];
    case mxSINGLE_CLASS: return ((float *)pm->data)[0];
    case mxINT32_CLASS: return ((int32_t *)pm->data)[0];
    case mxUINT32_CLASS: return ((uint32_t *)pm->data)[0];
    default: return 0;
  }
}
/* char arrays hold 16-bit characters; returns 1 when the string does not fit */
int mxGetString(const mxArray *pm, char *str, mwSize len) {
  if (!pm || pm->cls != mxCHAR_CLASS || len == 0) return 1;
  const size_t n = pm->m * pm->n;
  const uint16_t *c = (const uint16_t *)pm->data;
  size_t k = 0;
  for (; k < n && k + 1 < len; ++k) str[k] = (char)c[k];
  str[k] = 0;
  return k < n;
}

/* mxMalloc memory is released when the call ends, as MATLAB does */
#define MAXA 4096
static void *allocs[MAXA];
static int nallocs = 0;
void *mxMalloc(mwSize n) {
  void *p = malloc(n ? n : 1);
  if (nallocs < MAXA) allocs[nallocs++] = p;
  return p;
}
void mxFree(void *ptr) {
  for (int k = 0; k < nallocs; ++k)
    if (allocs[k] == ptr) {
      allocs[k] = allocs[--nallocs];
      break;
    }
  free(ptr);
}

/* ---- C MEX API --------------------------------------------------------- */
static jmp_buf jb;
static char err_id[128], err_msg[1024];
static void (*exit_fn)(void) = NULL;

void mexErrMsgIdAndTxt(const char *errorid, const char *fmt, ...) {
  snprintf(err_id, sizeof err_id, "%s", errorid);
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(err_msg, sizeof err_msg, fmt, ap);
  va_end(ap);
  longjmp(jb, 1);
}
int mexAtExit(void (*f)(void)) {
  exit_fn = f;
  return 0;
}

/* ---- driver API for the tests -------------------------------------------- */
mxArray *stub_numeric(int cls, size_t m, size_t n, const void *data) {
  mxArray *a = mxCreateNumericMatrix(m, n, (mxClassID)cls, mxREAL);
  if (data) memcpy(a->data, data, m * n * esize((mxClassID)cls));
  return a;
}
mxArray *stub_string(const char *s) {
  const size_t n = strlen(s);
  mxArray *a = mxCreateNumericMatrix(1, n, mxCHAR_CLASS, mxREAL);
  for (size_t k = 0; k < n; ++k) ((uint16_t *)a->data)[k] = (uint16_t)(unsigned char)s[k];
  return a;
}
mxArray *stub_struct(void) {
  mxArray *a = mxCreateNumericMatrix(1, 1, mxSTRUCT_CLASS, mxREAL);
  return a;
}
/* takes ownership of v */
int stub_set_field(mxArray *s, const char *name, mxArray *v) {
  if (s->cls != mxSTRUCT_CLASS || s->nf >= MAXF) return -1;
  s->fname[s->nf] = strdup(name);
  s->fval[s->nf++] = v;
  return 0;
}
int stub_info(const mxArray *a, int *cls, size_t *m, size_t *n, void **data) {
  *cls = (int)a->cls;
  *m = a->m;
  *n = a->n;
  *data = a->data;
  return 0;
}
void stub_free(mxArray *a) { mxDestroyArray(a); }

/* call the gateway; 0 on return, -1 when it raised an error (id and message
 * copied out); the outputs it created are in plhs */
int stub_call(int nlhs, mxArray **plhs, int nrhs, mxArray **prhs, char *id, size_t idn, char *msg,
              size_t msgn) {
  int rc = 0;
  nallocs = 0;
  if (setjmp(jb) == 0) {
    mexFunction(nlhs, plhs, nrhs, (const mxArray **)prhs);
  } else {
    snprintf(id, idn, "%s", err_id);
    snprintf(msg, msgn, "%s", err_msg);
    rc = -1;
  }
  while (nallocs > 0) free(allocs[--nallocs]);
  return rc;
}

/* MATLAB clearing the MEX function (or exiting) */
void stub_exit(void) {
  if (exit_fn) exit_fn();
  exit_fn = NULL;
}
