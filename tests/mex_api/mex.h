/* Declarations of the MATLAB MEX C API (R2018a interleaved-complex API) used by
 * gpe-matlab_amd/mex/gpe_mex.cpp, for a compile-only check of the gateway in this
 * repository's tests (MATLAB is not installed here).  Signatures follow MathWorks'
 * published MEX C API reference; nothing here is executed. */
#ifndef GPE_TEST_MEX_API_H
#define GPE_TEST_MEX_API_H
#include <stddef.h>
#include <stdbool.h>

typedef struct mxArray_tag mxArray;
typedef size_t mwSize;
typedef double mxDouble;
typedef struct { double real, imag; } mxComplexDouble;
typedef enum { mxUNKNOWN_CLASS = 0, mxDOUBLE_CLASS = 6, mxUINT64_CLASS = 15 } mxClassID;
typedef enum { mxREAL = 0, mxCOMPLEX = 1 } mxComplexity;

#ifdef __cplusplus
extern "C" {
#endif
void mexFunction(int nlhs, mxArray* plhs[], int nrhs, const mxArray* prhs[]);
void mexErrMsgIdAndTxt(const char* errorid, const char* errormsg, ...);
int mexAtExit(void (*exit_fcn)(void));
int mxGetString(const mxArray* pm, char* str, mwSize strlen);
char* mxArrayToString(const mxArray* array_ptr);
void mxFree(void* ptr);
bool mxIsChar(const mxArray* pm);
bool mxIsDouble(const mxArray* pm);
bool mxIsComplex(const mxArray* pm);
bool mxIsUint64(const mxArray* pm);
bool mxIsEmpty(const mxArray* pm);
size_t mxGetNumberOfElements(const mxArray* pm);
size_t mxGetM(const mxArray* pm);
size_t mxGetN(const mxArray* pm);
double mxGetScalar(const mxArray* pm);
void* mxGetData(const mxArray* pm);
mxDouble* mxGetDoubles(const mxArray* pa);
mxComplexDouble* mxGetComplexDoubles(const mxArray* pa);
mxArray* mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID classid, mxComplexity flag);
mxArray* mxCreateNumericArray(mwSize ndim, const mwSize* dims, mxClassID classid, mxComplexity flag);
mxArray* mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity flag);
mxArray* mxCreateDoubleScalar(double value);
mxArray* mxCreateString(const char* str);
void mxDestroyArray(mxArray* pm);
#ifdef __cplusplus
}
#endif
#endif
