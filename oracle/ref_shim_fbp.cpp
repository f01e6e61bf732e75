// ref_shim_fbp.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points onto the reference's own FBP functions.  Compiled
// together with Release_FBP_CPU/{filterSinogramTOMO,reconstructImageTOMO,
// misc,mallocMD}.cpp into oracle/_ref/libref_fbp.so by oracle/Makefile.  It
// only builds the row-pointer arrays the reference expects and forwards.
#include <stdlib.h>
#include <string.h>

#include "physParamsTOMO.h"
#include "mallocMD.h"
#include "misc.h"
#include "filterSinogramTOMO.h"
#include "reconstructImageTOMO.h"

extern "C" {

int ref_fbp_channels(void) { return End_Channel - Start_Channel; }

// F/filterSinogramTOMO.cpp:9 -> out [W][W]
void ref_const_map(int filter, double* out)
{
    const int W = End_Channel - Start_Channel;
    double** C = prepareConstMap(W, filter);
    for (int m = 0; m < W; m++) memcpy(out + (size_t)m * W, C[m], sizeof(double) * W);
    delete2DArray((void**)C, W, W, UNIT_FLOAT64);
}

// F/filterSinogramTOMO.cpp:205: proj, fltd float [H][W]
void ref_filter(int H, const float* proj, float* fltd, int filter, int in_mode)
{
    const int W = End_Channel - Start_Channel;
    double** C = prepareConstMap(W, filter);
    float** P = (float**)new2DArray(W, H, UNIT_FLOAT32);
    float** F = (float**)new2DArray(W, H, UNIT_FLOAT32);
    memcpy(P[0], proj, sizeof(float) * (size_t)W * H);
    double* w0 = (double*)new1DArray(W, UNIT_FLOAT64);
    double* w1 = (double*)new1DArray(W, UNIT_FLOAT64);
    double* w2 = (double*)new1DArray(W, UNIT_FLOAT64);
    prepareCorrectedDetectorInterpolation(W, w0, w1, w2);
    double* xs = (double*)new1DArray(H, UNIT_FLOAT64);
    double* ys = (double*)new1DArray(H, UNIT_FLOAT64);
    char label[4];
    strcpy(label, in_mode ? "IN" : "PR");
    filterSinogram(H, 1, 1, P, F, C, w0, w1, w2, xs, ys, label);
    memcpy(fltd, F[0], sizeof(float) * (size_t)W * H);
    delete2DArray((void**)C, W, W, UNIT_FLOAT64);
    delete2DArray((void**)P, W, H, UNIT_FLOAT32);
    delete2DArray((void**)F, W, H, UNIT_FLOAT32);
    free(w0); free(w1); free(w2); free(xs); free(ys);
}

// F/reconstructImageTOMO.cpp:11 (one slice): fltd [H][W] -> img [N][N]
void ref_backproject(int N, double pix, int H, const float* fltd, double* cosT,
                     double* sinT, double xfactor, float* img)
{
    const int W = End_Channel - Start_Channel;
    float** S = (float**)new2DArray(W, H, UNIT_FLOAT32);
    memcpy(S[0], fltd, sizeof(float) * (size_t)W * H);
    float*** R = (float***)new3DArray(N, N, 1, UNIT_FLOAT32);
    char** mask = prepareMaskData(N);
    char label[4];
    strcpy(label, "PR");
    reconstructImage(N, R, mask, W, H, S, cosT, sinT, 0.0, 0.1, 1, 0.0, 0.0, 1, xfactor, pix, label);
    memcpy(img, R[0][0], sizeof(float) * (size_t)N * N);
    delete2DArray((void**)S, W, H, UNIT_FLOAT32);
    delete3DArray((void***)R, N, N, 1, UNIT_FLOAT32);
    delete2DArray((void**)mask, N, N, UNIT_UINT8);
}

}  // extern "C"
