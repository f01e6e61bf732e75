// ref_shim_mbfpa.cpp -- TEST INFRASTRUCTURE ONLY.
//
// extern "C" entry points onto the reference's own MBFPA functions.  This
// file is compiled together with the reference sources where they lie
// (Release_MBFPA_CPU_human/{reprojection_TOMO,lengthproj_TOMO,
// prior_weight_production,misc_TOMO,mallocMD}.cpp) into oracle/_ref/libref_mbfpa.so by
// oracle/Makefile.  It only allocates the buffers the reference functions
// expect and forwards the call; no reference code is restated here.
#include <stdlib.h>
#include <string.h>

#include "physParamsTOMO.h"
#include "mallocMD.h"
#include "reprojection_TOMO.h"
#include "lengthproj_TOMO.h"
#include "misc_TOMO.h"
#include "virtual_projection.h"

extern "C" {

// M/prior_weight_production.cpp:180 (the element table there is fixed to 6)
int ref_attenuation(const char* path, int num_material, float* att, float* pdf)
{
    int NE = 0;
    attenuation_coefficient((char*)path, num_material, att, pdf, &NE, 1);
    return NE;
}

// One energy pass of M/reprojection_TOMO.cpp:13.  dens is float [M][N][N]
// (modified in place exactly as the reference does).  seg_counts, if not
// NULL, receives reprojection_sum [H][W].  Returns 0, or -1 on allocation
// failure.
int ref_reprojection(int N, float* dens, float* proj, double pix, int W, int H,
                     double* cosT, double* sinT, double xfactor, float* att,
                     float* pdf, int M, int NE, int ke, int* seg_counts)
{
    long depth = (N >= 256) ? 2L * N : 4L * N + 64;
    long cache = depth * (long)W * H;
    float* cacheM = (float*)calloc(cache, sizeof(float));
    int* cacheV = (int*)calloc(cache, sizeof(int));
    int* cacheS = (int*)calloc((long)W * H, sizeof(int));
    if (!cacheM || !cacheV || !cacheS) { free(cacheM); free(cacheV); free(cacheS); return -1; }
    double xalpha[1] = {0.0};
    char mask[1] = {1};
    reprojection_TOMO(1, N, dens, proj, pix, W, H, cosT, sinT, 0.0, 0.0, xfactor, xalpha,
                      mask, cacheM, cacheV, cacheS, 1, att, pdf, M, NE, ke);
    if (seg_counts) memcpy(seg_counts, cacheS, sizeof(int) * (size_t)W * H);
    free(cacheM); free(cacheV); free(cacheS);
    return 0;
}

// M/main.cpp:492-506: the whole energy loop, summed in float per bin as the
// reference driver does.  out is float [H][W].
int ref_reprojection_full(int N, float* dens, float* out, double pix, int W, int H,
                          double* cosT, double* sinT, double xfactor, float* att,
                          float* pdf, int M, int NE, int ke_begin, int ke_end)
{
    long depth = (N >= 256) ? 2L * N : 4L * N + 64;
    long cache = depth * (long)W * H;
    float* cacheM = (float*)calloc(cache, sizeof(float));
    int* cacheV = (int*)calloc(cache, sizeof(int));
    int* cacheS = (int*)calloc((long)W * H, sizeof(int));
    float* proj = (float*)calloc((long)W * H, sizeof(float));
    if (!cacheM || !cacheV || !cacheS || !proj) {
        free(cacheM); free(cacheV); free(cacheS); free(proj);
        return -1;
    }
    double xalpha[1] = {0.0};
    char mask[1] = {1};
    memset(out, 0, sizeof(float) * (size_t)W * H);
    for (int ke = ke_begin; ke < ke_end; ke++) {
        reprojection_TOMO(1, N, dens, proj, pix, W, H, cosT, sinT, 0.0, 0.0, xfactor, xalpha,
                          mask, cacheM, cacheV, cacheS, 1, att, pdf, M, NE, ke);
        for (long i = 0; i < (long)W * H; i++) out[i] += proj[i];
    }
    free(cacheM); free(cacheV); free(cacheS); free(proj);
    return 0;
}

}  // extern "C"

extern "C" {

// Segment lists the reference caches for a set of rays during one energy
// pass (M/reprojection_TOMO.cpp:160-163, :186-188; index = segment counter).
// For ray r = rays[i] (= view*W + channel), writes count[i] = the counter
// value (reprojection_sum) and, for c in 1..count[i], vox/len[i*stride + c-1]
// = the cached (vnum, rr) at counter c (0/0 where nothing was cached).
int ref_segments(int N, float* dens, double pix, int W, int H, double* cosT, double* sinT,
                 float* att, float* pdf, int M, int NE, const int* rays, int nrays,
                 int stride, int* count, int* vox, float* len)
{
    long depth = (N >= 256) ? 2L * N : 4L * N + 64;
    long cache = depth * (long)W * H;
    float* cacheM = (float*)calloc(cache, sizeof(float));
    int* cacheV = (int*)calloc(cache, sizeof(int));
    int* cacheS = (int*)calloc((long)W * H, sizeof(int));
    float* proj = (float*)calloc((long)W * H, sizeof(float));
    if (!cacheM || !cacheV || !cacheS || !proj) {
        free(cacheM); free(cacheV); free(cacheS); free(proj);
        return -1;
    }
    double xalpha[1] = {0.0};
    char mask[1] = {1};
    reprojection_TOMO(1, N, dens, proj, pix, W, H, cosT, sinT, 0.0, 0.0, 1.0, xalpha,
                      mask, cacheM, cacheV, cacheS, 1, att, pdf, M, NE, 0);
    for (int i = 0; i < nrays; i++) {
        const long r = rays[i];
        const int n = cacheS[r];
        count[i] = n;
        for (int c = 1; c <= n && c <= stride; c++) {
            vox[(long)i * stride + c - 1] = cacheV[(long)c * W * H + r];
            len[(long)i * stride + c - 1] = cacheM[(long)c * W * H + r];
        }
    }
    free(cacheM); free(cacheV); free(cacheS); free(proj);
    return 0;
}

}  // extern "C"

extern "C" {

// M/lengthproj_TOMO.cpp:14 (the npf pass of M/main.cpp:407-419): npf [N][N]
// must be zero on entry (clear_npf, M/main.cpp:330); proj [H][W] is the
// sinogram window; the mask is the reference's own prepareMaskData.
// Returns 0, or -1 on allocation failure.
int ref_lengthproj(int N, float* npf, float* proj, double pix, int W, int H, double* cosT, double* sinT,
                   double z, double couch)
{
    const size_t cache = (size_t)N * N * H;
    float* sm = (float*)calloc(cache, sizeof(float));
    float* xd = (float*)calloc(cache, sizeof(float));
    char* mask = prepareMaskData(N);
    if (!sm || !xd || !mask) { free(sm); free(xd); free(mask); return -1; }
    lengthproj_TOMO(0, N, npf, proj, pix, W, H, cosT, sinT, z, couch, mask, sm, xd, 1);
    free(sm); free(xd); free(mask);
    return 0;
}

}  // extern "C"

#include "projectionDataTOMO.h"

extern "C" {

// loadProjectionData (M/projectionDataTOMO.cpp:323-524) on a PROJECTION_DATA
// holding the fields loadData (:68-134) would have set: the class label, the
// sinogram width and derived height, the rotation count, sino_fact, the
// original sampling rate and the two file names (float data).  Copies the
// W x H' x rotations volume and mask out; returns the call's result and the
// height after sino_fact in *rows.  The reference writes projections.raw into
// the working directory.
int ref_load_measured(const char* inpr, int width, int height, int rotations, int sino_fact, double rate_orig,
                      const char* proj_file, const char* air_file, float* volume, unsigned char* mask, int* rows)
{
    PROJECTION_DATA* pd = newProjectionData();
    strncpy(pd->InPrClassString, inpr, sizeof(pd->InPrClassString) - 1);
    strncpy(pd->projImgDataType, "float", sizeof(pd->projImgDataType) - 1);
    strncpy(pd->projImgFileName, proj_file, sizeof(pd->projImgFileName) - 1);
    strncpy(pd->AirScanImgFileName, air_file, sizeof(pd->AirScanImgFileName) - 1);
    pd->projImgWidth = width;
    pd->projImgHeight = height;
    pd->totalProjNumber = rotations;
    pd->sino_fact = sino_fact;
    pd->DataSamplingRateOrig = rate_orig;
    pd->DataSamplingRate = rate_orig >= 200 ? rate_orig / 5 : rate_orig;
    pd->StartViewAngle = 0.0;
    pd->datasize = sizeof(float);
    pd->datatype = UNIT_FLOAT32;
    const int r = loadProjectionData(pd);
    *rows = pd->projImgHeight;
    const size_t n = (size_t)pd->projImgWidth * pd->projImgHeight * pd->usedProjNumber;
    memcpy(volume, pd->projVolume, n * sizeof(float));
    memcpy(mask, pd->mask_projVolume, n);
    deleteProjectionData(pd);
    return r;
}

}  // extern "C"
