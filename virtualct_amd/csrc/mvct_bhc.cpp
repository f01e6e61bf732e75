// mvct_bhc.cpp -- drop-in replacement for the beam-hardening tool
// (M/beam_hardening_correction.cpp:21-69, built there as a.out):
//
//   mvct_bhc <alpha> <beta>
//
// reads reprojection_float.raw from the working directory (float32; the
// reference hard-codes 609 x 800 values, here the file length decides),
// applies y' = |y| + alpha |y|^beta to y = -ln p on the GPU and writes
// reprojection_float_cor.raw; prints the minima before and after as the
// reference does.  Environment: VCT_DEVICE.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include <hip/hip_runtime.h>

#include "vct.h"

int main(int argc, char* argv[])
{
    if (argc != 3) {
        std::fprintf(stderr, "\n[usage]$ no alpha and beta \n\n");
        return 0;
    }
    const double alpha = std::atof(argv[1]);
    const double beta = std::atof(argv[2]);
    const char* name = "reprojection_float.raw";
    FILE* fp = std::fopen(name, "rb");
    if (!fp) {
        std::fprintf(stderr, "data file: %s not found\n", name);
        return 255;
    }
    std::fseek(fp, 0, SEEK_END);
    const long bytes = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    const int n = (int)(bytes / (long)sizeof(float));
    std::vector<float> p(n > 0 ? n : 1);
    const size_t got = n > 0 ? std::fread(p.data(), sizeof(float), n, fp) : 0;
    std::fclose(fp);
    if (n <= 0 || got != (size_t)n) {
        std::fprintf(stderr, "data file: %s is empty or short\n", name);
        return 1;
    }
    const char* env_dev = std::getenv("VCT_DEVICE");
    if (hipSetDevice(env_dev ? std::atoi(env_dev) : 0) != hipSuccess) {
        std::fprintf(stderr, "ERROR: no GPU\n");
        return 1;
    }
    float *d_in = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_in, sizeof(float) * n) != hipSuccess || hipMalloc(&d_out, sizeof(float) * n) != hipSuccess ||
        hipMemcpy(d_in, p.data(), sizeof(float) * n, hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "ERROR: device allocation or copy failed\n");
        return 1;
    }
    float before[2], after[2];
    vct_status st = vct_minmax(d_in, n, before, nullptr);
    if (st == VCT_OK) st = vct_beam_hardening(d_in, n, alpha, beta, d_out, nullptr);
    if (st == VCT_OK) st = vct_minmax(d_out, n, after, nullptr);
    if (st != VCT_OK) {
        std::fprintf(stderr, "ERROR: %s (%s)\n", vct_status_string(st), vct_last_error());
        return 1;
    }
    if (hipMemcpy(p.data(), d_out, sizeof(float) * n, hipMemcpyDeviceToHost) != hipSuccess) {
        std::fprintf(stderr, "ERROR: copy back failed\n");
        return 1;
    }
    (void)hipFree(d_in);
    (void)hipFree(d_out);
    std::printf("%lf %lf \n", (double)before[1], (double)after[1]);
    FILE* out = std::fopen("reprojection_float_cor.raw", "wb");
    if (!out) {
        std::fprintf(stderr, "ERROR: cannot write reprojection_float_cor.raw\n");
        return 1;
    }
    std::fwrite(p.data(), sizeof(float), n, out);
    std::fclose(out);
    return 0;
}
