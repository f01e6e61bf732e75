// vct_measured.hip -- measured (real-scanner) input of the MBFPA program:
// air-scan normalisation of the detector data (loadProjectionData,
// M/projectionDataTOMO.cpp:323-524) and the helical window that feeds the
// npf backprojection (M/main.cpp:352-371).
//
// Detector files hold 609 channels per row, of which the first 27 are
// skipped; air-scan files hold 1 + 5 blocks of 800 x 609 floats, block 0
// skipped and the next five averaged in float (sum of x/5, block order).
//   PR (pre-treatment): one 800 x 609 block per rotation; every sino_fact-th
//     row is kept, divided by the air average of the same cell, and the kept
//     rows are packed rotation after rotation.
//   IN (in-treatment): one 609-float row per view; the reference reads each
//     row into the first 609 floats of a zeroed 800 x 609 buffer and then
//     indexes that buffer at the view's row, so only view 0 of each rotation
//     sees data and the other views divide 0; the divisor is twice the air
//     average of row 0; values below 0.1 become 0 and leave the mask.
// Elementwise work, one thread per output value (HBM bound, a few MB).  The
// air average is recomputed per output (five loads) rather than staged, so
// the device entry points need no workspace.
#include "vct_internal.h"

#include <cstdio>
#include <cstring>
#include <vector>

namespace vct {

constexpr int kDetChannels = 609;
constexpr int kDetSkip = 27;
constexpr int kAirRows = 800;
constexpr int kAirBlocks = 5;
constexpr size_t kAirBlock = (size_t)kAirRows * kDetChannels;

struct MeasuredArgs {
    const float* air;   // [5][800][609], the averaged blocks
    const float* raw;   // PR [rot][800][609]; IN [rot][H][609]
    float* out;         // [rot*H][W]
    uint8_t* mask;      // [rot*H][W]
    uint32_t* max_bits; // IN: float bits of the largest positive value
    int in_mode, rot, H, W, nskip, R;  // R: rows kept per PR rotation
};

// sum over the five averaged blocks of air/5, in block order, from 0
__device__ __forceinline__ float air_average(const float* __restrict__ air, size_t cell)
{
    float s = 0.f;
#pragma unroll
    for (int b = 0; b < kAirBlocks; b++) s += air[b * kAirBlock + cell] / (float)kAirBlocks;
    return s;
}

__global__ __launch_bounds__(256) void k_measured_normalise(MeasuredArgs a)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t total = (int64_t)a.rot * a.H * a.W;
    if (o >= total) return;
    const int64_t q = o / a.W;
    const int c = (int)(o - q * a.W);
    if (a.in_mode == 0) {
        // PR: the packed kept rows fill the first rot*R rows; the rest stay 0
        if (q >= (int64_t)a.rot * a.R) {
            a.out[o] = 0.f;
            a.mask[o] = 0;
            return;
        }
        const int t = (int)(q / a.R), r = (int)(q - (int64_t)t * a.R);
        const size_t cell = (size_t)r * a.nskip * kDetChannels + kDetSkip + c;
        a.out[o] = a.raw[(size_t)t * kAirBlock + cell] / air_average(a.air, cell);
        a.mask[o] = 1;
        return;
    }
    const int t = (int)(q / a.H), j = (int)(q - (int64_t)t * a.H);
    const float x = j == 0 ? a.raw[(size_t)t * a.H * kDetChannels + kDetSkip + c] : 0.f;
    float d = x / (air_average(a.air, (size_t)kDetSkip + c) * 2.f);
    if (d > 0.f) atomicMax(a.max_bits, __float_as_uint(d));  // positive floats order as their bits
    uint8_t m = 1;
    if ((double)d < 0.1) {
        d = 0.f;
        m = 0;
    }
    a.out[o] = d;
    a.mask[o] = m;
}

__global__ void k_zero_u32(uint32_t* p) { *p = 0u; }

// window[n][c] = (1 - w) v[n][c] + w v[n + iphase*H][c] in FP64, n < H/2*2
__global__ __launch_bounds__(256) void k_helical_window(const float* __restrict__ vol, int rot, int H, int W,
                                                       float* __restrict__ win)
{
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = H / 2 * 2, use = H / 2 * 2;
    if (o >= (int64_t)rows * W) return;
    const int n = (int)(o / W), c = (int)(o - (int64_t)n * W);
    const int j = n;  // the window starts at view 0
    int iphase = 0;
    if (n < use / 2) iphase = 1;
    else if (n > use / 2) iphase = -1;
    if (j + iphase * H < H) iphase = 0;
    if (j + iphase * H > (rot - 1) * H) iphase = 0;
    const double w = 0.5 * fabs(n - (H / 2.0)) / (H / 2.0);
    win[o] = (float)((1.0 - w) * (double)vol[(size_t)j * W + c] +
                     w * (double)vol[(size_t)(j + iphase * H) * W + c]);
}

static vct_status check_measured(const vct_measured* m, int* R)
{
    if (!m) return set_error(VCT_ERR_ARG, "NULL argument");
    if (m->in_mode != 0 && m->in_mode != 1) return set_error(VCT_ERR_ARG, "in_mode %d is neither PR (0) nor IN (1)", m->in_mode);
    if (m->rotations < 1 || m->height < 1 || m->width < 1 || m->sino_fact < 1)
        return set_error(VCT_ERR_ARG, "rotations %d, height %d, width %d, sino_fact %d must be >= 1", m->rotations,
                         m->height, m->width, m->sino_fact);
    if (m->width + kDetSkip > kDetChannels)
        return set_error(VCT_ERR_ARG, "width %d + %d skipped channels exceeds the %d detector channels (the reference "
                         "then reads past its row buffers)", m->width, kDetSkip, kDetChannels);
    *R = (kAirRows + m->sino_fact - 1) / m->sino_fact;
    if (m->in_mode == 0 && *R > m->height)
        return set_error(VCT_ERR_ARG, "PR keeps %d rows per rotation but height is %d (the reference then writes past "
                         "its volume)", *R, m->height);
    if (m->in_mode == 1 && m->height > kAirRows)
        return set_error(VCT_ERR_ARG, "IN height %d exceeds the %d air-scan rows", m->height, kAirRows);
    if (m->skip_rows < 0) return set_error(VCT_ERR_ARG, "skip_rows %d < 0", m->skip_rows);
    return VCT_OK;
}

}  // namespace vct

using namespace vct;

extern "C" {

vct_status vct_measured_normalise(const vct_measured* m, const float* d_air, const float* d_raw, float* d_volume,
                                  uint8_t* d_mask, uint32_t* d_max_bits, void* stream)
{
    clear_error();
    int R = 0;
    VCT_CHECK(check_measured(m, &R));
    if (!d_air || !d_raw || !d_volume || !d_mask || !d_max_bits) return set_error(VCT_ERR_ARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    MeasuredArgs a{d_air, d_raw, d_volume, d_mask, d_max_bits, m->in_mode, m->rotations, m->height, m->width,
                   m->sino_fact, R};
    const int64_t total = (int64_t)m->rotations * m->height * m->width;
    hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, s, d_max_bits);
    hipLaunchKernelGGL(k_measured_normalise, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_helical_window(const float* d_volume, int rotations, int height, int width, float* d_window,
                              void* stream)
{
    clear_error();
    if (!d_volume || !d_window) return set_error(VCT_ERR_ARG, "NULL argument");
    if (rotations < 1 || height < 2 || width < 1)
        return set_error(VCT_ERR_ARG, "rotations %d, height %d, width %d invalid", rotations, height, width);
    const int64_t total = (int64_t)(height / 2 * 2) * width;
    hipLaunchKernelGGL(k_helical_window, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       d_volume, rotations, height, width, d_window);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_load_measured_host(vct_ctx* ctx, const vct_measured* m, const char* detector_path,
                                  const char* airscan_path, float* volume, uint8_t* mask, double* max_proj)
{
    clear_error();
    int R = 0;
    VCT_CHECK(check_measured(m, &R));
    if (!ctx || !detector_path || !airscan_path || !volume || !mask) return set_error(VCT_ERR_ARG, "NULL argument");
    // air scans: block 0 skipped, the next five kept (M/projectionDataTOMO.cpp:383-420)
    std::vector<float> air((size_t)kAirBlocks * kAirBlock);
    {
        FILE* fp = std::fopen(airscan_path, "rb");
        if (!fp) return set_error(VCT_ERR_IO, "data file: %s not found", airscan_path);
        std::vector<float> skip(kAirBlock);
        const size_t n0 = std::fread(skip.data(), sizeof(float), kAirBlock, fp);
        const size_t n1 = std::fread(air.data(), sizeof(float), air.size(), fp);
        std::fclose(fp);
        if (n0 != kAirBlock || n1 != air.size())
            return set_error(VCT_ERR_IO, "%s holds fewer than 6 air-scan blocks of %d x %d floats", airscan_path,
                             kAirRows, kDetChannels);
    }
    // detector rows (:434-497): PR one block per rotation; IN one row per
    // view, followed by skip_rows rows the reference reads past
    const size_t raw_n = m->in_mode == 0 ? (size_t)m->rotations * kAirBlock
                                         : (size_t)m->rotations * m->height * kDetChannels;
    std::vector<float> raw(raw_n);
    {
        FILE* fp = std::fopen(detector_path, "rb");
        if (!fp) return set_error(VCT_ERR_IO, "data file: %s not found", detector_path);
        bool ok = true;
        if (m->in_mode == 0) {
            ok = std::fread(raw.data(), sizeof(float), raw_n, fp) == raw_n;
        } else {
            std::vector<float> skip((size_t)kDetChannels * (m->skip_rows > 0 ? m->skip_rows : 1));
            for (size_t row = 0; ok && row < (size_t)m->rotations * m->height; row++) {
                ok = std::fread(raw.data() + row * kDetChannels, sizeof(float), kDetChannels, fp) == (size_t)kDetChannels;
                if (ok && m->skip_rows > 0)
                    ok = std::fread(skip.data(), sizeof(float), (size_t)kDetChannels * m->skip_rows, fp) ==
                         (size_t)kDetChannels * m->skip_rows;
            }
        }
        std::fclose(fp);
        if (!ok) return set_error(VCT_ERR_IO, "%s is shorter than the %s layout needs", detector_path,
                                  m->in_mode == 0 ? "PR" : "IN");
    }
    const size_t vol_n = (size_t)m->rotations * m->height * m->width;
    hipStream_t s = ctx->stream;
    float *d_air = nullptr, *d_raw = nullptr, *d_vol = nullptr;
    uint8_t* d_mask = nullptr;
    uint32_t* d_max = nullptr;
    uint32_t max_bits = 0;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_air, air.size() * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_raw, raw_n * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_vol, vol_n * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_mask, vol_n), "hipMalloc");
    hip(hipMalloc(&d_max, sizeof(uint32_t)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_air, air.data(), air.size() * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_raw, raw.data(), raw_n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) st = vct_measured_normalise(m, d_air, d_raw, d_vol, d_mask, d_max, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(volume, d_vol, vol_n * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipMemcpyAsync(mask, d_mask, vol_n, hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipMemcpyAsync(&max_bits, d_max, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    void* bufs[] = {d_air, d_raw, d_vol, d_mask, d_max};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (st != VCT_OK) return st;
    float mx = 0.f;
    std::memcpy(&mx, &max_bits, sizeof(float));
    if (max_proj) *max_proj = mx;
    // the reference stops the program here (:479-483)
    if (m->in_mode == 1 && mx > 1.0f) return set_error(VCT_ERR_ARG, "Maximum projection: %f > 1.0", (double)mx);
    return VCT_OK;
}

vct_status vct_helical_window_host(vct_ctx* ctx, const float* volume, int rotations, int height, int width,
                                   float* window)
{
    clear_error();
    if (!ctx || !volume || !window) return set_error(VCT_ERR_ARG, "NULL argument");
    if (rotations < 1 || height < 2 || width < 1)
        return set_error(VCT_ERR_ARG, "rotations %d, height %d, width %d invalid", rotations, height, width);
    const size_t vol_n = (size_t)rotations * height * width, win_n = (size_t)(height / 2 * 2) * width;
    hipStream_t s = ctx->stream;
    float *d_vol = nullptr, *d_win = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_vol, vol_n * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_win, win_n * sizeof(float)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_vol, volume, vol_n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) st = vct_helical_window(d_vol, rotations, height, width, d_win, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(window, d_win, win_n * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (d_vol) (void)hipFree(d_vol);
    if (d_win) (void)hipFree(d_win);
    return st;
}

}  // extern "C"
