// vct_internal.h -- internal types shared by the libvct translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "vct.h"

namespace vct {

// Thread-local error message + status helper.
vct_status set_error(vct_status s, const char* fmt, ...);
void clear_error();

#define VCT_HIP(call)                                                               \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            return ::vct::set_error(VCT_ERR_HIP, "%s failed: %s (%s:%d)", #call,    \
                                    hipGetErrorString(e_), __FILE__, __LINE__);     \
    } while (0)

#define VCT_CHECK(st)                         \
    do {                                      \
        vct_status s_ = (st);                 \
        if (s_ != VCT_OK) return s_;          \
    } while (0)

// Filter GEMM padding: K and W are padded to a multiple of kFilterPad.
constexpr int kFilterPad = 64;
// Maximum spectra evaluated by one spectral launch and maximum bins.
constexpr int kMaxBins = 1024;

// Tracer launch parameters (passed by value).
struct TraceArgs {
    const double* cosT;   // [V]
    const double* sinT;   // [V]
    const double* px0;    // [W] detector point before rotation (x)
    const double* py0;    // [W] detector point before rotation (y)
    const void* dens;     // packed [N][N][8] (u16 or f32)
    void* dens_t;         // orbit tracer: work copies of dens, [y][x] and [x][y], padded to N + 4
    float* L;             // [V][W][8] path integrals (forward)
    int32_t* counts;      // [V][W] optional segment counter
    const float* wts;     // [V][W][8] adjoint weights (adjoint)
    const float* g;       // [V][W] optional upstream gradient (adjoint)
    float* grad;          // [N][N][8] (adjoint)
    int* err;             // device error flag (step bound exceeded)
    int N, W, V, VT;      // VT = views traced = V/2*2
    int nxm;              // (int)(N*0.5)
    int max_steps;        // hard bound on marching steps per ray
    int rcheck;           // 1 when the reference's r >= rmax test can fire
    int exact;            // 1 forces the exact-division march
    int rot;              // forward: rays per orbit of k_trace_fwd_rot (1: one ray per lane)
    int block;            // lanes per workgroup of the tracer (64, 128 or 256)
    int adj_views;        // views per workgroup of the tile adjoint (0: automatic)
    int ray_offset;       // first ray handled by this launch
    int n_rays;           // rays handled by this launch
    double D;             // pixel pitch
    double invD;          // 1/D (fast path for the voxel index, with exact fallback)
    double rmax;          // N*0.5 (reference :37)
    double sid;           // source to isocentre
    double out_scale;     // factor applied to the accumulated sums (1/1000 for u16, then /xfactor)
};

struct SpectralArgs {
    const float* L;       // [V][W][8]
    const float* att;     // [S][M][NE]
    const float* pdf;     // [S][NE]
    float* raw;           // [S][V][W]
    uint32_t* stats;      // [S][2] = {max bits, min bits}
    float* dIdL;          // [S][V][W][8] optional
    int W, V, VT, NE, S, M;
};

struct DetectorArgs {
    const float* raw;       // [S][V*W]
    const uint32_t* stats;  // [S][2]
    float* out;             // [S][V*W]
    int n;                  // V*W
    int S;
    double sigma;
    uint32_t key0, key1;    // Philox key (seed)
    uint32_t stream_id;
};

struct FilterPrepArgs {
    const float* proj;    // [R][W]
    const double* coeff;  // [W] sid*cos((k - W/2) * dbeta)
    float* apk;           // packed A operand
    int R, W, Kp;
    int bhc;              // fuse the beam-hardening correction
    double bhc_alpha, bhc_beta;
};

struct FilterGemmArgs {
    const float* apk;     // packed A [R/32][Kp/8][64][4]
    const float* bpk;     // packed B [Wp/32][Kp/8][64][4]
    const float* toep;    // Toeplitz generator of the rows without edge renormalisation
    float* out;           // [R][W]
    int R, W, Kp, in_mode;
    int edge_lo, edge_hi; // output columns m <= edge_lo or m >= edge_hi are renormalised rows
};

struct BackprojArgs {
    const float* fltd;    // [S][V][W]
    const double* cosT;   // [V]
    const double* sinT;   // [V]
    float* img;           // [S][N][N]
    int N, W, V, use_views, cent, half;
    int exact;            // 1: views in the reference's order per pixel, no 4-fold sharing
    int max_k;            // most sinograms per lane of the multi-sinogram kernel (0: kBpSinos)
    double pix, sid, dbeta_ch, inv_dbeta_ch, dbeta_view, xfactor, center, mask_r2, mask_r;
};

// Backprojections of up to kMaxBpGroups geometries in one launch
// (launch_backproject_groups fills tiles_x and pref).
constexpr int kMaxBpGroups = 16;
struct BpGroups {
    BackprojArgs a[kMaxBpGroups];
    int S[kMaxBpGroups];        // sinograms of each group
    int tiles_x[kMaxBpGroups];  // tiles per image row
    int pref[kMaxBpGroups + 1]; // first workgroup of each group
    int n;
};

struct LengthprojArgs {
    const float* proj;    // [count][W] sinogram window
    const double* cosT;   // [V]
    const double* sinT;   // [V]
    float* npf;           // [N][N]
    int N, W, V;          // W: detector channels (the channel pitch)
    int pw;               // channels per window row (<= W): channels 0 .. pw-1
    int base;             // view of window row 0
    int count;            // window rows (V/2*2)
    double pix, rz, sid, det_src_iso, det_radius, iso_det, iso_channel;
    double delta;         // channel pitch in radians, max_alpha / ((W-1)/2)
};

}  // namespace vct

struct vct_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int* d_err = nullptr;          // kernel error flag
    void* scratch = nullptr;       // grows on demand
    size_t scratch_bytes = 0;
    // Buffers the scratch outgrew.  They stay allocated until vct_terminate,
    // so a kernel recorded with an older scratch pointer (a captured HIP
    // graph) never writes into freed memory.
    std::vector<void*> retired;
    void* noalloc_scratch = nullptr;  // this device's copy of the no-allocation forward's buffer
    int num_cu = 0;
    vct_status ensure_scratch(size_t bytes);
};

struct vct_plan {
    vct_ctx* ctx = nullptr;
    vct_geometry g{};
    int half = 0;       // (W-1)/2
    int cent = 0;       // W/2
    int nxm = 0;        // (int)(N*0.5)
    int VT = 0;         // views traced
    int Kp = 0;         // padded K for the filter GEMM
    int Wp = 0;         // padded W for the filter GEMM
    double* d_cos = nullptr;
    double* d_sin = nullptr;
    double* d_px0 = nullptr;
    double* d_py0 = nullptr;
    double* d_coeff = nullptr;
    float* d_bpk[3] = {nullptr, nullptr, nullptr};   // packed filter matrix per type
    float* d_toep[3] = {nullptr, nullptr, nullptr};  // Toeplitz generator per type
    void* d_dens_t = nullptr;  // the orbit tracer's density work copies (N, V even)
    int exact_trace = 0;  // VCT_OPT_EXACT_TRACE
    int max_steps = 0;    // VCT_OPT_MAX_STEPS (0: the automatic bound)
    int exact_bp = 0;     // VCT_OPT_EXACT_BACKPROJECT
    int trace_orbit = 4;  // VCT_OPT_TRACE_ORBIT
    int trace_block = 0;  // VCT_OPT_TRACE_BLOCK (0: 64)
    int bp_max_k = 0;     // VCT_OPT_BP_SINOS_PER_LANE (0: the kernel's limit)
    int adj_views = 0;    // VCT_OPT_ADJOINT_VIEWS (0: automatic)
};

namespace vct {
// kernels launchers (implemented in the .hip files)
vct_status launch_trace(const TraceArgs& a, int dens_type, int M, bool adjoint, hipStream_t s);
vct_status launch_spectral(const SpectralArgs& a, hipStream_t s);
vct_status launch_stats_init(uint32_t* stats, int S, hipStream_t s);
vct_status launch_detector(const DetectorArgs& a, hipStream_t s);
vct_status launch_minmax(const float* x, int n, uint32_t* stats, hipStream_t s);
vct_status launch_pack_i16(const int16_t* src, int M, int N, uint16_t* dst, hipStream_t s);
vct_status launch_pack_f32(const float* src, int M, int N, float* dst, hipStream_t s);
vct_status launch_filter(const FilterPrepArgs& p, const FilterGemmArgs& g, hipStream_t s);
vct_status launch_backproject(const BackprojArgs& a, int S, hipStream_t s);
// il: every group's sinograms interleaved in slices of K (4, 8, 12 or 16)
vct_status launch_backproject_groups(BpGroups& ga, int K, hipStream_t s, bool il = false);
vct_status launch_interleave_sinos(const float* src, int S, int VW, int K, float* dst, hipStream_t s);
vct_status launch_beam_hardening(const float* in, float* out, int n, double alpha, double beta, hipStream_t s);
vct_status launch_lengthproj(const LengthprojArgs& a, hipStream_t s);
vct_status plan_filter_matrix(vct_plan* plan, int filter, float** d_bpk, float** d_toep);
TraceArgs make_trace_args(const vct_plan* plan);
}  // namespace vct
