// vct_api.cpp -- host side of libvct: contexts, geometry plans, the
// spectrum/attenuation and filter-matrix tables, and the C ABI entry points.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vct_internal.h"

// Work buffer of vct_mbfpa_forward_noalloc (path integrals, raw sums and
// their statistics), one copy per GPU, held until the process ends.  Host
// code never uses the name itself (on the host it is a stand-in, not a device
// address): vct_init resolves each device's copy with hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace vct {

static thread_local std::string g_last_error;

// Device address of `scratch` per device, resolved once with that device
// current (vct_init).
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

static vct_status noalloc_scratch_for(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return set_error(VCT_ERR_UNSUPPORTED, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return set_error(VCT_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return VCT_OK;
}

vct_status set_error(vct_status s, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return s;
}

void clear_error() { g_last_error.clear(); }

static const double kPi = 3.141592653589793238462643383279;

static vct_status check_geometry(const vct_geometry* g)
{
    if (!g) return set_error(VCT_ERR_ARG, "geometry is NULL");
    if (g->recon_size < 2 || g->recon_size > 8192) return set_error(VCT_ERR_ARG, "recon_size %d outside 2..8192", g->recon_size);
    if (g->n_channels < 22 || g->n_channels > 8192) return set_error(VCT_ERR_ARG, "n_channels %d outside 22..8192", g->n_channels);
    if (g->n_views < 1 || g->n_views > 65536) return set_error(VCT_ERR_ARG, "n_views %d outside 1..65536", g->n_views);
    if (!(g->pixel_cm > 0.0) || !std::isfinite(g->pixel_cm)) return set_error(VCT_ERR_ARG, "pixel size must be > 0");
    if (!(g->xfactor != 0.0) || !std::isfinite(g->xfactor)) return set_error(VCT_ERR_ARG, "xfactor must be finite and non-zero");
    if ((long long)g->n_views * g->n_channels > (1LL << 28)) return set_error(VCT_ERR_ARG, "sinogram too large");
    if (!(g->sid_cm > 0.0) || !(g->det_radius_cm > 0.0)) return set_error(VCT_ERR_ARG, "distances must be > 0");
    return VCT_OK;
}

TraceArgs make_trace_args(const vct_plan* p)
{
    TraceArgs a{};
    const vct_geometry& g = p->g;
    a.cosT = p->d_cos;
    a.sinT = p->d_sin;
    a.px0 = p->d_px0;
    a.py0 = p->d_py0;
    a.err = p->ctx->d_err;
    a.N = g.recon_size;
    a.W = g.n_channels;
    a.V = g.n_views;
    a.VT = p->VT;
    a.nxm = p->nxm;
    a.D = g.pixel_cm;
    a.invD = 1.0 / g.pixel_cm;
    a.rmax = g.recon_size * 0.5 * 1.0;
    a.sid = g.sid_cm;
    // The r >= rmax test of the reference (M/reprojection_TOMO.cpp:150-155)
    // compares a radius in cm with N/2 and only fires when a segment midpoint
    // can lie farther than N/2 cm from the centre.
    const double corner = std::sqrt(2.0) * (p->nxm + 2) * g.pixel_cm;
    a.rcheck = corner >= a.rmax ? 1 : 0;
    // A ray crosses at most 2N/sqrt(2)+1 major planes plus the integer-cm
    // start plane; the bound below is generous and only guards the loop.
    a.max_steps = 4 * g.recon_size + (int)(4.0 / g.pixel_cm) + 64;
    if (p->max_steps > 0) a.max_steps = p->max_steps;
    a.ray_offset = 0;
    a.n_rays = g.n_views * g.n_channels;
    a.exact = p->exact_trace;
    a.block = p->trace_block > 0 ? p->trace_block : 64;
    a.adj_views = p->adj_views;
    a.dens_t = p->d_dens_t;
    // orbits of the half or quarter turn: every view traced (V even), a grid
    // centred on a voxel corner (N even), the guarded march (no exact option,
    // no r >= rmax test)
    a.rot = 1;
    if (!a.exact && !a.rcheck && g.recon_size % 2 == 0 && a.VT == g.n_views) {
        if (p->trace_orbit >= 4 && g.n_views % 4 == 0) a.rot = 4;
        else if (p->trace_orbit >= 2 && g.n_views % 2 == 0) a.rot = 2;
    }
    return a;
}

static vct_status plan_upload(const void* h, size_t bytes, void** d)
{
    VCT_HIP(hipMalloc(d, bytes));
    VCT_HIP(hipMemcpy(*d, h, bytes, hipMemcpyHostToDevice));
    return VCT_OK;
}

// prepareConstMap (F/filterSinogramTOMO.cpp:9-138) in FP64.
static vct_status const_map(const vct_geometry* g, int filter, double* C)
{
    if (filter != VCT_FILTER_RAMLAK && filter != VCT_FILTER_SHEPP_LOGAN)
        return set_error(VCT_ERR_ARG, "filter %d is not 1 (Ram-Lak) or 2 (Shepp-Logan)", filter);
    const int W = g->n_channels;
    const int half = (W - 1) / 2;
    const double arc = g->max_beta / half;
    std::vector<double> rowsum(W, 0.0);
    for (int m = 0; m < W; m++) {
        double sum = 0.0;
        for (int n = 0; n < W; n++) {
            const int d = m - n;
            double v;
            if (filter == VCT_FILTER_RAMLAK) {
                if (d == 0) v = 0.125 / (arc * arc);
                else if (d % 2 != 0) {
                    const double o = kPi * std::sin(d * arc);
                    v = -0.5 / o / o;
                } else v = 0.0;
            } else {
                const double o = kPi * kPi * arc;
                v = 0.25 / o * (1.0 / std::tan((d + 0.5) * arc) - 1.0 / std::tan((d - 0.5) * arc));
            }
            C[(size_t)m * W + n] = v;
            sum += v;
        }
        rowsum[m] = sum;
    }
    for (int m = 0; m < W; m++)
        if (m <= 20 || m >= W - 20)
            for (int n = 0; n < W; n++) C[(size_t)m * W + n] = C[(size_t)m * W + n] * rowsum[21] / rowsum[m];
    return VCT_OK;
}

// The filter operand B = C * dbeta, packed for the MFMA GEMM, and its
// Toeplitz generator.  Rows m of C that prepareConstMap does not renormalise
// (20 < m < W - 20) depend on m - k only, so toep[m - k + Kp - 1] =
// (float)(C[m][k] * dbeta) for those rows (the same double expression as the
// packed value); toep covers m - k in [-(Kp - 1), Wp - 1] and is 0 outside
// the matrix (padded k, whose A operand is 0).
vct_status plan_filter_matrix(vct_plan* p, int filter, float** d_bpk, float** d_toep)
{
    if (filter != VCT_FILTER_RAMLAK && filter != VCT_FILTER_SHEPP_LOGAN)
        return set_error(VCT_ERR_ARG, "filter %d is not 1 (Ram-Lak) or 2 (Shepp-Logan)", filter);
    if (!p->d_bpk[filter]) {
        const int W = p->g.n_channels;
        std::vector<double> C((size_t)W * W);
        VCT_CHECK(const_map(&p->g, filter, C.data()));
        const double scale = p->g.max_beta / p->half;
        std::vector<float> bpk((size_t)p->Wp * p->Kp, 0.f);
        for (int m = 0; m < W; m++)
            for (int k = 0; k < W; k++) {
                const size_t idx = ((((size_t)(m >> 5) * (p->Kp / 8) + ((k >> 1) >> 2)) * 64 + ((m & 31) + 32 * (k & 1))) * 4) + ((k >> 1) & 3);
                bpk[idx] = (float)(C[(size_t)m * W + k] * scale);
            }
        // generator from the inner rows: every difference d = m - k in
        // [-(W-1), W-1] occurs in some inner row when W > 41
        std::vector<float> toep((size_t)p->Wp + p->Kp, 0.f);
        std::vector<char> seen(toep.size(), 0);
        for (int m = 21; m < W - 20; m++)
            for (int k = 0; k < W; k++) {
                const size_t i = (size_t)(m - k + p->Kp - 1);
                const float v = (float)(C[(size_t)m * W + k] * scale);
                if (seen[i] && toep[i] != v)
                    return set_error(VCT_ERR_INTERNAL, "filter matrix row %d is not Toeplitz", m);
                toep[i] = v;
                seen[i] = 1;
            }
        void* d = nullptr;
        void* dt = nullptr;
        VCT_CHECK(plan_upload(bpk.data(), bpk.size() * sizeof(float), &d));
        vct_status st = plan_upload(toep.data(), toep.size() * sizeof(float), &dt);
        if (st != VCT_OK) {
            (void)hipFree(d);
            return st;
        }
        p->d_bpk[filter] = (float*)d;
        p->d_toep[filter] = (float*)dt;
    }
    *d_bpk = p->d_bpk[filter];
    *d_toep = p->d_toep[filter];
    return VCT_OK;
}

// A NULL stream is HIP's null stream (ordered with every blocking stream,
// e.g. PyTorch's default stream); the context stream serves only the
// host-buffer entry points.
static hipStream_t pick_stream(vct_plan*, void* stream)
{
    return (hipStream_t)stream;
}

static vct_status check_err_flag(vct_ctx* ctx, hipStream_t s)
{
    int h = 0;
    VCT_HIP(hipMemcpyAsync(&h, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    if (h) {
        VCT_HIP(hipMemsetAsync(ctx->d_err, 0, sizeof(int), s));
        VCT_HIP(hipStreamSynchronize(s));
        return set_error(VCT_ERR_INTERNAL, "a kernel reported error flags %#x (1: a ray exceeded its marching "
                                           "step bound; 2: a system-matrix row did not match its count)", h);
    }
    return VCT_OK;
}

}  // namespace vct

using namespace vct;

// Grow-only: a larger request allocates a new buffer and retires the old one
// (freed at vct_terminate), so pointers already handed to kernels, including
// kernels recorded in a HIP graph, stay valid.
vct_status vct_ctx::ensure_scratch(size_t bytes)
{
    if (bytes <= scratch_bytes) return VCT_OK;
    void* fresh = nullptr;
    VCT_HIP(hipMalloc(&fresh, bytes));
    if (scratch) retired.push_back(scratch);
    scratch = fresh;
    scratch_bytes = bytes;
    return VCT_OK;
}

extern "C" {

void vct_geometry_default(vct_geometry* g)
{
    if (!g) return;
    std::memset(g, 0, sizeof(*g));
    g->recon_size = 512;
    g->n_channels = 609;
    g->n_views = 800;
    g->pixel_cm = 0.1;
    g->start_view_deg = 0.0;
    g->sid_cm = 85.0;
    g->det_radius_cm = 142.6;
    g->det_src_iso_cm = 85.0;
    g->iso_channel = 304.5;
    g->max_alpha = 0.326086956;
    g->max_beta = 0.326086956;
    g->xfactor = 1.0;
}

const char* vct_status_string(vct_status s)
{
    switch (s) {
    case VCT_OK: return "ok";
    case VCT_ERR_ARG: return "invalid argument";
    case VCT_ERR_HIP: return "HIP runtime error";
    case VCT_ERR_NOMEM: return "out of memory";
    case VCT_ERR_IO: return "I/O error";
    case VCT_ERR_UNSUPPORTED: return "unsupported";
    case VCT_ERR_INTERNAL: return "internal limit";
    default: return "unknown status";
    }
}

const char* vct_last_error(void) { return g_last_error.c_str(); }

int vct_abi_version(void) { return VCT_ABI_VERSION; }

vct_status vct_init(int device, vct_ctx** out)
{
    clear_error();
    if (!out) return set_error(VCT_ERR_ARG, "ctx pointer is NULL");
    *out = nullptr;
    int n = 0;
    VCT_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return set_error(VCT_ERR_ARG, "device %d not present (%d devices)", device, n);
    VCT_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    VCT_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(VCT_ERR_UNSUPPORTED, "device %d is %s; the library is compiled for gfx950 only", device, prop.gcnArchName);
    void* noalloc = nullptr;
    VCT_CHECK(noalloc_scratch_for(device, &noalloc));  // with the device current
    vct_ctx* c = new vct_ctx();
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->noalloc_scratch = noalloc;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMalloc(&c->d_err, sizeof(int));
    if (e == hipSuccess) e = hipMemset(c->d_err, 0, sizeof(int));
    if (e != hipSuccess) {
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return set_error(VCT_ERR_HIP, "context setup failed: %s", hipGetErrorString(e));
    }
    *out = c;
    return VCT_OK;
}

vct_status vct_check_errors(vct_ctx* c, void* stream)
{
    clear_error();
    if (!c) return set_error(VCT_ERR_ARG, "ctx is NULL");
    VCT_HIP(hipSetDevice(c->device));
    return check_err_flag(c, (hipStream_t)stream);
}

vct_status vct_terminate(vct_ctx* c)
{
    clear_error();
    if (!c) return VCT_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->scratch) (void)hipFree(c->scratch);
    for (void* r : c->retired) (void)hipFree(r);
    if (c->d_err) (void)hipFree(c->d_err);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return VCT_OK;
}

vct_status vct_plan_create(vct_ctx* ctx, const vct_geometry* g, vct_plan** out)
{
    clear_error();
    if (!ctx || !out) return set_error(VCT_ERR_ARG, "NULL argument");
    VCT_CHECK(check_geometry(g));
    VCT_HIP(hipSetDevice(ctx->device));
    vct_plan* p = new vct_plan();
    p->ctx = ctx;
    p->g = *g;
    p->half = (g->n_channels - 1) / 2;
    p->cent = g->n_channels / 2;
    p->nxm = (int)(g->recon_size * 0.5);
    p->VT = g->n_views / 2 * 2;
    p->Kp = (g->n_channels + kFilterPad - 1) / kFilterPad * kFilterPad;
    p->Wp = (g->n_channels + 31) / 32 * 32;

    // View tables: angles as M/projectionDataTOMO.cpp:362-368, cos/sin from
    // the host libm as M/misc_TOMO.cpp:93-97.
    const int V = g->n_views, W = g->n_channels;
    std::vector<double> c(V), s(V), px0(W), py0(W), coeff(W);
    for (int i = 0; i < V; i++) {
        const double angle = (double)(360.0 / V * i);
        const double rad = (g->start_view_deg + angle) * kPi / 180.0;
        c[i] = std::cos(rad);
        s[i] = std::sin(rad);
    }
    // Channel detector points before rotation, M/reprojection_TOMO.cpp:61-71.
    for (int ch = 0; ch < W; ch++) {
        const double theta = (g->iso_channel - ch) * g->max_alpha / p->half;
        const double ox0 = g->det_radius_cm * std::sin(theta);
        const double oy0 = -g->det_radius_cm * std::cos(theta);
        px0[ch] = ox0;
        py0[ch] = oy0 + g->det_src_iso_cm;
    }
    // FBP channel pre-weight, F/filterSinogramTOMO.cpp:236.
    for (int k = 0; k < W; k++) coeff[k] = g->sid_cm * std::cos((k - p->cent) * (g->max_beta / p->half));

    vct_status st = VCT_OK;
    if (st == VCT_OK) st = plan_upload(c.data(), V * sizeof(double), (void**)&p->d_cos);
    if (st == VCT_OK) st = plan_upload(s.data(), V * sizeof(double), (void**)&p->d_sin);
    if (st == VCT_OK) st = plan_upload(px0.data(), W * sizeof(double), (void**)&p->d_px0);
    if (st == VCT_OK) st = plan_upload(py0.data(), W * sizeof(double), (void**)&p->d_py0);
    if (st == VCT_OK) st = plan_upload(coeff.data(), W * sizeof(double), (void**)&p->d_coeff);
    // the orbit tracer's work copies (any dens type), allocated with the
    // plan so that no trace call allocates
    if (st == VCT_OK && g->recon_size % 2 == 0 && V % 2 == 0) {
        // two padded copies of (N + 4)^2 voxels of up to 32 bytes
        const size_t np = (size_t)g->recon_size + 4;
        const size_t bytes = 2 * np * np * 32;
        if (hipMalloc(&p->d_dens_t, bytes) != hipSuccess) {
            (void)hipGetLastError();
            st = set_error(VCT_ERR_NOMEM, "plan work copy of %zu bytes", bytes);
        }
    }
    if (st != VCT_OK) {
        vct_plan_destroy(p);
        return st;
    }
    *out = p;
    return VCT_OK;
}

vct_status vct_plan_set_option(vct_plan* p, int option, int value)
{
    clear_error();
    if (!p) return set_error(VCT_ERR_ARG, "plan is NULL");
    switch (option) {
    case VCT_OPT_EXACT_TRACE: p->exact_trace = value ? 1 : 0; return VCT_OK;
    case VCT_OPT_EXACT_BACKPROJECT: p->exact_bp = value ? 1 : 0; return VCT_OK;
    case VCT_OPT_TRACE_ORBIT:
        if (value != 1 && value != 2 && value != 4) return set_error(VCT_ERR_ARG, "rays per orbit %d is not 1, 2 or 4", value);
        p->trace_orbit = value;
        return VCT_OK;
    case VCT_OPT_MAX_STEPS:
        if (value < 0) return set_error(VCT_ERR_ARG, "step bound %d must be >= 0", value);
        p->max_steps = value;
        return VCT_OK;
    case VCT_OPT_TRACE_BLOCK:
        if (value != 0 && value != 64 && value != 128 && value != 256)
            return set_error(VCT_ERR_ARG, "tracer workgroup %d is not 64, 128 or 256", value);
        p->trace_block = value;
        return VCT_OK;
    case VCT_OPT_BP_SINOS_PER_LANE:
        if (value != 0 && (value < 2 || value > 16))
            return set_error(VCT_ERR_ARG, "sinograms per lane %d outside 2..16", value);
        p->bp_max_k = value;
        return VCT_OK;
    case VCT_OPT_ADJOINT_VIEWS:
        if (value < 0 || value > 256) return set_error(VCT_ERR_ARG, "adjoint views per workgroup %d outside 0..256", value);
        p->adj_views = value;
        return VCT_OK;
    default: return set_error(VCT_ERR_ARG, "unknown option %d", option);
    }
}

vct_status vct_plan_destroy(vct_plan* p)
{
    if (!p) return VCT_OK;
    (void)hipSetDevice(p->ctx->device);
    (void)hipStreamSynchronize(p->ctx->stream);
    double* dd[] = {p->d_cos, p->d_sin, p->d_px0, p->d_py0, p->d_coeff};
    for (double* d : dd)
        if (d) (void)hipFree(d);
    for (float* b : p->d_bpk)
        if (b) (void)hipFree(b);
    for (float* b : p->d_toep)
        if (b) (void)hipFree(b);
    if (p->d_dens_t) (void)hipFree(p->d_dens_t);
    delete p;
    return VCT_OK;
}

vct_status vct_attenuation_coefficient(const char* path, int M, const double* Zin, const double* Ain,
                                       int nthin, float* att, float* pdf, int* NEout, double* energy_mev)
{
    // M/prior_weight_production.cpp:180-297 with the element table extended
    // to Mg and Si; same parsing, bin thinning, normalisation and
    // photoelectric + Klein-Nishina + pair-production formula.
    clear_error();
    static const double kZ[8] = {1, 6, 7, 8, 15, 20, 12, 14};
    static const double kA[8] = {1, 12, 14, 16, 31, 40, 24, 28};
    if (!path || !att || !pdf || !NEout) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || (M > 8 && (!Zin || !Ain)) || M > 64) return set_error(VCT_ERR_ARG, "num_material %d invalid", M);
    if (nthin < 1) return set_error(VCT_ERR_ARG, "nthin must be >= 1");
    const double* Z = Zin ? Zin : kZ;
    const double* A = Ain ? Ain : kA;
    FILE* fp = std::fopen(path, "r");
    if (!fp) return set_error(VCT_ERR_IO, "no photon energy file: %s", path);
    char row[1024];
    if (!std::fgets(row, sizeof(row), fp)) {
        std::fclose(fp);
        return set_error(VCT_ERR_IO, "empty spectrum file: %s", path);
    }
    std::vector<double> energy;
    std::vector<float> w;
    float d0 = 0.f, d1 = 0.f, d2 = 0.f;
    while (std::fgets(row, sizeof(row), fp)) {
        if (energy.size() >= (size_t)kMaxBins) {
            std::fclose(fp);
            return set_error(VCT_ERR_ARG, "spectrum has more than %d bins", kMaxBins);
        }
        // a line that does not parse keeps the previous values (sscanf semantics of :202)
        std::sscanf(row, "%f %f %f", &d0, &d1, &d2);
        energy.push_back((d1 + d0) * 0.5);
        w.push_back(d2);
    }
    std::fclose(fp);
    const int n = (int)energy.size();
    std::vector<float> pdf_t, e_t;
    double psum = 0.0, esum = 0.0, total = 0.0;
    for (int i = 0; i < n; i++) {
        psum += w[i];
        esum += energy[i];
        if ((i - nthin + 1) == (i - nthin + 1) / nthin * nthin) {
            pdf_t.push_back((float)psum);
            e_t.push_back((float)(esum / nthin));
            total += psum;
            psum = 0.0;
            esum = 0.0;
        }
    }
    const int NE = (int)pdf_t.size();
    if (NE == 0) return set_error(VCT_ERR_IO, "spectrum file has no bins: %s", path);
    const float sum = (float)total;
    for (int i = 0; i < NE; i++) pdf[i] = pdf_t[i] / sum;
    for (int iz = 0; iz < M; iz++) {
        const double z = Z[iz], a = A[iz];
        const double fact = 6.0221 * z / a * (2.81794 * 2.81794) * 0.001;
        for (int i = 0; i < NE; i++) {
            const double E = (double)e_t[i] / 0.511;
            const double photo = fact * 0.00000345 * (1.0 + 0.008 * z) * (z * z * z) / (E * E * E) * 1.0;
            const double l2e = std::log(1.0 + 2 * E);
            const double compton = fact * 2 * kPi *
                                   ((1.0 + E) / (E * E) * (2.0 * (1.0 + E) / (1.0 + 2 * E) - l2e / E) + l2e / (2 * E) -
                                    (1.0 + 3 * E) / ((1.0 + 2 * E) * (1.0 + 2 * E)));
            double pair = fact * 0.2545 * (E - 2.332) * z / 137;
            if (pair < 0) pair = 0;
            att[(size_t)NE * iz + i] = (float)(photo + compton + pair);
        }
    }
    if (energy_mev)
        for (int i = 0; i < NE; i++) energy_mev[i] = e_t[i];
    *NEout = NE;
    return VCT_OK;
}

vct_status vct_const_map(const vct_geometry* g, int filter, double* C)
{
    clear_error();
    VCT_CHECK(check_geometry(g));
    if (!C) return set_error(VCT_ERR_ARG, "C is NULL");
    return const_map(g, filter, C);
}

vct_status vct_pack_densities_i16(vct_plan* p, const int16_t* d_dens, int M, uint16_t* d_packed, void* stream)
{
    clear_error();
    if (!p || !d_dens || !d_packed) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    return launch_pack_i16(d_dens, M, p->g.recon_size, d_packed, pick_stream(p, stream));
}

vct_status vct_pack_densities_f32(vct_plan* p, const float* d_dens, int M, float* d_packed, void* stream)
{
    clear_error();
    if (!p || !d_dens || !d_packed) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    return launch_pack_f32(d_dens, M, p->g.recon_size, d_packed, pick_stream(p, stream));
}

vct_status vct_trace_paths(vct_plan* p, const void* d_packed, int dens_type, int M, float* d_L,
                           int32_t* d_counts, void* stream)
{
    clear_error();
    if (!p || !d_packed || !d_L) return set_error(VCT_ERR_ARG, "NULL argument");
    if (dens_type != VCT_DENS_U16 && dens_type != VCT_DENS_F32) return set_error(VCT_ERR_ARG, "bad dens_type %d", dens_type);
    TraceArgs a = make_trace_args(p);
    a.dens = d_packed;
    a.L = d_L;
    a.counts = d_counts;
    a.out_scale = (dens_type == VCT_DENS_U16 ? 1.0 / 1000.0 : 1.0) / p->g.xfactor;
    return launch_trace(a, dens_type, M, false, pick_stream(p, stream));
}

vct_status vct_spectral_sum(vct_plan* p, const float* d_L, int M, const float* d_att, const float* d_pdf, int NE,
                            int S, float* d_raw, uint32_t* d_stats, float* d_dIdL, void* stream)
{
    clear_error();
    if (!p || !d_L || !d_att || !d_pdf || !d_raw || !d_stats) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    hipStream_t s = pick_stream(p, stream);
    VCT_CHECK(launch_stats_init(d_stats, S, s));
    SpectralArgs a{};
    a.L = d_L;
    a.att = d_att;
    a.pdf = d_pdf;
    a.raw = d_raw;
    a.stats = d_stats;
    a.dIdL = d_dIdL;
    a.W = p->g.n_channels;
    a.V = p->g.n_views;
    a.VT = p->VT;
    a.NE = NE;
    a.S = S;
    a.M = M;
    return launch_spectral(a, s);
}

vct_status vct_detector(vct_plan* p, const float* d_raw, const uint32_t* d_stats, int S, double sigma, uint64_t seed,
                        uint32_t stream_id, float* d_out, void* stream)
{
    clear_error();
    if (!p || !d_raw || !d_stats || !d_out) return set_error(VCT_ERR_ARG, "NULL argument");
    if (!(sigma >= 0.0) || !std::isfinite(sigma)) return set_error(VCT_ERR_ARG, "sigma must be finite and >= 0");
    DetectorArgs a{};
    a.raw = d_raw;
    a.stats = d_stats;
    a.out = d_out;
    a.n = p->g.n_views * p->g.n_channels;
    a.S = S;
    a.sigma = sigma;
    a.key0 = (uint32_t)seed;
    a.key1 = (uint32_t)(seed >> 32);
    a.stream_id = stream_id;
    return launch_detector(a, pick_stream(p, stream));
}

size_t vct_forward_workspace_bytes(const vct_plan* p, int S)
{
    if (!p || S < 1) return 0;
    const size_t nr = (size_t)p->g.n_views * p->g.n_channels;
    const size_t L = nr * 8 * sizeof(float);
    const size_t raw = (size_t)S * nr * sizeof(float);
    const size_t st = 256;
    return ((L + 255) / 256 * 256) + ((raw + 255) / 256 * 256) + st;
}

vct_status vct_mbfpa_forward(vct_plan* p, const void* d_packed, int dens_type, int M, const float* d_att,
                             const float* d_pdf, int NE, int S, double sigma, uint64_t seed, uint32_t stream_id,
                             float* d_sino, void* d_work, void* stream)
{
    clear_error();
    if (!p || !d_packed || !d_att || !d_pdf || !d_sino) return set_error(VCT_ERR_ARG, "NULL argument");
    if (S < 1 || S > 32) return set_error(VCT_ERR_ARG, "n_spectra %d outside 1..32", S);
    hipStream_t s = pick_stream(p, stream);
    const size_t need = vct_forward_workspace_bytes(p, S);
    if (!d_work) {
        VCT_CHECK(p->ctx->ensure_scratch(need));
        d_work = p->ctx->scratch;
    }
    const size_t nr = (size_t)p->g.n_views * p->g.n_channels;
    char* w = (char*)d_work;
    float* L = (float*)w;
    w += (nr * 8 * sizeof(float) + 255) / 256 * 256;
    float* raw = (float*)w;
    w += ((size_t)S * nr * sizeof(float) + 255) / 256 * 256;
    uint32_t* stats = (uint32_t*)w;
    VCT_CHECK(vct_trace_paths(p, d_packed, dens_type, M, L, nullptr, s));
    VCT_CHECK(vct_spectral_sum(p, L, M, d_att, d_pdf, NE, S, raw, stats, nullptr, s));
    return vct_detector(p, raw, stats, S, sigma, seed, stream_id, d_sino, s);
}

size_t vct_filter_workspace_bytes(const vct_plan* p, int n_sino)
{
    if (!p || n_sino < 1) return 0;
    const size_t R = (size_t)n_sino * p->g.n_views;
    return (R + 31) / 32 * 32 * p->Kp * sizeof(float);
}

size_t vct_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

vct_status vct_mbfpa_forward_noalloc(vct_plan* p, const void* d_packed, int dens_type, int M, const float* d_att,
                                     const float* d_pdf, int NE, int S, double sigma, uint64_t seed,
                                     uint32_t stream_id, float* d_sino, void* stream)
{
    clear_error();
    if (!p) return set_error(VCT_ERR_ARG, "plan is NULL");
    if (S < 1 || S > 32) return set_error(VCT_ERR_ARG, "n_spectra %d outside 1..32", S);
    const size_t need = vct_forward_workspace_bytes(p, S);
    if (need > kNoAllocScratchBytes)
        return set_error(VCT_ERR_ARG,
                         "%d spectra of %d x %d rays need %zu bytes of work space; the no-allocation forward "
                         "holds %zu (use vct_mbfpa_forward with a workspace)",
                         S, p->g.n_views, p->g.n_channels, need, kNoAllocScratchBytes);
    if (!p->ctx->noalloc_scratch) return set_error(VCT_ERR_INTERNAL, "no-allocation scratch not resolved");
    return vct_mbfpa_forward(p, d_packed, dens_type, M, d_att, d_pdf, NE, S, sigma, seed, stream_id, d_sino,
                             p->ctx->noalloc_scratch, stream);
}

static vct_status filter_impl(vct_plan* p, const float* d_proj, int n_sino, int filter, int in_mode, int bhc,
                              double alpha, double beta, float* d_fltd, void* d_work, void* stream)
{
    clear_error();
    if (!p || !d_proj || !d_fltd) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_sino < 1) return set_error(VCT_ERR_ARG, "n_sino must be >= 1");
    float *bpk = nullptr, *toep = nullptr;
    VCT_CHECK(plan_filter_matrix(p, filter, &bpk, &toep));
    const long long R = (long long)n_sino * p->g.n_views;
    if (R > (1LL << 26)) return set_error(VCT_ERR_ARG, "too many sinogram rows");
    if (!d_work) {
        VCT_CHECK(p->ctx->ensure_scratch(vct_filter_workspace_bytes(p, n_sino)));
        d_work = p->ctx->scratch;
    }
    FilterPrepArgs pa{};
    pa.proj = d_proj;
    pa.coeff = p->d_coeff;
    pa.apk = (float*)d_work;
    pa.R = (int)R;
    pa.W = p->g.n_channels;
    pa.Kp = p->Kp;
    pa.bhc = bhc;
    pa.bhc_alpha = alpha;
    pa.bhc_beta = beta;
    FilterGemmArgs ga{};
    ga.apk = pa.apk;
    ga.bpk = bpk;
    ga.toep = toep;
    ga.out = d_fltd;
    ga.R = (int)R;
    ga.W = p->g.n_channels;
    ga.Kp = p->Kp;
    ga.in_mode = in_mode ? 1 : 0;
    // rows m <= 20 and m >= W - 20 are renormalised (F/filterSinogramTOMO.cpp:114-123)
    ga.edge_lo = 20;
    ga.edge_hi = p->g.n_channels - 20;
    return launch_filter(pa, ga, pick_stream(p, stream));
}

vct_status vct_filter_sinogram(vct_plan* p, const float* d_proj, int n_sino, int filter, int in_mode, float* d_fltd,
                               void* stream)
{
    return filter_impl(p, d_proj, n_sino, filter, in_mode, 0, 0.0, 0.0, d_fltd, nullptr, stream);
}

vct_status vct_filter_sinogram_ws(vct_plan* p, const float* d_proj, int n_sino, int filter, int in_mode,
                                  float* d_fltd, void* d_work, void* stream)
{
    if (!d_work) return set_error(VCT_ERR_ARG, "workspace is NULL");
    return filter_impl(p, d_proj, n_sino, filter, in_mode, 0, 0.0, 0.0, d_fltd, d_work, stream);
}

vct_status vct_filter_sinogram_bhc(vct_plan* p, const float* d_proj, int n_sino, int filter, int in_mode, double alpha,
                                   double beta, float* d_fltd, void* stream)
{
    if (!std::isfinite(alpha) || !std::isfinite(beta)) return set_error(VCT_ERR_ARG, "alpha/beta must be finite");
    return filter_impl(p, d_proj, n_sino, filter, in_mode, 1, alpha, beta, d_fltd, nullptr, stream);
}

vct_status vct_beam_hardening(const float* d_in, int n, double alpha, double beta, float* d_out, void* stream)
{
    clear_error();
    if (!d_in || !d_out) return set_error(VCT_ERR_ARG, "NULL argument");
    if (!std::isfinite(alpha) || !std::isfinite(beta)) return set_error(VCT_ERR_ARG, "alpha/beta must be finite");
    return launch_beam_hardening(d_in, d_out, n, alpha, beta, (hipStream_t)stream);
}

vct_status vct_minmax(const float* d_x, int n, float* h_maxmin, void* stream)
{
    clear_error();
    if (!d_x || !h_maxmin || n < 1) return set_error(VCT_ERR_ARG, "bad argument");
    hipStream_t s = (hipStream_t)stream;
    uint32_t* d_st = nullptr;
    VCT_HIP(hipMallocAsync((void**)&d_st, 2 * sizeof(uint32_t), s));
    vct_status st = launch_minmax(d_x, n, d_st, s);
    if (st == VCT_OK) {
        hipError_t e = hipMemcpyAsync(h_maxmin, d_st, 2 * sizeof(float), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) st = set_error(VCT_ERR_HIP, "minmax copy: %s", hipGetErrorString(e));
    }
    (void)hipFreeAsync(d_st, s);
    return st;
}

static BackprojArgs make_bp_args(const vct_plan* p, const float* d_fltd, float* d_img)
{
    const vct_geometry& g = p->g;
    BackprojArgs a{};
    a.fltd = d_fltd;
    a.cosT = p->d_cos;
    a.sinT = p->d_sin;
    a.img = d_img;
    a.N = g.recon_size;
    a.W = g.n_channels;
    a.V = g.n_views;
    a.use_views = (int)(g.n_views / 2 * 2);
    a.cent = p->cent;
    a.half = p->half;
    a.pix = g.pixel_cm;
    a.sid = g.sid_cm;
    a.dbeta_ch = g.max_beta / p->half;
    a.inv_dbeta_ch = 1.0 / a.dbeta_ch;
    a.dbeta_view = 2.0 * kPi / (g.n_views);
    a.xfactor = g.xfactor;
    a.center = (double)g.recon_size / 2.0 - 0.5;
    a.mask_r = (double)(g.recon_size - 1) / 2.0;
    a.mask_r2 = a.mask_r * a.mask_r;
    a.exact = p->exact_bp;
    a.max_k = p->bp_max_k;
    return a;
}

vct_status vct_backproject(vct_plan* p, const float* d_fltd, int n_sino, float* d_img, void* stream)
{
    clear_error();
    if (!p || !d_fltd || !d_img) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_sino < 1 || n_sino > 65535) return set_error(VCT_ERR_ARG, "n_sino %d invalid", n_sino);
    return launch_backproject(make_bp_args(p, d_fltd, d_img), n_sino, pick_stream(p, stream));
}

static vct_status backproject_groups(vct_plan* const* plans, const float* const* d_fltd, const int* n_sino,
                                     float* const* d_img, int n_groups, int sinos_per_lane, bool il, void* stream)
{
    if (!plans || !d_fltd || !n_sino || !d_img) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_groups < 1 || n_groups > kMaxBpGroups)
        return set_error(VCT_ERR_ARG, "n_groups %d outside 1..%d", n_groups, kMaxBpGroups);
    static_assert(sizeof(BpGroups) <= 4096, "kernel arguments");
    BpGroups ga{};
    ga.n = n_groups;
    int smax = 0;
    for (int k = 0; k < n_groups; k++) {
        if (!plans[k] || !d_fltd[k] || !d_img[k]) return set_error(VCT_ERR_ARG, "group %d: NULL argument", k);
        if (n_sino[k] < 1 || n_sino[k] > 65535) return set_error(VCT_ERR_ARG, "group %d: n_sino %d invalid", k, n_sino[k]);
        if (plans[k]->ctx != plans[0]->ctx) return set_error(VCT_ERR_ARG, "group %d: plan of another context", k);
        ga.a[k] = make_bp_args(plans[k], d_fltd[k], d_img[k]);
        ga.S[k] = n_sino[k];
        smax = std::max(smax, n_sino[k]);
    }
    int K = sinos_per_lane > 0 ? sinos_per_lane : 16;
    if (!il) K = std::min(K, std::max(smax, 2));
    return launch_backproject_groups(ga, K, (hipStream_t)stream, il);
}

vct_status vct_backproject_groups(vct_plan* const* plans, const float* const* d_fltd, const int* n_sino,
                                  float* const* d_img, int n_groups, int sinos_per_lane, void* stream)
{
    clear_error();
    return backproject_groups(plans, d_fltd, n_sino, d_img, n_groups, sinos_per_lane, false, stream);
}

static bool il_width(int k) { return k == 4 || k == 8 || k == 12 || k == 16; }

size_t vct_interleaved_bytes(const vct_plan* p, int n_sino, int k)
{
    if (!p || n_sino < 1 || !il_width(k)) return 0;
    return (size_t)((n_sino + k - 1) / k) * k * p->g.n_views * p->g.n_channels * sizeof(float);
}

vct_status vct_interleave_sinograms(const vct_plan* p, const float* d_fltd, int n_sino, int k, float* d_il,
                                    void* stream)
{
    clear_error();
    if (!p || !d_fltd || !d_il) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_sino < 1 || n_sino > 65535) return set_error(VCT_ERR_ARG, "n_sino %d invalid", n_sino);
    if (!il_width(k)) return set_error(VCT_ERR_ARG, "slices of %d sinograms (4, 8, 12 or 16)", k);
    return launch_interleave_sinos(d_fltd, n_sino, p->g.n_views * p->g.n_channels, k, d_il, (hipStream_t)stream);
}

vct_status vct_backproject_groups_il(vct_plan* const* plans, const float* const* d_il, const int* n_sino,
                                     float* const* d_img, int n_groups, int sinos_per_lane, void* stream)
{
    clear_error();
    if (!il_width(sinos_per_lane))
        return set_error(VCT_ERR_ARG, "sinos_per_lane %d: the interleaved slices hold 4, 8, 12 or 16", sinos_per_lane);
    return backproject_groups(plans, d_il, n_sino, d_img, n_groups, sinos_per_lane, true, stream);
}

vct_status vct_lengthproj(vct_plan* p, const float* d_proj, double z_cm, double couch_speed_cm, double iso_det_cm,
                          float* d_npf, void* stream)
{
    if (!p) return set_error(VCT_ERR_ARG, "NULL argument");
    return vct_lengthproj_window(p, d_proj, p->g.n_channels, z_cm, couch_speed_cm, iso_det_cm, d_npf, stream);
}

vct_status vct_lengthproj_window(vct_plan* p, const float* d_proj, int proj_width, double z_cm, double couch_speed_cm,
                                 double iso_det_cm, float* d_npf, void* stream)
{
    clear_error();
    if (!p || !d_proj || !d_npf) return set_error(VCT_ERR_ARG, "NULL argument");
    if (proj_width < 2 || proj_width > p->g.n_channels)
        return set_error(VCT_ERR_ARG, "window width %d outside 2..%d (the detector's channels)", proj_width,
                         p->g.n_channels);
    const vct_geometry& g = p->g;
    const int V = g.n_views;
    // view window of M/lengthproj_TOMO.cpp:35-49 (UsedReconAngle = 2)
    const double pitch = (double)(couch_speed_cm / V);
    const double s = z_cm / pitch - V / 2 + 0.00000001;
    if (!(s > -1.0) || !(s < 2147483647.0 - V))
        return set_error(VCT_ERR_ARG,
                         "view window starts at %g (z %g cm, couch %g cm/rotation): it must not precede view 0", s,
                         z_cm, couch_speed_cm);
    if (!(iso_det_cm > 0.0) || !std::isfinite(iso_det_cm))
        return set_error(VCT_ERR_ARG, "isocentre-to-detector distance must be > 0");
    const int start = (int)s;
    LengthprojArgs a{};
    a.proj = d_proj;
    a.cosT = p->d_cos;
    a.sinT = p->d_sin;
    a.npf = d_npf;
    a.N = g.recon_size;
    a.W = g.n_channels;
    a.pw = proj_width;
    a.V = V;
    a.base = start - (start / V) * V;
    a.count = V / 2 * 2;
    a.pix = g.pixel_cm;
    a.rz = z_cm;
    a.sid = g.sid_cm;
    a.det_src_iso = g.det_src_iso_cm;
    a.det_radius = g.det_radius_cm;
    a.iso_det = iso_det_cm;
    a.iso_channel = g.iso_channel;
    a.delta = g.max_alpha / p->half;
    return launch_lengthproj(a, pick_stream(p, stream));
}

vct_status vct_trace_adjoint(vct_plan* p, const void* d_packed, int dens_type, int M, const float* d_dIdL,
                             const float* d_g, float* d_grad, void* stream)
{
    clear_error();
    if (!p || !d_dIdL || !d_grad) return set_error(VCT_ERR_ARG, "NULL argument");
    if (d_packed && dens_type != VCT_DENS_F32)
        return set_error(VCT_ERR_ARG, "the adjoint sign test needs float densities (or pass NULL)");
    TraceArgs a = make_trace_args(p);
    a.dens = d_packed;
    a.wts = d_dIdL;
    a.g = d_g;
    a.out_scale = (dens_type == VCT_DENS_U16 ? 1.0 / 1000.0 : 1.0) / p->g.xfactor;
    a.grad = d_grad;
    return launch_trace(a, VCT_DENS_F32, M, true, pick_stream(p, stream));
}

// ------------------------------------------------------------ host drop-ins

vct_status vct_reprojection(vct_ctx* ctx, const vct_geometry* g, const float* dens, int M, const float* att,
                            const float* pdf, int NE, float* sino_full)
{
    clear_error();
    if (!ctx || !dens || !att || !pdf || !sino_full) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    if (NE < 1 || NE > kMaxBins) return set_error(VCT_ERR_ARG, "NE=%d outside 1..%d", NE, kMaxBins);
    vct_plan* p = nullptr;
    VCT_CHECK(vct_plan_create(ctx, g, &p));
    const size_t NN = (size_t)g->recon_size * g->recon_size;
    const size_t nr = (size_t)g->n_views * g->n_channels;
    hipStream_t s = ctx->stream;
    float *d_dens = nullptr, *d_pk = nullptr, *d_att = nullptr, *d_pdf = nullptr, *d_L = nullptr, *d_raw = nullptr;
    uint32_t* d_stats = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_dens, NN * M * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_pk, NN * 8 * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_att, (size_t)M * NE * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_pdf, (size_t)NE * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_L, nr * 8 * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_raw, nr * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_stats, 2 * sizeof(uint32_t)), "hipMalloc");
    if (st == VCT_OK) {
        hip(hipMemcpyAsync(d_dens, dens, NN * M * sizeof(float), hipMemcpyHostToDevice, s), "copy");
        hip(hipMemcpyAsync(d_att, att, (size_t)M * NE * sizeof(float), hipMemcpyHostToDevice, s), "copy");
        hip(hipMemcpyAsync(d_pdf, pdf, (size_t)NE * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    }
    if (st == VCT_OK) st = vct_pack_densities_f32(p, d_dens, M, d_pk, s);
    if (st == VCT_OK) st = vct_trace_paths(p, d_pk, VCT_DENS_F32, M, d_L, nullptr, s);
    if (st == VCT_OK) st = vct_spectral_sum(p, d_L, M, d_att, d_pdf, NE, 1, d_raw, d_stats, nullptr, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(sino_full, d_raw, nr * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (st == VCT_OK) st = check_err_flag(ctx, s);
    void* bufs[] = {d_dens, d_pk, d_att, d_pdf, d_L, d_raw, d_stats};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    vct_plan_destroy(p);
    return st;
}

vct_status vct_gaussian_noise(vct_ctx* ctx, float* sino, int n, double sigma, uint64_t seed, float* stats_out)
{
    // M/main.cpp:511-525: max-normalise then add noise, on the device.
    clear_error();
    if (!ctx || !sino || n < 1) return set_error(VCT_ERR_ARG, "bad argument");
    if (!(sigma >= 0.0) || !std::isfinite(sigma)) return set_error(VCT_ERR_ARG, "sigma must be finite and >= 0");
    hipStream_t s = ctx->stream;
    VCT_HIP(hipSetDevice(ctx->device));
    float *d_in = nullptr, *d_out = nullptr;
    uint32_t* d_stats = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_in, (size_t)n * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_out, (size_t)n * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_stats, 2 * sizeof(uint32_t)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_in, sino, (size_t)n * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    // projection_max / projection_min of M/main.cpp:512-518 and :53-60
    if (st == VCT_OK) st = launch_minmax(d_in, n, d_stats, s);
    if (st == VCT_OK) {
        DetectorArgs a{};
        a.raw = d_in;
        a.stats = d_stats;
        a.out = d_out;
        a.n = n;
        a.S = 1;
        a.sigma = sigma;
        a.key0 = (uint32_t)seed;
        a.key1 = (uint32_t)(seed >> 32);
        a.stream_id = 0;
        st = launch_detector(a, s);
    }
    if (st == VCT_OK) hip(hipMemcpyAsync(sino, d_out, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK && stats_out)
        hip(hipMemcpyAsync(stats_out, d_stats, 2 * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    void* bufs[] = {d_in, d_out, d_stats};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    return st;
}

vct_status vct_filter_sinogram_host(vct_ctx* ctx, const vct_geometry* g, const float* proj, int filter, int in_mode,
                                    float* fltd)
{
    clear_error();
    if (!ctx || !proj || !fltd) return set_error(VCT_ERR_ARG, "NULL argument");
    vct_plan* p = nullptr;
    VCT_CHECK(vct_plan_create(ctx, g, &p));
    const size_t nr = (size_t)g->n_views * g->n_channels;
    hipStream_t s = ctx->stream;
    float *d_p = nullptr, *d_f = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_p, nr * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_f, nr * sizeof(float)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_p, proj, nr * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) st = vct_filter_sinogram(p, d_p, 1, filter, in_mode, d_f, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(fltd, d_f, nr * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (d_p) (void)hipFree(d_p);
    if (d_f) (void)hipFree(d_f);
    vct_plan_destroy(p);
    return st;
}

vct_status vct_reconstruct_image_host(vct_ctx* ctx, const vct_geometry* g, const float* fltd, float* img)
{
    clear_error();
    if (!ctx || !fltd || !img) return set_error(VCT_ERR_ARG, "NULL argument");
    vct_plan* p = nullptr;
    VCT_CHECK(vct_plan_create(ctx, g, &p));
    const size_t nr = (size_t)g->n_views * g->n_channels;
    const size_t NN = (size_t)g->recon_size * g->recon_size;
    hipStream_t s = ctx->stream;
    float *d_f = nullptr, *d_i = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_f, nr * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_i, NN * sizeof(float)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_f, fltd, nr * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) st = vct_backproject(p, d_f, 1, d_i, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(img, d_i, NN * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (d_f) (void)hipFree(d_f);
    if (d_i) (void)hipFree(d_i);
    vct_plan_destroy(p);
    return st;
}

vct_status vct_lengthproj_host(vct_ctx* ctx, const vct_geometry* g, const float* proj, double z_cm,
                               double couch_speed_cm, double iso_det_cm, float* npf)
{
    if (!g) return set_error(VCT_ERR_ARG, "NULL argument");
    return vct_lengthproj_window_host(ctx, g, proj, g->n_channels, z_cm, couch_speed_cm, iso_det_cm, npf);
}

vct_status vct_lengthproj_window_host(vct_ctx* ctx, const vct_geometry* g, const float* proj, int proj_width,
                                      double z_cm, double couch_speed_cm, double iso_det_cm, float* npf)
{
    clear_error();
    if (!ctx || !g || !proj || !npf) return set_error(VCT_ERR_ARG, "NULL argument");
    if (proj_width < 2 || proj_width > g->n_channels)
        return set_error(VCT_ERR_ARG, "window width %d outside 2..%d (the detector's channels)", proj_width,
                         g->n_channels);
    vct_plan* p = nullptr;
    VCT_CHECK(vct_plan_create(ctx, g, &p));
    const size_t nr = (size_t)(g->n_views / 2 * 2) * proj_width;
    const size_t NN = (size_t)g->recon_size * g->recon_size;
    hipStream_t s = ctx->stream;
    float *d_p = nullptr, *d_n = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    hip(hipMalloc(&d_p, nr * sizeof(float)), "hipMalloc");
    hip(hipMalloc(&d_n, NN * sizeof(float)), "hipMalloc");
    if (st == VCT_OK) hip(hipMemcpyAsync(d_p, proj, nr * sizeof(float), hipMemcpyHostToDevice, s), "copy");
    if (st == VCT_OK) st = vct_lengthproj_window(p, d_p, proj_width, z_cm, couch_speed_cm, iso_det_cm, d_n, s);
    if (st == VCT_OK) hip(hipMemcpyAsync(npf, d_n, NN * sizeof(float), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (d_p) (void)hipFree(d_p);
    if (d_n) (void)hipFree(d_n);
    vct_plan_destroy(p);
    return st;
}

}  // extern "C"
