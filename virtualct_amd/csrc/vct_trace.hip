// vct_trace.hip -- MBFPA ray tracer, polyenergetic detector sum, detector
// noise and phantom packing for CDNA4 (gfx950).
//
// Design (64 adjacent channels of one view per wave):
//   k_trace_fwd  marches each fan ray once through the image in FP64 and
//                accumulates the per-element path integrals
//                L_m = sum_seg rho_m(voxel) * length in FP64 registers; the
//                segment decomposition (entry/exit clip, integer-cm start
//                plane, midpoint voxel, minor-axis crossing split) follows
//                M/reprojection_TOMO.cpp:53-283 operation for operation, so
//                it is bit-identical to the reference when compiled with
//                -ffp-contract=off.  The reference re-traces every ray once
//                per energy bin (M/main.cpp:492); here the trace is done once
//                and the spectrum is applied afterwards.
//   k_spectral   I_s = sum_E pdf_s(E) exp(-sum_m mu_s,m(E) L_m) with the
//                spectrum tables of spectrum s in LDS, plus the per-spectrum
//                max/min used by the normalisation and the noise clamp.
//   k_detector   max-normalisation + Gaussian noise (Philox4x32-10 +
//                Box-Muller) + the reference's clamps.
//   k_trace_fwd_rot  the quarter- (or half-) turn orbit form: one march
//                for the rays of views v, v + V/4, v + V/2, v + 3V/4, split
//                over eight lanes by step windows.
//   k_trace_adj_tile  the adjoint by image tiles: windowed marches of the
//                rays that cross a tile, w_m * length summed in LDS in 64-bit
//                fixed point, one float atomic per voxel at the end.
#include "vct_internal.h"
#include "vct_march.h"

#include <algorithm>
#include <climits>
#include <cstring>
#include <cmath>

namespace vct {

// ------------------------------------------------------------------ sinks

// Forward sinks accumulate L_m += rho_m(voxel) * length in FP64 (the u16
// sink per 32-bit word of two materials: the word as one integer and its high
// half, which resolves the low half to ~1e-11 relative).  add2
// takes the two candidate segments of one march step and issues both gathers
// before using either, each only on the lanes where it is a segment in the
// image: the gathers' texture path, not the FP64 arithmetic, is what the
// march keeps busiest (78% / 84% TA / TD busy with unmasked gathers).
template <int M>
struct SinkU16 {
    static constexpr bool kAdd2 = true;
    const char* __restrict__ vox;
    // word k of a voxel holds materials 2k (low half) and 2k+1 (high half);
    // accw[k] sums the whole word as one integer lo + 65536 hi and acch[k]
    // the high half, so material 2k is accw[k] - 65536 acch[k] at the end
    double accw[(M + 1) / 2], acch[M / 2];
    double acc[M];
    __device__ __forceinline__ void init(const void* p)
    {
        vox = reinterpret_cast<const char*>(p);
#pragma unroll
        for (int k = 0; k < (M + 1) / 2; k++) accw[k] = 0.0;
#pragma unroll
        for (int k = 0; k < M / 2; k++) acch[k] = 0.0;
    }
    __device__ __forceinline__ uint4 load(int v) const
    {
        // 32-bit byte offset (N*N*16 bytes < 4 GB)
        return *reinterpret_cast<const uint4*>(vox + (uint32_t)v * 16u);
    }
    __device__ __forceinline__ void accumulate(const uint4& q, double rr)
    {
        const uint32_t w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
        for (int k = 0; k < (M + 1) / 2; k++) {
            // odd M: the last word's high half is not a material of this call
            const uint32_t wk = (2 * k + 1 < M) ? w[k] : (w[k] & 0xffffu);
            accw[k] = __builtin_fma((double)wk, rr, accw[k]);
        }
#pragma unroll
        for (int k = 0; k < M / 2; k++) acch[k] = __builtin_fma((double)(w[k] >> 16), rr, acch[k]);
    }
    __device__ __forceinline__ void add(int v, double rr) { accumulate(load(v), rr); }
    // the two candidate segments of a committed step, each only on lanes
    // where it is a segment (ok1, ok2): the others issue no gather
    __device__ __forceinline__ void add2(bool ok1, int v1, double r1, bool ok2, int v2, double r2)
    {
        uint4 q1, q2;
        if (ok1) q1 = load(v1);
        if (ok2) q2 = load(v2);
        if (ok1) accumulate(q1, r1);
        if (ok2) accumulate(q2, r2);
    }
    __device__ __forceinline__ void finish()
    {
#pragma unroll
        for (int m = 0; m < M; m++)
            acc[m] = (m & 1) ? acch[m >> 1] : accw[m >> 1] - (m / 2 < M / 2 ? 65536.0 * acch[m >> 1] : 0.0);
    }
};

template <int M>
struct SinkF32 {
    static constexpr bool kAdd2 = true;
    const float4* __restrict__ vox;
    double acc[M];
    struct Vox {
        float4 lo, hi;
    };
    __device__ __forceinline__ void init(const void* p)
    {
        vox = reinterpret_cast<const float4*>(p);
#pragma unroll
        for (int m = 0; m < M; m++) acc[m] = 0.0;
    }
    __device__ __forceinline__ Vox load(int v) const
    {
        Vox q;
        q.lo = vox[2 * v];
        q.hi = M > 4 ? vox[2 * v + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
        return q;
    }
    __device__ __forceinline__ void accumulate(const Vox& q, double rr)
    {
        const float r[8] = {q.lo.x, q.lo.y, q.lo.z, q.lo.w, q.hi.x, q.hi.y, q.hi.z, q.hi.w};
#pragma unroll
        for (int m = 0; m < M; m++) {
            // a negative density contributes nothing (reference ph = 0, :155):
            // fmaxf gives 0 for negatives and NaN, and -0 adds the same zero
            const double rho = (double)fmaxf(r[m], 0.f);
            acc[m] = __builtin_fma(rho, rr, acc[m]);
        }
    }
    __device__ __forceinline__ void add(int v, double rr) { accumulate(load(v), rr); }
    __device__ __forceinline__ void add2(bool ok1, int v1, double r1, bool ok2, int v2, double r2)
    {
        Vox q1, q2;
        if (ok1) q1 = load(v1);
        if (ok2) q2 = load(v2);
        if (ok1) accumulate(q1, r1);
        if (ok2) accumulate(q2, r2);
    }
};

// Orbit sink: the segments of the ray of view v, taken also as the segments
// of the rays of views v + kV/R (k = 1..R-1, turned by k/R of a turn), each
// at the turned voxel.  Ray k accumulates rho(voxel) * length as the
// single-ray sinks do.  The march hands over, for every candidate segment,
// the floors (fa, fb) of its major and minor index values; the x and y floors
// (fx, fy) of ray k are those of ray 0 turned: a quarter turn maps (fx, fy)
// to (N-1-fy, fx).  Each ray then takes the reference's own rule: index
// trunc(q) = f + (f < 0), in the image when both lie in 0..N-1, i.e.
// f in -1..N-1 (a midpoint less than a pixel beyond the low edge counts into
// the edge voxel, one beyond the high edge does not).
// unsafe: a decision of the march that a turned ray might take otherwise did
// not clear the margin (the caller then traces rays 1..R-1 by themselves;
// ray 0 is right in any case).
__host__ __device__ __forceinline__ void swap_d(double& x, double& y)
{
    const double t = x;
    x = y;
    y = t;
}

template <int M, int DT, int R>
struct SinkRot {
    static_assert(R == 1 || R == 2 || R == 4, "orbits of 1, 2 or 4 rays");
    static constexpr bool kAdd2 = true;
    static constexpr bool kRot = true;
    static constexpr int NW = (M + 1) / 2;  // u16: 32-bit words of two materials
    // Bytes per voxel of the work copies: u16 with M <= 6 keeps the three
    // words that hold the materials, other u16 all four, f32 eight floats.
    static constexpr int VB = DT == VCT_DENS_U16 ? (M <= 6 ? 12 : 16) : 32;
    // Gathers read the copy of the densities whose rows run along the
    // minor axis of the slot's ray ([y][x] voxels when y is its major axis,
    // the transposed [x][y] copy when x is), so the lanes of a wave, which
    // hold neighbouring rays at the same major step, meet neighbouring voxels
    // of one row.  Slot k holds ray k (the base ray turned k quarter turns),
    // except that rays 1 and 3 trade slots when y is the base ray's major
    // axis; finish() puts them back.  Both copies are padded by two voxels
    // on every side (NP = N + 4 per row): row and column -1 repeat row and
    // column 0 (the reference truncates index -1 to voxel 0), rows and
    // columns -2, N and N + 1 are zero, so the committed steps' gathers need
    // no in-image test (a voxel outside the image reads zeros).
    const char* __restrict__ voxE;  // slots 0 and 2 (and 1 when R == 2)
    const char* __restrict__ voxO;  // slots 1 and 3 when R == 4
    bool xmajor;
    int N;
    int NP;   // row pitch of the copies in voxels, N + 4
    int ORG;  // index of voxel (0, 0), 2 NP + 2
    bool unsafe;
    double accw[R][NW], acch[R][M / 2];  // u16: word sums and high halves
    double acc[R][M];                    // f32: sums; u16: after finish()
    // pt: the two padded copies, [y][x] then [x][y]
    __host__ __device__ void init(const void* pt, bool xm, int n)
    {
        NP = n + 4;
        ORG = 2 * NP + 2;
        const char* P = reinterpret_cast<const char*>(pt);
        const char* T = P + (size_t)NP * NP * VB;
        voxE = xm ? T : P;
        voxO = xm ? P : T;
        xmajor = xm;
        N = n;
        unsafe = false;
        #pragma unroll
        for (int k = 0; k < R; k++) {
            #pragma unroll
            for (int w = 0; w < NW; w++) accw[k][w] = 0.0;
            #pragma unroll
            for (int w = 0; w < M / 2; w++) acch[k][w] = 0.0;
            #pragma unroll
            for (int m = 0; m < M; m++) acc[k][m] = 0.0;
        }
    }
    // slot k of a segment whose floors are (fa, fb) on the base ray's (major,
    // minor) axes: the (major, minor) coordinates of the turned ray, indexed
    // minor + NP * major in its copy, with the reference's truncation of
    // index -1 to voxel 0 (0 with rr = 0 when outside the image)
    __host__ __device__ __forceinline__ void voxels(int fa, int fb, double rr, int (&v)[R], double (&r)[R]) const
    {
        const int n1 = N - 1;
        const int ta = n1 - fa, tb = n1 - fb;
        #pragma unroll
        for (int k = 0; k < R; k++) {
            int ma, mi;
            if (R == 1 || k == 0) { ma = fa; mi = fb; }
            else if (R == 2 || k == 2) { ma = ta; mi = tb; }
            else if (k == 1) { ma = fa; mi = tb; }
            else { ma = ta; mi = fb; }
            const bool in = (unsigned)(ma + 1) <= (unsigned)N && (unsigned)(mi + 1) <= (unsigned)N;
            v[k] = in ? ORG + max(mi, 0) + mul24(max(ma, 0), NP) : 0;
            r[k] = in ? rr : 0.0;
        }
    }
    struct Vox {
        uint4 lo, hi;  // u16: lo = the 8 materials; f32: lo = materials 0-3, hi = 4-7
    };
    // the voxel at byte offset o of the copy of slot k
    __host__ __device__ __forceinline__ Vox load_at(int k, uint32_t o) const
    {
        const char* vox = (R == 4 && (k & 1)) ? voxO : voxE;
        Vox q;
        if constexpr (VB == 12) {
            const uint3 t = *reinterpret_cast<const uint3*>(vox + o);
            q.lo = make_uint4(t.x, t.y, t.z, 0u);
            q.hi = make_uint4(0u, 0u, 0u, 0u);
        } else if constexpr (VB == 16) {
            q.lo = *reinterpret_cast<const uint4*>(vox + o);
            q.hi = make_uint4(0u, 0u, 0u, 0u);
        } else {
            const uint4* f = reinterpret_cast<const uint4*>(vox + o);
            q.lo = f[0];
            q.hi = M > 4 ? f[1] : make_uint4(0u, 0u, 0u, 0u);
        }
        return q;
    }
    __host__ __device__ __forceinline__ Vox load(int k, int v) const { return load_at(k, (uint32_t)v * (uint32_t)VB); }
    __host__ __device__ __forceinline__ void accumulate(int k, const Vox& q, double rr)
    {
        if constexpr (DT == VCT_DENS_U16) {
            const uint32_t w[4] = {q.lo.x, q.lo.y, q.lo.z, q.lo.w};
            #pragma unroll
            for (int j = 0; j < NW; j++) {
                const uint32_t wj = (2 * j + 1 < M) ? w[j] : (w[j] & 0xffffu);
                accw[k][j] = __builtin_fma((double)wj, rr, accw[k][j]);
            }
            #pragma unroll
            for (int j = 0; j < M / 2; j++) acch[k][j] = __builtin_fma((double)(w[j] >> 16), rr, acch[k][j]);
        } else {
            const uint32_t b[8] = {q.lo.x, q.lo.y, q.lo.z, q.lo.w, q.hi.x, q.hi.y, q.hi.z, q.hi.w};
            #pragma unroll
            for (int m = 0; m < M; m++) {
                float rho;
                memcpy(&rho, &b[m], sizeof(float));
                acc[k][m] = __builtin_fma((double)fmaxf(rho, 0.f), rr, acc[k][m]);  // as SinkF32
            }
        }
    }
    // one candidate segment whose index values clear the margin (floors)
    __host__ __device__ __forceinline__ void add_fl(int fa, int fb, double rr)
    {
        int v[R];
        double r[R];
        voxels(fa, fb, rr, v, r);
        Vox q[R];
        #pragma unroll
        for (int k = 0; k < R; k++) q[k] = load(k, v[k]);
        #pragma unroll
        for (int k = 0; k < R; k++) accumulate(k, q[k], r[k]);
    }
    // one candidate segment of a general march step from the signed quotients
    // (ta, tb) of its index values t + N/2: ray k takes fl(+-t + N/2) on each
    // of its axes (minus where the turn flips the axis) and the reference's
    // trunc and in-image test
    __host__ __device__ __forceinline__ void add_t(double ta, double tb, double rr)
    {
        int v[R];
        double r[R];
        voxels_t(ta, tb, rr, v, r);
        Vox q[R];
        #pragma unroll
        for (int k = 0; k < R; k++) q[k] = load(k, v[k]);
        #pragma unroll
        for (int k = 0; k < R; k++) accumulate(k, q[k], r[k]);
    }
    __host__ __device__ __forceinline__ void voxels_t(double ta, double tb, double rr, int (&v)[R], double (&r)[R]) const
    {
        const double tx = xmajor ? ta : tb, ty = xmajor ? tb : ta;
        const double c = (double)(N / 2);
        #pragma unroll
        for (int k = 0; k < R; k++) {
            double gx, gy;
            if (R == 1 || k == 0) { gx = tx + c; gy = ty + c; }
            else if (R == 2 || k == 2) { gx = -tx + c; gy = -ty + c; }
            else if (k == 1) { gx = -ty + c; gy = tx + c; }
            else { gx = ty + c; gy = -tx + c; }
            const int ix = trunc_i32(gx), iy = trunc_i32(gy);
            const bool in = ix >= 0 && ix <= N - 1 && iy >= 0 && iy <= N - 1;
            // ray k goes to slot k, but rays 1 and 3 trade slots when y is the
            // base ray's major axis; its major axis is x when xmajor differs
            // from its own quarter-turn parity
            const bool odd = R == 4 && (k & 1);
            const int slot = (odd && !xmajor) ? 4 - k : k;
            const bool majx = xmajor != odd;
            v[slot] = in ? ORG + (majx ? iy + NP * ix : ix + NP * iy) : 0;
            r[slot] = in ? rr : 0.0;
        }
    }
    // byte offsets of the four slots of one candidate, its coordinates
    // clamped to [-2, N + 1] (values outside read zeros there too)
    __host__ __device__ __forceinline__ void offsets(int fa, int fb, uint32_t (&o)[R]) const
    {
        const int a = min(max(fa + 2, 0), N + 3), b = min(max(fb + 2, 0), N + 3);
        const int ta = N + 3 - a, tb = N + 3 - b;  // (N - 1 - f) + 2
        const uint32_t ra = (uint32_t)mul24(a, VB * NP), rt = (uint32_t)mul24(ta, VB * NP);
        const uint32_t cb = (uint32_t)mul24(b, VB), ct = (uint32_t)mul24(tb, VB);
        #pragma unroll
        for (int k = 0; k < R; k++) {
            if (R == 1 || k == 0) o[k] = ra + cb;
            else if (R == 2 || k == 2) o[k] = rt + ct;
            else if (k == 1) o[k] = ra + ct;
            else o[k] = rt + cb;
        }
    }
    // the two candidate segments of a committed step; the first only on lanes
    // whose step crosses a minor grid line (cross), so lanes without a
    // crossing issue no gathers for it (the gathers' texture path is what the
    // kernel keeps busiest); a segment of length 0 would add exactly nothing
    __host__ __device__ __forceinline__ void add2_fl(bool cross, int fa1, int fb1, double r1, int fa2, int fb2,
                                                     double r2)
    {
        uint32_t o1[R], o2[R];
        offsets(fa1, fb1, o1);
        offsets(fa2, fb2, o2);
        Vox q1[R], q2[R];
        if (cross) {
            #pragma unroll
            for (int k = 0; k < R; k++) q1[k] = load_at(k, o1[k]);
        }
        #pragma unroll
        for (int k = 0; k < R; k++) q2[k] = load_at(k, o2[k]);
        if (cross) {
            #pragma unroll
            for (int k = 0; k < R; k++) accumulate(k, q1[k], r1);
        }
        #pragma unroll
        for (int k = 0; k < R; k++) accumulate(k, q2[k], r2);
    }
    __host__ __device__ __forceinline__ void finish()
    {
        if (R == 4 && !xmajor) {
            // slots 1 and 3 hold rays 3 and 1
            #pragma unroll
            for (int w = 0; w < NW; w++) swap_d(accw[1 % R][w], accw[3 % R][w]);
            #pragma unroll
            for (int w = 0; w < M / 2; w++) swap_d(acch[1 % R][w], acch[3 % R][w]);
            #pragma unroll
            for (int m = 0; m < M; m++) swap_d(acc[1 % R][m], acc[3 % R][m]);
        }
        if constexpr (DT == VCT_DENS_U16) {
            #pragma unroll
            for (int k = 0; k < R; k++)
                #pragma unroll
                for (int m = 0; m < M; m++)
                    acc[k][m] = (m & 1) ? acch[k][m >> 1]
                                        : accw[k][m >> 1] - (m / 2 < M / 2 ? 65536.0 * acch[k][m >> 1] : 0.0);
        }
    }
};

// Adjoint sink of one image tile: a segment whose voxel lies in the tile adds
// w[m] * length to the tile's accumulator in LDS, [m][row][col] with a row
// pitch of TS + 1 (a wave's lanes are adjacent channels of one view, which
// meet a column of the tile at one step when x is the major axis).  FIXED:
// the accumulator is 64-bit fixed point, each contribution rounded to an
// integer after scaling by a power of two that keeps it below 2^30 in
// magnitude; the sums are then exact (and independent of the order of the
// adds), and ds_add_u64 runs about ten times faster than ds_add_f32 on
// gfx950 (0.3 vs 3.4 lane-operations per clock per CU, measured).  The float
// form serves workgroups whose weights are not all finite, so NaN and
// infinity reach the gradient as with float sums.
template <int M>
constexpr int adj_tile()
{
    return M <= 6 ? 32 : 28;  // M x TS x (TS + 1) x 8 B of LDS stays near 50 KB
}
constexpr int kAdjMaxViews = 256;  // views per workgroup

template <int M, bool FIXED>
struct SinkTile {
    static constexpr bool kAdd2 = false;
    static constexpr bool kAB = true;
    static constexpr int TS = adj_tile<M>(), PITCH = TS + 1, PLANE = TS * PITCH;
    unsigned long long* acc;  // [M][TS][PITCH], int64 or (FIXED false) float in the same memory
    int x0, y0;
    bool xmajor;
    float scale;  // 2^S (FIXED)
    float w[M];
    __device__ __forceinline__ void add_ab(int ia, int ib, double rr)
    {
        const int x = xmajor ? ia : ib, y = xmajor ? ib : ia;
        const uint32_t lx = (uint32_t)(x - x0), ly = (uint32_t)(y - y0);
        if (lx < (uint32_t)TS && ly < (uint32_t)TS) {
            const float r = (float)rr;
            const int off = (int)(ly * PITCH + lx);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const float c = w[m] * r;
                if constexpr (FIXED) {
                    const long long q = (long long)__float2int_rn(c * scale);
                    atomicAdd(acc + off + m * PLANE, (unsigned long long)q);
                } else {
                    atomicAdd(reinterpret_cast<float*>(acc) + off + m * PLANE, c);
                }
            }
        }
    }
};

// ---------------------------------------------------------------- kernels

template <int M, int DT, bool FAST>
__global__ __launch_bounds__(256) void k_trace_fwd(TraceArgs a)
{
    const int local = blockIdx.x * blockDim.x + threadIdx.x;
    if (local >= a.n_rays) return;
    const int ray = a.ray_offset + local;
    const int view = ray / a.W;
    const int ch = ray - view * a.W;
    float out[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int iks = 0;
    if (view < a.VT) {
        bool ovf = false;
        if (DT == VCT_DENS_U16) {
            SinkU16<M> sink;
            sink.init(a.dens);
            iks = march_ray<FAST>(a, view, ch, sink, ovf);
            sink.finish();
#pragma unroll
            for (int m = 0; m < M; m++) out[m] = (float)(sink.acc[m] * a.out_scale);
        } else {
            SinkF32<M> sink;
            sink.init(a.dens);
            iks = march_ray<FAST>(a, view, ch, sink, ovf);
#pragma unroll
            for (int m = 0; m < M; m++) out[m] = (float)(sink.acc[m] * a.out_scale);
        }
        if (ovf) atomicOr(a.err, 1);
    }
    float4* dst = reinterpret_cast<float4*>(a.L) + (size_t)ray * 2;
    dst[0] = make_float4(out[0], out[1], out[2], out[3]);
    dst[1] = make_float4(out[4], out[5], out[6], out[7]);
    if (a.counts) a.counts[ray] = iks;
}

// One ray of k_trace_fwd, written to L (and counts).
template <int M, int DT>
__device__ __forceinline__ void trace_one(const TraceArgs& a, int view, int ch)
{
    const int ray = view * a.W + ch;
    float out[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    int iks = 0;
    bool ovf = false;
    if (DT == VCT_DENS_U16) {
        SinkU16<M> sink;
        sink.init(a.dens);
        iks = march_ray<true>(a, view, ch, sink, ovf);
        sink.finish();
#pragma unroll
        for (int m = 0; m < M; m++) out[m] = (float)(sink.acc[m] * a.out_scale);
    } else {
        SinkF32<M> sink;
        sink.init(a.dens);
        iks = march_ray<true>(a, view, ch, sink, ovf);
#pragma unroll
        for (int m = 0; m < M; m++) out[m] = (float)(sink.acc[m] * a.out_scale);
    }
    if (ovf) atomicOr(a.err, 1);
    float4* dst = reinterpret_cast<float4*>(a.L) + (size_t)ray * 2;
    dst[0] = make_float4(out[0], out[1], out[2], out[3]);
    dst[1] = make_float4(out[4], out[5], out[6], out[7]);
    if (a.counts) a.counts[ray] = iks;
}

// Sum over the P lanes of a group (P a power of two dividing 64, groups
// aligned): every lane of the group gets the same sum.
template <int P>
__device__ __forceinline__ double group_sum(double x)
{
#pragma unroll
    for (int d = 1; d < P; d <<= 1) x += __shfl_xor(x, d);
    return x;
}
template <int P>
__device__ __forceinline__ int group_sum(int x)
{
#pragma unroll
    for (int d = 1; d < P; d <<= 1) x += __shfl_xor(x, d);
    return x;
}

// Orbits of R rays: the base ray (view v < V/R, channel) is marched once and
// gives the rays of views v + kV/R their path integrals from the turned
// voxels (SinkRot).  The march of one orbit is split over a group of P
// adjacent lanes: lane p runs steps [p len, (p + 1) len) of the windowed
// march (the last lane to the end), len = ceil(steps / P), and the group adds
// its partial sums and segment counts with cross-lane shuffles.  When a march
// piece reports a decision within the margin, or the ray needs the exact
// march, the group traces rays 1..R-1 (and the base ray, unless marched) one
// by one, one ray per lane.  Every ray's segments are the reference's; the
// orbit shares the base ray's lengths, which differ from a ray's own
// |da| * sec by rounding.  Every branch before a shuffle depends only on the
// orbit, so the P lanes of a group stay together through the shuffles.
template <int M, int DT, int R, int P>
__global__ __launch_bounds__(256) void k_trace_fwd_rot(TraceArgs a)
{
    static_assert(P >= 1 && P <= 64 && (P & (P - 1)) == 0, "P: a power of two");
    const int gl = blockIdx.x * blockDim.x + threadIdx.x;
    const int orbit = gl / P, piece = gl % P;
    const int Vq = a.V / R;
    if (orbit >= Vq * a.W) return;  // whole groups
    const int view = orbit / a.W;
    const int ch = orbit - view * a.W;
    RayConst r;
    bool safe = false;
    const bool hit = setup_ray(a, view, ch, r, &safe);
    bool done = false;
    if (safe && !hit) {
        // the whole orbit misses the image: no segment, path integrals 0
        for (int k = piece; k < R; k += P) {
            const int ray = (view + k * Vq) * a.W + ch;
            float4* dst = reinterpret_cast<float4*>(a.L) + (size_t)ray * 2;
            dst[0] = make_float4(0.f, 0.f, 0.f, 0.f);
            dst[1] = make_float4(0.f, 0.f, 0.f, 0.f);
            if (a.counts) a.counts[ray] = 0;
        }
        done = true;
    } else if (safe && r.dB != 0.0 && !a.rcheck && !a.exact) {
        SinkRot<M, DT, R> sink;
        sink.init(a.dens_t, r.xmajor, a.N);
        bool ovf = false;
        int k0 = 0, k1 = INT_MAX;
        if (P > 1) {
            const int len = (march_step_estimate(a, r) + P - 1) / P;
            k0 = piece * len;
            k1 = piece == P - 1 ? INT_MAX : k0 + len - 1;
        }
        int iks = march_loop<true>(a, r, sink, ovf, k0, k1);
        if (ovf) atomicOr(a.err, 1);
        if (P > 1) {
            iks = group_sum<P>(iks);
            sink.unsafe = group_sum<P>(sink.unsafe ? 1 : 0) != 0;
            if constexpr (DT == VCT_DENS_U16) {
#pragma unroll
                for (int k = 0; k < R; k++) {
#pragma unroll
                    for (int w = 0; w < SinkRot<M, DT, R>::NW; w++) sink.accw[k][w] = group_sum<P>(sink.accw[k][w]);
#pragma unroll
                    for (int w = 0; w < M / 2; w++) sink.acch[k][w] = group_sum<P>(sink.acch[k][w]);
                }
            } else {
#pragma unroll
                for (int k = 0; k < R; k++)
#pragma unroll
                    for (int m = 0; m < M; m++) sink.acc[k][m] = group_sum<P>(sink.acc[k][m]);
            }
        }
        sink.finish();
        // ray 0 is right in any case (its own march); the turned rays only
        // when no decision was within the margin
#pragma unroll
        for (int k = 0; k < R; k++) {
            if (k % P != piece || (k > 0 && sink.unsafe)) continue;
            const int ray = (view + k * Vq) * a.W + ch;
            float out[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < M; m++) out[m] = (float)(sink.acc[k][m] * a.out_scale);
            float4* dst = reinterpret_cast<float4*>(a.L) + (size_t)ray * 2;
            dst[0] = make_float4(out[0], out[1], out[2], out[3]);
            dst[1] = make_float4(out[4], out[5], out[6], out[7]);
            if (a.counts) a.counts[ray] = iks;
        }
        if (sink.unsafe) {
            for (int k = 1 + piece; k < R; k += P) trace_one<M, DT>(a, view + k * Vq, ch);
        }
        done = true;
    }
    if (!done) {
        for (int k = piece; k < R; k += P) trace_one<M, DT>(a, view + k * Vq, ch);
    }
}

// Adjoint by image tiles.  Workgroup (chunk, tile) takes the views of one
// chunk and one TS^2 tile of the image; for each view it finds the channels
// whose rays can cross the tile (box_channels), and each such ray runs only
// the march steps that can land in the tile (box_steps, the windowed
// march_loop), adding w[m] * length into the tile in LDS.  Each segment of
// each ray lands in exactly one tile, so the tiles of one chunk together add
// every segment once.  A first pass over the workgroup's rays takes the
// largest |w|, which sets the fixed-point scale (SinkTile); at the end the
// tile is added to grad with one float atomic per voxel and material (zero
// where the voxel's density of that material is negative, the tracer's sign
// test).  Work items are dealt to XCDs (workgroup g runs on XCD g % 8) in
// contiguous runs, tile fastest, so the workgroups of one XCD share view
// chunks and read their rays' weights through one L2.
template <int M>
__device__ __forceinline__ void adj_weights(const TraceArgs& a, int ray, float (&w)[M])
{
    const float gscale = (a.g ? a.g[ray] : 1.f) * (float)a.out_scale;
    const float4* wv = reinterpret_cast<const float4*>(a.wts) + (size_t)ray * 2;
    const float4 w0 = wv[0];
    const float4 w1 = M > 4 ? wv[1] : make_float4(0.f, 0.f, 0.f, 0.f);
    const float wa[8] = {w0.x, w0.y, w0.z, w0.w, w1.x, w1.y, w1.z, w1.w};
#pragma unroll
    for (int m = 0; m < M; m++) w[m] = wa[m] * gscale;
}

template <int M>
__global__ __launch_bounds__(256) void k_trace_adj_tile(TraceArgs a, int ntx, int ntiles, int cv, int per_xcd)
{
    using Fixed = SinkTile<M, true>;
    using Float = SinkTile<M, false>;
    constexpr int TS = Fixed::TS, PITCH = Fixed::PITCH, PLANE = Fixed::PLANE;
    __shared__ unsigned long long acc[M * PLANE];
    __shared__ int pref[kAdjMaxViews + 1];
    __shared__ int clo[kAdjMaxViews];
    __shared__ uint32_t wmax_bits, nonfinite;
    const int item = (int)(blockIdx.x & 7) * per_xcd + (int)(blockIdx.x >> 3);
    const int nchunk = (a.VT + cv - 1) / cv;
    if (item >= ntiles * nchunk) return;  // whole workgroup, before any barrier
    const int chunk = item / ntiles, tile = item - chunk * ntiles;
    const int x0 = (tile % ntx) * TS, y0 = (tile / ntx) * TS;
    const int v0 = chunk * cv;
    const int nv = min(cv, a.VT - v0);
    for (int i = threadIdx.x; i < M * PLANE; i += blockDim.x) acc[i] = 0ull;
    if (threadIdx.x == 0) {
        wmax_bits = 0u;
        nonfinite = 0u;
    }
    // the tile plus two voxels, in cm (x index i covers [(i - nxm) D, (i + 1 - nxm) D),
    // y index j covers (-(j + 1 - nxm) D, -(j - nxm) D])
    const double D = a.D;
    const double x_lo = (double)(x0 - 2 - a.nxm) * D, x_hi = (double)(x0 + TS + 2 - a.nxm) * D;
    const double y_lo = -(double)(y0 + TS + 2 - a.nxm) * D, y_hi = -(double)(y0 - 2 - a.nxm) * D;
    for (int i = threadIdx.x; i < nv; i += blockDim.x) {
        int lo, hi;
        box_channels(a, v0 + i, x_lo, x_hi, y_lo, y_hi, lo, hi);
        clo[i] = lo;
        pref[i + 1] = hi >= lo ? hi - lo + 1 : 0;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        pref[0] = 0;
        for (int i = 0; i < nv; i++) pref[i + 1] += pref[i];
    }
    __syncthreads();
    const int total = pref[nv];
    // view slot j with pref[j] <= i < pref[j + 1], and the ray of item i
    auto item_ray = [&](int i, int& view, int& ch) {
        int lo = 0, hi = nv - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pref[mid] <= i) lo = mid;
            else hi = mid - 1;
        }
        view = v0 + lo;
        ch = clo[lo] + (i - pref[lo]);
    };
    // pass 1: the largest |w| of the workgroup's rays (non-negative float bits
    // order as unsigned integers) and whether any weight is not finite
    {
        float wmax = 0.f;
        bool bad = false;
        for (int i = threadIdx.x; i < total; i += blockDim.x) {
            int view, ch;
            item_ray(i, view, ch);
            float w[M];
            adj_weights<M>(a, view * a.W + ch, w);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const float u = fabsf(w[m]);
                bad = bad || !(u <= 3.0e38f);
                wmax = fmaxf(wmax, u);
            }
        }
        if (wmax > 0.f) atomicMax(&wmax_bits, __float_as_uint(wmax));
        if (bad) atomicOr(&nonfinite, 1u);
    }
    __syncthreads();
    const float wmax = __uint_as_float(wmax_bits);
    // segment lengths are at most sqrt(2) max(1 cm, D): the first step runs from
    // the entry to the integer-cm start plane, the others one pitch
    // (weights so small that the scale would leave float's range, wl below
    // 2^-80, sum in float too, as non-finite ones do)
    const float wl = wmax * (float)(2.0 * fmax(1.0, D));
    if (nonfinite == 0u && !(wmax > 0.f)) return;  // nothing to add (uniform over the workgroup)
    const bool fixed = nonfinite == 0u && wl <= 3.0e38f && wl >= 0x1p-80f;
    int S = 0;
    if (fixed) {
        int e;
        frexpf(wl, &e);  // wl < 2^e, e >= -79
        S = 30 - e;      // <= 109
    }
    const float scale = ldexpf(1.f, S);
    bool ovf = false;
    for (int i = threadIdx.x; i < total; i += blockDim.x) {
        int view, ch;
        item_ray(i, view, ch);
        float w[M];
        adj_weights<M>(a, view * a.W + ch, w);
        bool any = false;
#pragma unroll
        for (int m = 0; m < M; m++) any = any || (w[m] != 0.f);
        if (!any) continue;
        RayConst r;
        if (!setup_ray(a, view, ch, r)) continue;
        int k0, k1;
        if (!box_steps(a, r, x_lo, x_hi, y_lo, y_hi, k0, k1)) continue;
        const bool fast = r.dB != 0.0 && !a.rcheck && !a.exact;
        if (fixed) {
            Fixed sink{acc, x0, y0, r.xmajor, scale, {}};
#pragma unroll
            for (int m = 0; m < M; m++) sink.w[m] = w[m];
            if (fast) march_loop<true>(a, r, sink, ovf, k0, k1);
            else march_loop<false>(a, r, sink, ovf, k0, k1);
        } else {
            Float sink{acc, x0, y0, r.xmajor, 1.f, {}};
#pragma unroll
            for (int m = 0; m < M; m++) sink.w[m] = w[m];
            if (fast) march_loop<true>(a, r, sink, ovf, k0, k1);
            else march_loop<false>(a, r, sink, ovf, k0, k1);
        }
    }
    if (ovf) atomicOr(a.err, 1);
    __syncthreads();
    const float* dens = reinterpret_cast<const float*>(a.dens);
    const double unscale = ldexp(1.0, -S);
    for (int i = threadIdx.x; i < TS * TS; i += blockDim.x) {
        const int lx = i % TS, ly = i / TS;
        const int x = x0 + lx, y = y0 + ly;
        if (x >= a.N || y >= a.N) continue;
        const size_t v = (size_t)y * a.N + x;
        const int off = ly * PITCH + lx;
#pragma unroll
        for (int m = 0; m < M; m++) {
            float val = fixed ? (float)((double)(long long)acc[off + m * PLANE] * unscale)
                              : reinterpret_cast<const float*>(acc)[off + m * PLANE];
            if (dens && !(dens[v * 8 + m] >= 0.f)) val = 0.f;
            if (val != 0.f) unsafeAtomicAdd(a.grad + v * 8 + m, val);
        }
    }
}

// Polyenergetic sum: one spectrum per blockIdx.y; tables in LDS as
// mu2[e][8] = -mu_m(E_e) * log2(e) and pdf[e].
// Rays per workgroup of k_spectral: each workgroup adds its max and min of
// I to the spectrum's two statistics words with one atomic each, and those
// atomics serialise on the one address (1904 workgroups of 256 rays took as
// long as the sums at 800 x 609), so workgroups are large.
constexpr int kSpecBlock = 1024;

template <int M, bool DI>
__global__ __launch_bounds__(kSpecBlock) void k_spectral(SpectralArgs a)
{
    extern __shared__ float lds[];
    float* mu2 = lds;
    float* pdf = lds + a.NE * 8;
    __shared__ uint32_t red[2][kSpecBlock / 64];
    const int s = blockIdx.y;
    const int NE = a.NE;
    const float kLog2e = 1.4426950408889634f;
    for (int i = threadIdx.x; i < NE * 8; i += blockDim.x) {
        const int e = i >> 3, m = i & 7;
        mu2[i] = (m < M) ? -a.att[((size_t)s * M + m) * NE + e] * kLog2e : 0.f;
    }
    for (int e = threadIdx.x; e < NE; e += blockDim.x) pdf[e] = a.pdf[(size_t)s * NE + e];
    __syncthreads();

    const int nr = a.V * a.W;
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = ray < nr;
    float I = 0.f;
    float d[M];
#pragma unroll
    for (int m = 0; m < M; m++) d[m] = 0.f;
    if (valid && ray / a.W < a.VT) {
        const float4* Lv = reinterpret_cast<const float4*>(a.L) + (size_t)ray * 2;
        const float4 l0 = Lv[0], l1 = Lv[1];
        const float Lm[8] = {l0.x, l0.y, l0.z, l0.w, l1.x, l1.y, l1.z, l1.w};
        for (int e = 0; e < NE; e++) {
            // the 8 table entries of bin e in two 16-byte LDS reads
            const float4 m0 = reinterpret_cast<const float4*>(mu2)[2 * e];
            const float4 m1 = reinterpret_cast<const float4*>(mu2)[2 * e + 1];
            const float mu[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            float x = 0.f;
#pragma unroll
            for (int m = 0; m < M; m++) x = fmaf(mu[m], Lm[m], x);
            // hardware exp2: only terms below 2^-126 (flushed) differ from exp2f
            const float t = pdf[e] * __builtin_amdgcn_exp2f(x);
            I += t;
            if (DI) {
#pragma unroll
                for (int m = 0; m < M; m++) d[m] = fmaf(t, mu[m], d[m]);
            }
        }
    }
    if (valid) {
        a.raw[(size_t)s * nr + ray] = I;
        if (DI) {
            const float kLn2 = 0.69314718055994531f;
            float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < M; m++) o[m] = d[m] * kLn2;
            float4* dst = reinterpret_cast<float4*>(a.dIdL) + ((size_t)s * nr + ray) * 2;
            dst[0] = make_float4(o[0], o[1], o[2], o[3]);
            dst[1] = make_float4(o[4], o[5], o[6], o[7]);
        }
    }
    // block max / min of I (non-negative floats order like their bits)
    uint32_t mx = valid ? __float_as_uint(I) : 0u;
    uint32_t mn = valid ? __float_as_uint(I) : 0x7f800000u;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[0][wave] = mx; red[1][wave] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
        for (int w = 1; w < nw; w++) { mx = max(mx, red[0][w]); mn = min(mn, red[1][w]); }
        atomicMax(a.stats + 2 * s, mx);
        atomicMin(a.stats + 2 * s + 1, mn);
    }
}

__global__ void k_stats_init(uint32_t* stats, int S)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S) {
        stats[2 * i] = 0u;
        stats[2 * i + 1] = 0x7f800000u;
    }
}

// Order-preserving uint image of a float (for atomic min/max of any sign).
__device__ __forceinline__ uint32_t ord_of(float f)
{
    const uint32_t b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float float_of_ord(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7fffffffu) : ~o);
}

__global__ void k_minmax_init(uint32_t* st)
{
    st[0] = 0u;           // max, ordered
    st[1] = 0xffffffffu;  // min, ordered
}

__global__ __launch_bounds__(256) void k_minmax(const float* __restrict__ x, int n, uint32_t* st)
{
    __shared__ uint32_t red[2][4];
    uint32_t mx = 0u, mn = 0xffffffffu;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t o = ord_of(x[i]);
        mx = max(mx, o);
        mn = min(mn, o);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
        mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[0][wave] = mx; red[1][wave] = mn; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mx = max(mx, red[0][w]); mn = min(mn, red[1][w]); }
        atomicMax(st, mx);
        atomicMin(st + 1, mn);
    }
}

// Converts the ordered images back to float bit patterns {max, min}.
__global__ void k_minmax_final(uint32_t* st)
{
    st[0] = __float_as_uint(float_of_ord(st[0]));
    st[1] = __float_as_uint(float_of_ord(st[1]));
}

// Philox4x32-10 (Salmon et al., SC'11).
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const uint32_t lo0 = 0xD2511F53u * c.x, hi0 = __umulhi(0xD2511F53u, c.x);
        const uint32_t lo1 = 0xCD9E8D57u * c.z, hi1 = __umulhi(0xCD9E8D57u, c.z);
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

__global__ __launch_bounds__(256) void k_detector(DetectorArgs a)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n * a.S) return;
    const int s = i / a.n;
    const int r = i - s * a.n;
    const float mx = __uint_as_float(a.stats[2 * s]);
    const float p = a.raw[i] / mx;  // M/main.cpp:521
    float o = p;
    if (a.sigma != 0.0) {
        // M/main.cpp:63-74 with Philox uniforms in (0,1) in place of rand()
        const uint4 u = philox4x32_10(make_uint4((uint32_t)r, a.stream_id + (uint32_t)s, 0u, 0u), a.key0, a.key1);
        const double u1 = ((double)u.x + 1.0) / 4294967297.0;
        const double u2 = ((double)u.y + 1.0) / 4294967297.0;
        const double z = sqrt(-2.0 * log(u1)) * sin(2.0 * 3.141592653589793238462643383279 * u2);
        const double nstar = (double)p + (0 + a.sigma * z);
        o = (float)nstar;
        if (nstar < 0) {
            const float pmin = __uint_as_float(a.stats[2 * s + 1]) / mx;
            o = (float)((double)pmin / 10);
        }
        if (nstar > 1.0) o = 1.0f;
    }
    a.out[i] = o;
}

__global__ __launch_bounds__(256) void k_pack_i16(const int16_t* __restrict__ src, int M, int NN,
                                                  uint16_t* __restrict__ dst)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NN) return;
    uint32_t w[4] = {0u, 0u, 0u, 0u};
    for (int m = 0; m < M; m++) {
        const int x = src[(size_t)m * NN + v];
        const uint32_t u = x > 0 ? (uint32_t)x : 0u;
        w[m >> 1] |= (m & 1) ? (u << 16) : u;
    }
    reinterpret_cast<uint4*>(dst)[v] = make_uint4(w[0], w[1], w[2], w[3]);
}

__global__ __launch_bounds__(256) void k_pack_f32(const float* __restrict__ src, int M, int NN,
                                                  float* __restrict__ dst)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= NN) return;
    float r[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int m = 0; m < M; m++) r[m] = src[(size_t)m * NN + v];
    float4* d = reinterpret_cast<float4*>(dst) + (size_t)v * 2;
    d[0] = make_float4(r[0], r[1], r[2], r[3]);
    d[1] = make_float4(r[4], r[5], r[6], r[7]);
}

// One voxel of the tracer's work copies: the first VB bytes of a packed
// voxel (VB = 12: the three words of u16 materials 0-5; 16: all four; 32:
// eight floats).
template <int VB>
struct alignas(VB % 16 == 0 ? 16 : 4) WorkVox {
    uint32_t w[VB / 4];
};

// The tracer's two work copies of the packed densities (IB bytes per voxel),
// padded by two voxels on every side (NP = N + 4 per row): yx[y + 2][x + 2]
// and xy[x + 2][y + 2] hold voxel (x, y) for x, y in [-1, N - 1], index -1
// reading voxel 0 as the reference's truncation does, and zeros elsewhere;
// by 32 x 32 tiles of the padded grid through LDS.
template <int IB, int VB>
__global__ __launch_bounds__(256) void k_pack_copies(const char* __restrict__ in, WorkVox<VB>* __restrict__ yx,
                                                     WorkVox<VB>* __restrict__ xy, int N)
{
    static_assert(VB % 4 == 0 && VB <= IB, "a work voxel is a prefix of the packed voxel");
    __shared__ WorkVox<VB> tile[32][33];
    const int NP = N + 4;
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;  // padded coordinates
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    for (int j = ty; j < 32; j += 8) {
        const int pr = r0 + j, pc = c0 + tx;
        const int y = pr - 2, x = pc - 2;
        WorkVox<VB> t;
#pragma unroll
        for (int q = 0; q < VB / 4; q++) t.w[q] = 0u;
        if (y >= -1 && y < N && x >= -1 && x < N)
            t = *reinterpret_cast<const WorkVox<VB>*>(in + ((size_t)max(y, 0) * N + max(x, 0)) * IB);
        tile[j][tx] = t;
        if (pr < NP && pc < NP) yx[(size_t)pr * NP + pc] = t;
    }
    __syncthreads();
    for (int i = ty; i < 32; i += 8) {
        const int pr = c0 + i, pc = r0 + tx;  // row x + 2, column y + 2
        if (pr < NP && pc < NP) xy[(size_t)pr * NP + pc] = tile[tx][i];
    }
}

template <int IB, int VB>
static void launch_pack_copies(const TraceArgs& a, hipStream_t s)
{
    const int NP = a.N + 4;
    const dim3 grid((unsigned)((NP + 31) / 32), (unsigned)((NP + 31) / 32));
    WorkVox<VB>* yx = reinterpret_cast<WorkVox<VB>*>(a.dens_t);
    hipLaunchKernelGGL((k_pack_copies<IB, VB>), grid, dim3(256), 0, s, reinterpret_cast<const char*>(a.dens), yx,
                       yx + (size_t)NP * NP, a.N);
}

// --------------------------------------------------------------- launchers

// Lanes per orbit: eight, so that the launch holds eight times as many waves
// as orbits/64 (121,800 orbits at 800 x 609 give ~15,200 waves); measured at
// 512^2 (u16, 64-lane workgroups): 2 lanes 0.621 ms, 4 0.538, 8 0.531, 16 0.600.
constexpr int kOrbitLanes = 8;

template <int M, int DT, int R>
static void launch_trace_rot(const TraceArgs& a, hipStream_t s)
{
    const long long lanes = (long long)(a.V / R) * a.W * kOrbitLanes;
    hipLaunchKernelGGL((k_trace_fwd_rot<M, DT, R, kOrbitLanes>), dim3((unsigned)((lanes + a.block - 1) / a.block)),
                       dim3(a.block), 0, s, a);
}

template <int M>
static vct_status launch_trace_m(const TraceArgs& a, int dens_type, bool adjoint, hipStream_t s)
{
    const dim3 block(a.block);
    const dim3 grid((a.n_rays + a.block - 1) / a.block);
    const bool whole = a.ray_offset == 0 && a.n_rays == a.V * a.W;
    if (!adjoint && whole && (a.rot == 4 || a.rot == 2)) {
        if (a.V % a.rot != 0 || a.N % 2 != 0 || a.VT != a.V || a.exact || a.rcheck || !a.dens_t)
            return set_error(VCT_ERR_INTERNAL, "orbit trace with V %d, N %d", a.V, a.N);
        // the orbit gathers read two padded work copies, [y][x] and [x][y]
        if (dens_type == VCT_DENS_U16) {
            if constexpr (M <= 6) launch_pack_copies<16, 12>(a, s);
            else launch_pack_copies<16, 16>(a, s);
        } else {
            launch_pack_copies<32, 32>(a, s);
        }
        if (dens_type == VCT_DENS_U16) {
            if (a.rot == 4) launch_trace_rot<M, VCT_DENS_U16, 4>(a, s);
            else launch_trace_rot<M, VCT_DENS_U16, 2>(a, s);
        } else {
            if (a.rot == 4) launch_trace_rot<M, VCT_DENS_F32, 4>(a, s);
            else launch_trace_rot<M, VCT_DENS_F32, 2>(a, s);
        }
    } else if (adjoint) {
        if (a.VT > 0) {
            // about 2048 workgroups: enough view chunks to fill the chip twice over
            const int ntx = (a.N + adj_tile<M>() - 1) / adj_tile<M>();
            const int ntiles = ntx * ntx;
            int nchunk = std::max(1, std::min(a.VT, (2048 + ntiles - 1) / ntiles));
            int cv = std::min((a.VT + nchunk - 1) / nchunk, kAdjMaxViews);
            if (a.adj_views > 0) cv = std::min(a.adj_views, kAdjMaxViews);
            nchunk = (a.VT + cv - 1) / cv;
            const int per_xcd = (ntiles * nchunk + 7) / 8;
            hipLaunchKernelGGL((k_trace_adj_tile<M>), dim3(per_xcd * 8), dim3(256), 0, s, a, ntx, ntiles, cv,
                               per_xcd);
        }
    } else if (dens_type == VCT_DENS_U16) {
        hipLaunchKernelGGL((k_trace_fwd<M, VCT_DENS_U16, true>), grid, block, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_trace_fwd<M, VCT_DENS_F32, true>), grid, block, 0, s, a);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_trace(const TraceArgs& a, int dens_type, int M, bool adjoint, hipStream_t s)
{
    if (a.n_rays <= 0) return VCT_OK;
    switch (M) {
    case 1: return launch_trace_m<1>(a, dens_type, adjoint, s);
    case 2: return launch_trace_m<2>(a, dens_type, adjoint, s);
    case 3: return launch_trace_m<3>(a, dens_type, adjoint, s);
    case 4: return launch_trace_m<4>(a, dens_type, adjoint, s);
    case 5: return launch_trace_m<5>(a, dens_type, adjoint, s);
    case 6: return launch_trace_m<6>(a, dens_type, adjoint, s);
    case 7: return launch_trace_m<7>(a, dens_type, adjoint, s);
    case 8: return launch_trace_m<8>(a, dens_type, adjoint, s);
    default: return set_error(VCT_ERR_ARG, "number of materials %d outside 1..8", M);
    }
}

template <int M>
static void launch_spectral_m(const SpectralArgs& a, hipStream_t s)
{
    const int nr = a.V * a.W;
    const dim3 grid((nr + kSpecBlock - 1) / kSpecBlock, a.S);
    const size_t lds = sizeof(float) * (size_t)a.NE * 9;
    if (a.dIdL)
        hipLaunchKernelGGL((k_spectral<M, true>), grid, dim3(kSpecBlock), lds, s, a);
    else
        hipLaunchKernelGGL((k_spectral<M, false>), grid, dim3(kSpecBlock), lds, s, a);
}

vct_status launch_spectral(const SpectralArgs& a, hipStream_t s)
{
    if (a.NE <= 0 || a.NE > kMaxBins) return set_error(VCT_ERR_ARG, "NE=%d outside 1..%d", a.NE, kMaxBins);
    if (a.S <= 0 || a.S > 65535) return set_error(VCT_ERR_ARG, "n_spectra=%d invalid", a.S);
    switch (a.M) {
    case 1: launch_spectral_m<1>(a, s); break;
    case 2: launch_spectral_m<2>(a, s); break;
    case 3: launch_spectral_m<3>(a, s); break;
    case 4: launch_spectral_m<4>(a, s); break;
    case 5: launch_spectral_m<5>(a, s); break;
    case 6: launch_spectral_m<6>(a, s); break;
    case 7: launch_spectral_m<7>(a, s); break;
    case 8: launch_spectral_m<8>(a, s); break;
    default: return set_error(VCT_ERR_ARG, "number of materials %d outside 1..8", a.M);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_stats_init(uint32_t* stats, int S, hipStream_t s)
{
    hipLaunchKernelGGL(k_stats_init, dim3((S + 63) / 64), dim3(64), 0, s, stats, S);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_minmax(const float* x, int n, uint32_t* stats, hipStream_t s)
{
    if (n <= 0) return set_error(VCT_ERR_ARG, "empty array");
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, s, stats);
    const int blocks = min((n + 255) / 256, 2048);
    hipLaunchKernelGGL(k_minmax, dim3(blocks), dim3(256), 0, s, x, n, stats);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(1), 0, s, stats);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_detector(const DetectorArgs& a, hipStream_t s)
{
    const long long total = (long long)a.n * a.S;
    if (total <= 0 || total > INT_MAX) return set_error(VCT_ERR_ARG, "detector size %lld invalid", total);
    hipLaunchKernelGGL(k_detector, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, a);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_pack_i16(const int16_t* src, int M, int N, uint16_t* dst, hipStream_t s)
{
    const int NN = N * N;
    hipLaunchKernelGGL(k_pack_i16, dim3((NN + 255) / 256), dim3(256), 0, s, src, M, NN, dst);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_pack_f32(const float* src, int M, int N, float* dst, hipStream_t s)
{
    const int NN = N * N;
    hipLaunchKernelGGL(k_pack_f32, dim3((NN + 255) / 256), dim3(256), 0, s, src, M, NN, dst);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // namespace vct
