// vct_fbp.hip -- fan-beam FBP for CDNA4 (gfx950).
//
//   k_filter_prep  q[r][k] = -ln p[r][k] * sid * cos((k - W/2) * dbeta), the
//                  reference's per-channel pre-weight (F/filterSinogramTOMO.cpp:236-242),
//                  written straight into the register layout of the MFMA A
//                  operand (zero padded to Kp).
//   k_filter_gemm  fltd = q . C'^T with v_mfma_f32_32x32x2f32, where
//                  C'[m][k] = C[m][k] * dbeta is the reference's filter
//                  matrix (prepareConstMap, F/filterSinogramTOMO.cpp:9-138;
//                  Shepp-Logan or Ram-Lak).  Its rows are Toeplitz except the
//                  41 edge rows that prepareConstMap renormalises, so most
//                  column tiles read C' from an LDS copy of its Toeplitz
//                  generator and only edge tiles read the packed matrix.  A
//                  workgroup owns a 32x32 output tile, one wave per quarter
//                  of K; the quarters meet in LDS.
//   k_backproject  pixel-driven fan backprojection
//                  (F/reconstructImageTOMO.cpp:46-100): a 16x16 pixel tile per
//                  workgroup, one lane per pixel, the view loop uniform over
//                  the wave, FP64 geometry and the reference's float
//                  accumulation order.
//   k_backproject_multi  the same for 2..16 sinograms of one geometry per
//                  lane, sharing each pixel-view's geometry.
//   k_backproject_sym  one sinogram by orbits of the quarter turn: each
//                  pixel-view geometry serves four pixels (views a multiple
//                  of 4, N even; sums in another order).
#include "vct_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace vct {

typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int kBpTile = 16;   // pixel tile edge
constexpr int kBpSinos = 16;  // most sinograms per lane in k_backproject_multi
constexpr long long kBpMinLanes = 1 << 18;  // lanes a backprojection launch keeps (4 waves per SIMD)

__device__ __forceinline__ int trunc_i32_bp(double v)
{
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

// Packed operand index: tile (32 rows or columns), K position k, row/col in
// tile rr.  Lane l = rr + 32*(k&1) holds step t = k>>1; four consecutive
// steps share one float4.
__host__ __device__ __forceinline__ size_t pack_index(int tile, int k, int rr, int Kp)
{
    const int t = k >> 1, hi = k & 1;
    const int q = t >> 2, j = t & 3;
    const int l = rr + 32 * hi;
    return ((((size_t)tile * (Kp / 8) + q) * 64 + l) * 4) + j;
}

// Single-precision log rounded from the double log: the correctly rounded
// logf up to double rounding, which is what the C library's logf returns on
// all but ~0.06% of sinogram values (the device logf differs on about a third).
__device__ __forceinline__ float logf_rounded(float x) { return (float)log((double)x); }

// One workgroup per 32 rows x 32 channels: the rows are read along the
// channels (coalesced), pre-weighted into LDS, and written out as the 4 x 64
// float4 of their packed operand block (contiguous), so the writes are whole
// lines instead of 4-byte scatters.
__global__ __launch_bounds__(256) void k_filter_prep(FilterPrepArgs a)
{
    __shared__ float t[32][33];
    const int kc = blockIdx.x, tr = blockIdx.y;
    for (int e = threadIdx.x; e < 32 * 32; e += blockDim.x) {
        const int rr = e >> 5, kk = e & 31;
        const int r = tr * 32 + rr, k = kc * 32 + kk;
        float v = 0.f;
        if (r < a.R && k < a.W) {
            const double p = a.proj[(size_t)r * a.W + k];
            double y;
            if (a.bhc) {
                // beam-hardening correction, M/beam_hardening_correction.cpp:51-55 (whose
                // log of a float is the single-precision overload)
                y = -(double)logf_rounded((float)p);
                y = fabs(y) + (a.bhc_alpha * pow(fabs(y), a.bhc_beta));
            } else {
                y = -log(p);
            }
            v = (float)(y * a.coeff[k]);
        }
        t[rr][kk] = v;
    }
    __syncthreads();
    // float4 (q, l) of the tile holds k = 8q + 2j + (l >> 5), j = 0..3, of
    // row l & 31 (pack_index)
    const int i = threadIdx.x, qq = i >> 6, l = i & 63, rr = l & 31, hi = l >> 5;
    const float* row = t[rr] + 8 * qq + hi;
    float4* out = reinterpret_cast<float4*>(a.apk) + ((size_t)tr * (a.Kp / 8) + kc * 4 + qq) * 64 + l;
    *out = make_float4(row[0], row[2], row[4], row[6]);
}

constexpr int kFiltWaves = 4;  // K quarters of one 32x32 output tile, one wave each
constexpr int kFiltU = 10;     // quads (4 MFMA steps each) a lane loads one chunk ahead

// One 32x32 output tile per workgroup; wave w sums the w-th quarter of K, and
// the quarters meet in LDS in a fixed order.  Operands are the pre-packed
// register order: lane l = i + 32*h holds row/column i at k = 8q + 2j + h in
// element j of float4 q.  Every lane loads a whole chunk of kFiltU float4 of A
// and B before the chunk's MFMAs run, so 20-40 loads are in flight per lane.
//
// B is the filter matrix.  For a column tile whose rows are not renormalised
// (edge_lo < m < edge_hi) it is Toeplitz, B[k][m] = toep[m - k + Kp - 1]: the
// workgroup copies the tile's window of the generator (Kp + 31 floats) into
// LDS and reads B from there, so only the few edge tiles read the packed
// matrix from memory.
//
// Tiles are dealt to XCDs in contiguous row-major ranges (workgroup g runs on
// XCD g % 8), so an XCD reads about an eighth of A, once, into its L2.
__global__ __launch_bounds__(256) void k_filter_gemm(FilterGemmArgs a, int ct, int ntiles, int per_xcd)
{
    extern __shared__ float tw[];  // Toeplitz window, Kp + 31 floats
    __shared__ float part[kFiltWaves - 1][16][64];
    const int xcd = blockIdx.x % 8, slot = blockIdx.x / 8;
    const int tile = xcd * per_xcd + slot;
    if (slot >= per_xcd || tile >= ntiles) return;  // the whole workgroup
    const int tr = tile / ct, tc = tile - tr * ct;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int Q = a.Kp / 8, Qw = Q / kFiltWaves;
    const int qb = wave * Qw, qe = qb + Qw;
    const bool toep = tc * 32 > a.edge_lo && tc * 32 + 31 < a.edge_hi;  // uniform over the workgroup
    if (toep) {
        for (int j = threadIdx.x; j < a.Kp + 31; j += blockDim.x) tw[j] = a.toep[32 * tc + j];
        __syncthreads();
    }
    const float4* A = reinterpret_cast<const float4*>(a.apk) + ((size_t)tr * Q) * 64 + lane;
    const float4* B = reinterpret_cast<const float4*>(a.bpk) + ((size_t)tc * Q) * 64 + lane;
    // window index of B[k = 0] for this lane: column m = 32 tc + (lane & 31), k = lane >> 5
    const int t0 = (lane & 31) - (lane >> 5) + a.Kp - 1;
    floatx16 acc;
#pragma unroll
    for (int i = 0; i < 16; i++) acc[i] = 0.f;
    float4 ca[kFiltU], cb[kFiltU];
#pragma unroll
    for (int u = 0; u < kFiltU; u++) {
        const int q = qb + u;
        if (q < qe) {
            ca[u] = A[(size_t)q * 64];
            if (toep) {
                const int t = t0 - 8 * q;
                cb[u] = make_float4(tw[t], tw[t - 2], tw[t - 4], tw[t - 6]);
            } else {
                cb[u] = B[(size_t)q * 64];
            }
        }
    }
    for (int qs = qb; qs < qe; qs += kFiltU) {
        float4 na[kFiltU], nb[kFiltU];
#pragma unroll
        for (int u = 0; u < kFiltU; u++) {
            const int q = qs + kFiltU + u;
            if (q < qe) {
                na[u] = A[(size_t)q * 64];
                if (toep) {
                    const int t = t0 - 8 * q;
                    nb[u] = make_float4(tw[t], tw[t - 2], tw[t - 4], tw[t - 6]);
                } else {
                    nb[u] = B[(size_t)q * 64];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kFiltU; u++) {
            if (qs + u < qe) {
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ca[u].x, cb[u].x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ca[u].y, cb[u].y, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ca[u].z, cb[u].z, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f32_32x32x2f32(ca[u].w, cb[u].w, acc, 0, 0, 0);
            }
        }
#pragma unroll
        for (int u = 0; u < kFiltU; u++) {
            ca[u] = na[u];
            cb[u] = nb[u];
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int i = 0; i < 16; i++) part[wave - 1][i][lane] = acc[i];
    }
    __syncthreads();
    if (wave == 0) {
        const int j = lane & 31;
        const int col = tc * 32 + j;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int row_in = 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
            const int row = tr * 32 + row_in;
            float v = acc[i] + part[0][i][lane];
#pragma unroll
            for (int w = 1; w < kFiltWaves - 1; w++) v += part[w][i][lane];
            if (a.in_mode && (v < 1000.0f || v > 300000.0f)) v = 0.f;  // :250-255
            if (row < a.R && col < a.W) a.out[(size_t)row * a.W + col] = v;
        }
    }
}

// atan(t) for |t| <= 0.5 as t * P(t^2), P of degree 12: the interpolant of
// atan(sqrt(u))/sqrt(u) at 13 Chebyshev nodes of [0, 0.25], computed in 50-digit
// arithmetic (relative error 4e-18 before rounding; <= 1.2 ulp evaluated in
// FP64 over |t| <= 0.5).  Masked pixels see |t| <= tan(asin(r_mask / sid))
// (0.32 for the reference geometry); larger |t| take the library atan.
__device__ __forceinline__ double atan_small(double t)
{
    const double u = t * t;
    double p = 0.010030841489968486;
    p = __builtin_fma(p, u, -0.027302779603074177);
    p = __builtin_fma(p, u, 0.041775850326747944);
    p = __builtin_fma(p, u, -0.05118531938647954);
    p = __builtin_fma(p, u, 0.05857435765165694);
    p = __builtin_fma(p, u, -0.06663667844183857);
    p = __builtin_fma(p, u, 0.07692057989976618);
    p = __builtin_fma(p, u, -0.09090895058092312);
    p = __builtin_fma(p, u, 0.11111110602305971);
    p = __builtin_fma(p, u, -0.14285714274692063);
    p = __builtin_fma(p, u, 0.19999999999875542);
    p = __builtin_fma(p, u, -0.3333333333333278);
    p = __builtin_fma(p, u, 1.0);
    return t * p;
}

// 1/x to ~1 ulp: hardware estimate + two Newton steps.
__device__ __forceinline__ double rcp_f64(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = __builtin_fma(-x, r, 1.0);
    r = __builtin_fma(r, e, r);
    e = __builtin_fma(-x, r, 1.0);
    return __builtin_fma(r, e, r);
}

// atan(t) for |t| <= 0.5 by a degree-8 polynomial in t^2 (Chebyshev
// interpolant of atan(t)/t on [0, 0.25]): error below 2.5e-13, a 2e-10
// channel shift of the detector coordinate at 0.326 rad over 304 channels.
__device__ __forceinline__ double atan_sym(double t)
{
    const double u = t * t;
    double p = 0.023298796324010433;
    p = __builtin_fma(p, u, -0.05296597132474749);
    p = __builtin_fma(p, u, 0.07378028372611917);
    p = __builtin_fma(p, u, -0.09047114852983332);
    p = __builtin_fma(p, u, 0.1110744933786942);
    p = __builtin_fma(p, u, -0.1428554012478447);
    p = __builtin_fma(p, u, 0.19999995815529828);
    p = __builtin_fma(p, u, -0.3333333329437766);
    p = __builtin_fma(p, u, 0.9999999999993997);
    return t * p;
}

// One lane per pixel, a 16x16 pixel tile per workgroup.  The view loop is the
// same on every lane of a wave, so each view's cos/sin and the base of its
// filtered-sinogram row are scalar loads, and four views are unrolled so that
// their geometry and gathers overlap.  The float accumulation adds the views
// in the reference's order (F/reconstructImageTOMO.cpp:57-95); the filtered
// sinogram (1.95 MB at 800 x 609) is read through L2 without staging.
//
// WIDE: some masked pixel may see |tan(gamma)| > 0.5, beyond the polynomial;
// those take the library atan.  The host picks the variant from the bound
// |tan(gamma)| <= r / sqrt(sid^2 - r^2) over the mask radius r; without the
// fallback the view loop has no branch.
template <bool WIDE>
__global__ __launch_bounds__(kBpTile * kBpTile) void k_backproject(BackprojArgs a)
{
    const int tx = threadIdx.x & (kBpTile - 1), ty = threadIdx.x / kBpTile;
    const int x = blockIdx.x * kBpTile + tx;
    const int y = blockIdx.y * kBpTile + ty;
    const int s = blockIdx.z;
    const float* fl = a.fltd + (size_t)s * a.V * a.W;
    const bool in_img = x < a.N && y < a.N;
    // prepareMaskData, F/misc.cpp:20-36
    const double dyc = (double)y - a.mask_r, dxc = (double)x - a.mask_r;
    const bool masked = in_img && (dyc * dyc + dxc * dxc <= a.mask_r2);
    const double xx = -(x - a.center) * a.pix;  // :52-53
    const double yy = -(y - a.center) * a.pix;
    float acc = 0.f;
    if (masked) {
#pragma unroll 4
        for (int v = 0; v < a.use_views; v++) {
            const double c = a.cosT[v], sn = a.sinT[v];
            const double X = xx * c + yy * sn;
            const double Y = -xx * sn + yy * c;
            const double den = a.sid - Y;
            const double L2 = den * den + X * X;
            const double tg = X * rcp_f64(den);
            double gam = atan_small(tg);
            if (WIDE && fabs(tg) > 0.5) gam = atan(X / den);
            const double fGAM = gam * a.inv_dbeta_ch + a.cent;
            const int GAM = trunc_i32_bp(fGAM);
            const double w = fGAM - GAM;
            const bool ok = GAM >= 0 && GAM < a.W - 1;
            const float* row = fl + (size_t)v * a.W + (ok ? GAM : 0);
            // the reference's sum is 1.0 * sum1 + 0.0 * 0 (:92), which only turns
            // -0 into +0; acc is never -0 (it starts at +0, and x + (-x) is +0),
            // and acc + (-0) = acc + (+0) = acc, so the add is left out
            const double sum = (1 - w) * row[0] + w * row[1];
            const float term = (float)(rcp_f64(L2) * sum * a.dbeta_view * a.xfactor);
            acc = ok ? acc + term : acc;
        }
    }
    if (in_img) a.img[((size_t)s * a.N + y) * a.N + x] = acc;
}

// The sinogram-independent part of one pixel-view: detector weights, 1/L^2,
// whether the bin pair is on the detector, and the row offset of the pair
// (col: its offset within the row).
struct BpView {
    double w, w0, rL2;
    float w1s, w0s;  // FW: the two interpolation weights times 1/L^2 and the view step, in float
    size_t off;
    int col;
    bool ok;
};

// FW: the float-weight form (the degree-8 atan, 1/L^2 in float), whose
// terms differ from the exact ones by ~1e-7 relative
template <bool WIDE, bool FW = false>
__device__ __forceinline__ BpView bp_view(const BackprojArgs& a, double xx, double yy, int v)
{
    const double c = a.cosT[v], sn = a.sinT[v];
    const double X = xx * c + yy * sn;
    const double Y = -xx * sn + yy * c;
    const double den = a.sid - Y;
    const double L2 = den * den + X * X;
    // FW: one Newton step on the reciprocal estimate (within ~11 ulp of the
    // quotient, a 1e-12 channel shift); the exact form takes two, which
    // reproduce the IEEE quotient
    double rden;
    if constexpr (FW) {
        const double r = __builtin_amdgcn_rcp(den);
        rden = __builtin_fma(r, __builtin_fma(-den, r, 1.0), r);
    } else {
        rden = rcp_f64(den);
    }
    const double tg = X * rden;
    double gam = FW ? atan_sym(tg) : atan_small(tg);
    if (WIDE && fabs(tg) > 0.5) gam = atan(X / den);
    const double fGAM = gam * a.inv_dbeta_ch + a.cent;
    const int GAM = trunc_i32_bp(fGAM);
    BpView g;
    g.w = fGAM - GAM;
    g.w0 = 1 - g.w;
    if constexpr (FW) {
        const float sf = __builtin_amdgcn_rcpf((float)L2) * (float)(a.dbeta_view * a.xfactor);
        g.w1s = (float)g.w * sf;
        g.w0s = (float)g.w0 * sf;
    } else {
        g.rL2 = rcp_f64(L2);
    }
    g.ok = GAM >= 0 && GAM < a.W - 1;
    g.col = g.ok ? GAM : 0;
    g.off = (size_t)v * a.W + g.col;
    return g;
}

// k_backproject for K >= 2 sinograms of the same geometry per lane.  The pixel-view geometry (X,
// Y, L^2, gamma, the detector bin and its weights) does not depend on the
// sinogram, so a lane computes it once and applies it to K sinograms, each
// with the single-sinogram expression; the images are bit-identical for
// every K.  Workgroup z covers sinograms K*z ... K*z + K - 1; past S the
// lane re-reads sinogram S - 1 and stores nothing.  bp_multi_tile is one
// tile (bx, by) of workgroup z = bz.  FW (float weights, the default unless
// the plan asks for the exact backprojection): the interpolation and
// distance weights of a term are computed in the float-weight form of
// bp_view and applied in float, as in k_backproject_sym (within ~1e-6 of
// the exact terms); otherwise every term is the single-sinogram FP64
// expression and the images equal k_backproject's bit for bit.
// IL (K a multiple of 4): a.fltd holds the sinograms interleaved in slices
// of K, [S/K][V][W][K] (k_interleave_sinos), so the K tap pairs of a
// pixel-view are 2K contiguous floats, read as K/2 16-byte loads instead of
// 2K 4-byte ones; the same values, so the same images.
template <bool WIDE, int K, bool FW, bool IL = false>
__device__ __forceinline__ void bp_multi_tile(const BackprojArgs& a, int S, int bx, int by, int bz)
{
    static_assert(!IL || K % 4 == 0, "interleaved slices hold a multiple of 4 sinograms");
    const double dbx = a.dbeta_view * a.xfactor;
    const int tx = threadIdx.x & (kBpTile - 1), ty = threadIdx.x / kBpTile;
    const int x = bx * kBpTile + tx;
    const int y = by * kBpTile + ty;
    const int s0 = bz * K;
    const float* fl[K];
#pragma unroll
    for (int k = 0; k < K; k++) fl[k] = a.fltd + (size_t)min(s0 + k, S - 1) * a.V * a.W;
    // IL: the tap pairs of bin pair off are floats [off * K, off * K + 2K) of
    // slice bz; off <= V * W - 2, so they stay inside the slice
    const float* slice = a.fltd + (size_t)bz * a.V * a.W * K;
    auto taps = [&](size_t off, float (&t0)[K], float (&t1)[K]) {
        if constexpr (IL) {
            const float4* q = reinterpret_cast<const float4*>(slice + off * K);
#pragma unroll
            for (int h = 0; h < K / 4; h++) {
                const float4 u = q[h], w = q[K / 4 + h];
                t0[4 * h] = u.x; t0[4 * h + 1] = u.y; t0[4 * h + 2] = u.z; t0[4 * h + 3] = u.w;
                t1[4 * h] = w.x; t1[4 * h + 1] = w.y; t1[4 * h + 2] = w.z; t1[4 * h + 3] = w.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < K; k++) {
                t0[k] = fl[k][off];
                t1[k] = fl[k][off + 1];
            }
        }
    };
    const bool in_img = x < a.N && y < a.N;
    // prepareMaskData, F/misc.cpp:20-36
    const double dyc = (double)y - a.mask_r, dxc = (double)x - a.mask_r;
    const bool masked = in_img && (dyc * dyc + dxc * dxc <= a.mask_r2);
    const double xx = -(x - a.center) * a.pix;  // :52-53
    const double yy = -(y - a.center) * a.pix;
    float acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.f;
    if (masked && a.use_views > 0) {
        // software pipeline: the geometry and gathers of view v + 1 are
        // issued before the terms of view v are summed (the last view loads
        // itself again), so the gathers' latency overlaps the arithmetic
        auto add = [&](const BpView& g, const float (&t0)[K], const float (&t1)[K]) {
#pragma unroll
            for (int k = 0; k < K; k++) {
                float term;
                if constexpr (FW) {
                    term = fmaf(g.w1s, t1[k], g.w0s * t0[k]);
                } else {
                    const double sum = g.w0 * t0[k] + g.w * t1[k];  // + 0.0 left out as above
                    term = (float)(g.rL2 * sum * a.dbeta_view * a.xfactor);
                }
                acc[k] = g.ok ? acc[k] + term : acc[k];
            }
        };
        // two views per iteration in two register sets: view v + 1 is
        // fetched while view v is summed, then view v + 2 while view v + 1
        // is (no copies between the sets); an odd count's last fetch
        // re-reads the last view and is not summed
        const int nv = a.use_views;
        BpView g = bp_view<WIDE, FW>(a, xx, yy, 0), gn;
        float r0[K], r1[K], n0[K], n1[K];
        taps(g.off, r0, r1);
#pragma unroll 1
        for (int v = 0; v < nv; v += 2) {
            gn = bp_view<WIDE, FW>(a, xx, yy, min(v + 1, nv - 1));
            taps(gn.off, n0, n1);
            add(g, r0, r1);
            if (v + 1 == nv) break;
            g = bp_view<WIDE, FW>(a, xx, yy, min(v + 2, nv - 1));
            taps(g.off, r0, r1);
            add(gn, n0, n1);
        }
    }
    if (in_img) {
#pragma unroll
        for (int k = 0; k < K; k++)
            if (s0 + k < S) a.img[((size_t)(s0 + k) * a.N + y) * a.N + x] = acc[k];
    }
}

template <bool WIDE, int K, bool FW>
__global__ __launch_bounds__(kBpTile * kBpTile) void k_backproject_multi(BackprojArgs a, int S)
{
    bp_multi_tile<WIDE, K, FW>(a, S, blockIdx.x, blockIdx.y, blockIdx.z);
}

// Several geometries in one launch: workgroup l belongs to the group g with
// pref[g] <= l < pref[g + 1] (a scalar search, uniform over the workgroup)
// and runs tile l - pref[g] of that group's tiles x slices, exactly as
// k_backproject_multi would.
template <bool WIDE, int K, bool FW, bool IL = false>
__global__ __launch_bounds__(kBpTile * kBpTile) void k_backproject_groups(BpGroups ga)
{
    const int l = blockIdx.x;
    int g = 0;
    while (g + 1 < ga.n && l >= ga.pref[g + 1]) g++;
    const int local = l - ga.pref[g];
    const int tx = ga.tiles_x[g];
    const int per_slice = tx * tx;
    const int bz = local / per_slice, r = local - bz * per_slice;
    bp_multi_tile<WIDE, K, FW, IL>(ga.a[g], ga.S[g], r % tx, r / tx, bz);
}

// Sinograms [S][VW] -> slices of K, dst[S/K][VW][K] (zeros past S): lane p
// of slice sl gathers the K values of pixel-row entry p (each read coalesced
// over the lanes) and stores them as K/4 16-byte stores.
template <int K>
__global__ __launch_bounds__(256) void k_interleave_sinos(const float* __restrict__ src, int S, int VW,
                                                          float* __restrict__ dst)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x, sl = blockIdx.y;
    if (p >= VW) return;
    float v[K];
#pragma unroll
    for (int k = 0; k < K; k++) {
        const int s = sl * K + k;
        v[k] = s < S ? src[(size_t)s * VW + p] : 0.f;
    }
    float4* d = reinterpret_cast<float4*>(dst + ((size_t)sl * VW + p) * K);
#pragma unroll
    for (int h = 0; h < K / 4; h++) d[h] = make_float4(v[4 * h], v[4 * h + 1], v[4 * h + 2], v[4 * h + 3]);
}

// One sinogram by orbits of the quarter turn.  With V views over 360 degrees
// (V a multiple of 4, Q = V/4) and a centred N x N grid (N even), pixel p at
// view v + Q has the geometry of pixel T(p) = (y, N-1-x) at view v, so the
// geometry G(p, v) of a base pixel p (x, y < N/2) serves the four pixels
// T^-k(p), k = 0..3, at views v + kQ (mod V):
//   T^0 = (x, y), T^-1 = (N-1-y, x), T^-2 = (N-1-x, N-1-y), T^-3 = (y, N-1-x).
// A workgroup holds 64 base pixels (an 8 x 8 tile of the quadrant); wave w
// runs views v in [wQ, (w+1)Q) with the software pipeline of
// k_backproject_multi, so each pixel's 4 partial sums (wave 0 .. 3) are added
// in that order through LDS.  The geometry follows the single-sinogram
// expression in FP64; the interpolation and distance weights of a term are
// applied in float, and the order of the float sum differs from the
// reference's ascending views (images within 1e-5 of the reference's).
template <bool WIDE>
__global__ __launch_bounds__(256) void k_backproject_sym(BackprojArgs a)
{
    __shared__ float part[3][4][64];
    // the view range is per wave: made scalar, so the four rows of a view
    // are scalar bases and their wrap a scalar test
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int lane = threadIdx.x & 63;
    const int h = a.N / 2, Q = a.V / 4;
    const int x = blockIdx.x * 8 + (lane & 7), y = blockIdx.y * 8 + (lane >> 3);
    const bool in_q = x < h && y < h;
    // prepareMaskData, F/misc.cpp:20-36 (the circle is invariant under T)
    const double dyc = (double)y - a.mask_r, dxc = (double)x - a.mask_r;
    const bool masked = in_q && (dyc * dyc + dxc * dxc <= a.mask_r2);
    const double xx = -(x - a.center) * a.pix;  // :52-53
    const double yy = -(y - a.center) * a.pix;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    const int vb = wave * Q, ve = vb + Q;
    if (masked && Q > 0) {
        // rows of view v + kQ (mod V); vb + kQ < V + 3Q, one wrap at most
        auto row_of = [&](int v, int k) {
            const int r = v + k * Q;
            return a.fltd + (size_t)(r >= a.V ? r - a.V : r) * a.W;
        };
        // the pixel-view geometry in the float-weight form: the sums already
        // run in another order than the reference's, and its roundings (~1e-7
        // relative per term) stay far below that difference
        auto fetch = [&](int v, const BpView& g, float (&t0)[4], float (&t1)[4]) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float* row = row_of(v, k) + g.col;
                t0[k] = row[0];
                t1[k] = row[1];
            }
        };
        auto add = [&](const BpView& g, const float (&t0)[4], const float (&t1)[4]) {
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const float term = fmaf(g.w1s, t1[k], g.w0s * t0[k]);
                acc[k] = g.ok ? acc[k] + term : acc[k];
            }
        };
        // two views per iteration in two register sets (as in bp_multi_tile)
        BpView g = bp_view<WIDE, true>(a, xx, yy, vb), gn;
        float r0[4], r1[4], n0[4], n1[4];
        fetch(vb, g, r0, r1);
#pragma unroll 1
        for (int v = vb; v < ve; v += 2) {
            const int v1 = min(v + 1, ve - 1);
            gn = bp_view<WIDE, true>(a, xx, yy, v1);
            fetch(v1, gn, n0, n1);
            add(g, r0, r1);
            if (v + 1 == ve) break;
            const int v2 = min(v + 2, ve - 1);
            g = bp_view<WIDE, true>(a, xx, yy, v2);
            fetch(v2, g, r0, r1);
            add(gn, n0, n1);
        }
    }
    if (wave > 0) {
#pragma unroll
        for (int k = 0; k < 4; k++) part[wave - 1][k][lane] = acc[k];
    }
    __syncthreads();
    if (wave == 0 && in_q) {
        const int n1 = a.N - 1;
        const int px[4] = {x, n1 - y, n1 - x, y};
        const int py[4] = {y, x, n1 - y, n1 - x};
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const float v = ((acc[k] + part[0][k][lane]) + part[1][k][lane]) + part[2][k][lane];
            a.img[(size_t)py[k] * a.N + px[k]] = v;
        }
    }
}

// The beam-hardening tool M/beam_hardening_correction.cpp:47-60 elementwise:
// y = -ln p (single-precision log, as the tool's C++ overload resolution
// gives), y' = |y| + alpha |y|^beta, out = exp(-y') in FP64, stored as float.
__global__ __launch_bounds__(256) void k_beam_hardening(const float* __restrict__ in, float* __restrict__ out, int n,
                                                        double alpha, double beta)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double yi = -(double)logf_rounded(in[i]);  // the tool's log of a float is logf
    yi = fabs(yi) + (alpha * pow(fabs(yi), beta));
    out[i] = (float)exp(-yi);
}

vct_status launch_beam_hardening(const float* in, float* out, int n, double alpha, double beta, hipStream_t s)
{
    if (n <= 0) return set_error(VCT_ERR_ARG, "empty sinogram");
    hipLaunchKernelGGL(k_beam_hardening, dim3((n + 255) / 256), dim3(256), 0, s, in, out, n, alpha, beta);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_filter(const FilterPrepArgs& p, const FilterGemmArgs& g, hipStream_t s)
{
    if (p.Kp % (8 * kFiltWaves) != 0)
        return set_error(VCT_ERR_INTERNAL, "filter K padding %d not a multiple of %d", p.Kp, 8 * kFiltWaves);
    if ((size_t)(p.Kp + 31) * sizeof(float) > 48 * 1024) return set_error(VCT_ERR_ARG, "too many channels for the filter");
    const int Rp = (p.R + 31) / 32 * 32;
    hipLaunchKernelGGL(k_filter_prep, dim3(p.Kp / 32, Rp / 32), dim3(256), 0, s, p);
    VCT_HIP(hipGetLastError());
    const int Wp = (g.W + 31) / 32 * 32;
    const int ct = Wp / 32;
    const long long ntiles = (long long)(Rp / 32) * ct;
    if (ntiles > INT_MAX / 2) return set_error(VCT_ERR_ARG, "filter GEMM too large");
    const int per_xcd = (int)((ntiles + 7) / 8);
    const size_t lds = (size_t)(g.Kp + 31) * sizeof(float);
    hipLaunchKernelGGL(k_filter_gemm, dim3(8 * per_xcd), dim3(64 * kFiltWaves), lds, s, g, ct, (int)ntiles, per_xcd);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

// k_backproject_multi with K sinograms per lane, for K = K0 ... kBpSinos.
template <bool WIDE, int K0>
static void launch_bp_multi(const BackprojArgs& a, int S, int K, hipStream_t s)
{
    if constexpr (K0 < kBpSinos) {
        if (K > K0) {
            launch_bp_multi<WIDE, K0 + 1>(a, S, K, s);
            return;
        }
    }
    const dim3 grid((a.N + kBpTile - 1) / kBpTile, (a.N + kBpTile - 1) / kBpTile, (S + K0 - 1) / K0);
    if (a.exact)
        hipLaunchKernelGGL((k_backproject_multi<WIDE, K0, false>), grid, dim3(kBpTile * kBpTile), 0, s, a, S);
    else
        hipLaunchKernelGGL((k_backproject_multi<WIDE, K0, true>), grid, dim3(kBpTile * kBpTile), 0, s, a, S);
}

template <bool WIDE, bool FW, int K0>
static void launch_bp_groups_k(const BpGroups& ga, int K, hipStream_t s)
{
    if constexpr (K0 < kBpSinos) {
        if (K > K0) {
            launch_bp_groups_k<WIDE, FW, K0 + 1>(ga, K, s);
            return;
        }
    }
    hipLaunchKernelGGL((k_backproject_groups<WIDE, K0, FW>), dim3(ga.pref[ga.n]), dim3(kBpTile * kBpTile), 0, s, ga);
}

vct_status launch_backproject_groups(BpGroups& ga, int K, hipStream_t s, bool il)
{
    if (ga.n < 1 || ga.n > kMaxBpGroups) return set_error(VCT_ERR_ARG, "%d groups (1..%d)", ga.n, kMaxBpGroups);
    if (K < 2) K = 2;
    if (K > kBpSinos) K = kBpSinos;
    bool wide = false, exact = false;
    long long total = 0;
    ga.pref[0] = 0;
    for (int g = 0; g < ga.n; g++) {
        const BackprojArgs& a = ga.a[g];
        exact = exact || a.exact;
        if (ga.S[g] < 1) return set_error(VCT_ERR_ARG, "group %d has no sinograms", g);
        const double r = a.mask_r * a.pix;
        wide = wide || !(r < a.sid) || r / std::sqrt(a.sid * a.sid - r * r) > 0.49;
        ga.tiles_x[g] = (a.N + kBpTile - 1) / kBpTile;
        total += (long long)ga.tiles_x[g] * ga.tiles_x[g] * ((ga.S[g] + K - 1) / K);
        if (total > INT_MAX) return set_error(VCT_ERR_ARG, "grouped backprojection too large");
        ga.pref[g + 1] = (int)total;
    }
    if (il) {
        const dim3 grid(ga.pref[ga.n]), blk(kBpTile * kBpTile);
#define VCT_BP_IL(KK)                                                                                   \
    case KK:                                                                                            \
        if (wide && exact) hipLaunchKernelGGL((k_backproject_groups<true, KK, false, true>), grid, blk, 0, s, ga);  \
        else if (wide) hipLaunchKernelGGL((k_backproject_groups<true, KK, true, true>), grid, blk, 0, s, ga);  \
        else if (exact) hipLaunchKernelGGL((k_backproject_groups<false, KK, false, true>), grid, blk, 0, s, ga); \
        else hipLaunchKernelGGL((k_backproject_groups<false, KK, true, true>), grid, blk, 0, s, ga);   \
        break;
        switch (K) {
            VCT_BP_IL(4) VCT_BP_IL(8) VCT_BP_IL(12) VCT_BP_IL(16)
        default: return set_error(VCT_ERR_ARG, "interleaved slices of %d sinograms (4, 8, 12 or 16)", K);
        }
#undef VCT_BP_IL
    } else if (wide && exact) launch_bp_groups_k<true, false, 2>(ga, K, s);
    else if (wide) launch_bp_groups_k<true, true, 2>(ga, K, s);
    else if (exact) launch_bp_groups_k<false, false, 2>(ga, K, s);
    else launch_bp_groups_k<false, true, 2>(ga, K, s);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status launch_interleave_sinos(const float* src, int S, int VW, int K, float* dst, hipStream_t s)
{
    if (S < 1 || VW < 1) return set_error(VCT_ERR_ARG, "empty sinograms");
    const dim3 grid((VW + 255) / 256, (S + K - 1) / K), blk(256);
    switch (K) {
    case 4: hipLaunchKernelGGL(k_interleave_sinos<4>, grid, blk, 0, s, src, S, VW, dst); break;
    case 8: hipLaunchKernelGGL(k_interleave_sinos<8>, grid, blk, 0, s, src, S, VW, dst); break;
    case 12: hipLaunchKernelGGL(k_interleave_sinos<12>, grid, blk, 0, s, src, S, VW, dst); break;
    case 16: hipLaunchKernelGGL(k_interleave_sinos<16>, grid, blk, 0, s, src, S, VW, dst); break;
    default: return set_error(VCT_ERR_ARG, "interleaved slices of %d sinograms (4, 8, 12 or 16)", K);
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

template <bool WIDE>
static void launch_bp_w(const BackprojArgs& a, int S, hipStream_t s)
{
    // the fewest workgroup slices of at most kBpSinos sinograms, split evenly
    // (S = 20: two slices of 10, not 16 + 4), but enough slices that the launch
    // has kBpMinLanes lanes: 256^2 pixels x 3 sinograms in one slice would
    // leave one wave per SIMD, and 3 slices of one sinogram run 2-3x faster
    const long long px = (long long)a.N * a.N;
    const long long floor_slices = (kBpMinLanes + px - 1) / px;
    const int cap = a.max_k >= 2 && a.max_k < kBpSinos ? a.max_k : kBpSinos;
    int slices = (S + cap - 1) / cap;
    if (slices < floor_slices) slices = (int)std::min<long long>(S, floor_slices);
    const int K = (S + slices - 1) / slices;
    if (S == 1 && !a.exact && a.V % 4 == 0 && a.N % 2 == 0 && a.use_views == a.V && a.N >= 2) {
        const int h = a.N / 2;
        const dim3 grid((h + 7) / 8, (h + 7) / 8);
        hipLaunchKernelGGL(k_backproject_sym<WIDE>, grid, dim3(256), 0, s, a);
        return;
    }
    if (K == 1 && (a.exact || S == 1)) {
        const dim3 grid((a.N + kBpTile - 1) / kBpTile, (a.N + kBpTile - 1) / kBpTile, S);
        hipLaunchKernelGGL(k_backproject<WIDE>, grid, dim3(kBpTile * kBpTile), 0, s, a);
        return;
    }
    // (several sinograms one per lane take the float-weight template too, so
    // an image does not depend on how the call was split)
    launch_bp_multi<WIDE, 1>(a, S, K, s);
}

vct_status launch_backproject(const BackprojArgs& a, int S, hipStream_t s)
{
    if (S <= 0) return set_error(VCT_ERR_ARG, "no sinograms to backproject");
    // largest |tan(gamma)| of a masked pixel: the tangent from the source to
    // the mask circle of radius r (pixel centres lie within mask_r pitches)
    const double r = a.mask_r * a.pix;
    const bool wide = !(r < a.sid) || r / std::sqrt(a.sid * a.sid - r * r) > 0.49;
    if (wide)
        launch_bp_w<true>(a, S, s);
    else
        launch_bp_w<false>(a, S, s);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // namespace vct
