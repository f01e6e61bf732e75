// vct_batch.hip -- batched execution over cached system matrices.
//
// A system matrix is the exact segment decomposition of one geometry, traced
// once with the same march as the on-the-fly tracer (vct_march.h), stored
//   ray-major   (CSR: row = ray,   entry = (voxel, length))  for the forward,
//   voxel-major (CSC: row = voxel, entry = (ray, length))    for the adjoint.
// The voxel-major copy comes from a stable radix sort of the ray-major
// entries by voxel, so each column lists its rays in increasing order and
// the gather adjoint is deterministic and atomic-free.
//
// Batched data is phantom-interleaved in rows of 8*B floats -- densities one
// row per voxel, path integrals and adjoint weights one row per ray -- each
// row in three phantom-major planes (see load_slot), so the B lanes that
// share a row read each plane of an entry's row as one dense run.  One persistent kernel
// (k_gather) serves both directions over up to kMaxGroups groups of
// phantoms (one group per geometry), pulling work items round-robin.
//
// Blocked two-stage form (VCT_SYSMAT_BLOCKED).  A ray crosses the whole
// image, so the rays in flight on one XCD sweep all N*N*B*32 bytes of the
// densities past its 4 MB L2 once per view; the voxels in flight in the
// adjoint start at different times and so read the rays of every view.
// The blocked form regroups the entries of each direction by a block key,
//   forward: the kBlockTile x kBlockTile image tile of the voxel,
//   adjoint: (view / kAdjViews, tile of the voxel, ranked along the rays),
// keeping row order inside a block (stable radix sort).  Stage 1 gathers
// one row per (block, row) run into a partial-sum workspace, block after
// block, so what a block reads (one tile of densities, or the rays of
// kAdjViews views that cross one tile) stays in the L2 of the XCD that owns
// it; stage 2 sums each row's partials in block order.  Both stages are
// ordinary k_gather launches and the result is deterministic.
//
// LDS-resident stage 1 (built with the blocked form unless
// VCT_SYSMAT_NO_TILE).  The runs of each LDS block, sorted by length into
// slices of 64 with interleaved entries, are summed by k_gather_tile, one
// workgroup per (block, phantom) with what the block reads of that phantom
// in LDS: the vector memory path then carries only the 8-byte entries, read
// once from HBM by the B workgroups of a block on one XCD.
//   forward: a block is a 64x64 tile, its densities in LDS;
//   adjoint: a block is one part of a 64x64 tile (1 to 16 parts, the
//     coarsest split whose tables fit) in one 50-view block, the weights of
//     the rays of those views that cross the part in LDS.
// Same runs, same order of terms: the partial sums are those of the k_gather
// stage 1, bit for bit.
#include <algorithm>
#include <cstring>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include "vct_internal.h"
#include "vct_march.h"

// Two-stage form of one direction (see above).
struct vct_blocked {
    uint32_t rows1 = 0;           // stage-1 rows: (block, row) runs, block-major
    uint32_t* ptr1 = nullptr;     // [rows1+1]
    uint2* ent1 = nullptr;        // [nseg] the direction's entries, regrouped
    uint32_t* ptr2 = nullptr;     // [rows+1]
    uint2* ent2 = nullptr;        // [rows1] (stage-1 row, 1.0f), block order per row
};

// LDS-resident form of a stage 1 (built with the blocked form unless
// VCT_SYSMAT_NO_TILE): the stage-1 runs of each LDS block sorted by length
// (longest first) in slices of 64, whose entries are interleaved so that lane
// l of a wave reads entry k of run l at slice_ptr + 64 k + l.
//   forward: a block is a 64x64 image tile; an entry holds its voxel's index
//     inside the tile (ly * 64 + lx) and the length; padding entries are
//     (kTilePadVoxel, 0.0f), which read zeroed LDS;
//   adjoint: a block is (view block, part of a 64x64 image tile), whose LDS
//     table holds, view after view, the rays of the view block that cross
//     the part (the channel range [clo, clo + n) of each view); an entry
//     holds its ray's slot in the table; slot 0 holds zeros and pads
//     (0, 0.0f).
struct vct_tiled {
    int ntiles = 0, nslices = 0, kmax = 0, ntx = 0;
    uint2* ell = nullptr;            // [sum slice_len * 64 + 64 * kTileUnroll]
    uint32_t* slice_ptr = nullptr;   // [nslices] first entry of the slice
    uint32_t* slice_len = nullptr;   // [nslices] entries per lane (adjoint: a multiple of kTileUnroll)
    int32_t* slice_runs = nullptr;   // [nslices * 64] stage-1 row of each lane, -1: none
    int32_t* tile_slices = nullptr;  // [ntiles + 1]
    int32_t* xcd_tiles = nullptr;    // [8 * kmax] tiles of each XCD, -1: none
    size_t n_ell = 0;
    // adjoint blocks only (slot_off == nullptr: forward)
    int32_t* slot_off = nullptr;     // [ntiles * (kAdjViews + 1)] first slot of each view, then the table size
    int32_t* slot_clo = nullptr;     // [ntiles * kAdjViews] first channel of each view
    int32_t* blk_v0 = nullptr;       // [ntiles] first view of the block's view block
    int slots_max = 0;               // largest table, zero slot included
};

struct vct_sysmat {
    vct_plan* plan = nullptr;
    int64_t nseg = 0;
    int nrays = 0;
    int NN = 0;
    float scale = 1.f;            // 1/xfactor: the gathers return path sums in the tracer's units
    uint32_t* row_ptr = nullptr;  // [nrays+1]
    uint2* csr = nullptr;         // [nseg] (voxel, float bits of length)
    uint32_t* col_ptr = nullptr;  // [NN+1]
    uint2* csc = nullptr;         // [nseg] (ray, float bits of length)
    vct_blocked fwd, adj;         // blocked forms (rows1 == 0: not built)
    vct_tiled tile;               // LDS-resident forward stage 1 (ell == nullptr: not built)
    vct_tiled atile;              // LDS-resident adjoint stage 1 (ell == nullptr: not built)
};

namespace vct {

constexpr int kMaxGroups = 16;

struct CountSink {
    static constexpr bool kAdd2 = false;
    int n = 0;
    __device__ __forceinline__ void add(int, double) { n++; }
};

struct FillSink {
    static constexpr bool kAdd2 = false;
    uint2* out;
    uint32_t* keys;   // voxel of each entry (optional)
    uint32_t* vals;   // entry index (optional)
    uint32_t* eray;   // ray of each entry (optional)
    uint32_t* colcnt; // entries per voxel (optional)
    uint32_t base, ray, k;
    __device__ __forceinline__ void add(int v, double rr)
    {
        const uint32_t e = base + k++;
        out[e] = make_uint2((uint32_t)v, __float_as_uint((float)rr));
        if (eray) eray[e] = ray;
        if (keys) {
            keys[e] = (uint32_t)v;
            vals[e] = e;
            atomicAdd(colcnt + v, 1u);
        }
    }
};

__global__ __launch_bounds__(256) void k_sm_count(TraceArgs a, uint32_t* counts)
{
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= a.n_rays) return;
    const int view = ray / a.W, ch = ray - view * a.W;
    CountSink sink;
    if (view < a.VT) {
        bool ovf = false;
        march_ray<true>(a, view, ch, sink, ovf);
        if (ovf) atomicOr(a.err, 1);
    }
    counts[ray] = (uint32_t)sink.n;
}

__global__ __launch_bounds__(256) void k_sm_fill(TraceArgs a, const uint32_t* row_ptr, uint2* out, uint32_t* keys,
                                                 uint32_t* vals, uint32_t* eray, uint32_t* colcnt)
{
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    if (ray >= a.n_rays) return;
    const int view = ray / a.W, ch = ray - view * a.W;
    if (view >= a.VT) return;
    FillSink sink{out, keys, vals, eray, colcnt, row_ptr[ray], (uint32_t)ray, 0u};
    bool ovf = false;
    march_ray<true>(a, view, ch, sink, ovf);
    if (ovf || sink.k != row_ptr[ray + 1] - row_ptr[ray]) atomicOr(a.err, 2);
}

__global__ __launch_bounds__(256) void k_csc_gather(const uint32_t* __restrict__ order, const uint32_t* __restrict__ eray,
                                                    const uint2* __restrict__ csr, uint2* __restrict__ csc, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = order[i];
    csc[i] = make_uint2(eray[e], csr[e].y);
}

// ------------------------------------------------------ blocked form build

constexpr int kBlockTile = 64;  // image tile edge (pixels) of the forward blocks
constexpr int kAdjTile = 64;    // image tile edge (pixels) of the adjoint blocks
constexpr int kAdjViews = 50;   // views per adjoint block

struct TileGrid {
    int N, T, ntx, ntiles;
    __host__ __device__ __forceinline__ uint32_t tile(uint32_t v) const
    {
        const uint32_t iy = v / (uint32_t)N, ix = v - iy * (uint32_t)N;
        return (iy / (uint32_t)T) * (uint32_t)ntx + ix / (uint32_t)T;
    }
};

static TileGrid tile_grid(int N, int T)
{
    TileGrid t;
    t.N = N;
    t.T = T;
    t.ntx = (N + T - 1) / T;
    t.ntiles = t.ntx * t.ntx;
    return t;
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* p, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = (uint32_t)i;
}

// forward block key: the tile of the entry's voxel
__global__ __launch_bounds__(256) void k_keys_fwd(const uint2* __restrict__ ent, int64_t n, TileGrid tg,
                                                  uint32_t* __restrict__ key)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) key[i] = tg.tile(ent[i].x);
}

// adjoint block key: (view block of the entry's ray, rank of the row's tile).
// Within a view block the tiles are ranked along the block's rays: down the
// tile columns when its middle view's central ray (source (sid sin t,
// sid cos t) through the origin) runs closer to the y axis, along the tile
// rows otherwise, so consecutive blocks share most of their rays and read
// them from L2 (28 -> 18 GB of HBM reads per adjoint launch at 512^2,
// 800 views, 16 phantoms).
__global__ __launch_bounds__(256) void k_keys_adj(const uint2* __restrict__ ent, const uint32_t* __restrict__ vox,
                                                  int64_t n, TileGrid tg, int W, int V, const double* __restrict__ cosT,
                                                  const double* __restrict__ sinT, uint32_t* __restrict__ key)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t vb = ent[i].x / (uint32_t)W / kAdjViews;
    const int mid = min((int)vb * kAdjViews + kAdjViews / 2, V - 1);
    const uint32_t v = vox[i];
    const uint32_t iy = v / (uint32_t)tg.N, ix = v - iy * (uint32_t)tg.N;
    const uint32_t ty = iy / (uint32_t)tg.T, tx = ix / (uint32_t)tg.T;
    const bool along_y = fabs(cosT[mid]) > fabs(sinT[mid]);
    key[i] = vb * (uint32_t)tg.ntiles + (along_y ? tx * (uint32_t)tg.ntx + ty : ty * (uint32_t)tg.ntx + tx);
}

// head[i] = 1 where a new (block, row) run starts in sorted order; head[n] = 0
__global__ __launch_bounds__(256) void k_run_heads(const uint32_t* __restrict__ key_s, const uint32_t* __restrict__ order,
                                                   const uint32_t* __restrict__ row_of, int64_t n,
                                                   uint32_t* __restrict__ head)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) { head[n] = 0u; return; }
    head[i] = (i == 0 || key_s[i] != key_s[i - 1] || row_of[order[i]] != row_of[order[i - 1]]) ? 1u : 0u;
}

__global__ __launch_bounds__(256) void k_stage1_fill(const uint32_t* __restrict__ order, const uint32_t* __restrict__ head,
                                                     const uint32_t* __restrict__ run, const uint2* __restrict__ ent,
                                                     const uint32_t* __restrict__ row_of, int64_t n,
                                                     uint2* __restrict__ ent1, uint32_t* __restrict__ ptr1,
                                                     uint32_t* __restrict__ row1, const uint32_t* __restrict__ key_s,
                                                     uint32_t* __restrict__ key1)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t e = order[i];
    ent1[i] = ent[e];
    if (head[i]) {
        const uint32_t p = run[i];
        ptr1[p] = (uint32_t)i;
        row1[p] = row_of[e];
        if (key1) key1[p] = key_s[i];
    }
}

__global__ __launch_bounds__(256) void k_count_rows(const uint32_t* __restrict__ row1, int64_t n,
                                                    uint32_t* __restrict__ cnt)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(cnt + row1[i], 1u);
}

__global__ __launch_bounds__(256) void k_stage2_fill(const uint32_t* __restrict__ order2, int64_t n,
                                                     uint2* __restrict__ ent2)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) ent2[i] = make_uint2(order2[i], 0x3f800000u);  // weight 1.0f
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) { *p = v; }

__global__ __launch_bounds__(256) void k_set_u32_n(uint32_t* p, int n, uint32_t v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------ gather SpMM

struct GatherGroup {
    const uint32_t* ptr;  // [rows+1]
    const uint2* ent;     // entries (index, float bits of value)
    const float* in;      // [index] rows of 8*B floats
    float* out;           // [row] rows of 8*B floats
    int B;
    int rows;
    int item_begin;       // first work item of this group
    float scale;          // applied to every output of the group (1/xfactor)
    int partial;          // 1: out is a stage-1 workspace, read back for M materials only
};

struct GatherArgs {
    GatherGroup g[kMaxGroups];
    int n_groups;
    int n_items;
    int rows_per_item;
    int chunked;          // 1: each XCD takes one contiguous eighth of the items
};

// Workgroups are dispatched round-robin over the XCDs, each with its own L2.
// Renumber them so that the workgroups resident on one XCD take consecutive
// work items: adjacent rays (or voxels) share most of their gathered data,
// which then stays in that XCD's L2 (a bijection when the grid is a multiple
// of kXcds; other grids keep the hardware order).
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_block()
{
    if (gridDim.x % kXcds != 0) return blockIdx.x;
    const int per = gridDim.x / kXcds;
    return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}

// Items [first, end) with stride step for this workgroup.  Default: rounds of
// gridDim.x items, each XCD taking a consecutive run of every round.
// chunked: XCD x takes items [x*n/8, (x+1)*n/8) for the whole launch, so a
// block of the blocked form (one tile's stage-1 rows) is read into one L2.
struct ItemRange {
    int first, end, step;
};

__device__ __forceinline__ ItemRange item_range(const GatherArgs& a)
{
    if (a.chunked && gridDim.x % kXcds == 0) {
        const int x = blockIdx.x % kXcds, per = gridDim.x / kXcds;
        const int lo = (int)((int64_t)a.n_items * x / kXcds);
        const int hi = (int)((int64_t)a.n_items * (x + 1) / kXcds);
        return {lo + (int)blockIdx.x / kXcds, hi, per};
    }
    return {xcd_block(), a.n_items, (int)gridDim.x};
}

__device__ __forceinline__ int group_of(const GatherArgs& a, int item)
{
    int gi = 0;
#pragma unroll 1
    for (int k = 1; k < a.n_groups; k++)
        if (item >= a.g[k].item_begin) gi = k;
    return gi;
}

// acc[m] += in[idx][b][m] * val for the U entries j, j+stride, ..., in that
// order: all U entry loads are issued first, then all U gathers, so a lane
// keeps U independent loads in flight.
template <int M, int U>
__device__ __forceinline__ void gather_step(const uint2* __restrict__ ent, const float4* __restrict__ in4, int B, int b,
                                            uint32_t j, uint32_t stride, double (&acc)[M])
{
    uint2 e[U];
#pragma unroll
    for (int u = 0; u < U; u++) e[u] = ent[j + u * stride];
    float4 lo[U], hi[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const size_t i = ((size_t)e[u].x * B + b) * 2;
        lo[u] = in4[i];
        hi[u] = M > 4 ? in4[i + 1] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const double v = __uint_as_float(e[u].y);
        const float x[8] = {lo[u].x, lo[u].y, lo[u].z, lo[u].w, hi[u].x, hi[u].y, hi[u].z, hi[u].w};
#pragma unroll
        for (int m = 0; m < M; m++) acc[m] = __builtin_fma((double)x[m], v, acc[m]);
    }
}

// Lane K of a row's LPR lanes -> every lane of the row (all of them active):
// a DPP row broadcast when a row is one 16-lane DPP row, a lane read when it
// is the whole wave, a cross-lane permute otherwise.
template <int LPR, int K>
__device__ __forceinline__ uint32_t row_bcast(uint32_t v)
{
    if constexpr (LPR == 16)
        return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x150 + K, 0xF, 0xF, false);  // row_newbcast:K
    else if constexpr (LPR == 64)
        return (uint32_t)__builtin_amdgcn_readlane((int)v, K);
    else
        return (uint32_t)__shfl((int)v, (int)(__lane_id() & ~(LPR - 1)) + K, 64);
}

template <int LPR, int K>
__device__ __forceinline__ uint2 entry_bcast(uint2 mine)
{
    return make_uint2(row_bcast<LPR, K>(mine.x), row_bcast<LPR, K>(mine.y));
}

// Row layout of the batched arrays.  Row r (a voxel, a ray or a stage-1 run)
// holds 8*B floats in three planes, each phantom-major:
//   [0, 4B)    materials 0-3 of phantom b at 4b,
//   [4B, 6B)   materials 4-5 at 4B + 2b,
//   [6B, 8B)   materials 6-7 at 6B + 2b,
// with the materials >= M zero.  The lanes of a row (one per phantom) read
// an entry's row as one 16-byte load per lane for M <= 4, plus one 8-byte
// load for M <= 6 and another for M <= 8, each plane a dense run of lines
// (B = 16: 256 + 128 + 128 bytes).  For B = 1 a row is the plain [8] of the
// single-phantom kernels.
template <int M>
__device__ __forceinline__ void load_slot(const float* __restrict__ in, size_t row, int B, int b, float (&x)[8])
{
    const float* r = in + row * (size_t)(8 * B);
    const float4 lo = *reinterpret_cast<const float4*>(r + 4 * b);
    x[0] = lo.x; x[1] = lo.y; x[2] = lo.z; x[3] = lo.w;
    x[4] = x[5] = x[6] = x[7] = 0.f;
    if constexpr (M > 4) {
        const float2 p = *reinterpret_cast<const float2*>(r + 4 * B + 2 * b);
        x[4] = p.x; x[5] = p.y;
    }
    if constexpr (M > 6) {
        const float2 p = *reinterpret_cast<const float2*>(r + 6 * B + 2 * b);
        x[6] = p.x; x[7] = p.y;
    }
}

// The planes of (row, b) that load_slot<M> reads (stage-1 partial sums: the
// zero planes of materials >= M are never read back, so not written).
template <int M>
__device__ __forceinline__ void store_slot_m(float* __restrict__ out, size_t row, int B, int b, const float (&x)[8])
{
    float* r = out + row * (size_t)(8 * B);
    *reinterpret_cast<float4*>(r + 4 * b) = make_float4(x[0], x[1], x[2], x[3]);
    if constexpr (M > 4) *reinterpret_cast<float2*>(r + 4 * B + 2 * b) = make_float2(x[4], x[5]);
    if constexpr (M > 6) *reinterpret_cast<float2*>(r + 6 * B + 2 * b) = make_float2(x[6], x[7]);
}

// All 8 values of (row, b), padding included.
__device__ __forceinline__ void store_slot(float* __restrict__ out, size_t row, int B, int b, const float (&x)[8])
{
    float* r = out + row * (size_t)(8 * B);
    *reinterpret_cast<float4*>(r + 4 * b) = make_float4(x[0], x[1], x[2], x[3]);
    *reinterpret_cast<float2*>(r + 4 * B + 2 * b) = make_float2(x[4], x[5]);
    *reinterpret_cast<float2*>(r + 6 * B + 2 * b) = make_float2(x[6], x[7]);
}

template <int M>
__device__ __forceinline__ void fma_slot(const float (&x)[8], uint32_t vbits, double (&acc)[M])
{
    const double v = __uint_as_float(vbits);
#pragma unroll
    for (int m = 0; m < M; m++) acc[m] = __builtin_fma((double)x[m], v, acc[m]);
}

template <int M, int LPR, int K>
__device__ __forceinline__ void gather_one(uint2 mine, int n, const float* __restrict__ in, int B, int b, bool act,
                                           double (&acc)[M])
{
    const uint2 e = entry_bcast<LPR, K>(mine);  // every lane of the row takes part
    if (act && K < n) {
        float x[8];
        load_slot<M>(in, e.x, B, b, x);
        fma_slot<M>(x, e.y, acc);
    }
}

// Entries U0..U0+3 of a chunk of n (the entry of chunk position k sits in
// lane k of the row, `mine`): broadcast by every lane of the row, then the
// lanes of real phantoms (act) issue the four loads before the FMAs, in
// entry order.  A group cut by the end of the row takes its entries one by
// one.
template <int M, int LPR, int U0>
__device__ __forceinline__ void gather_group(uint2 mine, int n, const float* __restrict__ in, int B, int b, bool act,
                                             double (&acc)[M])
{
    if (U0 + 4 <= n) {
        const uint2 e0 = entry_bcast<LPR, U0>(mine), e1 = entry_bcast<LPR, U0 + 1>(mine);
        const uint2 e2 = entry_bcast<LPR, U0 + 2>(mine), e3 = entry_bcast<LPR, U0 + 3>(mine);
        if (!act) return;
        float x0[8], x1[8], x2[8], x3[8];
        load_slot<M>(in, e0.x, B, b, x0);
        load_slot<M>(in, e1.x, B, b, x1);
        load_slot<M>(in, e2.x, B, b, x2);
        load_slot<M>(in, e3.x, B, b, x3);
        fma_slot<M>(x0, e0.y, acc);
        fma_slot<M>(x1, e1.y, acc);
        fma_slot<M>(x2, e2.y, acc);
        fma_slot<M>(x3, e3.y, acc);
    } else {
        // per-entry tail; the broadcasts stay outside any lane-dependent branch
        gather_one<M, LPR, U0>(mine, n, in, B, b, act, acc);
        gather_one<M, LPR, U0 + 1>(mine, n, in, B, b, act, acc);
        gather_one<M, LPR, U0 + 2>(mine, n, in, B, b, act, acc);
    }
}

// out[row][b][m] = scale * sum_{(idx,val) in row} in[idx][b][m] * val
// LPR lanes per row (one per phantom, B <= LPR), 256/LPR rows per item.  A
// row's entries are read in chunks of CH: lane k loads entry k of the chunk
// (one coalesced load, issued one chunk ahead) and the row broadcasts it, so
// the vector memory path carries each entry once instead of once per lane.
template <int M, int LPR>
__global__ __launch_bounds__(256) void k_gather(GatherArgs a)
{
    constexpr int CH = LPR < 16 ? LPR : 16;
    const int lane_row = threadIdx.x / LPR;
    const int b = threadIdx.x % LPR;
    const ItemRange r = item_range(a);
    for (int item = r.first; item < r.end; item += r.step) {
        const GatherGroup& G = a.g[group_of(a, item)];
        const int row = (item - G.item_begin) * a.rows_per_item + lane_row;
        // all LPR lanes of a row (one wave) take this branch together
        if (row >= G.rows) continue;
        const bool act = b < G.B;
        const uint32_t j0 = G.ptr[row], j1 = G.ptr[row + 1];
        double acc[M];
#pragma unroll
        for (int m = 0; m < M; m++) acc[m] = 0.0;
        uint2 mine = j0 + b < j1 && b < CH ? G.ent[j0 + b] : make_uint2(0u, 0u);
        for (uint32_t jb = j0; jb < j1; jb += CH) {
            const int n = (int)min(j1 - jb, (uint32_t)CH);
            const uint32_t jn = jb + CH;
            const uint2 next = jn + b < j1 && b < CH ? G.ent[jn + b] : make_uint2(0u, 0u);
            gather_group<M, LPR, 0>(mine, n, G.in, G.B, b, act, acc);
            if (n > 4) gather_group<M, LPR, 4>(mine, n, G.in, G.B, b, act, acc);
            if constexpr (CH > 8) {
                if (n > 8) gather_group<M, LPR, 8>(mine, n, G.in, G.B, b, act, acc);
                if (n > 12) gather_group<M, LPR, 12>(mine, n, G.in, G.B, b, act, acc);
            }
            mine = next;
        }
        if (!act) continue;
        float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int m = 0; m < M; m++) o[m] = (float)(acc[m] * G.scale);
        if (G.partial)
            store_slot_m<M>(G.out, row, G.B, b, o);
        else
            store_slot(G.out, row, G.B, b, o);
    }
}

// One phantom per group (B = 1, where a row is the plain [8]): the SPL lanes of a row split its entries
// (lane k takes entries j0+k, j0+k+SPL, ...: the row is read in contiguous
// SPL*8-byte runs) and combine their FP64 partial sums by a butterfly over
// the SPL lanes; 256/SPL rows per item.
template <int M, int SPL>
__global__ __launch_bounds__(256) void k_gather_split(GatherArgs a)
{
    static_assert(SPL >= 2 && SPL <= 64 && (SPL & (SPL - 1)) == 0, "SPL is a power of two <= 64");
    const int lane_row = threadIdx.x / SPL;
    const int k = threadIdx.x % SPL;
    const ItemRange r = item_range(a);
    for (int item = r.first; item < r.end; item += r.step) {
        const GatherGroup& G = a.g[group_of(a, item)];
        const int row = (item - G.item_begin) * a.rows_per_item + lane_row;
        // every lane of a row takes the same branch, and the SPL lanes of a
        // row lie in one wave, so the shuffles below see only active lanes
        if (row >= G.rows) continue;
        const uint32_t j0 = G.ptr[row], j1 = G.ptr[row + 1];
        double acc[M];
#pragma unroll
        for (int m = 0; m < M; m++) acc[m] = 0.0;
        const float4* in4 = reinterpret_cast<const float4*>(G.in);
        uint32_t j = j0 + k;
        for (; j + 3 * SPL < j1; j += 4 * SPL) gather_step<M, 4>(G.ent, in4, 1, 0, j, SPL, acc);
        for (; j < j1; j += SPL) gather_step<M, 1>(G.ent, in4, 1, 0, j, SPL, acc);
#pragma unroll
        for (int off = SPL / 2; off > 0; off >>= 1) {
#pragma unroll
            for (int m = 0; m < M; m++) acc[m] += __shfl_xor(acc[m], off, SPL);
        }
        if (k == 0) {
            float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < M; m++) o[m] = (float)(acc[m] * G.scale);
            float4* dst = reinterpret_cast<float4*>(G.out) + (size_t)row * 2;
            dst[0] = make_float4(o[0], o[1], o[2], o[3]);
            dst[1] = make_float4(o[4], o[5], o[6], o[7]);
        }
    }
}

// ---------------------------------------------------- batched spectral sum

constexpr int kBatchSpecGroups = 16;
struct BatchSpecArgs {
    const float* L;       // [rays] rows of 8*B floats
    const float* att;     // [S][M][NE]
    const float* pdf;     // [S][NE]
    const float* target;  // [B][S][rays] (loss mode)
    float* raw;           // [B][S][rays] (forward mode)
    uint32_t* stats;      // [B][S][2]
    float* w;             // [rays] rows of 8*B floats (loss mode)
    float* loss;          // [B] (loss mode)
    int rays, W, VT, B, S, M, NE;
    // forward mode over several geometries (ngroups > 0): phantom z of the
    // launch is phantom z - pref[g] of group g, whose rows (of 8 * Bg[g]
    // floats) start at Lg[g]; raw and stats are indexed by z
    int ngroups;
    const float* Lg[kBatchSpecGroups];
    int Bg[kBatchSpecGroups];
    int pref[kBatchSpecGroups + 1];
};

// The rows, phantom count and phantom index in its group of launch phantom z.
__device__ __forceinline__ void spec_slot(const BatchSpecArgs& a, int z, const float*& L, int& B, int& b)
{
    if (a.ngroups <= 0) {
        L = a.L;
        B = a.B;
        b = z;
        return;
    }
    int g = 0;
    while (g + 1 < a.ngroups && z >= a.pref[g + 1]) g++;
    L = a.Lg[g];
    B = a.Bg[g];
    b = z - a.pref[g];
}

// One past the highest bin with a nonzero probability among the bins this
// wave's lanes hold (bin e on this lane, consecutive bins on consecutive
// lanes), or 0; called by every active lane, it needs no atomics.
__device__ __forceinline__ int last_nonzero_bin(int e, float p)
{
    const unsigned long long m = __ballot(p != 0.f);
    const int base = e - (int)(threadIdx.x & 63);  // the bin of lane 0
    return m ? base + 64 - __clzll((long long)m) : 0;
}

// Forward mode: raw and per-(b,s) max/min, one (b,s) per blockIdx.y/z.
// Launched with 256 threads (four waves).
template <int M>
__global__ __launch_bounds__(256) void k_batch_spectral(BatchSpecArgs a)
{
    extern __shared__ float lds[];
    float* mu2 = lds;
    float* pdf = lds + a.NE * 8;
    __shared__ uint32_t red[2][4];
    __shared__ int ne_w[4];
    const int s = blockIdx.y, z = blockIdx.z, NE = a.NE;
    const float* Lrows;
    int B, b;
    spec_slot(a, z, Lrows, B, b);
    const float kLog2e = 1.4426950408889634f;
    for (int i = threadIdx.x; i < NE * 8; i += blockDim.x) {
        const int e = i >> 3, m = i & 7;
        mu2[i] = (m < M) ? -a.att[((size_t)s * M + m) * NE + e] * kLog2e : 0.f;
    }
    // bins past the spectrum's last nonzero probability (the zero padding of
    // spectra of unequal length) add exactly +0 and are skipped
    int last = 0;
    for (int e = threadIdx.x; e < NE; e += blockDim.x) {
        const float p = a.pdf[(size_t)s * NE + e];
        pdf[e] = p;
        last = max(last, last_nonzero_bin(e, p));
    }
    if ((threadIdx.x & 63) == 0) ne_w[threadIdx.x >> 6] = last;
    __syncthreads();
    const int ne = max(max(ne_w[0], ne_w[1]), max(ne_w[2], ne_w[3]));
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = ray < a.rays;
    float I = 0.f;
    if (valid && ray / a.W < a.VT) {
        float Lm[8];
        load_slot<M>(Lrows, ray, B, b, Lm);
        for (int e = 0; e < ne; e++) {
            const float4 m0 = reinterpret_cast<const float4*>(mu2)[2 * e];
            const float4 m1 = reinterpret_cast<const float4*>(mu2)[2 * e + 1];
            const float mu[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
            float x = 0.f;
#pragma unroll
            for (int m = 0; m < M; m++) x = fmaf(mu[m], Lm[m], x);
            I += pdf[e] * __builtin_amdgcn_exp2f(x);  // as k_spectral
        }
    }
    const size_t bs = (size_t)z * a.S + s;
    if (valid) a.raw[bs * a.rays + ray] = I;
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
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mx = max(mx, red[0][w]); mn = min(mn, red[1][w]); }
        atomicMax(a.stats + 2 * bs, mx);
        atomicMin(a.stats + 2 * bs + 1, mn);
    }
}

// Forward mode with every spectrum of phantom blockIdx.z in one block: the
// tables and the ray's path integrals are loaded once for all S spectra.
// Each (phantom, spectrum, ray) is summed exactly as in k_batch_spectral.
// Rays per workgroup of k_batch_spectral_all: each workgroup adds its max and
// min of each (phantom, spectrum) to that pair's statistics words with one
// atomic each, which serialise per address, so workgroups are large.
constexpr int kSpecAllBlock = 1024, kSpecAllWaves = kSpecAllBlock / 64;

template <int M>
__global__ __launch_bounds__(kSpecAllBlock) void k_batch_spectral_all(BatchSpecArgs a)
{
    extern __shared__ float lds[];
    const int S = a.S, NE = a.NE, z = blockIdx.z;
    const float* Lrows;
    int B, b;
    spec_slot(a, z, Lrows, B, b);
    float* mu2 = lds;                                          // [S][NE][8]
    float* pdf = lds + (size_t)S * NE * 8;                     // [S][NE]
    int* ne_w = reinterpret_cast<int*>(pdf + (size_t)S * NE);  // [S][kSpecAllWaves]
    __shared__ uint32_t red[2][kSpecAllWaves];
    const float kLog2e = 1.4426950408889634f;
    for (int i = threadIdx.x; i < S * NE * 8; i += blockDim.x) {
        const int s = i / (NE * 8), r = i - s * NE * 8, e = r >> 3, m = r & 7;
        mu2[i] = (m < M) ? -a.att[((size_t)s * M + m) * NE + e] * kLog2e : 0.f;
    }
    for (int s = 0; s < S; s++) {
        int last = 0;
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const float p = a.pdf[(size_t)s * NE + e];
            pdf[s * NE + e] = p;
            last = max(last, last_nonzero_bin(e, p));
        }
        if ((threadIdx.x & 63) == 0) ne_w[kSpecAllWaves * s + (threadIdx.x >> 6)] = last;
    }
    __syncthreads();
    const int ray = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = ray < a.rays;
    const bool hit = valid && ray / a.W < a.VT;
    float Lm[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    if (hit) load_slot<M>(Lrows, ray, B, b, Lm);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int s = 0; s < S; s++) {
        int ne = 0;
        for (int w = 0; w < kSpecAllWaves; w++) ne = max(ne, ne_w[kSpecAllWaves * s + w]);
        const float* mus = mu2 + (size_t)s * NE * 8;
        const float* ps = pdf + (size_t)s * NE;
        float I = 0.f;
        if (hit) {
            for (int e = 0; e < ne; e++) {
                const float4 m0 = reinterpret_cast<const float4*>(mus)[2 * e];
                const float4 m1 = reinterpret_cast<const float4*>(mus)[2 * e + 1];
                const float mu[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
                float x = 0.f;
#pragma unroll
                for (int m = 0; m < M; m++) x = fmaf(mu[m], Lm[m], x);
                I += ps[e] * __builtin_amdgcn_exp2f(x);  // as k_spectral
            }
        }
        const size_t bs = (size_t)z * S + s;
        if (valid) a.raw[bs * a.rays + ray] = I;
        uint32_t mx = valid ? __float_as_uint(I) : 0u;
        uint32_t mn = valid ? __float_as_uint(I) : 0x7f800000u;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mx = max(mx, (uint32_t)__shfl_xor((int)mx, off, 64));
            mn = min(mn, (uint32_t)__shfl_xor((int)mn, off, 64));
        }
        if (lane == 0) { red[0][wave] = mx; red[1][wave] = mn; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < (int)(blockDim.x >> 6); w++) { mx = max(mx, red[0][w]); mn = min(mn, red[1][w]); }
            atomicMax(a.stats + 2 * bs, mx);
            atomicMin(a.stats + 2 * bs + 1, mn);
        }
        __syncthreads();  // red is reused by the next spectrum
    }
}

// Loss mode: per (ray, phantom) over all spectra,
//   r_s = ln(I_s / I_max,s) - ln t_s,  loss_b += sum_s r_s^2,
//   w[ray][b][m] = sum_s (2 r_s / I_s) dI_s/dL_m,
// with I_max,s = sum_E pdf_s(E) (the value of every ray that misses the object).
template <int M>
__global__ __launch_bounds__(256) void k_batch_loss(BatchSpecArgs a, int lpr)
{
    extern __shared__ float lds[];
    const int NE = a.NE, S = a.S;
    float* mu2 = lds;                    // [S][NE][8]
    float* pdf = lds + (size_t)S * NE * 8;  // [S][NE]
    float* imax = pdf + (size_t)S * NE;     // [S]
    float* lsum = imax + S;                 // [lpr]
    int* ne_w = reinterpret_cast<int*>(lsum + lpr);  // [S][4]
    const float kLog2e = 1.4426950408889634f, kLn2 = 0.69314718055994531f;
    for (int i = threadIdx.x; i < S * NE * 8; i += blockDim.x) {
        const int s = i / (NE * 8), r = i - s * NE * 8, e = r >> 3, m = r & 7;
        mu2[i] = (m < M) ? -a.att[((size_t)s * M + m) * NE + e] * kLog2e : 0.f;
    }
    for (int i = threadIdx.x; i < lpr; i += blockDim.x) lsum[i] = 0.f;
    // bins past a spectrum's last nonzero probability add exactly +0 to I and
    // to every d[m], and are skipped
    for (int s = 0; s < S; s++) {
        int last = 0;
        for (int e = threadIdx.x; e < NE; e += blockDim.x) {
            const float p = a.pdf[(size_t)s * NE + e];
            pdf[s * NE + e] = p;
            last = max(last, last_nonzero_bin(e, p));
        }
        if ((threadIdx.x & 63) == 0) ne_w[4 * s + (threadIdx.x >> 6)] = last;
    }
    __syncthreads();
    if (threadIdx.x < S) {
        float t = 0.f;
        for (int e = 0; e < NE; e++) t += pdf[threadIdx.x * NE + e] * __builtin_amdgcn_exp2f(0.f);
        imax[threadIdx.x] = t;
    }
    __syncthreads();
    const int b = threadIdx.x % lpr;
    const int ray = blockIdx.x * (blockDim.x / lpr) + threadIdx.x / lpr;
    float wout[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    float l = 0.f;
    const bool valid = ray < a.rays && b < a.B;
    if (valid && ray / a.W < a.VT) {
        float Lm[8];
        load_slot<M>(a.L, ray, a.B, b, Lm);
        for (int s = 0; s < S; s++) {
            float I = 0.f, d[M];
#pragma unroll
            for (int m = 0; m < M; m++) d[m] = 0.f;
            const float* mus = mu2 + (size_t)s * NE * 8;
            const float* ps = pdf + (size_t)s * NE;
            const int ne = max(max(ne_w[4 * s], ne_w[4 * s + 1]), max(ne_w[4 * s + 2], ne_w[4 * s + 3]));
            for (int e = 0; e < ne; e++) {
                // the 8 table entries of bin e in two 16-byte LDS reads
                const float4 m0 = reinterpret_cast<const float4*>(mus)[2 * e];
                const float4 m1 = reinterpret_cast<const float4*>(mus)[2 * e + 1];
                const float mu[8] = {m0.x, m0.y, m0.z, m0.w, m1.x, m1.y, m1.z, m1.w};
                float x = 0.f;
#pragma unroll
                for (int m = 0; m < M; m++) x = fmaf(mu[m], Lm[m], x);
                // hardware exp2: only terms below 2^-126 (flushed) differ from exp2f
                const float t = ps[e] * __builtin_amdgcn_exp2f(x);
                I += t;
#pragma unroll
                for (int m = 0; m < M; m++) d[m] = fmaf(t, mu[m], d[m]);
            }
            const float tg = a.target[((size_t)b * S + s) * a.rays + ray];
            const float r = logf(I / imax[s]) - logf(tg);
            l += r * r;
            const float gI = 2.f * r / I;
#pragma unroll
            for (int m = 0; m < M; m++) wout[m] = fmaf(gI, d[m] * kLn2, wout[m]);
        }
    }
    if (valid) {
        store_slot(a.w, ray, a.B, b, wout);
        atomicAdd(lsum + b, l);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < min(lpr, a.B); i += blockDim.x) atomicAdd(a.loss + i, lsum[i]);
}

// ------------------------------------------- Barzilai-Borwein descent step
//
// The arrays hold n rows of 8*B floats (the layout of load_slot); slot
// k = (row k / B, phantom k % B) holds the 8 values of one voxel of one
// phantom.  The grid stride (kBBBlocksPerPhantom * B blocks of 256 threads)
// is a multiple of B, so every thread stays on one phantom; sums are FP64 in
// a fixed order (per thread, then per block over the threads of that phantom,
// then over the blocks), so the step is deterministic.
constexpr int kBBBlocksPerPhantom = 64;

struct BBArgs {
    float* x;          // densities (updated)
    float* x_prev;     // previous densities (overwritten with x)
    float* g_prev;     // previous gradient (overwritten with g)
    const float* g;    // gradient
    float* alpha;      // [B] step lengths
    double* part;      // [blocks][B][2] partial sums (s.y, s.s)
    int64_t slots;     // n * B
    int B;
    int first;         // 1: no previous iterate (plain projected step)
};

// This thread's phantom, first row and row stride (the grid stride is a
// multiple of B).
struct BBLane {
    int b;
    int64_t row, step, rows;
};

__device__ __forceinline__ BBLane bb_lane(const BBArgs& a)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    return {(int)(k % a.B), k / a.B, stride / a.B, a.slots / a.B};
}

__global__ __launch_bounds__(256) void k_bb_partial(BBArgs a)
{
    __shared__ double red[2][256];
    double sy = 0.0, ss = 0.0;
    const BBLane l = bb_lane(a);
    for (int64_t r = l.row; r < l.rows; r += l.step) {
        float x[8], xp[8], g[8], gp[8];
        load_slot<8>(a.x, r, a.B, l.b, x);
        load_slot<8>(a.x_prev, r, a.B, l.b, xp);
        load_slot<8>(a.g, r, a.B, l.b, g);
        load_slot<8>(a.g_prev, r, a.B, l.b, gp);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const float s = x[i] - xp[i], y = g[i] - gp[i];
            sy += (double)s * (double)y;
            ss += (double)s * (double)s;
        }
    }
    red[0][threadIdx.x] = sy;
    red[1][threadIdx.x] = ss;
    __syncthreads();
    // thread j sums the threads j, j+B, ... of the block: one phantom each
    if ((int)threadIdx.x < a.B && (int)threadIdx.x < (int)blockDim.x) {
        double t0 = 0.0, t1 = 0.0;
        for (int t = threadIdx.x; t < (int)blockDim.x; t += a.B) {
            t0 += red[0][t];
            t1 += red[1][t];
        }
        const int b = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) % a.B);
        a.part[((size_t)blockIdx.x * a.B + b) * 2] = t0;
        a.part[((size_t)blockIdx.x * a.B + b) * 2 + 1] = t1;
    }
}

// alpha_b = clamp(sy > 0 ? ss / max(sy, 1e-30) : 2 alpha_b, 1e-8, 1e2)
__global__ void k_bb_alpha(BBArgs a, int blocks)
{
    const int b = threadIdx.x;
    if (b >= a.B || a.first) return;
    double sy = 0.0, ss = 0.0;
    for (int k = 0; k < blocks; k++) {
        sy += a.part[((size_t)k * a.B + b) * 2];
        ss += a.part[((size_t)k * a.B + b) * 2 + 1];
    }
    const double bb = sy > 0.0 ? ss / fmax(sy, 1e-30) : 2.0 * (double)a.alpha[b];
    a.alpha[b] = (float)fmin(fmax(bb, 1e-8), 1e2);
}

// x_prev = x, g_prev = g, x = max(x - alpha_b g, 0)
__global__ __launch_bounds__(256) void k_bb_apply(BBArgs a)
{
    const BBLane l = bb_lane(a);
    const float al = a.alpha[l.b];
    for (int64_t r = l.row; r < l.rows; r += l.step) {
        float x[8], g[8], xn[8];
        load_slot<8>(a.x, r, a.B, l.b, x);
        load_slot<8>(a.g, r, a.B, l.b, g);
#pragma unroll
        for (int i = 0; i < 8; i++) xn[i] = fmaxf(x[i] - al * g[i], 0.f);
        store_slot(a.x_prev, r, a.B, l.b, x);
        store_slot(a.g_prev, r, a.B, l.b, g);
        store_slot(a.x, r, a.B, l.b, xn);
    }
}

__global__ void k_batch_stats_init(uint32_t* stats, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        stats[2 * i] = 0u;
        stats[2 * i + 1] = 0x7f800000u;
    }
}

__global__ void k_zero_f32(float* p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.f;
}


// ------------------------------------------- tile-resident forward stage 1

constexpr int kTileThreads = 1024;  // one workgroup per CU (LDS)
constexpr int kTileUnroll = 8;      // entries in flight per lane
// Tile-local voxel code of padding entries: "row 64", the zeroed LDS slots
// after the tile, so padding adds exact zeros whatever the tile holds (an inf
// or NaN density included).
constexpr uint32_t kTilePadVoxel = 64u * 64u;

// Adjoint blocks.  A block is one part of a 64x64 tile in one view block:
// the build splits each (view block, tile) into 1, 2, 4, 8 or 16 parts on a
// 4x4 grid of 16x16 cells, the coarsest split whose ray tables all hold at
// most kAdjSlotCap slots (slot 0, the zero slot of the padding entries,
// included).  A table of S slots takes S * MP floats of LDS: at M <= 6 two
// 512-lane workgroups fit one CU, so one loads its table while the other
// sums.
constexpr int kAtileThreads = 512;
constexpr int kAdjSlotCap = 3328;       // 78 KB at MP = 6
constexpr int kAdjCells = 16;           // 16x16 cells of a 64x64 tile
static_assert(kAdjTile == 64, "adjoint LDS blocks split 64x64 tiles into 4x4 cells of 16x16");
constexpr int kAdjPadSlot = 0;

__global__ __launch_bounds__(256) void k_fill_pad(uint2* __restrict__ p, size_t n, uint32_t pad)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = make_uint2(pad, 0u);
}

// Build: per stage-1 run, sort key tile << 16 | (0xffff - length) (longest
// first within a tile), and the first run of each tile (the runs are in tile
// order, so each tile has one first run; no counter is contended).
__global__ __launch_bounds__(256) void k_tile_run_keys(const uint32_t* __restrict__ ptr1, const uint2* __restrict__ ent1,
                                                       uint32_t rows1, TileGrid tg, uint32_t* __restrict__ key,
                                                       uint32_t* __restrict__ val, uint32_t* __restrict__ tile_first,
                                                       uint32_t* __restrict__ err)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows1) return;
    const uint32_t len = ptr1[r + 1] - ptr1[r];
    const uint32_t t = tg.tile(ent1[ptr1[r]].x);
    const uint32_t tp = r ? tg.tile(ent1[ptr1[r - 1]].x) : 0xffffffffu;
    if (len == 0u || len >= 0xffffu || (r && t < tp)) atomicOr(err, 1u);
    key[r] = (t << 16) | (0xffffu - min(len, 0xfffeu));
    val[r] = r;
    if (t != tp) tile_first[t] = r;
}

// Build: run counts from first runs (0xffffffff: a tile without runs), in
// place; one thread.
__global__ void k_first_to_count(uint32_t* __restrict__ c, int ntiles, uint32_t rows1)
{
    uint32_t next = rows1;
    for (int t = ntiles - 1; t >= 0; t--) {
        const uint32_t f = c[t];
        c[t] = f == 0xffffffffu ? 0u : next - f;
        if (f != 0xffffffffu) next = f;
    }
}

// Build: entries per lane of each slice = the length of its first (longest) run.
__global__ __launch_bounds__(256) void k_tile_slice_len(const uint32_t* __restrict__ ptr1, const uint32_t* __restrict__ order,
                                                        const int32_t* __restrict__ slice_tile,
                                                        const uint32_t* __restrict__ tile_begin,
                                                        const int32_t* __restrict__ tile_slices, int nslices,
                                                        uint32_t* __restrict__ slice_len)
{
    const int sl = blockIdx.x * blockDim.x + threadIdx.x;
    if (sl >= nslices) return;
    const int t = slice_tile[sl];
    const uint32_t r = order[tile_begin[t] + (uint32_t)(sl - tile_slices[t]) * 64u];
    slice_len[sl] = ptr1[r + 1] - ptr1[r];
}

// Build: the interleaved entries with tile-local voxel indices, one thread
// per run (in sorted order).
__global__ __launch_bounds__(256) void k_tile_fill(const uint32_t* __restrict__ ptr1, const uint2* __restrict__ ent1,
                                                   const uint32_t* __restrict__ key_s, const uint32_t* __restrict__ order,
                                                   uint32_t rows1, const uint32_t* __restrict__ tile_begin,
                                                   const int32_t* __restrict__ tile_slices,
                                                   const uint32_t* __restrict__ slice_ptr, TileGrid tg,
                                                   uint2* __restrict__ ell, int32_t* __restrict__ slice_runs,
                                                   uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows1) return;
    const uint32_t r = order[i], t = key_s[i] >> 16;
    const uint32_t p = i - tile_begin[t];
    const uint32_t sl = (uint32_t)tile_slices[t] + p / 64u, lane = p % 64u;
    slice_runs[sl * 64u + lane] = (int32_t)r;
    const uint32_t tx0 = (t % (uint32_t)tg.ntx) * 64u, ty0 = (t / (uint32_t)tg.ntx) * 64u;
    uint2* dst = ell + slice_ptr[sl] + lane;
    for (uint32_t j = ptr1[r], k = 0; j < ptr1[r + 1]; j++, k++) {
        const uint2 e = ent1[j];
        const uint32_t iy = e.x / (uint32_t)tg.N, ix = e.x - iy * (uint32_t)tg.N;
        const uint32_t ly = iy - ty0, lx = ix - tx0;  // wraps above 63 if outside the tile
        if (ly > 63u || lx > 63u) atomicOr(err, 2u);
        dst[(size_t)k * 64u] = make_uint2((ly & 63u) * 64u + (lx & 63u), e.y);
    }
}

// Adjoint build: the first stage-1 run of each blocked-form key (view block
// * tiles + rank of the 64x64 tile; the runs are in key order).
__global__ __launch_bounds__(256) void k_key_first(const uint32_t* __restrict__ key1, uint32_t rows1, uint32_t nkeys,
                                                   uint32_t* __restrict__ first, uint32_t* __restrict__ err)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows1) return;
    const uint32_t k = key1[r], kp = r ? key1[r - 1] : 0xffffffffu;
    if (k >= nkeys || (r && k < kp)) {
        atomicOr(err, 1u);
        return;
    }
    if (k != kp) first[k] = r;
}

__device__ __forceinline__ uint32_t adj_cell(uint32_t v, int N)
{
    const uint32_t iy = v / (uint32_t)N, ix = v - iy * (uint32_t)N;
    return ((iy & 63u) >> 4) * 4u + ((ix & 63u) >> 4);
}

// Adjoint build: first and last channel of each view in each 16x16 cell of a
// blocked-form key, over the entries of its runs; one workgroup per key.
__global__ __launch_bounds__(256) void k_atile_range(const uint32_t* __restrict__ ptr1, const uint2* __restrict__ ent1,
                                                     const uint32_t* __restrict__ row1,
                                                     const uint32_t* __restrict__ kbegin,
                                                     const uint32_t* __restrict__ kcount, int N, int W, int ntiles64,
                                                     int32_t* __restrict__ lo_out, int32_t* __restrict__ hi_out,
                                                     uint32_t* __restrict__ err)
{
    __shared__ int32_t lo[kAdjCells * kAdjViews], hi[kAdjCells * kAdjViews];
    const int K = blockIdx.x;
    for (int i = threadIdx.x; i < kAdjCells * kAdjViews; i += blockDim.x) {
        lo[i] = 0x7fffffff;
        hi[i] = -1;
    }
    __syncthreads();
    const uint32_t v0 = (uint32_t)(K / ntiles64) * kAdjViews;
    for (uint32_t q = threadIdx.x; q < kcount[K]; q += blockDim.x) {
        const uint32_t r = kbegin[K] + q, cell = adj_cell(row1[r], N);
        for (uint32_t j = ptr1[r]; j < ptr1[r + 1]; j++) {
            const uint32_t ray = ent1[j].x, v = ray / (uint32_t)W, c = ray - v * (uint32_t)W;
            const uint32_t jv = v - v0;  // wraps above kAdjViews - 1 outside the view block
            if (jv >= (uint32_t)kAdjViews) {
                atomicOr(err, 4u);
                continue;
            }
            atomicMin(&lo[cell * kAdjViews + jv], (int32_t)c);
            atomicMax(&hi[cell * kAdjViews + jv], (int32_t)c);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAdjCells * kAdjViews; i += blockDim.x) {
        lo_out[(size_t)K * kAdjCells * kAdjViews + i] = lo[i];
        hi_out[(size_t)K * kAdjCells * kAdjViews + i] = hi[i];
    }
}

// Adjoint build: the LDS block of each stage-1 run (cell_block[key * 16 +
// cell of the run's voxel]), its sort key block << 16 | (0xffff - length),
// and runs per block.
__global__ __launch_bounds__(256) void k_atile_run_keys(const uint32_t* __restrict__ ptr1,
                                                        const uint32_t* __restrict__ key1,
                                                        const uint32_t* __restrict__ row1, uint32_t rows1, int N,
                                                        const uint32_t* __restrict__ cell_block, uint32_t nblocks,
                                                        uint32_t* __restrict__ key, uint32_t* __restrict__ val,
                                                        uint32_t* __restrict__ cnt, uint32_t* __restrict__ err)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows1) return;
    const uint32_t len = ptr1[r + 1] - ptr1[r];
    const uint32_t bt = cell_block[(size_t)key1[r] * kAdjCells + adj_cell(row1[r], N)];
    if (len == 0u || len >= 0xffffu || bt >= nblocks) {
        atomicOr(err, 1u);
        key[r] = 0u;
    } else {
        key[r] = (bt << 16) | (0xffffu - len);
        atomicAdd(cnt + bt, 1u);
    }
    val[r] = r;
}

// Adjoint build: the interleaved entries with table slots, one thread per run
// (in sorted order).
__global__ __launch_bounds__(256) void k_atile_fill(const uint32_t* __restrict__ ptr1, const uint2* __restrict__ ent1,
                                                    const uint32_t* __restrict__ key_s,
                                                    const uint32_t* __restrict__ order, uint32_t rows1,
                                                    const uint32_t* __restrict__ begin,
                                                    const int32_t* __restrict__ tile_slices,
                                                    const uint32_t* __restrict__ slice_ptr,
                                                    const int32_t* __restrict__ slot_off,
                                                    const int32_t* __restrict__ slot_clo,
                                                    const int32_t* __restrict__ blk_v0, int W,
                                                    uint2* __restrict__ ell, int32_t* __restrict__ slice_runs,
                                                    uint32_t* __restrict__ err)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= rows1) return;
    const uint32_t r = order[i], K = key_s[i] >> 16;
    const uint32_t p = i - begin[K];
    const uint32_t sl = (uint32_t)tile_slices[K] + p / 64u, lane = p % 64u;
    slice_runs[sl * 64u + lane] = (int32_t)r;
    const uint32_t v0 = (uint32_t)blk_v0[K];
    const int32_t* off = slot_off + (size_t)K * (kAdjViews + 1);
    const int32_t* clo = slot_clo + (size_t)K * kAdjViews;
    uint2* dst = ell + slice_ptr[sl] + lane;
    for (uint32_t j = ptr1[r], k = 0; j < ptr1[r + 1]; j++, k++) {
        const uint2 e = ent1[j];
        const uint32_t v = e.x / (uint32_t)W, c = e.x - v * (uint32_t)W, jv = v - v0;
        uint32_t slot = (uint32_t)kAdjPadSlot;
        if (jv < (uint32_t)kAdjViews) {
            const int32_t d = (int32_t)c - clo[jv];
            if (d >= 0 && d < off[jv + 1] - off[jv]) slot = (uint32_t)(off[jv] + d);
            else atomicOr(err, 8u);
        } else {
            atomicOr(err, 4u);
        }
        dst[(size_t)k * 64u] = make_uint2(slot, e.y);
    }
}

struct TileArgs {
    const float* in;             // densities (forward) or ray weights (adjoint): rows of 8*B floats
    float* out;                  // stage-1 partial sums: rows of 8*B floats
    const uint2* ell;
    const uint32_t* slice_ptr;
    const uint32_t* slice_len;
    const int32_t* slice_runs;
    const int32_t* tile_slices;
    const int32_t* xcd_tiles;
    int kmax, ntx, N, B;
    const int32_t* slot_off;     // adjoint: ray tables of the blocks
    const int32_t* slot_clo;
    const int32_t* blk_v0;
    int W;
    const float* in_pm;          // adjoint: the weights phantom-major, [B][rays][8] (k_weights_pm)
    int rays;
};

// Adjoint: the weights of the batched layout copied phantom-major,
// out[b][ray][0..7] = the values load_slot<M> gives for (ray, b), so the
// table fill of one phantom reads consecutive rays from consecutive 32 bytes
// (in the batched rows, one phantom's values of consecutive rays are 8 * B
// floats apart: a cache line per ray).  Lanes of a ray read its row
// together; lpr lanes per ray, 256 / lpr rays per workgroup.
template <int M>
__global__ __launch_bounds__(256) void k_weights_pm(const float* __restrict__ in, int rays, int B, int lpr,
                                                   float* __restrict__ out)
{
    const int b = threadIdx.x % lpr;
    const int ray = blockIdx.x * (blockDim.x / lpr) + threadIdx.x / lpr;
    if (b >= B || ray >= rays) return;
    float v[8];
    load_slot<M>(in, ray, B, b, v);
    float4* d = reinterpret_cast<float4*>(out + ((size_t)b * rays + ray) * 8);
    d[0] = make_float4(v[0], v[1], v[2], v[3]);
    d[1] = make_float4(v[4], v[5], v[6], v[7]);
}

// Stage 1 with the block's data in LDS.  Workgroup g runs on XCD g % 8
// (round-robin dispatch) and takes phantom b of block xcd_tiles[xcd][k], so
// the B workgroups of a block share one L2 for its entries.
//   forward: the tile's M materials of phantom b go to LDS (64 rows of 64
//     voxels x MP floats, rows padded by 2 floats so that voxels one row
//     apart fall 2 banks apart);
//   adjoint: the M weights of phantom b of the block's rays go to LDS, slot
//     after slot (MP floats each), slot 0 zero.
// Each lane then sums one run of a slice in FP64, 8 entries in flight, and
// writes its partial sums as stage-1 row run, phantom b.  The sums are those
// of k_gather over the same runs, term for term in the same order.
// Several groups (one per geometry) in one launch: workgroup l belongs to the
// group gi with pref[gi] <= l < pref[gi + 1] (a scalar search).  Each group's
// range is a multiple of 8 workgroups, so l - pref[gi] keeps l's XCD.
struct TileGroups {
    TileArgs a[kMaxGroups];
    int pref[kMaxGroups + 1];
    int n;
};

template <int M, bool ADJ>
__global__ __launch_bounds__(ADJ ? kAtileThreads : kTileThreads) void k_gather_tile(TileGroups ga)
{
    int gi = 0;
    while (gi + 1 < ga.n && (int)blockIdx.x >= ga.pref[gi + 1]) gi++;
    const TileArgs& a = ga.a[gi];
    constexpr int MP = (M + 1) & ~1;  // floats per voxel or ray in LDS
    constexpr int ROW = 64 * MP + 2;  // floats per tile row (forward)
    constexpr int U = kTileUnroll;
    constexpr int LDSF = ADJ ? kAdjSlotCap * MP : 64 * ROW + 8;
    constexpr uint32_t PAD = ADJ ? (uint32_t)kAdjPadSlot : kTilePadVoxel;
    __shared__ float lds[LDSF];
    __shared__ int32_t tab[2 * kAdjViews + 2];  // adjoint: slot_off, slot_clo, first view of the block
    const int g = (int)blockIdx.x - ga.pref[gi], xcd = g % kXcds, slot = g / kXcds;
    const int k = slot / a.B, b = slot % a.B;
    if (k >= a.kmax) return;  // the whole workgroup
    const int t = a.xcd_tiles[xcd * a.kmax + k];
    if (t < 0) return;
    if constexpr (ADJ) {
        if (threadIdx.x <= kAdjViews) tab[threadIdx.x] = a.slot_off[(size_t)t * (kAdjViews + 1) + threadIdx.x];
        else if (threadIdx.x < 2 * kAdjViews + 1)
            tab[threadIdx.x] = a.slot_clo[(size_t)t * kAdjViews + threadIdx.x - (kAdjViews + 1)];
        else if (threadIdx.x == 2 * kAdjViews + 1)
            tab[threadIdx.x] = a.blk_v0[t];
        if (threadIdx.x < MP) lds[kAdjPadSlot * MP + threadIdx.x] = 0.f;
        __syncthreads();
        // slot s of view j (tab[j] <= s < tab[j + 1]) holds ray
        // (view block * kAdjViews + j) * W + clo_j + s - tab[j]; four slots
        // per lane in flight
        const int S = tab[kAdjViews], v0 = tab[2 * kAdjViews + 1];
        for (int s0 = 1 + (int)threadIdx.x; s0 < S; s0 += 4 * kAtileThreads) {
            float v[4][8];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int s = s0 + u * kAtileThreads;
                if (s < S) {
                    int lo = 0, hi = kAdjViews;
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (tab[mid] <= s) lo = mid;
                        else hi = mid;
                    }
                    const size_t ray = (size_t)(v0 + lo) * a.W + tab[kAdjViews + 1 + lo] + (s - tab[lo]);
                    const float4* src = reinterpret_cast<const float4*>(a.in_pm + ((size_t)b * a.rays + ray) * 8);
                    const float4 lo4 = src[0];
                    v[u][0] = lo4.x; v[u][1] = lo4.y; v[u][2] = lo4.z; v[u][3] = lo4.w;
                    const float4 hi4 = M > 4 ? src[1] : make_float4(0.f, 0.f, 0.f, 0.f);
                    v[u][4] = hi4.x; v[u][5] = hi4.y; v[u][6] = hi4.z; v[u][7] = hi4.w;
                }
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int s = s0 + u * kAtileThreads;
                if (s < S) {
                    float2* d = reinterpret_cast<float2*>(lds + s * MP);
#pragma unroll
                    for (int q = 0; q < MP / 2; q++) d[q] = make_float2(v[u][2 * q], v[u][2 * q + 1]);
                }
            }
        }
    } else {
        const int tx0 = (t % a.ntx) * 64, ty0 = (t / a.ntx) * 64;
        if (threadIdx.x < 8) lds[64 * ROW + threadIdx.x] = 0.f;
        for (int i = threadIdx.x; i < 64 * 64; i += blockDim.x) {
            const int ly = i >> 6, lx = i & 63, gy = ty0 + ly, gx = tx0 + lx;
            float v[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
            if (gy < a.N && gx < a.N) load_slot<M>(a.in, (size_t)gy * a.N + gx, a.B, b, v);
            float2* d = reinterpret_cast<float2*>(lds + ly * ROW + lx * MP);
#pragma unroll
            for (int q = 0; q < MP / 2; q++) d[q] = make_float2(v[2 * q], v[2 * q + 1]);
        }
    }
    __syncthreads();
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, nw = blockDim.x >> 6;
    for (int sl = a.tile_slices[t] + w; sl < a.tile_slices[t + 1]; sl += nw) {
        const int len = (int)a.slice_len[sl];
        const uint2* e = a.ell + a.slice_ptr[sl] + lane;
        double acc[MP];
#pragma unroll
        for (int m = 0; m < MP; m++) acc[m] = 0.0;
        uint2 q[U], nq[U];
        // padding entries (runs shorter than the slice, and in the forward
        // past len) are (PAD, 0.0f): zero LDS data times zero length.  An
        // adjoint len is a multiple of U, and the last prefetch reads the
        // next slice or the slack after the last one.
#pragma unroll
        for (int u = 0; u < U; u++) q[u] = ADJ || u < len ? e[u * 64] : make_uint2(PAD, 0u);
        for (int j = 0; j < len; j += U) {
#pragma unroll
            for (int u = 0; u < U; u++)
                nq[u] = ADJ || j + U + u < len ? e[(j + U + u) * 64] : make_uint2(PAD, 0u);
            float2 f[U][MP / 2];
#pragma unroll
            for (int u = 0; u < U; u++) {
                // 24-bit multiplies (full rate; v_mul_lo_u32 is quarter rate):
                // slots and tile rows are far below 2^24
                const float* base = ADJ ? lds + __umul24(q[u].x, (unsigned)MP)
                                        : lds + __umul24(q[u].x >> 6, (unsigned)ROW) + (q[u].x & 63) * MP;
                const float2* p = reinterpret_cast<const float2*>(base);
#pragma unroll
                for (int h = 0; h < MP / 2; h++) f[u][h] = p[h];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const double v = __uint_as_float(q[u].y);
#pragma unroll
                for (int h = 0; h < MP / 2; h++) {
                    acc[2 * h] = __builtin_fma((double)f[u][h].x, v, acc[2 * h]);
                    acc[2 * h + 1] = __builtin_fma((double)f[u][h].y, v, acc[2 * h + 1]);
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) q[u] = nq[u];
        }
        const int run = a.slice_runs[sl * 64 + lane];
        if (run >= 0) {
            float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int m = 0; m < M; m++) o[m] = (float)acc[m];
            store_slot_m<M>(a.out, (size_t)run, a.B, b, o);
        }
    }
}

}  // namespace vct

using namespace vct;

// ---------------------------------------------------------------- host API

static unsigned key_bits(uint64_t n_keys)
{
    unsigned b = 1;
    while ((1ull << b) < n_keys) b++;
    return b;
}

static void free_blocked(vct_blocked& f)
{
    void* bufs[] = {f.ptr1, f.ent1, f.ptr2, f.ent2};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    f = vct_blocked{};
}

// Regroup the n entries ent (row of entry i: row_of[i], entries in row order)
// by key[i] < n_keys into the two-stage form f over `rows` rows.  key is
// overwritten.  Stage-1 rows are the (key, row) runs in key-major order;
// stage 2 lists each row's runs in key order.  With run_info, the key and the
// row of each stage-1 run are handed out there ([rows1] each, the caller
// frees them).
struct RunInfo {
    uint32_t* key = nullptr;
    uint32_t* row = nullptr;
};

static vct_status build_blocked(const uint2* ent, const uint32_t* row_of, uint32_t* key, int64_t n, uint32_t n_keys,
                                int rows, vct_blocked& f, hipStream_t s, RunInfo* run_info = nullptr)
{
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    const size_t nz = (size_t)n;
    const dim3 g1((unsigned)((n + 255) / 256)), g1p((unsigned)((n + 256) / 256)), blk(256);
    uint32_t *vals = nullptr, *key_s = nullptr, *order = nullptr, *head = nullptr, *run = nullptr, *row1 = nullptr;
    uint32_t *cnt = nullptr, *iota2 = nullptr, *row1_s = nullptr, *order2 = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    f = vct_blocked{};
    // 1. stable sort of the entries by block key
    hip(hipMalloc(&vals, 4 * nz + 16), "hipMalloc vals");
    hip(hipMalloc(&key_s, 4 * nz + 16), "hipMalloc key_s");
    hip(hipMalloc(&order, 4 * nz + 16), "hipMalloc order");
    if (st == VCT_OK) hipLaunchKernelGGL(k_iota, g1, blk, 0, s, vals, n);
    const unsigned kb = key_bits(n_keys);
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(nullptr, tb, key, key_s, vals, order, nz, 0u, kb, s), "sort size");
    hip(hipMalloc(&tmp, tb + 16), "hipMalloc sort");
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(tmp, tb, key, key_s, vals, order, nz, 0u, kb, s), "sort");
    if (tmp) { (void)hipStreamSynchronize(s); (void)hipFree(tmp); tmp = nullptr; }
    if (vals) { (void)hipFree(vals); vals = nullptr; }
    // 2. run heads and their exclusive scan = stage-1 row of each run
    hip(hipMalloc(&head, 4 * nz + 16), "hipMalloc head");
    hip(hipMalloc(&run, 4 * nz + 16), "hipMalloc run");
    if (st == VCT_OK) hipLaunchKernelGGL(k_run_heads, g1p, blk, 0, s, key_s, order, row_of, n, head);
    tb = 0;
    if (st == VCT_OK) hip(rocprim::exclusive_scan(nullptr, tb, head, run, 0u, nz + 1, rocprim::plus<uint32_t>(), s), "scan size");
    hip(hipMalloc(&tmp, tb + 16), "hipMalloc scan");
    if (st == VCT_OK) hip(rocprim::exclusive_scan(tmp, tb, head, run, 0u, nz + 1, rocprim::plus<uint32_t>(), s), "scan");
    uint32_t rows1 = 0;
    if (st == VCT_OK) hip(hipMemcpyAsync(&rows1, run + n, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (tmp) { (void)hipFree(tmp); tmp = nullptr; }
    // 3. stage 1: regrouped entries, run starts, row of each run
    hip(hipMalloc(&f.ent1, 8 * nz + 16), "hipMalloc ent1");
    hip(hipMalloc(&f.ptr1, 4 * ((size_t)rows1 + 1)), "hipMalloc ptr1");
    hip(hipMalloc(&row1, 4 * ((size_t)rows1 + 1)), "hipMalloc row1");
    uint32_t* key1 = nullptr;
    if (run_info) hip(hipMalloc(&key1, 4 * ((size_t)rows1 + 1)), "hipMalloc key1");
    if (st == VCT_OK) {
        hipLaunchKernelGGL(k_stage1_fill, g1, blk, 0, s, order, head, run, ent, row_of, n, f.ent1, f.ptr1, row1, key_s,
                           key1);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(1), 0, s, f.ptr1 + rows1, (uint32_t)n);
        hip(hipGetLastError(), "k_stage1_fill");
    }
    void* done1[] = {key_s, order, head, run};
    for (void* b : done1)
        if (b) (void)hipFree(b);
    // 4. stage 2: the runs of each row in key order (stable sort by row)
    const dim3 g2((rows1 + 255) / 256);
    hip(hipMalloc(&cnt, 4 * ((size_t)rows + 1)), "hipMalloc cnt");
    hip(hipMalloc(&iota2, 4 * ((size_t)rows1 + 1)), "hipMalloc iota2");
    hip(hipMalloc(&row1_s, 4 * ((size_t)rows1 + 1)), "hipMalloc row1_s");
    hip(hipMalloc(&order2, 4 * ((size_t)rows1 + 1)), "hipMalloc order2");
    hip(hipMalloc(&f.ptr2, 4 * ((size_t)rows + 1)), "hipMalloc ptr2");
    hip(hipMalloc(&f.ent2, 8 * ((size_t)rows1 + 1)), "hipMalloc ent2");
    if (st == VCT_OK) {
        hipLaunchKernelGGL(k_zero_f32, dim3((rows + 256) / 256), blk, 0, s, (float*)cnt, rows + 1);
        if (rows1) {
            hipLaunchKernelGGL(k_count_rows, g2, blk, 0, s, row1, (int64_t)rows1, cnt);
            hipLaunchKernelGGL(k_iota, g2, blk, 0, s, iota2, (int64_t)rows1);
        }
        hip(hipGetLastError(), "k_count_rows");
    }
    tb = 0;
    const unsigned rb = key_bits((uint64_t)rows);
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(nullptr, tb, row1, row1_s, iota2, order2, (size_t)rows1, 0u, rb, s), "sort size");
    hip(hipMalloc(&tmp, tb + 16), "hipMalloc sort");
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(tmp, tb, row1, row1_s, iota2, order2, (size_t)rows1, 0u, rb, s), "sort");
    if (tmp) { (void)hipStreamSynchronize(s); (void)hipFree(tmp); tmp = nullptr; }
    tb = 0;
    if (st == VCT_OK) hip(rocprim::exclusive_scan(nullptr, tb, cnt, f.ptr2, 0u, (size_t)rows + 1, rocprim::plus<uint32_t>(), s),
        "scan size");
    hip(hipMalloc(&tmp, tb + 16), "hipMalloc scan");
    if (st == VCT_OK) hip(rocprim::exclusive_scan(tmp, tb, cnt, f.ptr2, 0u, (size_t)rows + 1, rocprim::plus<uint32_t>(), s), "scan");
    if (st == VCT_OK && rows1) {
        hipLaunchKernelGGL(k_stage2_fill, g2, blk, 0, s, order2, (int64_t)rows1, f.ent2);
        hip(hipGetLastError(), "k_stage2_fill");
    }
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (run_info && st == VCT_OK) {
        run_info->key = key1;
        run_info->row = row1;
        key1 = row1 = nullptr;
    }
    void* done2[] = {tmp, row1, key1, cnt, iota2, row1_s, order2};
    for (void* b : done2)
        if (b) (void)hipFree(b);
    f.rows1 = rows1;
    if (st != VCT_OK) free_blocked(f);
    return st;
}


static void free_tiled(vct_tiled& t)
{
    void* bufs[] = {t.ell,         t.slice_ptr, t.slice_len, t.slice_runs, t.tile_slices,
                    t.xcd_tiles,   t.slot_off,  t.slot_clo,  t.blk_v0};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    t = vct_tiled{};
}

// What the adjoint LDS blocks are built from: the key and the row (voxel) of
// each stage-1 run of the adjoint blocked form, over 64x64 tiles of an N x N
// image and views of W channels.
struct AdjBlocks {
    const RunInfo* runs;
    int N, W, nvb, ntiles64;
};

// The adjoint LDS blocks on the host: per blocked-form key (view block, tile)
// the coarsest split of the tile's 4x4 cells (1, 2, 4, 8 or 16 parts, halves
// and eighths cut along either axis, the one with the smaller largest table)
// whose ray tables hold at most kAdjSlotCap slots.  lo/hi: first and last
// channel of each (key, cell, view), hi < 0 where the cell has no ray of the
// view.  Fills the block of each (key, cell), and per block the slot offsets,
// first channels and first view.  Returns false if a cell alone does not fit.
struct AdjTables {
    std::vector<uint32_t> cell_block;
    std::vector<int32_t> off, clo, v0;
    int nblocks = 0, slots_max = 0;
};

static bool adj_tables(const std::vector<int32_t>& lo, const std::vector<int32_t>& hi, int nkeys, int ntiles64,
                       AdjTables& T)
{
    // part of cell (cy, cx) under each split, coarsest first; entries of a
    // level are alternatives
    auto part = [](int split, int cy, int cx) {
        switch (split) {
        case 0: return 0;                          // 64x64
        case 1: return cy >> 1;                    // 64 wide x 32 tall
        case 2: return cx >> 1;                    // 32 wide x 64 tall
        case 3: return (cy >> 1) * 2 + (cx >> 1);  // 32x32
        case 4: return cy * 2 + (cx >> 1);         // 32 wide x 16 tall
        case 5: return (cy >> 1) * 4 + cx;         // 16 wide x 32 tall
        default: return cy * 4 + cx;               // 16x16
        }
    };
    const int nparts[7] = {1, 2, 2, 4, 8, 8, 16};
    const int levels[5][2] = {{0, 0}, {1, 2}, {3, 3}, {4, 5}, {6, 6}};
    T = AdjTables{};
    T.cell_block.assign((size_t)nkeys * kAdjCells, 0u);
    std::vector<int32_t> plo(kAdjCells * kAdjViews), phi(kAdjCells * kAdjViews);
    // tables of the parts of one split of key k into plo/phi; returns the largest table
    auto tables = [&](int k, int split) {
        std::fill(plo.begin(), plo.end(), 0x7fffffff);
        std::fill(phi.begin(), phi.end(), -1);
        for (int c = 0; c < kAdjCells; c++) {
            const int q = part(split, c >> 2, c & 3);
            for (int j = 0; j < kAdjViews; j++) {
                const size_t i = ((size_t)k * kAdjCells + c) * kAdjViews + j;
                if (hi[i] < 0) continue;
                plo[q * kAdjViews + j] = std::min(plo[q * kAdjViews + j], lo[i]);
                phi[q * kAdjViews + j] = std::max(phi[q * kAdjViews + j], hi[i]);
            }
        }
        int worst = 0;
        for (int q = 0; q < nparts[split]; q++) {
            int n = kAdjPadSlot + 1;
            for (int j = 0; j < kAdjViews; j++)
                if (phi[q * kAdjViews + j] >= 0) n += phi[q * kAdjViews + j] - plo[q * kAdjViews + j] + 1;
            worst = std::max(worst, n);
        }
        return worst;
    };
    for (int k = 0; k < nkeys; k++) {
        int split = -1;
        for (int l = 0; l < 5 && split < 0; l++) {
            int best = -1, best_worst = 0;
            for (int a = 0; a < 2; a++) {
                const int sp = levels[l][a], w = tables(k, sp);
                if (w <= kAdjSlotCap && (best < 0 || w < best_worst)) {
                    best = sp;
                    best_worst = w;
                }
            }
            split = best;
        }
        if (split < 0) return false;
        tables(k, split);
        for (int c = 0; c < kAdjCells; c++)
            T.cell_block[(size_t)k * kAdjCells + c] = (uint32_t)(T.nblocks + part(split, c >> 2, c & 3));
        for (int q = 0; q < nparts[split]; q++) {
            int32_t o = kAdjPadSlot + 1;
            for (int j = 0; j < kAdjViews; j++) {
                const bool any = phi[q * kAdjViews + j] >= 0;
                T.off.push_back(o);
                T.clo.push_back(any ? plo[q * kAdjViews + j] : 0);
                o += any ? phi[q * kAdjViews + j] - plo[q * kAdjViews + j] + 1 : 0;
            }
            T.off.push_back(o);
            T.v0.push_back((k / ntiles64) * kAdjViews);
            T.slots_max = std::max(T.slots_max, (int)o);
        }
        T.nblocks += nparts[split];
    }
    return true;
}

// LDS-resident form of a stage 1 from its blocked form f: forward (adj ==
// nullptr, block = forward tile of tg) or adjoint (blocks from adj_tables).
// Leaves t empty, and returns VCT_OK, when a run does not fit the form (a
// run longer than 0xfffe entries, more than 65535 blocks, 2^32 interleaved
// entries or, in the adjoint, a 16x16 cell whose ray table exceeds
// kAdjSlotCap slots); HIP failures are errors.
static vct_status build_tiled(const vct_blocked& f, const TileGrid& tg, vct_tiled& t, hipStream_t s,
                              const AdjBlocks* adj = nullptr)
{
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    t = vct_tiled{};
    const uint32_t R = f.rows1;
    if (R == 0 || (!adj && tg.T != 64)) return VCT_OK;
    const dim3 gr((R + 255) / 256), blk(256);
    uint32_t herr = 0;
    // adjoint: channel ranges per (key, cell, view), then the blocks and
    // their ray tables
    AdjTables at;
    uint32_t* cell_block = nullptr;
    if (adj) {
        const int nkeys = adj->nvb * adj->ntiles64;
        const size_t nr = (size_t)nkeys * kAdjCells * kAdjViews;
        uint32_t *first = nullptr, *kbegin = nullptr, *kcount = nullptr, *e = nullptr;
        int32_t *lo = nullptr, *hi = nullptr;
        hip(hipMalloc(&first, 4 * (size_t)nkeys), "hipMalloc key first");
        hip(hipMalloc(&kbegin, 4 * (size_t)nkeys), "hipMalloc key begin");
        hip(hipMalloc(&kcount, 4 * (size_t)nkeys), "hipMalloc key count");
        hip(hipMalloc(&e, 16), "hipMalloc err");
        hip(hipMalloc(&lo, 4 * nr), "hipMalloc channel range");
        hip(hipMalloc(&hi, 4 * nr), "hipMalloc channel range");
        std::vector<uint32_t> hfirst(nkeys), hbegin(nkeys), hcount(nkeys);
        if (st == VCT_OK) {
            hipLaunchKernelGGL(k_zero_f32, dim3(1), blk, 0, s, (float*)e, 1);
            hipLaunchKernelGGL(k_set_u32_n, dim3((nkeys + 255) / 256), blk, 0, s, first, nkeys, 0xffffffffu);
            hipLaunchKernelGGL(k_key_first, gr, blk, 0, s, adj->runs->key, R, (uint32_t)nkeys, first, e);
            hip(hipGetLastError(), "k_key_first");
            hip(hipMemcpyAsync(hfirst.data(), first, 4 * (size_t)nkeys, hipMemcpyDeviceToHost, s), "copy");
            hip(hipMemcpyAsync(&herr, e, 4, hipMemcpyDeviceToHost, s), "copy");
            hip(hipStreamSynchronize(s), "sync");
        }
        uint32_t next = R;
        for (int k = nkeys - 1; k >= 0; k--) {
            hbegin[k] = hfirst[k] == 0xffffffffu ? next : hfirst[k];
            hcount[k] = next - hbegin[k];
            next = hbegin[k];
        }
        bool fits = st == VCT_OK && herr == 0 && next == 0;
        std::vector<int32_t> hlo(fits ? nr : 0), hhi(fits ? nr : 0);
        if (fits) {
            hip(hipMemcpyAsync(kbegin, hbegin.data(), 4 * (size_t)nkeys, hipMemcpyHostToDevice, s), "copy");
            hip(hipMemcpyAsync(kcount, hcount.data(), 4 * (size_t)nkeys, hipMemcpyHostToDevice, s), "copy");
            hipLaunchKernelGGL(k_atile_range, dim3(nkeys), blk, 0, s, f.ptr1, f.ent1, adj->runs->row, kbegin, kcount,
                               adj->N, adj->W, adj->ntiles64, lo, hi, e);
            hip(hipGetLastError(), "k_atile_range");
            hip(hipMemcpyAsync(hlo.data(), lo, 4 * nr, hipMemcpyDeviceToHost, s), "copy");
            hip(hipMemcpyAsync(hhi.data(), hi, 4 * nr, hipMemcpyDeviceToHost, s), "copy");
            hip(hipMemcpyAsync(&herr, e, 4, hipMemcpyDeviceToHost, s), "copy");
            hip(hipStreamSynchronize(s), "sync");
            fits = st == VCT_OK && herr == 0 && adj_tables(hlo, hhi, nkeys, adj->ntiles64, at) && at.nblocks <= 65535;
        }
        if (fits) {
            hip(hipMalloc(&cell_block, 4 * at.cell_block.size()), "hipMalloc cell blocks");
            hip(hipMalloc(&t.slot_off, 4 * at.off.size()), "hipMalloc slot offsets");
            hip(hipMalloc(&t.slot_clo, 4 * at.clo.size()), "hipMalloc slot channels");
            hip(hipMalloc(&t.blk_v0, 4 * at.v0.size()), "hipMalloc block views");
            if (st == VCT_OK) {
                hip(hipMemcpyAsync(cell_block, at.cell_block.data(), 4 * at.cell_block.size(), hipMemcpyHostToDevice, s),
                    "copy");
                hip(hipMemcpyAsync(t.slot_off, at.off.data(), 4 * at.off.size(), hipMemcpyHostToDevice, s), "copy");
                hip(hipMemcpyAsync(t.slot_clo, at.clo.data(), 4 * at.clo.size(), hipMemcpyHostToDevice, s), "copy");
                hip(hipMemcpyAsync(t.blk_v0, at.v0.data(), 4 * at.v0.size(), hipMemcpyHostToDevice, s), "copy");
                hip(hipStreamSynchronize(s), "sync");
            }
            t.slots_max = at.slots_max;
        }
        void* bufs[] = {first, kbegin, kcount, e, lo, hi};
        for (void* b : bufs)
            if (b) (void)hipFree(b);
        if (st != VCT_OK || !fits) {
            if (cell_block) (void)hipFree(cell_block);
            free_tiled(t);
            return st;
        }
    }
    const int nblocks = adj ? at.nblocks : tg.ntiles;
    if (nblocks > 65535) {
        if (cell_block) (void)hipFree(cell_block);
        free_tiled(t);
        return VCT_OK;
    }
    uint32_t *key = nullptr, *key_s = nullptr, *val = nullptr, *order = nullptr, *cnt = nullptr, *err = nullptr;
    uint32_t* tile_begin = nullptr;
    int32_t* slice_tile = nullptr;
    void* tmp = nullptr;
    size_t tb = 0;
    hip(hipMalloc(&key, 4 * (size_t)R), "hipMalloc tile keys");
    hip(hipMalloc(&key_s, 4 * (size_t)R), "hipMalloc tile keys");
    hip(hipMalloc(&val, 4 * (size_t)R), "hipMalloc tile order");
    hip(hipMalloc(&order, 4 * (size_t)R), "hipMalloc tile order");
    hip(hipMalloc(&cnt, 4 * (size_t)nblocks + 16), "hipMalloc tile counts");
    hip(hipMalloc(&err, 16), "hipMalloc tile err");
    if (st == VCT_OK) {
        hipLaunchKernelGGL(k_zero_f32, dim3(1), blk, 0, s, (float*)err, 1);
        if (adj) {
            hipLaunchKernelGGL(k_zero_f32, dim3((nblocks + 255) / 256), blk, 0, s, (float*)cnt, nblocks);
            hipLaunchKernelGGL(k_atile_run_keys, gr, blk, 0, s, f.ptr1, adj->runs->key, adj->runs->row, R, adj->N,
                               cell_block, (uint32_t)nblocks, key, val, cnt, err);
        } else {
            hipLaunchKernelGGL(k_set_u32_n, dim3((nblocks + 255) / 256), blk, 0, s, cnt, nblocks, 0xffffffffu);
            hipLaunchKernelGGL(k_tile_run_keys, gr, blk, 0, s, f.ptr1, f.ent1, R, tg, key, val, cnt, err);
            hipLaunchKernelGGL(k_first_to_count, dim3(1), dim3(1), 0, s, cnt, nblocks, R);
        }
        hip(hipGetLastError(), "tile run keys");
    }
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(nullptr, tb, key, key_s, val, order, (size_t)R, 0u, 32u, s), "sort size");
    hip(hipMalloc(&tmp, tb + 16), "hipMalloc sort");
    if (st == VCT_OK) hip(rocprim::radix_sort_pairs(tmp, tb, key, key_s, val, order, (size_t)R, 0u, 32u, s), "sort");
    // host: block begins (in sorted order), slices per block, slice -> block
    std::vector<uint32_t> n_t(nblocks);
    herr = 0;
    if (st == VCT_OK) hip(hipMemcpyAsync(n_t.data(), cnt, 4 * (size_t)nblocks, hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    std::vector<uint32_t> begin(nblocks);
    std::vector<int32_t> tslices(nblocks + 1, 0), stile;
    uint32_t acc = 0;
    for (int k = 0; k < nblocks; k++) {
        begin[k] = acc;
        acc += n_t[k];
        tslices[k + 1] = tslices[k] + (int32_t)((n_t[k] + 63) / 64);
        for (uint32_t q = 0; q < (n_t[k] + 63) / 64; q++) stile.push_back(k);
    }
    const int nslices = tslices[nblocks];
    bool ok = st == VCT_OK && herr == 0 && acc == R && nslices > 0;
    if (ok) {
        hip(hipMalloc(&tile_begin, 4 * (size_t)nblocks), "hipMalloc tile begin");
        hip(hipMalloc(&t.tile_slices, 4 * ((size_t)nblocks + 1)), "hipMalloc tile slices");
        hip(hipMalloc(&slice_tile, 4 * (size_t)nslices), "hipMalloc slice tile");
        hip(hipMalloc(&t.slice_len, 4 * (size_t)nslices), "hipMalloc slice len");
        hip(hipMalloc(&t.slice_ptr, 4 * (size_t)nslices), "hipMalloc slice ptr");
        hip(hipMalloc(&t.slice_runs, 4 * 64 * (size_t)nslices), "hipMalloc slice runs");
        if (st == VCT_OK) {
            hip(hipMemcpyAsync(tile_begin, begin.data(), 4 * (size_t)nblocks, hipMemcpyHostToDevice, s), "copy");
            hip(hipMemcpyAsync(t.tile_slices, tslices.data(), 4 * ((size_t)nblocks + 1), hipMemcpyHostToDevice, s),
                "copy");
            hip(hipMemcpyAsync(slice_tile, stile.data(), 4 * (size_t)nslices, hipMemcpyHostToDevice, s), "copy");
            hipLaunchKernelGGL(k_tile_slice_len, dim3((nslices + 255) / 256), blk, 0, s, f.ptr1, order, slice_tile,
                               tile_begin, t.tile_slices, nslices, t.slice_len);
            hip(hipGetLastError(), "k_tile_slice_len");
        }
    }
    // host: slice offsets, per-block work and the XCD block lists
    std::vector<uint32_t> slen(ok ? nslices : 0), sptr(ok ? nslices : 0);
    if (ok && st == VCT_OK) {
        hip(hipMemcpyAsync(slen.data(), t.slice_len, 4 * (size_t)nslices, hipMemcpyDeviceToHost, s), "copy");
        hip(hipStreamSynchronize(s), "sync");
    }
    uint64_t total = 0;
    std::vector<uint64_t> work(nblocks, 0);
    if (ok && st == VCT_OK) {
        for (int sl = 0; sl < nslices; sl++) {
            // adjoint slices hold a multiple of kTileUnroll entries per lane,
            // so the kernel reads them without bounds tests (the forward's
            // short runs near tile edges would pay more padding than it saves)
            if (adj) slen[sl] = (slen[sl] + kTileUnroll - 1) / kTileUnroll * kTileUnroll;
            sptr[sl] = (uint32_t)total;
            total += 64ull * slen[sl];
            work[stile[sl]] += 64ull * slen[sl];
        }
        ok = total < (1ull << 32);
    }
    if (ok && st == VCT_OK) {
        std::vector<std::vector<int32_t>> lists(kXcds);
        if (adj) {
            // consecutive blocks (the quarters of a tile, then the next tile
            // along the rays) share most of their rays: each XCD takes one
            // contiguous run of blocks of about an eighth of the work
            uint64_t run = 0;
            for (int k = 0; k < nblocks; k++) {
                if (n_t[k] == 0) continue;
                const int x = (int)std::min<uint64_t>(kXcds - 1, (run + work[k] / 2) * kXcds / total);
                lists[x].push_back(k);
                run += work[k];
            }
        } else {
            // largest tile first onto the least loaded XCD
            std::vector<int> idx(nblocks);
            for (int k = 0; k < nblocks; k++) idx[k] = k;
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return work[x] > work[y]; });
            std::vector<uint64_t> load(kXcds, 0);
            for (int k : idx) {
                if (n_t[k] == 0) continue;
                int best = 0;
                for (int x = 1; x < kXcds; x++)
                    if (load[x] < load[best]) best = x;
                lists[best].push_back(k);
                load[best] += work[k];
            }
        }
        int kmax = 0;
        for (auto& l : lists) kmax = std::max(kmax, (int)l.size());
        std::vector<int32_t> xt((size_t)kXcds * kmax, -1);
        for (int x = 0; x < kXcds; x++)
            for (size_t q = 0; q < lists[x].size(); q++) xt[(size_t)x * kmax + q] = lists[x][q];
        t.ntiles = nblocks;
        t.ntx = tg.ntx;
        t.nslices = nslices;
        t.kmax = kmax;
        t.n_ell = (size_t)total;
        hip(hipMalloc(&t.xcd_tiles, 4 * xt.size()), "hipMalloc xcd tiles");
        const size_t n_fill = (size_t)total + 64 * kTileUnroll;  // slack for the adjoint's last prefetch
        hip(hipMalloc(&t.ell, 8 * n_fill), "hipMalloc tile entries");
        if (st == VCT_OK) {
            hip(hipMemcpyAsync(t.xcd_tiles, xt.data(), 4 * xt.size(), hipMemcpyHostToDevice, s), "copy");
            hip(hipMemcpyAsync(t.slice_ptr, sptr.data(), 4 * (size_t)nslices, hipMemcpyHostToDevice, s), "copy");
            hip(hipMemcpyAsync(t.slice_len, slen.data(), 4 * (size_t)nslices, hipMemcpyHostToDevice, s), "copy");
            hipLaunchKernelGGL(k_fill_pad, dim3((unsigned)((n_fill + 255) / 256)), blk, 0, s, t.ell, n_fill,
                               adj ? (uint32_t)kAdjPadSlot : kTilePadVoxel);
            hip(hipMemsetAsync(t.slice_runs, 0xff, 4 * 64 * (size_t)nslices, s), "memset");
            hipLaunchKernelGGL(k_zero_f32, dim3(1), blk, 0, s, (float*)err, 1);
            if (adj)
                hipLaunchKernelGGL(k_atile_fill, gr, blk, 0, s, f.ptr1, f.ent1, key_s, order, R, tile_begin,
                                   t.tile_slices, t.slice_ptr, t.slot_off, t.slot_clo, t.blk_v0, adj->W, t.ell,
                                   t.slice_runs, err);
            else
                hipLaunchKernelGGL(k_tile_fill, gr, blk, 0, s, f.ptr1, f.ent1, key_s, order, R, tile_begin,
                                   t.tile_slices, t.slice_ptr, tg, t.ell, t.slice_runs, err);
            hip(hipGetLastError(), "tile fill");
            hip(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s), "copy");
            hip(hipStreamSynchronize(s), "sync");
            ok = herr == 0;
        }
    }
    void* bufs[] = {key, key_s, val, order, cnt, err, tile_begin, slice_tile, cell_block, tmp};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (st != VCT_OK || !ok || !t.ell) free_tiled(t);
    return st;
}

extern "C" {

vct_status vct_sysmat_build(vct_plan* p, int flags, vct_sysmat** out, void* stream)
{
    clear_error();
    if (!p || !out) return set_error(VCT_ERR_ARG, "NULL argument");
    if (flags & ~(VCT_SYSMAT_TRANSPOSE | VCT_SYSMAT_BLOCKED | VCT_SYSMAT_NO_TILE))
        return set_error(VCT_ERR_ARG, "unknown flags %#x", flags);
    const bool transpose = (flags & VCT_SYSMAT_TRANSPOSE) != 0;
    const bool blocked = (flags & VCT_SYSMAT_BLOCKED) != 0;
    const bool tiled = blocked && !(flags & VCT_SYSMAT_NO_TILE);
    *out = nullptr;
    hipStream_t s = (hipStream_t)stream;
    const vct_geometry& g = p->g;
    const int nrays = g.n_views * g.n_channels;
    const int NN = g.recon_size * g.recon_size;
    vct_sysmat* m = new vct_sysmat();
    m->plan = p;
    m->scale = (float)(1.0 / p->g.xfactor);
    m->nrays = nrays;
    m->NN = NN;
    uint32_t* counts = nullptr;
    uint32_t *keys = nullptr, *vals = nullptr, *keys2 = nullptr, *vals2 = nullptr, *eray = nullptr, *colcnt = nullptr;
    void* tmp = nullptr;
    vct_status st = VCT_OK;
    auto hip = [&](hipError_t e, const char* what) {
        if (st == VCT_OK && e != hipSuccess) st = set_error(VCT_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    };
    TraceArgs a = make_trace_args(p);
    hip(hipMalloc(&counts, sizeof(uint32_t) * (nrays + 1)), "hipMalloc counts");
    hip(hipMalloc(&m->row_ptr, sizeof(uint32_t) * (nrays + 1)), "hipMalloc row_ptr");
    if (st == VCT_OK) {
        hipLaunchKernelGGL(k_sm_count, dim3((nrays + 255) / 256), dim3(256), 0, s, a, counts);
        hip(hipGetLastError(), "k_sm_count");
        hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(1), 0, s, (float*)(counts + nrays), 1);
        size_t tb = 0;
        if (st == VCT_OK) hip(rocprim::exclusive_scan(nullptr, tb, counts, m->row_ptr, 0u, (size_t)nrays + 1, rocprim::plus<uint32_t>(), s),
            "scan size");
        hip(hipMalloc(&tmp, tb + 16), "hipMalloc scan");
        if (st == VCT_OK) hip(rocprim::exclusive_scan(tmp, tb, counts, m->row_ptr, 0u, (size_t)nrays + 1, rocprim::plus<uint32_t>(), s),
            "scan");
    }
    uint32_t total = 0;
    if (st == VCT_OK) hip(hipMemcpyAsync(&total, m->row_ptr + nrays, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (tmp) { (void)hipFree(tmp); tmp = nullptr; }
    m->nseg = total;
    if (st == VCT_OK) hip(hipMalloc(&m->csr, sizeof(uint2) * (size_t)total + 8), "hipMalloc csr");
    if (st == VCT_OK && (transpose || blocked))
        hip(hipMalloc(&eray, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc eray");
    if (st == VCT_OK && transpose) {
        hip(hipMalloc(&keys, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc keys");
        hip(hipMalloc(&vals, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc vals");
        hip(hipMalloc(&colcnt, sizeof(uint32_t) * (NN + 1)), "hipMalloc colcnt");
        if (st == VCT_OK)
            hipLaunchKernelGGL(k_zero_f32, dim3((NN + 256) / 256), dim3(256), 0, s, (float*)colcnt, NN + 1);
    }
    if (st == VCT_OK) {
        hipLaunchKernelGGL(k_sm_fill, dim3((nrays + 255) / 256), dim3(256), 0, s, a, m->row_ptr, m->csr, keys, vals, eray,
                           colcnt);
        hip(hipGetLastError(), "k_sm_fill");
    }
    if (st == VCT_OK && transpose) {
        hip(hipMalloc(&keys2, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc keys2");
        hip(hipMalloc(&vals2, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc vals2");
        unsigned end_bit = 1;
        while ((1u << end_bit) < (unsigned)NN) end_bit++;
        size_t tb = 0;
        if (st == VCT_OK) hip(rocprim::radix_sort_pairs(nullptr, tb, keys, keys2, vals, vals2, (size_t)total, 0u, end_bit, s), "sort size");
        hip(hipMalloc(&tmp, tb + 16), "hipMalloc sort");
        if (st == VCT_OK) hip(rocprim::radix_sort_pairs(tmp, tb, keys, keys2, vals, vals2, (size_t)total, 0u, end_bit, s), "sort");
        if (tmp) { (void)hipStreamSynchronize(s); (void)hipFree(tmp); tmp = nullptr; }
        hip(hipMalloc(&m->col_ptr, sizeof(uint32_t) * (NN + 1)), "hipMalloc col_ptr");
        tb = 0;
        if (st == VCT_OK) hip(rocprim::exclusive_scan(nullptr, tb, colcnt, m->col_ptr, 0u, (size_t)NN + 1, rocprim::plus<uint32_t>(), s),
            "scan size");
        hip(hipMalloc(&tmp, tb + 16), "hipMalloc scan");
        if (st == VCT_OK) hip(rocprim::exclusive_scan(tmp, tb, colcnt, m->col_ptr, 0u, (size_t)NN + 1, rocprim::plus<uint32_t>(), s), "scan");
        hip(hipMalloc(&m->csc, sizeof(uint2) * (size_t)total + 8), "hipMalloc csc");
        if (st == VCT_OK) {
            hipLaunchKernelGGL(k_csc_gather, dim3((unsigned)(((int64_t)total + 255) / 256)), dim3(256), 0, s, vals2, eray,
                               m->csr, m->csc, (int64_t)total);
            hip(hipGetLastError(), "k_csc_gather");
        }
    }
    if (st == VCT_OK && blocked && total > 0) {
        // keys (the per-entry voxel of the ray-major order) is free again and
        // holds the block keys; keys2 holds the voxel of each CSC entry
        const TileGrid tg = tile_grid(g.recon_size, kBlockTile), ta = tile_grid(g.recon_size, kAdjTile);
        const dim3 ge((unsigned)(((int64_t)total + 255) / 256));
        uint32_t* bkey = keys;
        if (!bkey) hip(hipMalloc(&bkey, sizeof(uint32_t) * (size_t)total + 8), "hipMalloc block keys");
        if (st == VCT_OK) {
            hipLaunchKernelGGL(k_keys_fwd, ge, dim3(256), 0, s, m->csr, (int64_t)total, tg, bkey);
            hip(hipGetLastError(), "k_keys_fwd");
        }
        if (st == VCT_OK) st = build_blocked(m->csr, eray, bkey, total, (uint32_t)tg.ntiles, nrays, m->fwd, s);
        if (st == VCT_OK && tiled) st = build_tiled(m->fwd, tg, m->tile, s);
        if (st == VCT_OK && transpose) {
            const int nvb = (g.n_views + kAdjViews - 1) / kAdjViews;
            hipLaunchKernelGGL(k_keys_adj, ge, dim3(256), 0, s, m->csc, keys2, (int64_t)total, ta, g.n_channels,
                               g.n_views, a.cosT, a.sinT, bkey);
            hip(hipGetLastError(), "k_keys_adj");
            RunInfo ri;
            if (st == VCT_OK)
                st = build_blocked(m->csc, keys2, bkey, total, (uint32_t)(nvb * ta.ntiles), NN, m->adj, s,
                                   tiled ? &ri : nullptr);
            if (st == VCT_OK && tiled) {
                const AdjBlocks ab{&ri, g.recon_size, g.n_channels, nvb, ta.ntiles};
                st = build_tiled(m->adj, ta, m->atile, s, &ab);
            }
            if (ri.key) (void)hipFree(ri.key);
            if (ri.row) (void)hipFree(ri.row);
        }
        if (!keys && bkey) (void)hipFree(bkey);
    }
    int herr = 0;
    if (st == VCT_OK) hip(hipMemcpyAsync(&herr, p->ctx->d_err, sizeof(int), hipMemcpyDeviceToHost, s), "copy");
    if (st == VCT_OK) hip(hipStreamSynchronize(s), "sync");
    if (st == VCT_OK && herr) st = set_error(VCT_ERR_INTERNAL, "system matrix build reported error %d", herr);
    void* bufs[] = {counts, keys, vals, keys2, vals2, eray, colcnt, tmp};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (st != VCT_OK) {
        vct_sysmat_destroy(m);
        return st;
    }
    *out = m;
    return VCT_OK;
}

vct_status vct_sysmat_destroy(vct_sysmat* m)
{
    if (!m) return VCT_OK;
    void* bufs[] = {m->row_ptr, m->csr, m->col_ptr, m->csc};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    free_blocked(m->fwd);
    free_blocked(m->adj);
    free_tiled(m->tile);
    free_tiled(m->atile);
    delete m;
    return VCT_OK;
}

int64_t vct_sysmat_segments(const vct_sysmat* m) { return m ? m->nseg : -1; }

vct_status vct_sysmat_row(const vct_sysmat* m, int ray, uint32_t* h_voxel, float* h_length, int capacity, int* n_out,
                          void* stream)
{
    clear_error();
    if (!m || !n_out) return set_error(VCT_ERR_ARG, "NULL argument");
    if (ray < 0 || ray >= m->nrays) return set_error(VCT_ERR_ARG, "ray %d outside 0..%d", ray, m->nrays - 1);
    hipStream_t s = (hipStream_t)stream;
    uint32_t be[2] = {0u, 0u};
    VCT_HIP(hipMemcpyAsync(be, m->row_ptr + ray, sizeof(be), hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    const int n = (int)(be[1] - be[0]);
    *n_out = n;
    if (n > capacity) return set_error(VCT_ERR_ARG, "row %d holds %d segments, capacity %d", ray, n, capacity);
    if (n == 0) return VCT_OK;
    if (!h_voxel || !h_length) return set_error(VCT_ERR_ARG, "NULL output");
    std::vector<uint2> e(n);
    VCT_HIP(hipMemcpyAsync(e.data(), m->csr + be[0], sizeof(uint2) * n, hipMemcpyDeviceToHost, s));
    VCT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) {
        h_voxel[i] = e[i].x;
        uint32_t bits = e[i].y;
        std::memcpy(&h_length[i], &bits, sizeof(float));
    }
    return VCT_OK;
}

size_t vct_sysmat_bytes(const vct_sysmat* m)
{
    if (!m) return 0;
    size_t b = sizeof(uint32_t) * (m->nrays + 1) + sizeof(uint2) * (size_t)m->nseg;
    if (m->csc) b += sizeof(uint32_t) * (m->NN + 1) + sizeof(uint2) * (size_t)m->nseg;
    const vct_blocked* fs[] = {&m->fwd, &m->adj};
    const int rows[] = {m->nrays, m->NN};
    for (int k = 0; k < 2; k++)
        if (fs[k]->rows1)
            b += sizeof(uint2) * (size_t)m->nseg + sizeof(uint32_t) * ((size_t)fs[k]->rows1 + 1) +
                 sizeof(uint32_t) * ((size_t)rows[k] + 1) + sizeof(uint2) * ((size_t)fs[k]->rows1 + 1);
    for (const vct_tiled* t : {&m->tile, &m->atile})
        if (t->ell)
            b += sizeof(uint2) * t->n_ell + sizeof(uint32_t) * 66 * (size_t)t->nslices +
                 sizeof(int32_t) * ((size_t)t->ntiles + 1 + (size_t)kXcds * t->kmax) +
                 (t->slot_off ? sizeof(int32_t) * (2 * kAdjViews + 1) * (size_t)t->ntiles : 0);
    return b;
}

// Workspace layout: the stage-1 partial sums (rows of 8 * B floats), then,
// with the adjoint's LDS-resident form, the phantom-major copy of the
// adjoint weights (k_weights_pm) at a 256-byte boundary.
static size_t pm_offset(const vct_sysmat* m, int B)
{
    const size_t rows1 = m->fwd.rows1 > m->adj.rows1 ? m->fwd.rows1 : m->adj.rows1;
    return (rows1 * (size_t)B * 8 * sizeof(float) + 255) / 256 * 256;
}

size_t vct_sysmat_workspace(const vct_sysmat* m, int n_phantoms)
{
    if (!m || n_phantoms < 1) return 0;
    const size_t pm = m->atile.ell ? (size_t)m->nrays * n_phantoms * 8 * sizeof(float) : 0;
    return pm_offset(m, n_phantoms) + pm;
}

// lanes_per_row 1: one phantom per group, each row split over kSplit lanes
constexpr int kSplit = 16;

static void add_group(GatherArgs& a, int& items, const uint32_t* ptr, const uint2* ent, int rows, const float* in,
                      float* out, int B, float scale, int partial = 0)
{
    GatherGroup& d = a.g[a.n_groups++];
    d.partial = partial;
    d.ptr = ptr;
    d.ent = ent;
    d.rows = rows;
    d.in = in;
    d.out = out;
    d.B = B;
    d.scale = scale;
    d.item_begin = items;
    items += (rows + a.rows_per_item - 1) / a.rows_per_item;
}

static vct_status run_gather(GatherArgs& a, int items, int M, int lanes, int blocks, hipStream_t s)
{
    a.n_items = items;
    if (items == 0) return VCT_OK;
    if (blocks <= 0) blocks = 1024;  // 4 workgroups per CU of MI355X
    blocks = blocks < items ? blocks : items;
    if (blocks >= kXcds) blocks -= blocks % kXcds;
#define VCT_GATHER_CASE(MM, LL) \
    if (M == MM && lanes == LL) hipLaunchKernelGGL((k_gather<MM, LL>), dim3(blocks), dim3(256), 0, s, a);
#define VCT_GATHER_M(MM)                                                                                   \
    if (M == MM && lanes == 1) hipLaunchKernelGGL((k_gather_split<MM, kSplit>), dim3(blocks), dim3(256), 0, s, a); \
    VCT_GATHER_CASE(MM, 8) VCT_GATHER_CASE(MM, 16) VCT_GATHER_CASE(MM, 32) VCT_GATHER_CASE(MM, 64)
    VCT_GATHER_M(1) VCT_GATHER_M(2) VCT_GATHER_M(3) VCT_GATHER_M(4)
    VCT_GATHER_M(5) VCT_GATHER_M(6) VCT_GATHER_M(7) VCT_GATHER_M(8)
#undef VCT_GATHER_M
#undef VCT_GATHER_CASE
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

// The LDS-resident stage 1 of a direction, if built.
static const vct_tiled* lds_form(const vct_sysmat* m, bool adjoint)
{
    const vct_tiled& t = adjoint ? m->atile : m->tile;
    if (!t.ell || (adjoint && t.slots_max > kAdjSlotCap)) return nullptr;
    return &t;
}

// Stage 1 of the groups over their LDS-resident forms in one launch: per
// group B workgroups per block, kmax blocks per XCD.
static vct_status run_gather_tile(const vct_batch_group* groups, const int* idx, const vct_tiled* const* forms, int n,
                                  bool adjoint, int M, hipStream_t s)
{
    static_assert(sizeof(TileGroups) <= 4096, "kernel arguments");
    TileGroups tg{};
    tg.n = n;
    long long total = 0;
    for (int q = 0; q < n; q++) {
        const vct_batch_group& G = groups[idx[q]];
        const vct_tiled& t = *forms[q];
        float* pm = nullptr;
        if (adjoint) {
            // the weights phantom-major, after the partial sums of the workspace
            pm = reinterpret_cast<float*>(reinterpret_cast<char*>(G.d_work) + pm_offset(G.sysmat, G.n_phantoms));
            int lpr = 1;
            while (lpr < G.n_phantoms) lpr <<= 1;
            const int rays = G.sysmat->nrays, per = 256 / lpr;
            const dim3 grid((rays + per - 1) / per);
#define VCT_PM_M(MM) \
    if (M == MM) hipLaunchKernelGGL(k_weights_pm<MM>, grid, dim3(256), 0, s, G.d_in, rays, G.n_phantoms, lpr, pm);
            VCT_PM_M(1) VCT_PM_M(2) VCT_PM_M(3) VCT_PM_M(4)
            VCT_PM_M(5) VCT_PM_M(6) VCT_PM_M(7) VCT_PM_M(8)
#undef VCT_PM_M
        }
        tg.a[q] = TileArgs{G.d_in, G.d_work, t.ell, t.slice_ptr, t.slice_len, t.slice_runs, t.tile_slices,
                           t.xcd_tiles, t.kmax, t.ntx, G.sysmat->plan->g.recon_size, G.n_phantoms, t.slot_off,
                           t.slot_clo, t.blk_v0, G.sysmat->plan->g.n_channels, pm, G.sysmat->nrays};
        tg.pref[q] = (int)total;
        total += (long long)kXcds * t.kmax * G.n_phantoms;  // a multiple of kXcds
        if (total > INT_MAX) return set_error(VCT_ERR_ARG, "stage-1 launch too large");
    }
    tg.pref[n] = (int)total;
    if (total == 0) return VCT_OK;
    const unsigned grid = (unsigned)total;
#define VCT_TILE_M(MM)                                                                                      \
    if (M == MM && !adjoint) hipLaunchKernelGGL((k_gather_tile<MM, false>), dim3(grid), dim3(kTileThreads), 0, s, tg); \
    if (M == MM && adjoint) hipLaunchKernelGGL((k_gather_tile<MM, true>), dim3(grid), dim3(kAtileThreads), 0, s, tg);
    VCT_TILE_M(1) VCT_TILE_M(2) VCT_TILE_M(3) VCT_TILE_M(4)
    VCT_TILE_M(5) VCT_TILE_M(6) VCT_TILE_M(7) VCT_TILE_M(8)
#undef VCT_TILE_M
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

// Groups with a workspace and a blocked form of the direction run in two
// launches (stage 1 into the workspace, then stage 2); the others gather
// directly in the first launch.
static vct_status launch_gather(const vct_batch_group* groups, int n_groups, int M, int lanes, int blocks,
                                bool adjoint, void* stream)
{
    if (!groups || n_groups < 1 || n_groups > kMaxGroups)
        return set_error(VCT_ERR_ARG, "n_groups %d outside 1..%d", n_groups, kMaxGroups);
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    if (lanes != 1 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
        return set_error(VCT_ERR_ARG, "lanes_per_row %d not 1/8/16/32/64", lanes);
    GatherArgs a{}, b{};
    a.rows_per_item = b.rows_per_item = 256 / (lanes == 1 ? kSplit : lanes);
    int items_a = 0, items_b = 0;
    int tiles[kMaxGroups], n_tiles = 0;
    const vct_tiled* tforms[kMaxGroups];
    for (int k = 0; k < n_groups; k++) {
        const vct_batch_group& G = groups[k];
        if (!G.sysmat || !G.d_in || !G.d_out) return set_error(VCT_ERR_ARG, "group %d has NULL pointers", k);
        if (G.n_phantoms < 1 || G.n_phantoms > lanes)
            return set_error(VCT_ERR_ARG, "group %d: %d phantoms with %d lanes per row", k, G.n_phantoms, lanes);
        const vct_sysmat* m = G.sysmat;
        if (adjoint && !m->csc) return set_error(VCT_ERR_ARG, "group %d: system matrix built without transpose", k);
        const vct_blocked& f = adjoint ? m->adj : m->fwd;
        const int rows = adjoint ? m->NN : m->nrays;
        if (G.d_work && f.rows1) {
            if (const vct_tiled* t = lds_form(m, adjoint)) {
                tforms[n_tiles] = t;
                tiles[n_tiles++] = k;  // stage 1 by k_gather_tile
            } else {
                add_group(a, items_a, f.ptr1, f.ent1, (int)f.rows1, G.d_in, G.d_work, G.n_phantoms, 1.f, 1);
                a.chunked = 1;
            }
            add_group(b, items_b, f.ptr2, f.ent2, rows, G.d_work, G.d_out, G.n_phantoms, m->scale);
        } else {
            add_group(a, items_a, adjoint ? m->col_ptr : m->row_ptr, adjoint ? m->csc : m->csr, rows, G.d_in, G.d_out,
                      G.n_phantoms, m->scale);
        }
    }
    hipStream_t s = (hipStream_t)stream;
    if (n_tiles) VCT_CHECK(run_gather_tile(groups, tiles, tforms, n_tiles, adjoint, M, s));
    VCT_CHECK(run_gather(a, items_a, M, lanes, blocks, s));
    return run_gather(b, items_b, M, lanes, blocks, s);
}

vct_status vct_batch_paths(const vct_batch_group* groups, int n_groups, int M, int lanes_per_row, int blocks,
                           void* stream)
{
    clear_error();
    return launch_gather(groups, n_groups, M, lanes_per_row, blocks, false, stream);
}

vct_status vct_batch_adjoint(const vct_batch_group* groups, int n_groups, int M, int lanes_per_row, int blocks,
                             void* stream)
{
    clear_error();
    return launch_gather(groups, n_groups, M, lanes_per_row, blocks, true, stream);
}

size_t vct_batch_bb_workspace(int n_phantoms)
{
    if (n_phantoms < 1 || n_phantoms > 64) return 0;
    return (size_t)kBBBlocksPerPhantom * n_phantoms * n_phantoms * 2 * sizeof(double);
}

vct_status vct_batch_bb_step(float* d_x, float* d_x_prev, float* d_g_prev, const float* d_g, float* d_alpha,
                             int n_vox, int n_phantoms, int first, double* d_work, void* stream)
{
    clear_error();
    if (!d_x || !d_x_prev || !d_g_prev || !d_g || !d_alpha || !d_work) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_vox < 1 || n_phantoms < 1 || n_phantoms > 64)
        return set_error(VCT_ERR_ARG, "n_vox %d / n_phantoms %d invalid (1..64 phantoms)", n_vox, n_phantoms);
    BBArgs a{d_x, d_x_prev, d_g_prev, d_g, d_alpha, d_work, (int64_t)n_vox * n_phantoms, n_phantoms, first ? 1 : 0};
    const int blocks = kBBBlocksPerPhantom * n_phantoms;
    hipStream_t s = (hipStream_t)stream;
    if (!a.first) {
        hipLaunchKernelGGL(k_bb_partial, dim3(blocks), dim3(256), 0, s, a);
        hipLaunchKernelGGL(k_bb_alpha, dim3(1), dim3(64), 0, s, a, blocks);
    }
    hipLaunchKernelGGL(k_bb_apply, dim3(blocks), dim3(256), 0, s, a);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

static vct_status check_spec(vct_plan* p, const float* L, int B, int M, const float* att, const float* pdf, int NE,
                             int S)
{
    if (!p || !L || !att || !pdf) return set_error(VCT_ERR_ARG, "NULL argument");
    if (M < 1 || M > 8) return set_error(VCT_ERR_ARG, "M=%d outside 1..8", M);
    if (NE < 1 || NE > kMaxBins) return set_error(VCT_ERR_ARG, "NE=%d outside 1..%d", NE, kMaxBins);
    if (B < 1 || B > 65535 || S < 1 || S > 64) return set_error(VCT_ERR_ARG, "B=%d / S=%d invalid", B, S);
    return VCT_OK;
}

vct_status vct_batch_spectral(vct_plan* p, const float* d_L, int B, int M, const float* d_att, const float* d_pdf,
                              int NE, int S, float* d_raw, uint32_t* d_stats, void* stream)
{
    clear_error();
    VCT_CHECK(check_spec(p, d_L, B, M, d_att, d_pdf, NE, S));
    if (!d_raw || !d_stats) return set_error(VCT_ERR_ARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    BatchSpecArgs a{};
    a.L = d_L; a.att = d_att; a.pdf = d_pdf; a.raw = d_raw; a.stats = d_stats;
    a.rays = p->g.n_views * p->g.n_channels; a.W = p->g.n_channels; a.VT = p->VT;
    a.B = B; a.S = S; a.M = M; a.NE = NE;
    hipLaunchKernelGGL(k_batch_stats_init, dim3((B * S + 63) / 64), dim3(64), 0, s, d_stats, B * S);
    // all spectra of a phantom per block while their tables fit in 64 KB of
    // LDS (two blocks per CU); else one (phantom, spectrum) per block
    const size_t lds_all = sizeof(float) * (size_t)S * NE * 9 + sizeof(int) * kSpecAllWaves * S;
    if (lds_all <= 64 * 1024) {
        const dim3 grid((a.rays + kSpecAllBlock - 1) / kSpecAllBlock, 1, B);
        switch (M) {
#define VCT_SPEC_CASE(MM) \
    case MM: hipLaunchKernelGGL((k_batch_spectral_all<MM>), grid, dim3(kSpecAllBlock), lds_all, s, a); break;
            VCT_SPEC_CASE(1) VCT_SPEC_CASE(2) VCT_SPEC_CASE(3) VCT_SPEC_CASE(4)
            VCT_SPEC_CASE(5) VCT_SPEC_CASE(6) VCT_SPEC_CASE(7) VCT_SPEC_CASE(8)
#undef VCT_SPEC_CASE
        }
    } else {
        const dim3 grid((a.rays + 255) / 256, S, B);
        const size_t lds = sizeof(float) * (size_t)NE * 9;
        switch (M) {
#define VCT_SPEC_CASE(MM) \
    case MM: hipLaunchKernelGGL((k_batch_spectral<MM>), grid, dim3(256), lds, s, a); break;
            VCT_SPEC_CASE(1) VCT_SPEC_CASE(2) VCT_SPEC_CASE(3) VCT_SPEC_CASE(4)
            VCT_SPEC_CASE(5) VCT_SPEC_CASE(6) VCT_SPEC_CASE(7) VCT_SPEC_CASE(8)
#undef VCT_SPEC_CASE
        }
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_batch_spectral_groups(vct_plan* p, const float* const* d_L, const int* n_phantoms, int n_groups, int M,
                                     const float* d_att, const float* d_pdf, int NE, int S, float* d_raw,
                                     uint32_t* d_stats, void* stream)
{
    clear_error();
    if (!p || !d_L || !n_phantoms || !d_raw || !d_stats) return set_error(VCT_ERR_ARG, "NULL argument");
    if (n_groups < 1 || n_groups > kBatchSpecGroups)
        return set_error(VCT_ERR_ARG, "n_groups %d outside 1..%d", n_groups, kBatchSpecGroups);
    BatchSpecArgs a{};
    a.ngroups = n_groups;
    a.pref[0] = 0;
    for (int g = 0; g < n_groups; g++) {
        VCT_CHECK(check_spec(p, d_L[g], n_phantoms[g], M, d_att, d_pdf, NE, S));
        a.Lg[g] = d_L[g];
        a.Bg[g] = n_phantoms[g];
        a.pref[g + 1] = a.pref[g] + n_phantoms[g];
    }
    const int B = a.pref[n_groups];
    if (B > 65535) return set_error(VCT_ERR_ARG, "%d phantoms in one launch", B);
    hipStream_t s = (hipStream_t)stream;
    a.att = d_att; a.pdf = d_pdf; a.raw = d_raw; a.stats = d_stats;
    a.rays = p->g.n_views * p->g.n_channels; a.W = p->g.n_channels; a.VT = p->VT;
    a.B = B; a.S = S; a.M = M; a.NE = NE;
    hipLaunchKernelGGL(k_batch_stats_init, dim3((B * S + 63) / 64), dim3(64), 0, s, d_stats, B * S);
    const size_t lds_all = sizeof(float) * (size_t)S * NE * 9 + sizeof(int) * kSpecAllWaves * S;
    if (lds_all <= 64 * 1024) {
        const dim3 grid((a.rays + kSpecAllBlock - 1) / kSpecAllBlock, 1, B);
        switch (M) {
#define VCT_SPEC_CASE(MM) \
    case MM: hipLaunchKernelGGL((k_batch_spectral_all<MM>), grid, dim3(kSpecAllBlock), lds_all, s, a); break;
            VCT_SPEC_CASE(1) VCT_SPEC_CASE(2) VCT_SPEC_CASE(3) VCT_SPEC_CASE(4)
            VCT_SPEC_CASE(5) VCT_SPEC_CASE(6) VCT_SPEC_CASE(7) VCT_SPEC_CASE(8)
#undef VCT_SPEC_CASE
        }
    } else {
        const dim3 grid((a.rays + 255) / 256, S, B);
        const size_t lds = sizeof(float) * (size_t)NE * 9;
        switch (M) {
#define VCT_SPEC_CASE(MM) \
    case MM: hipLaunchKernelGGL((k_batch_spectral<MM>), grid, dim3(256), lds, s, a); break;
            VCT_SPEC_CASE(1) VCT_SPEC_CASE(2) VCT_SPEC_CASE(3) VCT_SPEC_CASE(4)
            VCT_SPEC_CASE(5) VCT_SPEC_CASE(6) VCT_SPEC_CASE(7) VCT_SPEC_CASE(8)
#undef VCT_SPEC_CASE
        }
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

vct_status vct_batch_loss(vct_plan* p, const float* d_L, int B, int M, const float* d_att, const float* d_pdf, int NE,
                          int S, const float* d_target, float* d_w, float* d_loss, void* stream)
{
    clear_error();
    VCT_CHECK(check_spec(p, d_L, B, M, d_att, d_pdf, NE, S));
    if (!d_target || !d_w || !d_loss) return set_error(VCT_ERR_ARG, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    int lpr = 1;
    while (lpr < B) lpr <<= 1;
    if (lpr > 64) return set_error(VCT_ERR_UNSUPPORTED, "more than 64 phantoms in one loss launch");
    const size_t lds = sizeof(float) * ((size_t)S * NE * 9 + S + lpr) + sizeof(int) * 4 * S;
    if (lds > 160 * 1024) return set_error(VCT_ERR_UNSUPPORTED, "spectrum tables too large for LDS");
    BatchSpecArgs a{};
    a.L = d_L; a.att = d_att; a.pdf = d_pdf; a.target = d_target; a.w = d_w; a.loss = d_loss;
    a.rays = p->g.n_views * p->g.n_channels; a.W = p->g.n_channels; a.VT = p->VT;
    a.B = B; a.S = S; a.M = M; a.NE = NE;
    hipLaunchKernelGGL(k_zero_f32, dim3((B + 63) / 64), dim3(64), 0, s, d_loss, B);
    const int rays_per_block = 256 / lpr;
    const dim3 grid((a.rays + rays_per_block - 1) / rays_per_block);
    switch (M) {
#define VCT_LOSS_CASE(MM) \
    case MM: hipLaunchKernelGGL((k_batch_loss<MM>), grid, dim3(256), lds, s, a, lpr); break;
        VCT_LOSS_CASE(1) VCT_LOSS_CASE(2) VCT_LOSS_CASE(3) VCT_LOSS_CASE(4)
        VCT_LOSS_CASE(5) VCT_LOSS_CASE(6) VCT_LOSS_CASE(7) VCT_LOSS_CASE(8)
#undef VCT_LOSS_CASE
    }
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // extern "C"
