/*
 * vct.h -- C ABI of the MI355X VirtualCT hot path (libvct.so).
 *
 * The library implements the two hot paths of VirtualCT on CDNA4 (gfx950):
 * the material-based forward projection (MBFPA: per-ray elemental path
 * integrals, polyenergetic sum, max-normalisation, Gaussian detector noise)
 * and fan-beam FBP (Shepp-Logan / Ram-Lak filter as an MFMA GEMM and a
 * pixel-driven backprojector), plus the adjoint of the forward projection.
 *
 * Conventions
 *   - Every entry point returns a vct_status; the library never exits.
 *     vct_last_error() gives a message for the calling thread.
 *   - "d_" pointers are device (HBM) pointers; "h_"/unprefixed buffers of the
 *     host entry points are host pointers owned by the caller.
 *   - A stream argument is a hipStream_t passed as void* (NULL = HIP's
 *     null stream, which is ordered with all blocking streams such as
 *     PyTorch's default stream).  Calls on one stream are ordered; calls
 *     that share the context scratch buffer must not overlap across streams,
 *     nor may traces on one plan (they share the plan's density work copy).
 *     The scratch only grows: a larger request allocates a new buffer and
 *     the old one stays allocated until vct_terminate, so work recorded in a
 *     HIP graph never writes into freed memory.
 *   - Inputs are never modified (the reference zeroes negative material-0
 *     voxels in place, M/reprojection_TOMO.cpp:151; here a negative density
 *     contributes nothing, which gives the same sums without the write).
 *   - Sinograms are float32 [views][channels] (views = sinogram height,
 *     channels = width), images float32 [N][N], row index = -y, exactly the
 *     raw layouts of M/main.cpp:527-530 and F/main.cpp:187-190.
 *
 * Reference citations are relative to the VirtualCT repository; "M/" is
 * Release_MBFPA_CPU_human/ and "F/" is Release_FBP_CPU/.
 */
#ifndef VCT_H
#define VCT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VCT_ABI_VERSION 1

typedef int vct_status;
enum {
    VCT_OK = 0,
    VCT_ERR_ARG = 1,          /* invalid argument / shape                     */
    VCT_ERR_HIP = 2,          /* a HIP runtime call failed                    */
    VCT_ERR_NOMEM = 3,        /* device or host allocation failed             */
    VCT_ERR_IO = 4,           /* file missing or short                        */
    VCT_ERR_UNSUPPORTED = 5,  /* geometry/option outside what is implemented  */
    VCT_ERR_INTERNAL = 6      /* a kernel reported an internal limit (e.g. a  */
                              /* ray exceeded its step bound)                 */
};

/* Density storage of a phantom on the device (one voxel = 8 slots). */
enum {
    VCT_DENS_U16 = 0,  /* uint16 [N][N][8], partial density in mg/cm^3 (the
                          reference's int16 file value, /1000 -> g/cm^3)     */
    VCT_DENS_F32 = 1   /* float  [N][N][8], partial density in g/cm^3         */
};

enum { VCT_FILTER_RAMLAK = 1, VCT_FILTER_SHEPP_LOGAN = 2 };

typedef struct vct_ctx vct_ctx;
typedef struct vct_plan vct_plan;

/* Fan-beam geometry.  Defaults (vct_geometry_default) are the reference's
 * compile-time constants M/physParamsTOMO.h:24-62 = F/physParamsTOMO.h:23-51
 * and the input-txt values of M/IR_TOMO_input_virtual_projection_120kV.txt. */
typedef struct vct_geometry {
    int32_t recon_size;      /* N, image is N x N                        (512) */
    int32_t n_channels;      /* W, detector channels                     (609) */
    int32_t n_views;         /* V, views over 360 degrees                (800) */
    int32_t reserved0;
    double  pixel_cm;        /* pixel pitch                              (0.1) */
    double  start_view_deg;  /* starting view angle                        (0) */
    double  sid_cm;          /* source to isocentre                     (85.0) */
    double  det_radius_cm;   /* detector radius of curvature           (142.6) */
    double  det_src_iso_cm;  /* detector-arc centre to isocentre        (85.0) */
    double  iso_channel;     /* isocentre channel                      (304.5) */
    double  max_alpha;       /* projector fan half-angle         (0.326086956) */
    double  max_beta;        /* FBP fan half-angle               (0.326086956) */
    double  xfactor;         /* "Factor for output CT value"             (1)   */
} vct_geometry;

/* ------------------------------------------------------------ lifecycle */

/* Fill g with the reference defaults. */
void vct_geometry_default(vct_geometry* g);
const char* vct_status_string(vct_status s);
/* Message of the last error on this thread ("" if none). */
const char* vct_last_error(void);
int vct_abi_version(void);

/* Replaces initializeGPU (M/main.cpp:289-291; live in the vendor build):
 * binds a device and creates its streams and scratch. */
vct_status vct_init(int device, vct_ctx** ctx);
/* Replaces terminateGPU (M/main.cpp:670). */
vct_status vct_terminate(vct_ctx* ctx);

/* Device-side error flags (no counterpart in the reference, whose loops
 * have no step bound).  The device-pointer entry points enqueue work and
 * return without waiting, so a kernel that gives up (a ray that exceeds its
 * marching step bound, a system-matrix row that does not match its count)
 * only raises a flag in the context.  This call waits for `stream`, reads
 * and clears the flag and returns VCT_ERR_INTERNAL if it was set (the
 * results of the work since the last check are then invalid). */
vct_status vct_check_errors(vct_ctx* ctx, void* stream);

/* Device tables for one geometry (view cos/sin from glibc on the host, as
 * prepareTrigonometricTable M/misc_TOMO.cpp:81-100; per-channel detector
 * points M/reprojection_TOMO.cpp:60-71; filter matrices F/filterSinogramTOMO.cpp:9-138). */
vct_status vct_plan_create(vct_ctx* ctx, const vct_geometry* g, vct_plan** plan);
vct_status vct_plan_destroy(vct_plan* plan);

/* Plan options (no counterpart in the reference).
 * VCT_OPT_EXACT_TRACE = 1: march every ray with correctly rounded divisions
 * (the reference's operation sequence) instead of the default guarded
 * reciprocal march; both give the same segments, the default is faster.
 * VCT_OPT_MAX_STEPS = 2: marching steps after which a ray gives up and
 * raises the context's error flag (0, the default: 4N + 4/pixel + 64, more
 * than any ray of the geometry needs).
 * VCT_OPT_EXACT_BACKPROJECT = 4: 1 keeps, for a single sinogram, the
 * backprojection that adds each pixel's views in the reference's order
 * (>= 99% of pixels bit-identical to it for the same filtered input), and
 * the reference's FP64 terms for several sinograms per call.  By
 * default, when the view count is a multiple of 4 and N is even, one
 * sinogram is backprojected by orbits of the 90-degree rotation (pixel p
 * at view v + V/4 sees the geometry of pixel (y, N-1-x) at view v), which
 * computes each pixel-view geometry once for four pixels; its sums are in
 * another order (within 1e-5 relative of the reference's image).
 * VCT_OPT_TRACE_ORBIT = 5: rays per march of vct_trace_paths: 4 (default)
 * marches the ray of view v once for the rays of views v + kV/4 (the same
 * segments turned by quarter turns, when V is a multiple of 4 and N even),
 * 2 for views v and v + V/2 (V even), 1 marches every ray by itself.  The
 * segments are the reference's either way: every decision of the shared
 * march must clear the guarded march's margin for the turned rays too, or
 * those rays are marched by themselves; lengths may differ in the last
 * bits.  The exact option and geometries where the reference's r >= rmax
 * test can fire use one ray per march.
 * Launch shapes (results do not depend on them, beyond the order of float
 * adds where noted; an autotuner times them per shape):
 * VCT_OPT_TRACE_BLOCK = 6: lanes per workgroup of the tracer, 64, 128 or 256
 * (0: 64).
 * VCT_OPT_BP_SINOS_PER_LANE = 7: most sinograms one lane backprojects in a
 * multi-sinogram call, 2..16 (0: 16); images are bit-identical for every value.
 * VCT_OPT_ADJOINT_VIEWS = 8: views per workgroup of the adjoint, 1..256 (0: a
 * chunk that gives about 2048 workgroups); workgroups of other chunks add
 * their tiles in another order, which changes the float rounding of grad. */
enum {
    VCT_OPT_EXACT_TRACE = 1,
    VCT_OPT_MAX_STEPS = 2,
    VCT_OPT_EXACT_BACKPROJECT = 4,
    VCT_OPT_TRACE_ORBIT = 5,
    VCT_OPT_TRACE_BLOCK = 6,
    VCT_OPT_BP_SINOS_PER_LANE = 7,
    VCT_OPT_ADJOINT_VIEWS = 8
};
vct_status vct_plan_set_option(vct_plan* plan, int option, int value);

/* ------------------------------------------------------- host-side tables */

/* Replaces attenuation_coefficient (M/prior_weight_production.cpp:180-297):
 * reads a reference-format spectrum file ("Elo Ehi weight" per line in MeV
 * after one header line) and writes the normalised pdf[NE] and the mass
 * attenuation attenuation[M][NE] (cm^2/g).  Elements are, in order,
 * H C N O P Ca (the reference's table, :249-255) followed by Mg Si, so
 * num_material <= 8.  Capacity: NE <= 1024.  Z/A may be NULL for that
 * table or give num_material custom elements. */
vct_status vct_attenuation_coefficient(const char* spectrum_path, int num_material,
                                       const double* Z, const double* A, int nthin,
                                       float* attenuation, float* pdf, int* NE,
                                       double* energy_mev);

/* Replaces prepareConstMap (F/filterSinogramTOMO.cpp:9-138): C[W][W] in FP64. */
vct_status vct_const_map(const vct_geometry* g, int filter, double* C);

/* ---------------------------------------------- device-pointer hot path */

/* Pack a reference-layout int16 [M][N][N] phantom (values = 1000 x g/cm^3,
 * the Weight_input_*.raw format read at M/main.cpp:421-429) into the
 * voxel-interleaved uint16 [N][N][8] layout; negatives become 0. */
vct_status vct_pack_densities_i16(vct_plan* plan, const int16_t* d_dens, int M,
                                  uint16_t* d_packed, void* stream);
/* Same for float [M][N][N] g/cm^3 -> float [N][N][8]. */
vct_status vct_pack_densities_f32(vct_plan* plan, const float* d_dens, int M,
                                  float* d_packed, void* stream);

/* Ray tracer, replacing the geometry of reprojection_TOMO
 * (M/reprojection_TOMO.cpp:13-289): for every ray (view, channel) writes
 * the per-element path integrals d_L[V][W][8] (float; slot m = sum over the
 * ray's segments of rho_m * length, g/cm^2, divided by xfactor).  Each ray
 * is traced once, independent of the spectrum.  d_counts (optional, int32
 * [V][W]) receives the reference's per-ray segment counter
 * (reprojection_sum, :195/:280).  With orbits (VCT_OPT_TRACE_ORBIT), the
 * call first writes d_packed and its transpose, each padded by two voxels,
 * into a work buffer that the plan allocates with itself (64*(N+4)^2 bytes
 * when N and V are even), so that the gathers of rays along x read rows
 * too. */
vct_status vct_trace_paths(vct_plan* plan, const void* d_packed, int dens_type, int M,
                           float* d_L, int32_t* d_counts, void* stream);

/* Polyenergetic detector sum, replacing the energy loop M/main.cpp:492-506
 * and the per-bin output M/reprojection_TOMO.cpp:196-197:
 *   raw[s][ray] = sum_E pdf[s][E] * exp(-sum_m att[s][m][E] * L[ray][m])
 * for n_spectra spectra of NE bins each (att [S][M][NE], pdf [S][NE]).
 * d_stats[s] = {max, min} of raw[s] as float bit patterns (inputs to the
 * normalisation and the noise clamp).  d_dIdL (optional, float
 * [S][V][W][8]) receives d raw / d L_m for the adjoint. */
vct_status vct_spectral_sum(vct_plan* plan, const float* d_L, int M,
                            const float* d_att, const float* d_pdf, int NE, int n_spectra,
                            float* d_raw, uint32_t* d_stats, float* d_dIdL, void* stream);

/* Max-normalisation and detector noise, replacing M/main.cpp:511-525 and
 * gaussiannoise0 (:51-76): out = raw / max + sigma * z, clamped as the
 * reference (x < 0 -> (min/max)/10, x > 1 -> 1).  z is standard normal from
 * Philox4x32-10 keyed by seed, with counter (element, stream_id + s).
 * sigma = 0 gives the plain normalisation. */
vct_status vct_detector(vct_plan* plan, const float* d_raw, const uint32_t* d_stats,
                        int n_spectra, double sigma, uint64_t seed, uint32_t stream_id,
                        float* d_out, void* stream);

/* Whole MBFPA forward for one phantom and n_spectra spectra:
 * trace + spectral sum + detector.  d_work must hold
 * vct_forward_workspace_bytes(plan, n_spectra) bytes (or be NULL to use the
 * context scratch). */
size_t vct_forward_workspace_bytes(const vct_plan* plan, int n_spectra);
vct_status vct_mbfpa_forward(vct_plan* plan, const void* d_packed, int dens_type, int M,
                             const float* d_att, const float* d_pdf, int NE, int n_spectra,
                             double sigma, uint64_t seed, uint32_t stream_id,
                             float* d_sino, void* d_work, void* stream);

/* The same forward without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  Its
 * work buffer is one static device array per GPU of
 * vct_noalloc_scratch_bytes() bytes (64 MiB: the path integrals, V*W*32 B,
 * plus n_spectra*V*W*4 B of raw sums, e.g. up to 26 spectra at 800 x 609),
 * resolved by vct_init and held until the process ends; larger requests
 * return VCT_ERR_ARG.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (other streams, other threads, other contexts) race;
 * each call overwrites what it reads. */
size_t vct_noalloc_scratch_bytes(void);
vct_status vct_mbfpa_forward_noalloc(vct_plan* plan, const void* d_packed, int dens_type, int M,
                                     const float* d_att, const float* d_pdf, int NE, int n_spectra,
                                     double sigma, uint64_t seed, uint32_t stream_id, float* d_sino,
                                     void* stream);

/* FBP filter, replacing filterSinogram (F/filterSinogramTOMO.cpp:205-275):
 * fltd[v][m] = sum_k C[m][k] * (-ln proj[v][k] * sid * cos((k-W/2) dbeta)) * dbeta
 * as an FP32 MFMA GEMM over n_sino stacked sinograms.  in_mode != 0 applies
 * the "IN" threshold of :250-255. */
vct_status vct_filter_sinogram(vct_plan* plan, const float* d_proj, int n_sino, int filter,
                               int in_mode, float* d_fltd, void* stream);
/* The same with a caller-owned workspace (the packed GEMM operand) of
 * vct_filter_workspace_bytes(plan, n_sino) bytes instead of the context
 * scratch; callers that record the filter in a HIP graph pass their own. */
size_t vct_filter_workspace_bytes(const vct_plan* plan, int n_sino);
vct_status vct_filter_sinogram_ws(vct_plan* plan, const float* d_proj, int n_sino, int filter,
                                  int in_mode, float* d_fltd, void* d_work, void* stream);
/* Beam-hardening correction, replacing the standalone tool
 * M/beam_hardening_correction.cpp:21-69 (run between MBFPA and FBP):
 * y = -ln p, y' = |y| + alpha*|y|^beta, out = exp(-y'), elementwise over n
 * floats (in place allowed). */
vct_status vct_beam_hardening(const float* d_in, int n, double alpha, double beta, float* d_out,
                              void* stream);
/* The FBP filter with that correction fused into the channel pre-weight
 * (one pass instead of tool -> file -> filter). */
vct_status vct_filter_sinogram_bhc(vct_plan* plan, const float* d_proj, int n_sino, int filter,
                                   int in_mode, double alpha, double beta, float* d_fltd, void* stream);
/* Max and min of n floats on the device, returned to host h_maxmin[2]
 * (synchronous); the scans of M/main.cpp:512-518 and :53-60. */
vct_status vct_minmax(const float* d_x, int n, float* h_maxmin, void* stream);

/* The npf side output of the MBFPA program, replacing lengthproj_TOMO
 * (M/lengthproj_TOMO.cpp:14-185, called at M/main.cpp:407-419 and written
 * to npf_float.raw by print_npf_float, :78-91): for every pixel inside the
 * reconstruction circle, npf = sum over the views of the window of
 * proj(view, channel hit by the source->pixel ray) * chord length, float
 * accumulation in view order; pixels outside the circle get 0.
 * d_proj: float [V/2*2][W], row j = view (start + j) mod V of the window
 * that starts at start = (int)(z/(couch/V) - V/2 + 1e-8) (the sinogram the
 * reference interpolates for slice z, M/main.cpp:352-372).  iso_det_cm is
 * the isocentre-to-detector distance (57.6 cm, M/physParamsTOMO.h:29).
 * Rejects (VCT_ERR_ARG) a window that starts before view 0, where the
 * reference indexes its tables with negative views.  d_npf: float [N][N]. */
vct_status vct_lengthproj(vct_plan* plan, const float* d_proj, double z_cm, double couch_speed_cm,
                          double iso_det_cm, float* d_npf, void* stream);
/* The same for a window whose rows hold the first proj_width (<= the plan's
 * n_channels) channels of the plan's detector: measured data keep 582 or
 * fewer channels of the 609-channel detector, whose pitch and isocentre
 * channel still place the hits (the reference's compile-time End_Channel and
 * isocenter_channel, M/physParamsTOMO.h:43-47, with the measured width as
 * the row stride, M/lengthproj_TOMO.cpp:152-166).  A hit past the window's
 * last channel takes that channel's value; the reference reads the next
 * row's channels there (past its buffer on the last rows). */
vct_status vct_lengthproj_window(vct_plan* plan, const float* d_proj, int proj_width, double z_cm,
                                 double couch_speed_cm, double iso_det_cm, float* d_npf, void* stream);

/* Pixel-driven fan backprojection, replacing reconstructImage
 * (F/reconstructImageTOMO.cpp:11-105), n_sino images.  Several sinograms in
 * one call share the per-pixel geometry (up to 16 per pass), so one call with
 * n_sino = S is faster than S calls.  By default their terms apply the
 * interpolation and distance weights in float (within 1e-6 relative of the
 * reference's FP64 terms); with VCT_OPT_EXACT_BACKPROJECT each image equals a
 * single-sinogram call bit for bit. */
vct_status vct_backproject(vct_plan* plan, const float* d_fltd, int n_sino, float* d_img,
                           void* stream);
/* The backprojections of n_groups (<= 16) geometries in one launch, as one
 * vct_backproject call per group (images bit-identical when the plans share
 * VCT_OPT_EXACT_BACKPROJECT; if any asks for it, all groups get exact
 * terms): group k takes
 * n_sino[k] filtered sinograms at d_fltd[k] and writes d_img[k], with its
 * plan's geometry; sinos_per_lane (2..16, 0: 16) is the number of sinograms
 * a lane sums at once, shared by all groups.  Serves batches of phantoms of
 * uneven sizes (the reference reconstructs one image per run,
 * F/reconstructImageTOMO.cpp:11-105). */
vct_status vct_backproject_groups(vct_plan* const* plans, const float* const* d_fltd, const int* n_sino,
                                  float* const* d_img, int n_groups, int sinos_per_lane, void* stream);
/* The same with each group's filtered sinograms interleaved in slices of k
 * = sinos_per_lane (4, 8, 12 or 16): d_il[g] as vct_interleave_sinograms
 * writes it.  A lane then reads the taps of its k sinograms as contiguous
 * 16-byte loads; the images equal vct_backproject_groups' bit for bit.
 * vct_interleave_sinograms writes n_sino sinograms [n_sino][V][W] at d_fltd
 * as d_il[s / k][V][W][k] (zeros past n_sino), into
 * vct_interleaved_bytes(plan, n_sino, k) bytes (0 for an invalid k).  No
 * counterpart in the reference (it reconstructs one image per run,
 * F/reconstructImageTOMO.cpp:11-105); serves batched FBP. */
size_t vct_interleaved_bytes(const vct_plan* plan, int n_sino, int k);
vct_status vct_interleave_sinograms(const vct_plan* plan, const float* d_fltd, int n_sino, int k, float* d_il,
                                    void* stream);
vct_status vct_backproject_groups_il(vct_plan* const* plans, const float* const* d_il, const int* n_sino,
                                     float* const* d_img, int n_groups, int sinos_per_lane, void* stream);

/* Adjoint of vct_trace_paths weighted per ray:
 *   grad[v][m] += sum_rays w[ray][m] * sum_{segments of ray in voxel v} length
 * where w[ray][m] = d_dIdL[ray][m] * d_g[ray] (d_g may be NULL = 1).  grad is
 * float [N][N][8]; it is accumulated into (clear it first).  Segments whose
 * voxel has a negative density of element m contribute nothing to slot m. */
vct_status vct_trace_adjoint(vct_plan* plan, const void* d_packed, int dens_type, int M,
                             const float* d_dIdL, const float* d_g, float* d_grad,
                             void* stream);

/* ------------------------------------------------- batched execution */

/* Cached segment lists of one geometry (the "system matrix").  The reference
 * caches the same (voxel, length) pairs per segment counter during its trace
 * (reprojection_systemMatrixfloat / reprojection_vnum / reprojection_sum,
 * M/reprojection_TOMO.cpp:160-163, :186-188, :195) and replays them in
 * reprojection_TOMO_2 (:291-352); here they are built once on the device by
 * the exact tracer and stored ray-major (CSR).  Flags:
 *   VCT_SYSMAT_TRANSPOSE  also voxel-major (CSC, rays in increasing order
 *                         within each voxel), needed by vct_batch_adjoint;
 *   VCT_SYSMAT_BLOCKED    also the blocked two-stage form of each direction
 *                         built: entries regrouped by 64x64 image tile
 *                         (forward) and by 50-view block x tile (adjoint),
 *                         so that each block's reads stay in one L2; used by
 *                         the batched gathers of groups that pass a
 *                         workspace.  With it, the forward stage 1 of a
 *                         group reads the 64x64 tile of its phantom from LDS,
 *                         and the adjoint stage 1 the weights of the rays
 *                         that cross one part of a tile in a 50-view block
 *                         (runs sorted by length in interleaved 64-run
 *                         slices; the adjoint form is left out when a 16x16
 *                         image cell is crossed by more than 3328 rays of a
 *                         view block), unless
 *   VCT_SYSMAT_NO_TILE    is also given (the same sums, read through L2).
 * No counterpart in the reference. */
enum { VCT_SYSMAT_TRANSPOSE = 1, VCT_SYSMAT_BLOCKED = 2, VCT_SYSMAT_NO_TILE = 4 };
typedef struct vct_sysmat vct_sysmat;
vct_status vct_sysmat_build(vct_plan* plan, int flags, vct_sysmat** out, void* stream);
vct_status vct_sysmat_destroy(vct_sysmat* m);
int64_t vct_sysmat_segments(const vct_sysmat* m);
/* The cached segments of one ray (view * channels + channel) in march
 * order, copied to the host (waits for `stream`): voxel index x + y*N and
 * float length, i.e. the reference's reprojection_vnum /
 * reprojection_systemMatrixfloat entries of the ray's in-image segment
 * counters (M/reprojection_TOMO.cpp:160-163, :186-188).  *n_out receives
 * the count; VCT_ERR_ARG if it exceeds capacity. */
vct_status vct_sysmat_row(const vct_sysmat* m, int ray, uint32_t* h_voxel, float* h_length, int capacity,
                          int* n_out, void* stream);
size_t vct_sysmat_bytes(const vct_sysmat* m);
/* Bytes of the workspace a batched gather over this system matrix needs for
 * n_phantoms phantoms in the blocked form: the stage-1 partial sums and,
 * with the adjoint's LDS-resident form, a phantom-major copy of the adjoint
 * weights (n_rays * n_phantoms * 32 bytes).  One workspace serves both
 * directions. */
size_t vct_sysmat_workspace(const vct_sysmat* m, int n_phantoms);

/* One group of phantoms sharing a system matrix.  Batched arrays are
 * phantom-interleaved rows of 8*B floats: densities and gradients one row per
 * voxel (N*N rows), path integrals and adjoint weights one row per ray.  A
 * row holds three phantom-major planes: materials 0-3 of phantom b at 4b,
 * materials 4-5 at 4B + 2b, materials 6-7 at 6B + 2b; materials >= M are 0.
 * Below, x[row][b][m] names that element.  For B = 1 a row is the plain [8]. */
typedef struct vct_batch_group {
    const vct_sysmat* sysmat;
    const float* d_in;   /* paths: densities; adjoint: weights            */
    float* d_out;        /* paths: path integrals; adjoint: gradient (overwritten) */
    int32_t n_phantoms;  /* B <= lanes_per_row                           */
    int32_t reserved;
    float* d_work;       /* NULL, or vct_sysmat_workspace(sysmat, n_phantoms)
                            bytes: gather in the blocked form (partial sums are
                            rounded to float once); launches that share a
                            workspace must not overlap */
} vct_batch_group;

/* Batched replay of the cached segments, replacing reprojection_TOMO_2
 * (M/reprojection_TOMO.cpp:291-352) for many phantoms and geometries in one
 * persistent launch (up to 16 groups; lanes_per_row in {8,16,32,64}, or 1
 * when every group holds one phantom: each row is then split over 16 lanes;
 * blocks <= 0 picks the persistent grid).  Groups with a workspace gather in
 * two stages: stage 1 of the groups with an LDS-resident form in one launch
 * (in the adjoint after copying each group's weights phantom-major into its
 * workspace), the others' stage 1 in the persistent launch, then stage 2.
 * out[ray][b][m] = sum_seg in[vox][b][m]*len / xfactor (the tracer's units). */
vct_status vct_batch_paths(const vct_batch_group* groups, int n_groups, int M, int lanes_per_row,
                           int blocks, void* stream);
/* Batched transpose (no counterpart in the reference, whose iterative loop
 * backprojects with lengthproj_TOMO_2, M/lengthproj_TOMO.cpp:188-276):
 * out[vox][b][m] = sum_{rays through vox} in[ray][b][m] * len / xfactor,
 * atomic-free and deterministic. */
vct_status vct_batch_adjoint(const vct_batch_group* groups, int n_groups, int M, int lanes_per_row,
                             int blocks, void* stream);
/* Polyenergetic sum for B phantoms x S spectra (as vct_spectral_sum):
 * L [rays][B][8] -> raw [B][S][rays], stats [B][S][2]. */
vct_status vct_batch_spectral(vct_plan* plan, const float* d_L, int B, int M, const float* d_att,
                              const float* d_pdf, int NE, int S, float* d_raw, uint32_t* d_stats,
                              void* stream);
/* The same for n_groups (<= 16) path-integral arrays of one detector
 * geometry (the plan's views and channels): group g has n_phantoms[g]
 * phantoms in rows of 8 * n_phantoms[g] floats at d_L[g]; raw and stats hold
 * the phantoms of all groups in group order, [sum B][S][rays] and
 * [sum B][S][2], so one detector call then covers them all. */
vct_status vct_batch_spectral_groups(vct_plan* plan, const float* const* d_L, const int* n_phantoms, int n_groups,
                                     int M, const float* d_att, const float* d_pdf, int NE, int S, float* d_raw,
                                     uint32_t* d_stats, void* stream);
/* Dual-energy decomposition loss (no counterpart in the reference): for
 * targets t [B][S][rays] and normalisation I_max,s = sum_E pdf_s,
 *   loss[b] = sum_s sum_rays (ln(I/I_max) - ln t)^2,
 *   w[ray][b][m] = d loss / d L_m  (input of vct_batch_adjoint). */
vct_status vct_batch_loss(vct_plan* plan, const float* d_L, int B, int M, const float* d_att,
                          const float* d_pdf, int NE, int S, const float* d_target, float* d_w,
                          float* d_loss, void* stream);

/* Projected Barzilai-Borwein step of the batched decomposition (no
 * counterpart in the reference, whose iterative loop, M/IR_ImageTOMO.cpp:19-72,
 * is disabled) on phantom-interleaved arrays of n_vox rows:
 *   s = x - x_prev, y = g - g_prev, per phantom sy = sum s.y, ss = sum s.s
 *   (FP64, fixed order), alpha_b = clamp(sy > 0 ? ss / max(sy, 1e-30)
 *   : 2 alpha_b, 1e-8, 1e2) (skipped when first != 0),
 *   x_prev = x, g_prev = g, x = max(x - alpha_b g, 0).
 * d_work: vct_batch_bb_workspace(B) bytes; B <= 64.  Allocates nothing. */
size_t vct_batch_bb_workspace(int n_phantoms);
vct_status vct_batch_bb_step(float* d_x, float* d_x_prev, float* d_g_prev, const float* d_g, float* d_alpha,
                             int n_vox, int n_phantoms, int first, double* d_work, void* stream);

/* ------------------------------------------------ host-buffer drop-ins */

/* Replaces reprojectionOnGPU_TOMO (M/main.cpp:439-446) = the energy loop
 * over reprojection_TOMO (M/main.cpp:492-506): dens float [M][N][N] g/cm^3,
 * att [M][NE], pdf [NE]; sino_full float [V][W] = the un-normalised sum. */
vct_status vct_reprojection(vct_ctx* ctx, const vct_geometry* g, const float* dens, int M,
                            const float* att, const float* pdf, int NE, float* sino_full);
/* Replaces the normalisation and gaussiannoise0 (M/main.cpp:511-525),
 * in place on a host sinogram of n floats.  stats (optional, float[2])
 * receives {max, min} of the input (projection_min = min / max). */
vct_status vct_gaussian_noise(vct_ctx* ctx, float* sino, int n, double sigma, uint64_t seed,
                              float* stats);
/* Replaces filterSinogram (F/filterSinogramTOMO.cpp:205); host [V][W]. */
vct_status vct_filter_sinogram_host(vct_ctx* ctx, const vct_geometry* g, const float* proj,
                                    int filter, int in_mode, float* fltd);
/* Replaces reconstructImage (F/reconstructImageTOMO.cpp:11); host buffers. */
vct_status vct_reconstruct_image_host(vct_ctx* ctx, const vct_geometry* g, const float* fltd,
                                      float* img);
/* Replaces lengthproj_TOMO (M/lengthproj_TOMO.cpp:14) as vct_lengthproj,
 * host buffers proj [V/2*2][W] -> npf [N][N]. */
vct_status vct_lengthproj_host(vct_ctx* ctx, const vct_geometry* g, const float* proj, double z_cm,
                               double couch_speed_cm, double iso_det_cm, float* npf);
/* vct_lengthproj_window on host buffers, proj [V/2*2][proj_width]. */
vct_status vct_lengthproj_window_host(vct_ctx* ctx, const vct_geometry* g, const float* proj, int proj_width,
                                      double z_cm, double couch_speed_cm, double iso_det_cm, float* npf);

/* ------------------------------------------- measured (real-scanner) data */

/* Shape of measured detector data, as loadData derives it
 * (M/projectionDataTOMO.cpp:126-134, :331-333): detector rows of 609
 * channels of which the first 27 are skipped; air-scan files of 1 + 5 blocks
 * of 800 x 609 floats (block 0 skipped). */
typedef struct vct_measured {
    int32_t in_mode;    /* 0: "PR" pre-treatment, 1: "IN" in-treatment        */
    int32_t rotations;  /* totalProjNumber                                   */
    int32_t height;     /* views per rotation after sino_fact (projImgHeight) */
    int32_t width;      /* channels kept (projImgWidth), width + 27 <= 609   */
    int32_t sino_fact;  /* PR keeps rows k with k % sino_fact == 0           */
    int32_t skip_rows;  /* IN: rows after each view that are not used (4 when
                           the sampling rate is >= 200 Hz, else 0)          */
} vct_measured;

/* Air-scan normalisation, replacing loadProjectionData
 * (M/projectionDataTOMO.cpp:380-497) on device data:
 *   d_air [5][800][609] (the five averaged blocks);
 *   d_raw PR [rotations][800][609], IN [rotations][height][609];
 *   d_volume float and d_mask uint8 [rotations*height][width];
 *   d_max_bits: float bits of the largest IN value (set by the call).
 * PR: kept rows divided by the air average, packed rotation after rotation.
 * IN: only view 0 of each rotation holds data (the reference indexes its
 * one-row read buffer at the view's row, which reads zeros otherwise),
 * divided by twice the air average of row 0; values below 0.1 become 0 with
 * mask 0.  Rejects shapes for which the reference reads or writes past its
 * buffers (width + 27 > 609; PR with ceil(800/sino_fact) > height; IN with
 * height > 800).  Allocates nothing. */
vct_status vct_measured_normalise(const vct_measured* m, const float* d_air, const float* d_raw, float* d_volume,
                                  uint8_t* d_mask, uint32_t* d_max_bits, void* stream);
/* The helical window of M/main.cpp:352-371 (view 0, 360 degrees): rows
 * n < height/2*2 of d_volume [rotations*height][width] blended in FP64
 * with the same row one rotation later, weight 0.5*|n - height/2|/(height/2),
 * for the first half when that rotation exists -> d_window [height/2*2][width]
 * (the sinogram vct_lengthproj takes). */
vct_status vct_helical_window(const float* d_volume, int rotations, int height, int width, float* d_window,
                              void* stream);
/* loadProjectionData on files (the detector file and the air-scan file the
 * txt names): host volume and mask [rotations*height][width] and the largest
 * IN value.  Where the reference prints "Maximum projection ... > 1.0...
 * STOP" and exits, this returns VCT_ERR_ARG; a short file is VCT_ERR_IO (the
 * reference leaves stale buffer contents). */
vct_status vct_load_measured_host(vct_ctx* ctx, const vct_measured* m, const char* detector_path,
                                  const char* airscan_path, float* volume, uint8_t* mask, double* max_proj);
/* vct_helical_window on host buffers. */
vct_status vct_helical_window_host(vct_ctx* ctx, const float* volume, int rotations, int height, int width,
                                   float* window);

#ifdef __cplusplus
}
#endif
#endif /* VCT_H */
