/*
 * vct_oracle.h -- CPU oracle for the VirtualCT MBFPA + fan-beam FBP path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference
 * algorithm, used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product library (virtualct_amd/libvct.so)
 * never links, loads or calls anything in this directory.
 *
 * Two arithmetic modes are provided for every stage:
 *   *_ref : the reference's own rounding sequence (float densities, float
 *           rho*mu products, FP64 transmission sum, float per-bin outputs,
 *           float accumulation of the filter and backprojection), so results
 *           are bit-identical to the compiled reference on x86-64/glibc.
 *   *_f64 : the same mathematics carried in FP64 end to end (the "FP64 CPU
 *           oracle" the GPU path is also reported against).
 */
#ifndef VCT_ORACLE_H
#define VCT_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geom {
    int    N;            /* reconstruction size (square)                      */
    double pix;          /* reconstruction pixel size [cm]                     */
    int    W;            /* detector channels (sinogram width)                 */
    int    V;            /* views (sinogram height)                            */
    double start_deg;    /* starting view angle [deg]                          */
    double sid;          /* source to isocentre [cm]                    (85)   */
    double det_radius;   /* detector radius of curvature [cm]         (142.6)  */
    double det_src_iso;  /* detector-arc centre to isocentre [cm]       (85)   */
    double iso_channel;  /* isocentre channel                          (304.5) */
    double max_alpha;    /* fan half-angle used by the projector     (0.3261) */
    double max_beta;     /* fan half-angle used by the FBP filter/BP (0.3261) */
    double xfactor;      /* "Factor for output CT value" (an int in the txt)   */
} orc_geom;

void orc_geom_default(orc_geom* g);

/* view and channel tables (glibc libm, FP64) */
void orc_view_tables(const orc_geom* g, double* cosT, double* sinT);
void orc_channel_table(const orc_geom* g, double* px0, double* py0);

/* Segment list of one ray (view, ch).  Writes up to maxseg (voxel, length)
 * pairs for segments that land inside the image; *n_iks receives the
 * reference's per-ray segment counter (all candidate segments, including
 * those whose midpoint falls outside the image).  Returns the number of
 * in-image segments written, or -1 if maxseg was too small. */
int orc_trace_ray(const orc_geom* g, const double* cosT, const double* sinT,
                  const double* px0, const double* py0, int view, int ch,
                  int32_t* vox, double* rr, int maxseg, int* n_iks);

/* Same, also writing each segment's reference counter value (the index
 * under which M/reprojection_TOMO.cpp:160-163 caches it). */
int orc_trace_ray_iks(const orc_geom* g, const double* cosT, const double* sinT,
                      const double* px0, const double* py0, int view, int ch,
                      int32_t* vox, double* rr, int32_t* seg_iks, int maxseg, int* n_iks);

/* Per-ray candidate segment counts over the whole sinogram [V][W]. */
int64_t orc_segment_counts(const orc_geom* g, int32_t* counts);

/* Spectrum file -> normalised pdf[NE] and mass attenuation att[M][NE].
 * Z/A arrays give the elements (M entries).  Returns NE, or -1 on error. */
int orc_attenuation(const char* path, int M, const double* Z, const double* A,
                    int nthin, float* att, float* pdf, double* energy_mev,
                    int max_ne);

/* Forward projection.  dens: float [M][N][N] partial densities (g/cm^3).
 * out: float [V][W] = sum over energy bins (not normalised). */
void orc_forward_ref(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE, float* out);
void orc_forward_f64(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE, double* out,
                     double* Lpath /* optional [V][W][M], may be NULL */);

/* Max-normalise in place; returns the max.  */
float orc_normalize(float* sino, int n);
/* The reference detector noise model with glibc rand() (serial).
 * order: 0 = draw U for the log term first, 1 = draw U for the sin term first. */
void orc_gaussian_noise_glibc(float* sino, int n, double sigma, int order);

/* Philox4x32-10 (Salmon et al., SC'11): ctr[4], key[2] -> out[4]. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* Beam-hardening correction tool (elementwise, in place). */
void orc_beam_hardening(float* p, int n, double alpha, double beta, double* min_org, double* min_new);

/* FBP */
void orc_const_map(const orc_geom* g, int filter, double* C /* [W][W] */);
void orc_filter_ref(const orc_geom* g, const double* C, const float* proj,
                    float* fltd, int in_mode);
void orc_filter_f64(const orc_geom* g, const double* C, const float* proj,
                    double* fltd);
void orc_mask(int N, uint8_t* mask);
void orc_backproject_ref(const orc_geom* g, const float* fltd, float* img);
void orc_backproject_f64(const orc_geom* g, const double* fltd, double* img);

/* Adjoint of the (pre-normalisation) forward model in FP64:
 * grad[m][v] = sum_rays gI[ray] * dI[ray]/drho_m(v), where
 * I[ray] = sum_E pdf_E exp(-sum_m mu_m(E) L_m / xfactor). */
void orc_adjoint_f64(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE,
                     const double* gI /* [V][W] */, double* grad /* [M][N][N] */);

/* Chord-weighted pixel backprojection of a normalised sinogram, the npf
 * side output of the MBFPA program.  proj: float [V][W] (row j is the j-th
 * view of the window that starts at view (int)(z/(couch/V) - V/2 + 1e-8)).
 * npf: [N][N], masked pixels written, others set to 0.  iso_det is the
 * isocentre-to-detector distance (57.6 cm).  Returns 0, or -1 when the view
 * window starts before view 0 (the reference then indexes its tables with
 * negative views).  _ref accumulates in float as the reference, _f64 in
 * double. */
int orc_lengthproj_ref(const orc_geom* g, const float* proj, double z_cm, double couch_cm,
                       double iso_det, float* npf);
int orc_lengthproj_f64(const orc_geom* g, const float* proj, double z_cm, double couch_cm,
                       double iso_det, double* npf);
/* The same for a window whose rows hold the first pw <= W channels of the
 * W-channel detector (row stride pw; measured data keep fewer channels than
 * the detector has, M/lengthproj_TOMO.cpp:152-166 with End_Channel 609).
 * Returns -1 also for pw outside 2..W. */
int orc_lengthproj_window_ref(const orc_geom* g, const float* proj, int pw, double z_cm, double couch_cm,
                              double iso_det, float* npf);
int orc_lengthproj_window_f64(const orc_geom* g, const float* proj, int pw, double z_cm, double couch_cm,
                              double iso_det, double* npf);

/* Measured-data loader, M/projectionDataTOMO.cpp:323-524, on arrays holding
 * the file contents (detector of 609 channels, the first 27 skipped; air
 * scans of 1 + 5 blocks of 800 x 609, block 0 skipped).
 *   air: float [6][800][609]
 *   raw: in_mode 0 (PR): float [rot][800][609], one block per rotation;
 *        in_mode 1 (IN): float [rot][H][609], the rows the reference uses
 *        (without the 4 rows it skips after each row at >= 200 Hz)
 *   out: float [rot*H][W] and mask uint8 [rot*H][W] (H = rows after
 *        sino_fact); max_proj: the largest normalised value (IN).
 * Returns 0, or -1 for shapes the reference reads out of bounds. */
int orc_load_measured(int in_mode, int rot, int H, int W, int nskip, const float* air, const float* raw,
                      float* out, uint8_t* mask, double* max_proj);
/* Helical window, M/main.cpp:352-371: rows H/2*2 of the volume [rot*H][W]
 * blended with the same rows one rotation later (first half only, when that
 * rotation exists), weight 0.5*|n - H/2|/(H/2). */
void orc_helical_window(const float* volume, int rot, int H, int W, float* window);

#ifdef __cplusplus
}
#endif
#endif
