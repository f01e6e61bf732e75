/*
 * vct_oracle.c -- CPU oracle for the VirtualCT MBFPA + fan-beam FBP path.
 *
 * TEST INFRASTRUCTURE ONLY: a plain-C restatement of the reference algorithm
 * used as the checker by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg.  The product library never links or calls this file.
 *
 * Every function cites the reference file:line it restates ("M/" is
 * Release_MBFPA_CPU_human/, "F/" is Release_FBP_CPU/).  The *_ref functions
 * keep the reference's exact rounding sequence so that, compiled with
 * gcc -O2 -ffp-contract=off on x86-64/glibc, they are bit-identical to the
 * reference; the *_f64 functions carry the same mathematics in FP64.
 *
 * Pinning (see tests/test_oracle_*.py): trace + forward + normalise + noise
 * reproduce the shipped M/reprojection_float.raw bit for bit; filter + BP
 * reproduce the shipped F/FBP_virtual_projection_512x512_human.raw bit for
 * bit; the attenuation table, segment lists and small-geometry sinograms are
 * compared with the reference compiled from source (oracle/_ref); Philox is
 * checked against the published Random123 known-answer vectors.
 */
#include "vct_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <limits.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.141592653589793238462643383279

void orc_geom_default(orc_geom* g)
{
    /* M/physParamsTOMO.h:24-62, F/physParamsTOMO.h:23-51, M/IR_TOMO_input_*.txt */
    g->N = 512;
    g->pix = 0.1;
    g->W = 609;
    g->V = 800;
    g->start_deg = 0.0;
    g->sid = 85.0;
    g->det_radius = 142.6;
    g->det_src_iso = 85.0;
    g->iso_channel = 304.5;
    g->max_alpha = 0.326086956;
    g->max_beta = 0.326086956;
    g->xfactor = 1.0;
}

/* x86-64 cvttsd2si semantics: NaN and out-of-range give INT_MIN. */
static inline int trunc_i32(double v)
{
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

void orc_view_tables(const orc_geom* g, double* cosT, double* sinT)
{
    /* angles: M/projectionDataTOMO.cpp:362-368; tables: M/misc_TOMO.cpp:93-97 */
    for (int i = 0; i < g->V; i++) {
        double angle = (double)(360.0 / g->V * i);
        double rad = (g->start_deg + angle) * ORC_PI / 180.0;
        cosT[i] = cos(rad);
        sinT[i] = sin(rad);
    }
}

void orc_channel_table(const orc_geom* g, double* px0, double* py0)
{
    /* M/reprojection_TOMO.cpp:59-71 with the un-parenthesised macro
     * Delta_Equal_angle_spacing_alpha = Max_alpha/((End-Start-1)/2). */
    const int half = (g->W - 1) / 2;
    for (int c = 0; c < g->W; c++) {
        double theta = (g->iso_channel - c) * g->max_alpha / half;
        double ox0 = g->det_radius * sin(theta);
        double oy0 = -g->det_radius * cos(theta);
        px0[c] = ox0;
        py0[c] = oy0 + g->det_src_iso;
    }
}

typedef struct {
    int xmajor;
    int N, nxm, nym;
    double D, rmax;
    int32_t* vox;
    double* rr;
    int32_t* iks;  /* optional: reference segment counter of each segment */
    int maxseg, nseg, overflow;
} seg_sink;

/* One candidate segment (a0,b0)->(a1,b1) in (major, minor) coordinates:
 * M/reprojection_TOMO.cpp:141-166 and :168-189 (x-major), :228-274 (y-major). */
static void emit_segment(seg_sink* s, int iks, double a0, double b0, double a1, double b1)
{
    double xb0, yb0, xb1, yb1;
    if (s->xmajor) { xb0 = a0; yb0 = b0; xb1 = a1; yb1 = b1; }
    else           { xb0 = b0; yb0 = a0; xb1 = b1; yb1 = a1; }
    double rr = sqrt((xb0 - xb1) * (xb0 - xb1) + (yb0 - yb1) * (yb0 - yb1));
    int xmc = trunc_i32((xb0 + xb1) * 0.5 / s->D + s->nxm);
    int ymc = trunc_i32((-yb0 - yb1) * 0.5 / s->D + s->nym);
    if (xmc >= 0 && ymc >= 0 && xmc <= s->N - 1 && ymc <= s->N - 1) {
        double r = sqrt((xb0 + xb1) * (xb0 + xb1) + (-yb0 - yb1) * (-yb0 - yb1)) * 0.5;
        if (r >= s->rmax) return; /* ph = 0 for every material */
        if (s->nseg >= s->maxseg) { s->overflow = 1; return; }
        s->vox[s->nseg] = xmc + ymc * s->N;
        s->rr[s->nseg] = rr;
        if (s->iks) s->iks[s->nseg] = iks;
        s->nseg++;
    }
}

int orc_trace_ray(const orc_geom* g, const double* cosT, const double* sinT,
                  const double* px0t, const double* py0t, int view, int ch,
                  int32_t* vox, double* rrv, int maxseg, int* n_iks)
{
    return orc_trace_ray_iks(g, cosT, sinT, px0t, py0t, view, ch, vox, rrv, NULL, maxseg, n_iks);
}

int orc_trace_ray_iks(const orc_geom* g, const double* cosT, const double* sinT,
                      const double* px0t, const double* py0t, int view, int ch,
                      int32_t* vox, double* rrv, int32_t* seg_iks, int maxseg, int* n_iks)
{
    /* M/reprojection_TOMO.cpp:25-37 */
    const int N = g->N;
    const double D = g->pix;
    const int nxm = (int)(N * 0.5);
    const int nym = (int)(N * 0.5);
    /* :53-54 source, :67-74 detector point */
    const double c = cosT[view], s = sinT[view];
    const double sx = g->sid * s;
    const double sy = g->sid * c;
    const double px0 = px0t[ch], py0 = py0t[ch];
    const double px = +px0 * c + py0 * s;
    const double py = -px0 * s + py0 * c;

    /* :76-101 clip against the image square */
    double xx[4] = {0, 0, 0, 0}, yy[4] = {0, 0, 0, 0};
    int m = 0;
    const double Lx = (double)nxm * D, Ly = (double)nym * D;
    yy[m] = Ly;
    if (fabs((yy[m] - sy) / (py - sy) * (px - sx) + sx) < Lx) { xx[m] = (yy[m] - sy) / (py - sy) * (px - sx) + sx; m++; }
    yy[m] = -Ly;
    if (fabs((yy[m] - sy) / (py - sy) * (px - sx) + sx) < Lx) { xx[m] = (yy[m] - sy) / (py - sy) * (px - sx) + sx; m++; }
    xx[m] = Lx;
    if (fabs((xx[m] - sx) / (px - sx) * (py - sy) + sy) < Ly) { yy[m] = (xx[m] - sx) / (px - sx) * (py - sy) + sy; m++; }
    xx[m] = -Lx;
    if (fabs((xx[m] - sx) / (px - sx) * (py - sy) + sy) < Ly) { yy[m] = (xx[m] - sx) / (px - sx) * (py - sy) + sy; m++; }
    *n_iks = 0;
    if (m < 2) return 0; /* ray misses the image: no segments (transmission 1) */

    /* :103-108 entry point = the intersection closer to the source */
    double xi, yi, xf, yf;
    if ((xx[0] - sx) * (xx[0] - sx) + (yy[0] - sy) * (yy[0] - sy) <
        (xx[1] - sx) * (xx[1] - sx) + (yy[1] - sy) * (yy[1] - sy)) {
        xi = xx[0]; yi = yy[0]; xf = xx[1]; yf = yy[1];
    } else {
        xi = xx[1]; yi = yy[1]; xf = xx[0]; yf = yy[0];
    }

    seg_sink sink;
    sink.xmajor = fabs(xi - xf) > fabs(yi - yf); /* :113 */
    sink.N = N; sink.nxm = nxm; sink.nym = nym; sink.D = D;
    sink.rmax = N * 0.5 * 1.0; /* :37 */
    sink.vox = vox; sink.rr = rrv; sink.iks = seg_iks; sink.maxseg = maxseg; sink.nseg = 0; sink.overflow = 0;

    /* The two branches of the reference (:113-200 x-major, :201-283 y-major)
     * are the same march with the axes exchanged: a = major, b = minor. */
    double ai, af, bi, bf, sa, sb, pa, pb;
    if (sink.xmajor) { ai = xi; af = xf; bi = yi; bf = yf; sa = sx; sb = sy; pa = px; pb = py; }
    else             { ai = yi; af = yf; bi = xi; bf = xf; sa = sy; sb = sx; pa = py; pb = px; }

    const double pha = (af - ai) / fabs(ai - af);
    int ain;
    if ((ai < 0 && (af - ai) < 0) || (ai > 0 && (af - ai) > 0))
        ain = trunc_i32(trunc_i32(ai) + (af - ai) / fabs(af - ai));
    else
        ain = trunc_i32(ai);
    double aii = ain;
    double bii = (aii - sa) / (pa - sa) * (pb - sb) + sb;
    double ab0 = ai, bb0 = bi;
    int ik = 0, iks = 0, bma = 0, bmb;
    const double L = (double)nxm * D;
    while (aii <= L && aii >= -L && bii <= L && bii >= -L) {
        aii = ain + pha * ik * D;
        bmb = bma;
        double bj = bii;
        bii = (aii - sa) / (pa - sa) * (pb - sb) + sb;
        bma = trunc_i32(bii / D);
        int icheck = 0;
        double aiii = 0.0, biii = 0.0, chy = 0.0;
        if ((ik != 0 && bma != bmb) || (ik != 0 && (bii / fabs(bii) != bj / fabs(bj)))) {
            if ((bii < 0 && (bf - bi) < 0) || (bii > 0 && (bf - bi) > 0))
                biii = (double)(bma * D);
            else
                biii = (double)(bma * D) - (bf - bi) / fabs(bf - bi) * D;
            aiii = (biii - sb) / (pb - sb) * (pa - sa) + sa;
            chy = (aii - aiii) * (aii - aiii) + (bii - biii) * (bii - biii);
            icheck++;
        }
        if (icheck == 1 && chy != 0) {
            iks++;
            emit_segment(&sink, iks, ab0, bb0, aiii, biii);
            ab0 = aiii; bb0 = biii;
        }
        iks++;
        emit_segment(&sink, iks, ab0, bb0, aii, bii);
        ab0 = aii; bb0 = bii;
        ik++;
    }
    *n_iks = iks;
    return sink.overflow ? -1 : sink.nseg;
}

static int max_segments(const orc_geom* g) { return 4 * g->N + 64; }

int64_t orc_segment_counts(const orc_geom* g, int32_t* counts)
{
    double* cosT = (double*)malloc(sizeof(double) * g->V);
    double* sinT = (double*)malloc(sizeof(double) * g->V);
    double* px0 = (double*)malloc(sizeof(double) * g->W);
    double* py0 = (double*)malloc(sizeof(double) * g->W);
    orc_view_tables(g, cosT, sinT);
    orc_channel_table(g, px0, py0);
    int64_t total = 0;
    const int ms = max_segments(g);
#pragma omp parallel reduction(+ : total)
    {
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * ms);
        double* rr = (double*)malloc(sizeof(double) * ms);
#pragma omp for schedule(dynamic, 4)
        for (int v = 0; v < g->V; v++)
            for (int c = 0; c < g->W; c++) {
                int iks;
                orc_trace_ray(g, cosT, sinT, px0, py0, v, c, vox, rr, ms, &iks);
                counts[(size_t)v * g->W + c] = iks;
                total += iks;
            }
        free(vox);
        free(rr);
    }
    free(cosT); free(sinT); free(px0); free(py0);
    return total;
}

/* ---------------------------------------------------------------- tables */

int orc_attenuation(const char* path, int M, const double* ZZ, const double* AA,
                    int nthin, float* att, float* pdf, double* energy_mev, int max_ne)
{
    /* M/prior_weight_production.cpp:180-297 */
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char row[1024];
    float data0 = 0.f, data1 = 0.f, data2 = 0.f, sum = 0.f;
    int NNE = 0;
    double photonenergy[1024];
    float pdfv[1024];
    if (fgets(row, sizeof(row), fp) == NULL) { fclose(fp); return -1; }
    while (fgets(row, sizeof(row), fp) != NULL) {
        if (NNE >= 1024) { fclose(fp); return -1; }
        sscanf(row, "%f %f %f", &data0, &data1, &data2);
        photonenergy[NNE] = (data1 + data0) * 0.5;
        pdfv[NNE] = data2;
        sum += data2;
        NNE++;
    }
    fclose(fp);
    /* :210-235 bin thinning */
    double psum = 0.0, esum = 0.0, sum_thinout = 0.0;
    int NNE_thinout = 0;
    float pdf_thinout[1024], pe_thinout[1024];
    for (int i = 0; i < NNE; i++) {
        psum += pdfv[i];
        esum += photonenergy[i];
        if (i - nthin + 1 == (i - nthin + 1) / nthin * nthin) {
            pdf_thinout[(i - nthin + 1) / nthin] = (float)psum;
            pe_thinout[(i - nthin + 1) / nthin] = (float)(esum / nthin);
            sum_thinout += psum;
            psum = 0.0;
            esum = 0.0;
            NNE_thinout = (i - nthin + 1) / nthin + 1;
        }
    }
    NNE = NNE_thinout;
    sum = (float)sum_thinout;
    if (NNE > max_ne) return -1;
    for (int i = 0; i < NNE; i++) {
        pdfv[i] = pdf_thinout[i];
        photonenergy[i] = pe_thinout[i];
    }
    /* :236-244 */
    for (int i = 0; i < NNE; i++) pdfv[i] = pdfv[i] / sum;
    /* :258-295 */
    const double Ek = 0.0;
    for (int iz = 0; iz < M; iz++) {
        double Z = ZZ[iz], A = AA[iz];
        double fact = 6.0221 * Z / A * (2.81794 * 2.81794) * 0.001;
        for (int i = 0; i < NNE; i++) {
            double E = photonenergy[i] / 0.511;
            double photo = fact * 0.00000345 * (1.0 + 0.008 * Z) * (Z * Z * Z) / (E * E * E)
                           * (1.0 - Ek / (4.0 * E) - Ek * Ek / (1.21 * E));
            double compton = fact * 2 * ORC_PI * ((1.0 + E) / (E * E) * (2.0 * (1.0 + E) / (1.0 + 2 * E)
                                                   - log(1.0 + 2 * E) / E)
                                                  + log(1.0 + 2 * E) / (2 * E)
                                                  - (1.0 + 3 * E) / ((1.0 + 2 * E) * (1.0 + 2 * E)));
            double pair = fact * 0.2545 * (E - 2.332) * Z / 137;
            if (pair < 0) pair = 0;
            att[NNE * iz + i] = (float)(photo + compton + pair);
        }
    }
    for (int i = 0; i < NNE; i++) {
        pdf[i] = pdfv[i];
        if (energy_mev) energy_mev[i] = photonenergy[i];
    }
    return NNE;
}

/* --------------------------------------------------------------- forward */

typedef struct {
    double *cosT, *sinT, *px0, *py0;
} geo_tables;

static void tables_make(const orc_geom* g, geo_tables* t)
{
    t->cosT = (double*)malloc(sizeof(double) * g->V);
    t->sinT = (double*)malloc(sizeof(double) * g->V);
    t->px0 = (double*)malloc(sizeof(double) * g->W);
    t->py0 = (double*)malloc(sizeof(double) * g->W);
    orc_view_tables(g, t->cosT, t->sinT);
    orc_channel_table(g, t->px0, t->py0);
}

static void tables_free(geo_tables* t)
{
    free(t->cosT); free(t->sinT); free(t->px0); free(t->py0);
}

/* Views traced per sinogram: M/reprojection_TOMO.cpp:35-42 with thinout 1. */
static int views_traced(const orc_geom* g) { return g->V / 2 * 2; }

void orc_forward_ref(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE, float* out)
{
    /* M/main.cpp:492-506 around M/reprojection_TOMO.cpp:13-289.  The segment
     * list is independent of the energy bin, so each ray is traced once and
     * its segments replayed per bin in the reference's order (segment-major,
     * then material), which reproduces transp bit for bit. */
    geo_tables t;
    tables_make(g, &t);
    const size_t NN = (size_t)g->N * g->N;
    const int ms = max_segments(g);
    const int VT = views_traced(g);
    memset(out, 0, sizeof(float) * (size_t)g->V * g->W);
#pragma omp parallel
    {
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * ms);
        double* rr = (double*)malloc(sizeof(double) * ms);
#pragma omp for schedule(dynamic, 2)
        for (int v = 0; v < VT; v++)
            for (int c = 0; c < g->W; c++) {
                int iks;
                int ns = orc_trace_ray(g, t.cosT, t.sinT, t.px0, t.py0, v, c, vox, rr, ms, &iks);
                float full = 0.f;
                for (int ke = 0; ke < NE; ke++) {
                    double transp = 0.0;
                    for (int q = 0; q < ns; q++)
                        for (int mat = 0; mat < M; mat++) {
                            float rho = dens[mat * NN + vox[q]];
                            if (rho < 0) continue; /* ph = 0 */
                            transp += (double)(rho * att[mat * NE + ke]) * rr[q];
                        }
                    double intens = pdf[ke] * exp(-transp / g->xfactor);
                    full += (float)intens;
                }
                out[(size_t)v * g->W + c] = full;
            }
        free(vox);
        free(rr);
    }
    tables_free(&t);
}

void orc_forward_f64(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE, double* out,
                     double* Lpath)
{
    geo_tables t;
    tables_make(g, &t);
    const size_t NN = (size_t)g->N * g->N;
    const int ms = max_segments(g);
    const int VT = views_traced(g);
    memset(out, 0, sizeof(double) * (size_t)g->V * g->W);
#pragma omp parallel
    {
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * ms);
        double* rr = (double*)malloc(sizeof(double) * ms);
#pragma omp for schedule(dynamic, 2)
        for (int v = 0; v < VT; v++)
            for (int c = 0; c < g->W; c++) {
                int iks;
                int ns = orc_trace_ray(g, t.cosT, t.sinT, t.px0, t.py0, v, c, vox, rr, ms, &iks);
                double L[16] = {0};
                for (int q = 0; q < ns; q++)
                    for (int mat = 0; mat < M; mat++) {
                        double rho = dens[mat * NN + vox[q]];
                        if (rho < 0) continue;
                        L[mat] += rho * rr[q];
                    }
                double I = 0.0;
                for (int ke = 0; ke < NE; ke++) {
                    double e = 0.0;
                    for (int mat = 0; mat < M; mat++) e += (double)att[mat * NE + ke] * L[mat];
                    I += (double)pdf[ke] * exp(-e / g->xfactor);
                }
                out[(size_t)v * g->W + c] = I;
                if (Lpath)
                    for (int mat = 0; mat < M; mat++) Lpath[((size_t)v * g->W + c) * M + mat] = L[mat];
            }
        free(vox);
        free(rr);
    }
    tables_free(&t);
}

float orc_normalize(float* sino, int n)
{
    /* M/main.cpp:511-522 */
    float mx = 0;
    for (int i = 0; i < n; i++)
        if (mx < sino[i]) mx = sino[i];
    for (int i = 0; i < n; i++) sino[i] = sino[i] / mx;
    return mx;
}

static double uniform0(void) { return ((double)rand() + 1.0) / ((double)RAND_MAX + 2.0); }

void orc_gaussian_noise_glibc(float* p, int n, double sigma, int order)
{
    /* M/main.cpp:41-76; glibc rand() from its default state (seed 1). */
    srand(1);
    double projection_min = 10000000;
    for (int i = 0; i < n; i++)
        if (projection_min > p[i]) projection_min = p[i];
    for (int i = 0; i < n; i++) {
        double nstar = p[i];
        double u_log, u_sin;
        if (order == 0) { u_log = uniform0(); u_sin = uniform0(); }
        else            { u_sin = uniform0(); u_log = uniform0(); }
        double z = sqrt(-2.0 * log(u_log)) * sin(2.0 * ORC_PI * u_sin);
        double psn = 0 + sigma * z;
        nstar = nstar + psn;
        p[i] = (float)nstar;
        if (nstar < 0) p[i] = (float)(projection_min / 10);
        if (nstar > 1.0) p[i] = 1.0f;
    }
}

/* Philox4x32-10 from the published description (Salmon, Moraes, Dror, Shaw,
 * "Parallel random numbers: as easy as 1, 2, 3", SC'11). */
void orc_philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; r++) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_beam_hardening(float* p, int n, double alpha, double beta, double* min_org, double* min_new)
{
    /* M/beam_hardening_correction.cpp:47-60 */
    double min_nii = 1000000, min_o = 1000000;
    for (int i = 0; i < n; i++) {
        if (min_o > p[i]) min_o = p[i];
        /* log() of a float is the C++ float overload there, i.e. logf */
        double yi = -logf(p[i]);
        yi = fabs(yi) + (alpha * pow(fabs(yi), beta));
        p[i] = (float)exp(-yi);
        if (min_nii > p[i]) min_nii = p[i];
    }
    if (min_org) *min_org = min_o;
    if (min_new) *min_new = min_nii;
}

/* ------------------------------------------------------------------- FBP */

void orc_const_map(const orc_geom* g, int filter, double* C)
{
    /* F/filterSinogramTOMO.cpp:9-138 */
    const int W = g->W;
    const int half = (W - 1) / 2;
    const double arc = g->max_beta / half;
    double* sum = (double*)calloc(W, sizeof(double));
    memset(C, 0, sizeof(double) * (size_t)W * W);
    for (int m = 0; m < W; m++) {
        sum[m] = 0.0;
        for (int n = 0; n < W; n++) {
            int MN = m - n;
            double v = 0.0;
            if (m == n) {
                if (filter == 1) v = 0.125 / (arc * arc);
                else if (filter == 2) {
                    double oddterm = ORC_PI * ORC_PI * (arc);
                    v = 0.25 / oddterm * (1.0 / tan((0.5) * arc) - 1.0 / tan((-0.5) * arc));
                }
            } else if (MN % 2 != 0) {
                if (filter == 1) {
                    double oddterm = ORC_PI * sin(MN * arc);
                    v = -0.5 / oddterm / oddterm;
                } else if (filter == 2) {
                    double oddterm = ORC_PI * ORC_PI * arc;
                    v = 0.25 / oddterm * (1.0 / tan((MN + 0.5) * arc) - 1.0 / tan((MN - 0.5) * arc));
                }
            } else {
                if (filter == 1) v = 0.0;
                else if (filter == 2) {
                    double oddterm = ORC_PI * ORC_PI * arc;
                    v = 0.25 / oddterm * (1.0 / tan((MN + 0.5) * arc) - 1.0 / tan((MN - 0.5) * arc));
                }
            }
            C[(size_t)m * W + n] = v;
            sum[m] += v;
        }
    }
    /* :114-123 edge-row renormalisation */
    for (int m = 0; m < W; m++)
        for (int n = 0; n < W; n++)
            if ((0 <= m && m <= 20) || (W - 20 <= m && m <= W))
                C[(size_t)m * W + n] = C[(size_t)m * W + n] * sum[21] / sum[m];
    free(sum);
}

void orc_filter_ref(const orc_geom* g, const double* C, const float* proj, float* fltd, int in_mode)
{
    /* F/filterSinogramTOMO.cpp:225-256 */
    const int W = g->W;
    const int half = (W - 1) / 2;
    const int cent = W / 2;
    double* coeff = (double*)malloc(sizeof(double) * W);
    for (int k = 0; k < W; k++) coeff[k] = g->sid * cos((k - cent) * (g->max_beta / half));
#pragma omp parallel for schedule(static)
    for (int n = 0; n < g->V; n++) {
        for (int m = 0; m < W; m++) {
            float acc = 0.f;
            for (int k = 0; k < W; k++) {
                double projdat = proj[(size_t)n * W + k];
                acc = (float)((double)acc + C[(size_t)m * W + k] * (-log(projdat) * coeff[k]) * g->max_beta / half);
            }
            if (in_mode && (acc < 1000.0 || acc > 300000.0)) acc = 0.f;
            fltd[(size_t)n * W + m] = acc;
        }
    }
    free(coeff);
}

void orc_filter_f64(const orc_geom* g, const double* C, const float* proj, double* fltd)
{
    const int W = g->W;
    const int half = (W - 1) / 2;
    const int cent = W / 2;
    double* q = (double*)malloc(sizeof(double) * (size_t)g->V * W);
    for (int n = 0; n < g->V; n++)
        for (int k = 0; k < W; k++)
            q[(size_t)n * W + k] = -log((double)proj[(size_t)n * W + k]) * (g->sid * cos((k - cent) * (g->max_beta / half)));
#pragma omp parallel for schedule(static)
    for (int n = 0; n < g->V; n++)
        for (int m = 0; m < W; m++) {
            double acc = 0.0;
            for (int k = 0; k < W; k++) acc += C[(size_t)m * W + k] * q[(size_t)n * W + k];
            fltd[(size_t)n * W + m] = acc * g->max_beta / half;
        }
    free(q);
}

void orc_mask(int N, uint8_t* mask)
{
    /* F/misc.cpp:9-37, M/misc_TOMO.cpp:10-39 */
    double maskRadius = (double)(N - 1) / 2.0;
    double sq = maskRadius * maskRadius;
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++)
            mask[j * N + i] = (((double)j - maskRadius) * ((double)j - maskRadius) +
                               ((double)i - maskRadius) * ((double)i - maskRadius) <= sq) ? 1 : 0;
}

void orc_backproject_ref(const orc_geom* g, const float* fltd, float* img)
{
    /* F/reconstructImageTOMO.cpp:11-105 (single slice; z terms are dead) */
    const int N = g->N, W = g->W, V = g->V;
    const int half = (W - 1) / 2;
    const int cent = W / 2;
    const int use = (int)(V / 2 * 2);
    double* cosT = (double*)malloc(sizeof(double) * V);
    double* sinT = (double*)malloc(sizeof(double) * V);
    uint8_t* mask = (uint8_t*)malloc((size_t)N * N);
    orc_view_tables(g, cosT, sinT);
    orc_mask(N, mask);
    const double centerPos = (double)N / 2.0 - 0.5;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            float acc = 0.f;
            if (mask[y * N + x] != 0) {
                double xx = -(x - centerPos) * g->pix;
                double yy = -(y - centerPos) * g->pix;
                for (int th = 0; th < use; th++) {
                    double dbeta = 2.0 * ORC_PI / (V);
                    double X = xx * cosT[th] + yy * sinT[th];
                    double Y = -xx * sinT[th] + yy * cosT[th];
                    double L2 = (-Y + g->sid) * (-Y + g->sid) + X * X;
                    double fGAM = atan(X / (g->sid - Y)) / (g->max_beta / half) + cent;
                    int GAM = trunc_i32(fGAM);
                    double w = fGAM - GAM;
                    if (GAM >= 0 && GAM < W - 1) {
                        double sum1 = (1 - w) * fltd[(size_t)th * W + GAM] + w * fltd[(size_t)th * W + GAM + 1];
                        double sum = sum1 + 0.0;
                        acc += (float)(1.0 / L2 * sum * dbeta * g->xfactor);
                    }
                }
            }
            img[y * N + x] = acc;
        }
    free(cosT); free(sinT); free(mask);
}

void orc_backproject_f64(const orc_geom* g, const double* fltd, double* img)
{
    const int N = g->N, W = g->W, V = g->V;
    const int half = (W - 1) / 2;
    const int cent = W / 2;
    const int use = (int)(V / 2 * 2);
    double* cosT = (double*)malloc(sizeof(double) * V);
    double* sinT = (double*)malloc(sizeof(double) * V);
    uint8_t* mask = (uint8_t*)malloc((size_t)N * N);
    orc_view_tables(g, cosT, sinT);
    orc_mask(N, mask);
    const double centerPos = (double)N / 2.0 - 0.5;
#pragma omp parallel for schedule(dynamic, 4)
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            double acc = 0.0;
            if (mask[y * N + x] != 0) {
                double xx = -(x - centerPos) * g->pix;
                double yy = -(y - centerPos) * g->pix;
                for (int th = 0; th < use; th++) {
                    double dbeta = 2.0 * ORC_PI / (V);
                    double X = xx * cosT[th] + yy * sinT[th];
                    double Y = -xx * sinT[th] + yy * cosT[th];
                    double L2 = (-Y + g->sid) * (-Y + g->sid) + X * X;
                    double fGAM = atan(X / (g->sid - Y)) / (g->max_beta / half) + cent;
                    int GAM = trunc_i32(fGAM);
                    double w = fGAM - GAM;
                    if (GAM >= 0 && GAM < W - 1) {
                        double sum = (1 - w) * fltd[(size_t)th * W + GAM] + w * fltd[(size_t)th * W + GAM + 1];
                        acc += 1.0 / L2 * sum * dbeta * g->xfactor;
                    }
                }
            }
            img[y * N + x] = acc;
        }
    free(cosT); free(sinT); free(mask);
}

/* --------------------------------------------------------------- adjoint */

void orc_adjoint_f64(const orc_geom* g, const float* dens, int M,
                     const float* att, const float* pdf, int NE,
                     const double* gI, double* grad)
{
    geo_tables t;
    tables_make(g, &t);
    const size_t NN = (size_t)g->N * g->N;
    const int ms = max_segments(g);
    const int VT = views_traced(g);
    memset(grad, 0, sizeof(double) * NN * M);
    int nthreads = 1;
#ifdef _OPENMP
    nthreads = omp_get_max_threads();
#endif
    double* priv = (double*)calloc((size_t)nthreads * NN * M, sizeof(double));
#pragma omp parallel
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* my = priv + (size_t)tid * NN * M;
        int32_t* vox = (int32_t*)malloc(sizeof(int32_t) * ms);
        double* rr = (double*)malloc(sizeof(double) * ms);
#pragma omp for schedule(static)
        for (int v = 0; v < VT; v++)
            for (int c = 0; c < g->W; c++) {
                double gr = gI[(size_t)v * g->W + c];
                if (gr == 0.0) continue;
                int iks;
                int ns = orc_trace_ray(g, t.cosT, t.sinT, t.px0, t.py0, v, c, vox, rr, ms, &iks);
                double L[16] = {0};
                for (int q = 0; q < ns; q++)
                    for (int mat = 0; mat < M; mat++) {
                        double rho = dens[mat * NN + vox[q]];
                        if (rho < 0) continue;
                        L[mat] += rho * rr[q];
                    }
                double dIdL[16] = {0};
                for (int ke = 0; ke < NE; ke++) {
                    double e = 0.0;
                    for (int mat = 0; mat < M; mat++) e += (double)att[mat * NE + ke] * L[mat];
                    double tE = (double)pdf[ke] * exp(-e / g->xfactor);
                    for (int mat = 0; mat < M; mat++) dIdL[mat] -= tE * (double)att[mat * NE + ke] / g->xfactor;
                }
                for (int q = 0; q < ns; q++)
                    for (int mat = 0; mat < M; mat++) {
                        if (dens[mat * NN + vox[q]] < 0) continue;
                        my[(size_t)mat * NN + vox[q]] += gr * dIdL[mat] * rr[q];
                    }
            }
        free(vox);
        free(rr);
    }
    for (int tI = 0; tI < nthreads; tI++) {
        const double* p = priv + (size_t)tI * NN * M;
        for (size_t i = 0; i < NN * M; i++) grad[i] += p[i];
    }
    free(priv);
    tables_free(&t);
}

/* ------------------------------------------------------------ npf (lengthproj) */

/* View window of M/lengthproj_TOMO.cpp:35-49 (UsedReconAngle = 2,
 * M/physParamsTOMO.h:67): start = (int)(z/pitch - V/2 + 1e-8), V/2*2 views,
 * view index start + j - (start/V)*V, wrapped once.  Only start >= 0 keeps
 * every index inside [0, V). */
static int lp_window(int V, double z_cm, double couch_cm, int* base, int* count)
{
    double pitch = (double)(couch_cm / V);
    double s = z_cm / pitch - V / 2 + 0.00000001;
    if (!(s > -1.0) || !(s < 2147483647.0 - V)) return -1;
    int start = (int)s;
    *base = start - (start / V) * V;
    *count = V / 2 * 2;
    return 0;
}

/* One pixel-view term np * rr of M/lengthproj_TOMO.cpp:72-167 in the
 * reference's operation order; returns 0 when the reference adds nothing
 * (a NaN channel position). */
static int lp_term(const orc_geom* g, double rx, double ry, double rz, double c, double s,
                   const float* row, int pw, double iso_det, double* out)
{
    const double pix = g->pix;
    const double sx = g->sid * s, sy = g->sid * c, sz = rz;
    const double yslope = fabs(ry - sy), xslope = fabs(rx - sx);
    double rr;
    if (xslope > yslope) {
        double xx0 = rx - 0.5 * pix;
        double yy0 = (xx0 - sx) / (rx - sx) * (ry - sy) + sy;
        double zz0 = (xx0 - sx) / (rx - sx) * (rz - sz) + sz;
        double xx1 = rx + 0.5 * pix;
        double yy1 = (xx1 - sx) / (rx - sx) * (ry - sy) + sy;
        double zz1 = (xx1 - sx) / (rx - sx) * (rz - sz) + sz;
        rr = sqrt((xx0 - xx1) * (xx0 - xx1) + (yy0 - yy1) * (yy0 - yy1) + (zz0 - zz1) * (zz0 - zz1));
    } else {
        /* :94-101: the far end reuses yy0 (so xx1 == xx0 and rr = pix) */
        double yy0 = ry - 0.5 * pix;
        double xx0 = (yy0 - sy) / (ry - sy) * (rx - sx) + sx;
        double zz0 = (yy0 - sy) / (ry - sy) * (rz - sz) + sz;
        double yy1 = ry + 0.5 * pix;
        double xx1 = (yy0 - sy) / (ry - sy) * (rx - sx) + sx;
        double zz1 = (yy0 - sy) / (ry - sy) * (rz - sz) + sz;
        rr = sqrt((xx0 - xx1) * (xx0 - xx1) + (yy0 - yy1) * (yy0 - yy1) + (zz0 - zz1) * (zz0 - zz1));
    }
    /* detector hit on the arc of radius R about the source, :109-131 */
    const double dx = g->det_src_iso * s, dy = g->det_src_iso * c;
    const double R = g->det_radius;
    const double aa = (sy - ry) / (sx - rx);
    const double bb = -sx * (sy - ry) / (sx - rx) + sy;
    const double A = 1.0 + aa * aa;
    const double B = 2.0 * aa * bb - 2.0 * aa * dy - 2.0 * dx;
    const double C = dx * dx + bb * bb + dy * dy - 2.0 * bb * dy - R * R;
    const double px1 = (-B + sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
    const double py1 = aa * px1 + bb;
    const double px2 = (-B - sqrt(B * B - 4.0 * A * C)) / (2.0 * A);
    const double py2 = aa * px2 + bb;
    double px, py;
    if (px1 * px1 + py1 * py1 > px2 * px2 + py2 * py2) { px = px2; py = py2; }
    else { px = px1; py = py1; }
    /* channel position, :133-150 */
    const double ox = -iso_det * s, oy = -iso_det * c;
    const double dis = (px - ox) * (px - ox) + (py - oy) * (py - oy);
    const double angle = acos(1.0 - dis / (2.0 * R * R));
    const double rot_px = px * c - py * s;
    const double delta = g->max_alpha / ((g->W - 1) / 2);
    const double dj = rot_px > 0.0 ? g->iso_channel - angle / delta : g->iso_channel + angle / delta;
    /* the row holds the first pw channels of the W-channel detector (the
     * reference's End_Channel is fixed at 609 and the row stride is the
     * measured width; hits past the row's last channel take that channel,
     * where the reference reads into the next row) */
    const int last = pw - 1;
    double np;
    if (0 <= dj && dj < last) {   /* :152-158 */
        int i0 = (int)dj;
        double w = dj - (double)i0;
        np = (1.0 - w) * row[i0] + w * row[i0 + 1];
    } else if (0 > dj) {
        np = row[0];
    } else if (dj >= last) {
        np = row[last];
    } else {
        return 0;
    }
    *out = np * rr;
    return 1;
}

#define LP_BODY(ACC_T, ADD)                                                                 \
    int base, count;                                                                        \
    if (pw < 2 || pw > g->W) return -1;                                                     \
    if (lp_window(g->V, z_cm, couch_cm, &base, &count) != 0) return -1;                     \
    const int N = g->N, V = g->V, W = g->W;                                                 \
    double* cosT = (double*)malloc(sizeof(double) * V);                                     \
    double* sinT = (double*)malloc(sizeof(double) * V);                                     \
    uint8_t* mask = (uint8_t*)malloc((size_t)N * N);                                        \
    orc_view_tables(g, cosT, sinT);                                                         \
    orc_mask(N, mask);                                                                      \
    _Pragma("omp parallel for schedule(dynamic, 4)")                                        \
    for (int jy = 0; jy < N; jy++) {                                                        \
        double ry = N * g->pix * 0.5 - (jy + 0.5) * g->pix;                                 \
        for (int jx = 0; jx < N; jx++) {                                                    \
            ACC_T acc = 0;                                                                  \
            if (mask[jy * N + jx] != 0) {                                                   \
                double rx = -N * g->pix * 0.5 + (jx + 0.5) * g->pix;                        \
                for (int j = 0; j < count; j++) {                                           \
                    int jr = base + j;                                                      \
                    if (jr >= V) jr -= V;                                                   \
                    double t;                                                               \
                    if (lp_term(g, rx, ry, z_cm, cosT[jr], sinT[jr], proj + (size_t)j * pw, \
                                pw, iso_det, &t))                                           \
                        ADD;                                                                \
                }                                                                           \
            }                                                                               \
            npf[jy * N + jx] = acc;                                                         \
        }                                                                                   \
    }                                                                                       \
    free(cosT);                                                                             \
    free(sinT);                                                                             \
    free(mask);                                                                             \
    return 0;

int orc_lengthproj_window_ref(const orc_geom* g, const float* proj, int pw, double z_cm, double couch_cm,
                              double iso_det, float* npf)
{
    /* M/lengthproj_TOMO.cpp:14-185: float accumulator, views ascending */
    LP_BODY(float, acc += (float)t)
}

int orc_lengthproj_window_f64(const orc_geom* g, const float* proj, int pw, double z_cm, double couch_cm,
                              double iso_det, double* npf)
{
    LP_BODY(double, acc += t)
}

int orc_lengthproj_ref(const orc_geom* g, const float* proj, double z_cm, double couch_cm,
                       double iso_det, float* npf)
{
    return orc_lengthproj_window_ref(g, proj, g->W, z_cm, couch_cm, iso_det, npf);
}

int orc_lengthproj_f64(const orc_geom* g, const float* proj, double z_cm, double couch_cm,
                       double iso_det, double* npf)
{
    return orc_lengthproj_window_f64(g, proj, g->W, z_cm, couch_cm, iso_det, npf);
}
#undef LP_BODY

/* ------------------------------------------------------ measured-data loader */

enum { ORC_NC = 609, ORC_NS = 27, ORC_AH = 800, ORC_AN = 5 };

/* M/projectionDataTOMO.cpp:380-422 (air scans), :434-496 (sinogram). */
int orc_load_measured(int in_mode, int rot, int H, int W, int nskip, const float* air, const float* raw,
                      float* out, uint8_t* mask, double* max_proj)
{
    const size_t blk = (size_t)ORC_NC * ORC_AH;
    if (rot < 1 || H < 1 || W < 1 || nskip < 1 || W + ORC_NS > ORC_NC) return -1;
    memset(out, 0, sizeof(float) * (size_t)rot * H * W);
    memset(mask, 0, (size_t)rot * H * W);
    *max_proj = 0.0;
    if (in_mode == 0) {
        /* PR: rows k1 with k1 % nskip == 0 are kept, packed rotation after
         * rotation (:434-458) */
        if ((ORC_AH + nskip - 1) / nskip > H) return -1;
        float* avg = (float*)calloc(blk, sizeof(float));
        for (int b = 1; b <= ORC_AN; b++)
            for (size_t k = 0; k < blk; k++) avg[k] += air[b * blk + k] / ORC_AN;
        size_t kkk = 0;
        for (int t = 0; t < rot; t++)
            for (int k1 = 0; k1 < ORC_AH; k1++)
                for (int k2 = ORC_NS; k2 < ORC_NS + W; k2++) {
                    const size_t kk = (size_t)k1 * ORC_NC + k2;
                    const float d = raw[t * blk + kk] / avg[kk];
                    if (k1 / nskip * nskip == k1) {
                        out[kkk] = d;
                        mask[kkk] = 1;
                        kkk++;
                    }
                }
        free(avg);
        return 0;
    }
    /* IN: the air average is [H][W] (:404-417); each sinogram row is read into
     * the first 609 floats of a zeroed 800 x 609 buffer that is then indexed
     * at row j (:463-470), so only row 0 of each rotation holds data; values
     * below 0.1 are zeroed and unmasked (:486-491) */
    if (H > ORC_AH) return -1;
    float* avg = (float*)calloc((size_t)H * W, sizeof(float));
    for (int b = 1; b <= ORC_AN; b++) {
        size_t k = 0;
        for (int k1 = 0; k1 < H; k1++)
            for (int k2 = ORC_NS; k2 < ORC_NS + W; k2++) avg[k++] += air[b * blk + (size_t)k1 * ORC_NC + k2] / ORC_AN;
    }
    double pmax = 0.0;
    for (int t = 0; t < rot; t++)
        for (int j = 0; j < H; j++)
            for (int c = 0; c < W; c++) {
                const int i = ORC_NS + c;
                const float x = j == 0 ? raw[((size_t)t * H) * ORC_NC + i] : 0.0f;
                float d = x / (avg[c] * 2);
                if (pmax < d && i >= 0 && i < ORC_NC) pmax = d;
                uint8_t m = 1;
                if (d < 0.1) {
                    d = 0.0f;
                    m = 0;
                }
                const size_t o = ((size_t)t * H + j) * W + c;
                out[o] = d;
                mask[o] = m;
            }
    free(avg);
    *max_proj = pmax;
    return 0;
}

/* M/main.cpp:352-371 (start view 0, 360-degree window). */
void orc_helical_window(const float* volume, int rot, int H, int W, float* window)
{
    const int used = 2;
    const int end = H / used * 2, use = H / used * 2;
    for (int j = 0; j < end; j++) {
        const int n = j;
        int iphase = 0;
        if (n < use / 2) iphase = 1;
        else if (n > use / 2) iphase = -1;
        if (j + iphase * H < H) iphase = 0;
        if (j + iphase * H > (rot - 1) * H) iphase = 0;
        const double zw = 0.5 * fabs(n - (H / 2.0)) / (H / 2.0);
        for (int c = 0; c < W; c++)
            window[(size_t)n * W + c] =
                (float)((1.0 - zw) * volume[(size_t)j * W + c] + zw * volume[(size_t)(j + iphase * H) * W + c]);
    }
}
