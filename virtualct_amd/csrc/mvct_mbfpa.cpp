// mvct_mbfpa.cpp -- drop-in replacement for the MBFPA mvct.exe
// (M/main.cpp:214-675) running the forward projection on the GPU via libvct.
//
//   mvct_mbfpa <input.txt>
//
// Same argv, the same positional input-txt (M/projectionDataTOMO.cpp:68-135),
// the same files read from and written to the working directory, and the same
// summary lines on stdout/stderr:
//   reads  <input.txt>, the spectrum file it names, Weight_input_ICRU110.raw
//          (int16 [M][N][N], mg/cm^3; M/main.cpp:421-429)
//   writes reprojection_float.raw (float32 [V][W], max-normalised with
//          Gaussian detector noise; M/main.cpp:508-530) and npf_float.raw
//          (M/main.cpp:407-419: zeros when the input sinogram is "*")
// With the detector and air-scan files the txt names present, npf_float.raw
// is computed from them as the reference does (air-scan normalisation, PR or
// IN, M/projectionDataTOMO.cpp:323-524, which also writes projections.raw;
// helical window, M/main.cpp:352-371; npf for slice z = reconstruction
// start z); the reference's "Maximum projection ... STOP" ends the run with
// status 0 as there.
// Environment overrides (argv is unchanged): VCT_PHANTOM (density file),
// VCT_NUM_MATERIAL (6 or 8), VCT_SEED (noise seed), VCT_DEVICE, and
// VCT_NPF_SINOGRAM: a normalised float32 [V/2*2][W] sinogram window used for
// npf_float.raw in place of the measured files.
// The iterative-reconstruction leftovers (ReconImageRaw/, M/main.cpp:558-624)
// are not produced.  Exit status is 0 on success and on the reference's own
// soft failures (missing info file, usage), 1 on a library error.
#include <sys/stat.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vct.h"
#include "vct_txt.h"

namespace {

struct MbfpaInput {
    std::string inpr, dtype, proj_file, air_file, tgt, recon_file, spectrum;
    int prior = 0, width = 0, height = 0, offset = 0, recon_size = 0, slices = 0, xfactor = 1, fbp_slice = 0;
    int osem = 1, num_ite = 0, sino_fact = 1;
    double period = 0, rate = 0, start_angle = 0, couch_speed = 0, couch_start = 0, pix = 0;
    double recon_offset = 0, recon_step = 0, xr = 0, yr = 0, zr = 0, scan_z = 0, tv = 0, pc = 0, mc = 0, gd = 0;
    double noise = 0;
};

bool load(const char* path, MbfpaInput& in)
{
    vct::TxtReader r(path);
    if (!r.ok()) {
        std::fprintf(stderr, "info file: %s not found\n", path);
        return false;
    }
    r.skip(6);
    r.read_string(&in.inpr);
    r.read_int(&in.prior);
    r.read_int(&in.width);
    r.read_int(&in.height);
    r.read_string(&in.dtype);
    r.read_int(&in.offset);
    r.read_string(&in.proj_file);
    r.read_string(&in.air_file);
    r.read_double(&in.period);
    r.read_double(&in.rate);
    r.read_double(&in.start_angle);
    r.read_double(&in.couch_speed);
    r.read_double(&in.couch_start);
    r.read_int(&in.recon_size);
    r.read_double(&in.pix);
    r.read_string(&in.tgt);
    r.read_double(&in.recon_offset);
    r.read_double(&in.recon_step);
    r.read_int(&in.slices);
    r.read_int(&in.xfactor);
    r.read_string(&in.recon_file);
    r.read_int(&in.fbp_slice);
    r.read_double3(&in.xr, &in.yr, &in.zr);
    r.read_double(&in.scan_z);
    r.read_int(&in.osem);
    r.read_double3(&in.tv, &in.pc, &in.mc);
    r.read_double(&in.gd);
    r.read_int(&in.num_ite);
    r.read_int(&in.sino_fact);
    r.read_string(&in.spectrum);
    r.read_double(&in.noise);
    return true;
}

int fail(const char* what, vct_status st)
{
    std::fprintf(stderr, "ERROR: %s: %s (%s)\n", what, vct_status_string(st), vct_last_error());
    return 1;
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc != 2) {
        std::fprintf(stderr, "\n[usage]$ mvct <info filename>\n\n");
        return 0;
    }
    MbfpaInput in;
    if (!load(argv[1], in)) return 0;

    // derived sinogram height and angles, M/projectionDataTOMO.cpp:126-134, :332-333
    const double rate_orig = in.rate;
    double rate = in.rate;
    if (rate >= 200) rate = rate / 5;
    const int total_proj = (int)(in.height / (in.period * rate_orig));
    int H = (int)(in.period * rate);
    if (in.sino_fact != 0) H = H / in.sino_fact;
    const int W = in.width;
    if (in.prior == 1 || in.prior == 2) {
        std::fprintf(stderr, "ERROR: prior images (DICOM / MVCT) are not supported by this program\n");
        return 1;
    }
    std::fprintf(stderr, "%s\n", in.inpr == "PR" ? "Pre-treatment CT " : "In-treatment CT ");
    bool have_sino = false;
    if (FILE* fp = std::fopen(in.proj_file.c_str(), "rb")) {
        have_sino = true;
        std::fclose(fp);
    }
    if (!have_sino) std::fprintf(stderr, "data file: %s not found\n", in.proj_file.c_str());
    if (FILE* fp = std::fopen(in.recon_file.c_str(), "rb")) {
        std::fprintf(stderr, "Prior image: %s \n", in.recon_file.c_str());
        std::fclose(fp);
    } else {
        std::fprintf(stderr, "No prior image \n");
    }

    const char* env_m = std::getenv("VCT_NUM_MATERIAL");
    const int M = env_m ? std::atoi(env_m) : 6;
    const char* env_ph = std::getenv("VCT_PHANTOM");
    const std::string phantom_file = env_ph ? env_ph : "Weight_input_ICRU110.raw";
    const char* env_seed = std::getenv("VCT_SEED");
    const uint64_t seed = env_seed ? std::strtoull(env_seed, nullptr, 0) : 0x5EEDull;
    const char* env_dev = std::getenv("VCT_DEVICE");
    const int device = env_dev ? std::atoi(env_dev) : 0;
    if (M < 1 || M > 8) {
        std::fprintf(stderr, "ERROR: VCT_NUM_MATERIAL must be 1..8\n");
        return 1;
    }

    std::fprintf(stderr, "usedProj: %d SngrmH: %d SngrmW: %d\n", total_proj, H, W);
    const int used_recon_angle = 2;
    const double s_start = in.couch_start + in.couch_speed / used_recon_angle;
    const double s_end = in.couch_start + (total_proj + 1) * in.couch_speed - in.couch_speed / used_recon_angle;
    const double zstart = in.recon_offset;
    const double zend = in.recon_offset + (in.slices - 1) * in.recon_step;
    std::fprintf(stderr, "Allowable reconstraction range = %lf to %lf \n", s_start, s_end);
    std::fprintf(stderr, "Reconstructing z = %lf to %lf, Step = %lf [cm] ... \n", zstart, zend, in.recon_step);

    // spectrum -> pdf and mass attenuation, M/main.cpp:335-343
    std::vector<float> att((size_t)M * 1024), pdf(1024);
    std::vector<double> energy(1024);
    int NE = 0;
    vct_status st = vct_attenuation_coefficient(in.spectrum.c_str(), M, nullptr, nullptr, 1, att.data(), pdf.data(),
                                                &NE, energy.data());
    if (st != VCT_OK) {
        std::fprintf(stderr, "ERROR: no photon energy file \n");
        return fail("attenuation_coefficient", st);
    }
    {
        float norm = 0.f, mean = 0.f;
        for (int i = 0; i < NE; i++) {
            norm += pdf[i];
            mean += (float)energy[i] * pdf[i];
            std::printf("%d %f %f %f \n", i, energy[i], pdf[i], norm);
        }
        std::printf("mean x-ray energy = %f \n", mean);
    }
    std::printf("%d \n", H / used_recon_angle * 2);
    std::fprintf(stderr, "Iteration loop, %d step,\t\n", 0);

    vct_geometry g;
    vct_geometry_default(&g);
    g.recon_size = in.recon_size;
    g.pixel_cm = in.pix;
    g.n_channels = W;
    g.n_views = H;
    g.start_view_deg = in.start_angle;
    g.xfactor = in.xfactor;

    vct_ctx* ctx = nullptr;
    st = vct_init(device, &ctx);
    if (st != VCT_OK) return fail("vct_init", st);

    // npf side output, M/main.cpp:352-419: the measured volume of the txt's
    // detector and air-scan files (loadProjectionData, which loadData calls,
    // M/projectionDataTOMO.cpp:273, :323-524), its helical window (:352-371)
    // and the chord-weighted backprojection; zeros when either file is
    // missing (the reference's "*"); or the window VCT_NPF_SINOGRAM names
    const size_t NN = (size_t)in.recon_size * in.recon_size;
    const char* env_npf = std::getenv("VCT_NPF_SINOGRAM");
    const int rows = H / used_recon_angle * 2;
    std::vector<float> npf(NN, 0.f), win;
    auto stop = [&](const char* what, vct_status s2) {
        const int rc = fail(what, s2);  // before vct_terminate clears the message
        vct_terminate(ctx);
        return rc;
    };
    if (env_npf) {
        win.resize((size_t)rows * W);
        FILE* fp = std::fopen(env_npf, "rb");
        const size_t got = fp ? std::fread(win.data(), sizeof(float), win.size(), fp) : 0;
        if (fp) std::fclose(fp);
        if (got != win.size()) {
            std::fprintf(stderr, "ERROR: %s must hold float32 [%d][%d]\n", env_npf, rows, W);
            vct_terminate(ctx);
            return 1;
        }
    } else if (have_sino) {
        bool have_air = false;
        if (FILE* fp = std::fopen(in.air_file.c_str(), "rb")) {
            have_air = true;
            std::fclose(fp);
        } else {
            std::fprintf(stderr, "data file: %s not found\n", in.air_file.c_str());
        }
        if (have_air) {
            const bool in_mode = in.inpr != "PR";
            vct_measured m{in_mode ? 1 : 0, total_proj, H, W, in.sino_fact > 0 ? in.sino_fact : 1,
                           in_mode && rate_orig >= 200 ? 4 : 0};
            std::vector<float> vol((size_t)total_proj * H * W);
            std::vector<uint8_t> vmask(vol.size());
            double pmax = 0.0;
            if (in_mode) std::fprintf(stderr, "%d %d %d %d\n", total_proj, H, W, (int)(609 * sizeof(float)));
            st = vct_load_measured_host(ctx, &m, in.proj_file.c_str(), in.air_file.c_str(), vol.data(), vmask.data(),
                                        &pmax);
            if (st == VCT_ERR_ARG && in_mode && pmax > 1.0) {
                // the reference's own stop (M/projectionDataTOMO.cpp:479-483): exit status 0
                std::fprintf(stderr, "Maximum projection: %lf > 1.0... STOP \n", pmax);
                vct_terminate(ctx);
                return 0;
            }
            if (st != VCT_OK) return stop("load_measured", st);
            if (!in_mode)
                for (int t = 0; t < total_proj; t++)
                    std::fprintf(stderr, "Z scan step: %lf \n", in.couch_start + t * in.couch_speed);
            std::fprintf(stderr, "Maximum projection: %lf \n", pmax);
            if (FILE* fp = std::fopen("projections.raw", "wb")) {  // :513-521
                std::fwrite(vol.data(), sizeof(float), vol.size(), fp);
                std::fclose(fp);
            }
            win.resize((size_t)rows * W);
            st = vct_helical_window_host(ctx, vol.data(), total_proj, H, W, win.data());
            if (st != VCT_OK) return stop("helical_window", st);
        }
    }
    if (!win.empty()) {
        // the window keeps W channels of the 609-channel detector, whose pitch
        // and isocentre place the hits (End_Channel, isocenter_channel,
        // M/physParamsTOMO.h:43-47)
        vct_geometry gd = g;
        vct_geometry det;
        vct_geometry_default(&det);
        gd.n_channels = std::max(W, det.n_channels);
        st = vct_lengthproj_window_host(ctx, &gd, win.data(), W, zstart, in.couch_speed, 57.6, npf.data());
        if (st != VCT_OK) return stop("lengthproj", st);
    }
    if (FILE* fp = std::fopen("npf_float.raw", "wb")) {
        std::fwrite(npf.data(), sizeof(float), NN, fp);
        std::fclose(fp);
    }

    // densities, M/main.cpp:421-429 and initial_image_def :93-109
    std::vector<int16_t> ph((size_t)M * NN, 0);
    if (FILE* fp = std::fopen(phantom_file.c_str(), "rb")) {
        size_t got = std::fread(ph.data(), sizeof(int16_t), ph.size(), fp);
        std::fclose(fp);
        if (got != ph.size()) std::fprintf(stderr, "WARNING: %s is short (%zu of %zu values)\n", phantom_file.c_str(), got, ph.size());
    } else {
        std::fprintf(stderr, "ERROR: cannot open %s\n", phantom_file.c_str());
        vct_terminate(ctx);
        return 1;
    }
    std::vector<float> dens(ph.size());
    for (size_t i = 0; i < ph.size(); i++) dens[i] = (float)ph[i] / 1000;
    std::fprintf(stderr, "Iteration loop, %d step,\t\n", 1);

    std::vector<float> sino((size_t)W * H, 0.f);
    st = vct_reprojection(ctx, &g, dens.data(), M, att.data(), pdf.data(), NE, sino.data());
    if (st != VCT_OK) return fail("vct_reprojection", st);
    float stats[2] = {0.f, 0.f};
    st = vct_gaussian_noise(ctx, sino.data(), W * H, in.noise, seed, stats);
    if (st != VCT_OK) return fail("vct_gaussian_noise", st);
    std::printf("projection_min = %lf, used noise = %lf\n", (double)(stats[1] / stats[0]), in.noise);
    if (FILE* fp = std::fopen("reprojection_float.raw", "wb")) {
        std::fwrite(sino.data(), sizeof(float), sino.size(), fp);
        std::fclose(fp);
    } else {
        std::fprintf(stderr, "ERROR: cannot write reprojection_float.raw\n");
        vct_terminate(ctx);
        return 1;
    }
    vct_terminate(ctx);
    std::fprintf(stderr, "done.\n");
    return 0;
}
