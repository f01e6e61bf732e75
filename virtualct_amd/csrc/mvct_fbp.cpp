// mvct_fbp.cpp -- drop-in replacement for the FBP mvct.exe
// (F/main.cpp:31-227) running the filter and backprojection on the GPU.
//
//   mvct_fbp <input.txt>
//
// Same argv and positional input-txt (F/projectionDataTOMO.cpp:55-121):
// reads the float32 sinogram the txt names (PR: plain copy; IN: values below
// 50000 become 1, :250-292), writes the float32 image [slices][N][N] to the
// txt-named output (F/main.cpp:53, :187-190; every slice is the same, as the
// reference's z_image is 0), and prints the reference's summary lines.
// Environment: VCT_DEVICE.  Exit status 0 on success and on the reference's
// soft failures, 1 on a library error or an unsupported data type.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "vct.h"
#include "vct_txt.h"

namespace {

struct FbpInput {
    std::string inpr, dtype, proj_file, air_file, tgt, out_file;
    int filter = 2, width = 0, height = 0, offset = 0, recon_size = 0, slices = 1, xfactor = 1;
    double period = 0, rate = 0, start_angle = 0, couch_speed = 0, couch_start = 0, pix = 0;
    double recon_offset = 0, recon_step = 0;
};

bool load(const char* path, FbpInput& in)
{
    vct::TxtReader r(path);
    if (!r.ok()) {
        std::fprintf(stderr, "info file: %s not found\n", path);
        return false;
    }
    r.skip(6);
    r.read_string(&in.inpr);
    r.read_int(&in.filter);
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
    r.read_string(&in.out_file);
    return true;
}

int fail(const char* what, vct_status st)
{
    std::fprintf(stderr, "ERROR: %s: %s (%s)\n", what, vct_status_string(st), vct_last_error());
    return 1;
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char* argv[])
{
    if (argc != 2) {
        std::fprintf(stderr, "\n[usage]$ cbct <info filename>\n\n");
        return 0;
    }
    FbpInput in;
    if (!load(argv[1], in)) return 0;
    const int total_proj = (int)(in.height / (in.period * in.rate));  // :94-96
    const int H = (int)(in.period * in.rate);
    const int W = in.width;
    if (in.dtype != "float") {
        std::fprintf(stderr, "ERROR: sinogram data type '%s' is not supported (float only)\n", in.dtype.c_str());
        return 1;
    }
    if (H <= 0 || W <= 0 || total_proj < 1) {
        std::fprintf(stderr, "ERROR: sinogram %d x %d with %d rotations is empty\n", W, H, total_proj);
        return 1;
    }
    const bool in_mode = in.inpr == "IN";
    if (in_mode && total_proj != 1) {
        std::fprintf(stderr, "ERROR: IN sinograms with more than one rotation are not supported\n");
        return 1;
    }

    // sinogram, F/projectionDataTOMO.cpp:209-292 (the first rotation is used)
    std::vector<float> proj((size_t)W * H * total_proj, 0.f);
    FILE* fp = std::fopen(in.proj_file.c_str(), "rb");
    if (!fp) {
        std::fprintf(stderr, "data file: %s not found\n", in.proj_file.c_str());
    } else {
        for (int t = 0; t < (in_mode ? 1 : total_proj); t++) {
            size_t got = std::fread(proj.data() + (size_t)t * W * H, sizeof(float), (size_t)W * H, fp);
            (void)got;
            if (!in_mode) std::fprintf(stderr, "Z scan step: %lf \n", in.couch_start + t * in.couch_speed);
        }
        std::fclose(fp);
        if (in_mode)
            for (size_t i = 0; i < (size_t)W * H; i++)
                if (proj[i] < 50000.0f) proj[i] = 1.0f;
    }

    FILE* out = std::fopen(in.out_file.c_str(), "wb");
    if (!out) {
        std::fprintf(stderr, "ERROR: output file NOT writable\n");
        return 0;
    }
    // reconstruction range check, F/main.cpp:118-128
    const int used_recon_angle = 2;
    const double zstart = in.recon_offset;
    const double zend = in.recon_offset + (in.slices - 1) * in.recon_step;
    const double s_start = in.couch_start + in.couch_speed / used_recon_angle + in.couch_speed - 0.0000001;
    const double s_end = in.couch_start + total_proj * in.couch_speed - in.couch_speed / (used_recon_angle + 0.0000001);
    std::fprintf(stderr, "Allowable reconstraction range = %lf to %lf \n", s_start, s_end);
    std::fprintf(stderr, "Reconstructing z = %lf to %lf, Step = %f ... \n", zstart, zend, in.recon_step);
    if (zstart < s_start || zend > s_end) {
        std::fprintf(stderr, "\n Wrong reconstruction range !!-- STOP\n\n");
        std::fclose(out);
        return 0;
    }
    std::fprintf(stderr, "z slice num. = %d to %d, usedProjNum = %d\n", 1, 1, total_proj);

    vct_geometry g;
    vct_geometry_default(&g);
    g.recon_size = in.recon_size;
    g.pixel_cm = in.pix;
    g.n_channels = W;
    g.n_views = H;
    g.start_view_deg = in.start_angle;
    g.xfactor = in.xfactor;
    const char* env_dev = std::getenv("VCT_DEVICE");
    vct_ctx* ctx = nullptr;
    vct_status st = vct_init(env_dev ? std::atoi(env_dev) : 0, &ctx);
    if (st != VCT_OK) return fail("vct_init", st);

    auto t0 = std::chrono::steady_clock::now();
    std::vector<float> fltd((size_t)W * H);
    st = vct_filter_sinogram_host(ctx, &g, proj.data(), in.filter, in_mode ? 1 : 0, fltd.data());
    if (st != VCT_OK) return fail("filterSinogram", st);
    std::fprintf(stderr, "%s : %3.2lf sec.\n", "filterSinogram_TOMO", seconds_since(t0));
    t0 = std::chrono::steady_clock::now();
    const size_t NN = (size_t)in.recon_size * in.recon_size;
    std::vector<float> img(NN);
    st = vct_reconstruct_image_host(ctx, &g, fltd.data(), img.data());
    if (st != VCT_OK) return fail("reconstructImage", st);
    for (int z = 0; z < in.slices; z++)
        std::fprintf(stderr, "zz-c_position = %lf, z_image = %d %d %d \n",
                     in.recon_offset + z * in.recon_step - in.couch_start, 0, 0, 0);
    std::fprintf(stderr, "%s : %3.2lf sec.\n", "reconstructImage_TOMO", seconds_since(t0));
    for (int z = 0; z < in.slices; z++) std::fwrite(img.data(), sizeof(float), NN, out);
    std::fclose(out);
    vct_terminate(ctx);
    std::fprintf(stderr, "done.\n");
    return 0;
}
