// vct_npf.hip -- the npf side output of the MBFPA program for CDNA4 (gfx950):
// a pixel-driven, chord-weighted backprojection of a normalised sinogram
// (M/lengthproj_TOMO.cpp:14-185, called from M/main.cpp:407-419).
//
// One thread per pixel of a 16x16 tile.  Per view the thread forms the
// reference's chord length of the ray from the source through the pixel
// (:72-106), intersects that ray with the detector arc about the source
// (:109-131), converts the hit to a fractional channel with acos (:133-150),
// interpolates the view's row of the sinogram and adds sample * chord in
// float, views ascending (:152-170).  All geometry is FP64 in the
// reference's operation order (contraction off); the sinogram window
// (<= 2 MB) stays in L2/MALL.  The channel pitch and isocentre channel are
// the detector's (the plan's geometry); a window row may hold fewer channels
// (measured data keep 582 or fewer), and hits past its last channel take that
// channel's value (the reference, whose pitch is fixed by its 609-channel
// End_Channel, interpolates there with the narrower row as stride, i.e. reads
// the next row's first channels).  The reference also fills two per-pixel,
// per-view caches for its disabled iterative loop; they are not produced.
#include "vct_internal.h"

#include <cmath>

namespace vct {

constexpr int kNpfTile = 16;

__global__ __launch_bounds__(256) void k_lengthproj(LengthprojArgs a)
{
    const int jx = blockIdx.x * kNpfTile + (threadIdx.x & (kNpfTile - 1));
    const int jy = blockIdx.y * kNpfTile + threadIdx.x / kNpfTile;
    if (jx >= a.N || jy >= a.N) return;
    // prepareMaskData, M/misc_TOMO.cpp:21-36
    const double mr = (double)(a.N - 1) / 2.0;
    const bool inside = ((double)jy - mr) * ((double)jy - mr) + ((double)jx - mr) * ((double)jx - mr) <= mr * mr;
    float acc = 0.f;
    if (inside) {
        const double pix = a.pix;
        const double ry = a.N * pix * 0.5 - (jy + 0.5) * pix;   // :56
        const double rx = -a.N * pix * 0.5 + (jx + 0.5) * pix;  // :61
        const double rz = a.rz, R = a.det_radius;
        const int last = a.pw - 1;  // the window's last channel
        for (int j = 0; j < a.count; j++) {
            int jr = a.base + j;
            if (jr >= a.V) jr -= a.V;  // :70-71
            const double c = a.cosT[jr], s = a.sinT[jr];
            const double sx = a.sid * s, sy = a.sid * c, sz = rz;
            const double yslope = fabs(ry - sy), xslope = fabs(rx - sx);
            double rr;
            if (xslope > yslope) {  // :84-92
                const double q0 = (rx - 0.5 * pix - sx) / (rx - sx);
                const double q1 = (rx + 0.5 * pix - sx) / (rx - sx);
                const double xx0 = rx - 0.5 * pix, xx1 = rx + 0.5 * pix;
                const double yy0 = q0 * (ry - sy) + sy, yy1 = q1 * (ry - sy) + sy;
                const double zz0 = q0 * (rz - sz) + sz, zz1 = q1 * (rz - sz) + sz;
                rr = sqrt((xx0 - xx1) * (xx0 - xx1) + (yy0 - yy1) * (yy0 - yy1) + (zz0 - zz1) * (zz0 - zz1));
            } else {  // :94-101, whose far end reuses yy0: xx1 == xx0
                const double yy0 = ry - 0.5 * pix, yy1 = ry + 0.5 * pix;
                const double q0 = (yy0 - sy) / (ry - sy);
                const double xx0 = q0 * (rx - sx) + sx, xx1 = xx0;
                const double zz0 = q0 * (rz - sz) + sz, zz1 = zz0;
                rr = sqrt((xx0 - xx1) * (xx0 - xx1) + (yy0 - yy1) * (yy0 - yy1) + (zz0 - zz1) * (zz0 - zz1));
            }
            // detector hit, :109-131
            const double dx = a.det_src_iso * s, dy = a.det_src_iso * c;
            const double aa = (sy - ry) / (sx - rx);
            const double bb = -sx * (sy - ry) / (sx - rx) + sy;
            const double A = 1.0 + aa * aa;
            const double B = 2.0 * aa * bb - 2.0 * aa * dy - 2.0 * dx;
            const double C = dx * dx + bb * bb + dy * dy - 2.0 * bb * dy - R * R;
            const double root = sqrt(B * B - 4.0 * A * C);
            const double px1 = (-B + root) / (2.0 * A), py1 = aa * px1 + bb;
            const double px2 = (-B - root) / (2.0 * A), py2 = aa * px2 + bb;
            const bool second = px1 * px1 + py1 * py1 > px2 * px2 + py2 * py2;
            const double px = second ? px2 : px1, py = second ? py2 : py1;
            // fractional channel, :133-150
            const double ox = -a.iso_det * s, oy = -a.iso_det * c;
            const double dis = (px - ox) * (px - ox) + (py - oy) * (py - oy);
            const double angle = acos(1.0 - dis / (2.0 * R * R));
            const double rot_px = px * c - py * s;
            const double dj = rot_px > 0.0 ? a.iso_channel - angle / a.delta : a.iso_channel + angle / a.delta;
            const float* row = a.proj + (size_t)j * a.pw;
            double np;
            if (0 <= dj && dj < last) {  // :152-158
                const int i0 = (int)dj;
                const double w = dj - (double)i0;
                np = (1.0 - w) * row[i0] + w * row[i0 + 1];
            } else if (0 > dj) {
                np = row[0];
            } else if (dj >= last) {
                np = row[last];
            } else {
                continue;  // NaN channel: the reference adds nothing
            }
            acc += (float)(np * rr);
        }
    }
    a.npf[(size_t)jy * a.N + jx] = acc;
}

vct_status launch_lengthproj(const LengthprojArgs& a, hipStream_t s)
{
    const dim3 grid((a.N + kNpfTile - 1) / kNpfTile, (a.N + kNpfTile - 1) / kNpfTile);
    hipLaunchKernelGGL(k_lengthproj, grid, dim3(kNpfTile * kNpfTile), 0, s, a);
    VCT_HIP(hipGetLastError());
    return VCT_OK;
}

}  // namespace vct
