// vct_march.h -- the fan-ray march shared by the tracer, the adjoint and the
// system-matrix builder (device code, included by .hip files only).
//
// march_ray<false> evaluates the reference's segment decomposition
// (M/reprojection_TOMO.cpp:53-283) operation for operation in FP64 with
// correctly rounded divisions; compiled with -ffp-contract=off it produces
// the reference's segments bit for bit.
//
// march_ray<true> produces the same segments (same voxels, same count, same
// order) with far fewer FP64 divisions.  The per-step divisions by the ray's
// direction components are replaced by one fused multiply-add with a per-ray
// slope (reciprocal times the other component), which moves a computed
// coordinate by at most ~1.2e-13 cm (slope within 2 ulp, |offset| <= 143 cm,
// plus the rounding of a value below 2^7).  Every
// discrete decision taken from such a coordinate -- the loop bound +-L, the
// minor-axis cell index, the sign split at 0, the crossing test chy != 0 and
// the midpoint voxel index -- is first compared against a margin of 1e-9 (cm,
// or voxel units for indices, i.e. > 6000x the error bound for pixels down to
// 10 um); inside the margin the coordinates involved are recomputed with the
// reference's exact expression and the decision is taken on those, so every
// decision equals the reference's.  Segment lengths become |da| * sec(ray),
// which differs from the reference's sqrt by rounding only and never enters a
// decision.  Each step (plane point, minor cell, crossing, the voxels of its
// one or two segments) is first evaluated from the fast values with all
// margin tests folded into one flag and committed without branches when the
// flag holds; only a step that fails a test runs the general step, which
// makes the points involved exact.  The general step, given the same fast
// values and no failed test, takes the same decisions, so both routes give
// the same segments.
//
// Orbit sinks (kRot): with V views over 360 degrees, the ray of view v + kV/R
// is the ray of view v turned by k/R of a turn about the centre, so its
// segments are the turned voxels of view v's segments with the same lengths,
// provided every discrete decision of the march holds for the turned ray
// too.  Its own trig values differ from the exact turn by a few ulp, which
// moves its coordinates by ~1e-13; the march therefore reports to such a
// sink every decision that does not clear kMarchMargin (the general step,
// and in the set-up the clip tests, the entry choice, the major axis, the
// integer-cm start plane and the direction signs), and the caller then
// traces the turned rays one by one.  Such sinks take voxels as (major,
// minor) index pairs.
#pragma once

#include <climits>
#include <cmath>
#include <type_traits>

#include "vct_internal.h"

namespace vct {

constexpr double kMarchMargin = 1e-9;

// a * b for 0 <= a, b < 2^24 (the full-rate 24-bit multiply on the device;
// a 32-bit multiply is a quarter-rate instruction)
__host__ __device__ __forceinline__ int mul24(int a, int b)
{
#ifdef __HIP_DEVICE_COMPILE__
    return (int)__umul24((uint32_t)a, (uint32_t)b);
#else
    return a * b;
#endif
}

// Sink::kRot, false when the sink does not declare it.
template <class S, class = void>
struct sink_rot : std::false_type {};
template <class S>
struct sink_rot<S, std::void_t<decltype(S::kRot)>> : std::integral_constant<bool, S::kRot> {};

// Sink::kAB: the sink takes in-image segments as add_ab(major index, minor
// index, length) instead of add(voxel, length).
template <class S, class = void>
struct sink_ab : std::false_type {};
template <class S>
struct sink_ab<S, std::void_t<decltype(S::kAB)>> : std::integral_constant<bool, S::kAB> {};

// |v| and its distance to the nearest integer both clear the margin: trunc(v)
// and the sign of v do not change under a perturbation below the margin.
__host__ __device__ __forceinline__ bool trunc_safe(double v)
{
    const double u = fabs(v);
    const double f = u - floor(u);
    return u >= kMarchMargin && f >= kMarchMargin && f <= 1.0 - kMarchMargin;
}

// x86-64 cvttsd2si semantics (the reference's (int) casts): NaN and values
// outside int range give INT_MIN.
__host__ __device__ __forceinline__ int trunc_i32(double v)
{
    return (v > -2147483649.0 && v < 2147483648.0) ? (int)v : INT_MIN;
}

// trunc(fl(fl(h / D) + c)), the voxel index expressions of the reference
// (M/reprojection_TOMO.cpp:130,146-147,217), for an exactly known h.  The
// product h * (1/D) differs from the correctly rounded quotient by a few ulp,
// so its truncation can only differ within a few ulp of an integer; such
// values take the exact division.
template <bool FAST>
__host__ __device__ __forceinline__ int index_of(double h, double D, double invD, double c)
{
    if (FAST) {
        const double q = __builtin_fma(h, invD, c);
        const double t = trunc(q);
        const double f = fabs(q - t);
        if (fabs(f - 0.5) < 0.5 - kMarchMargin && fabs(q) < 1.0e6) return (int)t;
    }
    return trunc_i32(h / D + c);
}

// Index of a value known only to ~1e-13: returns false (undecided) when the
// value is within the margin of an integer boundary.
// The value is only trusted away from the boundaries, so the fused
// multiply-add (one rounding fewer) is as good as the separate operations,
// and |f - 0.5| < 0.5 - margin is the two-sided test f in (margin, 1 - margin).
// A NaN or infinite q fails that test (f is NaN), and every index whose value
// is used comes from a point within a few image widths of the image, far
// inside the int range, so no separate range test is needed.
__host__ __device__ __forceinline__ bool index_checked(double h, double invD, double c, int& out)
{
    const double q = __builtin_fma(h, invD, c);
    const double t = trunc(q);
    const double f = fabs(q - t);
    out = (int)t;
    return fabs(f - 0.5) < 0.5 - kMarchMargin;
}

// The same, also giving floor(q) (for orbit sinks: the reference's index is
// trunc(q), which is floor(q) + 1 for q in (-1, 0), and a turned ray's index
// value is N - q, so the turned indices follow from the floor).
__host__ __device__ __forceinline__ bool index_checked_fl(double h, double invD, double c, int& out, int& fl)
{
    const double q = __builtin_fma(h, invD, c);
    const double t = trunc(q);
    const double f = fabs(q - t);
    out = (int)t;
    fl = (int)t - (q < t ? 1 : 0);
    return fabs(f - 0.5) < 0.5 - kMarchMargin;
}

// (u/fabs(u) != v/fabs(v)) of M/reprojection_TOMO.cpp:134,221 for finite
// operands, without the divisions: true if either is zero or NaN (0/0 is
// NaN, which compares unequal) or the signs differ.
__host__ __device__ __forceinline__ bool sign_split(double u, double v)
{
    const bool up = u > 0.0, un = u < 0.0, vp = v > 0.0, vn = v < 0.0;
    return !(up || un) || !(vp || vn) || (up != vp);
}

// A point of the march in (major, minor) coordinates.  Plane points have an
// exact major coordinate, crossing points an exact minor coordinate, the
// entry point is exact; `exact` marks the other coordinate as exact too.
struct MarchPoint {
    double a, b;
    int kind;  // 0 entry, 1 plane, 2 crossing
    bool exact;
};

struct RayConst {
    bool xmajor;
    double ai, af, bi, bf;  // entry and exit, major/minor
    double sa, sb;          // source, major/minor
    double dA, dB;          // detector - source, major/minor
    double Lb;              // half width of the image square
};

// The march from the entry point (M/reprojection_TOMO.cpp:113-283, either
// branch) for one ray.  FAST selects reciprocal products with guarded
// decisions (see the file comment) over exact divisions.
// Window: only steps k_begin..k_end (the loop counter ik of :115) run.  Step
// k's segments lie between major planes k - 1 and k (step 0: from the entry
// to plane 0), and the state the march carries into step k is plane point
// k - 1 with its minor cell, so a window starts from that point, built as the
// general step builds it; consecutive windows then give the segments of the
// whole march (lengths equal, or within an ulp where a plane point was made
// exact for a decision of the step before), and the returned counter is
// meaningful for the whole march only.
template <bool FAST, class Sink>
__host__ __device__ __forceinline__ int march_loop(const TraceArgs& a, const RayConst& r, Sink& sink, bool& ovf,
                                          int k_begin = 0, int k_end = INT_MAX)
{
    constexpr bool kRot = sink_rot<Sink>::value;
    constexpr bool kTakesAB = sink_ab<Sink>::value;
    const bool xmajor = r.xmajor;
    const double ai = r.ai, af = r.af, bi = r.bi, bf = r.bf;
    const double sa = r.sa, sb = r.sb, dA = r.dA, dB = r.dB, Lb = r.Lb;
    const double D = a.D;
    const double rA = FAST ? 1.0 / dA : 0.0;
    const double rB = FAST ? 1.0 / dB : 0.0;
    double sec = 0.0;
    if (FAST) {
        const double t = dB / dA;
        sec = sqrt(1.0 + t * t);
    }
    // minor coordinate on a major plane and major coordinate on a minor
    // plane: exact (:121, :129, :137) and fast forms
    auto b_exact = [&](double a_) { return (a_ - sa) / dA * dB + sb; };
    auto a_exact = [&](double b_) { return (b_ - sb) / dB * dA + sa; };
    const double kAB = rA * dB, kBA = rB * dA;
    auto b_calc = [&](double a_) { return FAST ? __builtin_fma(a_ - sa, kAB, sb) : b_exact(a_); };
    auto a_calc = [&](double b_) { return FAST ? __builtin_fma(b_ - sb, kBA, sa) : a_exact(b_); };
    auto make_exact = [&](MarchPoint& p) {
        if (p.exact) return;
        if (p.kind == 1) p.b = b_exact(p.a);
        else if (p.kind == 2) p.a = a_exact(p.b);
        p.exact = true;
    };

    const double pha = (af - ai) / fabs(ai - af);
    int ain;
    if ((ai < 0 && (af - ai) < 0) || (ai > 0 && (af - ai) > 0))
        ain = trunc_i32(trunc_i32(ai) + (af - ai) / fabs(af - ai));
    else
        ain = trunc_i32(ai);
    const double bdir = (bf - bi) / fabs(bf - bi);
    const bool b_pos = (bf - bi) > 0;
    const bool b_neg = (bf - bi) < 0;

    // one candidate segment p0 -> p1, :141-166 / :168-189
    // The midpoint voxel (:146-149) in (major, minor) form: the x index
    // takes +x_mid = (x0 + x1) * 0.5, the y index (-y0 - y1) * 0.5, which is
    // exactly -(y0 + y1) * 0.5 (negation and halving are exact; only the sign
    // of a zero can differ, which no index sees).  So the major index takes
    // sA * a_mid and the minor one -sA * b_mid with sA = +1 when x is major,
    // and voxel = ia * stA + ib * stB; no per-segment selects between x and y.
    const double sA = xmajor ? 1.0 : -1.0;
    const int stA = xmajor ? 1 : a.N, stB = xmajor ? a.N : 1;
    // fast midpoint index: fma(p0 + p1, +-0.5/D, nxm) is the fma of the
    // reference's +-(p0 + p1) * 0.5 with 1/D (power-of-two scalings are exact)
    const double hA = 0.5 * sA * a.invD, hB = -0.5 * sA * a.invD;
    // a point whose major coordinate a turned ray reproduces exactly (negated
    // or equal): a plane, or an entry on a major-axis side of the square
    auto a_sym = [&](const MarchPoint& p) { return p.kind == 1 || (p.kind == 0 && fabs(p.a) == Lb); };
    auto emit = [&](MarchPoint& p0, MarchPoint& p1) {
        if constexpr (kRot) {
            // index values: the fast ones when both clear the margin, else the
            // reference's exact expressions.  A tie of the major index between
            // two points with exactly reproduced major coordinates resolves the
            // same way for every turned ray (their sums are exact negations or
            // equal); any other decision within the margin is left to the
            // turned rays' own marches.
            double qa = __builtin_fma(p0.a + p1.a, hA, (double)a.nxm);
            double qb = __builtin_fma(p0.b + p1.b, hB, (double)a.nxm);
            const bool ca = fabs(fabs(qa - trunc(qa)) - 0.5) < 0.5 - kMarchMargin;
            const bool cb = fabs(fabs(qb - trunc(qb)) - 0.5) < 0.5 - kMarchMargin;
            const double rr = FAST ? fabs(p1.a - p0.a) * sec
                                   : sqrt((p0.a - p1.a) * (p0.a - p1.a) + (p0.b - p1.b) * (p0.b - p1.b));
            if (!FAST || !(ca && cb)) {
                make_exact(p0);
                make_exact(p1);
                // the reference's index values are t + nxm with the signed
                // quotients t below; a turned ray whose coordinates on this
                // axis are these negated or equal computes +-t exactly (the
                // division is odd), so the sink gives it fl(+-t + nxm), its own
                // value: exact for the major axis between planes and major-side
                // entries, any other tie is left to the turned rays' marches
                const double ta = sA * ((p0.a + p1.a) * 0.5) / D;
                const double tb = -sA * ((p0.b + p1.b) * 0.5) / D;
                if (FAST && (!cb || !(ca || (a_sym(p0) && a_sym(p1))))) sink.unsafe = true;
                sink.add_t(ta, tb, rr);
            } else {
                sink.add_fl((int)floor(qa), (int)floor(qb), rr);
            }
        } else {
            int ia, ib;
            bool decided = false;
            if (FAST) {
                decided = index_checked(p0.a + p1.a, hA, (double)a.nxm, ia) &&
                          index_checked(p0.b + p1.b, hB, (double)a.nxm, ib);
                if (!decided) {
                    make_exact(p0);
                    make_exact(p1);
                }
            }
            if (!decided) {
                ia = index_of<FAST>(sA * ((p0.a + p1.a) * 0.5), D, a.invD, (double)a.nxm);
                ib = index_of<FAST>(-sA * ((p0.b + p1.b) * 0.5), D, a.invD, (double)a.nxm);
            }
            if (ia >= 0 && ib >= 0 && ia <= a.N - 1 && ib <= a.N - 1) {
                if (a.rcheck) {
                    // x^2 + y^2 of the reference with the terms in either order
                    const double r = sqrt((p0.a + p1.a) * (p0.a + p1.a) + (p0.b + p1.b) * (p0.b + p1.b)) * 0.5;
                    if (r >= a.rmax) return;
                }
                const double rr = FAST ? fabs(p1.a - p0.a) * sec
                                       : sqrt((p0.a - p1.a) * (p0.a - p1.a) + (p0.b - p1.b) * (p0.b - p1.b));
                if constexpr (kTakesAB) sink.add_ab(ia, ib, rr);
                else sink.add(mul24(ia, stA) + mul24(ib, stB), rr);
            }
        }
    };

    if constexpr (kRot) {
        // set-up decisions of this function: the integer-cm start plane and the
        // signs of the entry coordinate and of both directions (an entry on a
        // major-axis side is exactly +-Lb for every turned ray)
        if (!((fabs(ai) == Lb || trunc_safe(ai)) && fabs(af - ai) >= kMarchMargin && fabs(bf - bi) >= kMarchMargin))
            sink.unsafe = true;
    }
    MarchPoint p0{ai, bi, 0, true};
    MarchPoint pc{(double)ain, 0.0, 1, !FAST};
    pc.b = b_calc(pc.a);
    if (FAST && fabs(fabs(pc.b) - Lb) < kMarchMargin) {
        make_exact(pc);
        if constexpr (kRot) sink.unsafe = true;
    }
    int bma = 0;
    int iks = 0;

    const double nxm = (double)a.nxm;
    // 0 <= ia, ib < N in one unsigned comparison (negatives wrap above N)
    auto in_image = [&](int ia, int ib) {
        const uint32_t ua = (uint32_t)ia, ub = (uint32_t)ib;
        return (ua > ub ? ua : ub) < (uint32_t)a.N;
    };

    int ik_first = 0;
    if (k_begin > 0) {
        // windowed start: plane point k_begin - 1 as the general step makes it
        // (a committed fast step leaves the same point and cell: it commits
        // only when none of these tests fires)
        pc.a = ain + (pha * (double)(k_begin - 1)) * D;
        pc.b = b_calc(pc.a);
        pc.exact = !FAST;
        if (FAST) {
            const double ab = fabs(pc.b);
            if (ab < kMarchMargin || fabs(ab - Lb) < kMarchMargin) {
                make_exact(pc);
                if constexpr (kRot) sink.unsafe = true;
            }
            if (!index_checked(pc.b, a.invD, 0.0, bma)) {
                make_exact(pc);
                bma = trunc_i32(pc.b / D);
                if constexpr (kRot) sink.unsafe = true;
            }
        } else {
            bma = index_of<false>(pc.b, D, a.invD, 0.0);
        }
        p0 = pc;
        ik_first = k_begin;
    }

    double pik = pha * (double)ik_first;  // pha * ik, exact
    for (int ik = ik_first;; ik++, pik += pha) {
        if (!(pc.a <= Lb && pc.a >= -Lb && pc.b <= Lb && pc.b >= -Lb)) break;
        if (ik >= a.max_steps) { ovf = true; break; }
        if (ik > k_end) break;
        {
            const int bmb = bma;
            const double bj = pc.b;
            bool committed = false;
            if (FAST) {
                // The whole step from the fast values with every margin test
                // folded into one flag: when all of them pass, the general step
                // below would make no point exact and take exactly these
                // decisions with exactly these values, so the step is committed
                // here without its branches; otherwise the general step runs.
                const double na = ain + pik * D;
                const double nb = b_calc(na);
                const double nab = fabs(nb);
                int nbma;
                bool sure = index_checked(nb, a.invD, 0.0, nbma);
                sure &= (nab >= kMarchMargin) & (fabs(nab - Lb) >= kMarchMargin);
                const bool cand = (ik != 0) & ((nbma != bmb) | sign_split(nb, bj));
                const double xbb = (double)(nbma * D);
                const double xb = ((nb < 0 && b_neg) || (nb > 0 && b_pos)) ? xbb : xbb - bdir * D;
                const double xa = a_calc(xb);
                const double dxa = na - xa, dxb = nb - xb;
                sure &= !(cand & (fabs(dxa) < kMarchMargin) & (fabs(dxb) < kMarchMargin));
                // chy != 0 of :141: when the step is committed, cand implies
                // |dxa| or |dxb| >= margin, so chy > 0
                const bool cross = cand;
                int ia1, ib1, ia2, ib2, fa1, fb1, fa2, fb2;
                const bool ok1 = index_checked_fl(p0.a + xa, hA, nxm, ia1, fa1) &
                                 index_checked_fl(p0.b + xb, hB, nxm, ib1, fb1);
                const double s_a = cross ? xa : p0.a, s_b = cross ? xb : p0.b;
                const bool ok2 = index_checked_fl(s_a + na, hA, nxm, ia2, fa2) &
                                 index_checked_fl(s_b + nb, hB, nxm, ib2, fb2);
                sure &= (ok1 | !cross) & ok2;
                if (sure) {
                    const bool v1 = cross && in_image(ia1, ib1);
                    const bool v2 = in_image(ia2, ib2);
                    iks += cross ? 2 : 1;
                    if constexpr (kTakesAB) {
                        if (v1) sink.add_ab(ia1, ib1, fabs(xa - p0.a) * sec);
                        if (v2) sink.add_ab(ia2, ib2, fabs(na - s_a) * sec);
                    } else if constexpr (kRot) {
                        // every candidate segment with the floors of its index values; the
                        // sink turns them into each ray's voxel, a voxel outside the image
                        // reading zeros (no crossing: the first candidate is skipped)
                        sink.add2_fl(cross, fa1, fb1, fabs(xa - p0.a) * sec, fa2, fb2, fabs(na - s_a) * sec);
                    } else if constexpr (Sink::kAdd2) {
                        // (a voxel index is only used when its segment is in the image)
                        sink.add2(v1, mul24(ia1, stA) + mul24(ib1, stB), fabs(xa - p0.a) * sec,
                                  v2, mul24(ia2, stA) + mul24(ib2, stB), fabs(na - s_a) * sec);
                    } else {
                        if (v1) sink.add(mul24(ia1, stA) + mul24(ib1, stB), fabs(xa - p0.a) * sec);
                        if (v2) sink.add(mul24(ia2, stA) + mul24(ib2, stB), fabs(na - s_a) * sec);
                    }
                    bma = nbma;
                    pc = MarchPoint{na, nb, 1, false};
                    p0 = pc;
                    committed = true;
                }
            }
            if (!committed) {
                pc.a = ain + pik * D;  // :126
                pc.b = b_calc(pc.a);        // :129
                pc.exact = !FAST;
                if (FAST) {
                    const double ab = fabs(pc.b);
                    if (ab < kMarchMargin || fabs(ab - Lb) < kMarchMargin) {
                        make_exact(pc);
                        if constexpr (kRot) sink.unsafe = true;
                    }
                    if (!index_checked(pc.b, a.invD, 0.0, bma)) {
                        make_exact(pc);
                        bma = trunc_i32(pc.b / D);  // :130
                        if constexpr (kRot) sink.unsafe = true;
                    }
                } else {
                    bma = index_of<false>(pc.b, D, a.invD, 0.0);
                }
                const double bii = pc.b;
                if (ik != 0 && (bma != bmb || sign_split(bii, bj))) {  // :134
                    MarchPoint px{0.0, 0.0, 2, !FAST};
                    if ((bii < 0 && b_neg) || (bii > 0 && b_pos))
                        px.b = (double)(bma * D);
                    else
                        px.b = (double)(bma * D) - bdir * D;
                    px.a = a_calc(px.b);  // :137
                    if (FAST && fabs(pc.a - px.a) < kMarchMargin && fabs(pc.b - px.b) < kMarchMargin) {
                        make_exact(pc);
                        make_exact(px);
                        if constexpr (kRot) sink.unsafe = true;
                    }
                    const double chy = (pc.a - px.a) * (pc.a - px.a) + (pc.b - px.b) * (pc.b - px.b);
                    if (chy != 0) {  // :141
                        iks++;
                        emit(p0, px);
                        p0 = px;
                    }
                }
                iks++;
                emit(p0, pc);
                p0 = pc;
            }
        }
    }
    return iks;
}

// Entry, exit and march constants of ray (view, ch) (M/reprojection_TOMO.cpp:53-113);
// false when the ray misses the image.
__host__ __device__ __forceinline__ bool setup_ray(const TraceArgs& a, int view, int ch, RayConst& r,
                                                  bool* safe = nullptr)
{
    // source and detector point, M/reprojection_TOMO.cpp:53-74
    const double c = a.cosT[view], s = a.sinT[view];
    const double sx = a.sid * s;
    const double sy = a.sid * c;
    const double px0 = a.px0[ch], py0 = a.py0[ch];
    const double px = +px0 * c + py0 * s;
    const double py = -px0 * s + py0 * c;
    const double Lb = (double)a.nxm * a.D;
    const double D = a.D;

    // clip against the image square, :76-101 (only the first two complete
    // intersections are used)
    int m = 0;
    double h0x = 0.0, h0y = 0.0, h1x = 0.0, h1y = 0.0;
    // every decision clears the margin (for orbit sinks; NaN fails)
    bool ok = true;
    auto clear = [&](double d) { ok = ok && fabs(d) >= kMarchMargin; };
    {
        const double y = Lb;
        const double x = (y - sy) / (py - sy) * (px - sx) + sx;
        clear(fabs(x) - Lb);
        if (fabs(x) < Lb) { h0x = x; h0y = y; m = 1; }
    }
    {
        const double y = -Lb;
        const double x = (y - sy) / (py - sy) * (px - sx) + sx;
        clear(fabs(x) - Lb);
        if (fabs(x) < Lb) {
            if (m == 0) { h0x = x; h0y = y; } else { h1x = x; h1y = y; }
            m++;
        }
    }
    {
        const double x = Lb;
        const double y = (x - sx) / (px - sx) * (py - sy) + sy;
        clear(fabs(y) - Lb);
        if (fabs(y) < Lb) {
            if (m == 0) { h0x = x; h0y = y; } else if (m == 1) { h1x = x; h1y = y; }
            m++;
        }
    }
    {
        const double x = -Lb;
        const double y = (x - sx) / (px - sx) * (py - sy) + sy;
        clear(fabs(y) - Lb);
        if (fabs(y) < Lb) {
            if (m == 0) { h0x = x; h0y = y; } else if (m == 1) { h1x = x; h1y = y; }
            m++;
        }
    }
    if (safe) *safe = ok;
    if (m < 2) return false;  // the ray misses the image: no segment, transmission 1

    // entry = the intersection closer to the source, :103-108
    const double d0 = (h0x - sx) * (h0x - sx) + (h0y - sy) * (h0y - sy);
    const double d1 = (h1x - sx) * (h1x - sx) + (h1y - sy) * (h1y - sy);
    const bool first = d0 < d1;
    clear(d0 - d1);
    const double xi = first ? h0x : h1x, yi = first ? h0y : h1y;
    const double xf = first ? h1x : h0x, yf = first ? h1y : h0y;

    // dominant axis, :113.  a = major axis, b = minor axis; the reference's
    // two branches (:113-200, :201-283) are this march with x and y exchanged.
    const bool xmajor = fabs(xi - xf) > fabs(yi - yf);
    clear(fabs(xi - xf) - fabs(yi - yf));
    if (safe) *safe = ok;
    const double ai = xmajor ? xi : yi, af = xmajor ? xf : yf;
    const double bi = xmajor ? yi : xi, bf = xmajor ? yf : xf;
    const double sa = xmajor ? sx : sy, sb = xmajor ? sy : sx;
    const double pa = xmajor ? px : py, pb = xmajor ? py : px;

    r.xmajor = xmajor;
    r.ai = ai; r.af = af; r.bi = bi; r.bf = bf;
    r.sa = sa; r.sb = sb; r.dA = pa - sa; r.dB = pb - sb;
    r.Lb = Lb;
    return true;
}

// Steps k0..k1 of ray r whose segments can have their midpoint voxel in the
// box x_lo..x_hi, y_lo..y_hi (cm; the caller pads it by two voxels, which
// covers the reference's truncation of negative index values and the
// midpoint's distance from the box).  Step k's segments lie between major
// planes k - 1 and k at ain + pha * k * D (step 0 from the entry), so a
// midpoint at major coordinate u belongs to step ceil(pha * (u - ain) / D);
// the range of u where the ray is inside the box gives the steps, widened by
// one for rounding.  false: the ray has no such step.
__host__ __device__ __forceinline__ bool box_steps(const TraceArgs& a, const RayConst& r, double x_lo, double x_hi,
                                                   double y_lo, double y_hi, int& k0, int& k1)
{
    const double df = r.af - r.ai;
    if (!(df != 0.0)) {  // no major extent (or NaN): the whole march
        k0 = 0;
        k1 = INT_MAX;
        return true;
    }
    double lo = r.xmajor ? x_lo : y_lo, hi = r.xmajor ? x_hi : y_hi;
    const double b_lo = r.xmajor ? y_lo : x_lo, b_hi = r.xmajor ? y_hi : x_hi;
    if (r.dB != 0.0) {
        const double g = r.dA / r.dB;
        const double u = r.sa + (b_lo - r.sb) * g, w = r.sa + (b_hi - r.sb) * g;
        lo = fmax(lo, fmin(u, w));
        hi = fmin(hi, fmax(u, w));
    } else if (!(r.sb >= b_lo && r.sb <= b_hi)) {
        return false;
    }
    if (!(lo <= hi)) return false;
    // start plane of march_loop (M/reprojection_TOMO.cpp:115-119)
    const double pha = df / fabs(df);
    const int ain = ((r.ai < 0 && df < 0) || (r.ai > 0 && df > 0)) ? trunc_i32(trunc_i32(r.ai) + pha)
                                                                  : trunc_i32(r.ai);
    const double t1 = pha * (lo - ain) / a.D, t2 = pha * (hi - ain) / a.D;
    const double tmin = fmin(t1, t2), tmax = fmax(t1, t2);
    if (!(tmax < 1.0e9 && tmin > -1.0e9)) {
        k0 = 0;
        k1 = INT_MAX;
        return true;
    }
    k0 = (int)ceil(tmin) - 1;
    k1 = (int)ceil(tmax) + 1;
    if (k1 < 0) k1 = 0;  // step 0 runs from the entry, before plane 0
    if (k0 < 0) k0 = 0;
    return true;
}

// About the number of march steps of ray r: planes from the start plane to
// the exit, plus one (a split of the march into windows only needs the
// count roughly: the last window runs to the end).
__host__ __device__ __forceinline__ int march_step_estimate(const TraceArgs& a, const RayConst& r)
{
    const double df = r.af - r.ai;
    if (!(df != 0.0)) return 1;
    const int ain = ((r.ai < 0 && df < 0) || (r.ai > 0 && df > 0)) ? trunc_i32(trunc_i32(r.ai) + df / fabs(df))
                                                                  : trunc_i32(r.ai);
    const double n = fabs(r.af - (double)ain) / a.D;
    return n < 1.0e8 ? (int)n + 2 : 1;
}

// Channels lo..hi of view v whose rays can cross the box x_lo..x_hi,
// y_lo..y_hi (cm): in the view-0 frame the ray of channel ch leaves the
// source (0, sid) with tan = px0 / (sid - py0), monotone in ch, and a ray
// crosses a convex box seen from outside only between the tangents of its
// corners; one channel of margin on each side.  lo > hi: none.
__host__ __device__ __forceinline__ void box_channels(const TraceArgs& a, int v, double x_lo, double x_hi,
                                                      double y_lo, double y_hi, int& lo, int& hi)
{
    const double c = a.cosT[v], s = a.sinT[v];
    double tmin = 1.0e300, tmax = -1.0e300;
    bool ok = true;
    const double xs[2] = {x_lo, x_hi}, ys[2] = {y_lo, y_hi};
    for (int i = 0; i < 4; i++) {
        const double x = xs[i & 1], y = ys[i >> 1];
        const double x0 = c * x - s * y, y0 = s * x + c * y;  // rotated back to view 0
        const double den = a.sid - y0;
        if (!(den > 1.0e-6)) ok = false;
        const double t = x0 / den;
        tmin = fmin(tmin, t);
        tmax = fmax(tmax, t);
    }
    const int W = a.W;
    if (!ok) {
        lo = 0;
        hi = W - 1;
        return;
    }
    auto T = [&](int ch) { return a.px0[ch] / (a.sid - a.py0[ch]); };
    const bool dec = T(0) > T(W - 1);
    // U increasing in ch: first channel with U >= umin, and with U > umax
    const double umin = dec ? -tmax : tmin, umax = dec ? -tmin : tmax;
    auto first = [&](double u, bool strict) {
        int l = 0, h = W;  // answer in [l, h]
        while (l < h) {
            const int m = (l + h) >> 1;
            const double um = dec ? -T(m) : T(m);
            if (strict ? (um > u) : (um >= u)) h = m;
            else l = m + 1;
        }
        return l;
    };
    lo = first(umin, false) - 1;
    hi = first(umax, true);  // (last channel with U <= umax) + 1
    if (lo < 0) lo = 0;
    if (hi > W - 1) hi = W - 1;
}

// Trace ray (view, ch).  Calls sink.add(voxel, length) for every segment
// whose midpoint voxel lies in the image, in the reference's order, and
// returns the reference's per-ray segment counter (iks).  ovf is set when
// the step bound is hit (never for valid geometry).
template <bool FAST_IN, class Sink>
__host__ __device__ __forceinline__ int march_ray(const TraceArgs& a, int view, int ch, Sink& sink, bool& ovf)
{
    RayConst r;
    if (!setup_ray(a, view, ch, r)) return 0;
    // the fast path needs finite reciprocals and no r >= rmax test
    if (FAST_IN && r.dB != 0.0 && !a.rcheck && !a.exact) return march_loop<true>(a, r, sink, ovf);
    return march_loop<false>(a, r, sink, ovf);
}

}  // namespace vct
