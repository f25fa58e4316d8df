// Host-side conversion between the reference's FP64 state and the device's
// FP32 perturbation state (see vtcs_node.h for the representation).
//
// Reference state: dep (depth, +water), cel = c (cliffs_main.f:218-227),
// xvel, yvel.  Device state: s = sign(d) sqrt(g|d|), cp, u, v.
//
// Two predicates must come out of the conversion exactly as the FP64
// reference evaluates them, so that the wet/dry masks of the first step
// match bit for bit:
//   land = (c <= celmin)           cliffs.f:61     device: s+ + cp <= celmin32
//   deep = (d > ground)            cliffs.f:221    device: s > celmin32
// Where FP32 rounding flips one of them, the stored value is moved by ulps
// until the FP32 predicate agrees with the FP64 one.
#pragma once

#include <cmath>
#include <cstdint>

namespace cliffs {

struct Consts64 {
    double grav;     // (double)9.8f, global_modules.f:17
    double ground;
    double celmin;   // sqrt(grav*ground), cliffs_main.f:227
    float celmin32;
};

inline Consts64 make_consts(double ground) {
    Consts64 k;
    k.grav = (double)9.8f;
    k.ground = ground;
    k.celmin = std::sqrt(k.grav * ground);
    k.celmin32 = (float)k.celmin;
    return k;
}

inline float to_s(double d, const Consts64& k) {
    float s = (float)(d > 0 ? std::sqrt(k.grav * d) : -std::sqrt(-k.grav * d));
    const bool want = d > k.ground;
    for (int it = 0; it < 64 && ((s > k.celmin32) != want); ++it)
        s = std::nextafterf(s, want ? INFINITY : -INFINITY);
    return s;
}

inline float c0f(float s) { return s > 0.0f ? s : 0.0f; }

// Celerity perturbation for FP64 celerity c at FP64 depth d over device
// background s.  Wet cells keep the reference surface elevation
// eta = c^2/g - d exactly up to one FP32 rounding of cp.
inline float to_cp(double c, double d, float s, const Consts64& k) {
    const double c0 = c0f(s);
    float cp;
    const bool want_land = c <= k.celmin;
    if (want_land) {
        cp = (float)(c - c0);
    } else if (s > 0.0f) {
        const double geta = c * c - k.grav * d;
        const double ss = (double)s * (double)s;
        const double r = std::sqrt(std::fmax(ss + geta, 0.0));
        cp = (float)(geta / ((double)s + r));
    } else {
        cp = (float)c;
    }
    auto mismatch = [&](float q) { return ((c0f(s) + q) <= k.celmin32) != want_land; };
    auto ulp = [](float x) {
        x = std::fabs(x);
        return std::nextafterf(x, INFINITY) - x;
    };
    if (mismatch(cp)) {
        // within rounding of the threshold: start from it and step by one
        // ulp of the FP32 sum c0 + cp
        cp = k.celmin32 - c0f(s);
        const float step = std::fmax(ulp(c0f(s)), ulp(k.celmin32));
        for (int it = 0; it < 64 && mismatch(cp); ++it) cp += want_land ? -step : step;
    }
    return cp;
}

// Device (s, cp) back to an FP64 celerity whose surface elevation
// c^2/g - d equals the device's g*eta/g.  The FP64 predicate c <= celmin
// of the result equals the device's FP32 predicate.
inline double to_cel(float s, float cp, double d, const Consts64& k) {
    const float c32 = c0f(s) + cp;
    if (c32 <= k.celmin32) return std::fmin((double)c0f(s) + (double)cp, k.celmin);
    double c = (double)cp;
    if (s > 0.0f) {
        const double geta = (double)cp * (2.0 * (double)s + (double)cp);
        const double h = k.grav * d + geta;
        c = h > 0 ? std::sqrt(h) : (double)c0f(s) + (double)cp;
    }
    return std::fmax(c, std::nextafter(k.celmin, INFINITY));
}

}  // namespace cliffs
