// Default step-kernel choice (autotune.hpp). The table below is written by
// fastprox_amd/autotune.py from measurements on the device; without an entry
// the defaults are the measured best general choices.
//
// SPDX-License-Identifier: MIT
#include "autotune.hpp"

#include <cstdlib>

namespace fpr {

KernelChoice default_kernel(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B) {
    (void)img;
    (void)iso;
    (void)sw;
    (void)B;
    KernelChoice kc;
    if (k != 2) {
        // 3D: the streaming kernel over slabs of 32 slices ("generic" or
        // "s3:SLICES" in FASTPROX_ROCM_KERNEL3D).
        kc.stream = false;
        kc.TT = 1;
        const char *e = std::getenv("FASTPROX_ROCM_KERNEL3D");
        if (e && e[0] == 'g') {
            kc.generic = true;
            kc.RY = 4;
            return kc;
        }
        kc.s3 = true;
        kc.RY = 32;
        if (e && e[0] == 's') {
            const char *c = e;
            while (*c && *c != ':') ++c;
            if (*c == ':' && kernel_available_s3(std::atoi(c + 1))) kc.RY = std::atoi(c + 1);
        }
        return kc;
    }
    // FASTPROX_ROCM_KERNEL: "generic", "stream:TT,SEG" or "block:TT,RY".
    if (const char *e = std::getenv("FASTPROX_ROCM_KERNEL")) {
        if (e[0] == 'g') {
            kc.generic = true;
            kc.stream = false;
            kc.TT = 1;
            kc.RY = 4;
            return kc;
        }
        const bool stream = e[0] == 's';
        const char *c = e;
        while (*c && *c != ':') ++c;
        if (*c == ':') {
            int tt = std::atoi(c + 1), ry = 0;
            for (const char *p = c; *p; ++p)
                if (*p == ',') ry = std::atoi(p + 1);
            if (kernel_available(tsize, stream, tt, ry)) {
                kc.stream = stream;
                kc.TT = tt;
                kc.RY = ry;
                return kc;
            }
        }
    }
    kc.stream = true;
    kc.TT = 4;
    kc.RY = 32;
    return kc;
}

KernelChoice default_tail(int k, int tsize, double *frac) {
    KernelChoice kc;
    *frac = 0.0;
    if (k != 2) return kc;
    // FASTPROX_ROCM_TAIL: "stream:TT,SEG,FRAC" (FRAC 0: no tail kernel).
    if (const char *e = std::getenv("FASTPROX_ROCM_TAIL")) {
        const char *c = e;
        while (*c && *c != ':') ++c;
        if (*c == ':') {
            int v[2] = {0, 0}, n = 0;
            double f = 0.0;
            const char *p = c + 1;
            while (*p && n < 2) {
                v[n++] = std::atoi(p);
                while (*p && *p != ',') ++p;
                if (*p == ',') ++p;
            }
            if (*p) f = std::atof(p);
            if (n == 2 && kernel_available(tsize, true, v[0], v[1])) {
                kc.stream = true;
                kc.TT = v[0];
                kc.RY = v[1];
                *frac = f;
            }
        }
        return kc;
    }
    // Measured on 64 x 512^2 (iso, f32, tol 1e-6): 2598 images/s without a
    // tail kernel, 2716-2750 with segments of 12-16 rows below 30-80% live
    // tiles; neutral on the uneven batch.
    kc.stream = true;
    kc.TT = 4;
    kc.RY = 16;
    *frac = 0.5;
    return kc;
}

}  // namespace fpr
