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
        kc.generic = true;
        kc.stream = false;
        kc.TT = 1;
        kc.RY = 4;
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

}  // namespace fpr
