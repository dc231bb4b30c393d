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
        kc.TT = 1;
        kc.RY = 4;
        return kc;
    }
    if (const char *e = std::getenv("FASTPROX_ROCM_KERNEL")) {  // "generic" or "TT,RY"
        if (e[0] == 'g') {
            kc.generic = true;
            kc.TT = 1;
            kc.RY = 4;
            return kc;
        }
        int tt = std::atoi(e), ry = 0;
        for (const char *p = e; *p; ++p)
            if (*p == ',') ry = std::atoi(p + 1);
        if (kernel_available(tsize, tt, ry)) {
            kc.TT = tt;
            kc.RY = ry;
            return kc;
        }
    }
    if (tsize == 4) {
        kc.TT = 4;
        kc.RY = 32;
    } else {
        kc.TT = 2;
        kc.RY = 16;
    }
    return kc;
}

}  // namespace fpr
