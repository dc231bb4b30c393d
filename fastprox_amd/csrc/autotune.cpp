// The step-kernel choice (autotune.hpp): a tuning table measured on the
// device by fastprox_amd/autotune.py, read once from tuning.txt beside this
// library (or the file FASTPROX_ROCM_TUNING names). Each line is
//     k dtype_bytes iso spatial max_pixels kind TT RY tail_TT tail_RY tail_frac [max_images]
// (kind 0: the streaming 2D kernel, 1: generic, 2: block, 3: the 3D streaming
// kernel; max_images, when present and not 0, restricts the line to batches
// of at most that many images: a lone image fills the GPU only with short
// tiles); the entry for a batch of B images of n pixels is the first line of
// its class with n <= max_pixels that admits B (the last line of the class
// without an image limit beyond that). Without a table the built-in defaults
// below apply.
// FASTPROX_ROCM_KERNEL / FASTPROX_ROCM_KERNEL3D / FASTPROX_ROCM_TAIL override
// everything (for sweeps).
//
// SPDX-License-Identifier: MIT
#include "autotune.hpp"

#include <dlfcn.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace fpr {

namespace {

struct Entry {
    int k, tsize, iso, sw;
    long long max_px;
    int kind, TT, RY, tTT, tRY;
    double tfrac;
    long long max_images;  // 0: any batch
};

std::string table_path() {
    if (const char *e = std::getenv("FASTPROX_ROCM_TUNING")) return e;
    Dl_info info;
    if (dladdr((void *)&table_path, &info) && info.dli_fname) {
        std::string f = info.dli_fname;
        const size_t cut = f.find_last_of('/');
        return (cut == std::string::npos ? std::string() : f.substr(0, cut + 1)) + "tuning.txt";
    }
    return "tuning.txt";
}

std::vector<Entry> load_table() {
    std::vector<Entry> t;
    FILE *f = std::fopen(table_path().c_str(), "r");
    if (!f) return t;
    char line[512];
    while (std::fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n') continue;
        Entry e{};
        const int got = std::sscanf(line, "%d %d %d %d %lld %d %d %d %d %d %lf %lld", &e.k, &e.tsize, &e.iso, &e.sw,
                                    &e.max_px, &e.kind, &e.TT, &e.RY, &e.tTT, &e.tRY, &e.tfrac, &e.max_images);
        if (got == 11) e.max_images = 0;
        if (got >= 11) t.push_back(e);
    }
    std::fclose(f);
    return t;
}

const std::vector<Entry> &table() {
    static const std::vector<Entry> t = load_table();
    return t;
}

const Entry *lookup(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B) {
    long long n = 1;
    for (int a = 0; a < 3; ++a) n *= img ? (long long)img[a] : 1LL;
    const Entry *last = nullptr;
    for (const Entry &e : table()) {
        if (e.k != k || e.tsize != tsize || e.iso != (iso ? 1 : 0) || e.sw != (sw ? 1 : 0)) continue;
        if (e.max_images > 0 && B > e.max_images) continue;
        if (e.max_images == 0) last = &e;
        if (n <= e.max_px) return &e;
    }
    return last;
}

bool entry_kernel(const Entry &e, int tsize, int64_t n, KernelChoice &kc) {
    kc = KernelChoice{};
    switch (e.kind) {
        case 0:
        case 2:
            kc.stream = e.kind == 0;
            kc.TT = e.TT;
            kc.RY = e.RY;
            return e.k == 2 && kernel_available(tsize, kc.stream, e.TT, e.RY);
        case 1:
            kc.generic = true;
            kc.stream = false;
            kc.TT = 1;
            kc.RY = 4;
            return true;
        case 3:
            kc.stream = false;
            kc.s3 = true;
            kc.TT = e.TT < 1 ? 1 : e.TT;
            kc.RY = e.RY;
            return e.k == 3 && kernel_available_s3(e.RY, kc.TT, tsize, e.sw, n);
    }
    return false;
}

}  // namespace

KernelChoice default_kernel(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B) {
    KernelChoice kc;
    if (k != 2) {
        kc.stream = false;
        kc.TT = 1;
        if (const char *e = std::getenv("FASTPROX_ROCM_KERNEL3D")) {  // "generic" or "s3:SLICES[,TT]"
            if (e[0] == 'g') {
                kc.generic = true;
                kc.RY = 4;
                return kc;
            }
            const char *c = e;
            while (*c && *c != ':') ++c;
            if (*c == ':') {
                const int ry = std::atoi(c + 1);
                int tt = 1;
                for (const char *q = c; *q; ++q)
                    if (*q == ',') tt = std::atoi(q + 1);
                if (kernel_available_s3(ry, tt, tsize, sw, img[0] * img[1] * img[2])) {
                    kc.s3 = true;
                    kc.RY = ry;
                    kc.TT = tt;
                    return kc;
                }
            }
        }
        if (const Entry *t = lookup(k, tsize, img, iso, sw, B))
            if (entry_kernel(*t, tsize, img[0] * img[1] * img[2], kc)) return kc;
        kc = KernelChoice{};
        kc.stream = false;
        kc.s3 = true;
        kc.TT = 1;
        kc.RY = 32;
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
    if (const Entry *t = lookup(k, tsize, img, iso, sw, B))
        if (entry_kernel(*t, tsize, img[0] * img[1] * img[2], kc)) return kc;
    kc = KernelChoice{};
    kc.stream = true;
    kc.TT = 4;
    kc.RY = 32;
    return kc;
}

// 2D: the batch kernel of 64 images of 512^2 (autotuned; on MI355X TT 6 with
// 40-row segments). Measured on 2048 queued 512^2 requests and on the served
// trace (Poisson 200 requests/s for 20 s, 80% 2D of side 128..1024, 20% 3D of
// side 48..128, f32) when the table chose TT 6 with 32 rows, against 4
// iterations and 16 rows:
//   TT 6, 32 rows: 3,958 images/s (4,129 steady), trace p50 0.89 ms, p99 40 ms
//   TT 4, 32 rows: 3,759 (4,064),                 p50 0.81 ms, p99 40 ms
//   TT 4, 16 rows: 3,364 (3,624),                 p50 0.67 ms, p99 37 ms
// Short launches help a lone request, long ones a full server. 3D: slabs of
// 8 slices (the 3D requests set the trace's p99; the batch kernel's 16-slice
// slabs lengthen it), two iterations per launch for float32 (the same trace
// for 20 s: p99 37.3 -> 30.4 ms, the 3D requests' p99 54.2 -> 47.2 ms against
// one iteration per launch; 16-slice slabs 30.8 / 47.3 ms).
KernelChoice serve_kernel(int k, int tsize, bool iso) {
    const char *env = std::getenv(k == 2 ? "FASTPROX_ROCM_KERNEL" : "FASTPROX_ROCM_KERNEL3D");
    if (env || k == 2) {
        const int64_t img[3] = {k == 3 ? 128 : 1, 512, 512};
        return default_kernel(k, tsize, img, iso, false, 64);
    }
    KernelChoice kc;
    kc.TT = kernel_available_s3(8, 2, tsize, false) ? 2 : 1;
    kc.RY = 8;
    kc.stream = false;
    kc.s3 = true;
    return kc;
}

KernelChoice default_tail(int k, int tsize, const int64_t *img, bool iso, bool sw, int64_t B, double *frac) {
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
    if (const Entry *t = lookup(k, tsize, img, iso, sw, B)) {
        if (t->tfrac > 0.0 && kernel_available(tsize, true, t->tTT, t->tRY)) {
            kc.stream = true;
            kc.TT = t->tTT;
            kc.RY = t->tRY;
            *frac = t->tfrac;
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
