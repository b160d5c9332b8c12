// Drives the host-side code of libcgs under AddressSanitizer and
// UndefinedBehaviorSanitizer on a machine without a GPU: grid bookkeeping,
// sub-boxes, argument validation, and the filter tables that cgs_image_resize
// and cgs_image_fuse(local = 1) build before their first device allocation
// (which then fails for want of a device).  Built host-only (no device code)
// by tests/test_host_sanitize.py; a sanitizer report fails that test.
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "cgs.h"

static int g_fail = 0;

#define EXPECT(cond)                                                           \
    do {                                                                       \
        if (!(cond)) {                                                         \
            std::fprintf(stderr, "%s:%d: EXPECT(%s) failed\n", __FILE__, __LINE__, #cond); \
            g_fail = 1;                                                        \
        }                                                                      \
    } while (0)

static cgs_image fake(uintptr_t addr, int w, int h, int c, int pix) {
    cgs_image im;
    im.data = (const void *) addr;   // never dereferenced on the host
    im.width = w;
    im.height = h;
    im.channels = c;
    im.pixtype = pix;
    return im;
}

static void grids() {
    const float s[8] = {-0.1f, 0.9f, 0.9f, -0.05f, -8, -8, -1e-3f, -1e-3f};
    const float e[8] = {0.1f, 1.1f, 1.1f, 0.05f, 8, 8, 1e-3f, 1e-3f};
    for (int levels = 1; levels <= 7; ++levels) {
        float c[8];
        for (float &v : c) v = (float) levels;
        cgs_grid g;
        EXPECT(cgs_grid_init(&g, 8, s, e, c) == CGS_OK);
        float p[8];
        for (int64_t i = 0; i < g.num_elements; i += 997) EXPECT(cgs_grid_point(&g, i, p) == CGS_OK);
        EXPECT(cgs_grid_point(&g, g.num_elements, p) != CGS_OK);
        int32_t lo[8], cnt[8];
        for (int a = 0; a < 8; ++a) {
            lo[a] = levels / 2;
            cnt[a] = levels - levels / 2;
        }
        cgs_grid sub;
        EXPECT(cgs_grid_subbox(&g, lo, cnt, &sub) == CGS_OK);
        cnt[3] = levels + 1;
        EXPECT(cgs_grid_subbox(&g, lo, cnt, &sub) != CGS_OK);
    }
    const float frac[3] = {2.5f, 3.0f, 1.0f};
    cgs_grid g;
    EXPECT(cgs_grid_init(&g, 3, s, e, frac) == CGS_OK);
    const float huge[2] = {65536.0f, 65536.0f};
    EXPECT(cgs_grid_init(&g, 2, s, e, huge) == CGS_ERR_GRID_TOO_LARGE);
    EXPECT(cgs_grid_init(&g, 0, s, e, frac) != CGS_OK);
    EXPECT(cgs_grid_init(&g, 9, s, e, frac) != CGS_OK);
}

static void validation() {
    cgs_grid g;
    const float s[8] = {0}, e[8] = {1, 1, 1, 1, 1, 1, 1, 1}, c[8] = {2, 2, 2, 2, 2, 2, 2, 2};
    EXPECT(cgs_grid_init(&g, 8, s, e, c) == CGS_OK);
    cgs_search_opts o;
    cgs_default_opts(&o);
    cgs_image m = fake(0x1000, 64, 64, 1, CGS_PIX_U8), f = fake(0x2000, 64, 64, 1, CGS_PIX_U8);
    cgs_argmin am;
    o.cost = 99;
    EXPECT(cgs_search(&g, &m, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);
    cgs_default_opts(&o);
    o.mi_bins = 5;
    o.cost = CGS_COST_MI;
    EXPECT(cgs_search(&g, &m, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);
    cgs_default_opts(&o);
    o.tile_gpt = 3;
    EXPECT(cgs_search(&g, &m, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);
    cgs_default_opts(&o);
    cgs_image big = fake(0x3000, 8192, 4096, 1, CGS_PIX_F32);
    EXPECT(cgs_search(&g, &big, &f, &o, nullptr, &am, nullptr) == CGS_ERR_TOO_LARGE);
    EXPECT(cgs_search_noalloc(&g, &big, &f, &o, nullptr, &am, nullptr) != CGS_OK);
    cgs_image rgb = fake(0x4000, 64, 64, 3, CGS_PIX_U8);
    EXPECT(cgs_search(&g, &rgb, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);
    o.cost = CGS_COST_MI_ALT;
    EXPECT(cgs_search(&g, &m, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);   // needs NORM
    EXPECT(cgs_search(nullptr, &m, &f, &o, nullptr, &am, nullptr) == CGS_ERR_INVALID_ARGUMENT);
}

// Resize and fuse: every size pair builds its tables, then stops at the
// device allocation.
static void tables() {
    const int sides[] = {1, 2, 3, 5, 7, 16, 31, 64, 100, 127, 256, 333, 512, 700};
    for (int a : sides)
        for (int b : sides)
            for (int ch = 1; ch <= 4; ch += 3) {
                cgs_image in = fake(0x10000, a, b, ch, CGS_PIX_U8), out = fake(0x90000, b, a, ch, CGS_PIX_U8);
                (void) cgs_image_resize(&in, &out, nullptr);
            }
    const float I[8] = {1, 0, 0, 0, 1, 0, 0, 0};
    for (int a : sides)
        for (int b : sides) {
            cgs_image ref = fake(0x10000, a, b, 1, CGS_PIX_U8), mov = fake(0x20000, b, a, 1, CGS_PIX_U8);
            cgs_image fused = fake(0x30000, a, b, 3, CGS_PIX_U8);
            (void) cgs_image_fuse(&ref, &mov, I, I, 0, 1, &fused, nullptr);
        }
}

int main() {
    grids();
    validation();
    tables();
    std::printf("host_sanitize %s\n", g_fail ? "FAILED" : "ok");
    return g_fail;
}
