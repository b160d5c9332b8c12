// Drives the host-side code of libcgs under AddressSanitizer and
// UndefinedBehaviorSanitizer on a machine without a GPU: grid bookkeeping,
// sub-boxes, argument validation, and the filter tables that cgs_image_resize
// and cgs_image_fuse(local = 1) build before their first device allocation
// (which then fails for want of a device), and the PNG reader on good,
// truncated, corrupted and oversized files.  Built by
// tests/test_host_sanitize.py with the host code instrumented; a sanitizer
// report fails that test.
#include <unistd.h>

#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <cstring>
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

static std::vector<unsigned char> slurp(const char *path) {
    std::vector<unsigned char> b;
    FILE *f = std::fopen(path, "rb");
    if (!f) return b;
    int c;
    while ((c = std::fgetc(f)) != EOF) b.push_back((unsigned char) c);
    std::fclose(f);
    return b;
}

static void spill(const char *path, const std::vector<unsigned char> &b, size_t n) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return;
    std::fwrite(b.data(), 1, n, f);
    std::fclose(f);
}

static uint32_t crc32_png(const unsigned char *p, size_t n) {
    uint32_t c = 0xFFFFFFFFu;
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0xEDB88320u & (0u - (c & 1u)));
    }
    return c ^ 0xFFFFFFFFu;
}

// PNG reading: round trips through cgs_write_png at every channel count and
// req_comp, then every truncation, byte corruptions, and headers whose sizes
// are far larger than the data.
static void png() {
    char path[] = "/tmp/cgs_host_sanitize_XXXXXX";
    const int fd = mkstemp(path);
    EXPECT(fd >= 0);
    if (fd < 0) return;
    close(fd);
    for (int c = 1; c <= 4; ++c) {
        const int w = 37, h = 5;
        std::vector<unsigned char> px((size_t) w * h * c);
        for (size_t i = 0; i < px.size(); ++i) px[i] = (unsigned char) (i * 131 + 7);
        EXPECT(cgs_write_png(path, px.data(), w, h, c) == CGS_OK);
        for (int req = 0; req <= 4; ++req) {
            int32_t W = 0, H = 0, N = 0;
            EXPECT(cgs_read_png(path, req, nullptr, 0, &W, &H, &N) == CGS_OK && W == w && H == h && N == c);
            std::vector<unsigned char> out((size_t) w * h * (req ? req : c));
            EXPECT(cgs_read_png(path, req, out.data(), (int64_t) out.size(), nullptr, nullptr, nullptr) == CGS_OK);
            if (req == 0) EXPECT(out == px);
            EXPECT(cgs_read_png(path, req, out.data(), (int64_t) out.size() - 1, nullptr, nullptr, nullptr) != CGS_OK);
        }
    }
    const std::vector<unsigned char> good = slurp(path);
    std::vector<unsigned char> out(37 * 5 * 4);
    for (size_t n = 0; n < good.size(); ++n) {   // every truncation
        spill(path, good, n);
        (void) cgs_read_png(path, 0, out.data(), (int64_t) out.size(), nullptr, nullptr, nullptr);
    }
    uint32_t r = 12345;
    for (int t = 0; t < 2000; ++t) {   // corrupted bytes past the signature
        std::vector<unsigned char> b = good;
        for (int k = 0; k < 1 + t % 4; ++k) {
            r = r * 1664525u + 1013904223u;
            b[8 + (r >> 8) % (b.size() - 8)] ^= (unsigned char) (1u + (r >> 24) % 255u);
        }
        spill(path, b, b.size());
        int32_t W = 0, H = 0, N = 0;
        if (cgs_read_png(path, 0, nullptr, 0, &W, &H, &N) == CGS_OK && (int64_t) W * H * N <= (1 << 22)) {
            std::vector<unsigned char> o((size_t) W * H * N);
            (void) cgs_read_png(path, 0, o.data(), (int64_t) o.size(), nullptr, nullptr, nullptr);
        }
    }
    // headers far larger than their data (1 x 2^24 and 4096 x 4096, both
    // within the size rules): the decode runs out of data and fails cleanly
    for (uint32_t side : {1u << 24, 4096u}) {
        std::vector<unsigned char> b = good;
        const uint32_t w = side == 4096u ? 4096u : 1u, h = side;
        for (int k = 0; k < 4; ++k) {
            b[16 + k] = (unsigned char) (w >> (24 - 8 * k));
            b[20 + k] = (unsigned char) (h >> (24 - 8 * k));
        }
        const uint32_t crc = crc32_png(&b[12], 17);
        for (int k = 0; k < 4; ++k) b[29 + k] = (unsigned char) (crc >> (24 - 8 * k));
        spill(path, b, b.size());
        int32_t W = 0, H = 0, N = 0;
        if (cgs_read_png(path, 1, nullptr, 0, &W, &H, &N) == CGS_OK) {
            std::vector<unsigned char> o((size_t) W * H);
            EXPECT(cgs_read_png(path, 1, o.data(), (int64_t) o.size(), nullptr, nullptr, nullptr) != CGS_OK);
        }
    }
    std::remove(path);
}

int main() {
    grids();
    validation();
    tables();
    png();
    std::printf("host_sanitize %s\n", g_fail ? "FAILED" : "ok");
    return g_fail;
}
