// PNG reading: the matcher loads its two input images with stb_image v2.27's
// stbi_load(path, &w, &h, &n, CHANNELS) (src/cpu/multispectral_ImageMatcher.cu:
// 503-514; the PNG path is include/third_party/stb_image.h:5023-5229).
//
// The decoder follows the PNG specification (chunk stream, one zlib stream over
// the IDAT chunks, the five scanline filters, Adam7 interlacing).  What stb
// adds on top of the specification is kept as stb does it:
//   * zlib header checks of stbi__parse_zlib_header (:4361-4373); neither the
//     Adler-32 trailer nor the chunk CRCs are checked, and data past the last
//     scanline is ignored (:4614-4617);
//   * grey samples of 1, 2 or 4 bits scaled by 0xFF, 0x55, 0x11 (:4591, :4727);
//     palette indices are not scaled;
//   * tRNS: a grey or RGB file gains an alpha channel, 0 where the pixel equals
//     the key and full elsewhere (:5103-5113, :4851-4899); a palette file
//     expands to RGB, to RGBA when it has tRNS, or to req_comp channels when
//     req_comp >= 3 (:5171-5178, :4901-4936);
//   * the channel conversions of stbi__convert_format (:1735-1779), grey from
//     colour as (77 R + 150 G + 29 B) >> 8 (:1726-1729), computed on the 16-bit
//     samples of a 16-bit file (:1787-1791) before its 16 -> 8 step v >> 8
//     (:1170-1184);
//   * a CgBI chunk (Apple's variant) announces raw deflate data without the
//     zlib header; samples stay in file order (stb's BGR swap is off by default,
//     :4939).
// Colour and grey-alpha files of fewer than 8 bits per sample, which the
// specification forbids, are refused (stb decodes some of them).  Palette
// indices past the palette read black, opaque.  Host code only; zlib inflates.
#include <zlib.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "cgs.h"
#include "cgs_host.hpp"

using namespace cgs;

namespace {

constexpr uint32_t kMaxSide = 1u << 24;                 // STBI_MAX_DIMENSIONS
const uint8_t kDepthScale[9] = {0, 0xFF, 0x55, 0, 0x11, 0, 0, 0, 0x01};

struct Png {
    uint32_t w = 0, h = 0;
    int depth = 0, color = 0, interlace = 0;
    int img_n = 0;            // samples per pixel in the scanlines (1 for a palette file)
    int pal_n = 0;            // palette files: 3, or 4 with tRNS; 0 otherwise
    uint32_t pal_len = 0;
    uint8_t palette[256][4];
    bool has_trans = false;
    uint16_t key[3] = {0, 0, 0};
    bool raw_deflate = false;  // CgBI
    std::vector<uint8_t> idat;
};

uint32_t get32(const uint8_t *p) {
    return ((uint32_t) p[0] << 24) | ((uint32_t) p[1] << 16) | ((uint32_t) p[2] << 8) | p[3];
}

// Walks the chunks.  Stops after IEND, or at the first IDAT when only the
// header facts are wanted.  Returns an empty string or the error.
std::string parse(const std::vector<uint8_t> &f, Png &P, bool header_only) {
    static const uint8_t sig[8] = {137, 80, 78, 71, 13, 10, 26, 10};
    if (f.size() < 8 || std::memcmp(f.data(), sig, 8) != 0) return "not a PNG file";
    for (int i = 0; i < 256; ++i) {
        P.palette[i][0] = P.palette[i][1] = P.palette[i][2] = 0;
        P.palette[i][3] = 255;
    }
    bool first = true, seen_idat = false;
    size_t at = 8;
    for (;;) {
        if (f.size() - at < 8) return "truncated chunk header";
        const uint32_t len = get32(&f[at]);
        const uint8_t *type = &f[at + 4];
        if (len > f.size() - at - 8) return "truncated chunk";
        const uint8_t *d = &f[at + 8];
        auto is = [&](const char *t) { return std::memcmp(type, t, 4) == 0; };
        if (is("CgBI")) {
            P.raw_deflate = true;
        } else if (is("IHDR")) {
            if (!first) return "multiple IHDR";
            first = false;
            if (len != 13) return "bad IHDR length";
            P.w = get32(d);
            P.h = get32(d + 4);
            if (P.w > kMaxSide || P.h > kMaxSide) return "image too large";
            P.depth = d[8];
            P.color = d[9];
            if (P.depth != 1 && P.depth != 2 && P.depth != 4 && P.depth != 8 && P.depth != 16)
                return "bit depth must be 1, 2, 4, 8 or 16";
            if (P.color > 6 || (P.color & 1 && P.color != 3)) return "bad colour type";
            if (P.color == 3 && P.depth == 16) return "bad colour type";
            if (P.color != 0 && P.color != 3 && P.depth < 8) return "bit depth below 8 for a colour or alpha image";
            if (d[10]) return "bad compression method";
            if (d[11]) return "bad filter method";
            if (d[12] > 1) return "bad interlace method";
            P.interlace = d[12];
            if (!P.w || !P.h) return "0-pixel image";
            if (P.color == 3) {
                P.pal_n = 3;
                P.img_n = 1;
                if ((1u << 30) / P.w / 4 < P.h) return "image too large";
            } else {
                P.img_n = (P.color & 2 ? 3 : 1) + (P.color & 4 ? 1 : 0);
                if ((1u << 30) / P.w / (uint32_t) P.img_n < P.h) return "image too large";
            }
        } else if (is("PLTE")) {
            if (first) return "first chunk not IHDR";
            if (len > 256 * 3 || len % 3) return "bad PLTE";
            P.pal_len = len / 3;
            for (uint32_t i = 0; i < P.pal_len; ++i)
                for (int c = 0; c < 3; ++c) P.palette[i][c] = d[3 * i + c];
        } else if (is("tRNS")) {
            if (first) return "first chunk not IHDR";
            if (seen_idat) return "tRNS after IDAT";
            if (P.pal_n) {
                if (P.pal_len == 0) return "tRNS before PLTE";
                if (len > P.pal_len) return "bad tRNS length";
                P.pal_n = 4;
                for (uint32_t i = 0; i < len; ++i) P.palette[i][3] = d[i];
            } else {
                if (!(P.img_n & 1)) return "tRNS with an alpha channel";
                if (len != (uint32_t) P.img_n * 2) return "bad tRNS length";
                P.has_trans = true;
                for (int k = 0; k < P.img_n; ++k) {
                    const uint16_t v = (uint16_t) ((d[2 * k] << 8) | d[2 * k + 1]);
                    P.key[k] = P.depth == 16 ? v : (uint16_t) ((v & 255) * kDepthScale[P.depth]);
                }
            }
        } else if (is("IDAT")) {
            if (first) return "first chunk not IHDR";
            if (P.pal_n && !P.pal_len) return "no PLTE";
            if (header_only) return "";
            seen_idat = true;
            P.idat.insert(P.idat.end(), d, d + len);
        } else if (is("IEND")) {
            if (first) return "first chunk not IHDR";
            if (header_only) return "";
            if (!seen_idat) return "no IDAT";
            return "";
        } else {
            if (first) return "first chunk not IHDR";
            if (!(type[0] & 0x20)) return std::string("unknown critical chunk ") + std::string((const char *) type, 4);
        }
        at += 8 + (size_t) len;
        if (f.size() - at < 4) return "truncated chunk CRC";
        at += 4;
    }
}

// Inflates the IDAT stream into exactly `need` bytes (later data ignored).
std::string inflate_idat(const Png &P, size_t need, std::vector<uint8_t> &out) {
    const uint8_t *src = P.idat.data();
    size_t n = P.idat.size();
    if (!P.raw_deflate) {
        if (n < 2) return "bad zlib header";
        const int cmf = src[0], flg = src[1];
        if ((cmf * 256 + flg) % 31 != 0) return "bad zlib header";
        if (flg & 32) return "preset dictionary";
        if ((cmf & 15) != 8) return "bad compression";
        src += 2;
        n -= 2;
    }
    out.assign(need, 0);
    if (need == 0) return "";
    z_stream z;
    std::memset(&z, 0, sizeof z);
    if (inflateInit2(&z, -15) != Z_OK) return "inflateInit2 failed";
    size_t in_at = 0, out_at = 0;
    std::string err;
    while (out_at < need) {
        const size_t in_chunk = std::min<size_t>(n - in_at, 1u << 30);
        const size_t out_chunk = std::min<size_t>(need - out_at, 1u << 30);
        z.next_in = const_cast<Bytef *>(src + in_at);
        z.avail_in = (uInt) in_chunk;
        z.next_out = out.data() + out_at;
        z.avail_out = (uInt) out_chunk;
        const int r = inflate(&z, Z_NO_FLUSH);
        in_at += in_chunk - z.avail_in;
        out_at += out_chunk - z.avail_out;
        if (r == Z_STREAM_END || r == Z_BUF_ERROR) break;   // the stream ended, or its data ran out
        if (r != Z_OK) {
            err = std::string("corrupt deflate data") + (z.msg ? std::string(": ") + z.msg : std::string());
            break;
        }
    }
    inflateEnd(&z);
    if (!err.empty()) return err;
    if (out_at < need) return "not enough pixels";
    return "";
}

int paeth(int a, int b, int c) {
    const int p = a + b - c, pa = p > a ? p - a : a - p, pb = p > b ? p - b : b - p, pc = p > c ? p - c : c - p;
    if (pa <= pb && pa <= pc) return a;
    return pb <= pc ? b : c;
}

size_t row_bytes(const Png &P, uint32_t w) { return ((size_t) P.img_n * w * P.depth + 7) >> 3; }

// Unfilters one (sub)image of pw x ph pixels from raw (its rows, each led by
// its filter byte) and stores its samples at (x0 + i*dx, y0 + j*dy) of px.
std::string unfilter(const Png &P, const uint8_t *raw, uint32_t pw, uint32_t ph, uint32_t x0, uint32_t y0,
                     uint32_t dx, uint32_t dy, std::vector<uint16_t> &px) {
    const size_t rb = row_bytes(P, pw);
    const int fb = P.depth >= 8 ? P.img_n * P.depth / 8 : 1;
    std::vector<uint8_t> prev(rb, 0), cur(rb);
    const uint8_t scale = P.color == 0 ? kDepthScale[P.depth] : 1;
    const size_t per_row = (size_t) pw * P.img_n;
    for (uint32_t j = 0; j < ph; ++j) {
        const uint8_t *r = raw + j * (rb + 1);
        const int ft = r[0];
        ++r;
        if (ft > 4) return "invalid scanline filter";
        for (size_t k = 0; k < rb; ++k) {
            const int a = k >= (size_t) fb ? cur[k - fb] : 0;
            const int b = prev[k];
            const int c = k >= (size_t) fb ? prev[k - fb] : 0;
            int p = 0;
            switch (ft) {
                case 0: p = 0; break;
                case 1: p = a; break;
                case 2: p = b; break;
                case 3: p = (a + b) >> 1; break;
                default: p = paeth(a, b, c); break;
            }
            cur[k] = (uint8_t) (r[k] + p);
        }
        const size_t y = (size_t) y0 + (size_t) j * dy;
        for (size_t s = 0; s < per_row; ++s) {
            uint16_t v;
            if (P.depth == 16) {
                v = (uint16_t) ((cur[2 * s] << 8) | cur[2 * s + 1]);
            } else if (P.depth == 8) {
                v = cur[s];
            } else {
                const size_t bit = s * P.depth;
                const int shift = 8 - P.depth - (int) (bit & 7);
                v = (uint16_t) (((cur[bit >> 3] >> shift) & ((1 << P.depth) - 1)) * scale);
            }
            const size_t x = (size_t) x0 + (s / P.img_n) * dx;
            px[(y * P.w + x) * P.img_n + s % P.img_n] = v;
        }
        std::swap(prev, cur);
    }
    return "";
}

struct Pass {
    uint32_t x0, y0, dx, dy;
};
const Pass kAdam7[7] = {{0, 0, 8, 8}, {4, 0, 8, 8}, {0, 4, 4, 8}, {2, 0, 4, 4}, {0, 2, 2, 4}, {1, 0, 2, 2}, {0, 1, 1, 2}};

uint32_t pass_extent(uint32_t n, uint32_t o, uint32_t s) { return n > o ? (n - o + s - 1) / s : 0; }

std::string decode_samples(const Png &P, std::vector<uint16_t> &px) {
    size_t need = 0;
    if (!P.interlace) {
        need = (row_bytes(P, P.w) + 1) * P.h;
    } else {
        for (const Pass &q : kAdam7) {
            const uint32_t pw = pass_extent(P.w, q.x0, q.dx), ph = pass_extent(P.h, q.y0, q.dy);
            if (pw && ph) need += (row_bytes(P, pw) + 1) * ph;
        }
    }
    std::vector<uint8_t> raw;
    std::string err = inflate_idat(P, need, raw);
    if (!err.empty()) return err;
    px.assign((size_t) P.w * P.h * P.img_n, 0);
    if (!P.interlace) return unfilter(P, raw.data(), P.w, P.h, 0, 0, 1, 1, px);
    size_t off = 0;
    for (const Pass &q : kAdam7) {
        const uint32_t pw = pass_extent(P.w, q.x0, q.dx), ph = pass_extent(P.h, q.y0, q.dy);
        if (!pw || !ph) continue;
        err = unfilter(P, raw.data() + off, pw, ph, q.x0, q.y0, q.dx, q.dy, px);
        if (!err.empty()) return err;
        off += (row_bytes(P, pw) + 1) * ph;
    }
    return "";
}

// The channels stbi_load returns before any req_comp conversion, and the
// count it reports as the file's.
int file_channels(const Png &P) { return P.pal_n ? P.pal_n : P.img_n + (P.has_trans ? 1 : 0); }

int out_channels(const Png &P, int req_comp) { return req_comp ? req_comp : file_channels(P); }

}  // namespace

extern "C" cgs_status cgs_read_png(const char *path, int32_t req_comp, void *pixels, int64_t capacity,
                                   int32_t *width, int32_t *height, int32_t *channels_in_file) {
    clear_error();
    if (!path || req_comp < 0 || req_comp > 4 || (pixels && capacity < 0))
        return set_error(CGS_ERR_INVALID_ARGUMENT, "read_png: bad arguments");
    FILE *fp = std::fopen(path, "rb");
    if (!fp) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("read_png: cannot open ") + path);
    std::vector<uint8_t> f;
    uint8_t buf[1 << 16];
    size_t got;
    while ((got = std::fread(buf, 1, sizeof buf, fp)) > 0) f.insert(f.end(), buf, buf + got);
    const bool bad = std::ferror(fp) != 0;
    std::fclose(fp);
    if (bad) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("read_png: cannot read ") + path);
    Png P;
    std::string err = parse(f, P, pixels == nullptr);
    if (!err.empty()) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("read_png: ") + path + ": " + err);
    const int n_out = out_channels(P, req_comp);
    if (width) *width = (int32_t) P.w;
    if (height) *height = (int32_t) P.h;
    if (channels_in_file) *channels_in_file = file_channels(P);
    if (!pixels) return CGS_OK;
    const size_t npx = (size_t) P.w * P.h;
    if ((uint64_t) capacity < (uint64_t) npx * n_out)
        return set_error(CGS_ERR_INVALID_ARGUMENT, "read_png: pixel buffer of " + std::to_string(capacity) +
                                                       " bytes, the image needs " + std::to_string(npx * n_out));
    std::vector<uint16_t> px;
    err = decode_samples(P, px);
    if (!err.empty()) return set_error(CGS_ERR_INVALID_ARGUMENT, std::string("read_png: ") + path + ": " + err);

    // stb's intermediate image: the palette expanded, or the key's alpha added
    const int mid_n = P.pal_n ? (req_comp >= 3 ? req_comp : P.pal_n) : file_channels(P);
    const uint16_t full = P.depth == 16 ? 0xFFFF : 0xFF;
    std::vector<uint16_t> mid(npx * mid_n);
    for (size_t i = 0; i < npx; ++i) {
        uint16_t *o = &mid[i * mid_n];
        if (P.pal_n) {
            const uint8_t *e = P.palette[px[i]];
            for (int c = 0; c < mid_n; ++c) o[c] = e[c];
        } else {
            const uint16_t *s = &px[i * P.img_n];
            for (int c = 0; c < P.img_n; ++c) o[c] = s[c];
            if (P.has_trans) {
                const bool hit = P.img_n == 1 ? s[0] == P.key[0]
                                              : s[0] == P.key[0] && s[1] == P.key[1] && s[2] == P.key[2];
                o[P.img_n] = hit ? 0 : full;
            }
        }
    }
    // stbi__convert_format(16): mid_n -> n_out channels
    uint8_t *dst = (uint8_t *) pixels;
    const int down = P.depth == 16 ? 8 : 0;
    auto luma = [&](const uint16_t *s) {
        const uint32_t y = ((uint32_t) s[0] * 77 + (uint32_t) s[1] * 150 + (uint32_t) s[2] * 29) >> 8;
        return (uint16_t) (P.depth == 16 ? y & 0xFFFF : y & 0xFF);
    };
    for (size_t i = 0; i < npx; ++i) {
        const uint16_t *s = &mid[i * mid_n];
        uint16_t o[4] = {0, 0, 0, 0};
        if (n_out == mid_n) {
            for (int c = 0; c < n_out; ++c) o[c] = s[c];
        } else {
            switch (mid_n * 8 + n_out) {
                case 1 * 8 + 2: o[0] = s[0]; o[1] = full; break;
                case 1 * 8 + 3: o[0] = o[1] = o[2] = s[0]; break;
                case 1 * 8 + 4: o[0] = o[1] = o[2] = s[0]; o[3] = full; break;
                case 2 * 8 + 1: o[0] = s[0]; break;
                case 2 * 8 + 3: o[0] = o[1] = o[2] = s[0]; break;
                case 2 * 8 + 4: o[0] = o[1] = o[2] = s[0]; o[3] = s[1]; break;
                case 3 * 8 + 1: o[0] = luma(s); break;
                case 3 * 8 + 2: o[0] = luma(s); o[1] = full; break;
                case 3 * 8 + 4: o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; o[3] = full; break;
                case 4 * 8 + 1: o[0] = luma(s); break;
                case 4 * 8 + 2: o[0] = luma(s); o[1] = s[3]; break;
                case 4 * 8 + 3: o[0] = s[0]; o[1] = s[1]; o[2] = s[2]; break;
                default: return set_error(CGS_ERR_UNSUPPORTED, "read_png: unsupported channel conversion");
            }
        }
        for (int c = 0; c < n_out; ++c) dst[i * n_out + c] = (uint8_t) (o[c] >> down);
    }
    return CGS_OK;
}
