// cgs_matcher: the multispectral image matcher as a native executable over the
// C-ABI (include/cgs.h).  It is the counterpart of the reference's
// multispectral_ImageMatcher (src/cpu/multispectral_ImageMatcher.cu:452-1087),
// the program its MATLAB harness runs and parses (matlab/experiment_00.m:77-97):
// the same flags (:90-103), the same stdout protocol and the same
// imageMatcherResults.txt line.  The pipeline:
//
//  1. both images loaded as one-channel uint8 with stbi_load semantics
//     (cgs_read_png; :503-514);
//  2. when one image's area exceeds 1.5x the other's, that image resized to the
//     other's size (cgs_image_resize = stbir_resize_uint8, :516-557), its PNG
//     written, and the scale axes and the ground truth scaled by the ratios;
//  3. both images filtered with the 3x3 unsharp kernel (cgs_image_filter,
//     :620-621); the moving image quantised when MOD > 0 (cgs_image_quantize,
//     :626-631; the search reads the filtered copy made before it);
//  4. DEPTH coarse-to-fine calcMIAlt searches over the CLI's grid with the
//     reference's find_extrema tie order and the /4 window update (:711-807);
//  5. the 5 x 1 x 1 x 1 x 5 x 5 x 1 x 1 grid around the minimum, searched, and
//     covariance_by_value around its argmin (:818-934);
//  6. the answer H = parametersToHomographyNorm(minimum), its
//     homographyToParameters, the protocol lines and the results line with the
//     ground truth of --gt (:936-1068);
//  7. the output images: the moving image under H on the fixed image's grid
//     (-o) and the overlay of the filtered fixed image with the moving image
//     under H and under the ground truth (-f) (:1070-1076).
//
// Host arithmetic follows the reference's binary32 / binary64 expressions
// operation for operation (compiled with -ffp-contract=off, so nothing is
// fused); cos, sin and atan2 of a float are the binary64 functions rounded to
// binary32.  The Levenberg-Marquardt step is commented out in the reference
// (:957-959) and is not run.  --depth and --results-file are additions
// (defaults: the reference's DEPTH 1 and its file name).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cgs.h"

namespace {

const double kPi = 3.14159265;   // the reference's #define PI (a double)
const int kDim = 8;
const float kUnsharp[9] = {0, -1.0f, 0, -1.0f, 5.0f, -1.0f, 0, -1.0f, 0};
const float kCentralDiff5[5] = {1.0f / 12.0f, -2.0f / 3.0f, 0.0f, 2.0f / 3.0f, -1.0f / 12.0f};

[[noreturn]] void fail(const std::string &msg) {
    std::fflush(stdout);
    std::fprintf(stderr, "cgs_matcher: %s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
}

void check(cgs_status st, const char *what) {
    if (st != CGS_OK) fail(std::string(what) + ": " + cgs_last_error());
}

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(std::string(what) + ": " + hipGetErrorString(e));
}

// std::ostream's default rendering of a float: %g at precision 6.
std::string fmt(float v) {
    char b[64];
    std::snprintf(b, sizeof b, "%g", (double) v);
    return b;
}

// ------------------------------------------------------------------ options

struct Options {
    float mod = 0.0f;
    std::string gt, i_ref, i_mov;
    bool has_gt = false, has_ref = false, has_mov = false, debug = false, help = false;
    std::string output = "output_image.png", fused = "output_image_fused.png";
    std::string results = "imageMatcherResults.txt";
    int depth = 1;
};

const char *kHelp =
    "UNC Charlotte Machine Vision Lab CUDA-accelerated grid search code.\n"
    "Usage:\n"
    "  cgs_matcher [OPTION...]\n"
    "\n"
    "  -m, --mod arg           Intensity modifier (default: 0)\n"
    "      --gt arg            Ground Truth CSV file\n"
    "      --i_ref arg         Reference Image (image in the reference coordinate\n"
    "                          frame)\n"
    "      --i_mov arg         Moved Image (image in the measured coordinate frame)\n"
    "  -d, --debug             Enable debugging\n"
    "  -o, --output arg        Output file <output_image.png> (default:\n"
    "                          output_image.png)\n"
    "  -f, --fusedoutput arg   Fused output file <output_image_fused.png>\n"
    "                          (default: output_image_fused.png)\n"
    "      --depth arg         coarse-to-fine passes (default: 1)\n"
    "      --results-file arg  results line appended here (default:\n"
    "                          imageMatcherResults.txt)\n"
    "  -h, --help              Print usage\n";

float parse_float(const std::string &name, const std::string &text) {
    char *end = nullptr;
    const float v = std::strtof(text.c_str(), &end);
    if (text.empty() || *end != '\0') fail("option '" + name + "' takes a number, got '" + text + "'");
    return v;
}

Options parse_args(int argc, char **argv) {
    Options o;
    auto long_name = [](const std::string &s) -> std::string {
        if (s == "m") return "mod";
        if (s == "d") return "debug";
        if (s == "o") return "output";
        if (s == "f") return "fusedoutput";
        if (s == "h") return "help";
        return "";
    };
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i], name, value;
        bool inline_value = false;
        if (a.rfind("--", 0) == 0) {
            name = a.substr(2);
            const size_t eq = name.find('=');
            if (eq != std::string::npos) {
                value = name.substr(eq + 1);
                name = name.substr(0, eq);
                inline_value = true;
            }
        } else if (a.size() >= 2 && a[0] == '-') {
            name = long_name(a.substr(1, 1));
            if (name.empty()) fail("option '" + a.substr(1, 1) + "' does not exist");
            if (a.size() > 2) {
                value = a.substr(2);
                inline_value = true;
            }
        } else {
            fail("unexpected argument '" + a + "'");
        }
        if (name == "debug" || name == "help") {
            (name == "debug" ? o.debug : o.help) = true;
            continue;
        }
        if (!inline_value) {
            if (i + 1 >= argc) fail("option '" + name + "' is missing an argument");
            value = argv[++i];
        }
        if (name == "mod") o.mod = parse_float(name, value);
        else if (name == "gt") { o.gt = value; o.has_gt = true; }
        else if (name == "i_ref") { o.i_ref = value; o.has_ref = true; }
        else if (name == "i_mov") { o.i_mov = value; o.has_mov = true; }
        else if (name == "output") o.output = value;
        else if (name == "fusedoutput") o.fused = value;
        else if (name == "results-file") o.results = value;
        else if (name == "depth") {
            char *end = nullptr;
            const long d = std::strtol(value.c_str(), &end, 10);
            if (value.empty() || *end != '\0' || d < 0 || d > 64) fail("option 'depth' takes 0..64");
            o.depth = (int) d;
        } else {
            fail("option '" + name + "' does not exist");
        }
    }
    return o;
}

// ------------------------------------------------------------ device images

struct DevImage {
    cgs_image v{};
    void *ptr = nullptr;
    DevImage() = default;
    DevImage(int w, int h, int c) {
        hip_check(hipMalloc(&ptr, (size_t) w * h * c), "hipMalloc");
        v.data = ptr;
        v.width = w;
        v.height = h;
        v.channels = c;
        v.pixtype = CGS_PIX_U8;
    }
    DevImage(const DevImage &) = delete;
    DevImage &operator=(const DevImage &) = delete;
    DevImage(DevImage &&o) noexcept : v(o.v), ptr(o.ptr) { o.ptr = nullptr; }
    DevImage &operator=(DevImage &&o) noexcept {
        std::swap(v, o.v);
        std::swap(ptr, o.ptr);
        return *this;
    }
    ~DevImage() {
        if (ptr) (void) hipFree(ptr);
    }
    size_t bytes() const { return (size_t) v.width * v.height * v.channels; }
    std::vector<uint8_t> to_host() const {
        std::vector<uint8_t> h(bytes());
        hip_check(hipMemcpy(h.data(), ptr, h.size(), hipMemcpyDeviceToHost), "hipMemcpy");
        return h;
    }
};

std::vector<uint8_t> load_gray(const std::string &path, int &w, int &h) {
    int32_t W = 0, H = 0, n = 0;
    if (cgs_read_png(path.c_str(), 1, nullptr, 0, &W, &H, &n) != CGS_OK) return {};
    std::vector<uint8_t> px((size_t) W * H);
    if (cgs_read_png(path.c_str(), 1, px.data(), (int64_t) px.size(), nullptr, nullptr, nullptr) != CGS_OK) return {};
    w = W;
    h = H;
    return px;
}

bool ends_with_nocase(const std::string &s, const std::string &suffix) {
    if (s.size() < suffix.size()) return false;
    for (size_t i = 0; i < suffix.size(); ++i)
        if (std::tolower((unsigned char) s[s.size() - suffix.size() + i]) != std::tolower((unsigned char) suffix[i]))
            return false;
    return true;
}

// The writers' suffix rule (src/cpu/multispectral_ImageMatcher.cu:217-225).
void write_image(const std::string &path, const DevImage &img) {
    if (ends_with_nocase(path, ".png")) {
        const std::vector<uint8_t> h = img.to_host();
        check(cgs_write_png(path.c_str(), h.data(), img.v.width, img.v.height, img.v.channels), "cgs_write_png");
    } else if (ends_with_nocase(path, ".jpg")) {
        std::fprintf(stderr, "cgs_matcher: no JPEG encoder in this build; %s not written\n", path.c_str());
    } else {
        std::printf("Filename suffix has image format not recognized.\n");
    }
}

// --------------------------------------------------------- host arithmetic

// display_data and conv2_data (:368-391).
void display_data(const float *a, int ax, int ay) {
    for (int y = 0; y < ay; ++y) {
        for (int x = 0; x < ax; ++x) std::printf("%s, ", fmt(a[x + y * ax]).c_str());
        std::printf("\n");
    }
}

void conv2_data(const float *a, int ax, int ay, const float *h, int hx, int hy, float *c) {
    for (int ax0 = 0; ax0 < ax; ++ax0)
        for (int ay0 = 0; ay0 < ay; ++ay0)
            for (int hx0 = 0; hx0 < hx; ++hx0)
                for (int hy0 = 0; hy0 < hy; ++hy0)
                    if (ax0 - hx0 >= 0 && ax0 - hx0 < ax && ay0 - hy0 >= 0 && ay0 - hy0 < ay)
                        c[ax0 + ay0 * ax] += h[hx0 + hy0 * hx] * a[(ax0 - hx0) + (ay0 - hy0) * ax];
}

// The 5x5 arrays printed while the (unused) Sobel filter is built (:596-608).
void sobel_banner() {
    float z[25] = {0}, sob[25] = {0};
    z[2 * 5 + 0] = 1.0f;
    display_data(z, 5, 5);
    conv2_data(z, 5, 5, kCentralDiff5, 5, 1, sob);
    display_data(sob, 5, 5);
    z[2 * 5 + 0] = 0.0f;
    z[0 * 5 + 2] = 1.0f;
    display_data(z, 5, 5);
    conv2_data(z, 5, 5, kCentralDiff5, 1, 5, sob);
    z[0 * 5 + 2] = 0.0f;
    display_data(sob, 5, 5);
    for (int i = 0; i < 5; ++i) sob[i + i * 5] += kCentralDiff5[i];
    display_data(sob, 5, 5);
}

float cos32(float x) { return (float) std::cos((double) x); }
float sin32(float x) { return (float) std::sin((double) x); }

// parametersToHomographyNorm (src/gpu/cudaImageFunctions.cuh:181-214).
void parameters_to_homography_norm(const float *p, int mw, int mh, int fw, int fh, float *H) {
    const float x_ratio = (float) (fw / mw), y_ratio = (float) (fh / mh);
    const float cfx = (float) (fw / 2), cfy = (float) (fh / 2), cmx = (float) (mw / 2), cmy = (float) (mw / 2);
    const float sx = p[1] / x_ratio, sy = p[2] / y_ratio;
    const float c = cos32(p[0]), s = sin32(p[0]);
    H[0] = sx * c;
    H[1] = (sy * p[3]) * c - sy * s;
    H[2] = ((p[4] - cfx) + cmx) / (float) fw;
    H[3] = sx * s;
    H[4] = (sy * p[3]) * s + sy * c;
    H[5] = ((p[5] - cfy) + cmy) / (float) fh;
    H[6] = p[6];
    H[7] = p[7];
}

// homographyToParameters (src/gpu/cudaImageFunctions.cuh:288-313).
void homography_to_parameters(const float *h, float *p) {
    const float theta = (float) std::atan2((double) h[3], (double) h[0]);
    const float scale_x = (float) std::sqrt((double) (h[0] * h[0] + h[3] * h[3]));
    const float c = cos32(theta), s = sin32(theta);
    const float shear_sy = h[1] * c + h[4] * s;
    float scale_y;
    if (std::fabs((double) s) > 1.0e-6)
        scale_y = (shear_sy * c - h[1]) / s;
    else
        scale_y = (h[4] - shear_sy * s) / c;
    p[0] = theta;
    p[1] = scale_x;
    p[2] = scale_y;
    p[3] = shear_sy / scale_y;
    p[4] = h[2];
    p[5] = h[5];
    p[6] = h[6];
    p[7] = h[7];
}

// CudaMatrix::display of the grid (include/cudaTensor.cuh:395-415): its
// start, end, resolution and sample-count rows.
void display_grid(const cgs_grid &g, const char *name) {
    std::printf("Matrix %s 4 x %d elements of f\n\n", name, g.dim);
    const float *rows[4] = {g.start, g.end, g.resolution, g.count};
    for (const float *r : rows) {
        std::printf("{ ");
        for (int d = 0; d < g.dim; ++d) std::printf("%s%s", fmt(r[d]).c_str(), d + 1 < g.dim ? ", " : "");
        std::printf(" }\n");
    }
    std::printf("\n");
}

// Eigen's operator<< for a column vector: one coefficient per line, right-
// aligned to the widest.
std::string eigen_column(const float *v, int n) {
    std::vector<std::string> cells;
    size_t w = 0;
    for (int i = 0; i < n; ++i) {
        cells.push_back(fmt(v[i]));
        w = std::max(w, cells.back().size());
    }
    std::string out;
    for (int i = 0; i < n; ++i) {
        if (i) out += "\n";
        out += std::string(w - cells[i].size(), ' ') + cells[i];
    }
    return out;
}

// C's atoi, as the reference reads image numbers.
int atoi_c(const std::string &s) { return std::atoi(s.c_str()); }

// The image number in the fixed image's file name (:991-999): the characters
// after the last one outside "0123456789.png", up to the next '.'.  The
// reference copies them into a 5-byte buffer, which more than four characters
// overrun; here all of them are read.
int image_number(const std::string &name) {
    size_t i = name.find_last_not_of("0123456789.png");
    i = (i == std::string::npos) ? 0 : i + 1;
    const size_t j = name.find('.', i);
    return atoi_c(name.substr(i, j == std::string::npos ? std::string::npos : j - i));
}

// The ground-truth lookup (:988-1031).  When the file opens, the first line
// whose leading field is the image number gives eight values (std::stof),
// field after field, wrapping to the line's start past its last field; the
// first three are then scaled by the x factor and the next three by the y
// factor, whether or not a line matched.
void ground_truth(const std::string &path, const std::string &fixed_name, float sfx, float sfy, float *gth) {
    std::ifstream in(path.c_str());
    if (!in.is_open()) return;
    const int num = image_number(fixed_name);
    std::string line;
    while (std::getline(in, line)) {
        size_t comma = line.find_first_of(',');
        if (atoi_c(line.substr(0, comma)) != num) continue;
        comma += 1;
        for (int i = 0; i < 8; ++i) {
            size_t next = line.find(',', comma);
            try {
                gth[i] = std::stof(line.substr(comma, next == std::string::npos ? std::string::npos : next - comma));
            } catch (const std::exception &e) {
                fail("Ground truth file " + path + ": " + e.what() + " in a row for image " + std::to_string(num));
            }
            comma = ++next;
        }
        break;
    }
    for (int i = 0; i < 3; ++i) gth[i] *= sfx;
    for (int i = 3; i < 6; ++i) gth[i] *= sfy;
}

}  // namespace

int main(int argc, char **argv) {
    const Options o = parse_args(argc, argv);
    if (!o.has_ref) {
        std::fprintf(stderr, "No input reference image filename was provided. Exiting..\n");
        std::printf("%s\n", kHelp);
        return EXIT_FAILURE;
    }
    if (!o.has_mov) {
        std::fprintf(stderr, "No input moving image filename was provided. Exiting..\n");
        std::printf("%s\n", kHelp);
        return EXIT_FAILURE;
    }
    std::fprintf(stderr, "Output image filename is %s.\n", o.output.c_str());
    std::fprintf(stderr, "Fused Output image filename is %s.\n", o.fused.c_str());
    if (o.help) {
        std::printf("%s\n", kHelp);
        return EXIT_SUCCESS;
    }

    int xf = 0, yf = 0, xm = 0, ym = 0;
    std::vector<uint8_t> dataf = load_gray(o.i_ref, xf, yf);
    if (dataf.empty()) {
        std::fprintf(stderr, "Reference image %s failed to load! (%s)\n", o.i_ref.c_str(), cgs_last_error());
        return EXIT_FAILURE;
    }
    std::vector<uint8_t> datam = load_gray(o.i_mov, xm, ym);
    if (datam.empty()) {
        std::fprintf(stderr, "Moving image %s failed to load! (%s)\n", o.i_mov.c_str(), cgs_last_error());
        return EXIT_FAILURE;
    }

    int dev = 0;
    hip_check(hipSetDevice(0), "hipSetDevice");
    hipDeviceProp_t prop;
    hip_check(hipGetDevice(&dev), "hipGetDevice");
    hip_check(hipGetDeviceProperties(&prop, dev), "hipGetDeviceProperties");
    hipStream_t stream = nullptr;   // the null stream, as the reference's launches

    DevImage img_fix(xf, yf, 1), img_mov(xm, ym, 1);
    hip_check(hipMemcpy(img_fix.ptr, dataf.data(), dataf.size(), hipMemcpyHostToDevice), "hipMemcpy");
    hip_check(hipMemcpy(img_mov.ptr, datam.data(), datam.size(), hipMemcpyHostToDevice), "hipMemcpy");

    // the larger image (by area, over 1.5x) to the smaller one's size (:516-557)
    const float kMaxSizeDiscrepancy = 1.5f;
    float sfx = 1.0f, sfy = 1.0f;
    if ((float) (xf * yf) > kMaxSizeDiscrepancy * (float) xm * (float) ym) {
        sfx = (float) xm / (float) xf;
        sfy = (float) ym / (float) yf;
        std::fprintf(stderr, "Rescaling %s from (%d,%d) to (%d,%d)\n", o.i_ref.c_str(), xf, yf, xm, ym);
        DevImage r(xm, ym, 1);
        check(cgs_image_resize(&img_fix.v, &r.v, stream), "cgs_image_resize");
        img_fix = std::move(r);
        xf = xm;
        yf = ym;
        write_image("image_fixed_resized.png", img_fix);
    } else if ((float) (xm * ym) > kMaxSizeDiscrepancy * (float) xf * (float) yf) {
        sfx = (float) xf / (float) xm;
        sfy = (float) yf / (float) ym;
        std::fprintf(stderr, "Rescaling %s from (%d,%d) to (%d,%d)\n", o.i_mov.c_str(), xm, ym, xf, yf);
        DevImage r(xf, yf, 1);
        check(cgs_image_resize(&img_mov.v, &r.v, stream), "cgs_image_resize");
        img_mov = std::move(r);
        xm = xf;
        ym = yf;
        write_image("image_moved_resized.png", img_mov);
    }

    // helper_cuda.h's findCudaDevice line; a runtime that reports no device
    // name gets the architecture name instead
    std::string dev_name = prop.name;
    if (dev_name.empty()) dev_name = std::string(prop.gcnArchName).substr(0, std::string(prop.gcnArchName).find(':'));
    std::printf("GPU Device %d: \"%s\" with compute capability %d.%d\n\n", dev, dev_name.c_str(), prop.major,
                prop.minor);
    sobel_banner();

    DevImage fix_filtered(xf, yf, 1), mov_filtered(xm, ym, 1);
    check(cgs_image_filter(&img_fix.v, &fix_filtered.v, kUnsharp, 3, 3, stream), "cgs_image_filter");
    check(cgs_image_filter(&img_mov.v, &mov_filtered.v, kUnsharp, 3, 3, stream), "cgs_image_filter");
    if (o.mod > 0.0f) check(cgs_image_quantize(&img_mov.v, o.mod, stream), "cgs_image_quantize");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");

    // the coarse grid (:711-741)
    const float kMaxNonoverlap = 0.5f;
    float num[kDim] = {32.0f, 16.0f, 16.0f, 5.0f, (float) (xf + 1) / (20 * sfx), (float) (yf + 1) / (20 * sfy),
                       1.0f, 1.0f};
    float start[kDim] = {0.0f, 1.5f * sfx, 1.5f * sfy, -0.4f, (float) -xm * kMaxNonoverlap,
                         (float) -ym * kMaxNonoverlap, 0.0f, 0.0f};
    float end[kDim] = {(float) (2.0 * kPi - kPi / (double) num[0]), 5.0f * sfx, 5.0f * sfy, 0.2f,
                       (float) xf - (float) xm * kMaxNonoverlap, (float) yf - (float) ym * kMaxNonoverlap, 0.0f, 0.0f};

    std::string line = "input files," + o.i_ref + "," + o.i_mov + ",";
    line += "output fused image," + o.fused + ",";
    line += "DEPTH," + std::to_string(o.depth) + ",";
    line += "Scale Factor," + fmt(sfx) + "," + fmt(sfy) + ",";
    line += "start_point,";
    for (float v : start) line += fmt(v) + ",";
    line += "end_point,";
    for (float v : end) line += fmt(v) + ",";
    line += "num_samples,";
    for (float v : num) line += fmt(v) + ",";

    cgs_search_opts opts;
    cgs_default_opts(&opts);
    opts.cost = CGS_COST_MI_ALT;
    opts.sampler = CGS_SAMPLER_BILINEAR;
    opts.param_mode = CGS_PARAM_NORM;
    opts.mi_bins = 64;
    opts.tie_mode = CGS_TIE_REFERENCE;

    float best[kDim] = {0};
    void *result = nullptr;
    size_t result_cap = 0;
    auto ensure_result = [&](int64_t n) {
        const size_t bytes = (size_t) std::max<int64_t>(n, 1) * sizeof(float);
        if (bytes <= result_cap) return;
        if (result) hip_check(hipFree(result), "hipFree");
        hip_check(hipMalloc(&result, bytes), "hipMalloc");
        result_cap = bytes;
    };

    const auto t0 = std::chrono::system_clock::now();
    for (int pass = 0; pass < o.depth; ++pass) {
        cgs_grid g;
        check(cgs_grid_init(&g, kDim, start, end, num), "cgs_grid_init");
        display_grid(g, "translation_xy_grid");
        ensure_result(g.num_elements);
        cgs_argmin am;
        check(cgs_search(&g, &mov_filtered.v, &img_fix.v, &opts, (float *) result, &am, stream), "cgs_search");
        cgs_argmin mn;
        check(cgs_find_extrema((const float *) result, g.num_elements, CGS_TIE_REFERENCE, &mn, stream),
              "cgs_find_extrema");
        check(cgs_grid_point(&g, mn.index, best), "cgs_grid_point");
        std::printf("Minimum found at point p = { ");
        for (int d = 0; d < kDim; ++d) {
            std::printf("%s%s", fmt(best[d]).c_str(), d < kDim - 1 ? ", " : " ");
            start[d] = best[d] - (end[d] - start[d]) / 4;
            end[d] = best[d] + (end[d] - start[d]) / 4;
        }
        std::printf("}\n");
    }
    const std::chrono::duration<double> elapsed = std::chrono::system_clock::now() - t0;
    std::printf("Total Time Taken: %f\n", elapsed.count());

    // covariance around the minimum (:818-934)
    const float cnum[kDim] = {5, 1, 1, 1, 5, 5, 1, 1};
    const float cstart[kDim] = {(float) ((double) best[0] - 1.0f * kPi / 64.0f), best[1], best[2], best[3],
                                best[4] - 2.5f, best[5] - 2.5f, 0.0f, 0.0f};
    const float cend[kDim] = {(float) ((double) best[0] + 1.0f * kPi / 64.0f), best[1], best[2], best[3],
                              best[4] + 2.5f, best[5] + 2.5f, 0.0f, 0.0f};
    float covar[kDim * kDim] = {0};
    {
        cgs_grid g;
        check(cgs_grid_init(&g, kDim, cstart, cend, cnum), "cgs_grid_init");
        display_grid(g, "translation_xy_grid");
        ensure_result(g.num_elements);
        cgs_search_opts co = opts;
        co.tie_mode = CGS_TIE_LOWEST;
        cgs_argmin am;
        check(cgs_search(&g, &mov_filtered.v, &img_fix.v, &co, (float *) result, &am, stream), "cgs_search");
        cgs_argmin mn;
        check(cgs_covariance(&g, (const float *) result, covar, &mn, stream), "cgs_covariance");
    }
    std::printf("Covariance matrix:\n");
    for (int i = 0; i < kDim; ++i) {
        std::printf("[");
        for (int d = 0; d < kDim; ++d) std::printf("%e, ", (double) covar[i * kDim + d]);
        std::printf("]\n");
    }
    std::printf("x: %s\n", eigen_column(best, kDim).c_str());
    std::printf("x that minimizes the function: %s\n", eigen_column(best, kDim).c_str());

    float H[kDim], back[kDim];
    parameters_to_homography_norm(best, xm, ym, xf, yf, H);
    homography_to_parameters(H, back);
    std::printf("Answer[");
    for (int i = 0; i < kDim; ++i) std::printf("%s%s", fmt(H[i]).c_str(), i < kDim - 1 ? "," : "");
    std::printf("]\n");
    std::printf("Scale_factor[%s,%s]\n", fmt(sfx).c_str(), fmt(sfy).c_str());
    std::printf("Homography_to_parameters[");
    for (int i = 0; i < kDim; ++i) std::printf("%s%s", fmt(back[i]).c_str(), i < kDim - 1 ? "," : "");
    std::printf("]\n");
    std::printf("covar_matrix[");
    for (int i = 0; i < kDim * kDim; ++i) std::printf("%s%s", fmt(covar[i]).c_str(), i < kDim * kDim - 1 ? "," : "");
    std::printf("]\n");
    std::fflush(stdout);

    float gth[kDim];
    std::memcpy(gth, H, sizeof gth);
    if (o.has_gt) ground_truth(o.gt, o.i_ref, sfx, sfy, gth);
    line += "Homography,";
    for (float v : H) line += fmt(v) + ",";
    line += "GT Homography,";
    for (float v : gth) line += fmt(v) + ",";
    line += "%% Error,";
    for (int i = 0; i < kDim; ++i) line += fmt((H[i] - gth[i]) / gth[i]) + ",";
    {
        std::ofstream rf(o.results.c_str(), std::ios_base::app);
        if (!rf.is_open()) fail("cannot open " + o.results);
        rf << line << "\n";
    }

    // the output images (:1070-1076), both zero-filled first
    DevImage warped(xf, yf, 1), fused(xf, yf, 3);
    hip_check(hipMemset(warped.ptr, 0, warped.bytes()), "hipMemset");
    hip_check(hipMemset(fused.ptr, 0, fused.bytes()), "hipMemset");
    check(cgs_image_warp(&img_mov.v, &warped.v, H, stream), "cgs_image_warp");
    check(cgs_image_fuse(&fix_filtered.v, &img_mov.v, H, gth, 0, 1, &fused.v, stream), "cgs_image_fuse");
    hip_check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    write_image(o.output, warped);
    write_image(o.fused, fused);
    std::fflush(stdout);
    if (result) (void) hipFree(result);
    return EXIT_SUCCESS;
}
