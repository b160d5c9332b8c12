"""The multispectral image matcher on this build's kernels: the reference's CLI
pipeline (src/cpu/multispectral_ImageMatcher.cu:452-1087) that its MATLAB
harness runs and parses (matlab/experiment_00.m:84-97).

    python -m cugridsearch_amd.matcher --i_ref fixed.png --i_mov moving.png [-m MOD] [--gt GT.csv]

Pipeline, step for step with the reference:

1. load both images as one-channel uint8 (stb's luma (77 R + 150 G + 29 B) >> 8
   for colour files, include/third_party/stb_image.h:1726-1729);
2. when one image's area exceeds 1.5x the other's, resize it to the other's
   width and height with stb_image_resize's stbir_resize_uint8 (:518-557),
   byte for byte on the GPU (cgs_image_resize), and scale the grid's scale
   axes and the ground truth by the size ratios;
3. filter both images with the 3x3 unsharp kernel (:363-366, :620-621);
4. quantise the moving image by MOD when MOD > 0 (:626-631; the search uses
   the filtered image, made before the quantisation, so MOD does not change the
   result, as in the reference);
5. DEPTH coarse-to-fine passes of calcMIAlt over the 8-parameter grid of
   :711-741, each shrinking the window around its minimum (:775-804);
6. a 5 x 1 x 1 x 1 x 5 x 5 x 1 x 1 grid around the minimum, searched again,
   and covariance_by_value around its argmin (:818-896);
7. the stdout protocol: the grid displays (whose first four brace groups
   MATLAB reads as the grid rows), "Minimum found at point p = { ... }", the
   covariance matrix, and Answer[...] Scale_factor[...]
   Homography_to_parameters[...] covar_matrix[...] (:904-989), plus the
   results line appended to imageMatcherResults.txt (:486-490, :964-1062).

8. the output images: the moving image under the answer on the fixed
   image's grid (-o, transformImage_local) and the three-channel overlay of
   the filtered fixed image with the moving image under the answer and under
   the ground truth (-f, fuseAlignedImages_local) (:1072-1076).

The Levenberg-Marquardt refinement is commented out in the reference
(minParams are printed unchanged) and so is not run.
"""
import argparse
import math
import os
import sys
import time

import numpy as np

PI = 3.14159265                     # the reference's #define PI (a double literal)
MAX_SIZE_DISCREPANCY = np.float32(1.5)
MAX_NONOVERLAPPING_PCT = np.float32(0.5)
DEPTH = 1                           # the reference's #define DEPTH
UNSHARP_3X3 = np.array([[0, -1, 0], [-1, 5, -1], [0, -1, 0]], np.float32)
CENTRAL_DIFF_5 = np.array([np.float32(1) / np.float32(12), np.float32(-2) / np.float32(3), 0.0,
                           np.float32(2) / np.float32(3), np.float32(-1) / np.float32(12)], np.float32)
f32 = np.float32


# ------------------------------------------------------------------ images

def load_gray(path):
    """stbi_load(path, ..., req_comp = 1) of a PNG file through cgs_read_png: a
    one-channel uint8 (H, W) array (grey from colour as stb's luma
    (77 R + 150 G + 29 B) >> 8, include/third_party/stb_image.h:1726-1729)."""
    from . import api as A
    return A.read_png(path, 1)


# ------------------------------------------------------------- grid set-up

def coarse_grid(xf, yf, xm, ym, sfx=f32(1), sfy=f32(1)):
    """num_samples, start_point, end_point of the search grid
    (src/cpu/multispectral_ImageMatcher.cu:711-741), in the reference's
    binary32 / binary64 arithmetic."""
    num = [f32(32), f32(16), f32(16), f32(5),
           f32(f32(xf + 1) / f32(f32(20) * sfx)), f32(f32(yf + 1) / f32(f32(20) * sfy)), f32(1), f32(1)]
    start = [f32(0), f32(f32(1.5) * sfx), f32(f32(1.5) * sfy), f32(-0.4),
             f32(f32(-xm) * MAX_NONOVERLAPPING_PCT), f32(f32(-ym) * MAX_NONOVERLAPPING_PCT), f32(0), f32(0)]
    end = [f32(float(f32(2)) * PI - PI / float(num[0])), f32(f32(5) * sfx), f32(f32(5) * sfy), f32(0.2),
           f32(f32(xf) - f32(f32(xm) * MAX_NONOVERLAPPING_PCT)), f32(f32(yf) - f32(f32(ym) * MAX_NONOVERLAPPING_PCT)),
           f32(0), f32(0)]
    return num, start, end


def shrink_window(start, end, best):
    """The DEPTH loop's update (:796-797): start first, then end from the new
    start, each (end - start) / 4 in binary32."""
    start, end = list(start), list(end)
    for d in range(len(best)):
        start[d] = f32(best[d] - f32(f32(end[d] - start[d]) / f32(4)))
        end[d] = f32(best[d] + f32(f32(end[d] - start[d]) / f32(4)))
    return start, end


def covariance_grid(best):
    """The grid around the minimum (:818-851)."""
    num = [f32(5), f32(1), f32(1), f32(1), f32(5), f32(5), f32(1), f32(1)]
    dth = float(f32(1.0)) * PI / 64.0
    start = [f32(float(best[0]) - dth), best[1], best[2], best[3], f32(best[4] - f32(2.5)), f32(best[5] - f32(2.5)),
             f32(0), f32(0)]
    end = [f32(float(best[0]) + dth), best[1], best[2], best[3], f32(best[4] + f32(2.5)), f32(best[5] + f32(2.5)),
           f32(0), f32(0)]
    return num, start, end


# ------------------------------------------------------------ homographies

def _cos(x):
    return f32(math.cos(float(x)))


def _sin(x):
    return f32(math.sin(float(x)))


def parameters_to_homography_norm(p, mw, mh, fw, fh):
    """parametersToHomographyNorm (src/gpu/cudaImageFunctions.cuh:181-214) on
    the host, binary32, cos/sin as the binary64 values rounded to binary32."""
    x_ratio, y_ratio = f32(fw // mw), f32(fh // mh)
    cfx, cfy, cmx, cmy = f32(fw // 2), f32(fh // 2), f32(mw // 2), f32(mw // 2)
    with np.errstate(divide="ignore", invalid="ignore"):
        sx, sy = f32(p[1] / x_ratio), f32(p[2] / y_ratio)
        c, s = _cos(p[0]), _sin(p[0])
        return [f32(sx * c), f32(f32(f32(sy * p[3]) * c) - f32(sy * s)), f32(f32(f32(p[4] - cfx) + cmx) / f32(fw)),
                f32(sx * s), f32(f32(f32(sy * p[3]) * s) + f32(sy * c)), f32(f32(f32(p[5] - cfy) + cmy) / f32(fh)),
                f32(p[6]), f32(p[7])]


def homography_to_parameters(h):
    """homographyToParameters (src/gpu/cudaImageFunctions.cuh:288-313),
    binary32; atan2, cos, sin, sqrt as binary64 rounded to binary32."""
    h11, h12, h13, h21, h22, h23, h31, h32 = [f32(v) for v in h]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        theta = f32(math.atan2(float(h21), float(h11)))
        scale_x = f32(math.sqrt(float(f32(f32(h11 * h11) + f32(h21 * h21)))))
        c, s = _cos(theta), _sin(theta)
        shear_sy = f32(f32(h12 * c) + f32(h22 * s))
        if abs(float(s)) > 1.0e-6:
            scale_y = f32(f32(f32(shear_sy * c) - h12) / s)
        else:
            scale_y = f32(f32(h22 - f32(shear_sy * s)) / c)
        shear = f32(shear_sy / scale_y)
    return [theta, scale_x, scale_y, shear, h13, h23, h31, h32]


# ----------------------------------------------------------------- output

def g(v):
    """std::ostream's default rendering of a float (precision 6, %g), a NaN
    with its sign bit set included ("-nan", as glibc prints it)."""
    x = float(v)
    if math.isnan(x):
        return "-nan" if math.copysign(1.0, x) < 0 else "nan"
    return "%g" % x


def eigen_column(vals):
    """Eigen's operator<< for a column vector: one coefficient per line, each
    right-aligned to the widest."""
    cells = [g(v) for v in vals]
    w = max(len(c) for c in cells)
    return "\n".join(c.rjust(w) for c in cells)


def display_data(a, ax, ay):
    """display_data (:368-375)."""
    return "".join("".join(g(a[x + y * ax]) + ", " for x in range(ax)) + "\n" for y in range(ay))


def conv2_data(a, ax, ay, h, hx, hy, c):
    """conv2_data (:377-391), binary32 accumulation in place into c."""
    for ax0 in range(ax):
        for ay0 in range(ay):
            for hx0 in range(hx):
                for hy0 in range(hy):
                    if 0 <= ax0 - hx0 < ax and 0 <= ay0 - hy0 < ay:
                        c[ax0 + ay0 * ax] = f32(c[ax0 + ay0 * ax] + f32(h[hx0 + hy0 * hx] * a[(ax0 - hx0) + (ay0 - hy0) * ax]))


def sobel_banner():
    """The 5x5 arrays the reference prints while building its (unused) Sobel
    filter (:596-612)."""
    out = []
    z = np.zeros(25, np.float32)
    sob = np.zeros(25, np.float32)
    z[2 * 5 + 0] = 1
    out.append(display_data(z, 5, 5))
    conv2_data(z, 5, 5, CENTRAL_DIFF_5, 5, 1, sob)
    out.append(display_data(sob, 5, 5))
    z[2 * 5 + 0] = 0
    z[0 * 5 + 2] = 1
    out.append(display_data(z, 5, 5))
    conv2_data(z, 5, 5, CENTRAL_DIFF_5, 1, 5, sob)
    z[0 * 5 + 2] = 0
    out.append(display_data(sob, 5, 5))
    for i in range(5):
        sob[i + i * 5] = f32(sob[i + i * 5] + CENTRAL_DIFF_5[i])
    out.append(display_data(sob, 5, 5))
    return "".join(out)


# --------------------------------------------------------------- pipeline

def run(fixed, moved, mod=0.0, gt=None, depth=DEPTH, names=("fixed", "moving"), out=sys.stdout, log=sys.stderr,
        results_file=None, fused_name="output_image_fused.png", out_name=None):
    """Runs the matcher on two uint8 (H, W) arrays; writes the protocol to
    out and returns a dict of the results.  With out_name, also writes the
    output images to out_name and fused_name (:1072-1076)."""
    import torch
    from . import api as A
    from . import _lib as L

    yf, xf = fixed.shape
    ym, xm = moved.shape
    dev = torch.device("cuda")
    img_fix = A.CudaImage.from_numpy(fixed, dev)
    img_mov = A.CudaImage.from_numpy(moved, dev)
    # the larger image (by area, 1.5x) goes to the smaller one's size
    # (:516-557): int area against the binary32 1.5 * x * y
    sfx = sfy = f32(1)
    resized = None
    if f32(xf * yf) > f32(f32(MAX_SIZE_DISCREPANCY * f32(xm)) * f32(ym)):
        sfx, sfy = f32(f32(xm) / f32(xf)), f32(f32(ym) / f32(yf))
        log.write(f"Rescaling {names[0]} from ({xf},{yf}) to ({xm},{ym})\n")
        img_fix = A.resize_image(img_fix, xm, ym)
        xf, yf = xm, ym
        resized = ("image_fixed_resized.png", img_fix)
    elif f32(xm * ym) > f32(f32(MAX_SIZE_DISCREPANCY * f32(xf)) * f32(yf)):
        sfx, sfy = f32(f32(xf) / f32(xm)), f32(f32(yf) / f32(ym))
        log.write(f"Rescaling {names[1]} from ({xm},{ym}) to ({xf},{yf})\n")
        img_mov = A.resize_image(img_mov, xf, yf)
        xm, ym = xf, yf
        resized = ("image_moved_resized.png", img_mov)
    props = torch.cuda.get_device_properties(dev)
    out.write(f'GPU Device 0: "{props.name}" with compute capability {props.major}.{props.minor}\n\n')
    out.write(sobel_banner())
    fix_filtered = A.filter_image(img_fix, UNSHARP_3X3)
    mov_filtered = A.filter_image(img_mov, UNSHARP_3X3)
    if mod > 0.0:
        A.quantize_image(img_mov, mod)
    torch.cuda.synchronize()

    num, start, end = coarse_grid(xf, yf, xm, ym, sfx, sfy)
    line = [f"input files,{names[0]},{names[1]},", f"output fused image,{fused_name},", f"DEPTH,{depth},",
            f"Scale Factor,{g(sfx)},{g(sfy)},"]
    line.append("start_point," + "".join(g(v) + "," for v in start))
    line.append("end_point," + "".join(g(v) + "," for v in end))
    line.append("num_samples," + "".join(g(v) + "," for v in num))
    fn = A.calcMIAlt
    best = [f32(0)] * 8
    t0 = time.perf_counter()
    for _ in range(depth):
        grid = A.CudaGrid(8)
        grid.setStartPoint(start)
        grid.setEndPoint(end)
        grid.setNumSamples(num)
        out.write(grid.display("translation_xy_grid", echo=False) + "\n")
        # one cost per grid point: with fractional counts the reference's tensor
        # (dims truncated per axis) is smaller than the grid and its search writes
        # past it; here every point has its slot and the argmin covers them all
        res = A.CudaTensor([grid.numElements()], device=dev)
        searcher = A.CudaGridSearcher(grid, res)
        searcher.search_by_value(fn, mov_filtered, img_fix, tie_mode=L.TIE_REFERENCE)
        v, i = res.find_extrema(L.TIE_REFERENCE)   # the reference's find_extrema tie order
        p = grid.getGridPoint(i)
        best = [f32(x) for x in p]
        out.write("Minimum found at point p = { " + "".join(g(x) + (", " if d < 7 else " ") for d, x in enumerate(best))
                  + "}\n")
        start, end = shrink_window(start, end, best)
    elapsed = time.perf_counter() - t0
    out.write("Total Time Taken: %f\n" % elapsed)

    cnum, cstart, cend = covariance_grid(best)
    grid = A.CudaGrid(8)
    grid.setStartPoint(cstart)
    grid.setEndPoint(cend)
    grid.setNumSamples(cnum)
    out.write(grid.display("translation_xy_grid", echo=False) + "\n")
    res = A.CudaTensor([grid.numElements()], device=dev)
    searcher = A.CudaGridSearcher(grid, res)
    searcher.search_by_value(fn, mov_filtered, img_fix)
    covar = searcher.covariance_by_value().reshape(-1)
    out.write("Covariance matrix:\n")
    for r in range(8):
        out.write("[" + "".join("%e, " % float(covar[r * 8 + c]) for c in range(8)) + "]\n")
    out.write("x: " + eigen_column(best) + "\n")
    out.write("x that minimizes the function: " + eigen_column(best) + "\n")

    H = parameters_to_homography_norm(best, xm, ym, xf, yf)
    back = homography_to_parameters(H)
    out.write("Answer[" + ",".join(g(v) for v in H) + "]\n")
    out.write("Scale_factor[" + g(sfx) + "," + g(sfy) + "]\n")
    out.write("Homography_to_parameters[" + ",".join(g(v) for v in back) + "]\n")
    out.write("covar_matrix[" + ",".join(g(v) for v in covar) + "]\n")
    out.flush()

    gth = list(H)
    if gt:
        try:
            gth = ground_truth(gt, names[0], sfx, sfy, default=gth)
        except ValueError as e:   # std::stof throws there and the reference terminates
            raise GroundTruthError(f"Ground truth file {gt}: {e}") from e
    line.append("Homography," + "".join(g(v) + "," for v in H))
    line.append("GT Homography," + "".join(g(v) + "," for v in gth))
    with np.errstate(divide="ignore", invalid="ignore"):
        line.append("%% Error," + "".join(g(f32(f32(a - b) / b)) + "," for a, b in zip(H, gth)))
    if results_file:
        with open(results_file, "a") as fh:
            fh.write("".join(line) + "\n")
    # the output images (:1072-1076): the moving image under H on the fixed
    # image's grid, and the overlay of the filtered fixed image with the moving
    # image under H and under the ground truth
    warped = A.warp_image(img_mov, H, xf, yf)
    fused = A.fuse_images(fix_filtered, img_mov, H, gth, channel=0, local=True)
    if out_name is not None:
        write_image(out_name, warped, out)
        write_image(fused_name, fused, out)
    return dict(best=best, homography=H, parameters=back, covariance=covar.reshape(8, 8), seconds=elapsed,
                argmin=(v, i), fixed_filtered=fix_filtered, moving_filtered=mov_filtered, scale_factor=(sfx, sfy),
                resized=resized, warped=warped, fused=fused)


def write_image(path, img, out=sys.stdout):
    """The drivers' suffix rule (src/gpu/cudaImageFunctions.cuh:140-149): .png
    through cgs_write_png, .jpg at quality 95 (through PIL), anything else
    reported on stdout and not written."""
    from . import api as A
    low = path.lower()
    if low.endswith(".png"):
        A.write_png(path, img)
    elif low.endswith(".jpg"):
        from PIL import Image
        Image.fromarray(img.tensor.cpu().numpy()).save(path, quality=95)
    else:
        out.write("Filename suffix has image format not recognized.\n")


class GroundTruthError(ValueError):
    """A ground-truth row whose values do not parse."""


def image_number(fixed_name):
    """The image number the reference reads from the fixed image's file name
    (:993-1000): the characters after the last one outside "0123456789.png",
    up to the next '.', through atoi (the reference copies them into a
    5-byte buffer, which more than four characters overrun; here all of them
    are read)."""
    keep = set("0123456789.png")
    i = len(fixed_name) - 1
    while i >= 0 and fixed_name[i] in keep:
        i -= 1
    j = fixed_name.find(".", i + 1)
    return atoi(fixed_name[i + 1:] if j < 0 else fixed_name[i + 1:j])


def atoi(text):
    """C's atoi: leading white space, an optional sign, then decimal digits
    (0 when there are none)."""
    t = text.lstrip(" \t\n\v\f\r")
    sign, k = 1, 0
    if t[:1] in ("+", "-"):
        sign, k = (-1 if t[0] == "-" else 1), 1
    n = k
    while n < len(t) and t[n].isdigit() and t[n].isascii():
        n += 1
    return sign * int(t[k:n]) if n > k else 0


def stof(text):
    """std::stof: strtof on the text (the decimal rounded once, straight to
    binary32); ValueError when it holds no number."""
    import ctypes
    import errno
    libc = ctypes.CDLL(None, use_errno=True)
    libc.strtof.restype = ctypes.c_float
    libc.strtof.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p)]
    buf = ctypes.create_string_buffer(text.encode("latin-1"))
    start = ctypes.addressof(buf)
    end = ctypes.c_void_p()
    ctypes.set_errno(0)
    v = libc.strtof(start, ctypes.byref(end))
    if end.value == start:
        raise ValueError("stof: no conversion: %r" % text)
    if ctypes.get_errno() == errno.ERANGE:
        raise ValueError("stof: out of range: %r" % text)
    return f32(v)


def ground_truth(path, fixed_name, sfx, sfy, default):
    """The GT CSV lookup (:988-1031).  When the file opens: the first row whose
    leading field (atoi) is the fixed image's number gives eight values
    (stof), read field after field; past the row's last field the reader
    wraps to its start (the reference's std::string::npos + 1 = 0).  The first
    three are then scaled by the x factor and the next three by the y factor,
    whether or not a row matched.  A file that does not open leaves the
    default unscaled."""
    try:
        fh = open(path, newline="", encoding="latin-1")
    except OSError:
        return list(default)
    num = image_number(fixed_name)
    gth = [f32(v) for v in default]
    with fh:
        lines = fh.read().split("\n")
    if lines and lines[-1] == "":   # std::getline yields no line after a final newline
        lines.pop()
    for line in lines:
        comma = line.find(",")
        if atoi(line if comma < 0 else line[:comma]) != num:
            continue
        at = comma + 1 if comma >= 0 else 0
        for i in range(8):
            nxt = line.find(",", at)
            gth[i] = stof(line[at:] if nxt < 0 else line[at:nxt])
            at = 0 if nxt < 0 else nxt + 1
        break
    return [f32(v * sfx) for v in gth[:3]] + [f32(v * sfy) for v in gth[3:6]] + gth[6:8]


def main(argv=None):
    ap = argparse.ArgumentParser(prog="cugridsearch_amd.matcher",
                                 description="Multispectral image matcher (calcMIAlt grid search) on MI355X.")
    ap.add_argument("-m", "--mod", type=float, default=0.0, help="Intensity modifier")
    ap.add_argument("--gt", default=None, help="Ground Truth CSV file")
    ap.add_argument("--i_ref", required=True, help="Reference Image (image in the reference coordinate frame)")
    ap.add_argument("--i_mov", required=True, help="Moved Image (image in the measured coordinate frame)")
    ap.add_argument("-d", "--debug", action="store_true", help="Enable debugging")
    ap.add_argument("-o", "--output", default="output_image.png", help="Output file <output_image.png>")
    ap.add_argument("-f", "--fusedoutput", default="output_image_fused.png",
                    help="Fused output file <output_image_fused.png>")
    ap.add_argument("--depth", type=int, default=DEPTH, help="coarse-to-fine passes (the reference's DEPTH)")
    ap.add_argument("--results-file", default="imageMatcherResults.txt", help="results line appended here")
    args = ap.parse_args(argv)
    sys.stderr.write(f"Output image filename is {args.output}.\n")
    sys.stderr.write(f"Fused Output image filename is {args.fusedoutput}.\n")
    images = []
    for name, role in ((args.i_ref, "Reference"), (args.i_mov, "Moving")):
        try:
            images.append(load_gray(name))
        except Exception as e:   # the reference's "failed to load!" exit (:505-514)
            sys.stderr.write(f"{role} image {name} failed to load! ({e})\n")
            return 1
    fixed, moved = images
    try:
        r = run(fixed, moved, mod=args.mod, gt=args.gt, depth=args.depth, names=(args.i_ref, args.i_mov),
                results_file=args.results_file, fused_name=args.fusedoutput, out_name=args.output)
    except GroundTruthError as e:
        sys.stderr.write(f"{e}\n")
        return 1
    if r["resized"] is not None:   # the reference writes the resized image (:540, :559)
        from . import api as A
        name, img = r["resized"]
        A.write_png(name, img)
    return 0


if __name__ == "__main__":
    sys.exit(main())
