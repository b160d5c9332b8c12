"""Pins the oracle's line-traversal ("Alt") costs against closed forms and an
independent float64 numpy restatement built from the geometry (inverse
homography, border crossings) rather than from the reference's float
expressions.  CPU only."""
import math

import numpy as np
import pytest

from cugridsearch_amd import synth


def _centred_fp64(p, cx, cy):
    th, sx, sy, sh, tx, ty, kx, ky = [float(v) for v in p]
    C = np.array([[1, 0, cx], [0, 1, cy], [0, 0, 1]], float)
    Ci = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1]], float)
    K = np.array([[1, 0, 0], [0, 1, 0], [kx, ky, 1]], float)
    T = np.array([[1, 0, tx], [0, 1, ty], [0, 0, 1]], float)
    R = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]], float)
    Sh = np.array([[1, sh, 0], [0, 1, 0], [0, 0, 1]], float)
    Sc = np.diag([sx, sy, 1.0])
    H = C @ K @ T @ R @ Sh @ Sc @ Ci
    return H / H[2, 2]


def _rand_params(rng, t=5.0, k=1e-3):
    return np.array([rng.uniform(-0.1, 0.1), rng.uniform(0.9, 1.1), rng.uniform(0.9, 1.1), rng.uniform(-.05, .05),
                     rng.uniform(-t, t), rng.uniform(-t, t), rng.uniform(-k, k), rng.uniform(-k, k)], np.float32)


def _inv(H):
    Hi = np.linalg.inv(H)

    def f(x, y):
        q = Hi @ np.array([x, y, 1.0])
        return q[:2] / q[2]
    return f


def _crossings(p0, v, mw, mh, fw, fh, eps=0.0):
    """The reference's border rules (cudaImageFunctions.cuh:352-375) applied to
    float64 geometry: top/bottom crossings qualify with the crossing x in
    [0, mw) and lambda in [0, fw); left/right with y in [0, mh) and lambda in
    [0, fh); the direction of v picks start or end.  eps > 0 widens every
    interval, eps < 0 narrows it."""
    ls, le, inside = 0.0, float(fw), False
    with np.errstate(divide="ignore", invalid="ignore"):
        for border, coord, lim_c, lim_l in (("top", 0.0, mw, fw), ("bottom", float(mh), mw, fw),
                                            ("left", 0.0, mh, fh), ("right", float(mw), mh, fh)):
            if border in ("top", "bottom"):
                lam = (coord - p0[1]) / v[1]
                c = p0[0] + lam * v[0]
                starts = v[1] > 0 if border == "top" else v[1] < 0
            else:
                lam = (coord - p0[0]) / v[0]
                c = p0[1] + lam * v[1]
                starts = v[0] > 0 if border == "left" else v[0] < 0
            if -eps <= c < lim_c + eps and -eps <= lam < lim_l + eps:
                if starts:
                    ls = lam
                else:
                    le = lam
                inside = True
    return ls, le, inside


def _robust(rule, *args, eps):
    """The rule's outcome when it does not change within +-eps, else None."""
    a, b = rule(*args, eps=eps), rule(*args, eps=-eps)
    if a[2] != b[2] or abs(a[0] - b[0]) > 1e-6 or abs(a[1] - b[1]) > 1e-6:
        return None
    return rule(*args)


def test_line_is_the_chord_of_the_row(oracle):
    """p0 = H^-1(0, row) and p0 + W_f v = H^-1(W_f, row): the row's samples
    run along the chord between the inverse images of its two ends."""
    rng = np.random.default_rng(11)
    mw, mh, fw, fh = 40, 36, 44, 30
    for _ in range(40):
        p = _rand_params(rng)
        h = oracle.alt_homography(p, mw, mh, fw, fh)
        inv = _inv(_centred_fp64(p, mw / 2, mh / 2))
        for row in (0, 5, 17, 29):
            L = oracle.alt_line(h, mw, mh, fw, fh, row)
            assert np.allclose(L["p0"], inv(0, row), rtol=1e-5, atol=1e-4), (p, row)
            assert np.allclose(L["p0"] + fw * L["v"].astype(np.float64), inv(fw, row), rtol=1e-5, atol=2e-4)


def test_line_crossings_follow_the_border_rules(oracle):
    rng = np.random.default_rng(12)
    seen = set()
    total = robust = 0
    for mw, mh, fw, fh in ((40, 36, 44, 30), (32, 32, 32, 32), (50, 20, 30, 60)):
        for it in range(80):
            p = _rand_params(rng, t=12.0 if it % 2 else 45.0, k=3e-3)
            h = oracle.alt_homography(p, mw, mh, fw, fh)
            for row in range(0, fh, 3):
                L = oracle.alt_line(h, mw, mh, fw, fh, row)
                p0, v = L["p0"].astype(np.float64), L["v"].astype(np.float64)
                total += 1
                want = _robust(_crossings, p0, v, mw, mh, fw, fh, eps=1e-3)
                if want is None:
                    continue   # a crossing within rounding of a rule's bound
                ls, le, inside = want
                assert inside == L["inside"], (p, row, L, want)
                assert abs(ls - L["ls"]) <= 1e-4 * max(1.0, abs(ls)) and abs(le - L["le"]) <= 1e-4 * max(1.0, abs(le))
                robust += 1
                seen.add((inside, L["ls"] > 0, L["le"] < fw))
                if L["ls"] > 0:   # the start lies on a border of the moving image
                    q = p0 + L["ls"] * v
                    assert min(abs(q[0]), abs(q[0] - mw), abs(q[1]), abs(q[1] - mh)) < 1e-3
    assert (True, True, False) in seen and (True, False, True) in seen and (False, False, False) in seen
    assert robust >= 0.95 * total


def test_norm_line_affine_parametrises_the_row(oracle):
    """Normalised lines (calcMIAlt): for an affine H, p0 + lambda v is
    H^-1(-0.5 + lambda, row) exactly, and a start/end lies on a border of
    the normalised moving image [-0.5, 0.5]^2."""
    rng = np.random.default_rng(13)
    mw = mh = fw = fh = 48
    for _ in range(40):
        p = _rand_params(rng, t=6.0, k=0.0)
        h = oracle.alt_homography(p, mw, mh, fw, fh, norm=True).astype(np.float64)
        H = np.array([[h[0], h[1], h[2]], [h[3], h[4], h[5]], [h[6], h[7], 1.0]])
        inv = _inv(H)
        for row in (-0.5, -0.2, 0.0, 0.31):
            L = oracle.alt_line(h.astype(np.float32), mw, mh, fw, fh, row, norm=True)
            p0, v = L["p0"].astype(np.float64), L["v"].astype(np.float64)
            for lam in (0.0, 0.5, 1.0):
                assert np.allclose(p0 + lam * v, inv(-0.5 + lam, row), atol=1e-5), (p, row, lam)
            for lam in (L["ls"], L["le"]):
                if L["inside"] and 0 < lam < 1:
                    q = p0 + lam * v
                    assert min(abs(abs(q[0]) - 0.5), abs(abs(q[1]) - 0.5)) < 1e-5


def _crossings_norm(p0, v, eps=0.0):
    """parametricAssignValuesNorm's border rules on float64 geometry, in
    normalised coordinates: the moving image is (-0.5, 0.5)^2, lambda runs in
    [0, 1); the top border takes lambda >= 0, the others lambda > 0."""
    ls, le, inside = 0.0, 1.0, False
    with np.errstate(divide="ignore", invalid="ignore"):
        for border in ("top", "bottom", "left", "right"):
            if border in ("top", "bottom"):
                coord = -0.5 if border == "top" else 0.5
                lam = (coord - p0[1]) / v[1]
                c = p0[0] + lam * v[0]
                starts = v[1] > 0 if border == "top" else v[1] < 0
            else:
                coord = -0.5 if border == "left" else 0.5
                lam = (coord - p0[0]) / v[0]
                c = p0[1] + lam * v[1]
                starts = v[0] > 0 if border == "left" else v[0] < 0
            lam_ok = (lam >= -eps) if border == "top" else (lam > -eps)
            if -0.5 - eps < c < 0.5 + eps and lam_ok and lam < 1 + eps:   # eps < 0 narrows
                if starts:
                    ls = lam
                else:
                    le = lam
                inside = True
    return ls, le, inside


def test_norm_line_crossings_follow_the_border_rules(oracle):
    rng = np.random.default_rng(15)
    seen = set()
    total = robust = 0
    for mw, mh, fw, fh in ((48, 48, 48, 48), (40, 30, 90, 70)):
        for _ in range(60):
            p = _rand_params(rng, t=10.0, k=0.08)
            p[0] = rng.uniform(-0.4, 0.4)
            h = oracle.alt_homography(p, mw, mh, fw, fh, norm=True)
            for row in np.linspace(-0.5, 0.49, 23).astype(np.float32):
                L = oracle.alt_line(h, mw, mh, fw, fh, row, norm=True)
                p0, v = L["p0"].astype(np.float64), L["v"].astype(np.float64)
                total += 1
                want = _robust(_crossings_norm, p0, v, eps=1e-5)
                if want is None:
                    continue
                ls, le, inside = want
                assert inside == L["inside"], (p, row, L, want)
                assert abs(ls - L["ls"]) <= 1e-5 and abs(le - L["le"]) <= 1e-5, (p, row, L, want)
                robust += 1
                seen.add((inside, L["ls"] > 0, L["le"] < 1))
    assert {(True, True, False), (True, False, True), (True, True, True), (False, False, False)} <= seen
    assert robust >= 0.95 * total


def test_norm_homography_quirks(oracle):
    """parametersToHomographyNorm: integer size ratios and centres, the moving
    centre's y taken from the width (Q11)."""
    p = np.array([0.0, 1.0, 1.0, 0.0, 3.0, -2.0, 1e-4, -2e-4], np.float32)
    h = oracle.alt_homography(p, 40, 30, 90, 70, norm=True)
    # x_ratio = 90 // 40 = 2, y_ratio = 70 // 30 = 2; cfx = 45, cfy = 35, cmx = cmy = 20
    assert h[0] == np.float32(0.5) and h[4] == np.float32(0.5) and h[1] == 0 and h[3] == 0
    assert h[2] == np.float32(np.float32(3.0 - 45 + 20) / np.float32(90))
    assert h[5] == np.float32(np.float32(-2.0 - 35 + 20) / np.float32(70))
    assert h[6] == p[6] and h[7] == p[7]


def test_mi_alt_rows_and_area(oracle):
    for fh in (2, 3, 36, 255, 256, 512):
        rows = oracle.mi_alt_rows(fh)
        step = np.float32(1) / np.float32(fh - 1)
        assert rows[0] == np.float32(-0.5) and np.all(rows < 0.5) and np.float32(rows[-1] + step) >= 0.5
        assert np.all(np.diff(rows) > 0) and fh - 1 <= len(rows) <= fh + 1
        # the loop of cudaErrorFunction_mi.cuh:162 in numpy binary32
        want, r = [], np.float32(-0.5)
        while r < 0.5:
            want.append(r)
            r = np.float32(r + step)
        assert np.array_equal(rows, np.array(want, np.float32))
    # power-of-two area: every step exact
    assert oracle.mi_alt_area(64, 32, 2048) == np.float32(1.0)
    assert oracle.mi_alt_area(64, 32, 1000) == np.float32(1000 / 2048)
    # otherwise within the accumulated rounding of n binary32 adds
    n = 3000
    a = oracle.mi_alt_area(51, 47, n)
    assert abs(a - n / (51 * 47)) <= n * 2.0 ** -24


def _bilinear64(img, y, x):
    h, w = img.shape[:2]
    if not (0 <= x < w and 0 <= y < h):
        return 0.0
    flat = img.reshape(-1).astype(np.float64)
    fx, fy = math.floor(x), math.floor(y)
    b = fy * w + fx

    def px(i):
        return flat[i] if 0 <= i < flat.size else 0.0
    ax, ay = x - fx, y - fy
    top = (1 - ax) * px(b) + ax * px(b + 1)
    bot = (1 - ax) * px(b + w) + ax * px(b + w + 1)
    return (1 - ay) * top + ay * bot


def _alt_numpy(oracle, moved, fixed, p, cost, bins=64):
    """float64 restatement of the step-1 Alt functors from the geometry.
    A row that starts or ends on a border of the moving image puts a sample
    on that border, where binary32 and binary64 may gate it differently:
    slack[k] bounds what such samples can change in sum k."""
    mh, mw = moved.shape[:2]
    fh, fw = fixed.shape[:2]
    inv = _inv(_centred_fp64(p, mw / 2, mh / 2))
    s0 = s1 = s2 = mov = 0.0
    slack = np.zeros(4)
    steps = 0
    hist = np.zeros((bins, bins), np.int64)
    for row in range(fh):
        p0 = inv(0, row)
        v = (inv(fw, row) - p0) / fw
        ls, le, inside = _crossings(p0, v, mw, mh, fw, fh)
        if not inside:
            continue
        k = 0
        lam = ls
        while lam < le and lam < fw:
            X, Y = p0 + lam * v
            m, f = _bilinear64(moved, Y, X), _bilinear64(fixed, row, lam)
            if min(abs(X), abs(X - mw), abs(Y), abs(Y - mh)) < 1e-3:
                mx = max(abs(_bilinear64(moved, min(max(Y, 0), mh - 1e-6), min(max(X, 0), mw - 1e-6))), abs(m))
                slack += [max((mx - f) ** 2, f * f) if cost == oracle.COST_SQD_ALT else mx * f, 0.0, mx * mx, 1.0]
            s0 += (m - f) ** 2 if cost == oracle.COST_SQD_ALT else m * f
            s1 += f * f
            s2 += m * m
            hist[min(int(bins * f / 256), bins - 1), min(int(bins * m / 256), bins - 1)] += 1
            lam += 1.0
            k += 1
        mov += k * math.hypot(v[0], v[1])
        steps += k
    passed = mov / (mw * mh * float(p[2])) > 0.25 and steps / (fw * fh) > 0.35
    return s0, s1, s2, steps, passed, hist, slack, mov


@pytest.mark.parametrize("cost_name", ["COST_SQD_ALT", "COST_NCC_ALT", "COST_MI_STREAM_ALT"])
def test_step_alt_costs_match_numpy_geometry(oracle, cost_name):
    cost = getattr(oracle, cost_name)
    rng = np.random.default_rng(14)
    m, f, _ = synth.make_pair(15, 40, 36, sigma=2.0, perturb=3.0, dtype=np.uint8)
    checked = 0
    for _ in range(12):
        p = _rand_params(rng, t=6.0, k=2e-3)
        v, aux = oracle.cost_point(m, f, cost, oracle.BILINEAR, p, bins=64)
        s0, s1, s2, steps, passed, hist, slack, mov = _alt_numpy(oracle, m, f, p, cost)
        assert abs(aux[0] - steps) <= 2, (p, aux[0], steps)
        assert abs(aux[6] - mov) <= 1e-4 * mov + 2 * abs(aux[0] - steps) * 1.2, (p, aux[6], mov)
        if abs(aux[0] - steps) > 0 or bool(aux[8]) != passed:
            continue   # a lambda step on a rounding edge: the sums are not comparable
        checked += 1
        if cost == oracle.COST_MI_STREAM_ALT:
            _, _, oh = oracle.cost_point(m, f, cost, oracle.BILINEAR, p, bins=64, want_hist=True)
            assert np.abs(oh.astype(np.int64) - hist).sum() <= 0.002 * steps + 2 * slack[3]
        else:
            assert abs(aux[1] - s1) <= 1e-4 * s1
            assert abs(aux[2] - s2) <= 1e-4 * s2 + slack[2], (p, aux[2], s2, slack)
            assert abs(aux[3] - s0) <= 1e-4 * abs(s0) + slack[0], (p, aux[3], s0, slack)
            if passed and slack[3] == 0:
                want = s0 / math.sqrt(s1 * s2) if cost == oracle.COST_SQD_ALT else -s0 / math.sqrt(s1 * s2)
                assert abs(v - want) <= 1e-4 * abs(want)
            elif not passed:
                assert v == np.float32(123123.0)
    assert checked >= 8


def test_step_alt_closed_forms(oracle):
    m, f, _ = synth.make_pair(16, 40, 40, sigma=2.0, perturb=3.0, dtype=np.uint8)
    ident = [0, 1, 1, 0, 0, 0, 0, 0]
    # identity: every row starts on the left border (lambda = -0) and covers the image
    v, aux = oracle.cost_point(f, f, oracle.COST_NCC_ALT, oracle.BILINEAR, ident)
    assert v == np.float32(-1.0) and aux[0] == 40 * 40 and aux[6] == 1600 and aux[8] == 1
    v, _ = oracle.cost_point(f, f, oracle.COST_SQD_ALT, oracle.BILINEAR, ident)
    assert v == 0.0
    v, _ = oracle.cost_point(f, f, oracle.COST_MI_STREAM_ALT, oracle.BILINEAR, ident, bins=64)
    assert abs(v + 1.0) < 1e-6
    # NCC scale invariance (the reference's 1/255 cancels)
    p = [0.02, 1.01, 0.99, 0.01, 1.5, -0.7, 2e-4, -1e-4]
    fm = m.astype(np.float32)
    a, _ = oracle.cost_point(fm, f.astype(np.float32), oracle.COST_NCC_ALT, oracle.BILINEAR, p)
    b, _ = oracle.cost_point(fm * np.float32(0.5), f.astype(np.float32), oracle.COST_NCC_ALT, oracle.BILINEAR, p)
    assert abs(a - b) <= 1e-6 * abs(a)
    # too little overlap -> the sentinels: 123123 (calcNCCAlt, calcSQDAlt, calcMIstreamAlt) or 0 (calcNCCstreamAlt)
    far = [0, 1, 1, 0, 30, 0, 0, 0]
    for cost, want in ((oracle.COST_NCC_ALT, 123123.0), (oracle.COST_SQD_ALT, 123123.0),
                       (oracle.COST_MI_STREAM_ALT, 123123.0), (oracle.COST_NCC_STREAM_ALT, 0.0)):
        v, aux = oracle.cost_point(m, f, cost, oracle.BILINEAR, far, bins=64)
        assert v == np.float32(want) and aux[8] == 0 and 0 < aux[0] < 0.35 * 1600
    # calcNCCstreamAlt equals calcNCCAlt wherever the overlap test passes
    a, _ = oracle.cost_point(m, f, oracle.COST_NCC_ALT, oracle.BILINEAR, p)
    b, _ = oracle.cost_point(m, f, oracle.COST_NCC_STREAM_ALT, oracle.BILINEAR, p)
    assert a == b


def test_mi_alt_matches_its_parts(oracle):
    """calcMIAlt = -MI(counts) * fix_area(steps); counts total = steps *
    channels; identical images under a small shift give a large MI; no
    qualifying row at exact identity (every border test is strict there)."""
    m, f, _ = synth.make_pair(17, 48, 40, sigma=2.0, perturb=3.0, dtype=np.uint8)
    p = [0.01, 1.0, 1.0, 0.0, 0.7, 0.4, 0.0, 0.0]
    v, aux, hist = oracle.cost_point(m, f, oracle.COST_MI_ALT, oracle.BILINEAR, p, bins=64, want_hist=True)
    assert hist.sum() == aux[4] == aux[0]
    # the step count from the rows and lines, lambda stepped in numpy binary32
    h = oracle.alt_homography(p, 48, 40, 48, 40, norm=True)
    sc = np.float32(1) / np.float32(48 - 1)
    steps = 0
    for row in oracle.mi_alt_rows(40):
        L = oracle.alt_line(h, 48, 40, 48, 40, row, norm=True)
        lam = L["ls"]
        while L["inside"] and lam < L["le"] and lam < 1:
            steps += 1
            lam = np.float32(lam + sc)
    assert steps == aux[0] > 0.5 * 48 * 40
    assert aux[7] == oracle.mi_alt_area(48, 40, int(aux[0]))
    n = hist.sum()
    pxy = hist / n
    px, py = pxy.sum(1), pxy.sum(0)
    nz = pxy > 0
    mi = float((pxy[nz] * (np.log2(pxy[nz]) - np.log2(np.outer(px, py)[nz]))).sum())
    assert abs(aux[5] - mi) <= 1e-12 * abs(mi)
    assert v == np.float32(-mi * float(aux[7]))
    vs, _ = oracle.cost_point(f, f, oracle.COST_MI_ALT, oracle.BILINEAR, p, bins=64)
    assert vs < v < 0
    v0, aux0 = oracle.cost_point(f, f, oracle.COST_MI_ALT, oracle.BILINEAR, [0, 1, 1, 0, 0, 0, 0, 0], bins=64)
    assert aux0[0] == 0 and v0 == 0.0


def _areas_numpy(oracle, moved, fixed, p, cost):
    """The reference's binary32 area loops in numpy float32 scalars, on the
    oracle's own chords (orc_alt_line): calcSQDAlt / calcNCCAlt add |v| and 1
    at every step into one sum; calcNCCstreamAlt per row, once per step;
    calcMIstreamAlt per row, once per step and channel; the stream functors
    then add the row sums in row order."""
    mh, mw = moved.shape[:2]
    fh, fw = fixed.shape[:2]
    C = 1 if fixed.ndim == 2 else fixed.shape[2]
    h = oracle.alt_homography(p, mw, mh, fw, fh)
    f32 = np.float32
    per_row = cost in (oracle.COST_NCC_STREAM_ALT, oracle.COST_MI_STREAM_ALT)
    adds = C if cost == oracle.COST_MI_STREAM_ALT else 1
    mov, fix = f32(0), f32(0)
    for row in range(fh):
        L = oracle.alt_line(h, mw, mh, fw, fh, row)
        ls, le, vx, vy, inside = f32(L["ls"]), f32(L["le"]), f32(L["v"][0]), f32(L["v"][1]), L["inside"]
        rm, rf = f32(0), f32(0)
        if inside:
            mag = np.sqrt(f32(vx * vx) + f32(vy * vy), dtype=np.float32)
            lam = ls
            while lam < le and lam < f32(fw):
                for _ in range(adds):
                    if per_row:
                        rm, rf = f32(rm + mag), f32(rf + f32(1))
                    else:
                        mov, fix = f32(mov + mag), f32(fix + f32(1))
                lam = f32(lam + f32(1))
        if per_row:
            mov, fix = f32(mov + rm), f32(fix + rf)
    return mov, fix


@pytest.mark.parametrize("cost_name", ["COST_SQD_ALT", "COST_NCC_ALT", "COST_NCC_STREAM_ALT", "COST_MI_STREAM_ALT"])
def test_step_alt_area_recurrences(oracle, cost_name):
    """Moved and fixed areas (aux[6], aux[7]) are the reference's binary32
    loops bit for bit, on one- and three-channel pairs; the binary64 moved
    area (aux[9]) is within the loops' rounding bound of them."""
    cost = getattr(oracle, cost_name)
    rng = np.random.default_rng(21)
    for channels in (1, 3):
        planes = [synth.make_pair(40 + c, 48, 40, sigma=2.0, perturb=3.0, dtype=np.uint8) for c in range(channels)]
        m = planes[0][0] if channels == 1 else np.stack([q[0] for q in planes], -1)
        f = planes[0][1] if channels == 1 else np.stack([q[1] for q in planes], -1)
        for _ in range(6):
            p = _rand_params(rng, t=6.0, k=2e-3)
            _, aux = oracle.cost_point(m, f, cost, oracle.BILINEAR, p, bins=32)
            mov, fix = _areas_numpy(oracle, m, f, p, cost)
            assert np.float32(aux[6]) == mov and np.float32(aux[7]) == fix, (cost, channels, p, aux[6], mov, aux[7], fix)
            n = aux[7] + 40
            assert abs(aux[6] - aux[9]) <= n * np.spacing(np.float32(max(aux[9], 1.0) * 1.01)) / 2 + 1e-9


def test_mi_stream_alt_areas_count_channels(oracle):
    """calcMIstreamAlt adds both areas once per step and channel, so with 3
    channels an overlap of about 27% of the fixed image passes its 0.35
    fixed-area test (3 x 0.27) while the one-channel pair fails it."""
    m1, f1, _ = synth.make_pair(44, 40, 40, sigma=2.0, perturb=2.0, dtype=np.uint8)
    m3, f3 = np.stack([m1] * 3, -1), np.stack([f1] * 3, -1)
    p = [0, 1, 1, 0, 29.0, 0, 0, 0]
    v1, aux1 = oracle.cost_point(m1, f1, oracle.COST_MI_STREAM_ALT, oracle.BILINEAR, p, bins=32)
    v3, aux3 = oracle.cost_point(m3, f3, oracle.COST_MI_STREAM_ALT, oracle.BILINEAR, p, bins=32)
    assert aux1[0] == aux3[0] == 11 * 40                     # steps: 11 columns of 40 rows
    assert aux1[7] == 440 and aux3[7] == 3 * 440             # fixed areas
    assert v1 == np.float32(123123.0) and v3 != np.float32(123123.0)
    # calcNCCstreamAlt adds once per step: the 3-channel pair fails as well
    vn, auxn = oracle.cost_point(m3, f3, oracle.COST_NCC_STREAM_ALT, oracle.BILINEAR, p)
    assert auxn[7] == 440 and vn == 0.0


def test_sequential_area_binary32_decides_the_overlap(oracle):
    """On a 256 x 256 pair at scale 1.3, x-translations between 186.7673 and
    186.7873 leave the reference's binary32 moved area (the per-step sum of
    calcNCCAlt) on one side of the 0.25 threshold and the exact area on the
    other: the oracle's cost follows the binary32 sum (the sentinel or not)."""
    m, f, _ = synth.make_pair(77, 256, 256, sigma=4.0, perturb=4.0, dtype=np.float32)
    s = np.float32(1.3)
    denom = np.float32(np.float32(256 * 256) * s)
    flips = 0
    for tx in np.linspace(186.76, 186.795, 36).astype(np.float32):
        p = np.float32([0.02, s, s, 0, tx, 0.3, 0, 0])
        v, aux = oracle.cost_point(m, f, oracle.COST_NCC_ALT, oracle.BILINEAR, p)
        d32 = float(np.float32(np.float32(aux[6]) / denom)) > 0.25
        d64 = float(np.float32(np.float32(aux[9]) / denom)) > 0.25
        assert bool(aux[8]) == d32 and (v == np.float32(123123.0)) == (not d32)
        flips += d32 != d64
    assert flips >= 10, flips
