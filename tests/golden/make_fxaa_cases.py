"""Hand-worked FXAA cases (writes fxaa_cases.json).

The FXAA pass of the reference (kernels/ssaa.cuh:141-182; D13 and D14 of
the defect register) is, for pixel p with integer texel centres:

  luma(c) = (0.299 r + 0.587 g + 0.114 b) / 255
  NW, NE, SW, SE, M = luma of the four diagonal neighbours and of p
  dir   = (-((NW + NE) - (SW + SE)), (NW + SW) - (NE + SE))
  reduce = max((NW + NE + SW + SE) * 1/4 * 1/8, 1/128)      REDUCE_MUL, REDUCE_MIN
  dir   = clamp(dir / (min(|dir_x|, |dir_y|) + reduce), -8, 8)   SPAN_MAX
  A = (I(p - dir/6) + I(p + dir/6)) / 2
  B = A / 2 + (I(p - dir/2) + I(p + dir/2)) / 4
  out = A if luma(B) is outside [min, max] of the five lumas, else B

with I the bilinear lookup and every tap clamped to the image (D13).  Each
case below is a small image whose output at one pixel has a closed form,
worked out by hand from these lines (the derivation is in each function's
docstring).  The closed forms are evaluated here in double precision; the
oracle computes in float32, so the test allows 1e-5 relative.

Each case also lists what the output would be with one constant changed
(no /255 in luma, REDUCE_MIN = 1/64, REDUCE_MUL = 1/16, SPAN_MAX = 4, taps
not clamped at the border): the test checks that every such value is far
from the expected one, so a change of that constant fails the test.
"""
import json
import os

W_LUMA = (0.299, 0.587, 0.114)
HERE = os.path.dirname(os.path.abspath(__file__))


def luma(c, scale=255.0):
    return (W_LUMA[0] * c[0] + W_LUMA[1] * c[1] + W_LUMA[2] * c[2]) / scale


def diagonal_corner(colour, v, scale=255.0, rmin=1 / 128, rmul=1 / 32):
    """5x5 image, colour v*w on the pixels with x + y <= 3, zero elsewhere;
    pixel (2, 2).  Of the five lumas only NW = (1, 1) is lit, with luma l:
    dir = (-l, l), reduce = max(l * rmul, rmin), so dir = (-t, t) with
    t = l / (l + reduce) <= 1 (no span clamp).  The tap p + (a, -a), 0 < a < 1,
    weighs the lit texel (2, 1) by (1 - a) a and no other lit texel; the tap
    p + (-a, a) weighs (1, 2) by a (1 - a).  With a = t/6 for A and a = t/2
    for B: A = v w (t/6)(1 - t/6), B = A/2 + v w (t/2)(1 - t/2) / 2, and
    luma(B) lies in [0, l], so out = B."""
    l = luma([v * c for c in colour], scale)
    t = l / (l + max(l * rmul, rmin))
    F = 0.5 * (t / 6) * (1 - t / 6) + 0.5 * (t / 2) * (1 - t / 2)
    return [v * c * F for c in colour]


def diagonal_corner_image(colour, v):
    return [[[v * c for c in colour] if x + y <= 3 else [0.0, 0.0, 0.0] for x in range(5)] for y in range(5)]


def span_clamp(v, s, span=8.0, scale=255.0, rmin=1 / 128, rmul=1 / 32):
    """9x3 grey image: row 0 = v, row 1 = s x^2 at column x, row 2 = 0;
    pixel (4, 1).  NW = NE = l = v/255, SW = SE = 0: dir = (-2l, 0),
    min(|dir|) = 0 and reduce = max(2l * rmul, rmin).  For v >= 32,
    2l/32 >= 1/128, so dir_x = -2l / (l/16) = -32, clamped to -span.  All taps
    stay on row 1 (dir_y = 0, fy = 0), where the lookup interpolates
    linearly between the columns: with span 8 the taps are at x = 4 +- 4/3
    (A) and 4 +- 4 (B):
      A = (I(16/3) + I(8/3)) / 2 = (86/3 + 22/3) s / 2 = 18 s,
      B = A/2 + (I(8) + I(0)) / 4 = 9 s + 16 s = 25 s.
    v = 64, s = 1: luma(B) = 25/255 is inside [0, 64/255]: out = B = 25.
    v = 32, s = 2: M = 32 is the brightest luma, luma(B) = 50/255 is above
    it: out = A = 36."""
    l = v / scale
    red = max(2 * l * rmul, rmin)
    d = max(-span, min(span, -2 * l / red))

    def I(x):  # row 1, linear between columns, clamped to [0, 8]
        x = max(0.0, min(8.0, x))
        x0 = min(int(x // 1), 7)
        f = x - x0
        return s * ((1 - f) * x0 * x0 + f * (x0 + 1) * (x0 + 1))

    A = 0.5 * (I(4 - d / 6) + I(4 + d / 6))
    B = 0.5 * A + 0.25 * (I(4 - d / 2) + I(4 + d / 2))
    lums = [v / scale, v / scale, 0.0, 0.0, s * 16 / scale]
    out = A if (B / scale < min(lums) or B / scale > max(lums)) else B
    return [out] * 3


def span_clamp_image(v, s):
    return [[[float(v)] * 3 for x in range(9)], [[float(s * x * x)] * 3 for x in range(9)],
            [[0.0] * 3 for x in range(9)]]


def corner_clamp(v, scale=255.0, rmin=1 / 128, rmul=1 / 32, clamp=True):
    """3x3 grey image, only pixel (0, 0) = v; pixel (0, 0).  With taps
    clamped to the image, NW (at (-1, -1)) reads (0, 0) and M = (0, 0): the
    lumas are NW = M = l, NE = SW = SE = 0, dir = (-l, l), dir -> (-t, t) with
    t = l / (l + max(l * rmul, rmin)).  The tap (a, -a) reads rows -1 and 0,
    both row 0 after clamping: I = v (1 - a); the tap (-a, a) reads columns
    -1 and 0, both column 0: I = v (1 - a).  A = v (1 - t/6),
    B = A/2 + v (1 - t/2)/2 = v (1 - t/3), luma(B) in [0, l]: out = B.
    Without the clamp (zero outside the image) NW = 0, all four diagonal
    lumas vanish, dir = 0 and out = M = v."""
    if not clamp:
        return [v] * 3
    l = v / scale
    t = l / (l + max(l * rmul, rmin))
    return [v * (1 - t / 3)] * 3


def corner_clamp_image(v):
    return [[[float(v)] * 3 if (x, y) == (0, 0) else [0.0] * 3 for x in range(3)] for y in range(3)]


def main():
    cases = []
    colours = {"white": (1.0, 1.0, 1.0), "red": (1.0, 0.0, 0.0), "green": (0.0, 1.0, 0.0), "blue": (0.0, 0.0, 1.0)}
    for name, c in colours.items():
        altered = {"luma without /255": diagonal_corner(c, 1.0, scale=1.0),
                   "REDUCE_MIN 1/64": diagonal_corner(c, 1.0, rmin=1 / 64)}
        if name != "white":  # both weight sets sum to 1: white does not tell them apart
            altered["Rec.709 luma weights"] = diagonal_corner_709(c, 1.0)
        cases.append(dict(name=f"diagonal_corner_{name}_v1", image=diagonal_corner_image(c, 1.0), pixel=[2, 2],
                          expected=diagonal_corner(c, 1.0), altered=altered))
    # colours above 1 (the render's SSAA clamps to [0, 1]) are the only inputs on
    # which REDUCE_MUL acts: with luma / 255 <= 1/255 the floor 1/128 always wins
    cases.append(dict(
        name="diagonal_corner_white_v200", image=diagonal_corner_image((1.0, 1.0, 1.0), 200.0), pixel=[2, 2],
        expected=diagonal_corner((1.0, 1.0, 1.0), 200.0),
        altered={"REDUCE_MUL 1/16": diagonal_corner((1.0, 1.0, 1.0), 200.0, rmul=1 / 16),
                 "REDUCE_MIN only (no REDUCE_MUL)": diagonal_corner((1.0, 1.0, 1.0), 200.0, rmul=0.0)}))
    for v, s, sel in ((64, 1, "B"), (32, 2, "A")):
        cases.append(dict(
            name=f"span_clamp_v{v}_s{s}_selects_{sel}", image=span_clamp_image(v, s), pixel=[4, 1],
            expected=span_clamp(v, s),
            altered={"SPAN_MAX 4": span_clamp(v, s, span=4.0),
                     "SPAN_MAX 16": span_clamp(v, s, span=16.0)}))
    cases.append(dict(
        name="corner_clamp_v1", image=corner_clamp_image(1.0), pixel=[0, 0], expected=corner_clamp(1.0),
        altered={"taps not clamped (zero outside)": corner_clamp(1.0, clamp=False),
                 "luma without /255": corner_clamp(1.0, scale=1.0),
                 "REDUCE_MIN 1/64": corner_clamp(1.0, rmin=1 / 64)}))
    with open(os.path.join(HERE, "fxaa_cases.json"), "w") as f:
        json.dump({"source": "tests/golden/make_fxaa_cases.py", "cases": cases}, f, indent=1)
    print(f"{len(cases)} cases")


def diagonal_corner_709(colour, v):
    """diagonal_corner with the Rec.709 luma weights instead of the reference's."""
    global W_LUMA
    saved = W_LUMA
    W_LUMA = (0.2126, 0.7152, 0.0722)
    try:
        return diagonal_corner(colour, v)
    finally:
        W_LUMA = saved


if __name__ == "__main__":
    main()
