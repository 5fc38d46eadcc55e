"""Writes the committed golden fixtures under tests/golden/.

Run from the repository root in a container that has /root/reference:
    python tests/golden/make_golden.py
Outputs (all data, no reference source):
  reference_rows.tsv  the 512 rows of the reference's true_results.out,
                      re-labelled: the file prints element `index` of arrays
                      laid out [velocity slot][cell] (npx = 256 cells fastest,
                      determ.cu:146-160), so each row is velocity slot
                      index // 256 (box node (0, 0, slot)) at cell index % 256.
                      Flags: cell0 = DC of cell 0 comes from the reference's
                      defective 5x5 solve (collision_integral.cu:464-503);
                      race = IC perturbed by the missing barrier before the
                      a_shared reload (collision_integral.cu:87-106).
  box_area.json       sphere/box areas by 1-D quadrature (see box_area_quad).
  nu_closed_form.json hard-sphere collision frequency of the reference's
                      Maxwellian, closed form, at the 8 golden slots.
"""
import json
import math
import os

import numpy as np
from scipy import integrate, special

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/true_results.out"


def reference_rows():
    rows = np.loadtxt(REF)
    out = []
    for r in rows:
        index = int(r[0])
        slot, cell = index // 256, index % 256
        race = (slot in (5, 6) and 32 <= cell <= 87) or (slot == 7 and 32 <= cell <= 119)
        out.append((slot, cell, r[4], r[5], r[6], r[7], r[8], int(cell == 0), int(race)))
    with open(os.path.join(HERE, "reference_rows.tsv"), "w") as fh:
        fh.write("slot\tcell\tF\tDC\tIC\tCI\tACCURACY\tcell0\trace\n")
        for o in out:
            fh.write("%d\t%d\t%s\t%s\t%s\t%s\t%s\t%d\t%d\n" % (o[0], o[1], repr(float(o[2])), repr(float(o[3])), repr(float(o[4])),
                                                               repr(float(o[5])), repr(float(o[6])), o[7], o[8]))


def arc_angle(rho, A1, A2, A3, A4):
    """Angle of the circle x^2+y^2 = rho^2 (first quadrant) inside
    [A1,A2] x [A3,A4]: theta with rho cos in [A1,A2], rho sin in [A3,A4]."""
    if rho <= 0:
        return 0.0
    c = lambda x: min(1.0, max(-1.0, x))
    lo1, hi1 = math.acos(c(A2 / rho)), math.acos(c(A1 / rho))
    lo2, hi2 = math.asin(c(A3 / rho)), math.asin(c(A4 / rho))
    if A1 > rho or A3 > rho:
        return 0.0
    return max(0.0, min(hi1, hi2) - max(lo1, lo2))


def box_area_quad(du, dv, dw, i_n, j_n, k_n, g):
    """Area on the unit sphere of {w: g w in box}, box edges as in
    compute_matrix.cu:121-126, by Archimedes (dA = dz dtheta on the unit
    sphere) and adaptive quadrature over the height z."""
    u1 = du * (i_n - 1) if i_n > 0 else 0.0
    u2 = du * (i_n + 1)
    v1 = dv * (j_n - 1) if j_n > 0 else 0.0
    v2 = dv * (j_n + 1)
    w1 = dw * (k_n - 1) if k_n > 0 else 0.0
    w2 = dw * (k_n + 1)
    A1, A2, A3, A4 = u1 / g, u2 / g, v1 / g, v2 / g
    z0, z1 = w1 / g, min(1.0, w2 / g)
    if z0 >= z1:
        return 0.0
    brk = []
    for xa in (A1, A2):
        for ya in (A3, A4):
            s = 1 - xa * xa - ya * ya
            if s > 0:
                brk.append(math.sqrt(s))
    for a_ in (A1, A2, A3, A4):
        s = 1 - a_ * a_
        if s > 0:
            brk.append(math.sqrt(s))
    pts = sorted(set([z0, z1] + [b for b in brk if z0 < b < z1]))
    tot = 0.0
    for lo, hi in zip(pts[:-1], pts[1:]):
        val, _ = integrate.quad(lambda z: arc_angle(math.sqrt(max(0.0, 1 - z * z)), A1, A2, A3, A4), lo, hi,
                                epsabs=1e-14, epsrel=1e-13, limit=200)
        tot += val
    return tot


def box_areas():
    rng = np.random.default_rng(20260312)
    cases = []
    # systematic small cases on a unit lattice
    for g2 in (1, 2, 3, 5, 6, 9, 11, 13, 14, 17, 21, 26, 29, 41, 50):
        g = math.sqrt(g2)
        for (i_n, j_n, k_n) in [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 0), (2, 1, 1), (1, 2, 3), (3, 0, 2),
                                (4, 3, 0), (2, 2, 2), (5, 1, 4), (6, 3, 2), (0, 4, 5), (3, 3, 3)]:
            cases.append((1.0, 1.0, 1.0, i_n, j_n, k_n, g))
    # random grid spacings / radii
    for _ in range(150):
        d = float(rng.choice([0.625, 0.48, 0.4, 0.3, 1.0]))
        gi, gj, gk = (int(x) for x in rng.integers(0, 12, 3))
        g = math.sqrt((d * gi) ** 2 + (d * gj) ** 2 + (d * gk) ** 2)
        if g == 0:
            continue
        i_n, j_n, k_n = (int(x) for x in rng.integers(0, 14, 3))
        cases.append((d, d, d, i_n, j_n, k_n, g))
    out = []
    for c in cases:
        out.append({"du": c[0], "dv": c[1], "dw": c[2], "in": c[3], "jn": c[4], "kn": c[5], "g": c[6],
                    "area": box_area_quad(*c)})
    with open(os.path.join(HERE, "box_area.json"), "w") as fh:
        json.dump(out, fh, indent=0)


def nu_closed_form():
    """nu(v) = 4 pi int |v - v1| f(v1) dv1 for f = exp(-|v-u0|^2):
    n = pi^(3/2), 2T = 1, and the hard-sphere mean relative speed
    int f_M |v-v1| dv1 = n sqrt(2T) [exp(-x^2)/sqrt(pi) + (x + 1/(2x)) erf(x)],
    x = |v - u0| / sqrt(2T)."""
    u0 = np.array([1.75, 0.25, 0.0])
    d = 10.0 / 16
    out = []
    for s in range(8):
        v = np.array([-5 + d / 2, -5 + d / 2, -5 + d * (s + 0.5)])
        x = np.linalg.norm(v - u0)
        n = math.pi ** 1.5
        mean = n * (math.exp(-x * x) / math.sqrt(math.pi) + (x + 1 / (2 * x)) * special.erf(x))
        out.append({"slot": s, "nu": 4 * math.pi * mean})
    with open(os.path.join(HERE, "nu_closed_form.json"), "w") as fh:
        json.dump(out, fh, indent=1)


if __name__ == "__main__":
    reference_rows()
    box_areas()
    nu_closed_form()
    print("fixtures written to", HERE)
