"""Writes tests/golden/pair_integrals.json and tests/golden/self_integrals.json.

Independent high-precision values (mpmath, nested tanh-sinh quadrature at 25
digits) of the two integrals the HEM impedances are built from:

* the segment-pair integral ``I = L_r L_s * int_0^1 int_0^1 exp(-gamma r)/r dt du``
  with ``r`` the distance between points on the two segment axes, which is what
  ``integral(..., INTG_DOUBLE, ...)`` returns (src/electrode.c:176-197,272-277);
* the closed-form self term ``int_0^L int_0^L dx dy / sqrt((x-y)^2 + a^2)`` that
  ``self_integral`` evaluates (src/electrode.c:238-245).

Nothing here imports or runs the reference; the geometries are the pair types
that occur in the grids, rods and images of the workloads (touching collinear
and perpendicular segments, gaps, parallel neighbours, images, skew pairs) and
the propagation constants are those of the workloads' soils plus one large
synthetic value.

Run:  python tests/golden/make_golden.py [--self-only]
"""
import json
import os

import mpmath as mp

mp.mp.dps = 25
HERE = os.path.dirname(os.path.abspath(__file__))

TWO_PI = mp.mpf("6.283185307179586476925286766559")
MU0 = mp.mpf("1.256637061435917e-6")
EPS0 = mp.mpf("8.854187817620e-12")


def gamma_of(f, sigma, epsr):
    s = mp.mpc(0, TWO_PI * f)
    kappa = sigma + s * epsr * EPS0
    return mp.sqrt(s * MU0 * kappa)


GAMMAS = {
    "zero": mp.mpc(0),
    "c1_100Hz": gamma_of(100, mp.mpf("0.01"), 10),
    "c1_1MHz": gamma_of(1e6, mp.mpf("0.01"), 10),
    "c1_10MHz": gamma_of(1e7, mp.mpf("0.01"), 10),
    "c3_5MHz": gamma_of(5e6, mp.mpf("0.002"), 20),
    "synthetic": mp.mpc(2, 3),
}

Z = -0.5
GEOMS = {
    # name: (sender start, sender end, receiver start, receiver end)
    "collinear_touch": ([0, 0, Z], [0.5, 0, Z], [0.5, 0, Z], [1.0, 0, Z]),
    "collinear_touch_reversed": ([0.5, 0, Z], [0, 0, Z], [0.5, 0, Z], [1.0, 0, Z]),
    "perpendicular_touch": ([0, 0, Z], [0.5, 0, Z], [0.5, 0, Z], [0.5, 0.5, Z]),
    "perpendicular_touch_unequal": ([0, 0, Z], [0.5, 0, Z], [0.5, 0.2, Z], [0.5, 0, Z]),
    "vertical_touch": ([0, 0, -0.5], [0, 0, -1.0], [0, 0, -1.0], [0, 0, -1.5]),
    "collinear_gap": ([0, 0, Z], [0.5, 0, Z], [1.0, 0, Z], [1.5, 0, Z]),
    "parallel_2m": ([0, 0, Z], [0.5, 0, Z], [0, 2.0, Z], [0.5, 2.0, Z]),
    "parallel_shifted": ([0, 0, Z], [0.5, 0, Z], [0.25, 0.5, Z], [0.75, 0.5, Z]),
    "perpendicular_gap": ([0, 0, Z], [0.5, 0, Z], [0.5, 0.5, Z], [0.5, 1.0, Z]),
    "image_self": ([0, 0, Z], [0.5, 0, Z], [0, 0, -Z], [0.5, 0, -Z]),
    "image_perpendicular": ([0, 0, Z], [0.5, 0, Z], [0.5, 0, -Z], [0.5, 0.5, -Z]),
    "image_vertical": ([0, 0, -0.5], [0, 0, -1.0], [0, 0, 0.5], [0, 0, 1.0]),
    "far_20m": ([0, 0, Z], [0.5, 0, Z], [14.0, 14.5, Z], [14.0, 15.0, Z]),
    "skew": ([0, 0, -0.6], [0.5, 0.1, -0.7], [0.9, -0.3, -0.4], [1.1, 0.3, -0.9]),
}


def seg_len(a, b):
    return mp.sqrt(sum((mp.mpf(b[i]) - mp.mpf(a[i])) ** 2 for i in range(3)))


def pair_integral(P0, P1, Q0, Q1, g):
    P0 = [mp.mpf(x) for x in P0]
    P1 = [mp.mpf(x) for x in P1]
    Q0 = [mp.mpf(x) for x in Q0]
    Q1 = [mp.mpf(x) for x in Q1]

    def f(t, u):
        d = [Q0[i] + t * (Q1[i] - Q0[i]) - P0[i] - u * (P1[i] - P0[i]) for i in range(3)]
        r = mp.sqrt(d[0] ** 2 + d[1] ** 2 + d[2] ** 2)
        return mp.exp(-g * r) / r

    J = mp.quad(f, [0, 1], [0, 1])
    return J * seg_len(P0, P1) * seg_len(Q0, Q1)


def self_integral_mp(L, a):
    """int_0^L int_0^L g(x - y) dx dy = 2 int_0^L (L - u) g(u) du for even g;
    the 1-D form keeps the near-singular ridge x = y at an end point, with
    break points at the scale of the radius."""
    L = mp.mpf(L)
    a = mp.mpf(a)
    pts = [0] + [a * 10 ** k for k in range(0, 6) if a * 10 ** k < L] + [L]
    return 2 * mp.quad(lambda u: (L - u) / mp.sqrt(u * u + a * a), pts)


def main(only_self=False):
    pairs = []
    if only_self:
        GEOMS_RUN = {}
    else:
        GEOMS_RUN = GEOMS
    for gname, geo in GEOMS_RUN.items():
        for kname, g in GAMMAS.items():
            v = pair_integral(*geo, g)
            pairs.append({"geometry": gname, "sender": [geo[0], geo[1]],
                          "receiver": [geo[2], geo[3]], "gamma_name": kname,
                          "gamma": [float(g.real), float(g.imag)],
                          "integral": [float(v.real), float(v.imag)]})
            print(gname, kname, mp.nstr(v, 18), flush=True)
    if not only_self:
        with open(os.path.join(HERE, "pair_integrals.json"), "w") as fh:
            json.dump({"source": "mpmath nested tanh-sinh, dps=25", "pairs": pairs}, fh,
                      indent=1)
    selfs = []
    for L, a in [(0.5, 7e-3), (0.5, 5e-3), (0.5, 6e-3), (1.0, 0.012), (0.1, 0.01),
                 (2.0, 0.005)]:
        v = self_integral_mp(L, a)
        selfs.append({"length": L, "radius": a, "value": float(v)})
        print("self", L, a, mp.nstr(v, 18), flush=True)
    with open(os.path.join(HERE, "self_integrals.json"), "w") as fh:
        json.dump({"source": "mpmath tanh-sinh of the exact 1-D reduction, dps=25",
                   "cases": selfs}, fh,
                  indent=1)


if __name__ == "__main__":
    import sys
    main(only_self="--self-only" in sys.argv)
