"""Generate tests/golden/laws.json: known-answer vectors for the contact laws.

Values are computed with mpmath at 50 significant digits from the closed
forms the reference documents in its comments and tests, not from the
oracle's code:
  Hertz          F = 4/3 E* sqrt(R*) d^1.5 - 2 sqrt(m* 2E* sqrt(R* d)) |beta| v_n,
                 beta = ln e / sqrt(ln^2 e + pi^2), clipped at 0  (physics.rs:86-129)
  linear         F = E*R* d - 0.5 sqrt(m* E* R*) v_n, clipped at 0 (simulation.rs:404-408)
  JKR (inverse)  pick the contact radius a; d = a^2/R* - sqrt(8 pi g a / E*),
                 F = 4E* a^3/(3R*) - sqrt(8 pi g E* a^3)          (physics.rs:140-143)
  JKR pull-off   F < -(2/3) pi g R* is clipped to 0               (physics.rs:212-216)
  sJKR           F = 4/3 E* sqrt(R*) d^1.5 - [v_n > 0] 1.5 pi g R*, zero below
                 d_c = -sqrt(3 pi^2 g^2 R* / E*^2)                 (physics.rs:250-316)
  Coulomb        F = -mu |Fn| v_t / |v_t|                          (physics.rs:385-397)
Run: python tests/golden/make_laws.py
"""
import json
import os

import mpmath as mp

mp.mp.dps = 50
PI = mp.pi


def hertz(d, E, R, m, e, vn):
    if d <= 0:
        return mp.mpf(0)
    kn = mp.mpf(4) / 3 * E * mp.sqrt(R)
    fe = kn * d ** mp.mpf(1.5)
    lne = mp.log(e)
    beta = lne / mp.sqrt(lne ** 2 + PI ** 2)
    kt = 2 * E * mp.sqrt(R * d)
    cn = 2 * mp.sqrt(m * kt) * abs(beta)
    return max(fe - cn * vn, mp.mpf(0))


def linear(d, E, R, m, vn):
    kn = E * R
    gn = mp.mpf(0.5) * mp.sqrt(m * kn)
    return max(kn * d - gn * vn, mp.mpf(0))


def jkr_from_a(a, E, R, g):
    d = a ** 2 / R - mp.sqrt(8 * PI * g * a / E)
    f = 4 * E * a ** 3 / (3 * R) - mp.sqrt(8 * PI * g * E * a ** 3)
    return d, f


def sjkr(d, E, R, g, vn):
    dc = -mp.sqrt(3 * PI ** 2 * g ** 2 * R / E ** 2)
    if d < dc:
        return mp.mpf(0)
    fh = mp.mpf(4) / 3 * E * mp.sqrt(R) * d ** mp.mpf(1.5) if d > 0 else mp.mpf(0)
    return fh - (1.5 * PI * g * R if vn > 0 else 0)


def f(x):
    return float(x)


def main():
    out = {"hertz": [], "linear": [], "jkr_inverse": [], "sjkr": [], "coulomb": []}
    # Hertz: (delta, E*, R*, m*, e, v_n)
    for d, E, R, m, e, vn in [
        (1e-5, 5.4945054945e6, 2.5e-3, 6.5e-4, 0.5, -0.1),
        (3.2977e-4, 5.4945054945e6, 0.05, 1.0, 0.5, 0.0),
        (2e-6, 2.6666666667e7, 7.5e-4, 1e-5, 0.6, 0.02),
        (5e-7, 5.4945054945e7, 2.5e-5, 1e-9, 0.9, -0.05),
        (1e-4, 1e9, 1e-2, 1e-2, 0.2, 3.0),      # strong separation: clipped to 0
    ]:
        out["hertz"].append({"in": [d, E, R, m, e, vn],
                             "F": f(hertz(mp.mpf(d), mp.mpf(E), mp.mpf(R), mp.mpf(m), mp.mpf(e), mp.mpf(vn)))})
    for d, E, R, m, vn in [(1e-5, 5.4945054945e6, 2.5e-3, 6.5e-4, -0.1), (2e-6, 1e8, 1e-3, 1e-4, 0.5),
                           (1e-6, 1e7, 1e-3, 1e-3, 10.0)]:
        out["linear"].append({"in": [d, E, R, m, vn],
                              "F": f(linear(mp.mpf(d), mp.mpf(E), mp.mpf(R), mp.mpf(m), mp.mpf(vn)))})
    # JKR inverse: (a, E*, R*, gamma) -> (delta, F); a spans the stable branch
    for E, R, g in [(5.4945054945e6, 2.5e-3, 0.1), (5.4945054945e7, 2.5e-5, 0.05), (1e8, 1e-3, 0.01)]:
        a_eq = (9 * PI * g * R ** 2 / (2 * E)) ** (mp.mpf(1) / 3)
        for frac in (0.6, 0.8, 1.0, 1.5, 2.5, 4.0):
            a = a_eq * mp.mpf(frac)
            d, F = jkr_from_a(a, mp.mpf(E), mp.mpf(R), mp.mpf(g))
            out["jkr_inverse"].append({"in": [f(d), E, R, g], "a": f(a), "F": f(F),
                                       "fpull": f(-mp.mpf(2) / 3 * PI * g * R)})
    for d, E, R, g, vn in [(1e-6, 5.4945054945e4, 2.5e-3, 0.5, -0.01), (1e-6, 5.4945054945e4, 2.5e-3, 0.5, 0.01),
                           (-1e-6, 5.4945054945e4, 2.5e-3, 0.5, 0.01), (-1e-2, 5.4945054945e4, 2.5e-3, 0.5, 0.01),
                           (3e-7, 5.4945054945e7, 2.5e-5, 0.05, 0.2)]:
        out["sjkr"].append({"in": [d, E, R, g, vn],
                            "F": f(sjkr(mp.mpf(d), mp.mpf(E), mp.mpf(R), mp.mpf(g), mp.mpf(vn)))})
    for fn, vrel, n, mu in [(2.0, (0.3, -0.1, 0.2), (0.0, 1.0, 0.0), 0.5),
                            (1.5, (1.0, 2.0, 3.0), (0.6, 0.0, 0.8), 0.3)]:
        vr = [mp.mpf(c) for c in vrel]
        nn = [mp.mpf(c) for c in n]
        dotv = sum(a * b for a, b in zip(vr, nn))
        vt = [a - dotv * b for a, b in zip(vr, nn)]
        l = mp.sqrt(sum(c * c for c in vt))
        F = [-mp.mpf(mu) * fn * c / l for c in vt]
        out["coulomb"].append({"in": [fn, list(vrel), list(n), mu], "F": [f(c) for c in F]})
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "laws.json")
    with open(path, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
