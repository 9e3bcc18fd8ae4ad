"""Write the golden vectors of tests/golden/*.json with 40-digit mpmath.

Independent of both the oracle and the library: the Debye sum, the six-term
solvent basis, spherical Bessel functions and real spherical harmonics are
evaluated by brute force from their definitions (formulas of
scattering.rs:86-110, solvent.rs:441-488/504-558, spherical.rs:54-93) using the
published coefficient tables.  The structures are the small deterministic
fixtures of the reference's own tests:
  two_carbon   scattering.rs:219-244   (also examples/data/small.pdb)
  carbon_oxygen solvent.rs:623-648
  eight        spherical.rs:731-753
  helix24      spherical.rs:682-705
  spiral96     scattering.rs:313-338
  bench16/64   benches/debye_bench.rs:7-29

Run: python tests/golden/make_golden.py   (takes about a minute)
"""
from __future__ import annotations

import json
import math
import re
from pathlib import Path

import mpmath as mp

HERE = Path(__file__).resolve().parent
mp.mp.dps = 40

# Coefficient tables come from include/crabsaxs_tables.h (C syntax parsed).
TABLES = (HERE.parents[1] / "include" / "crabsaxs_tables.h").read_text()


def _table(name):
    m = re.search(name + r"\[[^\]]*\](?:\[[^\]]*\])?\s*=\s*\{(.*?)\};", TABLES, re.S)
    body = re.sub(r"/\*.*?\*/", "", m.group(1), flags=re.S)
    rows = re.findall(r"\{([^{}]*)\}", body)
    if rows:
        return [[mp.mpf(v) for v in r.split(",") if v.strip()] for r in rows]
    return [mp.mpf(v) for v in body.split(",") if v.strip()]


CM_A, CM_B, CM_C = _table("CSX_CM_A"), _table("CSX_CM_B"), _table("CSX_CM_C")
VOL = _table("CSX_EXCLUDED_VOLUME")
C, N, O, S, P = 1, 2, 3, 4, 5


def ff(el, q):
    s2 = (mp.mpf(q) / (4 * mp.pi)) ** 2
    return sum(a * mp.exp(-b * s2) for a, b in zip(CM_A[el], CM_B[el])) + CM_C[el]


def fraser(q, v):
    return v * mp.exp(-(v ** (mp.mpf(2) / 3)) * mp.mpf(q) ** 2 / (4 * mp.pi))


def amplitudes(el, q, t=mp.mpf(3), rho=mp.mpf("0.334")):
    v = VOL[el]
    radius = mp.cbrt(3 * v / (4 * mp.pi))
    shell = 4 * mp.pi * (radius + t / 2) ** 2 * t
    return ff(el, q), rho * fraser(q, v), rho * mp.mpf("0.1") * fraser(q, shell)


def sinc(x):
    return mp.mpf(1) if abs(x) < mp.mpf("1e-8") else mp.sin(x) / x


def structures():
    out = {}
    out["two_carbon"] = ([(0, 0, 0), (2, 0, 0)], [C, C])
    out["carbon_oxygen"] = ([(-1, 0, 0), (1, 0, 0)], [C, O])
    out["eight"] = ([(i * 1.4 - 4.9, math.sin(i * 0.9) * 2.0, math.cos(i * 0.7)) for i in range(8)],
                    [[C, N, O, S][i % 4] for i in range(8)])
    out["helix24"] = ([(7.0 * math.cos(i * 0.71), 7.0 * math.sin(i * 0.71), i * 0.45 - 5.0)
                       for i in range(24)], [[C, N, O, S][i % 4] for i in range(24)])
    sp = []
    for i in range(96):
        ang = i * 2.399963229728653
        rad = 4.0 + (i % 13)
        sp.append((rad * math.cos(ang), rad * math.sin(ang), (i - 47.5) * 0.31))
    out["spiral96"] = (sp, [[C, N, O, S, P][i % 5] for i in range(96)])
    for n in (16, 64):
        out[f"bench{n}"] = ([(12.0 * math.cos(i * 0.61803398875), 12.0 * math.sin(i * 0.61803398875),
                              i * 0.08) for i in range(n)], [[C, N, O, S][i % 4] for i in range(n)])
    return out


def debye(xyz, el, q):
    # Exact FP64 inputs are converted exactly (mpf of a float is exact).
    pts = [[mp.mpf(c) for c in p] for p in xyz]
    n = len(pts)
    r = {}
    for i in range(n):
        for j in range(i + 1, n):
            r[i, j] = mp.sqrt(sum((pts[i][a] - pts[j][a]) ** 2 for a in range(3)))
    res_i, res_b = [], []
    for qq in q:
        qm = mp.mpf(qq)
        amp = [amplitudes(e, qm) for e in el]
        t = [mp.mpf(0)] * 6
        for i in range(n):
            v, e, h = amp[i]
            t[0] += v * v
            t[1] += e * e
            t[2] += h * h
            t[3] += v * e
            t[4] += v * h
            t[5] += e * h
        for (i, j), rij in r.items():
            s = sinc(qm * rij)
            vi, ei, hi = amp[i]
            vj, ej, hj = amp[j]
            t[0] += 2 * vi * vj * s
            t[1] += 2 * ei * ej * s
            t[2] += 2 * hi * hj * s
            t[3] += (vi * ej + vj * ei) * s
            t[4] += (vi * hj + vj * hi) * s
            t[5] += (ei * hj + ej * hi) * s
        res_i.append(float(max(t[0], 0)))
        res_b.append([float(x) for x in t])
    return res_i, [list(col) for col in zip(*res_b)]


def bessel_table():
    cases = [(15, 0.5), (15, 20.0), (15, 35.0), (15, 45.0), (25, 60.0), (40, 100.0), (64, 3.0),
             (40, 0.01), (8, 7.5)]
    rows = []
    for L, x in cases:
        xm = mp.mpf(x)
        vals = [float(mp.sqrt(mp.pi / (2 * xm)) * mp.besselj(l + mp.mpf(1) / 2, xm))
                for l in range(L + 1)]
        rows.append({"l_max": L, "x": x, "j": vals})
    return rows


def harmonics_table():
    rows = []
    for L, ct, phi in [(8, 0.3, 1.1), (16, -0.77, -2.5), (24, 0.999, 0.2), (12, 0.0, 3.0)]:
        theta = mp.acos(mp.mpf(ct))
        vals = [mp.mpf(0)] * ((L + 1) ** 2)
        for l in range(L + 1):
            for m in range(0, l + 1):
                y = mp.spherharm(l, m, theta, mp.mpf(phi))
                if m == 0:
                    vals[l * l + l] = mp.re(y)
                else:
                    vals[l * l + l + m] = mp.sqrt(2) * mp.re(y)
                    vals[l * l + l - m] = mp.sqrt(2) * mp.im(y)
        rows.append({"l_max": L, "cos_theta": ct, "phi": phi, "y": [float(v) for v in vals]})
    return rows


def main():
    q_low = [0.0] + [k * 0.005 for k in range(1, 101)]
    q_high = [k / 50.0 for k in range(51)]
    data = {"q_low": q_low, "q_high": q_high, "structures": {}}
    for name, (xyz, el) in structures().items():
        entry = {"xyz": [list(p) for p in xyz], "element": el}
        for tag, q in (("low", q_low), ("high", q_high)):
            inten, basis = debye(xyz, el, q)
            entry[f"debye_{tag}"] = inten
            entry[f"basis_{tag}"] = basis
        data["structures"][name] = entry
        print(name, "done", flush=True)
    (HERE / "debye_golden.json").write_text(json.dumps(data))
    (HERE / "special_golden.json").write_text(json.dumps(
        {"bessel": bessel_table(), "harmonics": harmonics_table()}))


if __name__ == "__main__":
    main()
