"""Generate qs_table.h: the FP64-built saturation vapour pressure table of the
mixed-precision kernel.

e_sat(T) = E00 exp((d ln(T / TICE) + l (1 / TICE - 1 / T)) / RVGAS), the
closed form of the reference's wqs1 / iqs1 (microphysics_funcs.py:53-145),
with (d, l) = (DC_VAP, LV0) over water and (D2ICE, LI2) over ice.  In
u = 1 / T its logarithm is -l u / RVGAS plus a slowly varying term, so equal
intervals of u give a uniform relative accuracy.  Each of N intervals holds
the degree-DEG polynomial in s in [-1, 1] (s = -1 at the interval's start)
that interpolates e_sat at the DEG + 1 Chebyshev nodes, evaluated in FP64 with
numpy's exp and log and written as exact hexadecimal literals.  Relative
error against the closed form: < 2e-12 over water (100-1000 K), < 3e-14 over
ice (T_SAT_MIN - TICE); tests/test_qs_table.py checks it.

The constants are read from mp_common.h, so the table follows them.  Run by
pyshield_amd/build.py before compiling; `python gen_qs_table.py --check`
exits 1 when the committed header differs from what it would write.
"""

import os
import re
import sys

import numpy as np
from numpy.polynomial import chebyshev

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "qs_table.h")
N, DEG = 128, 7
T_LO_WATER, T_HI_WATER = 100.0, 1000.0


def constants():
    src = open(os.path.join(HERE, "mp_common.h")).read()
    env = {}
    for name, expr in re.findall(r"constexpr double (\w+)\s*=\s*([^;]+);", src):
        try:
            env[name] = float(eval(expr, {"__builtins__": {}}, env))
        except (NameError, SyntaxError, TypeError):
            pass
    return env


def esat(u, d, l, k):
    return k["E00"] * np.exp((d * np.log(1.0 / (u * k["TICE"])) + l * (1.0 / k["TICE"] - u)) / k["RVGAS"])


def table(u0, u1, d, l, k):
    edges = np.linspace(u0, u1, N + 1)
    nodes = np.cos((np.arange(DEG + 1) + 0.5) * np.pi / (DEG + 1))
    rows = []
    for a, b in zip(edges[:-1], edges[1:]):
        u = 0.5 * (a + b) + 0.5 * (b - a) * nodes
        c = chebyshev.chebfit(nodes, esat(u, d, l, k), DEG)
        rows.append(chebyshev.cheb2poly(c))
    return np.array(rows)


def render():
    k = constants()
    specs = [(1.0 / T_HI_WATER, 1.0 / T_LO_WATER, k["DC_VAP"], k["LV0"]),
             (1.0 / k["TICE"], 1.0 / k["T_SAT_MIN"], k["D2ICE"], k["LI2"])]
    tabs = [table(u0, u1, d, l, k) for u0, u1, d, l in specs]
    lines = [
        "// qs_table.h -- written by gen_qs_table.py from the constants of mp_common.h; do not edit.",
        "//",
        "// e_sat(T) = E00 exp((d ln(T / TICE) + l (1 / TICE - 1 / T)) / RVGAS) as a polynomial of",
        f"// degree {DEG} in s in [-1, 1] on each of {N} equal intervals of u = 1 / T, built in FP64 by",
        "// interpolation at Chebyshev nodes.  Table 0: over water (d = DC_VAP, l = LV0),",
        f"// T from {T_LO_WATER:g} K to {T_HI_WATER:g} K; table 1: over ice (d = D2ICE, l = LI2), T from T_SAT_MIN",
        "// to TICE.  Interval i of table t starts at u = QS_TAB_U0[t] + i / QS_TAB_RN[t].",
        "#pragma once",
        "",
        "namespace shield {",
        "",
        f"constexpr int QS_TAB_N = {N}, QS_TAB_DEG = {DEG};",
        f"constexpr double QS_TAB_U0[2] = {{{float(specs[0][0]).hex()}, {float(specs[1][0]).hex()}}};",
        f"constexpr double QS_TAB_U1[2] = {{{float(specs[0][1]).hex()}, {float(specs[1][1]).hex()}}};",
        "// intervals per unit of u",
        f"constexpr double QS_TAB_RN[2] = {{{float(N / (specs[0][1] - specs[0][0])).hex()}, "
        f"{float(N / (specs[1][1] - specs[1][0])).hex()}}};",
        f"constexpr double QS_TAB[2][{N}][{DEG + 1}] = {{",
    ]
    for t, tab in enumerate(tabs):
        lines.append("    {")
        for row in tab:
            lines.append("        {" + ", ".join(float(c).hex() for c in row) + "},")
        lines.append("    }," if t == 0 else "    },")
    lines += ["};", "", "}  // namespace shield", ""]
    return "\n".join(lines)


def main():
    text = render()
    if "--check" in sys.argv:
        ok = os.path.exists(HEADER) and open(HEADER).read() == text
        print("qs_table.h is current" if ok else "qs_table.h differs from the generator's output")
        sys.exit(0 if ok else 1)
    if not os.path.exists(HEADER) or open(HEADER).read() != text:
        with open(HEADER, "w") as f:
            f.write(text)
    print(HEADER)


if __name__ == "__main__":
    main()
