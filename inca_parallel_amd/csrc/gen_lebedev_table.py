"""Generate lebedev_orbits.h: the Lebedev-Laikov spherical rules as Oh-orbit
generators (code, a, b, v), one list per rule size.

Source of the numbers: scipy.integrate._lebedev (BSD, R. Parrish's
translation of D. N. Laikov's published tables; V. I. Lebedev and
D. N. Laikov, Doklady Mathematics 59 (1999) 477).  The decimal literals are
copied verbatim from that module so the doubles are those of the published
table.  Run:  python gen_lebedev_table.py > lebedev_orbits.h
"""
import re
import sys

import scipy.integrate._lebedev as leb

src = open(leb.__file__).read()
body = src[src.index("def get_lebedev_sphere"):src.index("def get_lebedev_recurrence_points")]
cases = re.split(r"\n\s*case (\d+):", body)
rules = []
for k in range(1, len(cases), 2):
    n = int(cases[k])
    txt = cases[k + 1]
    a = b = "0.0"
    orbits = []
    for line in txt.splitlines():
        line = line.strip()
        m = re.match(r"([abv]) = ([-+0-9.Ee]+)$", line)
        if m:
            if m.group(1) == "a":
                a = m.group(2)
            elif m.group(1) == "b":
                b = m.group(2)
            else:
                v = m.group(2)
            continue
        m = re.match(r"leb_tmp, start = get_lebedev_recurrence_points\((\d), start", line)
        if m:
            orbits.append((int(m.group(1)), a, b, v))
    size = {1: 6, 2: 12, 3: 8, 4: 24, 5: 24, 6: 48}
    assert sum(size[o[0]] for o in orbits) == n, n
    rules.append((n, orbits))

out = sys.stdout
out.write("/* Generated by gen_lebedev_table.py -- do not edit.\n"
          " * Lebedev-Laikov rules as Oh orbits {code, a, b, v}; weights v sum to 1.\n"
          " * code 1: (1,0,0)  2: (0,a,a)  3: (a,a,a)  4: (a,a,b) b=sqrt(1-2a^2)\n"
          " * 5: (a,b,0) b=sqrt(1-a^2)  6: (a,b,c) c=sqrt(1-a^2-b^2) */\n")
out.write("#pragma once\n")
out.write("struct LebOrbit { int code; double a, b, v; };\n")
out.write("static const int LEB_NRULES = %d;\n" % len(rules))
out.write("static const int LEB_SIZE[%d] = {%s};\n" % (len(rules), ", ".join(str(r[0]) for r in rules)))
offs = [0]
for r in rules:
    offs.append(offs[-1] + len(r[1]))
out.write("static const int LEB_FIRST[%d] = {%s};\n" % (len(offs), ", ".join(map(str, offs))))
out.write("static const LebOrbit LEB_ORBITS[%d] = {\n" % offs[-1])
for n, orbits in rules:
    out.write("    /* %d points */\n" % n)
    for c, a, b, v in orbits:
        out.write("    {%d, %s, %s, %s},\n" % (c, a, b, v))
out.write("};\n")
