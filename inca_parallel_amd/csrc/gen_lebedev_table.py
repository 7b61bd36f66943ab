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

# The numbers come from SciPy's source, which carries this notice.
SCIPY_NOTICE = """\
The orbit parameters below are taken from SciPy (scipy/integrate/_lebedev.py),
distributed under this licence:

Copyright (c) 2001-2002 Enthought, Inc. 2003-2024, SciPy Developers.
All rights reserved.

Redistribution and use in source and binary forms, with or without
modification, are permitted provided that the following conditions
are met:

1. Redistributions of source code must retain the above copyright
   notice, this list of conditions and the following disclaimer.

2. Redistributions in binary form must reproduce the above
   copyright notice, this list of conditions and the following
   disclaimer in the documentation and/or other materials provided
   with the distribution.

3. Neither the name of the copyright holder nor the names of its
   contributors may be used to endorse or promote products derived
   from this software without specific prior written permission.

THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS
"AS IS" AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT
LIMITED TO, THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR
A PARTICULAR PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT
OWNER OR CONTRIBUTORS BE LIABLE FOR ANY DIRECT, INDIRECT, INCIDENTAL,
SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT
LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES; LOSS OF USE,
DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND ON ANY
THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT
(INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE
OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
"""

out = sys.stdout
out.write("/* Generated by gen_lebedev_table.py -- do not edit.\n"
          " * Lebedev-Laikov rules as Oh orbits {code, a, b, v}; weights v sum to 1.\n"
          " * code 1: (1,0,0)  2: (0,a,a)  3: (a,a,a)  4: (a,a,b) b=sqrt(1-2a^2)\n"
          " * 5: (a,b,0) b=sqrt(1-a^2)  6: (a,b,c) c=sqrt(1-a^2-b^2)\n *\n")
for line in SCIPY_NOTICE.splitlines():
    out.write((" * " + line).rstrip() + "\n")
out.write(" */\n")
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
