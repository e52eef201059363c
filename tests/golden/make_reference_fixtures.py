"""Writes reference_constants.json: data read, as text, from two reference
sources so that the tests can pin the shared constants where the reference
is not mounted.

- scatterer_param_LUT.hpp:117-1560: the Kirkland parameter table (103 x 12
  values).  Stored: the SHA-256 of the table as little-endian float64 in the
  order [Z-1][a0 a1 a2 b0 b1 b2 c0 c1 c2 d0 d1 d2], and f_e(q^2) =
  sum a_i/(q^2 + b_i) + sum c_i exp(-d_i q^2) (projected_atomic_potential.cpp:209-225)
  of every Z at q^2 = 0, 0.37 and 4.0.
- scherzer.cpp:61-79 and :131-168: the leading constants of the ADF-STEM
  Scherzer helpers (defocus and alpha_max uncorrected; Cs3, defocus and
  alpha_max corrected to Cs5).

Run from the repository root: python tests/golden/make_reference_fixtures.py
"""
import hashlib
import json
import math
import os
import re
import struct

REF = "/root/reference/src"
HERE = os.path.dirname(os.path.abspath(__file__))


def kirkland_table():
    txt = open(os.path.join(REF, "scatterer_param_LUT.hpp")).read()
    pat = re.compile(r"\(get_paramPtr\(\s*(\d+)\s*\)\)->([abcd])\[([0-2])\]\s*=\s*([-+0-9.eE]+)\s*;")
    tab = {}
    for m in pat.finditer(txt):
        tab[(int(m.group(1)), m.group(2), int(m.group(3)))] = float(m.group(4))
    assert len(tab) == 103 * 12
    return [[tab[(z, k, i)] for k in "abcd" for i in range(3)] for z in range(1, 104)]


def fe(row, q2):
    a, b, c, d = row[0:3], row[3:6], row[6:9], row[9:12]
    return sum(ai / (q2 + bi) for ai, bi in zip(a, b)) + sum(
        ci * math.exp(-di * q2) for ci, di in zip(c, d))


def scherzer_constants():
    txt = open(os.path.join(REF, "scherzer.cpp")).read()
    out = {}
    for name in ("scherzer_defocus_adfstem_uncorrected", "scherzer_alphamax_adfstem_uncorrected",
                 "scherzer_Cs3_adfstem_correctedtoCs5", "scherzer_defocus_adfstem_correctedtoCs5",
                 "scherzer_alphamax_adfstem_correctedtoCs5"):
        body = txt[txt.index("TEM_NS::" + name + "("):]
        body = body[:body.index("\n}")]
        ret = re.search(r"return\s+\(?\s*([-+0-9.]+)(?:\s*/\s*([0-9.]+)\s*\))?\s*\*", body)
        val = float(ret.group(1)) / (float(ret.group(2)) if ret.group(2) else 1.0)
        out[name] = val
    return out


def main():
    tab = kirkland_table()
    blob = b"".join(struct.pack("<12d", *row) for row in tab)
    fixture = {
        "kirkland_sha256_f64le": hashlib.sha256(blob).hexdigest(),
        "kirkland_layout": "[Z-1][a0 a1 a2 b0 b1 b2 c0 c1 c2 d0 d1 d2], float64 little-endian",
        "fe_q2": [0.0, 0.37, 4.0],
        "fe": [[fe(row, q2) for q2 in (0.0, 0.37, 4.0)] for row in tab],
        "scherzer": scherzer_constants(),
        "sources": ["scatterer_param_LUT.hpp:117-1560", "projected_atomic_potential.cpp:209-225",
                    "scherzer.cpp:61-79", "scherzer.cpp:131-168"],
    }
    with open(os.path.join(HERE, "reference_constants.json"), "w") as f:
        json.dump(fixture, f, indent=1)
        f.write("\n")
    print(fixture["scherzer"], fixture["kirkland_sha256_f64le"])


if __name__ == "__main__":
    main()
