#!/usr/bin/env python3
"""Generate ``clblm_data.inc``: the numeric tables that define CLBLM's line-shape path.

The tables are data, not code.  They are read from the reference's Fortran
DATA/PARAMETER statements and written as C arrays of doubles.  Every literal is
parsed as an IEEE double, which is what the reference's ``-fdefault-real-8``
build does with a default-kind real literal.

Tables written (reference file:line):
  * AVRAT, CGAUSS, CF1, CF2, CF3, CER (102 zeta nodes each)
        src/clblm_src/clblm_ODLAY_CommonSub.f90:43-185
  * TIPS temperature grid Tdat (119 nodes)     src/clblm_src/clblm_ODLAY_TIPS.f90:13-27
  * TIPS Q(T) rows for molecules 1-7, every isotopologue the QT_* routine holds
        QT_H2O :424, QT_CO2 :615, QT_O3 :938, QT_N2O :1443, QT_CO :1609,
        QT_CH4 :1801, QT_O2 :1941
  * isotopologue masses for molecules 1-7      src/common/clblm_ConstParam.f90:260-285
  * MaxISOTPL_smass for molecules 1-7          src/common/clblm_ConstParam.f90:246-251

and ``clblm_contnm.inc``: the continuum tables of CONTNM's bands other than water
vapour (CO2; the N2 pure rotation, fundamental and first overtone; the O2 fundamental,
1.27 um, A-band, visible and far-UV continua; ozone's Chappuis and Hartley-Huggins
bands), listed at CONTNM_TABLES.

Run in a container that holds the reference sources:
    python clblm_amd/csrc/gen_tables.py /root/reference
The generated file is committed; the build never needs the reference.
"""
import json
import os
import re
import sys

NMOL = 7
QT_NAMES = ["QT_H2O", "QT_CO2", "QT_O3", "QT_N2O", "QT_CO", "QT_CH4", "QT_O2"]
NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[EeDd][-+]?\d+)?"


def _strip_comments(text):
    out = []
    for line in text.splitlines():
        i = line.find("!")
        out.append(line if i < 0 else line[:i])
    return "\n".join(out)


def _numbers(s):
    return [float(x.replace("D", "E").replace("d", "e")) for x in re.findall(NUM, s)]


def parse_param_array(src, name, n):
    m = re.search(r"::\s*" + name + r"\s*\(\s*1\s*:\s*%d\s*\)\s*=\s*\[(.*?)\]" % n, src, re.S)
    if not m:
        raise RuntimeError("array %s not found" % name)
    body = m.group(1).replace("&", " ")
    vals = _numbers(body)
    if len(vals) != n:
        raise RuntimeError("%s: expected %d values, got %d" % (name, n, len(vals)))
    return vals


def parse_tdat(src):
    m = re.search(r"Tdat\(NT\)\s*=\s*\[(.*?)\]", src, re.S)
    vals = _numbers(m.group(1).replace("&", " "))
    assert len(vals) == 119, len(vals)
    return vals


def parse_qt(src, name):
    start = re.search(r"Subroutine\s+" + name + r"\b", src).start()
    end = src.index("END SUBROUTINE", start)
    body = src[start:end]
    niso = int(re.search(r"QofT\(\s*(\d+)\s*,\s*119\)", body).group(1))
    rows = {}
    for m in re.finditer(r"data\s*\(QofT\(\s*(\d+)\s*,\s*J\)\s*,\s*J=1,119\)\s*/(.*?)/", body, re.S):
        iso = int(m.group(1))
        vals = _numbers(m.group(2).replace("&", " "))
        assert len(vals) == 119, (name, iso, len(vals))
        rows[iso] = vals
    assert sorted(rows) == list(range(1, niso + 1)), (name, sorted(rows))
    return [rows[i] for i in range(1, niso + 1)]


def parse_masses(src):
    masses = {}
    for m in re.finditer(r"data\s*\(isotpl_mass\(\s*(\d+)\s*,\s*i\)\s*,\s*i=1,(\d+)\)\s*/(.*?)/", src.replace("&", " "), re.S):
        mol, n = int(m.group(1)), int(m.group(2))
        vals = _numbers(m.group(3).replace("&", " "))
        assert len(vals) == n
        masses[mol] = vals
    return [masses[m] for m in range(1, NMOL + 1)]


def parse_maxiso(src):
    m = re.search(r"MaxISOTPL_smass\(MXMOL\)\s*=\s*\[(.*?)\]", src, re.S)
    vals = [int(x) for x in re.findall(r"\d+", m.group(1).replace("&", " "))]
    return vals[:NMOL]


def c_array(name, vals):
    body = ",\n  ".join(", ".join(repr(float(v)) for v in vals[i:i + 4]) for i in range(0, len(vals), 4))
    return "static const double %s[%d] = {\n  %s\n};\n" % (name, len(vals), body)


def main(ref_root, out_inc, out_json):
    common = _strip_comments(open(os.path.join(ref_root, "src/clblm_src/clblm_ODLAY_CommonSub.f90")).read())
    tips = _strip_comments(open(os.path.join(ref_root, "src/clblm_src/clblm_ODLAY_TIPS.f90")).read())
    const = _strip_comments(open(os.path.join(ref_root, "src/common/clblm_ConstParam.f90")).read())
    zeta = {n: parse_param_array(common, n, 102) for n in ["AVRAT", "CGAUSS", "CF1", "CF2", "CF3", "CER"]}
    tdat = parse_tdat(tips)
    qt = [parse_qt(tips, n) for n in QT_NAMES]
    masses = parse_masses(const)
    maxiso = parse_maxiso(const)
    for m in range(NMOL):
        assert len(qt[m]) >= maxiso[m], (m + 1, len(qt[m]), maxiso[m])
        assert len(masses[m]) == maxiso[m]

    lines = ["/* Generated by gen_tables.py from the reference's DATA statements; do not edit. */",
             "/* Values are the IEEE doubles nearest the reference's decimal literals. */",
             "#define CLBLM_NZETA 102", "#define CLBLM_NT 119", "#define CLBLM_NMOL_DATA %d" % NMOL, ""]
    for n, v in zeta.items():
        lines.append(c_array("CLBLM_" + n, v))
    lines.append(c_array("CLBLM_TIPS_TDAT", tdat))
    lines.append("static const int CLBLM_MAXISO[%d] = {%s};\n" % (NMOL, ", ".join(map(str, maxiso))))
    flat_mass, off = [], []
    for m in range(NMOL):
        off.append(len(flat_mass))
        flat_mass += masses[m]
    lines.append(c_array("CLBLM_ISO_MASS", flat_mass))
    lines.append("static const int CLBLM_ISO_OFFSET[%d] = {%s};\n" % (NMOL, ", ".join(map(str, off))))
    flat_q, qoff = [], []
    for m in range(NMOL):
        qoff.append(len(flat_q) // 119)
        for iso in range(maxiso[m]):
            flat_q += qt[m][iso]
    lines.append(c_array("CLBLM_TIPS_Q", flat_q))
    lines.append("/* row index of (mol m, iso 1) in CLBLM_TIPS_Q; rows are 119 long */")
    lines.append("static const int CLBLM_TIPS_ROW[%d] = {%s};\n" % (NMOL, ", ".join(map(str, qoff))))
    open(out_inc, "w").write("\n".join(lines))
    json.dump({"zeta": zeta, "tdat": tdat, "maxiso": maxiso, "mass": masses,
               "tips_q": [[qt[m][i] for i in range(maxiso[m])] for m in range(NMOL)]},
              open(out_json, "w"))


# ---- continuum tables (src/clblm_src/clblm_ODLAY_Continuum.f90) ------------------------
# The CONTNM bands of the thermal infrared, each tabulated on its own grid
# (V1S, V2S, DVS, NPTS):
#   XFACCO2 (CONTNM :145-236), FRNCO2's S, TDEP_BANDHEAD and T_EFF (:1333-2530),
#   xn2_r's C_296, sf_296, C_220, sf_220 (:2533-2681), n2_ver_1's xn2_272, xn2_228,
#   a_h2o (:2701-2953), o2_ver_1's xo2, xo2t (:7187-7311).
CONTNM_TABLES = [
    # (C name, subroutine, DATA name, count)
    ("CN_XFACCO2", "CONTNM", "XFACCO2", 500),
    ("CN_CO2_GRID", "FRNCO2", "V1S", 4),
    ("CN_CO2_S", "FRNCO2", "S", 5003),
    ("CN_CO2_TDEP", "FRNCO2", "tdep_bandhead", 25),
    ("CN_CO2_TEFF", "FRNCO2", "t_eff", 1),
    ("CN_N2R_GRID", "xn2_r", "V1S", 4),
    ("CN_N2R_C296", "xn2_r", "C_296", 73),
    ("CN_N2R_SF296", "xn2_r", "sf_296", 73),
    ("CN_N2R_C220", "xn2_r", "C_220", 73),
    ("CN_N2R_SF220", "xn2_r", "sf_220", 73),
    ("CN_N2F_GRID", "n2_ver_1", "V1S", 4),
    ("CN_N2F_X272", "n2_ver_1", "xn2_272", 230),
    ("CN_N2F_X228", "n2_ver_1", "xn2_228", 230),
    ("CN_N2F_AH2O", "n2_ver_1", "a_h2o", 230),
    ("CN_O2F_GRID", "o2_ver_1", "V1S", 4),
    ("CN_O2F_X", "o2_ver_1", "xo2", 103),
    ("CN_O2F_XT", "o2_ver_1", "xo2t", 103),
    ("CN_N2O_GRID", "n2_overtone1", "V1S", 4),
    ("CN_N2O_X", "n2_overtone1", "xn2", 201),
    ("CN_O2I1_GRID", "O2INF1", "V1S", 4),
    ("CN_O2I1_X", "O2INF1", "xo2inf1", 483),
    ("CN_O2I3_GRID", "O2INF3", "V1S", 4),
    ("CN_O2I3_X", "O2INF3", "xo2inf3", 261),
    ("CN_O2V_GRID", "O2_vis", "V1s", 4),
    ("CN_O2V_S", "O2_vis", "S", 1488),
    ("CN_O2U_GRID", "O2FUV", "V1S", 4),
    ("CN_O2U_S", "O2FUV", "S", 1512),
    ("CN_O3C_GRID", "XO3CHP", "V1S", 4),
    ("CN_O3C_X", "XO3CHP", "X", 3150),
    ("CN_O3C_Y", "XO3CHP", "Y", 3150),
    ("CN_O3C_Z", "XO3CHP", "Z", 3150),
    ("CN_O3T_GRID", "O3HHT0", "V1S", 4),
    ("CN_O3T0_S", "O3HHT0", "S", 2687),
    ("CN_O3T1_S", "O3HHT1", "S", 2687),
    ("CN_O3T2_S", "O3HHT2", "S", 2687),
    ("CN_O3UV_GRID", "O3HHUV", "V1S", 4),
    ("CN_O3UV_S", "O3HHUV", "S", 133),
]

# Arrays declared with a lower bound other than 1.
LOWER_BOUND = {"tdep_bandhead": 1196}

# The routines whose table grid must equal another's: O3HHT1 and O3HHT2 fill CT1(J)
# and CT2(J) on their own windows, which CONTNM indexes with O3HHT0's J.
SAME_GRID = [("O3HHT1", "O3HHT0"), ("O3HHT2", "O3HHT0")]


def subroutine_text(src, name):
    m = re.search(r"^\s*SUBROUTINE\s+" + name + r"\b.*?END\s+SUBROUTINE", src, re.S | re.I | re.M)
    if not m:
        raise RuntimeError("subroutine %s not found" % name)
    return _strip_comments(m.group(0))


def data_values(text, name):
    """The values of `DATA name/.../`, `DATA name(1:n)/.../` or `DATA (name(i),i=a,b)/.../`."""
    m = re.search(r"\bDATA\s+\(?\s*" + name + r"\b[^/]*/(.*?)/", text, re.S | re.I)
    if not m:
        raise RuntimeError("DATA %s not found" % name)
    return _numbers(m.group(1).replace("&", " "))


def _data_items(body):
    """The constants of a DATA value list, with Fortran's repeat form `r*value` expanded."""
    out = []
    for tok in body.replace("&", " ").split(","):
        tok = tok.strip()
        if not tok:
            continue
        m = re.fullmatch(r"(\d+)\s*\*\s*(%s)" % NUM, tok)
        if m:
            out.extend([float(m.group(2).replace("D", "E").replace("d", "e"))] * int(m.group(1)))
        elif re.fullmatch(NUM, tok):
            out.append(float(tok.replace("D", "E").replace("d", "e")))
        else:
            raise RuntimeError("DATA item %r not understood" % tok)
    return out


def data_array(text, name, n, lo=1):
    """Array `name(lo:lo+n-1)` from every DATA statement of `text` that fills part of it:
    `DATA name/.../`, `DATA name(a:b)/.../` and `DATA (name(I),I=a,b)/.../`.  Every element
    must be set exactly once."""
    vals = [None] * n
    pat = (r"\bDATA\s+(?:%s\s*(?:\(\s*(\d+)\s*:\s*(\d+)\s*\))?|\(\s*%s\s*\(\s*(\w+)\s*\)\s*,\s*\3\s*=\s*"
           r"(\d+)\s*,\s*(\d+)\s*\))\s*/(.*?)/") % (name, name)
    found = False
    for m in re.finditer(pat, text, re.S | re.I):
        found = True
        a, b = (m.group(1), m.group(2)) if m.group(1) else (m.group(4), m.group(5))
        a, b = (int(a), int(b)) if a else (lo, lo + n - 1)
        items = _data_items(m.group(6))
        if len(items) != b - a + 1 or a < lo or b > lo + n - 1:
            raise RuntimeError("DATA %s(%d:%d): %d values" % (name, a, b, len(items)))
        for k, v in enumerate(items):
            if vals[a - lo + k] is not None:
                raise RuntimeError("%s(%d) set twice" % (name, a + k))
            vals[a - lo + k] = v
    if not found or any(v is None for v in vals):
        raise RuntimeError("DATA %s: not every element set" % name)
    return vals


def contnm_tables(ref_root):
    src = open(os.path.join(ref_root, "src/clblm_src/clblm_ODLAY_Continuum.f90")).read()
    out = {}
    for cname, sub, dname, n in CONTNM_TABLES:
        text = subroutine_text(src, sub)
        if cname.endswith("_GRID") or n == 1:
            vals = data_values(text, dname)[:n]
        else:
            vals = data_array(text, dname, n, LOWER_BOUND.get(dname, 1))
        if len(vals) != n:
            raise RuntimeError("%s %s: expected %d values, got %d" % (sub, dname, n, len(vals)))
        out[cname] = vals
    for sub, like in SAME_GRID:
        a = data_values(subroutine_text(src, sub), "V1S")[:4]
        b = data_values(subroutine_text(src, like), "V1S")[:4]
        if a != b:
            raise RuntimeError("%s grid %s differs from %s grid %s" % (sub, a, like, b))
    return out


def main_contnm(ref_root, out_inc):
    t = contnm_tables(ref_root)
    lines = ["/* Generated by gen_tables.py from the DATA statements of the reference's",
             "   clblm_ODLAY_Continuum.f90; do not edit.  Values are the IEEE doubles nearest the",
             "   reference's decimal literals.  *_GRID: V1S, V2S, DVS, NPTS of the table. */", ""]
    for cname, _, _, _ in CONTNM_TABLES:
        lines.append(c_array(cname, t[cname]))
    open(out_inc, "w").write("\n".join(lines))


if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    repo = os.path.dirname(os.path.dirname(here))
    main(ref, os.path.join(here, "clblm_data.inc"), os.path.join(repo, "tests", "golden", "tables.json"))
    main_contnm(ref, os.path.join(here, "clblm_contnm.inc"))
