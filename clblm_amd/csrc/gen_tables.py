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

and ``clblm_contnm.inc``: the continuum tables of CONTNM's thermal-infrared bands
(CO2, N2 pure rotation and fundamental, O2 fundamental), listed at CONTNM_TABLES.

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
]


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


def contnm_tables(ref_root):
    src = open(os.path.join(ref_root, "src/clblm_src/clblm_ODLAY_Continuum.f90")).read()
    out = {}
    for cname, sub, dname, n in CONTNM_TABLES:
        vals = data_values(subroutine_text(src, sub), dname)[:n]
        if len(vals) != n:
            raise RuntimeError("%s %s: expected %d values, got %d" % (sub, dname, n, len(vals)))
        out[cname] = vals
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
