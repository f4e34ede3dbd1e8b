"""Generate spectral_data.h: the measured data tables the render core needs.

The tables are physical data, not code: the CIE 1931 2-degree colour
matching functions (1 nm, 360-830 nm), CIE D65 (5 nm, 300-830 nm), the 13
measured metal n,k grids (5 nm, 360-830 nm) and the four alloy data grids
(Au-Ag excess permittivity, Querry Cu / Cu85Zn15 / Cu70Zn30).  They are read
as numbers from the reference's data modules
(crates/argent-spectra/src/data/{cie,metals,alloys}.rs) and written as flat
C arrays, so the values are the same digits the reference renders with.

Run once, from the repository root:
    python argent_amd/csrc/tables/gen_tables.py /root/reference
The output is committed; nothing at build or run time reads the reference.
"""
import os
import re
import sys

NUM = r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eE][-+]?\d+)?"


def block(text, name):
    m = re.search(r"const\s+" + name + r"\s*:[^=]*=\s*\[(.*?)\];", text, re.S)
    if not m:
        raise SystemExit(f"table {name} not found")
    body = re.sub(r"//[^\n]*", "", m.group(1))
    return [float(x) for x in re.findall(NUM, body)]


def fmt(vals, per=8):
    out = []
    for i in range(0, len(vals), per):
        out.append("    " + ", ".join(repr(float(v)) + "f" for v in vals[i:i + per]) + ",")
    return "\n".join(out)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    d = os.path.join(ref, "crates/argent-spectra/src/data")
    cie = open(os.path.join(d, "cie.rs")).read()
    met = open(os.path.join(d, "metals.rs")).read()
    alo = open(os.path.join(d, "alloys.rs")).read()
    cmf = block(cie, "CMF_1931_2DEG")
    d65 = block(cie, "D65_300_830_5NM")
    assert len(cmf) == 471 * 3 and len(d65) == 107
    metals = ["AG", "AU", "CU", "PT", "PD", "RH", "NI", "TI", "AL", "CR", "FE", "ZN", "SN"]
    mt = {m: block(met, m) for m in metals}
    for m in metals:
        assert len(mt[m]) == 190, m
    alloys = ["AU_AG_EXCESS_50", "CU_QUERRY", "CU85ZN15_QUERRY", "CU70ZN30_QUERRY"]
    at = {a: block(alo, a) for a in alloys}
    for a in alloys:
        assert len(at[a]) == 190, a
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "spectral_data.h"), "w") as f:
        f.write("// Generated by gen_tables.py from measured data; do not edit.\n")
        f.write("// CIE 1931 2-degree observer (CIE 015:2018), CIE D65, metal n,k grids\n")
        f.write("// (Johnson & Christy, Rakic, Weaver, Querry; resampled to 5 nm), alloy data.\n")
        f.write("#pragma once\n\nnamespace ag_data {\n\n")
        f.write("// x, y, z at 360 + i nm, i = 0..470\n")
        f.write(f"static const float CMF_1931_2DEG[471 * 3] = {{\n{fmt(cmf, 9)}\n}};\n\n")
        f.write("// relative SPD at 300 + 5 i nm, i = 0..106\n")
        f.write(f"static const float D65_300_830_5NM[107] = {{\n{fmt(d65)}\n}};\n\n")
        f.write("// (n, k) at 360 + 5 i nm, i = 0..94; order Ag Au Cu Pt Pd Rh Ni Ti Al Cr Fe Zn Sn\n")
        f.write("static const float METAL_NK[13][190] = {\n")
        for m in metals:
            f.write(f"  {{ // {m}\n{fmt(mt[m], 10)}\n  }},\n")
        f.write("};\n\n")
        for a in alloys:
            f.write(f"static const float {a}[190] = {{\n{fmt(at[a], 10)}\n}};\n\n")
        f.write("}  // namespace ag_data\n")
    print("wrote spectral_data.h")


if __name__ == "__main__":
    main()
