#!/usr/bin/env python3
"""Generate gas_optics_tables.inc from the reference's Fortran table literals.

The correlated-k radiation needs the reference's lookup tables bit for bit:
pressure and temperature grids, shortwave/longwave absorption tables,
Rayleigh scattering, the Planck table and the per-g-point solar fractions
(src/mod_gas_optics.f90:25-132, 143-1841, 1853-1857, 1870-3568, 3581-4549;
src/mod_solar.f90:16-22).  All of them are written as *default-real*
literals in an FP64 context, so the Fortran compiler rounds each decimal
literal to IEEE binary32 and then widens it.  This script reproduces that
rounding exactly (decimal -> nearest binary32, ties to even, computed with
exact rationals so no double rounding can occur) and writes the values as
C99 hexadecimal float literals, one flat array per table in the Fortran
column-major order of the literal lists.

Run in a checkout that has the reference at /root/reference:
    python gen_gas_optics_tables.py /root/reference/src > gas_optics_tables.inc
The output is committed; the library and the oracle include it.
"""
import re
import struct
import sys
from fractions import Fraction

TABLES = [
    # (c name, fortran name, file, expected count)
    ("GO_PRESSURES", "pressures", "mod_gas_optics.f90", 53),
    ("GO_TEMPERATURES", "temperatures", "mod_gas_optics.f90", 6 * 53),
    ("GO_SW_ABS", "shortwave_absorption_lut", "mod_gas_optics.f90", 6 * 53 * 16),
    ("GO_RAYLEIGH", "rayleigh_scattering_lut", "mod_gas_optics.f90", 16),
    ("GO_LW_ABS", "longwave_absorption_lut", "mod_gas_optics.f90", 6 * 53 * 16),
    ("GO_PLANCK_T", "planckian_temperatures", "mod_gas_optics.f90", 231),
    ("GO_PLANCK", "planckian_lut", "mod_gas_optics.f90", 231 * 16),
    ("GO_SOLAR_FRACTION", "irradiance_fraction", "mod_solar.f90", 16),
]

NUM = re.compile(r"[-+]?(?:\d+\.\d*|\.\d+|\d+)(?:[eEdD][-+]?\d+)?")


def f32_bits(x: float) -> int:
    return struct.unpack("<I", struct.pack("<f", x))[0]


def f32_from_bits(b: int) -> float:
    return struct.unpack("<f", struct.pack("<I", b))[0]


def nearest_binary32(lit: str) -> float:
    """Correctly rounded decimal -> binary32 (round to nearest, ties even)."""
    exact = Fraction(lit.replace("d", "e").replace("D", "e"))
    guess = struct.unpack("<f", struct.pack("<f", float(exact)))[0]
    best = None
    b0 = f32_bits(guess)
    for b in (b0 - 1, b0, b0 + 1):
        c = f32_from_bits(b & 0xFFFFFFFF)
        err = abs(Fraction(c) - exact)
        key = (err, b & 1)  # ties go to the even significand
        if best is None or key < best[0]:
            best = (key, c)
    return best[1]


def extract(text: str, name: str) -> list:
    # the array constructor that follows "<name>(...) ... = [" or "reshape(["
    m = re.search(r"\b" + re.escape(name) + r"\s*\([^)]*\)\s*(?:&\s*)?=\s*(?:&\s*)?(?:reshape\s*\(\s*(?:&\s*)?)?\[", text)
    if m is None:
        raise SystemExit(f"table {name} not found")
    start = m.end()
    end = text.index("]", start)
    body = text[start:end]
    body = re.sub(r"!.*", "", body)
    return NUM.findall(body.replace("&", " "))


def main(src_dir: str) -> None:
    out = ["/* Generated by gen_gas_optics_tables.py from the reference's table literals.",
           " * Each value is the binary32 rounding of the reference's default-real literal",
           " * (the reference widens it to FP64 at use).  Flat Fortran column-major order. */",
           "#pragma once"]
    for cname, fname, fil, count in TABLES:
        with open(f"{src_dir}/{fil}") as fh:
            text = fh.read()
        lits = extract(text, fname)
        if len(lits) != count:
            raise SystemExit(f"{fname}: expected {count} values, found {len(lits)}")
        vals = [nearest_binary32(s) for s in lits]
        out.append(f"#define {cname}_N {count}")
        out.append(f"static const float {cname}[{count}] = {{")
        row = []
        for v in vals:
            row.append(float(v).hex() + "f")
            if len(row) == 4:
                out.append("    " + ", ".join(row) + ",")
                row = []
        if row:
            out.append("    " + ", ".join(row) + ",")
        out.append("};")
    sys.stdout.write("\n".join(out) + "\n")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src")
