"""Text writers of the reference's output bundle, value by value (oracle).

TEST INFRASTRUCTURE ONLY.  Restates core/io.py:44-118 of the reference: each
value as ``{:.<precision>g}`` followed by a tab, a newline after each x row,
an empty line after each component block of the field dump.
"""
from __future__ import annotations


def field_dat_text(field_flat, xlen, ylen, nfields, precision=32):
    fmt = f"{{:.{precision}g}}"
    out = []
    for a in range(nfields):
        for x in range(xlen):
            for y in range(ylen):
                out.append(fmt.format(float(field_flat[y + x * ylen + a * xlen * ylen])))
                out.append("\t")
            out.append("\n")
        out.append("\n")
    return "".join(out)


def density_dat_text(density_flat, a, xlen, ylen, precision=32):
    fmt = f"{{:.{precision}g}}"
    out = []
    for x in range(xlen):
        for y in range(ylen):
            out.append(fmt.format(float(density_flat[y + x * ylen + a * xlen * ylen])))
            out.append("\t")
        out.append("\n")
    return "".join(out)


def values_dat_text(values, precision=32):
    fmt = f"{{:.{precision}g}}"
    return "".join(fmt.format(float(v)) + "\t" for v in values)
