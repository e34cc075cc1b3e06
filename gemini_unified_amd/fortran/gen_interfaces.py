"""Writes geminiFKc.f90, the iso_c_binding interface module of the C ABI, from
include/gemini_fk.h: one bind(C) interface per entry point and one named
constant per integer #define.  Pointer arguments are type(c_ptr), value, so one
interface takes host arrays (c_loc) and, with GM_FLAG_DEVICE_PTRS, device
buffers.  Run after a change of the header; tests/test_fortran.py checks that
the committed module matches it."""
from __future__ import annotations

import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(HERE, "..", "..", "include", "gemini_fk.h")
OUT = os.path.join(HERE, "geminiFKc.f90")

SCALARS = {"int": "integer(c_int)", "unsigned": "integer(c_int)", "long long": "integer(c_long_long)",
           "double": "real(c_double)", "float": "real(c_float)"}


def parse(header_text):
    consts = re.findall(r"^#define (GM_\w+) (-?\d+)\b", header_text, flags=re.M)
    protos = []
    for name, params in re.findall(r"^int (gm_\w+)\(([^;]*?)\);", header_text, flags=re.M | re.S):
        args = []
        p = " ".join(params.split())
        if p not in ("", "void"):
            for a in p.split(","):
                a = a.strip().replace("const ", "")
                m = re.match(r"(.+?)\s*(\**)\s*(\w+)$", a)
                base, stars, arg = m.group(1).strip(), m.group(2), m.group(3)
                if stars == "**":
                    decl = "type(c_ptr)"                                  # gm_ctx** out
                elif stars == "*" and base == "char":
                    decl = "character(kind=c_char) :: {}(*)"
                elif stars == "*":
                    decl = "type(c_ptr), value"
                elif base == "gm_ctx":
                    raise ValueError(a)
                else:
                    decl = SCALARS[base] + ", value"
                if base == "gm_ctx" and stars == "*":
                    decl = "type(c_ptr), value"
                args.append((arg, decl))
        protos.append((name, args))
    return consts, protos


def wrap(line, indent):
    """Free-form lines of at most 110 characters, continued with &."""
    out, cur = [], line
    while len(cur) > 110:
        cut = cur.rfind(", ", 0, 108) + 1
        out.append(cur[:cut] + " &")
        cur = " " * (indent + 4) + cur[cut:].lstrip()
    out.append(cur)
    return out


def render(consts, protos):
    L = ["! iso_c_binding interface of libgemini_fk.so (include/gemini_fk.h).",
         "! Written by gen_interfaces.py from the header; do not edit by hand.",
         "module geminiFKc",
         "   use iso_c_binding",
         "   implicit none",
         "   public"]
    for name, val in consts:
        L.append(f"   integer(c_int), parameter :: {name} = {val}")
    L.append("   interface")
    for name, args in protos:
        names = ", ".join(a for a, _ in args)
        L += wrap(f"      integer(c_int) function {name}({names}) bind(C, name='{name}')", 6)
        L.append("         import")
        for a, decl in args:
            if "{}" in decl:
                L.append("         " + decl.format(a))
            else:
                L.append(f"         {decl} :: {a}")
        L.append(f"      end function {name}")
    L.append("   end interface")
    L.append("end module geminiFKc")
    return "\n".join(L) + "\n"


def generate():
    with open(HEADER) as fh:
        return render(*parse(fh.read()))


if __name__ == "__main__":
    text = generate()
    if "--check" in sys.argv:
        with open(OUT) as fh:
            sys.exit(0 if fh.read() == text else 1)
    with open(OUT, "w") as fh:
        fh.write(text)
    print(OUT)
