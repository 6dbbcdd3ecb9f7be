#!/usr/bin/env python3
"""Extracts the numeric tables of the reference's cell models into
tests/golden/ionic_tables.json.

Reads, as text, voomEP/src/IonicModel/{SetUpIonicConstants.h, Tusscher.h,
LR.cc} under /root/reference and records only numbers: the ten Tusscher and
Purkinje constant arrays, the ten Tusscher initial state (the active block,
not the commented-out one) and the Luo-Rudy initial table and constants.
Nothing of the reference is executed.  Run from the repository root:

    python tests/golden/make_ionic_tables.py
"""
import json
import os
import re
import sys

REF = "/root/reference/voomEP/src/IonicModel"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ionic_tables.json")

NUM = r"([-+]?[0-9]*\.?[0-9]+(?:[eE][-+]?[0-9]+)?)"


def strip_comments(text):
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return re.sub(r"//[^\n]*", "", text)


def function_body(text, name):
    i = text.index(name)
    j = text.index("{", i)
    depth, k = 0, j
    while True:
        if text[k] == "{":
            depth += 1
        elif text[k] == "}":
            depth -= 1
            if depth == 0:
                return text[j:k + 1]
        k += 1


def indexed_literals(body, var):
    vals = {}
    for m in re.finditer(var + r"\[\s*(\d+)\s*\]\s*=\s*" + NUM + r"\s*;", body):
        vals[int(m.group(1))] = float(m.group(2))
    n = max(vals) + 1 if vals else 0
    return [vals.get(i) for i in range(n)]


def main():
    const_h = strip_comments(open(os.path.join(REF, "SetUpIonicConstants.h")).read())
    tt = indexed_literals(function_body(const_h, "void SetUpTusscherParameters"), "Constants")
    pk = indexed_literals(function_body(const_h, "void SetUpPurkinjeParameters"), "Constants")
    tth = strip_comments(open(os.path.join(REF, "Tusscher.h")).read())
    tt_init = indexed_literals(function_body(tth, "Tusscher(Real *constants)"), "_State")
    lr = strip_comments(open(os.path.join(REF, "LR.cc")).read())
    lr_body = function_body(lr, "void LuoRudy::initialize")
    lr_init = indexed_literals(lr_body, "_in")
    lr_consts = {}
    for name in ["Gna", "Gsi", "c_K0", "c_Nai", "c_Na0", "c_Ki", "RT_F"]:
        m = re.search(r"\b" + name + r"\s*=\s*" + NUM + r"\s*;", lr_body)
        lr_consts[name] = float(m.group(1))
    xi = re.search(r"_Xi\s*=\s*" + NUM, tth)
    out = {
        "source": "voomEP/src/IonicModel (SetUpIonicConstants.h, Tusscher.h, LR.cc), numbers only",
        "tusscher_constants": tt,
        "tusscher_initial_state": tt_init,
        "tusscher_xi": float(xi.group(1)),
        "purkinje_constants_literal": pk,
        "luorudy_initial_state": lr_init,
        "luorudy_constants": lr_consts,
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {len(tt)} tT constants, {len(tt_init)} tT states, "
          f"{len(lr_init)} LR states", file=sys.stderr)


if __name__ == "__main__":
    main()
