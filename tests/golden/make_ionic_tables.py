#!/usr/bin/env python3
"""Extracts the numeric tables of the reference's cell models into
tests/golden/ionic_tables.json.

Reads, as text, voomEP/src/IonicModel/{SetUpIonicConstants.h, Tusscher.h,
LR.cc, Mahajan.cc, Purkinje.h, Purkinje.cc} under /root/reference and records
only numbers: the ten Tusscher and Purkinje constant arrays, the Purkinje
initial state, Xi, cell area and the centre and slope of each logistic
steady state of its gates, the ten Tusscher initial state (the
active block, not the commented-out one), the Luo-Rudy initial table and
constants, and the Mahajan constants, per-tissue-type G_tof / G_Ks and
per-type initial states.
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


def product(expr):
    """Value of a literal product such as 1.9*0.32 or 0.11*.85."""
    v = 1.0
    for f in expr.split("*"):
        v *= float(f)
    return v


def mahajan_tables(const_h, mj_cc):
    body = function_body(const_h, "void SetUpMahajanParameters")
    consts = indexed_literals(body, "Constants")
    per_type = {}
    for m in re.finditer(r"case\s+(\d+)\s*:\s*\{([^}]*)\}", body):
        vals = {}
        for a in re.finditer(r"Constants\[\s*(\d+)\s*\]\s*=\s*([0-9.eE*+-]+)\s*;", m.group(2)):
            vals[int(a.group(1))] = product(a.group(2))
        per_type[int(m.group(1))] = [vals[7], vals[10]]
    ctor = function_body(mj_cc, "Mahajan::Mahajan(Real *Constants, int pos)")
    init = {}
    for m in re.finditer(r"case\s*\(\s*(\d+)\s*\)\s*:\s*\{([^}]*)\}", ctor):
        init[int(m.group(1))] = indexed_literals(m.group(2), "_State")
    d = re.search(r"default\s*:\s*\{([^}]*)\}", ctor)
    xi = re.search(r"_Xi\s*=\s*" + NUM + r"\s*/\s*" + NUM, mj_cc)
    return {
        "mahajan_constants_literal": consts,
        "mahajan_gtof_gks_by_type": [per_type[t] for t in range(9)],
        "mahajan_initial_states": [init[t] for t in range(9)],
        "mahajan_initial_state_default": indexed_literals(d.group(1), "_State"),
        "mahajan_xi": float(xi.group(1)) / float(xi.group(2)),
    }


def arith(expr, C):
    """Value of an arithmetic expression of numbers, Constants[k], + - * /
    and parentheses (a small parser: the reference text is never executed)."""
    toks = re.findall(r"Constants\[\s*\d+\s*\]|[0-9]*\.?[0-9]+(?:[eE][-+]?\d+)?|[-+*/()]", expr)
    pos = [0]

    def peek():
        return toks[pos[0]] if pos[0] < len(toks) else None

    def take():
        pos[0] += 1
        return toks[pos[0] - 1]

    def atom():
        t = take()
        if t == "(":
            v = add()
            assert take() == ")"
            return v
        if t == "-":
            return -atom()
        if t.startswith("Constants"):
            return C[int(re.search(r"\d+", t).group(0))]
        return float(t)

    def mul():
        v = atom()
        while peek() in ("*", "/"):
            v = v * atom() if take() == "*" else v / atom()
        return v

    def add():
        v = mul()
        while peek() in ("+", "-"):
            v = v + mul() if take() == "+" else v - mul()
        return v

    v = add()
    assert pos[0] == len(toks), expr
    return v


def run_constants(body, pos):
    """The 35 constants a SetUp*Parameters(constants, pos) body leaves:
    assignments (= and *=) in order, the switch on pos taking its case or
    default block."""
    sw = re.search(r"switch\s*\(\s*pos\s*\)\s*\{", body)
    i = sw.end()
    depth, k = 1, i
    while depth:
        depth += {"{": 1, "}": -1}.get(body[k], 0)
        k += 1
    pre, inside, post = body[:sw.start()], body[i:k - 1], body[k:]
    cases = {int(m.group(1)): m.group(2)
             for m in re.finditer(r"case\s+(\d+)\s*:\s*\{([^}]*)\}", inside)}
    default = re.search(r"default\s*:\s*\{([^}]*)\}", inside).group(1)
    C = {}
    for part in (pre, cases.get(pos, default), post):
        for m in re.finditer(r"Constants\[\s*(\d+)\s*\]\s*(\*?=)\s*([^;]+);", part):
            idx, op, expr = int(m.group(1)), m.group(2), m.group(3)
            v = arith(expr, C)
            C[idx] = C[idx] * v if op == "*=" else v
    return [C[j] for j in range(35)]


def purkinje_tables():
    """Purkinje.h's initial state, Xi and cell surface area, and the
    (centre, slope) pair of every logistic steady state of Purkinje.cc's
    ComputeRates, x_inf = 1/(1 + exp((V + a)/b)), keyed by the state whose
    rate uses it (numbers only)."""
    h = strip_comments(open(os.path.join(REF, "Purkinje.h")).read())
    init = indexed_literals(function_body(h, "Purkinje(Real *constants)"), "_State")
    xi = float(re.search(r"_Xi\s*=\s*" + NUM, h).group(1))
    area = float(re.search(r"_SurfaceArea\s*=\s*" + NUM, h).group(1))
    cc = strip_comments(open(os.path.join(REF, "Purkinje.cc")).read())
    body = function_body(cc, "void Purkinje::ComputeRates")
    logistic = {}
    for m in re.finditer(r"_Algebraic\[(\d+)\]\s*=\s*1\.0+/\(1\.0+\+\(exp\(\(\(_State\[0\]\s*([+-])\s*"
                         + NUM + r"\)/" + NUM + r"\)\)\)\);", body):
        sign = -1.0 if m.group(2) == "-" else 1.0
        logistic[int(m.group(1))] = (sign * float(m.group(3)), float(m.group(4)))
    for m in re.finditer(r"_Algebraic\[(\d+)\]\s*=\s*_Algebraic\[(\d+)\]\s*;", body):
        if int(m.group(2)) in logistic:
            logistic[int(m.group(1))] = logistic[int(m.group(2))]
    gates = {}
    for m in re.finditer(r"_Rates\[(\d+)\]\s*=\s*\(_Algebraic\[(\d+)\]\s*-\s*_State\[(\d+)\]\)", body):
        g, k = int(m.group(1)), int(m.group(2))
        if g == int(m.group(3)) and k in logistic:
            gates[str(g)] = list(logistic[k])
    return {"purkinje_initial_state": init, "purkinje_xi": xi, "purkinje_area_cm2": area,
            "purkinje_logistic_gates": gates}


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
    mj_cc = strip_comments(open(os.path.join(REF, "Mahajan.cc")).read())
    out = {
        "source": "voomEP/src/IonicModel (SetUpIonicConstants.h, Tusscher.h, LR.cc), numbers only",
        "tusscher_constants": tt,
        "tusscher_initial_state": tt_init,
        "tusscher_xi": float(xi.group(1)),
        "purkinje_constants_literal": pk,
        "luorudy_initial_state": lr_init,
        "luorudy_constants": lr_consts,
    }
    out.update(mahajan_tables(const_h, mj_cc))
    out.update(purkinje_tables())
    fail = function_body(const_h, "void SetUpMahajan_failParameters")
    out["mahajan_fail_constants_by_type"] = [run_constants(fail, t) for t in range(9)]
    out["mahajan_fail_constants_default"] = run_constants(fail, 99)
    normal = function_body(const_h, "void SetUpMahajanParameters")
    out["mahajan_constants_by_type_evaluated"] = [run_constants(normal, t) for t in range(9)]
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {OUT}: {len(tt)} tT constants, {len(tt_init)} tT states, "
          f"{len(lr_init)} LR states", file=sys.stderr)


if __name__ == "__main__":
    main()
