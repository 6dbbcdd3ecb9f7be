"""Planted mistakes: the oracle's pins must catch plausible transcription errors.

Each mutation below changes one spot of a private copy of oracle/oracle_ionic.c
-- a digit dropped or moved, a sign flipped, a centre or slope off, a current
dropped, a stoichiometric factor or state index wrong -- builds that copy into
a temporary library, and runs the pin tests of the mutated cell model against
it (ORACLE_LIB points the oracle package at the copy).  Every mutation must
make at least one pin fail; the unmutated copy must pass them all.  The pins
are the invariants, closed forms, published half points, reference tables and
markers in tests/test_oracle_{ionic,gate_kinetics,lr_gate,mahajan,purkinje}.py.

Mutations are written against the oracle's own source text (test
infrastructure); each must occur exactly once there.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
ORACLE = os.path.join(ROOT, "oracle")
SRC = os.path.join(ORACLE, "oracle_ionic.c")

PINS = {
    "lr": ["test_oracle_ionic.py", "test_oracle_gate_kinetics.py", "test_oracle_lr_gate.py"],
    "tt": ["test_oracle_ionic.py", "test_oracle_gate_kinetics.py"],
    "mj": ["test_oracle_mahajan.py"],
    "pk": ["test_oracle_purkinje.py"],
}

# (model, what, original text, mutated text)
MUTATIONS = [
    # Luo-Rudy 1991 (LR.cc:102-183)
    ("lr", "alpha_m centre 47.13 -> 47.31",
     "else alpha_m = 0.32 * (v + 47.13)", "else alpha_m = 0.32 * (v + 47.31)"),
    ("lr", "beta_m slope 1/11 -> 1/11.1", "beta_m = 0.08 * exp(-v / 11.);",
     "beta_m = 0.08 * exp(-v / 11.1);"),
    ("lr", "alpha_h slope 6.8 -> 8.6", "alpha_h = 0.135 * exp(-(80 + v) / 6.8);",
     "alpha_h = 0.135 * exp(-(80 + v) / 8.6);"),
    ("lr", "beta_h second rate 0.35 -> 0.035", "3.1e5 * exp(0.35 * v)", "3.1e5 * exp(0.035 * v)"),
    ("lr", "beta_j (V<-40) slope 0.01052 -> 0.1052", "0.1212 * exp(-0.01052 * v)",
     "0.1212 * exp(-0.1052 * v)"),
    ("lr", "alpha_d centre sign", ".095 * exp(-.01 * (v - 5))", ".095 * exp(-.01 * (v + 5))"),
    ("lr", "beta_d amplitude .07 -> .007", "beta_d = .07 *", "beta_d = .007 *"),
    ("lr", "beta_f slope .02 -> .2", ".0065 * exp(-.02 * (v + 30))", ".0065 * exp(-.2 * (v + 30))"),
    ("lr", "beta_X centre 20 -> 2", ".0013 * exp(-.06 * (v + 20))", ".0013 * exp(-.06 * (v + 2))"),
    ("lr", "I_b reversal sign", "0.0392 * (v + 59.87)", "0.0392 * (v - 59.87)"),
    ("lr", "alpha_j centre 37.78 -> 37.87", "(v + 37.78)", "(v + 37.87)"),
    ("lr", "E_Na 54.4 -> 45.4", "(in[7] - 54.4)", "(in[7] - 45.4)"),
    ("lr", "E_si sign", "const double Es = 7.7 - 13.0287 * log(in[6]);",
     "const double Es = 7.7 + 13.0287 * log(in[6]);"),
    ("lr", "Ca_i uptake of I_si x10", "(-.0001 * i_si", "(-.001 * i_si"),
    # ten Tusscher 2006 (Tusscher.cc:4-96)
    ("tt", "f gate centre 20 -> 2", "A[7] = 1.0 / (1.0 + exp((S[0] + 20.0) / 7.0));",
     "A[7] = 1.0 / (1.0 + exp((S[0] + 2.0) / 7.0));"),
    ("tt", "Xr1 steady state slope 7 -> 17", "A[0] = 1.0 / (1.0 + exp((-26.0 - S[0]) / 7.0));",
     "A[0] = 1.0 / (1.0 + exp((-26.0 - S[0]) / 17.0));"),
    ("tt", "h gate centre 71.55 -> 75.15", "A[4] = 1.0 / pow(1.0 + exp((S[0] + 71.55) / 7.43), 2.0);",
     "A[4] = 1.0 / pow(1.0 + exp((S[0] + 75.15) / 7.43), 2.0);"),
    ("tt", "r gate tau constant 0.8 -> 8", "A[24] = 9.5 * exp(-pow(S[0] + 40.0, 2.0) / 1800.0) + 0.8;",
     "A[24] = 9.5 * exp(-pow(S[0] + 40.0, 2.0) / 1800.0) + 8.0;"),
    ("tt", "Na_i loses the 3 I_NaK", "3.0 * A[55] + 3.0 * A[56]", "2.0 * A[55] + 3.0 * A[56]"),
    ("tt", "I_Ks reversal uses the wrong state", "A[41] = RTF * log((C[10] + C[9] * C[11]) / (S[1] + C[9] * S[2]));",
     "A[41] = RTF * log((C[10] + C[9] * C[11]) / (S[2] + C[9] * S[1]));"),
    ("tt", "I_CaL driving potential 15 -> 1.5", "4.0 * (S[0] - 15.0)", "4.0 * (S[0] - 1.5)"),
    ("tt", "I_pK dropped from dV/dt", "A[56] + A[53] +\n           A[58] + A[57] + A[12]);",
     "A[56] + A[53] +\n           A[57] + A[12]);"),
    ("tt", "RyR recovery sign", "C[36] * (1.0 - S[18]);", "C[36] * (1.0 + S[18]);"),
    ("tt", "SS diffusion reversed", "A[61] = C[41] * (S[10] - S[3]);",
     "A[61] = C[41] * (S[3] - S[10]);"),
    # Mahajan 2008 (Mahajan.cc:168-644)
    ("mj", "L-type chain k5/k6 swapped", "const double k6 = fca * Ps / tau_ca, k5 = (1.0 - Ps) / tau_ca;",
     "const double k5 = fca * Ps / tau_ca, k6 = (1.0 - Ps) / tau_ca;"),
    ("mj", "Na_i loses a third of I_NaK", "i_na + 3.0 * i_nak", "i_na + 2.0 * i_nak"),
    ("mj", "tau_ys offset 30 -> 3", "const double tau_ys = 3000.0 / (1.0 + exp((V + 60.0) / 10.0)) + 30.0;",
     "const double tau_ys = 3000.0 / (1.0 + exp((V + 60.0) / 10.0)) + 3.0;"),
    ("mj", "SR leak sign", "(16.667 * cj - ci)", "(16.667 * cj + ci)"),
    ("mj", "spark rate 2.58079 -> 25.8079", "const double spark = 2.58079 *",
     "const double spark = 25.8079 *"),
    ("mj", "I_Ks reversal Na permeability 0.01833 -> 0.1833", "log((ko + 0.018330 * nao) / (ki + 0.018330 * nai))",
     "log((ko + 0.18330 * nao) / (ki + 0.018330 * nai))"),
    ("mj", "beta_m slope 11 -> 1.1", "const double bm = 0.08 * exp(-V / 11.0);",
     "const double bm = 0.08 * exp(-V / 1.10);"),
    ("mj", "alpha_h slope 6.8 -> 8.6", "exp((80.0 + V) / (-6.8))", "exp((80.0 + V) / (-8.6))"),
    # Purkinje (Purkinje.cc:5-210)
    ("pk", "I_CaT activation centre 47.8 -> 48.7", "A[1] = 1.0 / (1.0 + exp((S[0] + 47.8) / -5.5));",
     "A[1] = 1.0 / (1.0 + exp((S[0] + 48.7) / -5.5));"),
    ("pk", "I_Na inactivation slope 3.96 -> 3.69", "A[9] = 1.0 / (1.0 + exp((S[0] + 69.0) / 3.96));",
     "A[9] = 1.0 / (1.0 + exp((S[0] + 69.0) / 3.69));"),
    ("pk", "I_CaT inactivation tau 1.42271 -> 1.4271", "A[17] = 1.42271 * exp(-0.05119 * S[0]);",
     "A[17] = 1.4271 * exp(-0.05119 * S[0]);"),
    ("pk", "Na_i with 2 I_NaK", "-3.0 * A[52] + -3.0 * A[49]", "-2.0 * A[52] + -3.0 * A[49]"),
    ("pk", "I_Kr gate time constant factor 4 -> 2", "A[30] = 4.0 * A[22];", "A[30] = 2.0 * A[22];"),
    ("pk", "adaptive threshold 1.5 -> 15", "if (fabs(rates[0]) > 1.5) {", "if (fabs(rates[0]) > 15.0) {"),
    ("pk", "I_f Na reversal uses K", "A[64] = C[58] * log(C[4] / S[4]);",
     "A[64] = C[58] * log(C[5] / S[5]);"),
    ("pk", "Na/Ca exchanger scale 512.009 -> 51.2009", "A[49] = (512.009 * C[33] *",
     "A[49] = (51.2009 * C[33] *"),
    ("pk", "SR uptake sign", "R[3] = ((-1.0 * A[61] + -1.0 * A[67] + A[65] + A[63])",
     "R[3] = ((-1.0 * A[61] + -1.0 * A[67] + A[65] - A[63])"),
]


def _build(tmp, text):
    for f in ("oracle_fe.c", "oracle.h"):
        shutil.copy(os.path.join(ORACLE, f), os.path.join(tmp, f))
    with open(os.path.join(tmp, "oracle_ionic.c"), "w") as f:
        f.write(text)
    lib = os.path.join(tmp, "liboracle.so")
    subprocess.run(["gcc", "-O1", "-fPIC", "-fopenmp", "-std=gnu11", "-shared", "-o", lib,
                    os.path.join(tmp, "oracle_fe.c"), os.path.join(tmp, "oracle_ionic.c"), "-lm"],
                   check=True, capture_output=True)
    return lib


def _pins_pass(lib, model):
    env = dict(os.environ, ORACLE_LIB=lib, OMP_NUM_THREADS="1")
    files = [os.path.join(HERE, f) for f in PINS[model]]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-p", "no:cacheprovider",
                        *files], cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    return r.returncode == 0, r.stdout[-400:]


def _run(job):
    model, what, old, new = job
    text = open(SRC).read()
    with tempfile.TemporaryDirectory() as tmp:
        if old is not None:
            if text.count(old) != 1:
                return job, None, f"appears {text.count(old)} times"
            text = text.replace(old, new)
        ok, tail = _pins_pass(_build(tmp, text), model)
    return job, ok, tail


def test_planted_mistakes_fail_the_pins():
    jobs = [(m, "unmutated", None, None) for m in PINS] + MUTATIONS
    workers = max(1, min(8, len(os.sched_getaffinity(0))))
    with cf.ThreadPoolExecutor(workers) as ex:
        results = list(ex.map(_run, jobs))
    bad = []
    for (model, what, old, new), ok, tail in results:
        if ok is None:
            bad.append(f"{model}: '{what}': the original text {tail}")
        elif old is None and not ok:
            bad.append(f"{model}: the unmutated oracle fails its pins:\n{tail}")
        elif old is not None and ok:
            bad.append(f"{model}: '{what}' passes every pin")
        print(f"{model:3s} {what:45s} {'caught' if old is not None and not ok else 'ok' if ok else '-'}")
    assert not bad, "\n".join(bad)
