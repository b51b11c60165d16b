"""Independent reader of the CHEMKIN inputs (TEST INFRASTRUCTURE ONLY).

A second implementation, in Python, of what the reference's Species and
Reaction constructors build (src/ChemData.cpp:11-948, src/Element.cpp:13-31,
src/Constants.h:17-34): it fills the same `ostc_mech_tables` layout
(include/openstc/mech.h) so the oracle can run on tables that the product's
C++ reader did not produce, and the tests can compare the two readers.
"""
from __future__ import annotations

import ctypes as C
import math
import re

import numpy as np

MAX_SPEC, MAX_REAC, MAX_ELEM = 16, 32, 8
R_C = 1.98720425864083
K_BOLTZMANN = 1.380649e-23
ATOMIC_WEIGHT = {"H": 1.008, "HE": 4.002602, "C": 12.011, "N": 14.007, "O": 15.999, "AR": 39.95}


class MechTables(C.Structure):
    S, R, E = MAX_SPEC, MAX_REAC, MAX_ELEM
    _fields_ = [
        ("n_spec", C.c_int), ("n_reac", C.c_int), ("n_elem", C.c_int),
        ("spec_name", (C.c_char * 16) * S), ("elem_name", (C.c_char * 4) * E),
        ("elem_comp", (C.c_int * E) * S), ("mw", C.c_double * S),
        ("t_low", C.c_double * S), ("t_mid", C.c_double * S), ("t_high", C.c_double * S),
        ("low_coeff", (C.c_double * 7) * S), ("high_coeff", (C.c_double * 7) * S),
        ("lj_potent_inv", C.c_double * S), ("vis_coeff", C.c_double * S),
        ("wj_div_wi_to_one4th", (C.c_double * S) * S), ("sqrt_wi_div_wj_pl1_mul8", (C.c_double * S) * S),
        ("binary_diffusivity_coeff", (C.c_double * S) * S), ("kb_over_eps_jk", (C.c_double * S) * S),
        ("label", C.c_int * R), ("rev_type", C.c_int * R), ("order", C.c_int * R),
        ("stoi_f", (C.c_int * S) * R), ("stoi_b", (C.c_int * S) * R),
        ("A", C.c_double * R), ("b", C.c_double * R), ("Ea", C.c_double * R),
        ("A2", C.c_double * R), ("b2", C.c_double * R), ("Ea2", C.c_double * R),
        ("third_body_coeff", (C.c_double * S) * R),
        ("troe_alpha", C.c_double * R), ("troe_t3", C.c_double * R), ("troe_t1", C.c_double * R),
        ("troe_t2", C.c_double * R),
    ]


def _tokens(text: str):
    out = []
    for line in text.splitlines():
        for t in line.split():
            if t.startswith("!"):
                break
            out.append(t)
    return out


def _species_elements(text: str):
    tok = _tokens(text)
    i = next(k for k, t in enumerate(tok) if t in ("ELEMENTS", "ELEM")) + 1
    elems = []
    while tok[i].upper() not in ("SPECIES", "SPEC"):
        if tok[i].upper() != "END" and tok[i].upper() not in elems:
            elems.append(tok[i].upper())
        i += 1
    i += 1
    spec = []
    while i < len(tok) and tok[i].upper() not in ("REACTIONS", "REAC", "THERMO"):
        if tok[i].upper() != "END" and tok[i].upper() not in spec:
            spec.append(tok[i].upper())
        i += 1
    return spec, elems


def _therm(lines, spec, elems):
    """NASA-7 cards: name in columns 1-18, element counts from column 25,
    T_low/T_high/T_mid from column 46; then a1..a7 high, a1..a7 low."""
    ns = len(spec)
    data = {}
    start = next(k for k, l in enumerate(lines) if l.strip().upper().startswith("THERMO")) + 1
    while lines[start].strip() == "" or lines[start].startswith("!"):
        start += 1
    k = start + 1
    while k < len(lines):
        card = lines[k].lstrip().upper()
        if not card or card.startswith("!"):
            k += 1
            continue
        key = card.split()[0]
        if key == "END":
            k += 1
            continue
        if key in ("REACTIONS", "TRANSPORT", "TRAN") or len(data) >= ns:
            break
        name = card[:18].split()[0]
        if name not in spec:
            k += 4
            continue
        if name in data:
            k += 5
            continue
        tl, th = float(card[45:55]), float(card[55:65])
        tm_s = card[65:75].split()
        tm = float(tm_s[0]) if tm_s else 1000.0
        comp = [0] * len(elems)
        for r in range(4):
            f = card[24 + 5 * r: 29 + 5 * r].lstrip()
            if not f or f[0] == "0":
                break
            parts = f.split()
            if parts[0] in elems:
                comp[elems.index(parts[0])] = int(parts[1])
        c = [float(lines[k + 1 + q // 5][15 * (q % 5): 15 * (q % 5) + 15]) for q in range(14)]
        data[name] = (tl, tm, th, comp, c[:7], c[7:14])
        k += 4
    return data


def _tran(lines, spec):
    data = {}
    for l in lines:
        l = l.lstrip().upper()
        if not l or l.startswith("!"):
            continue
        f = l.split()
        if f[0].startswith("END") or f[0].startswith("REAC"):
            break
        if f[0] in spec and f[0] not in data:
            data[f[0]] = [float(v) for v in f[1:7]]
    return data


def _reactions(lines, spec):
    i = next(k for k, l in enumerate(lines) if l.lstrip().startswith("REAC")) + 1
    body = []
    for l in lines[i:]:
        u = l.lstrip().upper()
        if u.startswith("END"):
            break
        if not u or u.startswith("!"):
            continue
        body.append(u)
    recs = []
    for u in body:
        if "=" in u:
            recs.append({"line": u, "aux": []})
        else:
            recs[-1]["aux"].append(u.replace("/", " "))
    out = []
    open_dups = []
    for r in recs:
        line = r["line"]
        label = 1
        if "+M" in line:
            label = 4
            if "(+" in line:
                label = 5
                line = re.sub(r"\(\+[^)]*\)", "", line)
            else:
                line = line.replace("+M", "", 2)
        f = line.split()
        eq, A, b, Ea = f[0], float(f[1]), float(f[2]), float(f[3]) / R_C
        irrev = "=>" in eq and "<=>" not in eq
        lhs, rhs = re.split(r"<?=>?", eq, maxsplit=1)
        if irrev:
            label = 0

        def side(s):
            nu = [0] * len(spec)
            for term in s.split("+"):
                n = 1
                if term[0].isdigit():
                    n, term = int(term[0]), term[1:]
                nu[spec.index(term)] += n
            return nu

        rec = {"label": label, "rev": 0, "f": side(lhs), "b": side(rhs), "A": A, "beta": b, "Ea": Ea,
               "A2": 0.0, "b2": 0.0, "Ea2": 0.0, "tb": [1.0] * len(spec), "troe": [0.0] * 4, "dup": False}
        for a in r["aux"]:
            w = a.split()
            if w[0] in ("DUPLICATE", "DUP"):
                rec["dup"] = True
                rec["label"] = 3
            elif w[0] == "LOW":
                rec["A2"], rec["b2"], rec["Ea2"] = float(w[1]), float(w[2]), float(w[3]) / R_C
            elif w[0] == "REV":
                rec["A2"], rec["b2"], rec["Ea2"] = float(w[1]), float(w[2]), float(w[3]) / R_C
                rec["rev"] = 1
            elif w[0] == "TROE":
                vals = [float(v) for v in w[1:]]
                rec["troe"][:3] = vals[:3]
                rec["label"] = 6
                if len(vals) > 3:
                    rec["troe"][3] = vals[3]
                    rec["label"] = 7
            elif w[0] != "PLOG":
                for name, v in zip(w[0::2], w[1::2]):
                    rec["tb"][spec.index(name)] = float(v)
        if rec["dup"]:
            twin = next((d for d in open_dups if d["f"] == rec["f"] and d["b"] == rec["b"]), None)
            if twin is not None:
                twin["A2"], twin["b2"], twin["Ea2"] = rec["A"], rec["beta"], rec["Ea"]
                open_dups.remove(twin)
                continue
            open_dups.append(rec)
        out.append(rec)
    return out


def load(mech_file: str, therm_file: str, tran_file: str, reaction: int = 1) -> MechTables:
    text = open(mech_file).read()
    spec, elems = _species_elements(text)
    T = MechTables()
    ns = T.n_spec = len(spec)
    T.n_elem = len(elems)
    for i, s in enumerate(spec):
        T.spec_name[i].value = s.encode()
    for i, e in enumerate(elems):
        T.elem_name[i].value = e.encode()
    th = _therm(open(therm_file).read().splitlines(), spec, elems)
    tr = _tran(open(tran_file).read().splitlines(), spec)
    mw = np.zeros(ns)
    for i, s in enumerate(spec):
        tl, tm, thh, comp, hi, lo = th[s]
        T.t_low[i], T.t_mid[i], T.t_high[i] = tl, tm, thh
        for e in range(len(elems)):
            T.elem_comp[i][e] = comp[e]
        w = 0.0
        for e in sorted(elems):  # summation in element-name order
            w += ATOMIC_WEIGHT[e] * comp[elems.index(e)]
        mw[i] = T.mw[i] = w
        for q in range(7):
            T.high_coeff[i][q] = hi[q]
            T.low_coeff[i][q] = lo[q]
    geom, eps, sig, dip, pol = (np.array([tr[s][q] for s in spec]) for q in range(5))
    for i in range(ns):
        T.lj_potent_inv[i] = 1.0 / eps[i]
        T.vis_coeff[i] = 2.6693e-6 * math.sqrt(mw[i]) / (sig[i] * sig[i])
        for j in range(ns):
            T.wj_div_wi_to_one4th[i][j] = (mw[j] / mw[i]) ** 0.25
            T.sqrt_wi_div_wj_pl1_mul8[i][j] = 1.0 / math.sqrt(8 * (1 + mw[i] / mw[j]))

    def mu_star(i):
        if dip[i] == 0:
            return 0.0
        conv = 1e-18 * math.sqrt(1e-7) * (1e8 ** 1.5) / math.sqrt(K_BOLTZMANN)
        return conv * dip[i] / math.sqrt(eps[i] * sig[i] * sig[i] * sig[i])

    for j in range(ns):
        T.binary_diffusivity_coeff[j][j] = 0.0188324 / (math.sqrt(mw[j] * 0.5) * sig[j] * sig[j])
        for k in range(j + 1, ns):
            wjk = mw[j] * mw[k] / (mw[j] + mw[k])
            xi = 1.0
            if (dip[j] != 0) != (dip[k] != 0):
                p, n = (j, k) if dip[j] != 0 else (k, j)
                a_star, m_star = pol[n] / sig[n] ** 3, mu_star(p)
                xi = 1.0 + 0.25 * a_star * m_star * m_star * math.sqrt(eps[p] / eps[n])
            sjk = 0.5 * (sig[j] + sig[k]) * xi ** (-1.0 / 6.0)
            T.binary_diffusivity_coeff[j][k] = T.binary_diffusivity_coeff[k][j] = \
                0.0188324 / (math.sqrt(wjk) * sjk * sjk)
            T.kb_over_eps_jk[j][k] = T.kb_over_eps_jk[k][j] = 1.0 / (xi * xi * math.sqrt(eps[j] * eps[k]))
    for r in range(MAX_REAC):
        for s in range(MAX_SPEC):
            T.third_body_coeff[r][s] = 1.0
    if reaction == 1:
        recs = _reactions(open(mech_file).read().splitlines(), spec)
        T.n_reac = len(recs)
        for r, rec in enumerate(recs):
            T.label[r], T.rev_type[r] = rec["label"], rec["rev"]
            T.order[r] = sum(rec["b"]) - sum(rec["f"])
            for s in range(ns):
                T.stoi_f[r][s], T.stoi_b[r][s] = rec["f"][s], rec["b"][s]
                T.third_body_coeff[r][s] = rec["tb"][s]
            T.A[r], T.b[r], T.Ea[r] = rec["A"], rec["beta"], rec["Ea"]
            T.A2[r], T.b2[r], T.Ea2[r] = rec["A2"], rec["b2"], rec["Ea2"]
            T.troe_alpha[r], T.troe_t3[r], T.troe_t1[r], T.troe_t2[r] = rec["troe"]
    return T


def species_names(T: MechTables):
    return [T.spec_name[i].value.decode() for i in range(T.n_spec)]
