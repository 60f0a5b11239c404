"""Reader of generated mechanism files (a case's Chemistry.H) into the
library's VD_CHEM_GEN reaction table.

A Vidyut case compiles its mechanism in: py_utils/ceptr_noThermo (or an
earlier ceptr) writes Chemistry.H with one block per reaction inside
productionRate (e.g. test/Reactor/CO2_H2_Ar_mechanism_reduced/
Chemistry.H:913-6380, test/verification/He_RF_1d/Chemistry.H:447-600).
This module reads that generated code as data: the species (the *_ID
defines, CKSYMS_STR of Chemistry.cpp when given, CKCHRG, get_mw), the
g/RT polynomials of gibbs(), and for every reaction block its forward rate,
third body, reverse rate, stoichiometry and energy-exchange type.  Every
statement of a block must be one of the generated forms listed in
_RATE_FORMS and the block parser; anything else raises ValueError, so a
mechanism is never converted approximately.

    python -m vidyut3d_amd.mechgen Chemistry.H [Chemistry.cpp] -o mech.json

`load(path)` turns the JSON into a `vd_mech_desc` (transport left at its
defaults: the case's specMob/specDiff map onto vd_transport_fit by hand).
"""
import json
import math
import re
import sys

from . import abi

NA_GEN = 6.02214085774e23  # the Avogadro constant written into the fits

_UNUM = r"(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?"
_NUM = r"[-+]?" + _UNUM
_IDENT_OK = {
    "const", "amrex", "Real", "double", "int", "for", "if", "else", "std", "min", "exp", "pow", "log",
    "k_f", "qf", "qr", "qdot", "wdot", "sc", "Corr", "mixture", "g_RT", "refC", "refCinv", "tc", "invT",
    "logT", "Te", "TeeV", "Janev_sum", "Jfit_coefs", "Ffit_coefs", "Jfit_A", "Ffit_A", "Te_pow",
    "invTe_pow", "j", "rxntype", "eexci", "elidx", "comp_ener_exch", "enerExch", "Ue", "reaction",
}


def _num(s):
    return float(s)


def _strip_comments(text):
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    return re.sub(r"//[^\n]*", " ", text)


def _function_body(text, signature_re):
    m = re.search(signature_re, text)
    if not m:
        raise ValueError("no %s in the file" % signature_re)
    i = text.index("{", m.end())
    depth, j = 0, i
    while True:
        c = text[j]
        if c == "{":
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                return text[i + 1:j]
        j += 1


def _species(h_text, cpp_text):
    ids = {m.group(1): int(m.group(2)) for m in re.finditer(r"#define\s+(\w+)_ID\s+(\d+)", h_text)}
    ns = int(re.search(r"#define\s+NUM_GAS_SPECIES\s+(\d+)", h_text).group(1)) \
        if re.search(r"#define\s+NUM_GAS_SPECIES\s+(\d+)", h_text) else len(ids)
    names = [None] * ns
    if cpp_text:
        body = _function_body(cpp_text, r"void\s+CKSYMS_STR\s*\(")
        for m in re.finditer(r"kname\[(\d+)\]\s*=\s*\"([^\"]+)\"", body):
            names[int(m.group(1))] = m.group(2)
    for name, k in ids.items():
        if k < ns and names[k] is None:
            names[k] = name
    if any(n is None for n in names):
        raise ValueError("species names incomplete")
    body = _function_body(h_text, r"void\s+CKCHRG\s*\(\s*int\s+kcharge\[\]\s*\)")
    charge = [None] * ns
    for m in re.finditer(r"kcharge\[(\d+)\]\s*=\s*(-?\d+)", body):
        charge[int(m.group(1))] = int(m.group(2))
    body = _function_body(h_text, r"void\s+get_mw\s*\(\s*amrex::Real\s*\*\s*mw_new\s*\)")
    molwt = [None] * ns
    for m in re.finditer(r"mw_new\[(\d+)\]\s*=\s*(%s)" % _NUM, body):
        molwt[int(m.group(1))] = _num(m.group(2))
    if any(c is None for c in charge) or any(w is None for w in molwt):
        raise ValueError("charges or molecular weights incomplete")
    return names, ids, charge, molwt


_GIBBS_VARS = {"invT": 0, "": 1, "logT": 2, "T": 3, "T2": 4, "T3": 5, "T4": 6}


def _gibbs_terms(expr):
    """'+a * invT + b - c * logT ...' -> the 7 coefficients (signed)."""
    e = re.sub(r"\s+", "", expr)
    coef = [0.0] * 7
    pos, seen = 0, set()
    while pos < len(e):
        m = re.match(r"([-+]?)(%s)(?:\*(invT|logT|T2|T3|T4|T))?" % _UNUM, e[pos:])
        if not m or m.end() == 0:
            raise ValueError("g/RT term not understood at %r" % e[pos:pos + 40])
        v = _num(m.group(2))
        if m.group(1) == "-":
            v = -v
        var = m.group(3) or ""
        if var in seen:
            raise ValueError("repeated g/RT term %r" % var)
        seen.add(var)
        coef[_GIBBS_VARS[var]] = v
        pos += m.end()
    return coef


def _gibbs(h_text, ns):
    m = re.search(r"void\s+gibbs\s*\(\s*amrex::Real\s*\*\s*species\s*,\s*const\s+amrex::Real\s+(T|\*tc)\s*\)",
                  h_text)
    if not m:
        raise ValueError("no gibbs(species, T) in the file")
    body = _strip_comments(_function_body(h_text, re.escape(m.group(0))))
    if m.group(1) != "T":
        # older generators pass the temperature cache; the polynomial is the same
        pass
    out = [None] * ns
    # split into the flat part and the if/else ranges
    pos = 0
    tmid, branch = None, None
    tokens = re.finditer(r"if\s*\(\s*T\s*<\s*(%s)\s*\)\s*\{|\}\s*else\s*\{|\}|species\[(\d+)\]\s*=\s*([^;]+);" % _NUM,
                         body)
    for t in tokens:
        s = t.group(0)
        if s.startswith("if"):
            tmid, branch = _num(t.group(1)), "lo"
        elif s.startswith("}") and "else" in s:
            branch = "hi"
        elif s == "}":
            tmid, branch = None, None
        else:
            k = int(t.group(2))
            c = _gibbs_terms(t.group(3))
            if out[k] is None:
                out[k] = {"tmid": 0.0, "lo": c, "hi": c}
            if tmid is not None:
                out[k]["tmid"] = tmid
                out[k][branch] = c
            else:
                out[k]["lo"] = out[k]["hi"] = c
    for k in range(ns):
        if out[k] is None:
            out[k] = {"tmid": 0.0, "lo": [0.0] * 7, "hi": [0.0] * 7, "missing": True}
    return out


def _sp(tok, ids):
    tok = tok.strip()
    if tok.endswith("_ID"):
        return ids[tok[:-3]]
    return int(tok)


# forward-rate forms of a block, whitespace removed: (regex, builder)
def _rate_forms():
    N = _NUM
    return [
        (r"(%s)" % N, lambda m: dict(kind="arrh", A=_num(m[1]))),
        (r"(%s)\*exp\(\((%s)\)\*(?:logT|tc\[0\])\)" % (N, N),
         lambda m: dict(kind="arrh", A=_num(m[1]), beta=_num(m[2]))),
        (r"(%s)\*exp\(-\((%s)\)\*invT\)" % (N, N),
         lambda m: dict(kind="arrh", A=_num(m[1]), Ea=_num(m[2]))),
        (r"(%s)\*exp\(\((%s)\)\*(?:logT|tc\[0\])-\((%s)\)\*invT\)" % (N, N, N),
         lambda m: dict(kind="arrh", A=_num(m[1]), beta=_num(m[2]), Ea=_num(m[3]))),
        (r"(%s)\*(?:std::)?pow\(Te,(%s)\)" % (N, N), lambda m: dict(kind="tepow", A=_num(m[1]), beta=_num(m[2]))),
        (r"(%s)\*(?:std::)?pow\(Te/(%s),(%s)\)" % (N, N, N),
         lambda m: dict(kind="tepow", A=_num(m[1]), tref=_num(m[2]), beta=_num(m[3]))),
        (r"(%s)\*(?:std::)?pow\(TeeV,(%s)\)" % (N, N),
         lambda m: dict(kind="tepow", A=_num(m[1]), tref=11595.0, beta=_num(m[2]))),
        (r"(%s)\*(?:std::)?pow\(Te,(%s)\)/tc\[1\]" % (N, N),
         lambda m: dict(kind="tepow", A=_num(m[1]), beta=_num(m[2]), tgas_div=1)),
        (r"(%s)\*(?:std::)?pow\(Te,(%s)\)\*(?:std::)?exp\(-(%s)/Te\)" % (N, N, N),
         lambda m: dict(kind="tearrh", A=_num(m[1]), beta=_num(m[2]), Ea=_num(m[3]))),
    ]


_RATE_FORMS = _rate_forms()


def _fit(block, which):
    """(A, coefficients, loop count) of the J(anev) or F(fit) part."""
    m = re.search(r"%sfit_coefs=\{([^}]*)\};double%sfit_A=(%s);for\(intj=0;j<(\d+);j\+\+\)Janev_sum\+=%sfit_coefs"
                  r"\[j\]\*(Te_pow|invTe_pow)\[j\];" % (which, which, _NUM, which), block)
    if not m:
        raise ValueError("%sfit part not understood" % which)
    coefs = [_num(x) for x in m.group(1).split(",")]
    n = int(m.group(3))
    want = (9, "Te_pow") if which == "J" else (4, "invTe_pow")
    if (n, m.group(4)) != want or len(coefs) != n:
        raise ValueError("%sfit with %d terms over %s" % (which, n, m.group(4)))
    return _num(m.group(2)), coefs


def _eval_const(e):
    """A literal or a product of literals (the fits' caps)."""
    parts = e.split("*")
    v = _num(parts[0])
    for p in parts[1:]:
        v = v * _num(p)
    return v


def _parse_block(label, eq, raw, ids, ns):
    text = _strip_comments(raw)
    idents = set(re.findall(r"[A-Za-z_]\w*", re.sub(r"(?<![\w.])(?:\d+\.?\d*|\.\d+)(?:[eE][-+]?\d+)?", " ", text)))
    unknown = {x for x in idents if x not in _IDENT_OK and not x.endswith("_ID")}
    if unknown:
        raise ValueError("reaction %s: unknown identifiers %s" % (label, sorted(unknown)))
    b = re.sub(r"\s+", "", text)
    r = {"label": label, "equation": eq, "kind": "arrh", "A": 0.0, "beta": 0.0, "Ea": 0.0, "tref": 1.0,
         "tgas_div": 0, "fit": [0.0] * 9, "A2": 0.0, "fit2": [0.0] * 4, "fit_switch_ev": 0.0, "kmax": 0.0}
    if "fit_coefs" in b:
        mk = re.search(r"k_f=(std::min\()?(J|F)fit_A\*exp\(Janev_sum\)\*(%s)(?:,(%s(?:\*%s)?)\))?;" % (_NUM, _NUM, _NUM),
                       b)
        if not mk or _num(mk.group(3)) != NA_GEN:
            raise ValueError("reaction %s: fitted k_f not understood" % label)
        piece = re.search(r"if\(TeeV<(%s)\)\{" % _NUM, b)
        if piece:
            A, c = _fit(b, "J")
            A2, c2 = _fit(b, "F")
            if b.index("Jfit_coefs=") > b.index("}else{") or b.index("Ffit_coefs=") < b.index("}else{"):
                raise ValueError("reaction %s: piecewise fit in an unknown order" % label)
            r.update(kind="janev_ffit", A=A, fit=c, A2=A2, fit2=c2 + [0.0] * (4 - len(c2)),
                     fit_switch_ev=_num(piece.group(1)))
            if len(re.findall(r"k_f=", b)) != 2:
                raise ValueError("reaction %s: piecewise fit without two k_f" % label)
        elif "Jfit_coefs=" in b:
            A, c = _fit(b, "J")
            r.update(kind="janev", A=A, fit=c)
        else:
            A, c = _fit(b, "F")
            r.update(kind="ffit", A=A, fit=c + [0.0] * (9 - len(c)))
        if mk.group(4):
            if not mk.group(1):
                raise ValueError("reaction %s: cap without std::min" % label)
            r["kmax"] = _eval_const(mk.group(4))
    else:
        mk = re.findall(r"constamrex::Realk_f=([^;]+);", b)
        if len(mk) != 1:
            raise ValueError("reaction %s: %d k_f definitions" % (label, len(mk)))
        for rx, build in _RATE_FORMS:
            mm = re.fullmatch(rx, mk[0])
            if mm:
                r.update(build(mm))
                break
        else:
            raise ValueError("reaction %s: rate %r not understood" % (label, mk[0]))
    # third body
    corr = re.findall(r"constamrex::RealCorr=([^;]+);", b)
    if corr and corr != ["mixture"]:
        raise ValueError("reaction %s: third body %r not understood" % (label, corr))
    r["third_body"] = 1 if corr else 0
    # forward progress: qf = [Corr *] k_f * (product of sc[...])
    mq = re.search(r"constamrex::Realqf=(Corr\*)?k_f\*(\(.*?\));constamrex::Realqr", b)
    if not mq or bool(mq.group(1)) != bool(corr):
        raise ValueError("reaction %s: qf not understood" % label)
    prod = mq.group(2)
    if re.sub(r"sc\[\w+\]|[()*]", "", prod):
        raise ValueError("reaction %s: qf factor %r not understood" % (label, prod))
    factors = [_sp(t, ids) for t in re.findall(r"sc\[(\w+)\]", prod)]
    react = []
    for s in factors:
        if react and react[-1][0] == s:
            react[-1][1] += 1.0
        else:
            react.append([s, 1.0])
    r["react"] = react
    # reverse progress
    mr = re.search(r"constamrex::Realqr=([^;]+);", b)
    q = mr.group(1)
    r.update(reversible=0, g=[], refc_pow=0, rev=[])
    if q != "0.0" and not q.endswith("*(0.0)"):
        mrev = re.fullmatch(r"(Corr\*)?k_f\*exp\(-\(([^()]*)\)\)(?:\*\((refC|refCinv)\))?\*(\(.*\))", q)
        if not mrev or bool(mrev.group(1)) != bool(corr):
            raise ValueError("reaction %s: qr %r not understood" % (label, q))
        terms = re.findall(r"([-+]?)(?:(%s)\*)?g_RT\[(\d+)\]" % _UNUM, mrev.group(2))
        if re.sub(r"([-+]?)(?:(%s)\*)?g_RT\[(\d+)\]" % _UNUM, "", mrev.group(2)):
            raise ValueError("reaction %s: g/RT sum %r not understood" % (label, mrev.group(2)))
        g = []
        for sign, c, k in terms:
            v = _num(c) if c else 1.0
            g.append([int(k), -v if sign == "-" else v])
        rprod = mrev.group(4)
        if re.sub(r"sc\[\w+\]|[()*]", "", rprod):
            raise ValueError("reaction %s: qr factor %r not understood" % (label, rprod))
        rev = []
        for s in [_sp(t, ids) for t in re.findall(r"sc\[(\w+)\]", rprod)]:
            if rev and rev[-1][0] == s:
                rev[-1][1] += 1.0
            else:
                rev.append([s, 1.0])
        r.update(reversible=1, g=g, refc_pow={None: 0, "refC": 1, "refCinv": -1}[mrev.group(3)], rev=rev)
    elif q != "0.0":
        # the generator's reverse rate times (0.0): kept as data (its g/RT
        # sum and refC power), not evaluated
        mz = re.fullmatch(r"(Corr\*)?k_f\*exp\(-\(([^()]*)\)\)(?:\*\((refC|refCinv)\))?\*\(0\.0\)", q)
        if not mz:
            raise ValueError("reaction %s: zeroed qr %r not understood" % (label, q))
        terms = re.findall(r"([-+]?)(?:(%s)\*)?g_RT\[(\d+)\]" % _UNUM, mz.group(2))
        r["reverse_zeroed"] = {"g": [[int(k), -(_num(c) if c else 1.0) if sg == "-" else (_num(c) if c else 1.0)]
                                     for sg, c, k in terms],
                               "refc_pow": {None: 0, "refC": 1, "refCinv": -1}[mz.group(3)]}
    # the wdot updates as written (a species on both sides may take two)
    net = []
    for m in re.finditer(r"wdot\[(\w+)\](\+=|-=)(?:(%s)\*)?qdot;" % _UNUM, b):
        v = _num(m.group(3)) if m.group(3) else 1.0
        net.append([_sp(m.group(1), ids), v if m.group(2) == "+=" else -v])
    if len(net) != b.count("wdot["):
        raise ValueError("reaction %s: wdot update not understood" % label)
    r["wdot"] = net
    # energy exchange
    r.update(rxntype=0, eexci=0.0, elidx=0)
    if "comp_ener_exch(" in b:
        if not re.search(r"comp_ener_exch\(qf,qr,sc,k_f,rxntype,eexci,elidx,enerExch,Ue,tc\[1\],Te\);", b):
            raise ValueError("reaction %s: comp_ener_exch call not understood" % label)
        r["rxntype"] = int(re.search(r"intrxntype=(\d+);", b).group(1))
        r["eexci"] = _num(re.search(r"amrex::Realeexci=(%s);" % _NUM, b).group(1))
        r["elidx"] = _sp(re.search(r"intelidx=(\w+);", b).group(1), ids)
    return r


def _production_blocks(h_text):
    body = _function_body(h_text, r"productionRate\s*\(\s*amrex::Real\s*\*\s*wdot")
    # the reaction blocks are the top-level { ... } groups that hold a reaction comment
    out, depth, start = [], 0, None
    for i, c in enumerate(body):
        if c == "{":
            if depth == 0:
                start = i
            depth += 1
        elif c == "}":
            depth -= 1
            if depth == 0:
                seg = body[start:i + 1]
                m = re.search(r"//\s*reaction\s+(\d\w*):\s*([^\n]*)", seg)
                if m:
                    out.append((m.group(1), m.group(2).strip(), seg))
                elif "qdot" in seg:
                    raise ValueError("a block of productionRate without a reaction comment")
    # a block comment may introduce a second reaction in the same braces
    for label, eq, seg in out:
        if len(re.findall(r"//\s*reaction\s+\d\w*:", seg)) != 1:
            raise ValueError("reaction %s: more than one reaction in one block" % label)
    return out


# comp_ener_exch branches the generated files carry, whitespace removed
_ENER_BRANCHES = {
    (1, 2): "*enerExch-=(eexci*eV2J)*NA*(qf-qr)*sefact;",
    (3,): "*enerExch-=mee*NA*(qf-qr);",
    (4,): "*enerExch-=mee*NA*(qf-qr);",
    (5,): "amrex::Realqq=sc[E_ID]*sc[elidx]*kf;*enerExch-=3.0*kB*(wt[E_ID]/wt[elidx])*(Te-T)*qq*NA;",
    (7,): "*enerExch+=mee*NA*(qf-qr);",
}
_ENER_PRELUDE = ["amrex::RealeV2J=1.602176620800000e-19;", "amrex::RealNA=6.02214085774e23;",
                 "amrex::RealkB=1.380649e-23;", "amrex::Realne=sc[E_ID]*NA;", "amrex::Realmee=Ue/ne;",
                 "amrex::Realsefact=(tanh((Te-T)/0.5)+1.0)/2.0;"]


def _ener_types(h_text):
    """The rxntypes whose energy exchange this file's comp_ener_exch takes
    (commented-out branches excluded), each checked against its formula."""
    m = re.search(r"void\s+comp_ener_exch\s*\(", h_text)
    if not m:
        return []
    body = re.sub(r"\s+", "", _strip_comments(_function_body(h_text, re.escape(m.group(0)))))
    for line in _ENER_PRELUDE:
        if line not in body:
            raise ValueError("comp_ener_exch: %r not found" % line)
    out = []
    for cond, stmt in re.findall(r"if\(([^)]*)\)\{([^}]*)\}", body):
        types = tuple(int(x) for x in re.findall(r"rxntype==(\d+)", cond))
        if _ENER_BRANCHES.get(types) != stmt:
            raise ValueError("comp_ener_exch branch %s not understood" % cond)
        out.extend(types)
    return sorted(out)


def _check_prelude(h_text, reactions):
    """The definitions the rate forms rely on, as productionRate writes them."""
    body = re.sub(r"\s+", "", _strip_comments(_function_body(h_text, r"productionRate\s*\(\s*amrex::Real\s*\*\s*wdot")))
    need = ["constamrex::RealinvT=1.0/tc[1];", "constamrex::RealrefC=101325/8.31446*invT;",
            "constamrex::RealrefCinv=1/refC;"]
    kinds = {r["kind"] for r in reactions}
    if any(r["kind"] == "arrh" and (r["beta"] != 0.0) for r in reactions):
        need += ["constamrex::ReallogT=log(T/300.0);", "tc[0]=log(T/300.0);"]
    if kinds & {"janev", "ffit", "janev_ffit"} or any(r["tref"] == 11595.0 for r in reactions):
        need += ["doubleTeeV=Te/11595.0;", "doublelogTe=log(TeeV);", "doubleinvTeeV=(Te==0)?1.0:1.0/(TeeV);",
                 "doubleTe_pow[]={pow(logTe,0),pow(logTe,1),pow(logTe,2),pow(logTe,3),pow(logTe,4),pow(logTe,5),"
                 "pow(logTe,6),pow(logTe,7),pow(logTe,8)};",
                 "doubleinvTe_pow[]={pow(invTeeV,1),pow(invTeeV,2),pow(invTeeV,3),pow(invTeeV,4)};"]
    need += ["amrex::Realne=sc[E_ID]*6.02214085774e23;", "amrex::RealUe=1.5*Te*ne*1.380649e-23;"]
    for line in need:
        if line not in body:
            raise ValueError("productionRate: %r not found" % line)


def parse(h_text, cpp_text=None, source=""):
    names, ids, charge, molwt = _species(h_text, cpp_text)
    ns = len(names)
    reactions = [_parse_block(label, eq, seg, ids, ns) for label, eq, seg in _production_blocks(h_text)]
    _check_prelude(h_text, reactions)
    ener = _ener_types(h_text)
    rev_species = sorted({k for r in reactions for k, _ in r["g"]})
    gibbs = _gibbs(h_text, ns) if rev_species else [{"tmid": 0.0, "lo": [0.0] * 7, "hi": [0.0] * 7}] * ns
    for k in rev_species:
        if gibbs[k].get("missing"):
            raise ValueError("species %s has no g/RT but takes part in a reverse rate" % names[k])
    if "E" not in ids:
        raise ValueError("no electron (E_ID)")
    return {"source": source, "species": names, "charge": charge, "molwt": molwt, "e_idx": ids["E"],
            "energy_exchange_types": ener,
            "gibbs": [{"tmid": g["tmid"], "lo": g["lo"], "hi": g["hi"]} for g in gibbs],
            "reactions": reactions}


_KINDS = {"arrh": 0, "tepow": 1, "tearrh": 2, "janev": 3, "ffit": 4, "janev_ffit": 5}


def to_desc(mech, desc=None):
    """Fill a vd_mech_desc (VD_CHEM_GEN) from a parsed or loaded mechanism.
    Transport fields are left as they are (zero mobilities by default)."""
    d = desc if desc is not None else abi.vd_mech_desc()
    ns = len(mech["species"])
    if ns > abi.VD_MAX_SPECIES or len(mech["reactions"]) > abi.VD_MAX_REACTIONS:
        raise ValueError("mechanism of %d species / %d reactions beyond the library's limits"
                         % (ns, len(mech["reactions"])))
    d.nspec, d.nreac, d.e_idx = ns, len(mech["reactions"]), mech["e_idx"]
    d.chem_form = abi.VD_CHEM_GEN
    for k, name in enumerate(mech["species"]):
        d.names[k].value = name.encode()[:15]
        d.charge[k] = mech["charge"][k]
        d.molwt[k] = mech["molwt"][k]
        g = mech["gibbs"][k]
        d.gibbs_tmid[k] = g["tmid"]
        for q in range(7):
            d.gibbs_lo[k][q] = g["lo"][q]
            d.gibbs_hi[k][q] = g["hi"][q]
    for i, r in enumerate(mech["reactions"]):
        R = d.reac[i]
        R.rate_kind = _KINDS[r["kind"]]
        R.A, R.beta, R.Ea, R.tref, R.tgas_div = r["A"], r["beta"], r["Ea"], r["tref"], r["tgas_div"]
        for q in range(9):
            R.fit[q] = r["fit"][q]
        R.A2 = r["A2"]
        for q in range(4):
            R.fit2[q] = r["fit2"][q]
        R.fit_switch_ev, R.kmax = r["fit_switch_ev"], r["kmax"]
        R.third_body, R.reversible, R.refc_pow = r["third_body"], r["reversible"], r["refc_pow"]
        if len(r["react"]) > abi.VD_MAX_REACTANTS or len(r["rev"]) > abi.VD_MAX_PRODUCTS or \
                len(r["g"]) > abi.VD_GEN_NG:
            raise ValueError("reaction %s beyond the library's per-reaction limits" % r["label"])
        R.nreact = len(r["react"])
        for q, (s, nu) in enumerate(r["react"]):
            R.react_sp[q], R.react_nu[q] = s, nu
        # products = net change + reactants (a species on both sides keeps both)
        net = {}
        for s, nu in r["wdot"]:
            net[s] = net.get(s, 0.0) + nu
        prod = dict(net)
        for s, nu in r["react"]:
            prod[s] = prod.get(s, 0.0) + nu
        prod = [(s, nu) for s, nu in sorted(prod.items()) if nu != 0.0]
        if any(nu < 0 for _, nu in prod) or len(prod) > abi.VD_MAX_PRODUCTS:
            raise ValueError("reaction %s: products not representable" % r["label"])
        R.nprod = len(prod)
        for q, (s, nu) in enumerate(prod):
            R.prod_sp[q], R.prod_nu[q] = s, nu
        R.ng = len(r["g"])
        for q, (s, c) in enumerate(r["g"]):
            R.g_sp[q], R.g_nu[q] = s, c
        R.nrev = len(r["rev"])
        for q, (s, nu) in enumerate(r["rev"]):
            R.rev_sp[q], R.rev_nu[q] = s, nu
        # a type this file's comp_ener_exch does not take exchanges nothing
        rt = r["rxntype"] if r["rxntype"] in mech["energy_exchange_types"] else 0
        R.rxntype, R.eexci, R.elidx = rt, r["eexci"], r["elidx"]
        if len(r["wdot"]) > abi.VD_GEN_NW:
            raise ValueError("reaction %s: more wdot updates than the library holds" % r["label"])
        R.nwd = len(r["wdot"])
        for q, (s, nu) in enumerate(r["wdot"]):
            R.wd_sp[q], R.wd_nu[q] = s, nu
        R.temp_is_te = 1 if r["kind"] != "arrh" else 0
    return d


def dump(mech, path):
    """JSON with one species field and one reaction per line."""
    lines = ["{"]
    keys = [k for k in mech if k not in ("gibbs", "reactions")]
    for k in keys:
        lines.append("%s: %s," % (json.dumps(k), json.dumps(mech[k])))
    lines.append('"gibbs": [')
    lines.append(",\n".join(json.dumps(g) for g in mech["gibbs"]))
    lines.append("],")
    lines.append('"reactions": [')
    lines.append(",\n".join(json.dumps(r) for r in mech["reactions"]))
    lines.append("]}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def builtin(name):
    """Path of a mechanism converted from a reference case and shipped with
    the package (co2_h2_ar_reduced, he_rf_1d)."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "mechanisms", name + ".json")


def load(path):
    with open(path) as f:
        return json.load(f)


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("chemistry_h")
    ap.add_argument("chemistry_cpp", nargs="?")
    ap.add_argument("-o", "--out", required=True)
    a = ap.parse_args(argv)
    h = open(a.chemistry_h).read()
    cpp = open(a.chemistry_cpp).read() if a.chemistry_cpp else None
    src = a.chemistry_h
    src = src[src.index("test/"):] if "test/" in src else src.split("/")[-1]
    mech = parse(h, cpp, source=src)
    dump(mech, a.out)
    print("%d species, %d reactions -> %s" % (len(mech["species"]), len(mech["reactions"]), a.out))


if __name__ == "__main__":
    main(sys.argv[1:])
