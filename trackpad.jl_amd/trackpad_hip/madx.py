"""MAD-X lattice files: read a sequence or line into a :class:`Lattice`, write one back.

The reference has no lattice reader: its migration plan lists a MAD-X parser as future work
(MIGRATION_PLAN.md:87-97), and lattices are built in Julia from element constructors
(test/verify_dba.jl:20-60).  This module reads the subset of MAD-X that describes such
lattices and hands the result to :func:`compile_table`:

* statements end with ``;``; comments are ``//`` or ``!`` to the end of the line and
  ``/* ... */``; names and keywords are case-insensitive;
* variables: ``name = expr;`` (evaluated at once) and ``name := expr;`` (deferred: evaluated
  when the lattice is built, with the values then current); ``const`` / ``real`` prefixes;
  expressions with ``+ - * / ^``, parentheses, ``elem->attr`` references, the constants
  ``pi, twopi, degrad, raddeg, e, clight, emass, pmass`` and the functions ``sqrt, exp, log,
  log10, sin, cos, tan, asin, acos, atan, sinh, cosh, tanh, abs, floor, ceil, round``;
* element classes DRIFT, QUADRUPOLE, SEXTUPOLE, OCTUPOLE, SBEND, RBEND, MULTIPOLE, RFCAVITY,
  MARKER (MONITOR, HMONITOR, VMONITOR, INSTRUMENT, PLACEHOLDER as markers or drifts),
  HKICKER, VKICKER, KICKER, SOLENOID, with the attributes ``l, k1, k2, k3, angle, e1, e2,
  fint, fintx, hgap, knl, ksl, volt, freq, harmon, lag, kick, hkick, vkick, ks`` and PTC's
  ``nst`` (integration steps); ``label: parent, ...`` derives from an element or class;
  ``elem->attr = expr`` changes an element;
* ``label: LINE = (a, b, 2*c, -d, (e, f));`` and ``label: SEQUENCE, l = L, refer = entry |
  centre | exit; elem, at = s [, from = other]; ...; ENDSEQUENCE;`` (gaps become drifts);
* ``BEAM, particle = electron | positron | proton, energy = E | pc = P | gamma = G [, mass =
  M];`` (GeV) and ``USE, sequence | period = name;``; ``CALL, file = "...";`` includes a file;
  other commands (``option``, ``title``, ``twiss`` ...) are ignored.

Conversions (MAD-X -> the reference's element types):
* thick quadrupole / sextupole / octupole -> Quadrupole(l, k1) / Sextupole(l, k2) /
  Octupole(l, k3) with num_int_steps = nst (default 10, the reference's default);
* SBEND -> ``bend`` (SBend by default, or ExactSBend / ExactSBendLinearEdge) with fint1 =
  fint, fint2 = fintx (default fint), gap = 2 hgap, k1 / k2 into polynom_b; RBEND -> the same
  with e1, e2 increased by angle/2 and the length taken as the chord: arc = l (angle/2) /
  sin(angle/2); for ExactSBend the Forest fringe parameter is gk = fint hgap;
* MULTIPOLE knl, ksl -> ThinMultipole with polynom_b[n] = knl[n]/n!, polynom_a[n] = ksl[n]/n!
  (the reference's kick, src/tracking.jl:284-305, with L = 1 for integrated strengths);
* RFCAVITY volt [MV], freq [MHz] (or harmon: f = harmon beta c / C of the used line), lag
  [2 pi] -> RFCavity(l, volt, freq, h = harmon, philag = 2 pi lag, energy = the beam's
  energy): the reference's cavity kicks only when ``energy > 0`` (src/tracking.jl:508);
* HKICKER / VKICKER / KICKER -> Corrector(l, hkick, vkick); SOLENOID -> Solenoid(l, ks);
* BEAM -> Beam(E - m): the reference's Beam takes the kinetic energy in eV
  (src/lattice.jl:50-54), MAD-X the total energy in GeV.
Skew thick multipoles (k1s, k2s, k3s, ks on a bend) and classes outside this list are refused
with :class:`MadxError`, so a lattice is never silently changed.

:func:`write_madx` writes a lattice of these element types as MAD-X (a LINE, every value
printed so that it reads back to the same double), so ``read_madx(write_madx(lat, beam),
energy=beam.energy)`` compiles to the same table; the one exception is a thin-multipole
coefficient that no knl[n] reproduces after the division by n! (then it is one ulp away).
"""

from __future__ import annotations

import math
import os
import re
from typing import Dict, List, Optional, Tuple, Union

import numpy as np

from .elements import (C_LIGHT, Corrector, Drift, ExactSBend, ExactSBendLinearEdge, Marker, Octupole, Quadrupole,
                       RFCavity, SBend, Sextupole, Solenoid, ThinMultipole, get_length)
from .lattice import M_ELECTRON, M_PROTON, Beam, Lattice, total_length

__all__ = ["read_madx", "write_madx", "MadxError"]


class MadxError(ValueError):
    """A MAD-X input this reader does not accept (syntax, unknown name, unsupported class)."""


_CONST = {"pi": math.pi, "twopi": 2.0 * math.pi, "degrad": 180.0 / math.pi, "raddeg": math.pi / 180.0,
          "e": math.e, "clight": C_LIGHT, "emass": M_ELECTRON * 1e-9, "pmass": M_PROTON * 1e-9}
_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "log10": math.log10, "sin": math.sin,
          "cos": math.cos, "tan": math.tan, "asin": math.asin, "acos": math.acos, "atan": math.atan,
          "sinh": math.sinh, "cosh": math.cosh, "tanh": math.tanh, "abs": abs, "floor": math.floor,
          "ceil": math.ceil, "round": round}
_CLASSES = {"drift", "quadrupole", "sextupole", "octupole", "sbend", "rbend", "multipole", "rfcavity", "marker",
            "monitor", "hmonitor", "vmonitor", "instrument", "placeholder", "hkicker", "vkicker", "kicker",
            "solenoid"}
_IGNORED = {"option", "title", "set", "show", "value", "twiss", "select", "print", "system", "assign", "exit",
            "stop", "return", "survey", "emit", "ptc_create_universe", "ptc_end", "esave", "save", "seqedit",
            "flatten", "endedit", "help", "echo", "removefile", "resbeam", "savebeta", "create", "write", "fill",
            "setvars", "setvars_lin", "readtable", "match", "endmatch", "vary", "constraint", "lmdif", "simplex"}


_RESERVED = (_CLASSES | _IGNORED | set(_CONST) | set(_FUNCS) |
             {"line", "sequence", "endsequence", "beam", "use", "call", "const", "real", "int"})

# --------------------------------------------------------------------------------------
# Expressions
# --------------------------------------------------------------------------------------
_TOK = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|([A-Za-z_][\w.]*(?:->[A-Za-z_]\w*)?)"
                  r"|(\*\*|[-+*/^(),]))")


def _tokens(src: str):
    pos, out = 0, []
    src = src.strip()
    while pos < len(src):
        m = _TOK.match(src, pos)
        if not m or m.end() == pos:
            raise MadxError(f"cannot read expression {src!r} at {src[pos:]!r}")
        num, name, op = m.groups()
        out.append(("n", float(num)) if num is not None else ("v", name.lower()) if name is not None else ("o", op))
        pos = m.end()
    return out


class _Expr:
    """Recursive-descent evaluation of a MAD-X arithmetic expression (no Python eval)."""

    def __init__(self, src: str, lookup):
        self.t = _tokens(src)
        self.i = 0
        self.lookup = lookup
        self.src = src

    def _peek(self):
        return self.t[self.i] if self.i < len(self.t) else (None, None)

    def _next(self):
        tok = self._peek()
        self.i += 1
        return tok

    def value(self) -> float:
        v = self._sum()
        if self.i != len(self.t):
            raise MadxError(f"unexpected {self._peek()[1]!r} in {self.src!r}")
        return v

    def _sum(self):
        v = self._prod()
        while self._peek() in (("o", "+"), ("o", "-")):
            op = self._next()[1]
            w = self._prod()
            v = v + w if op == "+" else v - w
        return v

    def _prod(self):
        v = self._unary()
        while self._peek() in (("o", "*"), ("o", "/")):
            op = self._next()[1]
            w = self._unary()
            v = v * w if op == "*" else v / w
        return v

    def _unary(self):
        if self._peek() in (("o", "-"), ("o", "+")):
            op = self._next()[1]
            v = self._unary()
            return -v if op == "-" else v
        return self._power()

    def _power(self):
        v = self._atom()
        if self._peek() in (("o", "^"), ("o", "**")):
            self._next()
            v = v ** self._unary()  # right-associative
        return v

    def _atom(self):
        kind, val = self._next()
        if kind == "n":
            return val
        if kind == "o" and val == "(":
            v = self._sum()
            if self._next() != ("o", ")"):
                raise MadxError(f"missing ')' in {self.src!r}")
            return v
        if kind == "v":
            if self._peek() == ("o", "("):
                if val not in _FUNCS:
                    raise MadxError(f"unknown function {val!r} in {self.src!r}")
                self._next()
                arg = self._sum()
                if self._next() != ("o", ")"):
                    raise MadxError(f"missing ')' in {self.src!r}")
                return float(_FUNCS[val](arg))
            return float(self.lookup(val))
        raise MadxError(f"cannot read expression {self.src!r}")


# --------------------------------------------------------------------------------------
# Statements
# --------------------------------------------------------------------------------------
def _strip_comments(text: str) -> str:
    text = re.sub(r"/\*.*?\*/", " ", text, flags=re.S)
    out = []
    for line in text.splitlines():
        # a comment starts at // or ! outside a quoted string
        q, cut = None, len(line)
        for i, ch in enumerate(line):
            if q:
                if ch == q:
                    q = None
            elif ch in "\"'":
                q = ch
            elif ch == "!" or line.startswith("//", i):
                cut = i
                break
        out.append(line[:cut])
    return "\n".join(out)


def _split_top(s: str, sep: str = ",") -> List[str]:
    """Split at sep outside (), {} and quotes."""
    parts, depth, q, cur = [], 0, None, []
    for ch in s:
        if q:
            cur.append(ch)
            if ch == q:
                q = None
            continue
        if ch in "\"'":
            q = ch
        elif ch in "({":
            depth += 1
        elif ch in ")}":
            depth -= 1
        if ch == sep and depth == 0:
            parts.append("".join(cur))
            cur = []
        else:
            cur.append(ch)
    parts.append("".join(cur))
    return [p.strip() for p in parts]


def _attrs(items: List[str]) -> Dict[str, Tuple[str, bool]]:
    """``name = value`` / ``name := value`` items -> {name: (value text, deferred)}; a bare
    name is a flag set to 1."""
    out = {}
    for it in items:
        if not it:
            continue
        m = re.match(r"^([A-Za-z_][\w.]*)\s*(:=|=)\s*(.*)$", it, flags=re.S)
        if m:
            out[m.group(1).lower()] = (m.group(3).strip(), m.group(2) == ":=")
        elif re.match(r"^-?[A-Za-z_][\w.]*$", it):
            neg = it.startswith("-")
            out[it.lstrip("-").lower()] = ("0" if neg else "1", False)
        else:
            raise MadxError(f"cannot read attribute {it!r}")
    return out


class _Model:
    def __init__(self):
        self.vars: Dict[str, Tuple[str, bool, Optional[float]]] = {}
        self.elems: Dict[str, dict] = {}     # label -> {"class": base class, "attrs": {...}}
        self.lines: Dict[str, str] = {}      # label -> body text of a LINE
        self.seqs: Dict[str, dict] = {}      # label -> {"l", "refer", "items": [(label, attrs)]}
        self.beam: Dict[str, Tuple[str, bool]] = {}
        self.use: Optional[str] = None
        self.directives: Dict[str, str] = {}
        self._stack: List[str] = []

    # -- evaluation ---------------------------------------------------------------------
    def lookup(self, name: str) -> float:
        if "->" in name:
            el, at = name.split("->")
            return self.attr_value(el, at)
        if name in self.vars:
            text, deferred, val = self.vars[name]
            if val is not None and not deferred:
                return val
            return self._eval_guarded(text, name)
        if name in _CONST:
            return _CONST[name]
        raise MadxError(f"undefined name {name!r}")

    def _eval_guarded(self, text: str, what: str) -> float:
        if what in self._stack:
            raise MadxError(f"circular definition of {what!r}")
        self._stack.append(what)
        try:
            return _Expr(text, self.lookup).value()
        finally:
            self._stack.pop()

    def eval(self, text: str) -> float:
        return _Expr(text, self.lookup).value()

    def resolved(self, label: str) -> Tuple[str, Dict[str, Tuple[str, bool]]]:
        """(base class, merged attributes) of an element, following its parents."""
        chain, lab = [], label
        seen = set()
        while lab not in _CLASSES:
            if lab not in self.elems:
                raise MadxError(f"undefined element or class {lab!r}")
            if lab in seen:
                raise MadxError(f"circular element definition {lab!r}")
            seen.add(lab)
            chain.append(self.elems[lab]["attrs"])
            lab = self.elems[lab]["parent"]
        merged: Dict[str, Tuple[str, bool]] = {}
        for a in reversed(chain):
            merged.update(a)
        return lab, merged

    def attr_value(self, label: str, attr: str, default: Optional[float] = None) -> float:
        _, at = self.resolved(label)
        if attr not in at:
            if default is None:
                if attr == "l":
                    return 0.0
                raise MadxError(f"{label} has no attribute {attr!r}")
            return default
        return self._eval_guarded(at[attr][0], f"{label}->{attr}")

    def attr_list(self, label: str, attr: str) -> List[float]:
        _, at = self.resolved(label)
        if attr not in at:
            return []
        text = at[attr][0].strip()
        if not (text.startswith("{") and text.endswith("}")):
            return [self.eval(text)]
        body = text[1:-1].strip()
        return [self.eval(x) for x in _split_top(body)] if body else []

    # -- statements ---------------------------------------------------------------------
    def statement(self, st: str, base_dir: str, seq: Optional[str]) -> Optional[str]:
        s = st.strip()
        if not s:
            return seq
        low = s.lower()
        if seq is not None:
            if low == "endsequence":
                return None
            self._sequence_item(s, seq)
            return seq
        m = re.match(r"^(?:(const|real|int)\s+)?([A-Za-z_][\w.]*)\s*(:=|=)\s*(.+)$", s, flags=re.S)
        if m and not re.match(r"^[A-Za-z_][\w.]*\s*:\s*line\b", s, flags=re.I):
            name, op, text = m.group(2).lower(), m.group(3), m.group(4).strip()
            deferred = op == ":="
            self.vars[name] = (text, deferred, None if deferred else self.eval(text))
            return None
        m = re.match(r"^([A-Za-z_][\w.]*)->([A-Za-z_]\w*)\s*(:=|=)\s*(.+)$", s, flags=re.S)
        if m:
            el, at = m.group(1).lower(), m.group(2).lower()
            if el not in self.elems:
                raise MadxError(f"undefined element {el!r}")
            self.elems[el]["attrs"][at] = (m.group(4).strip(), m.group(3) == ":=")
            return None
        m = re.match(r"^([A-Za-z_][\w.]*)\s*:\s*(.+)$", s, flags=re.S)
        if m:
            return self._labelled(m.group(1).lower(), m.group(2).strip())
        head, *rest = _split_top(s)
        cmd = head.lower()
        if cmd == "beam":
            self.beam.update(_attrs(rest))
        elif cmd == "use":
            at = _attrs(rest)
            key = "sequence" if "sequence" in at else "period" if "period" in at else None
            if key is None:
                raise MadxError("USE needs sequence= or period=")
            self.use = at[key][0].lower()
        elif cmd == "call":
            at = _attrs(rest)
            if "file" not in at:
                raise MadxError("CALL needs file=")
            path = at["file"][0].strip("\"'")
            self.read_text(_read_file(os.path.join(base_dir, path)), os.path.dirname(os.path.join(base_dir, path)))
        elif cmd in self.elems and rest:
            self.elems[cmd]["attrs"].update(_attrs(rest))  # "elem, attr = value": change the element
        elif cmd in _IGNORED or cmd == "endsequence":
            pass
        else:
            raise MadxError(f"cannot read statement {s[:60]!r}")
        return None

    def _labelled(self, label: str, body: str) -> Optional[str]:
        m = re.match(r"^line\s*=\s*\((.*)\)$", body, flags=re.S | re.I)
        if m:
            self.lines[label] = m.group(1)
            return None
        if label in _RESERVED:
            raise MadxError(f"{label!r} is a keyword, not a label")
        head, *rest = _split_top(body)
        cls = head.lower()
        if cls == "sequence":
            at = _attrs(rest)
            refer = at.get("refer", ("centre", False))[0].lower()
            if refer == "center":
                refer = "centre"
            if refer not in ("entry", "centre", "exit"):
                raise MadxError(f"sequence {label}: refer = {refer!r}")
            self.seqs[label] = {"l": at.get("l"), "refer": refer, "items": []}
            return label
        self.elems[label] = {"parent": cls, "attrs": _attrs(rest)}
        self.resolved(label)  # the parent must exist
        return None

    def _sequence_item(self, s: str, seq: str) -> None:
        m = re.match(r"^([A-Za-z_][\w.]*)\s*:\s*(.+)$", s, flags=re.S)
        if m:  # an element defined in place
            label = m.group(1).lower()
            head, *rest = _split_top(m.group(2))
            at = _attrs(rest)
            place = {k: at.pop(k) for k in ("at", "from") if k in at}
            self.elems[label] = {"parent": head.lower(), "attrs": at}
            self.resolved(label)
        else:
            head, *rest = _split_top(s)
            label = head.lower()
            place = _attrs(rest)
            self.resolved(label)
        if "at" not in place:
            raise MadxError(f"sequence {seq}: {label} has no at=")
        self.seqs[seq]["items"].append((label, place))

    def read_text(self, text: str, base_dir: str) -> None:
        for m in re.finditer(r"trackpad:\s*(\w+)\s*=\s*(\w+)", text):
            self.directives[m.group(1).lower()] = m.group(2)
        seq = None
        for st in _split_top(_strip_comments(text), ";"):
            seq = self.statement(st, base_dir, seq)
        if seq is not None:
            raise MadxError(f"sequence {seq} has no ENDSEQUENCE")

    # -- flattening ---------------------------------------------------------------------
    def expand_line(self, label: str, depth: int = 0) -> List[str]:
        if depth > 100:
            raise MadxError("lines nested too deeply (recursive line?)")
        out: List[str] = []
        for item in _split_top(self.lines[label]):
            if not item:
                continue
            rep, inv = 1, False
            m = re.match(r"^(\d+)\s*\*\s*(.+)$", item, flags=re.S)
            if m:
                rep, item = int(m.group(1)), m.group(2).strip()
            if item.startswith("-"):
                inv, item = True, item[1:].strip()
            if item.startswith("(") and item.endswith(")"):
                self.lines["__anon%d_%d" % (depth, len(self.lines))] = item[1:-1]
                sub = self.expand_line("__anon%d_%d" % (depth, len(self.lines) - 1), depth + 1)
            elif item.lower() in self.lines:
                sub = self.expand_line(item.lower(), depth + 1)
            else:
                self.resolved(item.lower())
                sub = [item.lower()]
            if inv:
                sub = sub[::-1]
            out += sub * rep
        return out

    def flatten_sequence(self, label: str) -> List[Union[str, float]]:
        """Labels of a sequence with drift lengths (floats) for the gaps between them."""
        sq = self.seqs[label]
        pos = {}
        placed = []
        for lab, place in sq["items"]:
            L = self.attr_value(lab, "l", 0.0)
            cls, _ = self.resolved(lab)
            if cls == "rbend":  # the chord is given; the element occupies its arc length
                ang = self.attr_value(lab, "angle", 0.0)
                L = L * (ang / 2) / math.sin(ang / 2) if ang != 0.0 else L
            at = self.eval(place["at"][0])
            if "from" in place:
                ref = place["from"][0].lower()
                if ref not in pos:
                    raise MadxError(f"sequence {label}: from={ref} names no earlier element")
                at += pos[ref]
            centre = {"entry": at + L / 2, "centre": at, "exit": at - L / 2}[sq["refer"]]
            pos[lab] = centre
            placed.append((centre - L / 2, L, lab))
        out: List[Union[str, float]] = []
        s = 0.0
        for start, L, lab in placed:
            gap = start - s
            if gap < -1e-9:
                raise MadxError(f"sequence {label}: {lab} overlaps the element before it by {-gap:g} m")
            if gap > 1e-12:
                out.append(gap)
            out.append(lab)
            s = start + L
        total = self.eval(sq["l"][0]) if sq["l"] else s
        if total - s > 1e-12:
            out.append(total - s)
        return out


def _read_file(path: str) -> str:
    with open(path) as f:
        return f.read()


# --------------------------------------------------------------------------------------
# Conversion to the reference's element types
# --------------------------------------------------------------------------------------
_BENDS = {"SBend": SBend, "ExactSBend": ExactSBend, "ExactSBendLinearEdge": ExactSBendLinearEdge}


def _beam(model: _Model, energy: Optional[float]) -> Beam:
    b = model.beam
    part = b.get("particle", ("electron", False))[0].strip("\"'").lower()
    if "mass" in b:
        mass = model.eval(b["mass"][0]) * 1e9
    elif part in ("electron", "positron"):
        mass = M_ELECTRON
    elif part in ("proton", "antiproton"):
        mass = M_PROTON
    else:
        raise MadxError(f"BEAM particle {part!r}: give mass=")
    if energy is not None:
        return Beam(float(energy), mass=mass)
    if "energy" in b:
        etot = model.eval(b["energy"][0]) * 1e9
    elif "pc" in b:
        pc = model.eval(b["pc"][0]) * 1e9
        etot = math.sqrt(pc * pc + mass * mass)
    elif "gamma" in b:
        etot = model.eval(b["gamma"][0]) * mass
    else:
        raise MadxError("BEAM needs energy=, pc= or gamma= (or pass energy= to read_madx)")
    if etot <= mass:
        raise MadxError(f"BEAM total energy {etot:g} eV is not above the rest mass")
    return Beam(etot - mass, mass=mass)


def _element(model: _Model, lab: str, bend_cls, beam: Beam, circumference: float):
    cls, at = model.resolved(lab)
    name = lab.upper()
    g = lambda a, d=0.0: model.attr_value(lab, a, d)
    for skew in ("k1s", "k2s", "k3s", "tilt"):
        if skew in at and g(skew) != 0.0:
            raise MadxError(f"{name}: {skew} is not supported (the reference's magnets have no such field)")
    L = g("l")
    nst = int(round(g("nst", 10.0)))
    if cls == "drift":
        return Drift(L, name=name)
    if cls in ("marker", "monitor", "hmonitor", "vmonitor", "instrument", "placeholder"):
        return Drift(L, name=name) if L != 0.0 else Marker(name=name)
    if cls == "quadrupole":
        return Quadrupole(L, g("k1"), num_int_steps=nst, name=name)
    if cls == "sextupole":
        return Sextupole(L, g("k2"), num_int_steps=nst, name=name)
    if cls == "octupole":
        return Octupole(L, g("k3"), num_int_steps=nst, name=name)
    if cls in ("sbend", "rbend"):
        ang, e1, e2 = g("angle"), g("e1"), g("e2")
        if cls == "rbend":
            if ang != 0.0:
                L = L * (ang / 2) / math.sin(ang / 2)
            e1, e2 = e1 + ang / 2, e2 + ang / 2
        fint = g("fint")
        fintx = g("fintx", fint)
        hgap = g("hgap")
        pb = np.zeros(4)
        pb[1], pb[2] = g("k1"), g("k2") / 2
        mo = 2 if pb[2] != 0.0 else 1 if pb[1] != 0.0 else 0
        kw = dict(num_int_steps=nst, polynom_b=pb, max_order=mo, name=name)
        if bend_cls is ExactSBend:
            return ExactSBend(L, ang, e1, e2, gk=fint * hgap, **kw)
        return bend_cls(L, ang, e1, e2, fint1=fint, fint2=fintx, gap=2 * hgap, **kw)
    if cls == "multipole":
        knl, ksl = model.attr_list(lab, "knl"), model.attr_list(lab, "ksl")
        n = max(len(knl), len(ksl), 1)
        if n > 11:
            raise MadxError(f"{name}: multipoles above order 10 are not supported")
        pb, pa = np.zeros(max(n, 4)), np.zeros(max(n, 4))
        for i, v in enumerate(knl):
            pb[i] = v / math.factorial(i)
        for i, v in enumerate(ksl):
            pa[i] = v / math.factorial(i)
        return ThinMultipole(L, pa, pb, max_order=n - 1, name=name)
    if cls == "rfcavity":
        harmon = g("harmon", 0.0)
        if "freq" in at:
            freq = g("freq") * 1e6
        elif harmon > 0.0:
            freq = harmon * beam.beta * C_LIGHT / circumference
        else:
            raise MadxError(f"{name}: RFCAVITY needs freq= or harmon=")
        return RFCavity(L, g("volt") * 1e6, freq, 0.0, h=harmon if harmon > 0.0 else 1.0,
                        philag=2.0 * math.pi * g("lag"), energy=beam.energy, name=name)
    if cls in ("hkicker", "vkicker", "kicker"):
        hk = g("kick") if cls == "hkicker" else g("hkick")
        vk = g("kick") if cls == "vkicker" else g("vkick")
        return Corrector(L, hk, vk, name=name)
    if cls == "solenoid":
        return Solenoid(L, g("ks"), name=name)
    raise MadxError(f"{name}: element class {cls.upper()} is not supported")


def read_madx(source: str, *, use: Optional[str] = None, bend: Optional[str] = None,
              energy: Optional[float] = None) -> Tuple[Lattice, Beam]:
    """Read a MAD-X file (a path, or the text itself if it holds a ``;``) and return
    ``(Lattice, Beam)`` of the used sequence or line.  ``use`` overrides the file's USE;
    ``bend`` is "SBend" (default), "ExactSBend" or "ExactSBendLinearEdge" (a ``! trackpad:
    bend = ...`` comment in the file sets it too); ``energy`` (kinetic, eV) overrides BEAM."""
    if ";" in source or "\n" in source:
        text, base = source, os.getcwd()
    else:
        text, base = _read_file(source), os.path.dirname(os.path.abspath(source))
    model = _Model()
    model.read_text(text, base)
    name = (use or model.use or "").lower()
    if not name:
        cands = list(model.seqs) + list(model.lines)
        if len(cands) != 1:
            raise MadxError("no USE statement and not exactly one sequence or line: pass use=")
        name = cands[0]
    bend = bend or model.directives.get("bend", "SBend")
    if bend not in _BENDS:
        raise MadxError(f"bend model {bend!r} not in {sorted(_BENDS)}")
    beam = _beam(model, energy)
    if name in model.seqs:
        flat = model.flatten_sequence(name)
    elif name in model.lines:
        flat = model.expand_line(name)
    else:
        raise MadxError(f"no sequence or line named {name!r}")
    # the circumference, for cavities given by harmonic number
    circ = 0.0
    for it in flat:
        if isinstance(it, float):
            circ += it
        else:
            cls, _ = model.resolved(it)
            L = model.attr_value(it, "l", 0.0)
            if cls == "rbend":
                ang = model.attr_value(it, "angle", 0.0)
                L = L * (ang / 2) / math.sin(ang / 2) if ang != 0.0 else L
            circ += L
    els = []
    k = 0
    for it in flat:
        if isinstance(it, float):
            k += 1
            els.append(Drift(it, name=f"DRIFT_{k}"))
        else:
            els.append(_element(model, it, _BENDS[bend], beam, circ))
    return Lattice(els, name=name.upper()), beam


# --------------------------------------------------------------------------------------
# Writer
# --------------------------------------------------------------------------------------
def _num(v: float) -> str:
    return repr(float(v))


def _scaled(target: float, to_file, from_file) -> str:
    """A double v, printed, such that the reader's from_file(v) gives target bit for bit
    (unit changes such as MV -> V are not exact in floating point): the nearest candidates to
    to_file(target) are tried in turn.  Where no double maps onto target (a division by n!
    skips some doubles), the candidate that lands nearest, one ulp away."""
    target = float(target)
    v = float(to_file(target))
    best, err = v, abs(from_file(v) - target)
    lo = hi = v
    for _ in range(64):
        if err == 0.0:
            break
        lo, hi = np.nextafter(lo, -np.inf), np.nextafter(hi, np.inf)
        for c in (float(lo), float(hi)):
            e = abs(from_file(c) - target)
            if e < err:
                best, err = c, e
    return _num(best)


def _def(e) -> Tuple[str, str]:
    """(class, attribute text) of an element, every value printed exactly."""
    t = type(e).__name__
    nst = f", nst = {e.num_int_steps}" if hasattr(e, "num_int_steps") and e.num_int_steps != 10 else ""
    if t == "Drift":
        return "drift", f"l = {_num(e.L)}"
    if t == "Marker":
        return "marker", ""
    if t in ("Quadrupole", "Sextupole", "Octupole"):
        k = {"Quadrupole": "k1", "Sextupole": "k2", "Octupole": "k3"}[t]
        if np.any(e.polynom_a != 0) or np.any(e.polynom_b != 0) or e.max_order != {"k1": 1, "k2": 2, "k3": 3}[k]:
            raise MadxError(f"{e.name}: multipole errors on a thick magnet have no MAD-X attribute here")
        return t.lower(), f"l = {_num(e.L)}, {k} = {_num(getattr(e, k))}{nst}"
    if t in _BENDS:
        pb = e.polynom_b
        if np.any(e.polynom_a != 0) or np.any(pb[3:] != 0) or pb[0] != 0:
            raise MadxError(f"{e.name}: bend multipoles beyond k1, k2 cannot be written")
        if e.max_order != (2 if pb[2] != 0 else 1 if pb[1] != 0 else 0):
            raise MadxError(f"{e.name}: max_order {e.max_order} is not the one k1/k2 imply")
        if (e.fringe_bend_entrance, e.fringe_bend_exit) != (1, 1):
            raise MadxError(f"{e.name}: edge method {e.fringe_bend_entrance}/{e.fringe_bend_exit} has no MAD-X "
                            "attribute (only the default 1 is written)")
        s = f"l = {_num(e.L)}, angle = {_num(e.angle)}, e1 = {_num(e.e1)}, e2 = {_num(e.e2)}"
        if pb[1] != 0:
            s += f", k1 = {_num(pb[1])}"
        if pb[2] != 0:
            s += f", k2 = {_num(2 * pb[2])}"
        if t == "ExactSBend":
            if e.gk != 0.0:
                s += f", fint = {_num(e.gk)}, hgap = 1.0"
        elif e.gap != 0.0 or e.fint1 != 0.0 or e.fint2 != 0.0:
            s += f", fint = {_num(e.fint1)}, fintx = {_num(e.fint2)}, hgap = {_num(e.gap / 2)}"
        return "sbend", s + nst
    if t == "ThinMultipole":
        n = e.max_order + 1
        if np.any(e.polynom_a[n:] != 0) or np.any(e.polynom_b[n:] != 0):
            raise MadxError(f"{e.name}: coefficients beyond max_order cannot be written")

        def coef(v, i):
            f = float(math.factorial(i))
            return _scaled(v, lambda x: x * f, lambda x: x / f)

        at = lambda v, i: float(v[i]) if i < len(v) else 0.0
        knl = ", ".join(coef(at(e.polynom_b, i), i) for i in range(n))
        ksl = ", ".join(coef(at(e.polynom_a, i), i) for i in range(n))
        return "multipole", f"l = {_num(e.L)}, knl = {{{knl}}}, ksl = {{{ksl}}}"
    if t == "RFCavity":
        if e.lag != 0.0:
            raise MadxError(f"{e.name}: a lag in metres has no MAD-X attribute (MAD-X's lag is a phase)")
        mega = (lambda x: x / 1e6, lambda x: x * 1e6)
        phase = (lambda x: x / (2 * math.pi), lambda x: 2.0 * math.pi * x)
        return "rfcavity", (f"l = {_num(e.L)}, volt = {_scaled(e.volt, *mega)}, freq = {_scaled(e.freq, *mega)}, "
                            f"harmon = {_num(e.h)}, lag = {_scaled(e.philag, *phase)}")
    if t == "Corrector":
        return "kicker", f"l = {_num(e.L)}, hkick = {_num(e.xkick)}, vkick = {_num(e.ykick)}"
    if t == "Solenoid":
        return "solenoid", f"l = {_num(e.L)}, ks = {_num(e.ks)}"
    raise MadxError(f"{e.name}: {t} has no MAD-X counterpart here")


def write_madx(lat: Lattice, beam: Beam, path: Optional[str] = None, *, name: Optional[str] = None) -> str:
    """MAD-X text of ``lat`` as a LINE (identical elements share one definition) with its
    BEAM; written to ``path`` if given.  Reading it back with ``read_madx(text,
    energy=beam.energy)`` gives the same elements, bit for bit (the bend model is recorded in
    a ``! trackpad: bend = ...`` comment)."""
    bends = {type(e).__name__ for e in lat.elements if type(e).__name__ in _BENDS}
    if len(bends) > 1:
        raise MadxError(f"one bend model per file: {sorted(bends)}")
    line = (name or lat.name or "RING").upper()
    defs: Dict[Tuple[str, str], str] = {}
    labels: List[str] = []
    used = set()
    out = [f"! written by trackpad_hip.write_madx: {len(lat)} elements",
           f"! trackpad: bend = {bends.pop() if bends else 'SBend'}"]
    part = "electron" if beam.mass == M_ELECTRON else "proton" if beam.mass == M_PROTON else None
    etot = (beam.energy + beam.mass) / 1e9
    bl = f"beam, particle = {part}" if part else f"beam, particle = ion, mass = {_num(beam.mass / 1e9)}"
    out.append(f"{bl}, energy = {_num(etot)};")
    for e in lat.elements:
        key = _def(e)
        if key not in defs:
            base = re.sub(r"\W", "_", e.name.upper()) or "E"
            if not re.match(r"[A-Z_]", base):
                base = "E_" + base
            lab, k = base, 1
            while lab in used or lab.lower() in _RESERVED:
                k += 1
                lab = f"{base}_{k}"
            used.add(lab)
            defs[key] = lab
            out.append(f"{lab}: {key[0]}" + (f", {key[1]}" if key[1] else "") + ";")
        labels.append(defs[key])
    body = []
    for i in range(0, len(labels), 8):
        body.append(", ".join(labels[i:i + 8]))
    out.append(f"{line}: line = (" + ",\n    ".join(body) + ");")
    out.append(f"use, period = {line};")
    text = "\n".join(out) + "\n"
    if path is not None:
        with open(path, "w") as f:
            f.write(text)
    return text
