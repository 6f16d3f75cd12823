"""Readers and writers for DGM's on-disk input and restart formats.

These mirror the reference's file handling so that a DGM run directory
(``root.inp``, ``root.msh``, ``root.con``, ``root.ic``) can drive this engine
and its restart files can be compared with ``dgm_diff`` semantics:

* ``root.inp``  ``key = value  # comment`` parameter table
  (src/Domain.cpp:236-277, src/Table.hpp)
* ``root.msh``  ``**MESH DATA**`` / ``ne nsd`` / per element ``ELEMENT id TYPE``
  followed by one line of vertex coordinates per space dimension
  (src/Field.cpp:2600-2700, src/Mesh.cpp)
* ``root.con``  ``**CONNECTIVITY DATA**`` / ``ne`` / rows ``E eid sid leid lsid``
  (src/Field.cpp:3478-3556)
* ``root.ic``   initial-condition code and parameters (src/Euler.cpp:880-1190)
* restart files: old 10-line text header or a JSON header, then per field
  ``glen, nModes[ne], disp[ne], modal FP64 coefficients``
  (src/Domain.cpp:1473-1639, src/vField.cpp:362-445, 1082-1201)
"""
from __future__ import annotations

import json
import re
import struct
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

_NVERT = {"QUAD": 4, "NODALQUAD": 4, "HEX": 8, "NODALHEX": 8}


def read_inp(path: str) -> Dict[str, object]:
    """Parse a ``root.inp`` parameter file into a dict of numbers/strings."""
    params: Dict[str, object] = {}
    with open(path) as f:
        for line in f:
            line = line.split("#", 1)[0].strip()
            if "=" not in line:
                continue
            k, v = (s.strip() for s in line.split("=", 1))
            try:
                params[k] = int(v)
            except ValueError:
                try:
                    params[k] = float(v)
                except ValueError:
                    params[k] = v
    return params


@dataclass
class Mesh:
    """Element vertices ``vert[e, dim, v]`` and ``.con`` rows ``(eid, sid, leid, lsid)``."""
    nsd: int
    vert: np.ndarray
    con: np.ndarray
    porder: Optional[np.ndarray] = None   # per-element order from root.ord

    @property
    def ne(self) -> int:
        return int(self.vert.shape[0])


def read_msh(path: str) -> np.ndarray:
    with open(path) as f:
        f.readline()
        hdr = f.readline().split()
        ne, nsd = int(hdr[0]), int(hdr[1])
        verts = []
        for _ in range(ne):
            toks = f.readline().split()
            etype = toks[2].upper()
            nv = _NVERT.get(etype)
            if nv is None:
                raise ValueError(f"{path}: unsupported element type {etype}")
            rows = [[float(x) for x in f.readline().split()] for _ in range(nsd)]
            verts.append(rows)
    return np.asarray(verts, dtype=np.float64)


def read_con(path: str) -> np.ndarray:
    rows: List[List[int]] = []
    with open(path) as f:
        f.readline()
        f.readline()
        for line in f:
            toks = line.split()
            if len(toks) >= 5 and toks[0] == "E":
                rows.append([int(t) for t in toks[1:5]])
    return np.asarray(rows, dtype=np.int32).reshape(-1, 4)


def read_mesh(root: str) -> Mesh:
    vert = read_msh(root + ".msh")
    con = read_con(root + ".con")
    return Mesh(nsd=vert.shape[1], vert=vert, con=con)


def read_ord(path: str, ne: int, p: int) -> np.ndarray:
    """Per-element polynomial order from ``root.ord`` (Field::Field,
    src/Field.cpp:2290-2320): a title line, then rows ``eid p q``; elements
    not listed keep the default order ``p``.  A nonzero ``q`` sets the element's
    quadrature points; only the default q = p + 2 is supported here."""
    orders = np.full(ne, int(p), dtype=np.int32)
    with open(path) as f:
        f.readline()
        for line in f:
            toks = line.split()
            if len(toks) < 3:
                continue
            e, pl, ql = int(toks[0]), int(toks[1]), int(toks[2])
            if e < 0 or e >= ne:
                raise ValueError(f"{path}: illegal element number {e}")
            if ql not in (0, pl + 2):
                raise ValueError(f"{path}: element {e} asks for q = {ql} quadrature points; "
                                 f"only the default q = p + 2 is supported")
            orders[e] = pl
    return orders


def read_bc(path: str) -> List[tuple]:
    """Boundary sides of ``root.bc``: rows (tag, eid, sid) after the count line
    (Field::read_bc, src/Field.cpp:4469-4495).  A side listed here is a
    boundary even where ``root.con`` also links it: Field::initialize_bc
    relinks it to its boundary side (src/Field.cpp:4426-4462)."""
    rows: List[tuple] = []
    with open(path) as f:
        f.readline()
        n = int(f.readline().split()[0])
        for _ in range(n):
            toks = f.readline().split()
            rows.append((toks[0], int(toks[1]), int(toks[2])))
    return rows


def read_crv(path: str):
    """Curved sides of ``root.crv`` in the old text format (Curve::read_old and
    Curve::read_sides, src/Curve.cpp:185-330): a header line, the number of
    curve types, each as ``code tag [parameters]`` (Arc: the radius), then the
    number of curved sides and rows ``eid sid tag``.  Returns (types, sides):
    types maps each tag to (code, parameters), sides lists (eid, sid, tag).  A
    later row for the same element replaces the earlier one, as the
    reference keeps one curve per element."""
    with open(path) as f:
        f.readline()
        toks = f.read().split()
    pos = 0
    ntypes = int(toks[pos]); pos += 1
    while not toks[pos][0].isalpha():
        pos += 1
    while toks[pos].lower() in ("number", "of", "curve", "type(s)", "types"):
        pos += 1
    types = {}
    for _ in range(ntypes):
        code = toks[pos].lower(); tag = toks[pos + 1]; pos += 2
        params = []
        if code in ("arc", "circle"):
            params = [float(toks[pos])]; pos += 1
        elif code not in ("straight", "nonaffine", "skewed", "affine"):
            raise ValueError(f"{path}: unsupported curve type {code}")
        types[tag] = (code, params)
    nsides = int(toks[pos]); pos += 1
    while pos < len(toks) and not toks[pos].lstrip("-").isdigit():
        pos += 1
    sides = []
    for _ in range(nsides):
        sides.append((int(toks[pos]), int(toks[pos + 1]), toks[pos + 2])); pos += 3
    return types, sides


def boundary_con(con: np.ndarray, bc_rows, types: Dict[str, int]) -> np.ndarray:
    """Connectivity with the boundary sides of ``bc_rows`` turned into rows
    (eid, sid, -1, type), ``types`` mapping each tag to its boundary type."""
    bset = {(e, s): types[t] for t, e, s in bc_rows}
    out = [r for r in con.tolist() if (r[0], r[1]) not in bset]
    out += [[e, s, -1, t] for (e, s), t in sorted(bset.items())]
    return np.asarray(out, dtype=np.int32).reshape(-1, 4)


_FUNCS = {"sin": np.sin, "cos": np.cos, "tan": np.tan, "asin": np.arcsin, "acos": np.arccos,
          "atan": np.arctan, "sinh": np.sinh, "cosh": np.cosh, "tanh": np.tanh, "exp": np.exp,
          "log": np.log, "log10": np.log10, "sqrt": np.sqrt, "abs": np.abs, "fabs": np.abs,
          "atan2": np.arctan2, "pow": np.power}
_CONSTS = {"PI": np.pi, "pi": np.pi}


def _tokens(text: str):
    import re
    spec = r"\s*(?:(\d+\.?\d*(?:[eE][-+]?\d+)?|\.\d+(?:[eE][-+]?\d+)?)|([A-Za-z_]\w*)|(:=|[-+*/^(),;{}]))"
    pos, out = 0, []
    text = text.rstrip()
    while pos < len(text):
        m = re.compile(spec).match(text, pos)
        if not m or m.end() == pos:
            raise ValueError(f"bad character in function text near {text[pos:pos + 20]!r}")
        num, name, op = m.groups()
        out.append(("n", float(num)) if num else (("v", name) if name else ("o", op)))
        pos = m.end()
        while pos < len(text) and text[pos].isspace():
            pos += 1
    return out


class _ExprParser:
    """Arithmetic expressions of the reference's function definitions
    (src/Parser.y): + - * / ^ (right associative, above unary minus),
    parentheses, the usual elementary functions, PI, and the variables x, y,
    z, t.  Builds a tree of tuples; nothing in the text is executed."""

    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self):
        return self.t[self.i] if self.i < len(self.t) else ("o", None)

    def take(self, op=None):
        tok = self.peek()
        if op is not None and tok != ("o", op):
            raise ValueError(f"expected {op!r}, found {tok[1]!r}")
        self.i += 1
        return tok

    def expr(self):
        f = self.term()
        while self.peek() in (("o", "+"), ("o", "-")):
            f = ("bin", self.take()[1], f, self.term())
        return f

    def term(self):
        f = self.unary()
        while self.peek() in (("o", "*"), ("o", "/")):
            f = ("bin", self.take()[1], f, self.unary())
        return f

    def unary(self):
        if self.peek() == ("o", "-"):
            self.take()
            return ("neg", self.unary())
        if self.peek() == ("o", "+"):
            self.take()
            return self.unary()
        return self.power()

    def power(self):
        f = self.atom()
        if self.peek() == ("o", "^"):
            self.take()
            return ("bin", "^", f, self.unary())
        return f

    def atom(self):
        kind, val = self.take()
        if kind == "n":
            return ("num", val)
        if kind == "v":
            if self.peek() == ("o", "("):
                if val not in _FUNCS:
                    raise ValueError(f"unknown function {val}")
                self.take("(")
                args = [self.expr()]
                while self.peek() == ("o", ","):
                    self.take()
                    args.append(self.expr())
                self.take(")")
                if len(args) != (2 if val in ("atan2", "pow") else 1):
                    raise ValueError(f"{val} takes {2 if val in ('atan2', 'pow') else 1} arguments")
                return ("call", val, args)
            if val in _CONSTS:
                return ("num", _CONSTS[val])
            if val not in ("x", "y", "z", "t"):
                raise ValueError(f"unknown variable {val}")
            return ("var", val)
        if val == "(":
            f = self.expr()
            self.take(")")
            return f
        raise ValueError(f"unexpected {val!r}")


_BINOPS = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide, "^": np.power}


def eval_tree(node, env):
    """Value of an expression tree at env (x, y, z, t arrays or scalars)."""
    kind = node[0]
    if kind == "num":
        return node[1]
    if kind == "var":
        return env[node[1]]
    if kind == "neg":
        return -eval_tree(node[1], env)
    if kind == "bin":
        return _BINOPS[node[1]](eval_tree(node[2], env), eval_tree(node[3], env))
    return _FUNCS[node[1]](*[eval_tree(a, env) for a in node[2]])


# postfix instruction set of dgx_function (include/dgx.h): opcode in the low
# 8 bits, constant-table index above them for DGX_OP_CONST
OPCODES = {"const": 0, "x": 1, "y": 2, "z": 3, "t": 4, "+": 5, "-": 6, "*": 7, "/": 8, "^": 9,
           "neg": 10, "sin": 11, "cos": 12, "tan": 13, "asin": 14, "acos": 15, "atan": 16,
           "sinh": 17, "cosh": 18, "tanh": 19, "exp": 20, "log": 21, "log10": 22, "sqrt": 23,
           "abs": 24, "fabs": 24, "atan2": 25, "pow": 9}


def emit_tree(node, code, consts):
    """Append the postfix program of an expression tree to code."""
    kind = node[0]
    if kind == "num":
        consts.append(float(node[1]))
        code.append(OPCODES["const"] | ((len(consts) - 1) << 8))
    elif kind == "var":
        code.append(OPCODES[node[1]])
    elif kind == "neg":
        emit_tree(node[1], code, consts)
        code.append(OPCODES["neg"])
    elif kind == "bin":
        emit_tree(node[2], code, consts)
        emit_tree(node[3], code, consts)
        code.append(OPCODES[node[1]])
    else:
        for a in node[2]:
            emit_tree(a, code, consts)
        code.append(OPCODES[node[1]])


def compile_function(comps: Dict[str, object]):
    """Postfix programs of a read_functions block, one per component, in the
    layout of dgx_function: (start [ncomp + 1], code, consts)."""
    code: List[int] = []
    consts: List[float] = []
    start = [0]
    for tree in comps.values():
        emit_tree(tree, code, consts)
        start.append(len(code))
    return (np.asarray(start, dtype=np.int32), np.asarray(code, dtype=np.int32),
            np.asarray(consts, dtype=np.float64))


def read_functions(path: str) -> Dict[str, Dict[str, object]]:
    """Function definitions of a boundary-type or source file (``root.bct``,
    ``root.src``): blocks ``function NAME { var := expression; ... }``, the
    form FuncBC and FuncSource read (src/BC.hpp, src/Source.hpp:190-230).
    Returns {NAME: {var: expression tree}} (eval_tree, compile_function); the
    components keep their order in the block."""
    with open(path) as f:
        text = f.read().replace("\\\n", " ")
    toks = _tokens(text)
    out: Dict[str, Dict[str, object]] = {}
    i = 0
    while i < len(toks):
        if toks[i] == ("v", "function") or toks[i] == ("v", "Function"):
            name = toks[i + 1][1]
            if toks[i + 2] != ("o", "{"):
                raise ValueError(f"{path}: expected '{{' after function {name}")
            i += 3
            comps: Dict[str, object] = {}
            while toks[i] != ("o", "}"):
                var = toks[i][1]
                if toks[i + 1] != ("o", ":="):
                    raise ValueError(f"{path}: expected ':=' after {var}")
                j = i + 2
                while toks[j] != ("o", ";"):
                    j += 1
                p = _ExprParser(toks[i + 2:j])
                comps[var] = p.expr()
                if p.i != j - (i + 2):
                    raise ValueError(f"{path}: trailing tokens in {var}")
                i = j + 1
            out[name] = comps
            i += 1
        else:
            i += 1
    return out


def read_bc_types(path: str) -> Dict[str, tuple]:
    """Boundary types of ``root.bct`` with numeric parameters: blocks
    ``code tag { v0 v1 ... }`` as Euler3d::make_BC and Navier_Stokes3d::make_BC
    read them (src/Euler3d.cpp:849-868, src/Navier_Stokes3d.cpp:431-455;
    e.g. ``Isothermal_Wall T { Tw Eref }``, ``State S { r ru rv rw rE }``).
    Returns {tag: (lower-case code, [values])}.  ``function`` blocks are left
    to read_functions."""
    with open(path) as f:
        toks = f.read().replace("{", " { ").replace("}", " } ").split()
    out: Dict[str, tuple] = {}
    i = 0
    while i < len(toks):
        code = toks[i].lower()
        if code == "function":
            depth, i = 0, i + 2
            while i < len(toks):
                depth += (toks[i] == "{") - (toks[i] == "}")
                i += 1
                if depth == 0:
                    break
            continue
        if i + 2 >= len(toks) or toks[i + 2] != "{":
            raise ValueError(f"{path}: expected 'code tag {{ ... }}' at {toks[i:i + 3]}")
        tag = toks[i + 1]
        j = i + 3
        vals: List[float] = []
        while toks[j] != "}":
            vals.append(float(toks[j]))
            j += 1
        out[tag] = (code, vals)
        i = j + 1
    return out


def field_function(comps: Dict[str, object]):
    """(t, x, y) -> array [ncomp, n] of a read_functions block."""
    def fn(t, x, y):
        env = {"t": t, "x": x, "y": y, "z": np.zeros_like(x)}
        return np.stack([np.broadcast_to(np.asarray(eval_tree(f, env), dtype=np.float64), np.shape(x))
                         for f in comps.values()])
    return fn


@dataclass
class InitialCondition:
    code: str
    values: List[float] = field(default_factory=list)
    file: str = ""               # Restart: the restart file to start from


def read_ic(path: str) -> InitialCondition:
    """Read the first IC block of ``root.ic`` (the reference reads only that one)."""
    with open(path) as f:
        f.readline()
        code = f.readline().split()[0]
        lc = code.lower()
        if lc == "vortex":
            nv = int(f.readline().split()[0])
            vals: List[float] = [float(nv)]
            for _ in range(nv):
                vals += [float(x) for x in f.readline().split()[:5]]
            return InitialCondition(code, vals)
        if lc == "acousticpulse":
            return InitialCondition(code, [float(x) for x in f.readline().split()[:3]])
        if lc == "given":
            vals = []
            for line in f:
                m = re.match(r"\s*[A-Za-z]+\s*:=\s*([-+0-9.eE]+)\s*;", line)
                if m:
                    vals.append(float(m.group(1)))
            return InitialCondition(code, vals)
        if lc == "mmsinitialcondition":                  # no parameters (Shallow::set_ic)
            return InitialCondition(code, [])
        if lc == "restart":                              # Euler3d::set_ic (src/Euler3d.cpp:957-965)
            return InitialCondition(code, [], f.readline().split()[0])
    raise ValueError(f"{path}: unsupported IC code {code}")


# ---------------------------------------------------------------------------
# Restart files
# ---------------------------------------------------------------------------
_TYPES = {"int32": "i", "int64": "q", "uint32": "I", "uint64": "Q", "float64": "d", "float32": "f"}


@dataclass
class Restart:
    header: Dict[str, object]
    nmodes: np.ndarray           # per-element modal counts
    data: np.ndarray             # [field][e][max nmodes] as float64; with variable order each
                                 # element's nmodes[e] modes come first, zeros after


def read_restart(path: str, ne: int) -> Restart:
    """Decode a DGM restart file (``.rst``/``.ref``) whose fields share ``ne`` elements."""
    raw = open(path, "rb").read()
    if raw[:1] == b"{":
        depth, pos = 0, 0
        for pos, ch in enumerate(raw):
            if ch == ord("{"):
                depth += 1
            elif ch == ord("}"):
                depth -= 1
                if depth == 0:
                    break
        end = pos + 1
        while end < len(raw) and raw[end:end + 1] in (b"\n", b"\r"):
            end += 1
        header = json.loads(raw[:pos + 1].decode())
        types = header.get("Types", {})
        gfmt = _TYPES[types.get("GlobalSize", "int32")]
        lfmt = _TYPES[types.get("LocalSize", "int32")]
        sfmt = _TYPES[types.get("Scalar", "float64")]
        nf = int(header["Number of fields"])
        off = end
    else:
        lines = raw.split(b"\n", 10)
        hdr = [ln.decode(errors="replace") for ln in lines[:10]]
        off = sum(len(ln) + 1 for ln in lines[:10])
        nel, nsd, lmax = (int(x) for x in hdr[3].split()[:3])
        header = {"Nel": nel, "Nsd": nsd, "Lmax": lmax, "Step": int(hdr[4].split()[0]),
                  "Time": float(hdr[5].split()[0]), "Time step": float(hdr[6].split()[0]),
                  "Viscosity": float(hdr[7].split()[0]),
                  "Format": hdr[9].split()[0]}
        names = hdr[8].split()
        nf = len(names) - 2 if names[-2:] == ["Field", "names"] else len(names)
        header["Field names"] = names[:nf]
        gfmt, lfmt, sfmt = "i", "i", "d"
    gsz, lsz, ssz = (struct.calcsize(c) for c in (gfmt, lfmt, sfmt))
    fields = []
    nmodes = None
    for _ in range(nf):
        off += gsz                                             # global length
        nm = np.frombuffer(raw, dtype="<" + lfmt, count=ne, offset=off).astype(np.int64)
        off += ne * lsz + ne * gsz                             # lengths, displacements
        tot = int(nm.sum())
        vals = np.frombuffer(raw, dtype="<" + sfmt, count=tot, offset=off).astype(np.float64)
        off += tot * ssz
        if nmodes is None:
            nmodes = nm
        if np.all(nm == nm[0]):
            fields.append(vals.reshape(ne, int(nm[0])))
        else:   # variable order: each element's modes first in a slot of the largest count
            pad = np.zeros((ne, int(nm.max())))
            starts = np.concatenate([[0], np.cumsum(nm)[:-1]])
            for e in range(ne):
                pad[e, :nm[e]] = vals[starts[e]:starts[e] + nm[e]]
            fields.append(pad)
    if off != len(raw):
        raise ValueError(f"{path}: {len(raw) - off} trailing bytes after {nf} fields")
    return Restart(header=header, nmodes=nmodes, data=np.stack(fields))


def write_restart(path: str, data: np.ndarray, header: Dict[str, object], nmodes=None) -> None:
    """Write a JSON-header restart file (int32 sizes, float64 modes) readable by DGM.

    ``data`` is ``[field][e][nmodes]``; with variable order ``nmodes[e]`` gives
    each element's count, its modes first in the slot.
    """
    nf, ne, nm = data.shape
    if nmodes is not None and not np.all(np.asarray(nmodes) == nm):
        nmodes = np.asarray(nmodes, dtype=np.int64)
        hdr = dict(header)
        hdr.setdefault("Basis type", "Normalized")
        hdr["Nel"] = ne
        hdr["Number of fields"] = nf
        hdr["Types"] = {"ByteOrder": "LittleEndian", "GlobalSize": "int32",
                        "LocalSize": "int32", "Scalar": "float64"}
        disp = np.concatenate([[0], np.cumsum(nmodes)[:-1]])
        with open(path, "wb") as f:
            f.write((json.dumps(hdr, indent=3, sort_keys=True) + "\n").encode())
            for v in range(nf):
                f.write(struct.pack("<i", int(nmodes.sum())))
                f.write(nmodes.astype("<i4").tobytes())
                f.write(disp.astype("<i4").tobytes())
                for e in range(ne):
                    f.write(np.ascontiguousarray(data[v, e, :nmodes[e]], dtype="<f8").tobytes())
        return
    hdr = dict(header)
    hdr.setdefault("Basis type", "Normalized")
    hdr["Nel"] = ne
    hdr["Number of fields"] = nf
    hdr["Types"] = {"ByteOrder": "LittleEndian", "GlobalSize": "int32",
                    "LocalSize": "int32", "Scalar": "float64"}
    with open(path, "wb") as f:
        f.write((json.dumps(hdr, indent=3, sort_keys=True) + "\n").encode())
        for v in range(nf):
            f.write(struct.pack("<i", ne * nm))
            f.write(np.full(ne, nm, dtype="<i4").tobytes())
            f.write((np.arange(ne, dtype="<i4") * nm).tobytes())
            f.write(np.ascontiguousarray(data[v], dtype="<f8").tobytes())
