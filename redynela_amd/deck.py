"""Reader and runner for the reference's `.dyn` input decks (the subset that
builds and solves one explicit domain).

The reference reads decks with a bison/flex interpreter
(Sources/Libraries/interpretor/InputDyn.y, InputDyn.l) that builds a
Structure from its scripting classes (interpretor/*.class) and then calls
Structure::solve (FE/Structure.C:809-897).  This module restates that
language for the constructs the Taylor sample decks use, with the semantics
of the reference's interpreter:

* declarations `Real a = expr;`, `Type name;`; assignments `=`, `+=`, `-=`,
  `*=`, `/=`, `++`, `--`;
* `for (init; cond; step) stmt`, `if (cond) stmt`, blocks `{ ... }`;
* expressions with `+ - * / ^`, comparisons, `&& || !`, unary minus, `PI`,
  strings and the functions abs, int (truncation), round (int(x + 0.5)),
  sqrt, sin/cos/tan/asin/acos/atan (in degrees unless `#pragma radians`,
  as InputDyn.C:341 defaults), sinh, cosh, tanh, exp, erf, ln, log (base 10);
* `END PREPROCESSOR;`.

Supported classes and methods (the rest raise DeckError naming the
construct): Structure (createNode, createElement, setDefaultElement,
addMaterial, attachMaterialToElements, attachInitialBC, attachConstantBC,
setSaveTimes, addSolver), NodeSet and ElementSet (add(i), add(i, j),
add(i, j, step): FE/Select.C:151-212), Select, ElQua4nAx, ElHex8n3D,
Material, IsoHardElastoplastic, IsoHardJohnsonCook, BoundaryRestrain,
BoundarySpeed, ExplicitSolverCH (Courant or Pulse time step), HistoryFile
(nodal fields and the element fields stress, strain, strainInc,
plasticStrain) and
Structure::addHistoryFile.

`load(path)` returns a Deck (mesh arrays, material, BCs, solver and save
times); `Deck.run()` runs the Structure::solve loop over the save times on
the GPU and hands each saved state to a callback.
"""
import math
import os
import re

import numpy as np

from . import _lib
from ._lib import DnlgError


class DeckError(DnlgError):
    pass


# ---------------------------------------------------------------------------
# tokens
# ---------------------------------------------------------------------------
_TOKEN = re.compile(r"""
    (?P<ws>[ \t\r\n]+) |
    (?P<lcomment>//[^\n]*) |
    (?P<bcomment>/\*.*?\*/) |
    (?P<pragma>\#[^\n]*) |
    (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?) |
    (?P<str>"[^"]*") |
    (?P<id>[A-Za-z_][A-Za-z_0-9]*) |
    (?P<op>\+\+|--|\+=|-=|\*=|/=|<=|>=|==|!=|&&|\|\||[-+*/^<>=!(){};,.])
""", re.S | re.X)


def tokenize(text):
    toks, pos = [], 0
    line = 1
    while pos < len(text):
        m = _TOKEN.match(text, pos)
        if not m:
            raise DeckError("line %d: unexpected character %r" % (line, text[pos]))
        kind = m.lastgroup
        val = m.group(kind)
        if kind == "pragma":
            toks.append(("pragma", val[1:].strip(), line))
        elif kind not in ("ws", "lcomment", "bcomment"):
            toks.append((kind, val, line))
        line += val.count("\n")
        pos = m.end()
    toks.append(("eof", None, line))
    return toks


# ---------------------------------------------------------------------------
# parser: statements and expressions as nested tuples
# ---------------------------------------------------------------------------
# precedence levels of InputDyn.y:104-110: comparisons, && and || share one
# left-associative level; then + -, * /, ^; unary minus binds tightest
_BINARY = [("==", "!=", "<", "<=", ">", ">=", "&&", "||"), ("+", "-"), ("*", "/"), ("^",)]
_FUNCS = {"abs", "int", "round", "sqrt", "cos", "sin", "tan", "cosh", "sinh", "tanh", "acos", "asin", "atan",
          "exp", "erf", "ln", "log"}


class Parser:
    def __init__(self, toks):
        self.t, self.i = toks, 0

    def peek(self, k=0):
        return self.t[self.i + k]

    def take(self, val=None):
        tok = self.t[self.i]
        if val is not None and tok[1] != val:
            raise DeckError("line %d: expected %r, found %r" % (tok[2], val, tok[1]))
        self.i += 1
        return tok

    def program(self):
        out = []
        while self.peek()[0] != "eof":
            if self.peek()[1] == "END" and self.peek(1)[1] == "PREPROCESSOR":
                self.i += 2
                if self.peek()[1] == ";":
                    self.i += 1
                break
            out.append(self.statement())
        return out

    def statement(self):
        kind, val, line = self.peek()
        if kind == "pragma":
            self.i += 1
            return ("pragma", val, line)
        if val == ";":
            self.i += 1
            return ("nop",)
        if val == "{":
            self.i += 1
            body = []
            while self.peek()[1] != "}":
                if self.peek()[0] == "eof":
                    raise DeckError("line %d: unterminated block" % line)
                body.append(self.statement())
            self.i += 1
            return ("block", body)
        if kind == "id" and val == "for":
            self.i += 1
            self.take("(")
            init = self.expression() if self.peek()[1] != ";" else None
            self.take(";")
            cond = self.expression() if self.peek()[1] != ";" else None
            self.take(";")
            step = self.expression() if self.peek()[1] != ")" else None
            self.take(")")
            return ("for", init, cond, step, self.statement(), line)
        if kind == "id" and val == "if":
            self.i += 1
            self.take("(")
            cond = self.expression()
            self.take(")")
            then = self.statement()
            other = None
            if self.peek()[1] == "else":
                self.i += 1
                other = self.statement()
            return ("if", cond, then, other, line)
        # declaration: Type name [= expr] ;   (Type starts with an upper-case letter)
        if kind == "id" and val[0].isupper() and self.peek(1)[0] == "id" and self.peek(2)[1] in (";", "=", ","):
            self.i += 1
            names = []
            while True:
                name = self.take()[1]
                init = None
                if self.peek()[1] == "=":
                    self.i += 1
                    init = self.expression()
                names.append((name, init))
                if self.peek()[1] == ",":
                    self.i += 1
                    continue
                break
            self.take(";")
            return ("decl", val, names, line)
        e = self.expression()
        if self.peek()[1] == ";":
            self.i += 1
        return ("expr", e, line)

    def expression(self):
        # assignment (right associative)
        kind, val, line = self.peek()
        if kind == "id" and self.peek(1)[1] in ("=", "+=", "-=", "*=", "/="):
            self.i += 2
            return ("assign", self.peek(-1)[1], val, self.expression(), line)
        return self.binary(0)

    def binary(self, level):
        if level == len(_BINARY):
            return self.unary()
        lhs = self.binary(level + 1)
        while self.peek()[1] in _BINARY[level] and self.peek()[0] == "op":
            op = self.take()[1]
            rhs = self.binary(level + 1)
            lhs = ("bin", op, lhs, rhs)
        return lhs

    def unary(self):
        kind, val, line = self.peek()
        if val == "!" and kind == "op":
            # NOT shares the comparisons' level (InputDyn.y:105, 1560): its
            # operand runs over every tighter operator, so !a + b is !(a + b)
            # while !a < b is (!a) < b
            self.i += 1
            return ("un", val, self.binary(1))
        if val in ("-", "+") and kind == "op":  # UMINUS binds tightest: -2^2 is (-2)^2
            self.i += 1
            return ("un", val, self.unary())
        if val in ("++", "--") and kind == "op":
            self.i += 1
            return ("incr", self.take()[1], 1 if val == "++" else -1, line)
        return self.postfix()

    def postfix(self):
        kind, val, line = self.take()
        if kind == "num":
            return ("num", float(val))
        if kind == "str":
            return ("str", val[1:-1])
        if val == "(":
            e = self.expression()
            self.take(")")
            return e
        if kind != "id":
            raise DeckError("line %d: unexpected %r" % (line, val))
        if val == "PI":
            return ("num", math.pi)
        if val.lower() in _FUNCS and self.peek()[1] == "(":
            self.i += 1
            a = self.expression()
            self.take(")")
            return ("fn", val.lower(), a, line)
        node = ("var", val, line)
        if self.peek()[1] in ("++", "--"):
            op = self.take()[1]
            return ("incr", val, 1 if op == "++" else -1, line)
        while self.peek()[1] == ".":
            self.i += 1
            meth = self.take()[1]
            self.take("(")
            args = []
            while self.peek()[1] != ")":
                args.append(self.expression())
                if self.peek()[1] == ",":
                    self.i += 1
            self.take(")")
            node = ("call", node, meth, args, line)
        return node


# ---------------------------------------------------------------------------
# scripting classes (interpretor/*.class) mapped onto the model
# ---------------------------------------------------------------------------
def _arith(op, a, b):
    """double arithmetic as the interpreter's C++ does it: IEEE results
    (inf, nan) instead of Python's exceptions."""
    with np.errstate(all="ignore"):
        a, b = np.float64(a), np.float64(b)
        if op == "+":
            return float(a + b)
        if op == "-":
            return float(a - b)
        if op == "*":
            return float(a * b)
        if op == "/":
            return float(a / b)
        return float(np.power(a, b))


def _idx(v):
    """Real -> Indice, as the interpreter's argument conversion (truncation)."""
    return int(v)


class Obj:
    kind = "object"

    def call(self, deck, meth, args, line):
        f = getattr(self, "m_" + meth, None)
        if f is None:
            raise DeckError("line %d: %s.%s is not supported by this reader" % (line, self.kind, meth))
        return f(deck, *args)


class ElementType(Obj):
    def __init__(self, kind):
        self.kind = kind
        self.dim, self.nen = {"ElQua4nAx": (2, 4), "ElHex8n3D": (3, 8)}[kind]


class NodeSet(Obj):
    kind = "NodeSet"

    def __init__(self):
        self.items = []

    def m_add(self, deck, a, b=None, inc=1):
        if isinstance(a, NodeSet):
            self.items += a.items
            return
        a = _idx(a)
        b = a if b is None else _idx(b)
        inc = _idx(inc)
        for i in range(a, b + 1, inc):
            if i not in deck.node_index:
                raise DeckError("NodeSet::add: Node %d not exist in current grid and domain" % i)
            self.items.append(i)


class ElementSet(NodeSet):
    kind = "ElementSet"

    def m_add(self, deck, a, b=None, inc=1):
        if isinstance(a, ElementSet):
            self.items += a.items
            return
        a = _idx(a)
        b = a if b is None else _idx(b)
        for i in range(a, b + 1, _idx(inc)):
            if i not in deck.elem_index:
                raise DeckError("ElementSet::add: Element %d not exist in current grid and domain" % i)
            self.items.append(i)


class Select(Obj):
    kind = "Select"


class Hardening(Obj):
    def __init__(self, kind):
        self.kind = kind
        self.A, self.B, self.n = 0.0, 0.0, 1.0
        self.C, self.m, self.depsp0, self.Tm, self.T0 = 0.0, 1.0, 1.0, 0.0, 0.0

    def m_setYieldStress(self, deck, v):
        self.A = float(v)

    def m_setHardParameter(self, deck, v):
        self.B = float(v)

    def m_setHardExponent(self, deck, v):
        self.n = float(v)

    def m_defineLaw(self, deck, A, B, C, n, m, epsp0, Tm, T0):  # IsoHardJohnsonCook
        self.A, self.B, self.C, self.n, self.m, self.depsp0, self.Tm, self.T0 = A, B, C, n, m, epsp0, Tm, T0


class MaterialObj(Obj):
    kind = "Material"

    def __init__(self):
        self.young = self.poisson = self.density = 0.0
        self.init_temp = 0.0  # Material::T0 (FE/Material.C:49), written as nodal T in result files
        self.hardening = None

    def m_setYoung(self, deck, v):
        self.young = float(v)

    def m_setPoisson(self, deck, v):
        self.poisson = float(v)

    def m_setDensity(self, deck, v):
        self.density = float(v)

    def m_setColor(self, deck, *a):  # display only
        pass

    def m_setInitTemperature(self, deck, v):  # FE/Material.C:162-167
        self.init_temp = float(v)

    def m_setHardening(self, deck, h):
        if not isinstance(h, Hardening):
            raise DeckError("Material::setHardening: %s is not supported by this reader" % getattr(h, "kind", h))
        self.hardening = h


class BoundaryRestrain(Obj):
    kind = "BoundaryRestrain"

    def __init__(self):
        self.v = (0, 0, 0)

    def m_set(self, deck, x, y, z):
        self.v = (float(x), float(y), float(z))


class BoundarySpeed(BoundaryRestrain):
    kind = "BoundarySpeed"


class SolverObj(Obj):
    kind = "ExplicitSolverCH"

    def __init__(self):
        self.start, self.stop = 0.0, 1e300
        self.rho_b, self.safety, self.freq = 0.8182, 0.9, 10
        self.method = "Pulse"  # FE/ExplicitSolverCH.C:83

    def m_setTimes(self, deck, a, b):
        self.start, self.stop = float(a), float(b)

    def m_setTimeStepMethod(self, deck, name):  # FE/Solver.C:312-327
        if name not in ("Courant", "Pulse"):
            raise DeckError("setTimeStepMethod: Unknown method %s" % name)
        self.method = name

    def m_setDissipation(self, deck, v):
        if v < 0.0 or v > 1.0:
            raise DeckError("ExplicitSolverCH::setDissipation: dissipation value out of range [0.0, 1.0]")
        self.rho_b = float(v)

    def m_setTimeStepSafetyFactor(self, deck, v):
        self.safety = float(v)

    def m_setComputeTimeStepFrequency(self, deck, v):
        self.freq = _idx(v)


class HistoryFileObj(Obj):
    """HistoryFile (interpretor/HistoryFile.class): append(NodeSet, String[, Indice]),
    setName(String), setSaveTime(Real) or (Real, Real, Real)."""
    kind = "HistoryFile"

    def __init__(self):
        self.fileName = ""
        self.save = None
        self.items = []  # (node numbers, field, comp)

    def m_append(self, deck, ns, field, comp=0):
        from ._lib import HF_FIELDS
        if not isinstance(ns, NodeSet):
            raise DeckError("HistoryFile::append: not a NodeSet")
        if field not in HF_FIELDS:
            raise DeckError("getNodalValue: undefined field %s" % field)
        kind = HF_FIELDS[field][1]
        comp = _idx(comp)
        if kind == "tensor":
            if comp != 0 and not (1 <= comp // 10 <= 3 and 1 <= comp % 10 <= 3):
                raise DeckError("Node::getNodalValue: No sense for component %d of %s" % (comp, field))
        elif comp < 0 or (kind and comp > 3) or (not kind and comp > 1):
            raise DeckError("Node::getNodalValue::get: No sense for component %d of %s" % (comp, field))
        self.items.append((list(ns.items), field, comp))

    def m_setName(self, deck, name):
        self.fileName = name

    def m_setSaveTime(self, deck, *a):
        if len(a) not in (1, 3):
            raise DeckError("HistoryFile::setSaveTime takes 1 or 3 arguments")
        self.save = tuple(float(v) for v in a)


class StructureObj(Obj):
    kind = "Structure"

    def m_createNode(self, deck, num, x, y, z):
        num = _idx(num)
        if num in deck.node_index:  # FE/Domain.C:1542-1546
            raise DeckError("Domain::add: Node %d already exists in the node list of this grid" % num)
        deck.node_index[num] = len(deck.nodes)
        deck.nodes.append((float(x), float(y), float(z)))

    def m_setDefaultElement(self, deck, el):
        if not isinstance(el, ElementType):
            raise DeckError("Structure::setDefaultElement: not an element type")
        deck.elem_type = el

    def m_createElement(self, deck, num, *nodes):
        if deck.elem_type is None:
            raise DeckError("Structure::createElement: no default element")
        if len(nodes) != deck.elem_type.nen:
            raise DeckError("Structure::createElement: %s takes %d nodes" % (deck.elem_type.kind, deck.elem_type.nen))
        num = _idx(num)
        if num in deck.elem_index:  # FE/Domain.C:1578-1583
            raise DeckError("Domain::add: Element %d already exists in the element list of this grid" % num)
        ids = []
        for n in nodes:
            n = _idx(n)
            if n not in deck.node_index:
                raise DeckError("Structure::createElement: node %d does not exist" % n)
            ids.append(deck.node_index[n])
        deck.elem_index[num] = len(deck.elements)
        deck.elements.append(ids)

    def m_addMaterial(self, deck, mat):
        deck.materials.append(mat)

    def m_attachMaterialToElements(self, deck, mat, es):
        deck.material_sets.append((mat, list(es.items)))

    def m_attachInitialBC(self, deck, bc, ns):
        if not isinstance(bc, BoundarySpeed):
            raise DeckError("Structure::attachInitialBC: only BoundarySpeed is supported")
        deck.initial.append((bc.v, list(ns.items)))

    def m_attachConstantBC(self, deck, bc, ns):
        if not isinstance(bc, (BoundaryRestrain, BoundarySpeed)):
            raise DeckError("Structure::attachConstantBC: unsupported condition")
        deck.constant.append((bc, bc.v, list(ns.items)))

    def m_setSaveTimes(self, deck, a, b, c):
        deck.save = (float(a), float(b), float(c))

    def m_addSolver(self, deck, s):
        deck.solver = s

    def m_addHistoryFile(self, deck, hf):
        if not isinstance(hf, HistoryFileObj):
            raise DeckError("Structure::addHistoryFile: not a HistoryFile")
        deck.history_files.append(hf)


_CLASSES = {"Structure": StructureObj, "Select": Select, "NodeSet": NodeSet, "ElementSet": ElementSet,
            "Material": MaterialObj, "BoundaryRestrain": BoundaryRestrain, "BoundarySpeed": BoundarySpeed,
            "HistoryFile": HistoryFileObj,
            "ExplicitSolverCH": SolverObj,
            "IsoHardElastoplastic": lambda: Hardening("IsoHardElastoplastic"),
            "IsoHardJohnsonCook": lambda: Hardening("IsoHardJohnsonCook"),
            "ElQua4nAx": lambda: ElementType("ElQua4nAx"), "ElHex8n3D": lambda: ElementType("ElHex8n3D")}


# ---------------------------------------------------------------------------
# the deck: interpretation and the resulting model
# ---------------------------------------------------------------------------
class Deck:
    def __init__(self):
        self.vars = {}
        self.degrees = True
        self.nodes, self.node_index = [], {}
        self.elements, self.elem_index = [], {}
        self.elem_type = None
        self.materials, self.material_sets = [], []
        self.initial, self.constant = [], []
        self.save = None
        self.solver = None
        self.history_files = []

    # -- interpretation ----------------------------------------------------
    def execute(self, stmts):
        for s in stmts:
            self._stmt(s)

    def _stmt(self, s):
        k = s[0]
        if k == "nop":
            return
        if k == "pragma":
            w = s[1].split()
            if w and w[0] == "pragma" and len(w) > 1:
                w = w[1:]
            if w and w[0].lower() == "radians":
                self.degrees = False
            elif w and w[0].lower() == "degrees":
                self.degrees = True
            return
        if k == "block":
            self.execute(s[1])
        elif k == "decl":
            _, typ, names, line = s
            for name, init in names:
                if typ == "Real":
                    self.vars[name] = float(self.eval(init)) if init is not None else 0.0
                elif typ in _CLASSES:
                    if init is not None:
                        raise DeckError("line %d: %s %s cannot be initialised" % (line, typ, name))
                    self.vars[name] = _CLASSES[typ]()
                else:
                    raise DeckError("line %d: class %s is not supported by this reader" % (line, typ))
        elif k == "expr":
            self.eval(s[1])
        elif k == "for":
            _, init, cond, step, body, line = s
            if init is not None:
                self.eval(init)
            guard = 0
            while cond is None or self.eval(cond):
                self._stmt(body)
                if step is not None:
                    self.eval(step)
                guard += 1
                if guard > 100000000:
                    raise DeckError("line %d: runaway loop" % line)
        elif k == "if":
            _, cond, then, other, line = s
            if self.eval(cond):
                self._stmt(then)
            elif other is not None:
                self._stmt(other)

    def _var(self, name, line):
        if name not in self.vars:
            raise DeckError("line %d: unknown name %s" % (line, name))
        return self.vars[name]

    def eval(self, e):
        k = e[0]
        if k == "num" or k == "str":
            return e[1]
        if k == "var":
            return self._var(e[1], e[2])
        if k == "assign":
            _, op, name, rhs, line = e
            v = float(self.eval(rhs))
            if op == "=":
                self.vars[name] = v
            else:
                cur = self._var(name, line)
                self.vars[name] = _arith(op[0], cur, v)
            return self.vars[name]
        if k == "incr":
            _, name, d, line = e
            self.vars[name] = self._var(name, line) + d
            return self.vars[name]
        if k == "un":
            v = self.eval(e[2])
            return -v if e[1] == "-" else (v if e[1] == "+" else float(not v))
        if k == "bin":
            op, a = e[1], self.eval(e[2])
            if op == "&&":
                return float(bool(a) and bool(self.eval(e[3])))
            if op == "||":
                return float(bool(a) or bool(self.eval(e[3])))
            b = self.eval(e[3])
            if op in "+-*/^":
                return _arith(op, a, b)
            return float({"==": a == b, "!=": a != b, "<": a < b, "<=": a <= b, ">": a > b, ">=": a >= b}[op])
        if k == "fn":
            _, name, arg, line = e
            x = float(self.eval(arg))
            d2r = math.pi / 180.0 if self.degrees else 1.0
            f = {"abs": abs, "int": lambda v: float(int(v)), "round": lambda v: float(int(v + 0.5)),
                 "sqrt": math.sqrt, "exp": math.exp, "erf": math.erf, "ln": math.log, "log": math.log10,
                 "sinh": lambda v: math.sinh(d2r * v), "cosh": lambda v: math.cosh(d2r * v),
                 "tanh": lambda v: math.tanh(d2r * v), "sin": lambda v: math.sin(d2r * v),
                 "cos": lambda v: math.cos(d2r * v), "tan": lambda v: math.tan(d2r * v),
                 "asin": lambda v: math.asin(v) / d2r, "acos": lambda v: math.acos(v) / d2r,
                 "atan": lambda v: math.atan(v) / d2r}[name]
            return f(x)
        if k == "call":
            _, target, meth, args, line = e
            obj = self.eval(target)
            if not isinstance(obj, Obj):
                raise DeckError("line %d: %s has no method %s" % (line, target[1], meth))
            return obj.call(self, meth, [self.eval(a) for a in args], line)
        raise DeckError("unknown expression %r" % (k,))

    # -- the model ---------------------------------------------------------
    def model(self):
        """mesh dict (dim, x, conn, bc, v0, speed_mask, speed), material,
        solver, save times; raises DeckError on what the GPU path lacks."""
        if self.elem_type is None or not self.elements:
            raise DeckError("deck defines no elements")
        if self.solver is None:
            raise DeckError("deck adds no solver")
        mats = {id(m): m for m, _ in self.material_sets}
        if len(mats) != 1:
            raise DeckError("one material for all elements is supported (the deck attaches %d)" % len(mats))
        mat = next(iter(mats.values()))
        covered = set()
        for _, els in self.material_sets:
            covered.update(els)
        if len(covered) != len(self.elements):
            raise DeckError("elements without a material: %d" % (len(self.elements) - len(covered)))
        # a Domain keeps its nodes and elements sorted by number
        # (FE/Domain.C:1548-1555, 1585-1592): that is the order they are
        # computed in and written to result files
        nn = len(self.nodes)
        node_nums = sorted(self.node_index)
        rank = np.empty(nn, np.int64)
        rank[[self.node_index[k] for k in node_nums]] = np.arange(nn)
        x = np.array([self.nodes[self.node_index[k]] for k in node_nums], dtype=np.float64).reshape(nn, 3)
        conn = rank[np.array([self.elements[self.elem_index[k]] for k in sorted(self.elem_index)])].astype(np.int32)
        bc = np.zeros(nn, np.uint8)
        v0 = np.zeros((nn, 3))
        smask = np.zeros(nn, np.uint8)
        speed = np.zeros((nn, 3))
        # Domain::initSolve (FE/Domain.C:1141-1146) applies applyInitial_V0 of a
        # node's initial conditions, then of its constant ones
        # (FE/Boundary.C:144-161): BoundarySpeed sets the whole speed vector
        # (FE/Boundary.C:378-384), BoundaryRestrain nothing (FE/Boundary.C:600-603)
        for v, nodes in self.initial:
            for n in nodes:
                v0[rank[self.node_index[n]]] = v
        for obj, v, nodes in self.constant:
            if isinstance(obj, BoundarySpeed):
                for n in nodes:
                    v0[rank[self.node_index[n]]] = v
        # in the time loop BoundarySpeed::applyConstant imposes only the
        # components whose speed is not zero (FE/Boundary.C:478-524), and a
        # node's constant conditions apply in the order they were attached
        # (FE/Boundary.C:200-215); the kernels apply speeds before restraints,
        # so a speed attached after a restraint on the same component is refused
        restrained_first = set()
        for obj, v, nodes in self.constant:
            for n in nodes:
                i = rank[self.node_index[n]]
                if isinstance(obj, BoundarySpeed):
                    for c in range(3):
                        if v[c] != 0.0:
                            if (i, c) in restrained_first:
                                raise DeckError("Structure::attachConstantBC: a BoundarySpeed attached after a "
                                                "BoundaryRestrain on the same component of node %d is not supported" % n)
                            smask[i] |= 1 << c
                            speed[i, c] = v[c]
                else:
                    for c in range(3):
                        if v[c] == 1:
                            bc[i] |= 1 << c
                            restrained_first.add((i, c))
        self._rank = {num: int(rank[self.node_index[num]]) for num in node_nums}
        mesh = dict(dim=self.elem_type.dim, x=x, conn=conn, bc=bc, v0=v0)
        if smask.any():
            mesh["speed_mask"], mesh["speed"] = smask, speed
        return mesh, mat, self.solver

    def result_base(self):
        """the deck's file name before its extension (FE/Structure.C:72-74)."""
        if getattr(self, "source", None) is None:
            raise DeckError("a deck read from text has no file name for its results")
        base, ext = os.path.splitext(self.source)
        return base if ext == ".dyn" else self.source

    def run(self, form="fp64", on_save=None, results=None, restart=None):
        """Structure::solve on the GPU (FP64 by default): the save-time loop of
        FE/Structure.C:809-897 with saveResults' test (Structure.C:786-799)
        and startSaveTime defaulting to saveTime (Structure.C:364-366).
        on_save(t, domain) is called where the reference writes a result
        file; results=<base> (or True: the deck's name) writes those files, `<base>.NNN.bin` in the
        reference's layout (redynela_amd.results), which needs the strain
        fields and so runs with strain tracking; it also writes the domain's
        `<base>_defaultDomain.plot` history (redynela_amd.results.PlotFile).
        restart=N continues from result file N of <base>, as DynELA_solve's
        restart option does (Sources/Applications/DynELA_solve/main.C:106-130):
        the state is read back (redynela_amd.results.restore), the next save
        is at (N + 1) saveTime and the files go on from N + 1.
        Returns the Domain."""
        from . import domain as D
        from . import results as RF
        mesh, mat, sol = self.model()
        if results is True:
            results = self.result_base()
        writer = RF.ResultFiles(results, mat.density, mat.init_temp) if results is not None else None
        plot = RF.PlotFile(results + "_defaultDomain.plot") if results is not None else None
        h = mat.hardening
        hard = D.IsoHardElastoplastic(h.A, h.B, h.n) if h is not None else None
        m = D.Material(mat.young, mat.poisson, mat.density, hard)
        s = D.ExplicitSolverCH()
        s.setTimeStepMethod(sol.method)
        s.setDissipation(sol.rho_b)
        s.setTimeStepSafetyFactor(sol.safety)
        s.setComputeTimeStepFrequency(sol.freq)
        s.setTimes(sol.start, sol.stop)
        s.setPrecision(form)
        dom = D.Domain(mesh, m, s, speed_mask=mesh.get("speed_mask"), speed=mesh.get("speed"),
                       track_strain=writer is not None)
        if plot is not None:
            dom.set_option(_lib.OPT_PLOT_HISTORY, 1)
        hist = []
        for hfo in self.history_files:
            hf = D.HistoryFile()
            hf.setName(hfo.fileName)
            if hfo.save is None:
                raise DeckError("HistoryFile %s has no save time" % hfo.fileName)
            hf.setSaveTime(*hfo.save)
            for nums, field, comp in hfo.items:
                hf.append([self._rank[n] for n in nums], field, comp, numbers=nums)
            dom.addHistoryFile(hf, mat.init_temp)
            if not hfo.fileName:  # FE/HistoryFile.C:166-171
                raise DeckError("HistoryFile::open: No file name specified")
            where = os.path.dirname(results) if results is not None else os.path.dirname(self.source or "")
            hist.append((hf, RF.HistoryFileWriter(os.path.join(where, hfo.fileName), hf.columns)))
        save = self.save if self.save is not None else (0.0, sol.stop, sol.stop)
        resume = None
        if restart is not None:
            if writer is None:
                raise DeckError("restart needs the result files' base name (results=)")
            t0 = RF.restore(dom, "%s.%03d.bin" % (results, int(restart)))
            writer.count = int(restart) + 1
            # the restored Current field's ro and T (0 when the file was written at odd parity)
            writer.density, writer.T0 = dom.restored_density_T
            resume = (t0, (int(restart) + 1) * save[2])

        def solve(up_time):
            dom.solve(up_time)
            if plot is not None:
                plot.append(dom.plot_rows())
            for hf, w in hist:
                w.append(dom.history_rows(hf))
            return dom.time()["t"]

        def saved(t):
            if writer is not None:
                writer(t, dom)
            if on_save is not None:
                on_save(t, dom)

        structure_solve(sol.stop, save, solve, saved, resume=resume)
        return dom


def structure_solve(end_time, save_times, solve, save, resume=None):
    """The one-domain loop of Structure::solve (FE/Structure.C:833-883).

    save_times = (start, stop, every) as Structure::setSaveTimes takes them,
    with a start of 0 meaning `every` (FE/Structure.C:364-366); solve(t)
    advances the domain to t and returns its current time; save(t) is
    called where saveResults writes a file, under its test
    currentTime >= nextSaveTime and currentTime != previousSaveTime
    (FE/Structure.C:786-799).  resume = (current time, next save time) of a
    restarted run.  Returns the list of save times."""
    start, _, every = save_times
    st = dict(next=start if start != 0 else every, prev=-1.0, cur=0.0)
    if resume is not None:
        st["cur"], st["next"] = resume
    times = []

    def save_results():
        if st["cur"] >= st["next"] and st["cur"] != st["prev"]:
            save(st["cur"])
            times.append(st["cur"])
            st["prev"] = st["cur"]
            st["next"] += every

    next_time = 0.0
    while next_time < end_time:
        next_time = st["next"]
        if st["cur"] >= next_time:  # Domain::solve returns False (FE/Domain.C:1262-1266)
            raise DeckError("Solver Error: Unable to solve problem upto time = %10.3E" % next_time)
        st["cur"] = solve(next_time)
        save_results()
    save_results()
    return times


def load(path_or_text, text=False):
    """Interpret a deck (a file path, or the text itself with text=True) up
    to END PREPROCESSOR and return the Deck."""
    src = path_or_text if text else open(path_or_text, encoding="latin-1").read()
    d = Deck()
    d.source = None if text else path_or_text
    d.execute(Parser(tokenize(src)).program())
    return d
