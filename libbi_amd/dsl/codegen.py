"""HIP code generator for models written in LibBi's model language.

The reference turns a `.bi` file into C++/CUDA through its Perl front end and
templates: one class per block and per action (share/tt/cpp/model.hpp.tt,
share/tt/cpp/block/*.hpp.tt, share/tt/cpp/action/*.hpp.tt), compiled into
the program that runs the filter.  Here a model becomes one device functor
with the interface of the hand-written models (libbi_amd/csrc/models.hpp):
`rhs` for the integrators, and `transition`, `obs`, `obs_sample` and `init`
for the filter kernels, which are instantiated on it when the library is
built.

What a generated model supports (checked here, with an error that names the
line otherwise):
  * scalar variables (no `dim`): `param` and `input` (the model's parameter
    slots, in that order), `state`, `noise`, `obs`; `const` and `inline`;
  * `sub initial`: `~` and `<-` actions on state variables;
  * `sub transition(delta)`: `~` and `<-` actions and at most one
    `ode(alg, h, atoler, rtoler)` block of `dx/dt = ...` equations, run in the
    order written; discrete actions fire at multiples of delta
    (share/src/bi/state/Schedule.hpp:230-292), the ODE covers every interval;
  * `sub observation`: `~` actions on obs variables;
  * distributions gaussian/normal, log_gaussian/log_normal and uniform, with
    positional or named arguments (lib/Bi/Action/gaussian.pm, uniform.pm);
  * the ODE algorithms 'RK4' and 'RK5(4)' (share/tt/cpp/block/ode.hpp.tt).
Blocks the bootstrap filter does not run (`parameter`, `proposal_*`,
`lookahead_*`, `bridge`) are parsed and skipped.

Random draws are keyed Philox counters: the k-th normal draw of a block comes
from block k/4, word pair k%4 of the (particle, step, tag, batch) stream, as
the hand-written models draw theirs, so a model written here reproduces the
built-in model with the same equations bit for bit.
"""
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from .parser import (Action, Binary, Block, BiSyntaxError, Call, Model, Num, Ref, Str, Ternary,
                     Unary, parse, parse_file)

MAX_SLOTS = 16  # bihip_model::par
NORMAL_DISTS = {"gaussian": False, "normal": False, "log_gaussian": True, "log_normal": True}
DIST_ARGS = {"gaussian": ("mean", "std"), "uniform": ("lower", "upper")}
DIST_DEFAULTS = {"mean": 0.0, "std": 1.0, "lower": 0.0, "upper": 1.0}
UNIFORM_BLOCK = 0x40000000  # Philox block offset of uniform draws
# functions of the model language and their C++ names (overloaded for float and double)
FUNCS = {"exp": "exp", "log": "log", "sqrt": "sqrt", "pow": "pow", "abs": "fabs", "sin": "sin",
         "cos": "cos", "tan": "tan", "asin": "asin", "acos": "acos", "atan": "atan",
         "atan2": "atan2", "sinh": "sinh", "cosh": "cosh", "tanh": "tanh", "floor": "floor",
         "ceil": "ceil", "round": "round", "min": "fmin", "max": "fmax", "mod": "fmod",
         "erf": "erf", "erfc": "erfc", "lgamma": "lgamma", "tgamma": "tgamma", "exp10": "exp10",
         "log10": "log10", "cbrt": "cbrt"}
PY_FUNCS = {"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "pow": math.pow, "abs": abs,
            "sin": math.sin, "cos": math.cos, "tan": math.tan, "asin": math.asin,
            "acos": math.acos, "atan": math.atan, "atan2": math.atan2, "sinh": math.sinh,
            "cosh": math.cosh, "tanh": math.tanh, "floor": math.floor, "ceil": math.ceil,
            "min": min, "max": max, "mod": math.fmod, "erf": math.erf, "erfc": math.erfc,
            "lgamma": math.lgamma, "tgamma": math.gamma, "log10": math.log10,
            "exp10": lambda v: 10.0 ** v, "cbrt": lambda v: math.copysign(abs(v) ** (1 / 3), v)}
ALGS = {"RK4": 0, "RK5(4)": 1}  # BIHIP_RK4, BIHIP_DOPRI5


class GenError(ValueError):
    """A model that parses but uses something the generator does not support."""


@dataclass
class Draw:
    target: str
    dist: str  # gaussian | uniform
    log: bool
    a: object  # mean / lower
    b: object  # std / upper
    index: int  # among the block's draws of the same kind
    line: int


@dataclass
class Step:
    kind: str  # draw | set
    target: str
    expr: object = None
    draw: Optional[Draw] = None


@dataclass
class GenModel:
    name: str
    params: List[str]
    inputs: List[str]
    states: List[str]
    noises: List[str]
    obs: List[str]
    consts: Dict[str, float]
    inlines: Dict[str, object]
    initial: List[Step]
    pre: List[Step]  # transition actions before the ode block
    ode: Dict[str, object]  # state -> derivative
    post: List[Step]  # transition actions after the ode block
    observation: List[Draw]
    delta: float = 1.0
    integrator: int = 0
    h: float = 1.0
    atol: float = 1e-3
    rtol: float = 1e-3
    has_ode: bool = False
    normals: Dict[str, int] = field(default_factory=dict)   # block -> normal draws
    uniforms: Dict[str, int] = field(default_factory=dict)  # block -> uniform draws

    @property
    def slots(self):
        return self.params + self.inputs

    @property
    def cname(self):
        return "Gen_" + self.name


# ---------------------------------------------------------------- analysis
def _err(line, msg):
    raise GenError(f"line {line}: {msg}")


def _args(call, names, line):
    """Positional then named arguments of an action (Bi::Action::process_args)."""
    out = {}
    if len(call.args) > len(names):
        _err(line, f"{call.name} takes at most {len(names)} positional arguments")
    for n, a in zip(names, call.args):
        out[n] = a
    for n, a in call.named:
        if n == "log" and call.name in NORMAL_DISTS:
            continue
        if n not in names:
            _err(line, f"{call.name} has no argument {n!r}")
        if n in out:
            _err(line, f"argument {n!r} of {call.name} given twice")
        out[n] = a
    for n in names:
        out.setdefault(n, Num(DIST_DEFAULTS[n], repr(DIST_DEFAULTS[n]), False))
    return out


def _flatten(items):
    for it in items:
        if isinstance(it, list):
            yield from it
        else:
            yield it


def _const_eval(e, consts, line):
    if isinstance(e, Num):
        return float(e.value)
    if isinstance(e, Ref) and e.index is None and e.name in consts:
        return consts[e.name]
    if isinstance(e, Unary):
        v = _const_eval(e.arg, consts, line)
        return -v if e.op == "-" else (v if e.op == "+" else float(not v))
    if isinstance(e, Binary):
        a, b = _const_eval(e.left, consts, line), _const_eval(e.right, consts, line)
        ops = {"+": a + b, "-": a - b, "*": a * b, "**": None, "%": None}
        if e.op == "/":
            return a / b
        if e.op == "**":
            return math.pow(a, b)
        if e.op == "%":
            return math.fmod(a, b)
        if e.op in ops:
            return ops[e.op]
        cmp = {"<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b, "==": a == b, "!=": a != b,
               "&&": bool(a) and bool(b), "||": bool(a) or bool(b)}
        return float(cmp[e.op])
    if isinstance(e, Ternary):
        return (_const_eval(e.then, consts, line) if _const_eval(e.cond, consts, line)
                else _const_eval(e.other, consts, line))
    if isinstance(e, Call) and e.name in PY_FUNCS and not e.named:
        return float(PY_FUNCS[e.name](*[_const_eval(a, consts, line) for a in e.args]))
    _err(line, "a constant expression is required here")


class _Analyser:
    def __init__(self, m: Model):
        self.m = m
        if m.dims:
            _err(m.dims[0].line, "dimensions (dim) are not supported by the generator; "
                                 "vector models are written by hand (Lorenz-96)")
        for t, s in m.vars:
            if s.dims:
                _err(s.line, f"{t} {s.name} has dimensions, which the generator does not support")
            if t in ("state_aux_", "param_aux_"):
                _err(s.line, f"{t} variables are not supported")
        self.kind = {}
        for t, s in m.vars:
            if s.name in self.kind:
                _err(s.line, f"{s.name} declared twice")
            self.kind[s.name] = t
        self.consts = {}
        for n, e in m.consts:
            self.consts[n] = _const_eval(e, self.consts, 0)
        self.inlines = dict(m.inlines)

    def draw(self, a: Action, block, counts, local_consts):
        if not isinstance(a.expr, Call):
            _err(a.line, "the right side of '~' must be a distribution")
        name = a.expr.name
        if name in NORMAL_DISTS:
            args = _args(a.expr, DIST_ARGS["gaussian"], a.line)
            log = NORMAL_DISTS[name]
            for n, v in a.expr.named:
                if n == "log":
                    log = bool(_const_eval(v, self.consts, a.line))
            kind = "gaussian"
            k = counts.setdefault("normal", 0)
            counts["normal"] += 1
            d = Draw(a.target, kind, log, args["mean"], args["std"], k, a.line)
        elif name == "uniform":
            args = _args(a.expr, DIST_ARGS["uniform"], a.line)
            k = counts.setdefault("uniform", 0)
            counts["uniform"] += 1
            d = Draw(a.target, "uniform", False, args["lower"], args["upper"], k, a.line)
        else:
            _err(a.line, f"distribution {name!r} is not supported "
                         "(gaussian, normal, log_gaussian, log_normal, uniform)")
        self.check_refs(d.a, block, a.line, local_consts)
        self.check_refs(d.b, block, a.line, local_consts)
        return d

    def check_refs(self, e, block, line, local_consts, seen=()):
        """Every name resolves, to a variable the block may read (Bi::Block::*::validate)."""
        if isinstance(e, Ref):
            if e.index is not None:
                _err(line, f"indexed reference {e.name}[...] needs dimensions, not supported")
            n = e.name
            if n in local_consts or n in self.consts:
                return
            if n in self.inlines:
                if n in seen:
                    _err(line, f"inline {n} refers to itself")
                return self.check_refs(self.inlines[n], block, line, local_consts, seen + (n,))
            if n not in self.kind:
                _err(line, f"unknown name {n!r}")
            t = self.kind[n]
            allowed = {"initial": ("param", "input", "state"),
                       "transition": ("param", "input", "state", "noise"),
                       "ode": ("param", "input", "state", "noise"),
                       "observation": ("param", "input", "state")}[block]
            if t not in allowed:
                _err(line, f"{t} variable {n} cannot be read in the {block} block")
        elif isinstance(e, Unary):
            self.check_refs(e.arg, block, line, local_consts, seen)
        elif isinstance(e, Binary):
            self.check_refs(e.left, block, line, local_consts, seen)
            self.check_refs(e.right, block, line, local_consts, seen)
        elif isinstance(e, Ternary):
            for x in (e.cond, e.then, e.other):
                self.check_refs(x, block, line, local_consts, seen)
        elif isinstance(e, Call):
            if e.name not in FUNCS:
                _err(line, f"unknown function {e.name!r}")
            if e.named:
                _err(line, f"function {e.name} takes no named arguments")
            for x in e.args:
                self.check_refs(x, block, line, local_consts, seen)
        elif isinstance(e, Str):
            _err(line, "a string cannot appear in an expression")

    def local_decls(self, b: Block, local_consts):
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                kind, n, e = it
                if kind == "const":
                    local_consts[n] = _const_eval(e, {**self.consts, **local_consts}, b.line)
                else:
                    self.inlines[n] = e

    def steps(self, b: Block, block, counts, local_consts, targets):
        out = []
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                continue
            if isinstance(it, Block):
                _err(it.line, f"nested block {it.spec.name or '{...}'} is not supported here")
            a = it
            if a.target_dims is not None:
                _err(a.line, f"indexed target {a.target}[...] needs dimensions, not supported")
            t = self.kind.get(a.target)
            if t not in targets:
                _err(a.line, f"{a.target} cannot be the target of an action in the {block} block")
            if a.op == "~":
                out.append(Step("draw", a.target, draw=self.draw(a, block, counts, local_consts)))
            elif a.op == "<-":
                self.check_refs(a.expr, block, a.line, local_consts)
                out.append(Step("set", a.target, expr=a.expr))
            else:
                _err(a.line, "an ODE equation must be inside an ode block")
        return out

    def run(self) -> GenModel:
        m = self.m
        consts_all = dict(self.consts)
        g = GenModel(m.name, m.names_of("param"), m.names_of("input"), m.names_of("state"),
                     m.names_of("noise"), m.names_of("obs"), consts_all, self.inlines,
                     [], [], {}, [], [])
        if not g.states:
            raise GenError("the model has no state variables")
        if len(g.slots) > MAX_SLOTS:
            raise GenError(f"{len(g.slots)} params and inputs; at most {MAX_SLOTS}")
        for b in m.blocks:
            if b.spec.name not in ("initial", "transition", "observation", "parameter",
                                   "proposal_parameter", "proposal_initial",
                                   "lookahead_transition", "lookahead_observation", "bridge"):
                _err(b.line, f"unknown block {b.spec.name!r}")
        # sub initial
        b = m.block("initial")
        if b is not None:
            lc = {}
            self.local_decls(b, lc)
            counts = {}
            g.initial = self.steps(b, "initial", counts, lc, ("state",))
            g.consts.update(lc)
            g.normals["initial"] = counts.get("normal", 0)
            g.uniforms["initial"] = counts.get("uniform", 0)
        # sub transition(delta = 1.0)
        b = m.block("transition")
        if b is not None:
            g.delta = self.block_arg(b, "delta", 0, 1.0)
            if not g.delta > 0:
                _err(b.line, "delta must be positive")
            lc = {}
            self.local_decls(b, lc)
            counts = {}
            cur = g.pre
            for it in _flatten(b.items):
                if isinstance(it, tuple):
                    continue
                if isinstance(it, Block):
                    if it.spec.name != "ode":
                        _err(it.line, f"block {it.spec.name or '{...}'} inside transition "
                                      "is not supported (only ode)")
                    if g.has_ode:
                        _err(it.line, "at most one ode block per transition")
                    self.ode_block(it, g, lc)
                    cur = g.post
                    continue
                cur.extend(self.steps(Block(b.spec, [it], b.line), "transition", counts, lc,
                                      ("state", "noise")))
            g.consts.update(lc)
            g.normals["transition"] = counts.get("normal", 0)
            g.uniforms["transition"] = counts.get("uniform", 0)
        # sub observation
        b = m.block("observation")
        if b is not None:
            lc = {}
            self.local_decls(b, lc)
            counts = {}
            for s in self.steps(b, "observation", counts, lc, ("obs",)):
                if s.kind != "draw":
                    _err(0, "only '~' actions belong in the observation block")
                g.observation.append(s.draw)
            g.consts.update(lc)
            g.normals["observation"] = counts.get("normal", 0)
            g.uniforms["observation"] = counts.get("uniform", 0)
            seen = [d.target for d in g.observation]
            if len(set(seen)) != len(seen):
                _err(b.line, "an obs variable is the target of two actions")
        return g

    def block_arg(self, b, name, pos, default):
        if len(b.spec.args) > pos:
            return _const_eval(b.spec.args[pos], self.consts, b.line)
        for n, v in b.spec.named:
            if n == name:
                return _const_eval(v, self.consts, b.line)
        return default

    def ode_block(self, b: Block, g: GenModel, lc):
        names = ("alg", "h", "atoler", "rtoler")
        vals = {}
        for n, v in zip(names, b.spec.args):
            vals[n] = v
        for n, v in b.spec.named:
            if n not in names:
                _err(b.line, f"ode has no argument {n!r}")
            vals[n] = v
        alg = vals.get("alg", Str("RK4(3)"))
        if not isinstance(alg, Str):
            _err(b.line, "alg must be a string")
        if alg.value not in ALGS:
            _err(b.line, f"ode alg {alg.value!r} is not built: give alg = 'RK4' or 'RK5(4)'"
                         + (" (the default is 'RK4(3)')" if "alg" not in vals else ""))
        g.integrator = ALGS[alg.value]
        env = {**self.consts, **lc}
        g.h = _const_eval(vals["h"], env, b.line) if "h" in vals else 1.0
        g.atol = _const_eval(vals["atoler"], env, b.line) if "atoler" in vals else 1e-3
        g.rtol = _const_eval(vals["rtoler"], env, b.line) if "rtoler" in vals else 1e-3
        g.has_ode = True
        self.local_decls(b, lc)
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                continue
            if isinstance(it, Block) or it.op != "=":
                _err(getattr(it, "line", b.line), "an ode block holds only dx/dt = ... equations")
            if it.target_dims is not None:
                _err(it.line, "indexed ODE targets need dimensions, not supported")
            if self.kind.get(it.target) != "state":
                _err(it.line, f"{it.target} is not a state variable")
            if it.target in g.ode:
                _err(it.line, f"two equations for d{it.target}/dt")
            self.check_refs(it.expr, "ode", it.line, lc)
            g.ode[it.target] = it.expr


def analyse(model: Model) -> GenModel:
    return _Analyser(model).run()


# ---------------------------------------------------------------- emission
def _lit(v):
    if math.isinf(v):
        return "__builtin_inf()" if v > 0 else "(-__builtin_inf())"
    r = repr(float(v))
    return r


class _Emitter:
    """Expressions in one arithmetic type: 'T' (the state type) or 'double'."""

    def __init__(self, g: GenModel, ty: str, xname: str, xdouble=False):
        self.g, self.ty, self.x, self.xdouble = g, ty, xname, xdouble

    def num(self, v):
        return f"T({_lit(v)})" if self.ty == "T" else _lit(v)

    def expr(self, e, local=()):
        g = self.g
        if isinstance(e, Num):
            return self.num(e.value)
        if isinstance(e, Ref):
            n = e.name
            if n in g.consts:
                return self.num(g.consts[n])
            if n in g.inlines and n not in local:
                return "(" + self.expr(g.inlines[n], local + (n,)) + ")"
            if n in g.slots:
                v = f"p_{n}"
                return f"T({v})" if self.ty == "T" else v
            if n in g.states:
                v = f"{self.x}[{g.states.index(n)}]"
                if self.ty == "double" and not self.xdouble:
                    return f"(double){v}"
                return v
            if n in g.noises:
                v = f"w_{n}"
                return f"(double){v}" if self.ty == "double" else v
            raise GenError(f"cannot read {n!r} here")
        if isinstance(e, Unary):
            a = self.expr(e.arg, local)
            return f"({e.op}{a})"
        if isinstance(e, Binary):
            a, b = self.expr(e.left, local), self.expr(e.right, local)
            if e.op == "**":
                return f"pow({a}, {b})"
            if e.op == "%":
                return f"fmod({a}, {b})"
            if e.op in ("<", ">", "<=", ">=", "==", "!=", "&&", "||"):
                return f"{self.ty}(({a}) {e.op} ({b}))"
            return f"({a} {e.op} {b})"
        if isinstance(e, Ternary):
            return (f"(({self.expr(e.cond, local)}) != {self.num(0.0)} ? "
                    f"{self.expr(e.then, local)} : {self.expr(e.other, local)})")
        if isinstance(e, Call):
            # a function of constants is evaluated here, correctly rounded in double
            # (the reference evaluates constant expressions when it generates code)
            try:
                return self.num(_const_eval(e, g.consts, 0))
            except GenError:
                pass
            args = ", ".join(self.expr(a, local) for a in e.args)
            return f"{FUNCS[e.name]}({args})"
        raise GenError(f"cannot emit {e!r}")


def _draw_value(em: _Emitter, d: Draw, zname: str, uname: str):
    """Value of a draw (Bi::Action::gaussian / uniform `sample`, share/tt/cpp/action/*.hpp.tt)."""
    a, b = em.expr(d.a), em.expr(d.b)
    if d.dist == "gaussian":
        z = zname if em.ty == "T" else f"(double){zname}"
        v = f"({a} + ({b} * {z}))"
        return f"exp({v})" if d.log else v
    u = uname if em.ty == "T" else f"(double){uname}"
    return f"({a} + (({b} - {a}) * {u}))"


def _zname(d):
    return f"zq{d.index // 4}[{d.index % 4}]"


def generate_model(g: GenModel) -> str:
    """The device functor of one model."""
    S = len(g.states)
    out = []
    w = out.append
    w(f"// ---- model {g.name}: generated from its .bi file")
    w("template <class T>")
    w(f"struct {g.cname} {{")
    w("  using real_t = T;")
    w(f"  static constexpr int D = {S}, LPP = 1, DL = {S}, NOBS = {len(g.obs)};")
    w("  static constexpr bool NOISE_FIRST = false, HAS_FORCING = false, GENERATED = true;")
    w(f"  static constexpr double DELTA = {_lit(g.delta)};")
    for n in g.slots:
        w(f"  double p_{n};")
    for n in g.noises:
        w(f"  T w_{n};")
    init = ", ".join([f"p_{n}(a.par[{i}])" for i, n in enumerate(g.slots)]
                     + [f"w_{n}(T(0))" for n in g.noises])
    w(f"  __device__ __forceinline__ explicit {g.cname}(const ModelArgs& a)"
      + (f" : {init}" if init else "") + " {}")
    # rhs: the ode block (share/tt/cpp/action/ode_.hpp.tt); states without an equation stay put
    em = _Emitter(g, "T", "x")
    w("  template <class G>")
    w("  __device__ __forceinline__ void rhs(const G&, const T (&x)[DL], T (&k)[DL]) const {")
    for j, s in enumerate(g.states):
        w(f"    k[{j}] = {em.expr(g.ode[s]) if s in g.ode else 'T(0)'};")
    w("  }")

    # discrete part of the transition: actions in the order written
    def discrete(steps, ind):
        lines = []
        for st in steps:
            tgt = (f"x[{g.states.index(st.target)}]" if st.target in g.states
                   else f"w_{st.target}")
            if st.kind == "draw":
                lines.append(f"{ind}{tgt} = {_draw_value(em, st.draw, _zname(st.draw), f'u{st.draw.index}')};")
            else:
                lines.append(f"{ind}{tgt} = {em.expr(st.expr)};")
        return lines

    nn_t, nu_t = g.normals.get("transition", 0), g.uniforms.get("transition", 0)
    nbt = (nn_t + 3) // 4
    w("  // one interval [t1, t2]: discrete actions at multiples of DELTA, the ode in between")
    w("  template <int INTEG, class G>")
    w("  __device__ __forceinline__ int transition(const G& grp, T (&x)[DL], double t1, double t2,")
    w("                                            const ModelArgs& ma, uint64_t seed,")
    w("                                            uint32_t batch, uint32_t p, uint32_t step) {")
    w("    int n = 0;")
    w("    double sa = gen_scaled(t1, DELTA);")
    w("    const double s2 = gen_scaled(t2, DELTA);")
    w("    double ta = t1;")
    w("    uint32_t fire = 0;")
    w("    while (sa < s2) {")
    w("      const double sb = fmin(floor(sa) + 1.0, s2);")
    w("      const double tb = sb == s2 ? t2 : sb * DELTA;")
    w("      const bool on = sa == floor(sa);")
    has_discrete = bool(g.pre or g.post)
    if has_discrete:
        w("      if (on) {")
        for b in range(nbt):
            w(f"        normals4<T>(seed, p, step, TAG_NOISE, batch, fire * {nbt}u + {b}u, zq{b});")
        for k in range(nu_t):
            w(f"        u{k} = uniform_word<T>(philox_block(seed, p, step, TAG_NOISE, batch, "
              f"{UNIFORM_BLOCK:#x}u + fire * {nu_t}u + {k}u));")
        for ln in discrete(g.pre, "        "):
            w(ln)
        w("      }")
    if g.has_ode:
        w("      if constexpr (INTEG == BIHIP_RK4) n += integrate_rk4(*this, grp, x, ta, tb, ma.h0);")
        w("      else n += integrate_dopri5(*this, grp, x, ta, tb, ma);")
    if g.post:
        w("      if (on) {")
        for ln in discrete(g.post, "        "):
            w(ln)
        w("      }")
    w("      fire += on ? 1u : 0u;")
    w("      sa = sb;")
    w("      ta = tb;")
    w("    }")
    w("    return n;")
    w("  }")
    # the draw buffers are members, shared by the discrete parts before and after the ode
    decl = [f"  T zq{b}[4];" for b in range(nbt)] + [f"  T u{k};" for k in range(nu_t)]
    idx = out.index("  // one interval [t1, t2]: discrete actions at multiples of DELTA, the ode in between")
    out[idx:idx] = decl

    # observation log-density (share/tt/cpp/action/gaussian.hpp.tt:48-73, uniform.hpp.tt:38-52)
    eo = _Emitter(g, "double", "x")
    w("  __device__ __forceinline__ double obs(const T (&x)[DL], const double* y) const {")
    w("    double lp = 0.0;")
    for d in g.observation:
        k = g.obs.index(d.target)
        a, b = eo.expr(d.a), eo.expr(d.b)
        w("    {")
        w(f"      const double xy = y[{k}];")
        if d.dist == "gaussian":
            w(f"      const double mu = {a};")
            w(f"      const double sg = {b};")
            if d.log:
                w("      const double z = (log(xy) - mu) / sg;")
                w("      lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sg * xy);")
            else:
                w("      if (sg == 0.0) {")
                w("        lp = xy == mu ? __builtin_inf() : -__builtin_inf();")
                w("      } else {")
                w("        const double z = (xy - mu) / sg;")
                w("        lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sg);")
                w("      }")
        else:
            w(f"      const double mn = {a}, mx = {b};")
            w("      if (xy >= mn && xy <= mx) lp += -log(mx - mn);")
            w("      else lp = -__builtin_inf();")
        w("    }")
    w("    return lp;")
    w("  }")

    # observation samples (Model::observationSamples), y has NOBS entries
    no, nuo = g.normals.get("observation", 0), g.uniforms.get("observation", 0)
    w("  __device__ __forceinline__ void obs_sample(const T (&x)[DL], uint64_t seed, uint32_t batch,")
    w("                                             uint32_t p, uint32_t step, double* y) const {")
    for b in range((no + 3) // 4):
        w(f"    T zq{b}[4];")
        w(f"    normals4<T>(seed, p, step, TAG_OBS, batch, {b}u, zq{b});")
    for k in range(nuo):
        w(f"    const T u{k} = uniform_word<T>(philox_block(seed, p, step, TAG_OBS, batch, "
          f"{UNIFORM_BLOCK:#x}u + {k}u));")
    for j in range(len(g.obs)):
        w(f"    y[{j}] = __builtin_nan(\"\");")
    for d in g.observation:
        w(f"    y[{g.obs.index(d.target)}] = {_draw_value(eo, d, _zname(d), f'u{d.index}')};")
    w("  }")

    # initial states, in double (share/tt/cpp/block/initial.hpp.tt)
    ni, nui = g.normals.get("initial", 0), g.uniforms.get("initial", 0)
    ei = _Emitter(g, "double", "xd", xdouble=True)
    w("  __device__ __forceinline__ void init(double (&xd)[DL], uint64_t seed, uint32_t batch,")
    w("                                       uint32_t p) const {")
    for j in range(S):
        w(f"    xd[{j}] = 0.0;")
    for b in range((ni + 3) // 4):
        w(f"    T zq{b}[4];")
        w(f"    normals4<T>(seed, p, 0u, TAG_INIT, batch, {b}u, zq{b});")
    for k in range(nui):
        w(f"    const T u{k} = uniform_word<T>(philox_block(seed, p, 0u, TAG_INIT, batch, "
          f"{UNIFORM_BLOCK:#x}u + {k}u));")
    for st in g.initial:
        j = g.states.index(st.target)
        if st.kind == "draw":
            w(f"    xd[{j}] = {_draw_value(ei, st.draw, _zname(st.draw), f'u{st.draw.index}')};")
        else:
            w(f"    xd[{j}] = {ei.expr(st.expr)};")
    w("  }")
    w("};")
    return "\n".join(out)


def _cstr_list(names):
    return "{" + ", ".join(f'"{n}"' for n in names) + (", " if names else "") + "nullptr}"


def generate(models: List[GenModel]) -> str:
    """The header compiled into the library: every functor, their host-side
    descriptions and the dispatch over model ids."""
    out = ["// Generated by libbi_amd.dsl from the model files in libbi_amd/bi; do not edit.",
           "#pragma once", "", "namespace bihip {", ""]
    for g in models:
        out.append(generate_model(g))
        out.append("")
    out.append(f"constexpr int GEN_COUNT = {len(models)};")
    out.append("static const char* const GEN_NONE[] = {nullptr};")
    for i, g in enumerate(models):
        out.append(f"static const char* const GEN_{i}_PARAMS[] = {_cstr_list(g.params)};")
        out.append(f"static const char* const GEN_{i}_INPUTS[] = {_cstr_list(g.inputs)};")
        out.append(f"static const char* const GEN_{i}_STATES[] = {_cstr_list(g.states)};")
        out.append(f"static const char* const GEN_{i}_OBS[] = {_cstr_list(g.obs)};")
        out.append(f"static const char* const GEN_{i}_NOISES[] = {_cstr_list(g.noises)};")
    out.append("static const GenInfo GEN_MODELS[] = {")
    for i, g in enumerate(models):
        out.append(f'    {{"{g.name}", {len(g.states)}, {len(g.obs)}, {g.integrator}, '
                   f"{_lit(g.h)}, {_lit(g.atol)}, {_lit(g.rtol)}, {_lit(g.delta)}, "
                   f"GEN_{i}_PARAMS, GEN_{i}_INPUTS, GEN_{i}_STATES, GEN_{i}_OBS, "
                   f"GEN_{i}_NOISES}},")
    if not models:
        out.append('    {"", 0, 0, 0, 0.0, 0.0, 0.0, 1.0, GEN_NONE, GEN_NONE, GEN_NONE, GEN_NONE, '
                   "GEN_NONE},")
    out.append("};")
    out.append("")
    out.append("// F::template run<M>() for the generated model of m")
    out.append("template <class F>")
    out.append("int with_generated(const bihip_model* m, F& f) {")
    out.append("  const bool f64 = m->precision == BIHIP_F64;")
    out.append("  switch (m->model - BIHIP_MODEL_GENERATED) {")
    for i, g in enumerate(models):
        out.append(f"    case {i}:")
        out.append(f"      return f64 ? f.template run<{g.cname}<double>>() "
                   f": f.template run<{g.cname}<float>>();")
    out.append("  }")
    out.append('  return set_error(BIHIP_EINVAL, "unknown generated model");')
    out.append("}")
    out.append("")
    out.append("}  // namespace bihip")
    return "\n".join(out) + "\n"


def load_models(paths) -> List[GenModel]:
    out = []
    for p in sorted(paths):
        try:
            g = analyse(parse_file(p))
        except (BiSyntaxError, GenError) as e:
            raise type(e)(f"{p}: {e}") from None
        if any(g.name == o.name for o in out):
            raise GenError(f"{p}: a model named {g.name} already exists")
        out.append(g)
    return out


def analyse_text(text: str) -> GenModel:
    return analyse(parse(text))
