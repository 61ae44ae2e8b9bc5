"""HIP code generator for models written in LibBi's model language.

The reference turns a `.bi` file into C++/CUDA through its Perl front end and
templates: one class per block and per action (share/tt/cpp/model.hpp.tt,
share/tt/cpp/block/*.hpp.tt, share/tt/cpp/action/*.hpp.tt), compiled into
the program that runs the filter.  Here a model becomes one device functor
with the interface of the hand-written models (libbi_amd/csrc/models.hpp):
`rhs` for the integrators, and `transition`, `obs`, `obs_sample` and `init`
for the filter kernels, which are instantiated on it when the library is
built.  One lane holds one particle's whole state.

What a generated model supports (checked here, with an error that names the
line otherwise):
  * `dim` declarations (`dim n(size = 8, boundary = 'cyclic')`; boundaries
    'none', 'cyclic' and 'extended') and variables over at most one
    dimension; `param` and `input` (the model's parameter slots, in that
    order), `state`, `noise`, `obs`; `const` and `inline`;
  * actions on whole vectors (`x <- x + w`, element by element) or on
    elements named by a dimension alias (`dx[i]/dt = x[i-1]*(x[i+1] - x[i-2])`),
    with constant index arithmetic on the aliases; a vector action reads the
    values from before it (its elements are computed, then stored);
  * `sub initial`: `~` and `<-` actions on state variables;
  * `sub transition(delta)`: `~` and `<-` actions and at most one
    `ode(alg, h, atoler, rtoler)` block of `dx/dt = ...` equations, run in the
    order written; discrete actions fire at multiples of delta
    (share/src/bi/state/Schedule.hpp:230-292), the ODE covers every interval;
  * `sub observation`: `~` actions on obs variables, with the log-densities of
    gaussian/normal, log_gaussian/log_normal, uniform, exponential, gamma,
    inverse_gamma, beta, poisson, negbin and binomial (lib/Bi/Action/*.pm
    arguments, positional or named);
  * draws (initial, transition, observation samples) from all of them:
    gaussian from normal words, uniform and exponential from a uniform, the
    others by the rejection samplers of libbi_amd/csrc/gen_samplers.hpp, each
    draw with a keyed Philox stream of its own;
  * `sub parameter`: priors with constant arguments (gaussian, log_gaussian,
    uniform, exponential, gamma, inverse_gamma, beta), written as a host table
    for SMC^2 and PMMH (generate_priors);
  * the ODE algorithms 'RK4' and 'RK5(4)' (share/tt/cpp/block/ode.hpp.tt).
Blocks the samplers here do not run (`proposal_*`, `lookahead_*`, `bridge`)
are parsed and skipped.

Random draws are keyed Philox counters: the k-th normal draw of a block
(counting vector elements) comes from block k/4, word pair k%4 of the
(particle, step, tag, batch) stream, as the hand-written models draw theirs,
so a model written here reproduces the built-in model with the same
equations bit for bit.
"""
import math
import re
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from .parser import (Action, Binary, Block, BiSyntaxError, Call, Model, Num, Ref, Str, Ternary,
                     Unary, parse, parse_file)

MAX_SLOTS = 16  # bihip_model::par
MAX_STATE = 64  # state elements of one particle (one lane's registers)
NORMAL_DISTS = {"gaussian": False, "normal": False, "log_gaussian": True, "log_normal": True}
# (argument name, default) in position order (lib/Bi/Action/*.pm)
DIST_ARGS = {"gaussian": (("mean", 0.0), ("std", 1.0)),
             "uniform": (("lower", 0.0), ("upper", 1.0)),
             "exponential": (("lambda", 1.0),),
             "gamma": (("shape", 1.0), ("scale", 1.0)),
             "inverse_gamma": (("shape", 1.0), ("scale", 1.0)),
             "beta": (("alpha", 1.0), ("beta", 1.0)),
             "poisson": (("rate", 1.0),),
             "negbin": (("mean", 1.0), ("shape", 1.0)),
             "binomial": (("size", 1.0), ("prob", 0.5))}
# samplers: normal words (gaussian), one uniform word pair (uniform, exponential), or a
# keyed stream of words per draw (the rejection samplers of gen_samplers.hpp)
SAMPLED = ("gaussian", "uniform", "exponential", "gamma", "inverse_gamma", "beta", "poisson",
           "negbin", "binomial")
STREAM_DISTS = ("gamma", "inverse_gamma", "beta", "poisson", "negbin", "binomial")
STREAM_BLOCK = 0x20000000  # Philox block of the first stream draw; 4096 blocks per draw
# priors of the parameter block, sampled and evaluated on the host (libbi_amd/csrc/prior.cpp)
PRIOR_CODE = {"gaussian": 0, "log_gaussian": 1, "uniform": 2, "exponential": 3, "gamma": 4,
              "inverse_gamma": 5, "beta": 6}
PRIOR_NAMES = ("gaussian", "log_gaussian", "uniform", "exponential", "gamma", "inverse_gamma",
               "beta")
UNIFORM_BLOCK = 0x40000000  # Philox block offset of uniform draws
BOUNDARIES = ("none", "cyclic", "extended")
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
KIND_CODE = {"param": 0, "input": 1, "state": 2, "obs": 3, "noise": 4}  # BIHIP_VAR_*


class GenError(ValueError):
    """A model that parses but uses something the generator does not support."""


@dataclass
class Var:
    name: str
    kind: str  # param | input | state | noise | obs
    dim: Optional[str]  # None for a scalar
    size: int
    offset: int  # first element in its layout (par slots, state, obs values, noise)


@dataclass
class Draw:
    dist: str  # a key of DIST_ARGS
    log: bool  # log_gaussian / log_normal
    args: list  # argument expressions in position order
    index: int  # first draw of the block's draws of the same kind (one per element)


@dataclass
class Step:
    kind: str  # draw | set | ode
    target: str
    alias: Optional[str]  # dimension alias bound to the element index (x[i])
    expr: object = None
    draw: Optional[Draw] = None
    line: int = 0


@dataclass
class GenModel:
    name: str
    vars: Dict[str, Var]
    dims: Dict[str, tuple]  # name -> (size, boundary)
    consts: Dict[str, float]
    inlines: Dict[str, object]
    initial: List[Step] = field(default_factory=list)
    pre: List[Step] = field(default_factory=list)  # transition actions before the ode block
    ode: List[Step] = field(default_factory=list)
    post: List[Step] = field(default_factory=list)  # transition actions after the ode block
    observation: List[Step] = field(default_factory=list)
    delta: float = 1.0
    integrator: int = 0
    h: float = 1.0
    atol: float = 1e-3
    rtol: float = 1e-3
    has_ode: bool = False
    normals: Dict[str, int] = field(default_factory=dict)   # block -> normal draws
    uniforms: Dict[str, int] = field(default_factory=dict)  # block -> uniform draws
    streams: Dict[str, int] = field(default_factory=dict)   # block -> stream draws
    # sub parameter: (distribution, first argument, second argument) per param element
    priors: Dict[int, tuple] = field(default_factory=dict)
    lpp: int = 1  # lanes per particle

    def names(self, kind):
        return [v.name for v in self.vars.values() if v.kind == kind]

    def size(self, kind):
        return sum(v.size for v in self.vars.values() if v.kind == kind)

    @property
    def params(self):
        return self.names("param")

    @property
    def inputs(self):
        return self.names("input")

    @property
    def states(self):
        return self.names("state")

    @property
    def noises(self):
        return self.names("noise")

    @property
    def obs(self):
        return self.names("obs")

    @property
    def D(self):
        return self.size("state")

    @property
    def NOBS(self):
        return self.size("obs")

    @property
    def cname(self):
        return "Gen_" + self.name


# ---------------------------------------------------------------- analysis
def _err(line, msg):
    raise GenError(f"line {line}: {msg}")


def _args(call, spec, line):
    """Positional then named arguments of an action (Bi::Action::process_args)."""
    names = [n for n, _ in spec]
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
    for n, d in spec:
        out.setdefault(n, Num(d, repr(d), False))
    return [out[n] for n in names]


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
        if e.op == "/":
            return a / b
        if e.op == "**":
            return math.pow(a, b)
        if e.op == "%":
            return math.fmod(a, b)
        arith = {"+": a + b, "-": a - b, "*": a * b}
        if e.op in arith:
            return arith[e.op]
        cmp = {"<": a < b, ">": a > b, "<=": a <= b, ">=": a >= b, "==": a == b, "!=": a != b,
               "&&": bool(a) and bool(b), "||": bool(a) or bool(b)}
        return float(cmp[e.op])
    if isinstance(e, Ternary):
        return (_const_eval(e.then, consts, line) if _const_eval(e.cond, consts, line)
                else _const_eval(e.other, consts, line))
    if isinstance(e, Call) and e.name in PY_FUNCS and not e.named:
        return float(PY_FUNCS[e.name](*[_const_eval(a, consts, line) for a in e.args]))
    _err(line, "a constant expression is required here")


def _spec_arg(spec, name, pos, consts, default):
    if len(spec.args) > pos:
        v = spec.args[pos]
    else:
        v = dict(spec.named).get(name)
    if v is None:
        return default
    if isinstance(v, Str):
        return v.value
    return _const_eval(v, consts, spec.line)


class _Analyser:
    def __init__(self, m: Model):
        self.m = m
        self.consts = {}
        for n, e in m.consts:
            self.consts[n] = _const_eval(e, self.consts, 0)
        self.dims = {}
        for d in m.dims:
            size = _spec_arg(d, "size", 0, self.consts, None)
            if size is None or size != int(size) or size < 1:
                _err(d.line, f"dim {d.name} needs a positive integer size")
            bnd = _spec_arg(d, "boundary", 1, self.consts, "none")
            if bnd not in BOUNDARIES:
                _err(d.line, f"dim {d.name}: boundary must be one of {BOUNDARIES}")
            self.dims[d.name] = (int(size), bnd)
        self.vars = {}
        offs = {}
        for t, s in m.vars:
            if t in ("state_aux_", "param_aux_"):
                _err(s.line, f"{t} variables are not supported")
            if s.name in self.vars or s.name in self.dims:
                _err(s.line, f"{s.name} declared twice")
            if len(s.dims) > 1:
                _err(s.line, f"{s.name}: variables over more than one dimension are not supported")
            dim = s.dims[0] if s.dims else None
            if dim is not None and dim not in self.dims:
                _err(s.line, f"{s.name}: unknown dimension {dim!r}")
            size = self.dims[dim][0] if dim else 1
            lay = "par" if t in ("param", "input") else t
            self.vars[s.name] = Var(s.name, t, dim, size, 0)
            offs.setdefault(lay, []).append(s.name)
        # par slots: params then inputs; other kinds in declaration order
        for lay, kinds in (("par", ("param", "input")), ("state", ("state",)),
                           ("obs", ("obs",)), ("noise", ("noise",))):
            o = 0
            for k in kinds:
                for v in self.vars.values():
                    if v.kind == k:
                        v.offset = o
                        o += v.size
        self.inlines = dict(m.inlines)

    # ---- expressions
    def check_refs(self, e, block, line, lc, aliases, elem_dim, seen=()):
        """Every name resolves, to a variable the block may read
        (Bi::Block::*::validate); indices are constant expressions of the aliases."""
        if isinstance(e, Ref):
            n = e.name
            if n in aliases:
                if e.index is not None:
                    _err(line, f"dimension alias {n} cannot be indexed")
                return
            if n in lc or n in self.consts:
                if e.index is not None:
                    _err(line, f"constant {n} cannot be indexed")
                return
            if n in self.inlines:
                if n in seen:
                    _err(line, f"inline {n} refers to itself")
                if e.index is not None:
                    _err(line, f"inline {n} cannot be indexed")
                return self.check_refs(self.inlines[n], block, line, lc, aliases, elem_dim,
                                       seen + (n,))
            if n not in self.vars:
                _err(line, f"unknown name {n!r}")
            v = self.vars[n]
            allowed = {"parameter": (),
                       "initial": ("param", "input", "state"),
                       "transition": ("param", "input", "state", "noise"),
                       "ode": ("param", "input", "state", "noise"),
                       "observation": ("param", "input", "state")}[block]
            if v.kind not in allowed:
                _err(line, f"{v.kind} variable {n} cannot be read in the {block} block")
            if e.index is not None:
                if v.dim is None:
                    _err(line, f"{n} has no dimension to index")
                if len(e.index) != 1:
                    _err(line, f"{n} takes one index")
                env = {**self.consts, **lc, **{a: 0.0 for a in aliases}}
                _const_eval(e.index[0], env, line)  # a constant expression of the aliases
            elif v.dim is not None and v.dim != elem_dim:
                _err(line, f"vector {n} read without an index outside an action over "
                           f"its dimension {v.dim}")
        elif isinstance(e, Unary):
            self.check_refs(e.arg, block, line, lc, aliases, elem_dim, seen)
        elif isinstance(e, Binary):
            self.check_refs(e.left, block, line, lc, aliases, elem_dim, seen)
            self.check_refs(e.right, block, line, lc, aliases, elem_dim, seen)
        elif isinstance(e, Ternary):
            for x in (e.cond, e.then, e.other):
                self.check_refs(x, block, line, lc, aliases, elem_dim, seen)
        elif isinstance(e, Call):
            if e.name not in FUNCS:
                _err(line, f"unknown function {e.name!r}")
            if e.named:
                _err(line, f"function {e.name} takes no named arguments")
            for x in e.args:
                self.check_refs(x, block, line, lc, aliases, elem_dim, seen)
        elif isinstance(e, Str):
            _err(line, "a string cannot appear in an expression")

    def target(self, a: Action, block, targets):
        """(variable, alias) of an action's left side."""
        v = self.vars.get(a.target)
        if v is None or v.kind not in targets:
            _err(a.line, f"{a.target} cannot be the target of an action in the {block} block")
        alias = None
        if a.target_dims is not None:
            if v.dim is None:
                _err(a.line, f"{a.target} has no dimension to index")
            if len(a.target_dims) != 1 or not isinstance(a.target_dims[0], Ref) \
                    or a.target_dims[0].index is not None:
                _err(a.line, f"the target {a.target}[...] takes one dimension alias")
            alias = a.target_dims[0].name
            if alias in self.vars or alias in self.consts or alias in self.dims:
                _err(a.line, f"dimension alias {alias} hides a declared name")
        return v, alias

    def draw(self, a: Action, block, counts, lc, v, alias):
        if not isinstance(a.expr, Call):
            _err(a.line, "the right side of '~' must be a distribution")
        name = a.expr.name
        aliases = (alias,) if alias else ()
        log = False
        if name in NORMAL_DISTS:
            dist = "gaussian"
            log = NORMAL_DISTS[name]
            for n, val in a.expr.named:
                if n == "log":
                    log = bool(_const_eval(val, self.consts, a.line))
        elif name in DIST_ARGS:
            dist = name
        else:
            _err(a.line, f"distribution {name!r} is not supported "
                         f"({', '.join(sorted(set(DIST_ARGS) | set(NORMAL_DISTS)))})")
        args = _args(a.expr, DIST_ARGS[dist], a.line)
        kind = ("normal" if dist == "gaussian" else
                "stream" if dist in STREAM_DISTS else "uniform")
        k = counts.setdefault(kind, 0)
        counts[kind] += v.size
        d = Draw(dist, log, args, k)
        for x in d.args:
            self.check_refs(x, block, a.line, lc, aliases, v.dim)
        return d

    def local_decls(self, b: Block, lc):
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                kind, n, e = it
                if kind == "const":
                    lc[n] = _const_eval(e, {**self.consts, **lc}, b.line)
                else:
                    self.inlines[n] = e

    def step(self, a: Action, block, counts, lc, targets):
        v, alias = self.target(a, block, targets)
        aliases = (alias,) if alias else ()
        if a.op == "~":
            return Step("draw", v.name, alias, draw=self.draw(a, block, counts, lc, v, alias),
                        line=a.line)
        if a.op == "<-":
            self.check_refs(a.expr, block, a.line, lc, aliases, v.dim)
            return Step("set", v.name, alias, expr=a.expr, line=a.line)
        _err(a.line, "an ODE equation must be inside an ode block")

    def steps(self, b: Block, block, counts, lc, targets):
        out = []
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                continue
            if isinstance(it, Block):
                _err(it.line, f"nested block {it.spec.name or '{...}'} is not supported here")
            out.append(self.step(it, block, counts, lc, targets))
        return out

    def run(self) -> GenModel:
        m = self.m
        g = GenModel(m.name, self.vars, self.dims, dict(self.consts), self.inlines)
        if not g.states:
            raise GenError("the model has no state variables")
        if g.size("param") + g.size("input") > MAX_SLOTS:
            raise GenError(f"{g.size('param') + g.size('input')} param and input elements; "
                           f"at most {MAX_SLOTS}")
        if g.D > MAX_STATE:
            raise GenError(f"{g.D} state elements; at most {MAX_STATE} per particle")
        for b in m.blocks:
            if b.spec.name not in ("initial", "transition", "observation", "parameter",
                                   "proposal_parameter", "proposal_initial",
                                   "lookahead_transition", "lookahead_observation", "bridge"):
                _err(b.line, f"unknown block {b.spec.name!r}")
        # sub parameter: the prior of the params, constant arguments (SMC^2 and PMMH)
        b = m.block("parameter")
        if b is not None:
            lc = {}
            self.local_decls(b, lc)
            for it in _flatten(b.items):
                if isinstance(it, tuple):
                    continue
                if isinstance(it, Block) or it.op != "~":
                    _err(getattr(it, "line", b.line),
                         "the parameter block takes '~' actions only")
                v, alias = self.target(it, "parameter", ("param",))
                d = self.draw(it, "parameter", {}, lc, v, alias)
                if d.dist not in PRIOR_CODE:
                    _err(it.line, f"{d.dist} is not a prior here ({', '.join(PRIOR_NAMES)})")
                for i in range(v.size):
                    env = {**self.consts, **lc, **({alias: float(i)} if alias else {})}
                    args = [_const_eval(x, env, it.line) for x in d.args]
                    code = PRIOR_CODE["log_gaussian" if d.dist == "gaussian" and d.log
                                      else d.dist]
                    if v.offset + i in g.priors:
                        _err(it.line, f"{v.name} has two priors")
                    g.priors[v.offset + i] = (code, args[0], args[1] if len(args) > 1 else 0.0)
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
            g.streams["initial"] = counts.get("stream", 0)
        # sub transition(delta = 1.0)
        b = m.block("transition")
        if b is not None:
            g.delta = _spec_arg(b.spec, "delta", 0, self.consts, 1.0)
            if not (isinstance(g.delta, float) and g.delta > 0):
                _err(b.line, "delta must be a positive number")
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
                cur.append(self.step(it, "transition", counts, lc, ("state", "noise")))
            g.consts.update(lc)
            g.normals["transition"] = counts.get("normal", 0)
            g.uniforms["transition"] = counts.get("uniform", 0)
            g.streams["transition"] = counts.get("stream", 0)
        # sub observation
        b = m.block("observation")
        if b is not None:
            lc = {}
            self.local_decls(b, lc)
            counts = {}
            seen = []
            for s in self.steps(b, "observation", counts, lc, ("obs",)):
                if s.kind != "draw":
                    _err(s.line, "only '~' actions belong in the observation block")
                if s.target in seen:
                    _err(s.line, f"{s.target} is the target of two actions")
                seen.append(s.target)
                g.observation.append(s)
            g.consts.update(lc)
            g.normals["observation"] = counts.get("normal", 0)
            g.uniforms["observation"] = counts.get("uniform", 0)
            g.streams["observation"] = counts.get("stream", 0)
        return g

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
        done = set()
        for it in _flatten(b.items):
            if isinstance(it, tuple):
                continue
            if isinstance(it, Block) or it.op != "=":
                _err(getattr(it, "line", b.line), "an ode block holds only dx/dt = ... equations")
            v, alias = self.target(it, "ode", ("state",))
            if v.name in done:
                _err(it.line, f"two equations for d{v.name}/dt")
            done.add(v.name)
            self.check_refs(it.expr, "ode", it.line, lc, (alias,) if alias else (), v.dim)
            g.ode.append(Step("ode", v.name, alias, expr=it.expr, line=it.line))


def analyse(model: Model) -> GenModel:
    g = _Analyser(model).run()
    g.lpp = _lane_groups(g)
    generate_model(g)  # index ranges are resolved element by element when emitting
    return g


# ---------------------------------------------------------------- lane groups
class _NotLanes(Exception):
    pass


def _index_affine(g, e, alias, line):
    """(slope, intercept) of an index expression in the alias (alias absent: slope 0)."""
    f = lambda v: _const_eval(e, {**g.consts, **({alias: float(v)} if alias else {})}, line)
    b = f(0)
    if alias is None:
        return 0, int(b)
    a = f(1) - b
    if f(2) != 2 * a + b or a != int(a) or b != int(b):
        raise _NotLanes()
    return int(a), int(b)


def _refs(e, out):
    """Every Ref in an expression tree (inlines expanded by the caller)."""
    if isinstance(e, Ref):
        out.append(e)
    elif isinstance(e, Unary):
        _refs(e.arg, out)
    elif isinstance(e, Binary):
        _refs(e.left, out)
        _refs(e.right, out)
    elif isinstance(e, Ternary):
        for x in (e.cond, e.then, e.other):
            _refs(x, out)
    elif isinstance(e, Call):
        for x in e.args:
            _refs(x, out)
    return out


def _expand(g, e, seen=()):
    """The expression with inlines substituted."""
    if isinstance(e, Ref) and e.index is None and e.name in g.inlines and e.name not in seen:
        return _expand(g, g.inlines[e.name], seen + (e.name,))
    if isinstance(e, Unary):
        return Unary(e.op, _expand(g, e.arg, seen))
    if isinstance(e, Binary):
        return Binary(e.op, _expand(g, e.left, seen), _expand(g, e.right, seen))
    if isinstance(e, Ternary):
        return Ternary(*[_expand(g, x, seen) for x in (e.cond, e.then, e.other)])
    if isinstance(e, Call):
        return Call(e.name, [_expand(g, x, seen) for x in e.args], e.named)
    return e


def _step_exprs(st):
    return [st.expr] if st.kind != "draw" else list(st.draw.args)


def _lane_groups(g: GenModel) -> int:
    """Lanes per particle.  A model whose state is one vector over a cyclic
    dimension of at least 16 elements is split over a group of 8, 4 or 2
    lanes (the largest that leaves slices of at least 8), as the hand-written
    Lorenz-96 functor is, when its ODE reads the state at constant offsets of
    its alias (at most a slice away: the halo comes from the neighbouring
    lanes), its other actions work element by element, and each observation
    element reads the state at one index affine in its own alias.  Otherwise
    one lane holds the whole state."""
    states = [v for v in g.vars.values() if v.kind == "state"]
    if len(states) != 1 or states[0].dim is None:
        return 1
    x = states[0]
    D = x.size
    if g.dims[x.dim][1] != "cyclic" or D < 16:
        return 1
    lpp = next((n for n in (8, 4, 2) if D % n == 0 and D // n >= 8), 1)
    if lpp == 1:
        return 1
    DL = D // lpp
    try:
        # the ode: x at offsets |c| <= DL of the alias; noise element by element
        for st in g.ode + g.pre + g.post + g.initial:
            v = g.vars[st.target]
            ode = st.kind == "ode"
            if v.dim is not None and v.dim != x.dim:
                raise _NotLanes()
            for e in _step_exprs(st):
                for r in _refs(_expand(g, e), []):
                    u = g.vars.get(r.name)
                    if u is None:
                        continue  # an alias or a constant
                    if u.kind in ("param", "input"):
                        if u.dim is not None:
                            raise _NotLanes()
                        continue
                    if v.dim is None:  # a scalar target reads no element of a vector
                        if u.dim is not None:
                            raise _NotLanes()
                        continue
                    if u.dim is None:
                        continue
                    if u.dim != x.dim:
                        raise _NotLanes()
                    if r.index is None:
                        continue
                    a, c = _index_affine(g, r.index[0], st.alias, st.line)
                    if a != 1 or (c != 0 and not (ode and u.kind == "state" and abs(c) <= DL)):
                        raise _NotLanes()
        # observations: one state index affine in the obs element, inside the state
        for st in g.observation:
            v = g.vars[st.target]
            idx = set()
            for e in st.draw.args:
                for r in _refs(_expand(g, e), []):
                    u = g.vars.get(r.name)
                    if u is None or u.kind in ("param", "input"):
                        if u is not None and u.dim is not None:
                            raise _NotLanes()
                        continue
                    if r.index is None:
                        raise _NotLanes()
                    idx.add(_index_affine(g, r.index[0], st.alias, st.line))
            if len(idx) > 1:
                raise _NotLanes()
            a, b = idx.pop() if idx else (0, 0)
            if a < 0 or not all(0 <= a * j + b < D for j in range(v.size)):
                raise _NotLanes()
            if v.dim is not None and st.alias is None and idx:
                raise _NotLanes()
    except (_NotLanes, GenError):
        return 1
    return lpp


# ---------------------------------------------------------------- emission
def _lit(v):
    if math.isinf(v):
        return "__builtin_inf()" if v > 0 else "(-__builtin_inf())"
    return repr(float(v))


class _Emitter:
    """Expressions in one arithmetic type, 'T' (the state type) or 'double', for
    one element: bind maps dimension aliases to indices, elem is the element of
    an action over a whole vector (its unindexed vectors read that element)."""

    fold_bind = True  # aliases are numbers known when generating

    def __init__(self, g: GenModel, ty: str, xname: str, xdouble=False):
        self.g, self.ty, self.x, self.xdouble = g, ty, xname, xdouble

    def num(self, v):
        return f"T({_lit(v)})" if self.ty == "T" else _lit(v)

    def element(self, v: Var, e: Ref, bind, elem, line):
        if e.index is None:
            return 0 if v.dim is None else elem
        env = {**self.g.consts, **{k: float(i) for k, i in bind.items()}}
        f = _const_eval(e.index[0], env, line)
        if f != int(f):
            _err(line, f"index of {v.name} is not an integer")
        i = int(f)
        size, bnd = self.g.dims[v.dim]
        if bnd == "cyclic":
            return i % size
        if bnd == "extended":
            return min(max(i, 0), size - 1)
        if not 0 <= i < size:
            _err(line, f"index {i} of {v.name} is outside its dimension {v.dim} (size {size})")
        return i

    def expr(self, e, bind=None, elem=None, line=0, local=()):
        g = self.g
        bind = bind or {}
        if isinstance(e, Num):
            return self.num(e.value)
        if isinstance(e, Ref):
            n = e.name
            if n in g.consts:
                return self.num(g.consts[n])
            if n in g.inlines and n not in local and n not in bind:
                return "(" + self.expr(g.inlines[n], bind, elem, line, local + (n,)) + ")"
            return self.ref(e, bind, elem, line)
        return self.other(e, bind, elem, line, local)

    def ref(self, e, bind, elem, line):
        g, n = self.g, e.name
        if n in bind:
            return self.num(float(bind[n]))
        v = g.vars.get(n)
        if v is None:
            raise GenError(f"line {line}: cannot read {n!r} here")
        i = self.element(v, e, bind, elem, line)
        if v.kind in ("param", "input"):
            r = f"p_{n}" + (f"[{i}]" if v.dim else "")
            return f"pt_{n}" + (f"[{i}]" if v.dim else "") if self.ty == "T" else r
        if v.kind == "state":
            r = f"{self.x}[{v.offset + i}]"
            return f"(double){r}" if self.ty == "double" and not self.xdouble else r
        if v.kind == "noise":
            r = f"w_{n}" + (f"[{i}]" if v.dim else "")
            return f"(double){r}" if self.ty == "double" else r
        raise GenError(f"line {line}: cannot read {v.kind} {n} here")

    def other(self, e, bind, elem, line, local):
        g = self.g
        if isinstance(e, Unary):
            return f"({e.op}{self.expr(e.arg, bind, elem, line, local)})"
        if isinstance(e, Binary):
            a = self.expr(e.left, bind, elem, line, local)
            b = self.expr(e.right, bind, elem, line, local)
            if e.op == "**":
                return f"pow({a}, {b})"
            if e.op == "%":
                return f"fmod({a}, {b})"
            if e.op in ("<", ">", "<=", ">=", "==", "!=", "&&", "||"):
                return f"{self.ty}(({a}) {e.op} ({b}))"
            return f"({a} {e.op} {b})"
        if isinstance(e, Ternary):
            c = self.expr(e.cond, bind, elem, line, local)
            return (f"(({c}) != {self.num(0.0)} ? {self.expr(e.then, bind, elem, line, local)}"
                    f" : {self.expr(e.other, bind, elem, line, local)})")
        if isinstance(e, Call):
            # a function of constants is evaluated here, correctly rounded in double
            # (the reference evaluates constant expressions when it generates code)
            try:
                env = dict(g.consts)
                if self.fold_bind:
                    env.update({k: float(i) for k, i in bind.items()})
                return self.num(_const_eval(e, env, line))
            except GenError:
                pass
            args = ", ".join(self.expr(a, bind, elem, line, local) for a in e.args)
            return f"{FUNCS[e.name]}({args})"
        raise GenError(f"line {line}: cannot emit {e!r}")


def _logdensity(d: Draw, a, lsg=None):
    """Statements adding the log-density of xy to lp, in double
    (share/tt/cpp/action/*.hpp.tt `logdensity`, share/src/bi/pdf/functor.hpp);
    lsg names log(sd) of a gaussian when it is the same for every element."""
    if d.dist == "gaussian" and d.log:
        return [f"const double mu = {a[0]};", f"const double sg = {a[1]};",
                "const double z = (log(xy) - mu) / sg;",
                "lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - log(sg * xy);"]
    if d.dist == "gaussian":
        return [f"const double mu = {a[0]};", f"const double sg = {a[1]};",
                "if (sg == 0.0) {",
                "  lp = xy == mu ? __builtin_inf() : -__builtin_inf();",
                "} else {",
                "  const double z = (xy - mu) / sg;",
                f"  lp += -0.5 * (z * z) - HALF_LOG_TWO_PI - {lsg or 'log(sg)'};",
                "}"]
    if d.dist == "uniform":
        return [f"const double mn = {a[0]}, mx = {a[1]};",
                "if (xy >= mn && xy <= mx) lp += -log(mx - mn);",
                "else lp = -__builtin_inf();"]
    if d.dist == "exponential":
        return [f"const double la = {a[0]};", "lp += log(la) - la * xy;"]
    if d.dist == "gamma":
        return [f"const double sh = {a[0]}, sc = {a[1]};",
                "lp += (sh - 1.0) * log(xy) - xy / sc - (lgamma(sh) + sh * log(sc));"]
    if d.dist == "inverse_gamma":  # the gamma log-density of 1/xy with scale 1/sc, - 2 log xy
        return [f"const double sh = {a[0]}, sc = 1.0 / ({a[1]});",
                "const double ix = 1.0 / xy;",
                "lp += (sh - 1.0) * log(ix) - ix / sc - (lgamma(sh) + sh * log(sc)) - "
                "2.0 * log(xy);"]
    if d.dist == "beta":
        return [f"const double al = {a[0]}, be = {a[1]};",
                "lp += (al - 1.0) * log(xy) + (be - 1.0) * log(1.0 - xy) - "
                "(lgamma(al) + lgamma(be) - lgamma(al + be));"]
    if d.dist == "poisson":
        return [f"const double la = {a[0]};",
                "if (la == 0.0) lp += xy == 0.0 ? 0.0 : -__builtin_inf();",
                "else lp += xy * log(la) - la - lgamma(xy + 1.0);"]
    if d.dist == "negbin":
        return [f"const double mu = {a[0]}, k = {a[1]};",
                "if (xy == 0.0 && k == 0.0) {",
                "} else if (xy == 0.0) {",
                "  lp += k * (k < mu ? log(k / (k + mu)) : log1p(-mu / (k + mu)));",
                "} else if (xy < 1e-10 * k) {",
                "  const double q = k < mu ? log(k / (1.0 + k / mu)) : log(mu / (1.0 + mu / k));",
                "  lp += xy * q - mu - lgamma(xy + 1.0) + log1p(xy * (xy - 1.0) / (2.0 * k));",
                "} else {",
                "  double r = lgamma(k + xy + 1.0) - lgamma(k + 1.0) - lgamma(xy + 1.0);",
                "  if (k > 0.0) r += k * log(k / (k + mu));",
                "  r += xy * log(mu / (k + mu));",
                "  lp += r + log(k / (k + xy));",
                "}"]
    if d.dist == "binomial":
        return [f"const double n = {a[0]}, pr = {a[1]};",
                "if (!(xy == 0.0 && n == 0.0)) {",
                "  double r = lgamma(n + 1.0) - lgamma(xy + 1.0) - lgamma(n - xy + 1.0);",
                "  if (xy > 0.0) r += xy * log(pr);",
                "  if (xy < n) r += (n - xy) * log1p(-pr);",
                "  lp += r;",
                "}"]
    raise GenError(f"no log-density for {d.dist}")


def _hoist(args_by_elem, key, w, ind, local=None):
    """Arguments that are the same for every element are computed once: returns
    the argument strings to use and the name of log(sd) when the second
    argument of a gaussian is invariant (None otherwise).  An argument that
    names the variable `local` (declared per element) is never hoisted."""
    first = args_by_elem[0]
    same = [all(a[i] == first[i] for a in args_by_elem) and
            (local is None or not re.search(rf"\b{local}\b", first[i]))
            for i in range(len(first))]
    if len(args_by_elem) < 2:
        return first, [False] * len(first)
    names = []
    for i, (arg, inv) in enumerate(zip(first, same)):
        if inv:
            w(f"{ind}const double h{key}_{i} = {arg};")
        names.append(f"h{key}_{i}" if inv else None)
    return names, same


def _elements(g: GenModel, st: Step):
    """(element, bind) pairs an action covers."""
    v = g.vars[st.target]
    for i in range(v.size):
        yield i, ({st.alias: i} if st.alias else {})


def _zname(k):
    return f"zq{k // 4}[{k % 4}]"


def _draw_value(em: _Emitter, d: Draw, i, bind, line, sctx=None):
    """Value of element i of a draw (the actions' `sample`,
    share/tt/cpp/action/*.hpp.tt).  sctx = (step, tag, first stream index) of
    the block's keyed streams for the rejection samplers."""
    a = [em.expr(x, bind, i, line) for x in d.args]
    if d.dist in STREAM_DISTS:
        step, tag, first = sctx
        ad = ", ".join(f"(double)({x})" for x in a)
        v = (f"gen_{d.dist}({ad}, DrawStream(seed, p, {step}, {tag}, batch, "
             f"{STREAM_BLOCK:#x}u + (({first} + {d.index + i}u) << 12)))")
        return f"T({v})" if em.ty == "T" else v
    if d.dist == "gaussian":
        z = _zname(d.index + i)
        z = z if em.ty == "T" else f"(double){z}"
        v = f"({a[0]} + ({a[1]} * {z}))"
        return f"exp({v})" if d.log else v
    u = f"u{d.index + i}"
    u = u if em.ty == "T" else f"(double){u}"
    if d.dist == "exponential":  # inverse distribution function
        one = em.num(1.0)
        return f"((-log({one} - {u})) / {a[0]})"
    return f"({a[0]} + (({a[1]} - {a[0]}) * {u}))"


def _target(g: GenModel, name, i, xname):
    v = g.vars[name]
    if v.kind == "state":
        return f"{xname}[{v.offset + i}]"
    return f"w_{name}" + (f"[{i}]" if v.dim else "")


def _actions(g, steps, em, xname, ind, tmp_ty, sctx=None):
    """Statements of discrete actions, in order; a vector action computes all
    its elements before it stores them."""
    lines = []
    for st in steps:
        v = g.vars[st.target]
        vals = []
        for i, bind in _elements(g, st):
            if st.kind == "draw":
                vals.append(_draw_value(em, st.draw, i, bind, st.line, sctx))
            else:
                vals.append(em.expr(st.expr, bind, i, st.line))
        if v.dim is None:
            lines.append(f"{ind}{_target(g, st.target, 0, xname)} = {vals[0]};")
        else:
            lines.append(f"{ind}{{")
            lines.append(f"{ind}  {tmp_ty} t_[{v.size}];")
            for i, val in enumerate(vals):
                lines.append(f"{ind}  t_[{i}] = {val};")
            for i in range(v.size):
                lines.append(f"{ind}  {_target(g, st.target, i, xname)} = t_[{i}];")
            lines.append(f"{ind}}}")
    return lines


def _draw_buffers(g, block, step, tag, ind, w, member=False):
    nn, nu = g.normals.get(block, 0), g.uniforms.get(block, 0)
    for b in range((nn + 3) // 4):
        if not member:
            w(f"{ind}T zq{b}[4];")
        w(f"{ind}normals4<T>(seed, p, {step}, {tag}, batch, {b}u, zq{b});")
    for k in range(nu):
        w(f"{ind}{'' if member else 'const T '}u{k} = uniform_word<T>(philox_block(seed, p, "
          f"{step}, {tag}, batch, {UNIFORM_BLOCK:#x}u + {k}u));")


class _LaneEmitter(_Emitter):
    """Expressions of local element l of a lane's slice: the state at offset c
    of the alias is x[l + c] or a halo value from a neighbouring lane (named
    by halo()), vectors over the state's dimension are the lane's slices, an
    alias used as a number is the element's global index lo + l.  In an
    observation the alias is the runtime obs index j_ and the state element
    is the lane's x[l]."""
    fold_bind = False

    def __init__(self, g, ty, xname, DL, halo=None, obs=False, xdouble=False):
        super().__init__(g, ty, xname, xdouble)
        self.DL, self.halo, self.obs = DL, halo, obs

    def ref(self, e, bind, elem, line):
        g, n = self.g, e.name
        l = elem
        if n in bind:
            v = "j_" if self.obs else f"(lo + {l})"
            return f"T({v})" if self.ty == "T" else f"(double)({v})"
        v = g.vars.get(n)
        if v is None:
            raise GenError(f"line {line}: cannot read {n!r} here")
        if v.kind in ("param", "input"):
            return f"pt_{n}" if self.ty == "T" else f"p_{n}"
        if v.kind == "noise":
            r = f"w_{n}" + (f"[{l}]" if v.dim else "")
            return f"(double){r}" if self.ty == "double" else r
        # the state
        if self.obs:
            q = 0
        elif e.index is None:
            q = l
        else:
            alias = next(iter(bind)) if bind else None
            _, c = _index_affine(g, e.index[0], alias, line)
            q = l + c
        if 0 <= q < self.DL or self.obs:
            r = f"{self.x}[{l if self.obs else q}]"
        else:
            r = self.halo(q)
        return f"(double){r}" if self.ty == "double" and not self.xdouble else r


def _generate_lanes(g: GenModel) -> str:
    """The device functor of a model split over lane groups (g.lpp lanes per
    particle, DL = D / lpp state elements per lane)."""
    x = next(v for v in g.vars.values() if v.kind == "state")
    D, L = x.size, g.lpp
    DL = D // L
    out = []
    w = out.append
    w(f"// ---- model {g.name}: generated from its .bi file ({L} lanes per particle)")
    w("template <class T>")
    w(f"struct {g.cname} {{")
    w("  using real_t = T;")
    w(f"  static constexpr int D = {D}, LPP = {L}, DL = {DL}, NOBS = {g.NOBS};")
    w("  static constexpr bool NOISE_FIRST = false, GENERATED = true;")
    w(f"  static constexpr bool HAS_FORCING = {'true' if _forcing(g) else 'false'};")
    w(f"  static constexpr int MIN_WAVES = sizeof(T) == 8 ? {_min_waves(g, DL)} : 1;")
    w(f"  static constexpr double DELTA = {_lit(g.delta)};")
    slots = [v for v in g.vars.values() if v.kind in ("param", "input")]
    noises = [v for v in g.vars.values() if v.kind == "noise"]
    for v in slots:
        w(f"  double p_{v.name};")
        w(f"  T pt_{v.name};")
    for v in noises:
        w(f"  T w_{v.name}" + (f"[{DL}];" if v.dim else ";"))
    nn_t, nu_t = g.normals.get("transition", 0), g.uniforms.get("transition", 0)
    nbt = (nn_t + 3) // 4
    for b in sorted({st.draw.index // 4 for st in g.pre + g.post
                     if st.kind == "draw" and st.draw.dist == "gaussian"
                     and g.vars[st.target].dim is None}):
        w(f"  T zq{b}[4];")
    init = [f"p_{v.name}(a.par[{v.offset}]), pt_{v.name}((T)a.par[{v.offset}])" for v in slots]
    init += [f"w_{v.name}(T(0))" for v in noises if not v.dim]
    w(f"  __device__ __forceinline__ explicit {g.cname}(const ModelArgs& a)"
      + (f" : {', '.join(init)}" if init else "") + " {")
    for v in noises:
        if v.dim:
            w(f"    for (int i = 0; i < DL; ++i) w_{v.name}[i] = T(0);")
    w("  }")

    # rhs with the halo from the neighbouring lanes of the group
    need = {}

    def halo(q):
        if q < 0:
            name = f"hp{DL + q}"
            need[name] = f"grp.prev(x[{DL + q}])"
        else:
            name = f"hn{q - DL}"
            need[name] = f"grp.next(x[{q - DL}])"
        return name

    em = _LaneEmitter(g, "T", "x", DL, halo)
    forced = _forcing(g)
    for fname, forms in (("rhs", None), ("rhs_unforced", forced)):
        if fname == "rhs_unforced" and forced is None:
            break
        need.clear()
        rhs = ["T(0)"] * DL
        for n, st in enumerate(g.ode):
            e = st.expr if forms is None else forms[1][n]
            for l in range(DL):
                rhs[l] = em.expr(e, {st.alias: l} if st.alias else {}, l, st.line)
        w("  template <class G>")
        w(f"  __device__ __forceinline__ void {fname}(const G& grp, const T (&x)[DL], "
          "T (&k)[DL]) const {")
        w("    const int lo = grp.g * DL;")
        w("    (void)lo;")
        for name, val in sorted(need.items()):
            w(f"    const T {name} = {val};")
        for l in range(DL):
            w(f"    k[{l}] = {rhs[l]};")
        w("  }")
    if forced is not None:
        w(f"  __device__ __forceinline__ T forcing() const {{ return pt_{forced[0]}; }}")

    # discrete actions on the lane's elements (vector actions read the values from before them)
    def actions(steps, ind, ty, xname, sctx):
        lines = []
        el = _LaneEmitter(g, ty, xname, DL, xdouble=(ty == "double"))
        for st in steps:
            v = g.vars[st.target]
            n = DL if v.dim else 1
            vals = []
            for l in range(n):
                bind = {st.alias: l} if st.alias else {}
                if st.kind == "draw":
                    vals.append(_lane_draw(el, st.draw, l, bind, st.line, sctx, v.dim is not None))
                else:
                    vals.append(el.expr(st.expr, bind, l, st.line))
            tgt = (lambda l: f"{xname}[{l}]") if v.kind == "state" else (
                (lambda l: f"w_{v.name}[{l}]") if v.dim else (lambda l: f"w_{v.name}"))
            if v.dim is None:
                lines.append(f"{ind}{tgt(0)} = {vals[0]};")
            else:
                lines.append(f"{ind}{{")
                if st.kind == "draw" and st.draw.dist == "gaussian":
                    step, tag, _, _, nfirst, _ = sctx
                    lines.append(f"{ind}  T zs_[{DL}];  // normals {st.draw.index} + lo .. of the firing")
                    lines.append(f"{ind}  slice_normals<T, {DL}>(seed, p, {step}, {tag}, batch, "
                                 f"{nfirst} + {st.draw.index}u + (uint32_t)lo, zs_);")
                lines.append(f"{ind}  {ty} t_[{DL}];")
                for l, val in enumerate(vals):
                    lines.append(f"{ind}  t_[{l}] = {val};")
                for l in range(DL):
                    lines.append(f"{ind}  {tgt(l)} = t_[{l}];")
                lines.append(f"{ind}}}")
        return lines

    nst = g.streams.get("transition", 0)
    tctx = ("step", "TAG_NOISE", "TAG_NOISE", f"fire * {nst}u", f"fire * {4 * nbt}u",
            f"fire * {nu_t}u")
    w("  template <int INTEG, bool ONE = false, class G>")
    w("  __device__ __forceinline__ int transition(const G& grp, T (&x)[DL], double t1, double t2,")
    w("                                            const ModelArgs& ma, uint64_t seed,")
    w("                                            uint32_t batch, uint32_t p, uint32_t step) {")
    w("    const int lo = grp.g * DL;")
    w("    (void)lo;")
    w("    int n = 0;")
    outer, body = w, []
    w = body.append
    # the normal words of scalar draws (vector draws take their slice of the firing's normals)
    sblocks = sorted({st.draw.index // 4 for st in g.pre + g.post
                      if st.kind == "draw" and st.draw.dist == "gaussian"
                      and g.vars[st.target].dim is None})
    if g.pre or g.post:
        w("      if (on) {")
        for b in sblocks:
            w(f"        normals4<T>(seed, p, step, TAG_NOISE, batch, fire * {nbt}u + {b}u, zq{b});")
        for ln in actions(g.pre, "        ", "T", "x", tctx):
            w(ln)
        w("      }")
    if g.has_ode:
        w("      if constexpr (INTEG == BIHIP_RK4) n += integrate_rk4(*this, grp, x, ta, tb, ma.h0);")
        w("      else n += integrate_dopri5(*this, grp, x, ta, tb, ma);")
    if g.post:
        w("      if (on) {")
        for ln in actions(g.post, "        ", "T", "x", tctx):
            w(ln)
        w("      }")
    w = outer
    _emit_interval_loop(w, body)

    # the noise variables after the transition, into the step's noise rows
    # (element e of the noise layout at z + e ld): a vector's slice from each
    # lane, a scalar (the same in every lane) from lane 0
    w("  template <class G>")
    w("  __device__ __forceinline__ void noise_out(const G& grp, T* z, long ld, uint32_t p) const {")
    w("    const int lo = grp.g * DL;")
    w("    (void)lo;")
    for v in noises:
        if v.dim:
            w(f"    for (int i = 0; i < DL; ++i) z[(long)({v.offset} + lo + i) * ld + p] = w_{v.name}[i];")
        else:
            w(f"    if (grp.g == 0) z[(long){v.offset} * ld + p] = w_{v.name};")
    w("  }")

    # observation terms of the lane's state elements: element s = lo + l is read
    # by obs element j when s = a j + b
    eo = _LaneEmitter(g, "double", "x", DL, obs=True)
    w("  __device__ __forceinline__ double obs(const T (&x)[DL], const double* y, int lo) const {")
    w("    double lp = 0.0;")
    for k_, st in enumerate(g.observation):
        d, v = st.draw, g.vars[st.target]
        idx = set()
        for e in d.args:
            for r in _refs(_expand(g, e), []):
                u = g.vars.get(r.name)
                if u is not None and u.kind == "state":
                    idx.add(_index_affine(g, r.index[0], st.alias, st.line))
        a, b = idx.pop() if idx else (0, 0)
        bind = {st.alias: 0} if st.alias else {}
        per = [[eo.expr(x_, bind, l, st.line) for x_ in d.args] for l in range(DL)]
        names, same = _hoist(per, k_, w, "    ", local="j_")
        lsg = None
        if d.dist == "gaussian" and not d.log and same[1]:
            lsg = f"l{k_}"
            w(f"    const double {lsg} = log({names[1]});")
        for l in range(DL):
            args = [nm if sm else a_ for nm, sm, a_ in zip(names, same, per[l])]
            w("    {")
            w(f"      const int s_ = lo + {l};")
            if a == 0:
                # every element j reads the same state element b (or none):
                # the lane that holds b adds all of them
                w(f"      if (s_ == {b}) {{")
                w(f"       for (int j_ = 0; j_ < {v.size}; ++j_) {{")
            else:
                w(f"      const int j_ = (s_ - {b}) / {a};")
                w(f"      if ((s_ - {b}) % {a} == 0 && j_ >= 0 && j_ < {v.size}) {{")
            w(f"        const double xy = y[{v.offset} + j_];")
            for ln in _logdensity(d, args, lsg):
                w("        " + ln)
            if a == 0:
                w("       }")
            w("      }")
            w("    }")
    w("    return lp;")
    w("  }")
    _emit_obs_sample(g, w)

    # initial states of the lane's elements, in double
    w("  __device__ __forceinline__ void init(double (&xd)[DL], uint64_t seed, uint32_t batch,")
    w("                                       uint32_t p, int lo) const {")
    w("    for (int i = 0; i < DL; ++i) xd[i] = 0.0;")
    ictx = ("0u", "TAG_INIT", "TAG_INIT", "0u", "0u", "0u")
    for ln in actions(g.initial, "    ", "double", "xd", ictx):
        w(ln)
    w("  }")
    w("};")
    return "\n".join(out)


def _lane_draw(em, d: Draw, l, bind, line, sctx, vector):
    """Draw of local element l (vector: global element lo + l; scalar: the same
    draw in every lane).  sctx = (step, tag, tag, first stream, first normal,
    first uniform) of the firing."""
    step, tag, _, sfirst, nfirst, ufirst = sctx
    a = [em.expr(x_, bind, l, line) for x_ in d.args]
    k = f"{d.index}u + (uint32_t)(lo + {l})" if vector else f"{d.index}u"
    if d.dist in STREAM_DISTS:
        ad = ", ".join(f"(double)({x_})" for x_ in a)
        v = (f"gen_{d.dist}({ad}, DrawStream(seed, p, {step}, {tag}, batch, "
             f"{STREAM_BLOCK:#x}u + (({sfirst} + {k}) << 12)))")
        return f"T({v})" if em.ty == "T" else v
    if d.dist == "gaussian":
        z = f"zs_[{l}]" if vector else _zname(d.index)
        z = z if em.ty == "T" else f"(double){z}"
        v = f"({a[0]} + ({a[1]} * {z}))"
        return f"exp({v})" if d.log else v
    u = (f"uniform_word<T>(philox_block(seed, p, {step}, {tag}, batch, {UNIFORM_BLOCK:#x}u + "
         f"{ufirst} + {k}))")
    u = u if em.ty == "T" else f"(double){u}"
    if d.dist == "exponential":
        one = em.num(1.0)
        return f"((-log({one} - {u})) / {a[0]})"
    return f"({a[0]} + (({a[1]} - {a[0]}) * {u}))"


def _forcing(g: GenModel):
    """(F, {equation index: g}) when every state element's equation is
    g + F with F one scalar param or input: the integrators then carry the
    constant forcing themselves in FP32 (HAS_FORCING, as the hand-written
    Lorenz-96 functor does); otherwise None."""
    if not g.ode or sum(g.vars[st.target].size for st in g.ode) != g.D:
        return None
    name, rest = None, {}
    for n, st in enumerate(g.ode):
        e = st.expr
        if not (isinstance(e, Binary) and e.op == "+" and isinstance(e.right, Ref)
                and e.right.index is None):
            return None
        v = g.vars.get(e.right.name)
        if v is None or v.kind not in ("param", "input") or v.dim is not None:
            return None
        if name not in (None, v.name):
            return None
        name, rest[n] = v.name, e.left
    return name, rest


def _emit_interval_loop(w, body):
    """The rest of transition(): `body` (the actions and the ode of one
    sub-interval [ta, tb], firing when `on`, the firing's count `fire`) run
    over the sub-intervals of [t1, t2] that the multiples of DELTA cut, or
    once when the caller knows [t1, t2] to be one sub-interval that starts at
    a multiple (ONE): the loop's state then costs no registers."""
    w("    if constexpr (ONE) {")
    w("      const bool on = true;")
    w("      const uint32_t fire = 0;")
    w("      const double ta = t1, tb = t2;")
    w("      (void)on;")
    w("      (void)fire;")
    w("      (void)ta;")
    w("      (void)tb;")
    for ln in body:
        w(ln)
    w("      return n;")
    w("    }")
    w("    double sa = gen_scaled(t1, DELTA);")
    w("    const double s2 = gen_scaled(t2, DELTA);")
    w("    double ta = t1;")
    w("    uint32_t fire = 0;")
    w("    while (sa < s2) {")
    w("      const double sb = fmin(floor(sa) + 1.0, s2);")
    w("      const double tb = sb == s2 ? t2 : sb * DELTA;")
    w("      const bool on = sa == floor(sa);")
    for ln in body:
        w(ln)
    w("      fire += on ? 1u : 0u;")
    w("      sa = sb;")
    w("      ta = tb;")
    w("    }")
    w("    return n;")
    w("  }")


def _min_waves(g: GenModel, DL: int) -> int:
    """Waves per SIMD asked of the compiler for the double instance.  A
    double slice of DL elements through DOPRI5 holds about nine arrays of DL
    (two registers each) plus some 70 registers of controller and addresses,
    RK4 about six: when that estimate falls just above 256 (at most 300), a
    cap at 256 registers costs a few spills and doubles the waves per SIMD
    (the 40-element Lorenz-96 slice of 10: 268 registers uncapped, 256 with 8
    spills, 206 -> 136 us per step).  Float instances keep the compiler's
    choice (capped at 128 the same model spilled 14 and slowed)."""
    arrays = 9 if g.integrator == ALGS["RK5(4)"] else 6
    return 2 if 2 * arrays * DL + 70 <= 300 else 1


def _emit_obs_sample(g: GenModel, w):
    """Observation samples (Model::observationSamples) from a particle's whole
    state x[D] (one thread per particle); y has NOBS entries."""
    eo = _Emitter(g, "double", "x")
    w("  __device__ __forceinline__ void obs_sample(const T (&x)[D], uint64_t seed, uint32_t batch,")
    w("                                             uint32_t p, uint32_t step, double* y) const {")
    _draw_buffers(g, "observation", "step", "TAG_OBS", "    ", w)
    for j in range(g.NOBS):
        w(f"    y[{j}] = __builtin_nan(\"\");")
    for st in g.observation:
        v = g.vars[st.target]
        for i, bind in _elements(g, st):
            w(f"    y[{v.offset + i}] = "
              f"{_draw_value(eo, st.draw, i, bind, st.line, ('step', 'TAG_OBS', '0u'))};")
    w("  }")


def generate_model(g: GenModel) -> str:
    """The device functor of one model."""
    if g.lpp > 1:
        return _generate_lanes(g)
    S, NO = g.D, g.NOBS
    out = []
    w = out.append
    w(f"// ---- model {g.name}: generated from its .bi file")
    w("template <class T>")
    w(f"struct {g.cname} {{")
    w("  using real_t = T;")
    w(f"  static constexpr int D = {S}, LPP = 1, DL = {S}, NOBS = {NO};")
    w("  static constexpr bool NOISE_FIRST = false, GENERATED = true;")
    w(f"  static constexpr bool HAS_FORCING = {'true' if _forcing(g) else 'false'};")
    w(f"  static constexpr int MIN_WAVES = sizeof(T) == 8 ? {_min_waves(g, S)} : 1;")
    w(f"  static constexpr double DELTA = {_lit(g.delta)};")
    slots = [v for v in g.vars.values() if v.kind in ("param", "input")]
    noises = [v for v in g.vars.values() if v.kind == "noise"]
    for v in slots:  # in double, and in the state type for the transition's arithmetic
        w(f"  double p_{v.name}" + (f"[{v.size}];" if v.dim else ";"))
        w(f"  T pt_{v.name}" + (f"[{v.size}];" if v.dim else ";"))
    for v in noises:
        w(f"  T w_{v.name}" + (f"[{v.size}];" if v.dim else ";"))
    nn_t, nu_t = g.normals.get("transition", 0), g.uniforms.get("transition", 0)
    nbt = (nn_t + 3) // 4
    # draw buffers of the transition: members, shared by its parts before and after the ode
    for b in range(nbt):
        w(f"  T zq{b}[4];")
    for k in range(nu_t):
        w(f"  T u{k};")
    init = [f"p_{v.name}(a.par[{v.offset}]), pt_{v.name}((T)a.par[{v.offset}])"
            for v in slots if not v.dim]
    init += [f"w_{v.name}(T(0))" for v in noises if not v.dim]
    w(f"  __device__ __forceinline__ explicit {g.cname}(const ModelArgs& a)"
      + (f" : {', '.join(init)}" if init else "") + " {")
    for v in slots:
        if v.dim:
            w(f"    for (int i = 0; i < {v.size}; ++i) p_{v.name}[i] = a.par[{v.offset} + i];")
            w(f"    for (int i = 0; i < {v.size}; ++i) pt_{v.name}[i] = (T)a.par[{v.offset} + i];")
    for v in noises:
        if v.dim:
            w(f"    for (int i = 0; i < {v.size}; ++i) w_{v.name}[i] = T(0);")
    w("  }")

    # rhs: the ode block (share/tt/cpp/action/ode_.hpp.tt); states without an equation stay put
    em = _Emitter(g, "T", "x")
    forced = _forcing(g)
    for fname, forms in (("rhs", None), ("rhs_unforced", forced)):
        if fname == "rhs_unforced" and forced is None:
            break
        rhs = ["T(0)"] * S
        for n, st in enumerate(g.ode):
            v = g.vars[st.target]
            e = st.expr if forms is None else forms[1][n]
            for i, bind in _elements(g, st):
                rhs[v.offset + i] = em.expr(e, bind, i, st.line)
        w("  template <class G>")
        w(f"  __device__ __forceinline__ void {fname}(const G&, const T (&x)[DL], "
          "T (&k)[DL]) const {")
        for j in range(S):
            w(f"    k[{j}] = {rhs[j]};")
        w("  }")
    if forced is not None:
        w(f"  __device__ __forceinline__ T forcing() const {{ return pt_{forced[0]}; }}")

    nst = g.streams.get("transition", 0)
    tsctx = ("step", "TAG_NOISE", f"fire * {nst}u")
    w("  // one interval [t1, t2]: discrete actions at multiples of DELTA, the ode in between")
    w("  template <int INTEG, bool ONE = false, class G>")
    w("  __device__ __forceinline__ int transition(const G& grp, T (&x)[DL], double t1, double t2,")
    w("                                            const ModelArgs& ma, uint64_t seed,")
    w("                                            uint32_t batch, uint32_t p, uint32_t step) {")
    w("    int n = 0;")
    outer, body = w, []
    w = body.append
    if g.pre or g.post:
        w("      if (on) {")
        for b in range(nbt):
            w(f"        normals4<T>(seed, p, step, TAG_NOISE, batch, fire * {nbt}u + {b}u, zq{b});")
        for k in range(nu_t):
            w(f"        u{k} = uniform_word<T>(philox_block(seed, p, step, TAG_NOISE, batch, "
              f"{UNIFORM_BLOCK:#x}u + fire * {nu_t}u + {k}u));")
        for ln in _actions(g, g.pre, em, "x", "        ", "T", tsctx):
            w(ln)
        w("      }")
    if g.has_ode:
        w("      if constexpr (INTEG == BIHIP_RK4) n += integrate_rk4(*this, grp, x, ta, tb, ma.h0);")
        w("      else n += integrate_dopri5(*this, grp, x, ta, tb, ma);")
    if g.post:
        w("      if (on) {")
        for ln in _actions(g, g.post, em, "x", "        ", "T", tsctx):
            w(ln)
        w("      }")
    w = outer
    _emit_interval_loop(w, body)

    # the noise variables after the transition, into the step's noise rows
    # (element e of the noise layout at z + e ld)
    w("  template <class G>")
    w("  __device__ __forceinline__ void noise_out(const G&, T* z, long ld, uint32_t p) const {")
    for v in noises:
        if v.dim:
            w(f"    for (int i = 0; i < {v.size}; ++i) z[(long)({v.offset} + i) * ld + p] = w_{v.name}[i];")
        else:
            w(f"    z[(long){v.offset} * ld + p] = w_{v.name};")
    w("  }")

    # observation log-density (share/tt/cpp/action/gaussian.hpp.tt:48-73, uniform.hpp.tt:38-52)
    eo = _Emitter(g, "double", "x")
    w("  __device__ __forceinline__ double obs(const T (&x)[DL], const double* y, int) const {")
    w("    double lp = 0.0;")
    for ka, st in enumerate(g.observation):
        d, v = st.draw, g.vars[st.target]
        elems = list(_elements(g, st))
        per = [[eo.expr(x, bind, i, st.line) for x in d.args] for i, bind in elems]
        names, same = _hoist(per, ka, w, "    ")
        lsg = None
        if d.dist == "gaussian" and not d.log and len(same) > 1 and same[1]:
            lsg = f"l{ka}"
            w(f"    const double {lsg} = log({names[1]});")
        for (i, bind), args in zip(elems, per):
            args = [nm if sm else a for nm, sm, a in zip(names, same, args)]
            w("    {")
            w(f"      const double xy = y[{v.offset + i}];")
            for ln in _logdensity(d, args, lsg):
                w("      " + ln)
            w("    }")
    w("    return lp;")
    w("  }")

    _emit_obs_sample(g, w)

    # initial states, in double (share/tt/cpp/block/initial.hpp.tt)
    ei = _Emitter(g, "double", "xd", xdouble=True)
    w("  __device__ __forceinline__ void init(double (&xd)[DL], uint64_t seed, uint32_t batch,")
    w("                                       uint32_t p, int) const {")
    for j in range(S):
        w(f"    xd[{j}] = 0.0;")
    _draw_buffers(g, "initial", "0u", "TAG_INIT", "    ", w)
    for ln in _actions(g, g.initial, ei, "xd", "    ", "double", ("0u", "TAG_INIT", "0u")):
        w(ln)
    w("  }")
    w("};")
    return "\n".join(out)


def _cstr_list(names):
    return "{" + ", ".join(f'"{n}"' for n in names) + (", " if names else "") + "nullptr}"


def _int_list(vals):
    return "{" + ", ".join(str(v) for v in vals) + (", " if vals else "") + "0}"


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
    out.append("static const int GEN_NONE_SIZES[] = {0};")
    for i, g in enumerate(models):
        for kind in KIND_CODE:
            vs = [v for v in g.vars.values() if v.kind == kind]
            out.append(f"static const char* const GEN_{i}_{kind.upper()}[] = "
                       f"{_cstr_list([v.name for v in vs])};")
            out.append(f"static const int GEN_{i}_{kind.upper()}_SIZES[] = "
                       f"{_int_list([v.size for v in vs])};")
        sv = [v for v in g.vars.values() if v.kind == "state"]
        out.append(f"static const char* const GEN_{i}_STATE_DIMS[] = "
                   f"{_cstr_list([v.dim or '' for v in sv])};")
        nv = [v for v in g.vars.values() if v.kind == "noise"]
        out.append(f"static const char* const GEN_{i}_NOISE_DIMS[] = "
                   f"{_cstr_list([v.dim or '' for v in nv])};")
    out.append("static const GenInfo GEN_MODELS[] = {")
    for i, g in enumerate(models):
        names = ", ".join(f"GEN_{i}_{k.upper()}" for k in KIND_CODE)
        sizes = ", ".join(f"GEN_{i}_{k.upper()}_SIZES" for k in KIND_CODE)
        out.append(f'    {{"{g.name}", {g.D}, {g.NOBS}, {g.integrator}, {g.lpp}, '
                   f"{_lit(g.h)}, {_lit(g.atol)}, {_lit(g.rtol)}, {_lit(g.delta)}, "
                   f"{{{names}}}, {{{sizes}}}, GEN_{i}_STATE_DIMS, GEN_{i}_NOISE_DIMS}},")
    if not models:
        none = ", ".join(["GEN_NONE"] * len(KIND_CODE))
        nsz = ", ".join(["GEN_NONE_SIZES"] * len(KIND_CODE))
        out.append(f'    {{"", 0, 0, 0, 1, 0.0, 0.0, 0.0, 1.0, {{{none}}}, {{{nsz}}}, GEN_NONE, '
                   'GEN_NONE},')
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


def generate_priors(models: List[GenModel]) -> str:
    """Host data for the priors of the parameter blocks: per model, one
    {distribution, a, b} per param element (distribution -1: no prior)."""
    out = ["// Generated by libbi_amd.dsl from the model files in libbi_amd/bi; do not edit.",
           "#pragma once", "", "namespace bihip {", ""]
    for i, g in enumerate(models):
        n = g.size("param")
        rows = []
        for k in range(n):
            c, a, b = g.priors.get(k, (-1, 0.0, 0.0))
            rows.append(f"{{{c}, {_lit(a)}, {_lit(b)}}}")
        if not rows:
            rows = ["{-1, 0.0, 0.0}"]
        out.append(f"static const GenPrior GEN_{i}_PRIOR[] = {{{', '.join(rows)}}};")
    out.append("static const GenPrior* const GEN_PRIORS[] = {"
               + ", ".join(f"GEN_{i}_PRIOR" for i in range(len(models))) + ("" if models else "nullptr") + "};")
    out.append("static const int GEN_PRIOR_COUNT[] = {"
               + ", ".join(str(g.size("param")) for g in models) + ("" if models else "0") + "};")
    out.append(f"constexpr int GEN_PRIOR_MODELS = {len(models)};")
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
