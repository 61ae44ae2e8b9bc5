"""Parser for LibBi's model language (the `.bi` files read by `libbi
--model-file`).

Follows the tokens of share/bi.lex and the grammar of share/bi.yp: a model
holds `dim`, `const`, `inline` and variable declarations (`state`, `noise`,
`obs`, `param`, `input`, `state_aux_`, `param_aux_`) and `sub` blocks; a block
holds actions (`x ~ dist(...)`, `x <- expr`, `dx/dt = expr`), nested blocks
(`ode(...) { ... }`) and local `const`/`inline` declarations.  Expressions
are C-like with `**` for powers (share/bi.yp:330-420).  Statements end at a
`;` or where the next token cannot continue them, as in the reference's
LALR grammar.

The result is a tree of plain dataclasses; the code generator
(libbi_amd/dsl/codegen.py) checks what it supports.
"""
import re
from dataclasses import dataclass, field
from typing import List, Optional, Tuple


class BiSyntaxError(ValueError):
    """A model file that does not parse; the message names the line."""


# ---------------------------------------------------------------- tokens (share/bi.lex)
_TOKEN_RE = re.compile(r"""
    (?P<ws>[ \t\r\n]+)
  | (?P<comment>//[^\n]*|/\*.*?\*/)
  | (?P<string>'(?:\\.|[^\\'])*'|"(?:\\.|[^\\"])*")
  | (?P<number>(?:[0-9]+\.[0-9]*|[0-9]*\.[0-9]+)(?:[eE][+-]?[0-9]+)?|[0-9]+(?:[eE][+-]?[0-9]+)?)
  | (?P<ident>[A-Za-z_][A-Za-z_0-9]*)
  | (?P<op>\.\*\*|\*\*|<-|&&|\|\||<=|>=|==|!=|\.\*|\./|\.\+|\.-|[;,:=()\[\]{}!~\-+*/%<>^|?])
""", re.VERBOSE | re.DOTALL)

KEYWORDS = {"model", "sub", "dim", "const", "inline", "state", "state_aux_", "noise", "input",
            "obs", "param", "param_aux_", "do", "then", "dt"}
VAR_TYPES = ("state", "state_aux_", "noise", "input", "obs", "param", "param_aux_")


@dataclass
class Token:
    kind: str  # ident, keyword, number, string, op, eof
    text: str
    line: int


def tokenize(text: str) -> List[Token]:
    out, pos, line = [], 0, 1
    while pos < len(text):
        m = _TOKEN_RE.match(text, pos)
        if m is None:
            raise BiSyntaxError(f"line {line}: unexpected character {text[pos]!r}")
        kind = m.lastgroup
        s = m.group(kind)
        if kind == "ident" and s in KEYWORDS:
            kind = "keyword"
        if kind not in ("ws", "comment"):
            out.append(Token(kind, s, line))
        line += s.count("\n")
        pos = m.end()
    out.append(Token("eof", "", line))
    return out


# ---------------------------------------------------------------- tree
@dataclass
class Num:
    value: float
    text: str
    is_int: bool


@dataclass
class Str:
    value: str


@dataclass
class Ref:
    name: str
    index: Optional[list] = None  # indexed references (dimensions) are kept for errors


@dataclass
class Call:
    name: str
    args: list
    named: List[Tuple[str, object]]


@dataclass
class Unary:
    op: str
    arg: object


@dataclass
class Binary:
    op: str
    left: object
    right: object


@dataclass
class Ternary:
    cond: object
    then: object
    other: object


@dataclass
class Spec:
    name: str
    dims: List[str] = field(default_factory=list)
    args: list = field(default_factory=list)
    named: List[Tuple[str, object]] = field(default_factory=list)
    line: int = 0


@dataclass
class Action:
    target: str
    op: str  # '~', '<-', '='  ('=' is an ODE equation dtarget/dt = expr)
    expr: object
    line: int
    target_dims: Optional[list] = None


@dataclass
class Block:
    spec: Spec
    items: list  # Action | Block | ('const'|'inline', name, expr)
    line: int = 0


@dataclass
class Model:
    name: str
    dims: List[Spec] = field(default_factory=list)
    consts: List[Tuple[str, object]] = field(default_factory=list)
    inlines: List[Tuple[str, object]] = field(default_factory=list)
    vars: List[Tuple[str, Spec]] = field(default_factory=list)  # (type, spec) in order
    blocks: List[Block] = field(default_factory=list)

    def names_of(self, vtype):
        return [s.name for t, s in self.vars if t == vtype]

    def block(self, name):
        for b in self.blocks:
            if b.spec.name == name:
                return b
        return None


# ---------------------------------------------------------------- parser
class _Parser:
    def __init__(self, text):
        self.toks = tokenize(text)
        self.i = 0

    # token helpers
    def peek(self, k=0):
        return self.toks[min(self.i + k, len(self.toks) - 1)]

    def next(self):
        t = self.toks[self.i]
        self.i += 1
        return t

    def error(self, msg, tok=None):
        tok = tok or self.peek()
        what = tok.text if tok.kind != "eof" else "end of file"
        raise BiSyntaxError(f"line {tok.line}: {msg} (at {what!r})")

    def at(self, text, kind=None):
        t = self.peek()
        return t.text == text and (kind is None or t.kind == kind) and t.kind != "string"

    def accept(self, text):
        if self.at(text):
            return self.next()
        return None

    def expect(self, text):
        if not self.at(text):
            self.error(f"expected {text!r}")
        return self.next()

    def ident(self):
        t = self.peek()
        if t.kind != "ident":
            self.error("expected an identifier")
        return self.next().text

    def semi(self):
        self.accept(";")

    # model := 'model' spec '{' definitions '}'
    def model(self):
        self.expect("model")
        spec = self.spec()
        m = Model(spec.name)
        self.expect("{")
        while not self.at("}"):
            self.definition(m)
        self.expect("}")
        if self.peek().kind != "eof":
            self.error("text after the model")
        return m

    def definition(self, m):
        t = self.peek()
        if t.kind == "keyword" and t.text == "dim":
            self.next()
            m.dims.extend(self.spec_list())
        elif t.kind == "keyword" and t.text in ("const", "inline"):
            self.next()
            for name, e in self.assign_list():
                (m.consts if t.text == "const" else m.inlines).append((name, e))
        elif t.kind == "keyword" and t.text in VAR_TYPES:
            self.next()
            for s in self.spec_list(array=True):
                m.vars.append((t.text, s))
        elif t.kind == "keyword" and t.text == "sub":
            self.next()
            m.blocks.append(self.block())
        else:
            self.error("expected a declaration or a sub block")

    def assign_list(self):
        out = []
        while True:
            name = self.ident()
            self.expect("=")
            out.append((name, self.expression()))
            self.semi()
            if not self.accept(","):
                return out

    def spec_list(self, array=False):
        out = []
        while True:
            out.append(self.spec(array))
            self.semi()
            if not self.accept(","):
                return out

    # spec := IDENT ['[' dims ']'] ['(' args ')']
    def spec(self, array=True):
        line = self.peek().line
        name = self.ident()
        dims = []
        if array and self.accept("["):
            dims.append(self.ident())
            while self.accept(","):
                dims.append(self.ident())
            self.expect("]")
        args, named = [], []
        if self.accept("("):
            args, named = self.arguments(")")
        return Spec(name, dims, args, named, line)

    def arguments(self, close):
        args, named = [], []
        if self.accept(close):
            return args, named
        while True:
            if self.peek().kind == "ident" and self.peek(1).text == "=" and self.peek(2).text != "=":
                name = self.next().text
                self.next()
                named.append((name, self.expression()))
            else:
                if named:
                    self.error("positional argument after a named one")
                args.append(self.expression())
            if self.accept(close):
                return args, named
            self.expect(",")

    # block := spec '{' block_definitions '}'  |  '{' ... '}'
    def block(self):
        line = self.peek().line
        spec = self.spec(array=False) if not self.at("{") else Spec("", line=line)
        self.expect("{")
        items = []
        while not self.at("}"):
            items.append(self.block_item())
        self.expect("}")
        return Block(spec, items, line)

    def block_item(self):
        t = self.peek()
        if t.kind == "keyword" and t.text in ("const", "inline"):
            self.next()
            return [(t.text, n, e) for n, e in self.assign_list()]
        if t.kind == "keyword" and t.text in ("dim",) + VAR_TYPES:
            self.error("declarations inside blocks are not supported")
        if t.kind == "keyword" and t.text == "do":
            self.error("'do ... then' blocks are not supported")
        if t.kind != "ident":
            self.error("expected an action or a block")
        # an action starts IDENT ['[' ... ']'] then '~', '<-' or '/' dt '='
        j = 1
        if self.peek(1).text == "[":
            depth = 0
            while True:
                tk = self.peek(j)
                if tk.kind == "eof":
                    self.error("unterminated '['")
                depth += tk.text == "["
                depth -= tk.text == "]"
                j += 1
                if depth == 0:
                    break
        nxt = self.peek(j).text
        if nxt in ("~", "<-") or (nxt == "/" and self.peek(j + 1).text == "dt"):
            return self.action()
        return self.block()

    def action(self):
        line = self.peek().line
        target = self.ident()
        dims = None
        if self.accept("["):
            dims = [self.expression()]
            while self.accept(","):
                dims.append(self.expression())
            self.expect("]")
        if self.accept("~"):
            op = "~"
        elif self.accept("<-"):
            op = "<-"
        else:
            # 'dx/dt = ...': the identifier is the variable's name after a leading d
            if not target.startswith("d") or len(target) < 2:
                self.error("an ODE equation is written dx/dt = ...")
            target = target[1:]
            self.expect("/")
            self.expect("dt")
            self.expect("=")
            op = "="
        e = self.expression()
        self.semi()
        return Action(target, op, e, line, dims)

    # expressions: share/bi.yp:330-420 (C precedence, '**' binds tighter than '*')
    def expression(self):
        return self.conditional()

    def conditional(self):
        c = self.binary(0)
        if self.accept("?"):
            a = self.conditional()
            self.expect(":")
            b = self.conditional()
            return Ternary(c, a, b)
        return c

    _LEVELS = [("||",), ("&&",), ("==", "!="), ("<", ">", "<=", ">="), ("+", "-", ".+", ".-"),
               ("*", "/", "%", ".*", "./"), ("**", ".**")]

    def binary(self, level):
        if level == len(self._LEVELS):
            return self.unary()
        e = self.binary(level + 1)
        while self.peek().kind == "op" and self.peek().text in self._LEVELS[level]:
            op = self.next().text.lstrip(".")  # element-wise forms equal the scalar ones
            e = Binary(op, e, self.binary(level + 1))
        return e

    def unary(self):
        if self.peek().kind == "op" and self.peek().text in ("+", "-", "!"):
            op = self.next().text
            return Unary(op, self.unary())
        return self.postfix()

    def postfix(self):
        t = self.peek()
        if t.kind == "number":
            self.next()
            is_int = re.fullmatch(r"[0-9]+", t.text) is not None
            return Num(float(t.text), t.text, is_int)
        if t.kind == "string":
            self.next()
            return Str(t.text[1:-1])
        if t.kind == "ident":
            name = self.next().text
            if self.accept("("):
                args, named = self.arguments(")")
                return Call(name, args, named)
            if self.accept("["):
                idx = [self.expression()]
                while self.accept(","):
                    idx.append(self.expression())
                self.expect("]")
                return Ref(name, idx)
            return Ref(name)
        if self.accept("("):
            e = self.conditional()
            self.expect(")")
            return e
        self.error("expected an expression")


def parse(text: str) -> Model:
    """Parses the text of a `.bi` model file."""
    return _Parser(text).model()


def parse_file(path) -> Model:
    with open(path) as f:
        return parse(f.read())
