"""User-defined fluids compiled into the GPU kernels.

The reference registers ``Fluid`` classes (abcmb/species.py:11-249) with
``Model(user_species=...)`` (abcmb/model_specs.py:89-116) and JAX traces their
methods into its XLA program.  Here a model's user fluids are traced once, on
the host, with the symbolic values of this module, and turned into C++:

  * ``Expr``: a scalar expression over ln a, a, k, aH, the conformal time of
    the initial conditions and the cosmology's parameters (hash-consed, so
    equal subexpressions are one node);
  * ``Lin``: a linear form over the state rows y[i] and the metric
    derivatives h', eta', with ``Expr`` coefficients.  The reference's
    perturbation equations are linear in y; a product of two forms raises.

From the traced methods come, for the ODE kernels' row-descriptor form
(every row: up to 5 terms coefficient x background scalar x y[column], plus
static couplings to h' and eta'):

  * each user row's terms, with coefficients that depend on k only kept as
    static numbers and the others computed once per evaluation time into
    background scalar slots (the kernels' ``SC_FL0 + s``);
  * its h' / eta' coefficients (static in time and parameters: a fluid whose
    metric couplings vary in time is refused);
  * the rows' weights in h' (rho delta) and eta' ((rho + P) theta), again
    slots; their shear and matter-density terms for the output table;
  * the summed background density and pressure, and the initial conditions.

``FluidSet.generate`` writes ``fluid_dev.h`` (the interface of
``csrc/fluid_none_dev.h``, templated on the dual-number width) and
``fluid_host.h`` (``csrc/fluid_none_host.h``); ``build`` compiles the kernel
and host libraries against them into ``abcmb_amd/_fluids/<signature>/``.
The same classes are also called with numbers (derived parameters, output
dictionaries, the CPU oracle's callbacks).
"""
from __future__ import annotations

import hashlib
import json
import math
import os

import numpy as np

from . import constants as cnst

HERE = os.path.dirname(os.path.abspath(__file__))
FLUID_ROOT = os.path.join(HERE, "_fluids")
MAX_TERMS = 5          # row-descriptor terms per row (pert_core.h Row)
MAX_ROWS = 26          # the user block must fit the kernels' block width
MAX_PARAMS = 8         # ABCMB_FLUID_PARAMS

# params keys that are AbcmbCosmo fields (abcmb/main.py:368-612 -> include/abcmb.h)
PARAM_FIELDS = {
    "h": "h", "omega_b": "omega_b", "omega_cdm": "omega_cdm", "omega_Lambda": "omega_Lambda",
    "A_s": "A_s", "n_s": "n_s", "TCMB0": "TCMB0", "YHe": "YHe",
    "N_nu_massless": "N_ur", "T_nu_massless": "T_ur", "N_nu_massive": "N_ncdm",
    "T_nu_massive": "T_ncdm", "m_nu_massive": "m_ncdm",
    "R_nu": "R_nu", "R_b": "R_b", "om": "om", "omega_r": "omega_r", "omega_m": "omega_m",
}

# dependence flags of an expression
T, K, P, H, TAU = 1, 2, 4, 8, 16


class Traced:
    """Marker base of the symbolic values (``abcmb_amd.xp`` dispatches on it)."""


_TABLE: dict = {}


def _intern(op, args, deps):
    key = (op,) + tuple(a.nid if isinstance(a, Expr) else ("c", a) for a in args)
    e = _TABLE.get(key)
    if e is None:
        e = Expr.__new__(Expr)
        e.op, e.args, e.deps, e.nid = op, tuple(args), deps, len(_TABLE)
        _TABLE[key] = e
    return e


def const(v) -> "Expr":
    v = float(v)
    if not math.isfinite(v):
        raise ValueError(f"non-finite constant {v} in a fluid expression")
    return _intern("const", (v,), 0)


def sym(name, deps) -> "Expr":
    return _intern("sym", (name,), deps)


def lift(x):
    if isinstance(x, (Expr, Lin)):
        return x
    if isinstance(x, (bool, np.bool_)):
        raise TypeError("a boolean in a fluid expression")
    if isinstance(x, (int, float, np.integer, np.floating)) or (isinstance(x, np.ndarray) and x.ndim == 0):
        return const(float(x))
    raise TypeError(f"cannot use {type(x).__name__} in a traced fluid expression")


def _isc(e, v=None):
    return isinstance(e, Expr) and e.op == "const" and (v is None or e.args[0] == v)


class Expr(Traced):
    """A scalar expression node (see the module docstring)."""
    __slots__ = ("op", "args", "deps", "nid")

    # ---- construction with folding
    @staticmethod
    def add(a, b):
        if _isc(a) and _isc(b):
            return const(a.args[0] + b.args[0])
        if _isc(a, 0.0):
            return b
        if _isc(b, 0.0):
            return a
        return _intern("add", (a, b), a.deps | b.deps)

    @staticmethod
    def mul(a, b):
        if _isc(a) and _isc(b):
            return const(a.args[0] * b.args[0])
        if _isc(a, 0.0) or _isc(b, 0.0):
            return const(0.0)
        if _isc(a, 1.0):
            return b
        if _isc(b, 1.0):
            return a
        return _intern("mul", (a, b), a.deps | b.deps)

    @staticmethod
    def div(a, b):
        if _isc(b, 0.0):
            raise ZeroDivisionError("division by zero in a fluid expression")
        if _isc(a) and _isc(b):
            return const(a.args[0] / b.args[0])
        if _isc(a, 0.0):
            return const(0.0)
        if _isc(b, 1.0):
            return a
        return _intern("div", (a, b), a.deps | b.deps)

    @staticmethod
    def neg(a):
        if _isc(a):
            return const(-a.args[0])
        if a.op == "neg":
            return a.args[0]
        return _intern("neg", (a,), a.deps)

    @staticmethod
    def powc(a, c):
        c = float(c)
        if c == 0.0:
            return const(1.0)
        if c == 1.0:
            return a
        if _isc(a):
            return const(a.args[0] ** c)
        return _intern("powc", (a, c), a.deps)

    @staticmethod
    def fn(name, a):
        if _isc(a):
            return const({"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "tanh": math.tanh,
                          "cos": math.cos}[name](a.args[0]))
        if name == "exp" and a.op == "sym" and a.args[0] == "lna":
            return A
        if name == "log" and a.op == "sym" and a.args[0] == "a":
            return LNA
        return _intern(name, (a,), a.deps)

    # ---- operators
    def __add__(self, o):
        o = lift(o)
        return o.__radd__(self) if isinstance(o, Lin) else Expr.add(self, o)

    def __radd__(self, o):
        return Expr.add(lift(o), self)

    def __sub__(self, o):
        o = lift(o)
        return (-o).__radd__(self) if isinstance(o, Lin) else Expr.add(self, Expr.neg(o))

    def __rsub__(self, o):
        return Expr.add(lift(o), Expr.neg(self))

    def __mul__(self, o):
        o = lift(o)
        return o.__rmul__(self) if isinstance(o, Lin) else Expr.mul(self, o)

    def __rmul__(self, o):
        return Expr.mul(lift(o), self)

    def __truediv__(self, o):
        o = lift(o)
        if isinstance(o, Lin):
            raise NotImplementedError("division by a state-dependent value: the equations must be linear in y")
        return Expr.div(self, o)

    def __rtruediv__(self, o):
        return Expr.div(lift(o), self)

    def __neg__(self):
        return Expr.neg(self)

    def __pos__(self):
        return self

    def __pow__(self, p):
        p = lift(p)
        if isinstance(p, Lin):
            raise NotImplementedError("a power with a state-dependent exponent")
        if _isc(p):
            return Expr.powc(self, p.args[0])
        return Expr.fn("exp", Expr.mul(p, Expr.fn("log", self)))

    def __rpow__(self, b):
        return lift(b).__pow__(self)

    def exp(self):
        return Expr.fn("exp", self)

    def log(self):
        return Expr.fn("log", self)

    def sqrt(self):
        return Expr.fn("sqrt", self)

    def tanh(self):
        return Expr.fn("tanh", self)

    def cos(self):
        return Expr.fn("cos", self)

    def __bool__(self):
        raise TypeError("the truth value of a traced fluid expression (branch on Python values only)")

    def eval(self, env: dict, memo: dict = None) -> float:
        """The value with the symbols bound by ``env`` (names, ("field", f), ("fluid", j))."""
        memo = {} if memo is None else memo
        if self.nid in memo:
            return memo[self.nid]
        op, a = self.op, self.args
        if op == "const":
            v = a[0]
        elif op == "sym":
            v = env[a[0]]
        else:
            x = [t.eval(env, memo) if isinstance(t, Expr) else t for t in a]
            v = {"add": lambda: x[0] + x[1], "mul": lambda: x[0] * x[1], "div": lambda: x[0] / x[1],
                 "neg": lambda: -x[0], "powc": lambda: x[0] ** x[1], "exp": lambda: math.exp(x[0]),
                 "log": lambda: math.log(x[0]), "sqrt": lambda: math.sqrt(x[0]), "tanh": lambda: math.tanh(x[0]),
                 "cos": lambda: math.cos(x[0])}[op]()
        memo[self.nid] = v
        return v

    def __float__(self):
        if _isc(self):
            return self.args[0]
        raise TypeError("float() of a traced fluid expression")

    def __hash__(self):
        return self.nid

    def __eq__(self, o):
        return self is o

    def __repr__(self):
        return f"Expr({self.op}, {self.args!r})"


LNA = sym("lna", T)
A = sym("a", T)
KS = sym("k", K)
AH = sym("aH", T | H)
TAUS = sym("tau", TAU)
KH = sym("kh", T | K | P)   # h' weight base of rho delta: 2 K_G a^2 / aH^2
KE = sym("ke", T | K | P)   # eta' weight base of (rho + P) theta: K_G a^2 / (aH k^2)


class Lin(Traced):
    """sum_v coef_v * v + const, v in {('y', i), 'hp', 'ep'}."""
    __slots__ = ("terms", "const")

    def __init__(self, terms, const_=None):
        self.terms = {v: c for v, c in terms.items() if not _isc(c, 0.0)}
        self.const = const(0.0) if const_ is None else const_

    @staticmethod
    def mk(terms, const_):
        out = Lin(terms, const_)
        return out.const if not out.terms else out

    def __add__(self, o):
        o = lift(o)
        if isinstance(o, Lin):
            t = dict(self.terms)
            for v, c in o.terms.items():
                t[v] = Expr.add(t[v], c) if v in t else c
            return Lin.mk(t, Expr.add(self.const, o.const))
        return Lin.mk(self.terms, Expr.add(self.const, o))

    __radd__ = __add__

    def __neg__(self):
        return Lin.mk({v: Expr.neg(c) for v, c in self.terms.items()}, Expr.neg(self.const))

    def __sub__(self, o):
        return self + (-lift(o))

    def __rsub__(self, o):
        return (-self) + lift(o)

    def __mul__(self, o):
        o = lift(o)
        if isinstance(o, Lin):
            raise NotImplementedError("a product of two state-dependent values: the equations must be linear in y")
        return Lin.mk({v: Expr.mul(c, o) for v, c in self.terms.items()}, Expr.mul(self.const, o))

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = lift(o)
        if isinstance(o, Lin):
            raise NotImplementedError("division by a state-dependent value: the equations must be linear in y")
        return Lin.mk({v: Expr.div(c, o) for v, c in self.terms.items()}, Expr.div(self.const, o))

    def __pos__(self):
        return self

    def _nonlinear(self, *_):
        raise NotImplementedError("a nonlinear function of the state: the equations must be linear in y")

    exp = log = sqrt = tanh = cos = __pow__ = __rpow__ = __rtruediv__ = _nonlinear

    def __bool__(self):
        raise TypeError("the truth value of a state-dependent fluid expression")


class TVec(list, Traced):
    """``jnp.array`` of traced values."""


def as_lin(x):
    """Any traced or numeric value as (terms, const)."""
    x = lift(x)
    if isinstance(x, Lin):
        return dict(x.terms), x.const
    return {}, x


# ------------------------------------------------------------------ symbolic arguments
class _State:
    """The state vector y: y[i] is the linear form 1 * y_i."""

    def __init__(self, n):
        self.n = n

    def __getitem__(self, i):
        if isinstance(i, slice):
            return [self[j] for j in range(*i.indices(self.n))]
        i = int(i)
        if not 0 <= i < self.n:
            raise IndexError(f"state row {i} outside 0..{self.n - 1}")
        return Lin({("y", i): const(1.0)})

    def __len__(self):
        return self.n


class _Params:
    """params[name]: a cosmology field, or a parameter slot of the user fluids."""

    def __init__(self, fset):
        self.fset = fset

    def __getitem__(self, name):
        if name == "H0":
            return self["h"] * cnst.H0_over_h
        if name in PARAM_FIELDS:
            return sym(("field", PARAM_FIELDS[name]), P)
        return sym(("fluid", self.fset.param_slot(name)), P)

    def get(self, name, default=None):
        return self[name]

    def __contains__(self, name):
        return True


class _BG:
    """The background object of y_prime's args: aH only (abcmb/background.py:168-187)."""

    def aH(self, lna, params=None):
        if lna is not LNA:
            raise NotImplementedError("BG.aH at a time other than the row's own ln a")
        return AH

    def __getattr__(self, name):
        raise NotImplementedError(f"BG.{name} in a user fluid: only BG.aH(lna, params) is available")


class _NumBG:
    """aH for the numeric calls (the caller's value at its ln a)."""

    def __init__(self, lna, aH):
        self.lna, self.aHv = lna, aH

    def aH(self, lna, params=None):
        return self.aHv


# ------------------------------------------------------------------ the fluid set
class FluidSet:
    """The user fluids of one model, traced and turned into kernel code.

    ``species``: every species instance of the model in the reference's order
    (first_idx set); ``user``: the user fluids among them, whose rows are the
    last ``n_rows`` rows of the state from ``first_row``.
    """

    def __init__(self, species, user, first_row: int, n_state: int, species_off: int):
        self.species = tuple(species)
        self.user = tuple(user)
        self.first_row = first_row
        self.n_state = n_state
        self.n_rows = n_state - first_row
        self.species_off = species_off
        self.species_dict = {s.name: i for i, s in enumerate(self.species)}
        self.params = []     # fluid parameter names, in slot order
        self._trace()
        self.dev_src, self.host_src, self.signature = self._generate()

    def param_slot(self, name):
        if name not in self.params:
            if len(self.params) >= MAX_PARAMS:
                raise NotImplementedError(f"more than {MAX_PARAMS} parameters in the user fluids")
            self.params.append(name)
        return self.params.index(name)

    # ---- tracing
    def _trace(self):
        if self.n_rows > MAX_ROWS:
            raise NotImplementedError(f"{self.n_rows} user-fluid rows: at most {MAX_ROWS}")
        Pp, Y, bg = _Params(self), _State(self.n_state), _BG()
        args = (bg, Pp, self.species, self.species_dict)
        lo, hi = self.first_row, self.n_state

        def user_col(v, what):
            if v in ("hp", "ep"):
                raise NotImplementedError(f"{what} depends on the metric derivatives")
            i = v[1]
            if not lo <= i < hi:
                raise NotImplementedError(
                    f"{what} reads state row {i} of a standard species: user fluids may couple to user "
                    "fluids and to the metric only (the kernels' block solve)")
            return i - lo

        rho = pres = rho_m = const(0.0)
        wd, wt, wsg, wm = ({} for _ in range(4))   # per local row: rho delta / (rho+P) theta / sigma / matter
        rows, ics = [], []
        for f in self.user:
            r = lift(f.rho(LNA, Pp))
            p = lift(f.P(LNA, Pp))
            if isinstance(r, Lin) or isinstance(p, Lin) or (r.deps | p.deps) & (K | H | TAU):
                raise NotImplementedError(f"{f.name}: rho and P may depend on ln a and the parameters only")
            rho, pres = Expr.add(rho, r), Expr.add(pres, p)
            if f.is_matter:
                rho_m = Expr.add(rho_m, r)
            ne = int(f.num_equations)
            if ne == 0:
                continue
            for what, fn, acc in (("rho_delta", f.rho_delta, wd), ("rho_plus_P_theta", f.rho_plus_P_theta, wt),
                                  ("rho_plus_P_sigma", f.rho_plus_P_sigma, wsg)):
                terms, c0 = as_lin(fn(LNA, Y, Pp))
                if not _isc(c0, 0.0):
                    raise NotImplementedError(f"{f.name}.{what} has a term without the state")
                for v, c in terms.items():
                    if c.deps & (K | H | TAU):
                        raise NotImplementedError(f"{f.name}.{what}: coefficients of ln a and parameters only")
                    j = user_col(v, f"{f.name}.{what}")
                    acc[j] = Expr.add(acc.get(j, const(0.0)), c)
                    if what == "rho_delta" and f.is_matter:
                        wm[j] = Expr.add(wm.get(j, const(0.0)), c)
            out = f.y_prime(KS, LNA, Lin({"hp": const(1.0)}), Lin({"ep": const(1.0)}), Y, args)
            out = list(out)
            if len(out) != ne:
                raise ValueError(f"{f.name}.y_prime returned {len(out)} entries for {ne} equations")
            for i, e in enumerate(out):
                terms, c0 = as_lin(e)
                if not _isc(c0, 0.0):
                    raise NotImplementedError(f"{f.name}.y_prime[{i}] has a term without the state")
                rows.append((f.name, i, terms))
            ini = list(f.y_ini(KS, TAUS, Pp))
            if len(ini) != ne:
                raise ValueError(f"{f.name}.y_ini returned {len(ini)} entries for {ne} equations")
            for i, e in enumerate(ini):
                e = lift(e)
                if isinstance(e, Lin) or e.deps & (T | H):
                    raise NotImplementedError(f"{f.name}.y_ini[{i}]: a function of k, tau and the parameters only")
                ics.append(e)
        if len(rows) != self.n_rows:
            raise ValueError(f"user fluids declare {self.n_rows} rows, their y_prime returned {len(rows)}")

        self.slots = []
        slot_of = {}

        def slot(e):
            if e.nid not in slot_of:
                slot_of[e.nid] = len(self.slots)
                self.slots.append(e)
            return slot_of[e.nid]

        self.row_terms, self.row_metric = [], []
        for j, (name, i, terms) in enumerate(rows):
            tl, uh, ue = [], const(0.0), const(0.0)
            for v, c in sorted(terms.items(), key=lambda t: (t[0] if isinstance(t[0], str) else "y%06d" % t[0][1])):
                if v in ("hp", "ep"):
                    if c.deps & ~K:
                        raise NotImplementedError(
                            f"{name}.y_prime[{i}]: its coefficient of {'h' if v == 'hp' else 'eta'}' depends on "
                            "time or the parameters; the kernels take these couplings as constants (of k)")
                    if v == "hp":
                        uh = c
                    else:
                        ue = c
                    continue
                col = user_col(v, f"{name}.y_prime[{i}]")
                if c.deps & ~K:
                    tl.append((col, const(1.0), slot(c)))
                else:
                    tl.append((col, c, -1))
            if len(tl) > MAX_TERMS:
                raise NotImplementedError(f"{name}.y_prime[{i}] has {len(tl)} terms: at most {MAX_TERMS}")
            self.row_terms.append(tl)
            self.row_metric.append((uh, ue))
        self.h_slot = [slot(Expr.mul(KH, wd[j])) if j in wd else -1 for j in range(self.n_rows)]
        self.e_slot = [slot(Expr.mul(KE, wt[j])) if j in wt else -1 for j in range(self.n_rows)]
        self.rho, self.pres, self.rho_m = rho, pres, rho_m
        self.wd, self.wt, self.wsg, self.wm = wd, wt, wsg, wm
        self.ics = ics

    # ---- C++
    def _generate(self):
        dev = _Emitter("dev")
        host = _Emitter("host")
        n, ns = self.n_rows, len(self.slots)
        d = []
        d.append("constexpr bool kActive = true;")
        d.append(f"constexpr int kRows = {n};")
        d.append(f"constexpr int kSlots = {ns};")
        d.append("constexpr const char* kSignature = \"@SIG@\";")
        d.append("")
        d.append("template <int N, class Cos>")
        d.append("__device__ __forceinline__ void bg(const Cos& P, double lna, double a, D<N>& rho, D<N>& pres, "
                 "D<N>& rho_m) {")
        d += dev.block([("rho", self.rho), ("pres", self.pres), ("rho_m", self.rho_m)], assign=True)
        d.append("}")
        d.append("")
        d.append("template <int N, class Cos>")
        d.append("__device__ __forceinline__ void slots(const Cos& P, double lna, double a, double k, const D<N>& aH, "
                 "const D<N>& kh, const D<N>& ke, D<N>* s) {")
        d += dev.block([(f"s[{i}]", e) for i, e in enumerate(self.slots)], assign=True)
        d.append("}")
        d.append("")
        d.append("template <class Add>")
        d.append("__device__ __forceinline__ void row(int j, double k, Add add, double& uh, double& ue, int& hs, "
                 "int& es) {")
        d.append("  switch (j) {")
        for j in range(n):
            d.append(f"    case {j}: {{")
            items = [(None, c) for (_, c, _) in self.row_terms[j]] + [(None, self.row_metric[j][0]),
                                                                      (None, self.row_metric[j][1])]
            lines, vals = dev.block_values([e for _, e in items], scalar=True)
            d += ["    " + ln for ln in lines]
            for t, (col, _, sl) in enumerate(self.row_terms[j]):
                d.append(f"      add({col}, {vals[t]}, {sl});")
            d.append(f"      uh = {vals[-2]};")
            d.append(f"      ue = {vals[-1]};")
            d.append(f"      hs = {self.h_slot[j]};")
            d.append(f"      es = {self.e_slot[j]};")
            d.append("      break;")
            d.append("    }")
        d.append("    default: break;")
        d.append("  }")
        d.append("}")
        d.append("")
        d.append("template <int N, class Cos>")
        d.append("__device__ __forceinline__ D<N> ic(int j, const Cos& P, double k, const D<N>& tau) {")
        d.append("  switch (j) {")
        for j, e in enumerate(self.ics):
            lines, vals = dev.block_values([e])
            d.append(f"    case {j}: {{")
            d += ["    " + ln for ln in lines]
            d.append(f"      return D<N>(0.0) + {vals[0]};")
            d.append("    }")
        d.append("    default: return D<N>(0.0);")
        d.append("  }")
        d.append("}")
        d.append("")
        d.append("template <int N, class Cos, class Rows>")
        d.append("__device__ __forceinline__ void sums(const Cos& P, double lna, double a, const Rows& y, D<N>& srd, "
                 "D<N>& srt, D<N>& srs, D<N>& srdm, D<N>& srm) {")
        outs = []
        for tag, w in (("srd", self.wd), ("srt", self.wt), ("srs", self.wsg), ("srdm", self.wm)):
            for j in sorted(w):
                outs.append((f"{tag} = {tag} + {{}} * y.fl({j});", w[j]))
        outs.append(("srm = srm + {};", self.rho_m))
        lines, vals = dev.block_values([e for _, e in outs])
        d += lines
        for (fmt, _), v in zip(outs, vals):
            d.append("  " + fmt.format(v))
        d.append("}")
        h = []
        h.append("constexpr bool kActive = true;")
        h.append("constexpr const char* kSignature = \"@SIG@\";")
        h.append("// C: AbcmbCosmo, or the producer's cosmology over dual numbers; S: its scalar")
        h.append("template <class C, class S>")
        h.append("inline void bg(const C& c, double lna, double a, S& rho, S& pres) {")
        h.append("  using std::cos; using std::exp; using std::log; using std::pow; using std::sqrt; using std::tanh;")
        h += host.block([("rho", self.rho), ("pres", self.pres)], assign=True)
        h.append("}")
        names = ", ".join(s.name for s in self.user)
        head = ("// Generated by abcmb_amd/fluidgen.py from the user fluids {" + names + "}: do not edit.\n"
                "// Parameter slots (AbcmbCosmo.fluid_par): " + (", ".join(self.params) or "none") + ".\n")
        dev_src = head + "#pragma once\n\nnamespace fluid {\n" + "\n".join(d) + "\n}  // namespace fluid\n"
        host_src = head + "#pragma once\n#include <cmath>\n\nnamespace fluid_host {\n" + "\n".join(h) + \
            "\n}  // namespace fluid_host\n"
        sig = hashlib.sha256((dev_src + host_src).encode()).hexdigest()[:16]
        return dev_src.replace("@SIG@", sig), host_src.replace("@SIG@", sig), sig

    # ---- numbers
    def fluid_par(self, params: dict) -> np.ndarray:
        out = np.zeros(MAX_PARAMS)
        for i, nm in enumerate(self.params):
            if nm not in params:
                raise KeyError(f"parameter {nm!r} of the user fluids is missing from params")
            out[i] = float(params[nm])
        return out

    @property
    def directory(self):
        return os.path.join(FLUID_ROOT, self.signature)

    @property
    def hip_path(self):
        return os.path.join(self.directory, "libabcmb_hip.so")

    @property
    def host_path(self):
        return os.path.join(self.directory, "libabcmb_host.so")

    def write(self):
        os.makedirs(self.directory, exist_ok=True)
        for name, src in (("fluid_dev.h", self.dev_src), ("fluid_host.h", self.host_src)):
            path = os.path.join(self.directory, name)
            if not os.path.exists(path) or open(path).read() != src:
                with open(path, "w") as fh:
                    fh.write(src)
        meta = dict(signature=self.signature, fluids=[s.name for s in self.user], params=self.params,
                    rows=self.n_rows, slots=len(self.slots))
        with open(os.path.join(self.directory, "meta.json"), "w") as fh:
            json.dump(meta, fh, indent=1)

    def _stamp(self, kind):
        """Hash of everything a library of this set is compiled from (sources,
        headers, tables and the generated header), so that a source change
        rebuilds it whatever the files' times."""
        from . import build as B
        h = hashlib.sha256((self.host_src if kind == "host" else self.dev_src).encode())
        for path in B.sources(kind):
            with open(path, "rb") as fh:
                h.update(fh.read())
        return h.hexdigest()

    def _fresh(self, kind):
        lib = self.host_path if kind == "host" else self.hip_path
        stamp = os.path.join(self.directory, kind + ".stamp")
        return os.path.exists(lib) and os.path.exists(stamp) and open(stamp).read() == self._stamp(kind)

    def built(self, hip: bool = True) -> bool:
        return self._fresh("host") and (not hip or self._fresh("hip"))

    def build(self, force: bool = False, hip: bool = True):
        """Compile the host producer (seconds) and, with ``hip``, the kernel library
        (minutes: the ODE kernels are compiled again) for this fluid set, under
        ``abcmb_amd/_fluids/<signature>/``, unless built from the same sources."""
        from . import build as B
        self.write()
        for kind, run in (("host", lambda: B.build_fluid_variant(self.directory, hip=False)),
                          ("hip", lambda: B.build_fluid_variant(self.directory, host=False))):
            if kind == "hip" and not hip:
                continue
            if force or not self._fresh(kind):
                run()
                with open(os.path.join(self.directory, kind + ".stamp"), "w") as fh:
                    fh.write(self._stamp(kind))


class _Emitter:
    """C++ for a set of expressions: one ``const auto`` temporary per shared node."""

    FN_DEV = {"exp": "fluid::fexp", "log": "fluid::flog", "sqrt": "fluid::fsqrt", "tanh": None, "cos": None}
    # host code is a template over its scalar (double, or dual numbers with the
    # parameter derivatives): unqualified calls, std:: brought in by using-declarations
    FN_HOST = {"exp": "exp", "log": "log", "sqrt": "sqrt", "tanh": "tanh", "cos": "cos"}

    def __init__(self, style):
        self.style = style

    def _sym(self, name):
        if isinstance(name, tuple):
            kind, v = name
            if kind == "field":
                return f"P(F_{v.upper()})" if self.style == "dev" else f"c.{v}"
            return f"P(F_FLUID0 + {v})" if self.style == "dev" else f"c.fluid_par[{v}]"
        if self.style == "host" and name not in ("lna", "a"):
            raise NotImplementedError(f"{name} in a host expression")
        return name

    @staticmethod
    def _num(v):
        s = repr(float(v))
        return s if ("." in s or "e" in s or "inf" in s) else s + ".0"

    def block_values(self, exprs, scalar=False):
        """(lines, value strings) computing ``exprs``; temporaries for the shared nodes."""
        uses, order = {}, []

        def visit(e):
            if e.op in ("const", "sym"):
                return
            uses[e.nid] = uses.get(e.nid, 0) + 1
            if uses[e.nid] > 1:
                return
            for a in e.args:
                if isinstance(a, Expr):
                    visit(a)
            order.append(e)

        for e in exprs:
            visit(e)
        names, lines = {}, []

        def code(e):
            if e.op == "const":
                return self._num(e.args[0])
            if e.op == "sym":
                return self._sym(e.args[0])
            if e.nid in names:
                return names[e.nid]
            return self._node(e, code)

        for e in order:
            if uses[e.nid] > 1:
                nm = f"t{len(names)}"
                lines.append(f"  const auto {nm} = {self._node(e, code)};")
                names[e.nid] = nm
        return lines, [code(e) for e in exprs]

    def _node(self, e, code):
        op, a = e.op, e.args
        if op == "add":
            if a[1].op == "neg":
                return f"({code(a[0])} - {code(a[1].args[0])})"
            return f"({code(a[0])} + {code(a[1])})"
        if op == "mul":
            return f"({code(a[0])} * {code(a[1])})"
        if op == "div":
            return f"({code(a[0])} / {code(a[1])})"
        if op == "neg":
            return f"(-{code(a[0])})"
        if op == "powc":
            x, p = code(a[0]), a[1]
            if p == 2.0:
                return f"({x} * {x})"
            if p == 3.0:
                return f"({x} * {x} * {x})"
            if p == 4.0:
                return f"(({x} * {x}) * ({x} * {x}))"
            if p == -1.0:
                return f"(1.0 / {x})"
            if p == 0.5:
                return f"{'fluid::fsqrt' if self.style == 'dev' else 'sqrt'}({x})"
            return f"{'fluid::fpow' if self.style == 'dev' else 'pow'}({x}, {self._num(p)})"
        fns = self.FN_DEV if self.style == "dev" else self.FN_HOST
        if op in fns:
            if fns[op] is None:
                raise NotImplementedError(f"{op} in a fluid's kernel expression")
            return f"{fns[op]}({code(a[0])})"
        raise AssertionError(op)

    def block(self, outs, assign=True):
        lines, vals = self.block_values([e for _, e in outs])
        zero = "D<N>(0.0) + " if self.style == "dev" else ""
        return lines + [f"  {nm} = {zero}{v};" for (nm, _), v in zip(outs, vals)]


def numeric_call(f, method, *args):
    """A fluid method called with numbers (the jnp stand-in goes to numpy)."""
    return getattr(f, method)(*args)
