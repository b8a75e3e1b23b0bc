"""ctypes wrapper of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` use this module, and only as the checker.  The functions
restate the reference's hot path in FP64 C (``oracle/abcmb_oracle.c``, with
the reference file:line of each piece cited there).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ABCMB_ORACLE_LIB points the checker at another build of the same source (a
# mutated copy, to see that the pins fail on a planted mistake)
LIB_PATH = os.environ.get("ABCMB_ORACLE_LIB", os.path.join(_HERE, "liboracle.so"))
_lib = None
_P = C.c_void_p
_D = C.c_double
_I = C.c_int


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        from abcmb_amd import abi
        L = C.CDLL(LIB_PATH)
        L.oracle_pert.argtypes = [_P, _P, _I, C.POINTER(abi.AbcmbOdeSettings), _P, _P, _P]
        L.oracle_pert_scripted.argtypes = [_P, _P, _P, _I, C.POINTER(abi.AbcmbOdeSettings), _I, _P, _P, _I, _P,
                                           _P]
        L.oracle_pert_tangent.argtypes = [_P, _P, _P, _P, _I, _P, _I, C.POINTER(abi.AbcmbOdeSettings), _P, _P, _P]
        L.oracle_rhs.argtypes = [C.POINTER(abi.AbcmbCosmo), C.POINTER(abi.AbcmbOdeSettings), _D, _D, _P, _P]
        L.oracle_ic.argtypes = [C.POINTER(abi.AbcmbCosmo), C.POINTER(abi.AbcmbOdeSettings), _D, _D, _P]
        L.oracle_bg.argtypes = [C.POINTER(abi.AbcmbCosmo), _D, _P]
        L.oracle_kvaerno5_tableau.argtypes = [_P, _P, _P, _P]
        L.oracle_spline.argtypes = [_P, _P, _I, _P, _I, _P]
        L.oracle_sources.argtypes = [_P, _P, _I, _I, _P, _P, _P, _I, _P, _P]
        L.oracle_los_cl.argtypes = [_P, _I, _P, _D, _P, _I, _P, _P, _P, _I, _I, _P, _I, _D, _D, _D, _P, _P]
        L.oracle_los_cl_branch.argtypes = [_P, _I, _P, _D, _P, _I, _P, _P, _P, _I, _I, _P, _I, _D, _D, _D, _P, _P,
                                           _P]
        L.oracle_pk.argtypes = [_P, _P, _I, _I, _P, _P, _I, _D, _D, _D, _P]
        L.oracle_clpp.argtypes = [C.POINTER(abi.AbcmbCosmo), _P, _P, _P, _P, _I, _D, _P, _I, _P]
        L.oracle_wigner.argtypes = [_I, _I, _P, _I, _I, _P]
        L.oracle_lens.argtypes = [_P, _P, _I, _P, _P, _I, _P]
        L.oracle_set_threads.argtypes = [_I]
        L.oracle_pk_z.argtypes = [_P, _P, _I, _I, _P, _P, _I, _D, _D, _D, _D, _I, _D, _D, _D, _P]
        L.oracle_set_fluids.argtypes = [_P]
        _lib = L
    return _lib


_BG_CB = C.CFUNCTYPE(None, _P, _P, _D, C.POINTER(_D))
_SUMS_CB = C.CFUNCTYPE(None, _P, _P, _D, C.POINTER(_D), _I, C.POINTER(_D))
_RHS_CB = C.CFUNCTYPE(None, _P, _P, _D, _D, _D, _D, _D, C.POINTER(_D), _I, C.POINTER(_D))
_IC_CB = C.CFUNCTYPE(None, _P, _P, _D, _D, C.POINTER(_D))


class _OracleFluids(C.Structure):
    _fields_ = [("n_rows", _I), ("ctx", _P), ("bg", _BG_CB), ("sums", _SUMS_CB), ("rhs", _RHS_CB), ("ic", _IC_CB)]


class _NumBG:
    """The background argument of y_prime in the numeric calls: aH as the oracle has it."""

    def __init__(self, aH):
        self.v = aH

    def aH(self, lna, params=None):
        return self.v


class fluids:
    """``with fluids(model, params):`` -- the oracle evaluates the model's user
    fluids (abcmb_amd.species subclasses) by calling their methods with
    numbers, as the reference calls them (abcmb/perturbations.py:201-248,
    :318-404; abcmb/main.py:540-594 for the background), through C callbacks.
    The oracle runs single-threaded meanwhile.  The parameters the methods see
    are ``params`` with the cosmology fields (and the user fluids' parameter
    slots) of the cosmology the oracle evaluates, so that the tangent-linear
    integration's neighbouring cosmologies move them."""

    def __init__(self, model, params):
        from abcmb_amd import constants, fluidgen
        self.model, self.p = model, dict(params)
        fs = model.fluids
        self.user = [f for f in fs.user] if fs is not None else []
        self.rows = [f for f in self.user if int(f.num_equations) > 0]
        self.n_rows = model.n_fluid
        self.names = list(fs.params) if fs is not None else []
        self.fields = fluidgen.PARAM_FIELDS
        self.H0 = constants.H0_over_h
        self.args_tail = (model.species_objs, {o.name: i for i, o in enumerate(model.species_objs)})
        self._cache = {}

    def _params(self, cptr):
        from abcmb_amd import abi
        c = abi.AbcmbCosmo.from_address(cptr)
        key = bytes(c)[:abi.COSMO_DOUBLES * 8]
        p = self._cache.get(key)
        if p is None:
            p = dict(self.p)
            for name, fld in self.fields.items():
                p[name] = getattr(c, fld)
            p["H0"] = p["h"] * self.H0
            for j, nm in enumerate(self.names):
                p[nm] = c.fluid_par[j]
            if len(self._cache) > 64:
                self._cache.clear()
            self._cache[key] = p
        return p

    def _bg(self, ctx, cptr, lna, out):
        p = self._params(cptr)
        r = pr = m = 0.0
        for f in self.user:
            rv = float(f.rho(lna, p))
            r += rv
            pr += float(f.P(lna, p))
            if f.is_matter:
                m += rv
        out[0], out[1], out[2] = r, pr, m

    def _sums(self, ctx, cptr, lna, y, n, out):
        p = self._params(cptr)
        yv = np.ctypeslib.as_array(y, (n,))
        s = [0.0, 0.0, 0.0, 0.0]
        for f in self.rows:
            d = float(f.rho_delta(lna, yv, p))
            s[0] += d
            s[1] += float(f.rho_plus_P_theta(lna, yv, p))
            s[2] += float(f.rho_plus_P_sigma(lna, yv, p))
            if f.is_matter:
                s[3] += d
        for i in range(4):
            out[i] = s[i]

    def _rhs(self, ctx, cptr, k, lna, aH, hp, ep, y, n, dy):
        p = self._params(cptr)
        yv = np.ctypeslib.as_array(y, (n,))
        args = (_NumBG(aH), p) + self.args_tail
        vals = np.concatenate([np.asarray(f.y_prime(k, lna, hp, ep, yv, args), dtype=np.float64)
                               for f in self.rows])
        for i, v in enumerate(vals):
            dy[i] = v

    def _ic(self, ctx, cptr, k, tau, y):
        p = self._params(cptr)
        vals = np.concatenate([np.asarray(f.y_ini(k, tau, p), dtype=np.float64) for f in self.rows])
        for i, v in enumerate(vals):
            y[i] = v

    def _guard(self, fn, out_len):
        """A callback that records an exception (raised again on exit) and hands
        the oracle NaNs instead of returning half-written outputs."""
        def cb(*a):
            try:
                fn(*a)
            except Exception as e:   # noqa: BLE001 -- re-raised in __exit__
                if self.error is None:
                    self.error = e
                out, n = a[-1], out_len(a)
                for i in range(n):
                    out[i] = float("nan")
        return cb

    def __enter__(self):
        self.error = None
        self._cbs = (_BG_CB(self._guard(self._bg, lambda a: 3)), _SUMS_CB(self._guard(self._sums, lambda a: 4)),
                     _RHS_CB(self._guard(self._rhs, lambda a: self.n_rows)),
                     _IC_CB(self._guard(self._ic, lambda a: self.n_rows)))
        self._st = _OracleFluids(self.n_rows, None, *self._cbs)
        self._threads = lib().oracle_max_threads()
        lib().oracle_set_threads(1)
        lib().oracle_set_fluids(C.byref(self._st))
        return self

    def __exit__(self, *exc):
        lib().oracle_set_fluids(None)
        lib().oracle_set_threads(self._threads)
        if self.error is not None and exc[0] is None:
            raise RuntimeError("a user-fluid callback of the oracle failed") from self.error
        return False


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def kvaerno5_tableau():
    A = np.zeros(49); b = np.zeros(7); bh = np.zeros(7); c = np.zeros(7)
    lib().oracle_kvaerno5_tableau(_p(A), _p(b), _p(bh), _p(c))
    return A.reshape(7, 7), b, bh, c


def rhs(cosmo, settings, k, lna, y):
    y = _c(y); dy = np.zeros_like(y)
    lib().oracle_rhs(C.byref(cosmo), C.byref(settings), k, lna, _p(y), _p(dy))
    return dy


def ic(cosmo, settings, k, lna):
    y = np.zeros(settings.n)
    lib().oracle_ic(C.byref(cosmo), C.byref(settings), k, lna, _p(y))
    return y


def bg(cosmo, lna):
    out = np.zeros(6)
    lib().oracle_bg(C.byref(cosmo), lna, _p(out))
    return dict(zip(("aH", "tau", "tau_c", "cs2", "xe", "Tm"), out))


def pert(cosmos_arr, items_arr, settings, want_states=False):
    """cosmos_arr: ctypes array of AbcmbCosmo (host pointers, pt set); items_arr: ctypes array."""
    n = len(items_arr)
    status = np.zeros(n, np.int32)
    steps = np.zeros(n, np.int32)
    states = None
    if want_states:
        ns = cosmos_arr[0].n_save
        states = np.zeros((n, ns, settings.n))
    lib().oracle_pert(C.cast(cosmos_arr, C.c_void_p), C.cast(items_arr, C.c_void_p), n,
                      C.byref(settings), _p(status), _p(steps), _p(states))
    return status, steps, states


SCRIPT_W = 9   # doubles per recorded step: hh, accepted, chord iterations of stages 1..6, t + hh


def pert_scripted(cosmos_arr, items_arr, settings, mode, scripts, lens, jac_cosmos=None):
    """oracle_pert that records (mode 1) or replays each item's step script.

    Mode 2 replays the recorded chord-iteration counts, mode 3 solves every
    replayed stage to rounding.  scripts: float64 [n_items][cap][SCRIPT_W],
    lens: int32 [n_items] (written when recording); jac_cosmos: cosmologies
    whose chord matrices the replay uses (default: the integrated ones).
    """
    n = len(items_arr)
    status = np.zeros(n, np.int32)
    steps = np.zeros(n, np.int32)
    assert scripts.dtype == np.float64 and scripts.flags.c_contiguous and scripts.shape[0] == n
    assert lens.dtype == np.int32 and lens.size == n
    jc = C.cast(jac_cosmos, C.c_void_p) if jac_cosmos is not None else None
    lib().oracle_pert_scripted(C.cast(cosmos_arr, C.c_void_p), jc, C.cast(items_arr, C.c_void_p), n,
                               C.byref(settings), mode, _p(scripts), _p(lens), scripts.shape[1],
                               _p(status), _p(steps))
    return status, steps


def pert_tangent(base_arr, plus_arr, minus_arr, h, items_arr, settings, dpt):
    """Tangent-linear tables: values into base_arr[0].pt, derivatives into dpt [nd][21][n_save][nk]."""
    n = len(items_arr)
    h = _c(h)
    status = np.zeros(n, np.int32)
    steps = np.zeros(n, np.int32)
    assert dpt.dtype == np.float64 and dpt.flags.c_contiguous and dpt.shape[0] == h.size
    lib().oracle_pert_tangent(C.cast(base_arr, C.c_void_p), C.cast(plus_arr, C.c_void_p),
                              C.cast(minus_arr, C.c_void_p), _p(h), h.size, C.cast(items_arr, C.c_void_p), n,
                              C.byref(settings), _p(dpt), _p(status), _p(steps))
    return status, steps


def spline(x, y, xq):
    x = _c(x); y = _c(y); xq = _c(xq); out = np.zeros(xq.size)
    lib().oracle_spline(_p(x), _p(y), x.size, _p(xq), xq.size, _p(out))
    return out


def sources(pt, k, n_save, bgv, lna_save, kt, scales=(1, 1, 1, 1)):
    pt = _c(pt); k = _c(k); bgv = _c(bgv); lna_save = _c(lna_save); kt = _c(kt)
    sc = _c(scales)
    out = np.zeros((4, n_save - 1, kt.size))
    lib().oracle_sources(_p(pt), _p(k), k.size, n_save, _p(bgv), _p(lna_save), _p(kt), kt.size, _p(sc), _p(out))
    return out


def los_cl(src, tau_los, tau0, kt, bes, rng, bes_l, ell_idx, A_s, n_s, k_pivot, want_delta=False):
    src = _c(src); tau_los = _c(tau_los); kt = _c(kt); bes = _c(bes); rng = _c(rng)
    bes_l = _c(bes_l, np.int32); ell_idx = _c(ell_idx, np.int32)
    n_ell = ell_idx.size
    cl = np.zeros((3, n_ell))
    delta = np.zeros((4, n_ell, kt.size)) if want_delta else None
    lib().oracle_los_cl(_p(src), src.shape[1], _p(tau_los), tau0, _p(kt), kt.size, _p(bes), _p(rng), _p(bes_l),
                        bes.shape[1], bes.shape[2], _p(ell_idx), n_ell, A_s, n_s, k_pivot, _p(cl), _p(delta))
    return (cl, delta) if want_delta else cl


def los_cl_branch(src, tau_los, tau0, kt, bes, rng, bes_l, ell_idx, A_s, n_s, k_pivot, chi_branch):
    """los_cl with the Bessel branches (zero / table / asymptotic form) picked at
    chi_branch[s] * k instead of (tau0 - tau_los[s]) * k."""
    src = _c(src); tau_los = _c(tau_los); kt = _c(kt); bes = _c(bes); rng = _c(rng)
    bes_l = _c(bes_l, np.int32); ell_idx = _c(ell_idx, np.int32); chi_branch = _c(chi_branch)
    assert chi_branch.size == src.shape[1]
    n_ell = ell_idx.size
    cl = np.zeros((3, n_ell))
    lib().oracle_los_cl_branch(_p(src), src.shape[1], _p(tau_los), tau0, _p(kt), kt.size, _p(bes), _p(rng),
                               _p(bes_l), bes.shape[1], bes.shape[2], _p(ell_idx), n_ell, A_s, n_s, k_pivot, _p(cl),
                               None, _p(chi_branch))
    return cl


def pk(pt, k, n_save, lna_save, kq, A_s, n_s, k_pivot):
    pt = _c(pt); k = _c(k); lna_save = _c(lna_save); kq = _c(kq); out = np.zeros(kq.size)
    lib().oracle_pk(_p(pt), _p(k), k.size, n_save, _p(lna_save), _p(kq), kq.size, A_s, n_s, k_pivot, _p(out))
    return out


def pk_z(pt, k, n_save, lna_save, kq, A_s, n_s, k_pivot, z, kind, omega_b, omega_cdm, omega_m):
    """P(k, z): kind 0 total matter, kind 1 CDM + baryons."""
    pt = _c(pt); k = _c(k); lna_save = _c(lna_save); kq = _c(kq); out = np.zeros(kq.size)
    lib().oracle_pk_z(_p(pt), _p(k), k.size, n_save, _p(lna_save), _p(kq), kq.size, A_s, n_s, k_pivot,
                      -np.log1p(z), kind, omega_b, omega_cdm, omega_m, _p(out))
    return out


def clpp(cosmo, pt, lna_ax, aH_ax, tau_ax, chi_rec, ells):
    pt = _c(pt); lna_ax = _c(lna_ax); aH_ax = _c(aH_ax); tau_ax = _c(tau_ax); ells = _c(ells)
    out = np.zeros(ells.size)
    lib().oracle_clpp(C.byref(cosmo), _p(pt), _p(lna_ax), _p(aH_ax), _p(tau_ax), lna_ax.size, chi_rec,
                      _p(ells), ells.size, _p(out))
    return out


def wigner(m, n, mu, n_l):
    mu = _c(mu); out = np.zeros((mu.size, n_l))
    lib().oracle_wigner(m, n, _p(mu), mu.size, n_l, _p(out))
    return out


def lens(cl_in, clpp_, mu, w):
    cl_in = _c(cl_in); clpp_ = _c(clpp_); mu = _c(mu); w = _c(w)
    n_l = clpp_.size
    out = np.zeros((4, n_l))
    lib().oracle_lens(_p(cl_in), _p(clpp_), n_l, _p(mu), _p(w), mu.size, _p(out))
    return out


def set_threads(n: int):
    """OpenMP threads of the oracle's parallel loops."""
    if lib().oracle_set_threads(int(n)) != 0:
        raise ValueError(f"bad thread count {n}")


def max_threads() -> int:
    return int(lib().oracle_max_threads())
