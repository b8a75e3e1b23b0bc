"""The jax.numpy subset that user-defined fluids are written with.

A fluid's methods (``abcmb_amd.species``) run twice in this library: once on
symbolic inputs, when ``Model`` turns them into the kernels' code
(``abcmb_amd.fluidgen``), and with plain numbers, for the derived parameters,
the output dictionaries and the CPU checks.  The reference's fluids call
``jax.numpy``; here they call this module instead (``import abcmb_amd.xp as
jnp``), whose functions dispatch on the argument: traced values build
expressions, numbers go to numpy.
"""
from __future__ import annotations

import math

import numpy as np

pi = math.pi
e = math.e


def _traced(*xs):
    from .fluidgen import Traced
    return any(isinstance(x, Traced) for x in xs)


def exp(x):
    return x.exp() if _traced(x) else np.exp(x)


def log(x):
    return x.log() if _traced(x) else np.log(x)


def sqrt(x):
    return x.sqrt() if _traced(x) else np.sqrt(x)


def tanh(x):
    return x.tanh() if _traced(x) else np.tanh(x)


def cos(x):
    return x.cos() if _traced(x) else np.cos(x)


def power(x, p):
    return x ** p if _traced(x, p) else np.power(x, p)


def array(seq, dtype=None):
    """A vector of traced values (a list) when any entry is traced, else numpy."""
    items = list(seq)
    if _traced(*items):
        from .fluidgen import TVec
        return TVec(items)
    return np.array(items, dtype=dtype if dtype is not None else np.float64)


def where(cond, a, b):
    """Python / numpy conditions select directly; traced conditions are not supported
    (the generated code has no branches on the state)."""
    if _traced(cond):
        raise NotImplementedError("jnp.where on a traced condition: fluids may branch on Python values only")
    if isinstance(cond, (bool, np.bool_)) or np.ndim(cond) == 0:
        return a if bool(cond) else b
    return np.where(cond, a, b)


def concatenate(parts):
    out = []
    for p in parts:
        out.extend(list(p))
    return array(out)


def zeros(n):
    return np.zeros(n)
