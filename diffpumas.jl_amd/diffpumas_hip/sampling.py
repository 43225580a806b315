"""Batched samplers of the reference's exported API that the transport path
no longer calls: sample_mixture_material (src/Physics.jl:1331-1365) and the
legacy sample_del_event (src/Straggling.jl:217-317).  Sample i draws from
Philox keyed (seed, i) on the GPU."""
import ctypes as C

import numpy as np

from . import _lib
from ._lib import f64, ptr


def sample_mixture_material(physics, materials, fractions, energies, seed=0, ctx=None):
    """Material index (0-based) of an interaction at each energy in the
    mixture (materials, fractions)."""
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    m = np.ascontiguousarray(materials, dtype=np.int32)
    fr = f64(fractions)
    if m.shape != fr.shape:
        raise _lib.ArgumentError("materials and fractions must have the same length")
    e = f64(energies).reshape(-1)
    out = np.zeros(len(e), np.int32)
    _lib.check(ctx.lib.dp_sample_mixture_material(ctx.h, h, len(m), ptr(m, C.c_int32), ptr(fr),
                                                  len(e), ptr(e), int(seed),
                                                  ptr(out, C.c_int32)))
    return out


def sample_del_event(physics, material, ki, kf, dX, Xtot, ratio, seed=0, backward=False,
                     ctx=None):
    """(occurred, X_del, k_del, process) per step, arrays broadcast to one
    length; process 1 brems, 2 pair, 3 photonuclear, 0 when no event."""
    ctx = ctx or _lib.default_context()
    h = physics.device_handle(ctx)
    cols = np.broadcast_arrays(*(np.asarray(x, dtype=np.float64) for x in (ki, kf, dX, Xtot, ratio)))
    cols = [np.ascontiguousarray(c.reshape(-1)) for c in cols]
    n = len(cols[0])
    occ = np.zeros(n, np.uint8)
    X, k = np.zeros(n), np.zeros(n)
    proc = np.zeros(n, np.int32)
    _lib.check(ctx.lib.dp_sample_del_event(ctx.h, h, int(material), n, *[ptr(c) for c in cols],
                                           1 if backward else 0, int(seed), ptr(occ, C.c_uint8),
                                           ptr(X), ptr(k), ptr(proc, C.c_int32)))
    return occ.astype(bool), X, k, proc
