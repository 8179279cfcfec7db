"""Autotuner of the launch settings of an ensemble run.

The library's kernels give the same bits under every setting (cl_tuning: lanes per group of the orbit kernel, the
two single-frame passes of a two-frame map chain, the wave-specialised pipeline of small QR ensembles), so the fastest
one can be chosen by measurement alone. `tune` times short launches of the caller's own ensemble (its system, chain,
orbit count and element type) under each candidate on the device, picks the fastest and sets it on the library
context. tau is a parameter of the results (MEGNO and SALI depend on it), so it is tuned only over the values the
caller allows; a QR spectrum alone is tau-invariant up to rounding within tau_max.

No reference counterpart: the reference plans its launches from a measured one-wave throughput record
(ext/ChaoslabCUDAExt.jl:42-62, src/Tangent/propagate.jl:144-193).
"""
import ctypes
import itertools
import json
import os
import time

import numpy as np

from . import _lib as L
from .tangent import _context, make_chain, ELEMENT_TYPES

BLOCKS = (64, 128, 256)


def set_tuning(block=0, split=0, pipe=0, chain_qr=0):
    """Set the context's launch settings (0: the built-in rule; split/pipe 1 on, 2 off; chain_qr 1 the coupled chain's
    blocked HouseholderQR on the matrix cores, 2 its lane-butterfly sweep)."""
    lib, ctx = _context()
    t = L.cl_tuning(int(block), int(split), int(pipe), int(chain_qr))
    rc = lib.cl_set_tuning(ctx, ctypes.byref(t))
    if rc:
        raise L.HIPError(rc, lib.cl_last_error(ctx).decode())


def get_tuning():
    lib, ctx = _context()
    t = L.cl_tuning()
    lib.cl_get_tuning(ctx, ctypes.byref(t))
    return dict(block=t.block, split=t.split, pipe=t.pipe, chain_qr=t.chain_qr)


def _candidates(sysd, nf, pipe_ok, taus):
    maps = sysd.kind in (L.CL_SYS_STANDARD_MAP, L.CL_SYS_WRAPPED_STANDARD_MAP, L.CL_SYS_STANDARD_MAP_K)
    splits = (1, 2) if nf > 1 and maps else (0,)
    pipes = (1, 2) if pipe_ok else (0,)
    return [dict(block=b, split=s, pipe=p, tau=t) for b, s, p, t in itertools.product(BLOCKS, splits, pipes, taus)]


def tune(sys, x0, *, k, tau, reducers, reducers2=(), trial_steps=None, reps=2, tau_allowed=None, elt="f64",
         seed=1, log=None, cache=None):
    """Time every candidate setting on short launches of this ensemble and set the fastest on the context.

    x0: the ensemble's starts (m, s); trial_steps: steps per timed launch (default about 2e8 orbit-steps / m, at least
    2 tau); tau_allowed: the tau values the caller accepts (default: tau only). Returns dict(best, default, table,
    gain) with times in ns per orbit-step; `gain` is the default setting's time over the best one's. With `cache` (a
    JSON path) the choice is read from it when the key (device, system, chain, orbits, k, element type, taus) is there,
    and written after a measurement."""
    import torch
    from .tangent import device_name
    lib, ctx = _context()
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x0.shape
    d = sys.dimension
    E = ELEMENT_TYPES[elt]
    chain, allr = make_chain(reducers, reducers2)
    nf = 2 if reducers2 else 1
    sysd = sys.descriptor()
    taus = sorted(set([int(tau)] + [int(t) for t in (tau_allowed or ())]))
    key = json.dumps([device_name(), repr(sys), [type(r).__name__ for r, _ in allr], m, k, elt, taus])
    if cache and os.path.exists(cache):
        db = json.load(open(cache))
        if key in db:
            best = db[key]["best"]
            set_tuning(best["block"], best["split"], best["pipe"], get_tuning()["chain_qr"])
            return dict(db[key], cached=True)
    words = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), E, d, k) // 8
    dev = torch.device("cuda")
    xm = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    dx = torch.empty_like(xm)
    W = torch.empty((d * k, m), dtype=torch.float64, device=dev)
    W2 = torch.empty_like(W) if nf > 1 else None
    st = torch.empty((max(words, 1), m), dtype=torch.int64, device=dev)
    code = torch.zeros(m, dtype=torch.int8, device=dev)
    h = torch.zeros(m, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream()
    if trial_steps is None:
        trial_steps = max(2 * max(taus), int(2e8 // max(m, 1)))
    lcm = int(np.lcm.reduce(taus))
    trial_steps = max(lcm, trial_steps // lcm * lcm)
    pipe_ok = nf == 1 and d == 2 and len(reducers) == 1 and type(reducers[0]).__name__ == "QRSpectrum" and m <= 16384
    old = get_tuning()

    def run(c):
        set_tuning(c["block"], c["split"], c["pipe"], old["chain_qr"])
        dx.copy_(xm)
        assert lib.cl_initial_deviations_device(ctx, seed, 1, m, d, k, W.data_ptr(), stream.cuda_stream) == 0
        if W2 is not None:
            W2.copy_(W)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), E, m, s, d, k, 0, trial_steps,
                                     trial_steps, c["tau"], dx.data_ptr(), W.data_ptr(),
                                     W2.data_ptr() if W2 is not None else None, st.data_ptr(), code.data_ptr(),
                                     h.data_ptr(), stream.cuda_stream)
        e1.record(stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
        return e0, e1

    table = []
    t0 = time.perf_counter()
    try:
        for c in _candidates(sysd, nf, pipe_ok, taus) + [dict(block=0, split=0, pipe=0, tau=int(tau))]:
            run(c)  # warm-up of this route
            evs = [run(c) for _ in range(reps)]
            torch.cuda.synchronize()
            ms = min(a.elapsed_time(b) for a, b in evs)
            table.append(dict(c, ns_per_orbit_step=ms * 1e6 / (m * trial_steps)))
    finally:
        set_tuning(old["block"], old["split"], old["pipe"], old["chain_qr"])
    default = table[-1]
    best = min(table[:-1], key=lambda r: r["ns_per_orbit_step"])
    set_tuning(best["block"], best["split"], best["pipe"], old["chain_qr"])
    out = dict(best=best, default=default, gain=default["ns_per_orbit_step"] / best["ns_per_orbit_step"],
               table=table, trial_steps=trial_steps, orbits=m, seconds=round(time.perf_counter() - t0, 2))
    if log:
        log(f"autotune: best {best} ({out['gain']:.3f}x the built-in rule) in {out['seconds']} s")
    if cache:
        db = json.load(open(cache)) if os.path.exists(cache) else {}
        db[key] = {k_: v for k_, v in out.items() if k_ != "cached"}
        json.dump(db, open(cache, "w"))
    return out
