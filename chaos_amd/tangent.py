"""`propagate`: the host mirror of Chaoslab's ensemble entry point on the MI355X library.

`propagate(sys, x0, N; k, tau, seed, reducers, W0, ...)` keeps the reference's argument meaning,
validation and error behaviour (src/Tangent/propagate.jl:279-290): N must be a positive multiple of
tau, 1 <= k <= d, one W0 matrix per orbit; orbit ids are 1..len(x0); a non-finite state or frame,
or a failed reducer, raises NumericalHealthError for the lowest failing orbit id
(src/Tangent/propagate.jl:110-113). The work runs on the GPU through the C ABI; torch provides the
device buffers and the stream. There is no CPU path: without the library or a device this raises.

Extension over the reference: `reducers2` runs a second chain on a second frame of the same orbit
(same seed frame, same Jacobians), so a replacing indicator chain (MEGNO/FLI/OFLI/SALI) and a QR
spectrum can share one pass; each chain's results are bitwise those of a separate run.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L


class NumericalHealthError(Exception):
    """The reference's NumericalHealthError (src/Foundation/Foundation.jl:21-36)."""

    def __init__(self, what, value, threshold, step, orbit, backend="gpu", column=0):
        super().__init__(f"NumericalHealthError: {what} = {value} (threshold {threshold}) at step {step}, "
                         f"orbit {orbit}, backend {backend}")
        self.what, self.value, self.threshold = what, value, threshold
        self.step, self.orbit, self.backend, self.column = step, orbit, backend, column


@dataclass
class PropagationResult:
    """Final points `x` (m, s), frames `W` (m, d, k), per-reducer summaries `finals` (a list with one
    dict of per-orbit arrays per reducer, chain order), and the run `provenance`."""
    x: np.ndarray
    W: np.ndarray
    finals: list
    provenance: dict
    W2: np.ndarray = None
    code: np.ndarray = None
    hstep: np.ndarray = None
    states: np.ndarray = field(default=None, repr=False)
    workspace: np.ndarray = field(default=None, repr=False)   # (m, entries) UInt32 of a workspace reducer


_ctx = None


def _context():
    global _ctx
    import torch
    if not torch.cuda.is_available():
        raise L.HIPError(2, "no functional HIP device (torch.cuda.is_available() is False)")
    lib = L.load()
    if _ctx is None:
        torch.cuda.init()
        c = lib.cl_create(torch.cuda.current_device())
        if not c:
            raise L.HIPError(2, lib.cl_last_error(None).decode())
        _ctx = c
    return lib, _ctx


def make_chain(reducers, reducers2=(), reset_frame=False):
    ch = L.cl_chain()
    ch.flags = L.CL_CHAIN_RESET_FRAME if reset_frame else 0
    allr = [(r, 0) for r in reducers] + [(r, 1) for r in reducers2]
    if len(allr) > L.CL_MAX_REDUCERS:
        raise ValueError(f"at most {L.CL_MAX_REDUCERS} reducers per run")
    ch.n = len(allr)
    for i, (r, f) in enumerate(allr):
        ch.r[i] = r.descriptor(f)
    return ch, allr


def state_layout(allr, d, k):
    """Byte offset and numpy dtype of each reducer's state in the chain's state tuple."""
    out, off = [], 0
    for r, _ in allr:
        dt = r.dtype(d, k)
        out.append((off, dt))
        off += dt.itemsize
    return out, off


def decode_states(raw, allr, d, k):
    """Per-reducer structured arrays from the (m, bytes) uint8 state rows."""
    lay, total = state_layout(allr, d, k)
    assert raw.shape[1] == total, (raw.shape, total)
    return [np.ascontiguousarray(raw[:, o:o + dt.itemsize]).view(dt)[:, 0] for o, dt in lay]


def finals_of(states, allr, d, k, ws=None):
    """reducer_final of each reducer; the workspace reducer gets the orbit rows of the workspace `ws`."""
    out = []
    for (r, _), st in zip(allr, states):
        if r.kind == L.CL_RED_GALI:
            out.append(r.final(st, d))
        elif hasattr(r, "workspace"):
            out.append(r.final(st, ws))
        else:
            out.append(r.final(st))
    return out


def workspace_entries(allr, d, k):
    """_chain_workspace (src/Tangent/orbit.jl:8-18): the entries of the one workspace reducer, 0 without one."""
    ws = [r.workspace(d, k) for r, _ in allr if hasattr(r, "workspace")]
    if len(ws) > 1:
        raise ValueError(f"at most one reducer of a chain may request a workspace, got {len(ws)}")
    return ws[0] if ws else 0


def check_health(code, hstep, x, W, states, allr, orbit0=1):
    """Raise NumericalHealthError for the lowest failing orbit (src/Tangent/propagate.jl:111-113)."""
    bad = np.nonzero(code != 0)[0]
    if bad.size == 0:
        return
    i = int(bad[0])
    c = int(code[i])
    orbit = orbit0 + i
    if c == 3:
        for (r, _), st in zip(allr, states):
            if st["code"][i] != 0:
                what, v, thr, step, col = r.error(st[i])
                raise NumericalHealthError(what, v, thr, step, orbit, "gpu", col)
    vals = x[i] if c == 1 else W[i]
    vals = np.asarray(vals).ravel()
    nf = vals[~np.isfinite(vals)]
    first = float(nf[0]) if nf.size else float("nan")
    raise NumericalHealthError("state" if c == 1 else "deviation matrix", first, math.inf, int(hstep[i]), orbit, "gpu")


ELEMENT_TYPES = {"f64": L.CL_ELT_F64, "f32tan": L.CL_ELT_F64_F32TAN, "dd": L.CL_ELT_DD}
PRECISION_LABELS = {"f64": "Float64", "f32tan": "Float64 orbit, Float32 tangents",
                    "dd": "compensated orbit (value + correction), Float64 tangents"}


def propagate(sys, x0, N, *, k, tau=1, seed, reducers=(), reducers2=(), W0=None, launch_steps=None,
              check=True, keep_states=False, reset_frame=False, elt="f64"):
    """Propagate every orbit of `x0` (shape (m, s)) for N steps with k deviation vectors on the GPU.
    reset_frame=True restarts the frame from the unit columns after every handoff (the indicator recorder's window
    kinds, src/Experiments/run_indicator_device.jl:66-80); W0 is then the unit columns.
    elt="f32tan" carries the frames in Float32 between handoffs (FP64 orbit and Jacobian, FP64 reducers: the
    orthonormalisation and the compensated log sums), on the standard-map family and the coupled chain. elt="dd" carries the orbit of the
    Henon-Heiles Gauss map as a compensated pair (value, correction), the reference's compensated summation; x0 may
    give the values only (the corrections start at 0) and the returned x has both halves."""
    import torch
    if elt not in ELEMENT_TYPES:
        raise ValueError(f"elt must be one of {sorted(ELEMENT_TYPES)}, got {elt!r}")
    E = ELEMENT_TYPES[elt]
    N, tau, k = int(N), int(tau), int(k)
    if not (tau >= 1 and N >= 1 and N % tau == 0):
        raise ValueError(f"N must be a positive multiple of tau, got N = {N}, tau = {tau}")
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    if x0.ndim != 2:
        raise ValueError("x0 must have shape (orbits, state length)")
    m, s = x0.shape
    d = sys.dimension
    if elt == "dd" and s == sys.state_length:
        x0 = np.ascontiguousarray(np.concatenate([x0, np.zeros_like(x0)], axis=1))
        m, s = x0.shape
    if s < d:
        raise ValueError(f"state length {s} is below the system dimension {d}")
    if elt == "dd" and s == 2 * sys.state_length:
        pass
    elif s != sys.state_length:
        raise ValueError(f"state length {s} does not match {sys!r} (needs {sys.state_length})")
    if not (1 <= k <= d):
        raise ValueError(f"k must be in 1:{d}, got {k}")
    if W0 is not None and len(W0) != m:
        raise ValueError("W0 needs one matrix per orbit")
    if launch_steps is not None and not (launch_steps >= 1 and launch_steps % tau == 0):
        raise ValueError(f"forced steps per launch must be a positive multiple of tau, got L = {launch_steps}, tau = {tau}")
    for r in list(reducers) + list(reducers2):
        r.check(d, k)
    lib, ctx = _context()
    sysd = sys.descriptor()
    for r in list(reducers) + list(reducers2):
        if getattr(r, "sys", None) is not None and bytes(r.sys.descriptor()) != bytes(sysd):
            raise ValueError(f"{type(r).__name__} monitors {r.sys!r}, not the propagated {sys!r}")
    chain, allr = make_chain(reducers, reducers2, reset_frame)
    if reset_frame:
        W0 = np.broadcast_to(np.eye(d)[:, :k], (m, d, k))
    nws = workspace_entries(allr, d, k)
    nbytes = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), E, d, k)
    if nbytes < 0:
        raise L.HIPError(-nbytes, "no kernel instance for this system, k and reducer chain")
    words = nbytes // 8
    nf = 2 if reducers2 else 1
    dev = torch.device("cuda")
    if sysd.kind == L.CL_SYS_STANDARD_MAP_CHAIN and nws:
        raise L.HIPError(3, "chain kernels run no workspace reducer")
    if sysd.kind == L.CL_SYS_STANDARD_MAP_CHAIN:
        if E not in (L.CL_ELT_F64, L.CL_ELT_F64_F32TAN):
            raise L.HIPError(3, "chain kernels are built for Float64 and Float32 tangents")
        if reducers2:
            raise L.HIPError(3, "chain kernels carry one frame")
        return _propagate_orbit_major(lib, ctx, sys, sysd, chain, allr, x0, N, k, tau, seed, W0, launch_steps,
                                      words, check, keep_states, elt=elt)
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    dW = torch.empty((d * k, m), dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    if W0 is None:
        rc = lib.cl_initial_deviations_device(ctx, int(seed) & (2 ** 64 - 1), 1, m, d, k, dW.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    else:
        w = np.asarray(W0, dtype=np.float64).reshape(m, d, k)
        dW.copy_(torch.from_numpy(np.ascontiguousarray(w.transpose(2, 1, 0).reshape(k * d, m))))
    dW2 = dW.clone() if nf > 1 else None
    dst = torch.zeros((max(words, 1), m), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    # the workspace [entries][m] UInt32 (held as int32: torch has no uint32 arithmetic; the bits are the same)
    dws = torch.empty((nws, m), dtype=torch.int32, device=dev) if nws else None
    L_ = N if launch_steps is None else min(int(launch_steps), N)
    n0, launches = 0, 0
    while n0 < N:
        steps = min(L_, N - n0)
        if dws is not None:
            rc = lib.cl_propagate_device_ws(ctx, ctypes.byref(sysd), ctypes.byref(chain), E, m, s, d, k,
                                            n0, steps, N, tau, dx.data_ptr(), dW.data_ptr(),
                                            dW2.data_ptr() if dW2 is not None else None, dst.data_ptr(),
                                            dcode.data_ptr(), dh.data_ptr(), dws.data_ptr(), stream)
        else:
            rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), E, m, s, d, k, n0,
                                         steps, N, tau, dx.data_ptr(), dW.data_ptr(),
                                         dW2.data_ptr() if dW2 is not None else None, dst.data_ptr(),
                                         dcode.data_ptr(), dh.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
        n0 += steps
        launches += 1
    torch.cuda.synchronize()
    x = dx.cpu().numpy().T.copy()
    W = dW.cpu().numpy().reshape(k, d, m).transpose(2, 1, 0).copy()
    W2 = dW2.cpu().numpy().reshape(k, d, m).transpose(2, 1, 0).copy() if dW2 is not None else None
    raw = dst[:words].cpu().numpy().T.copy().view(np.uint8).reshape(m, words * 8)
    code = dcode.cpu().numpy()
    hstep = dh.cpu().numpy()
    states = decode_states(raw, allr, d, k)
    ws = dws.cpu().numpy().view(np.uint32).T.copy() if dws is not None else None
    if check:
        check_health(code, hstep, x, W, states, allr)
    prov = dict(backend="gpu", device=device_name(), backend_fallback_reason="", precision=PRECISION_LABELS[elt],
                seed=seed, tau=tau, k=k, chunk_size=m, gpu_chunks=1, gpu_chunk_size=m, launch_steps=L_,
                launches_per_chunk=launches)
    for r, _ in allr:
        prov.update(r.provenance())
    return PropagationResult(x=x, W=W, finals=finals_of(states, allr, d, k, ws), provenance=prov, W2=W2, code=code,
                             hstep=hstep, states=raw if keep_states else None, workspace=ws)


def _propagate_orbit_major(lib, ctx, sys, sysd, chain, allr, x0, N, k, tau, seed, W0, launch_steps, words, check,
                           keep_states, Ns=None, order=None, elt="f64"):
    """The chain system: device buffers in the orbit-major layout (x[i][s], W[i][k][d], states[i][words])."""
    import torch
    E = ELEMENT_TYPES[elt]
    m, s = x0.shape
    d = sys.dimension
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(x0.copy()).to(dev)
    dW = torch.empty((m, k, d), dtype=torch.float64, device=dev)
    if W0 is None:
        rc = lib.cl_initial_deviations_device(ctx, int(seed) & (2 ** 64 - 1), 1, m, d, k, dW.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    else:
        w = np.asarray(W0, dtype=np.float64).reshape(m, d, k)
        dW.copy_(torch.from_numpy(np.ascontiguousarray(w.transpose(0, 2, 1))))
    dst = torch.zeros((m, max(words, 1)), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    launches = 0
    if Ns is not None:
        Na = np.ascontiguousarray(np.asarray(Ns, dtype=np.int64))
        oa = None if order is None else np.ascontiguousarray(np.asarray(order, dtype=np.int64))
        rc = lib.cl_propagate_ragged(ctx, ctypes.byref(sysd), ctypes.byref(chain), E, m, s, d, k,
                                     Na.ctypes.data, None if oa is None else oa.ctypes.data, tau, dx.data_ptr(),
                                     dW.data_ptr(), dst.data_ptr(), dcode.data_ptr(), dh.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
        launches, L_ = 1, int(Na.max()) if m else 0
    else:
        L_ = N if launch_steps is None else min(int(launch_steps), N)
        n0 = 0
        while n0 < N:
            steps = min(L_, N - n0)
            rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), E, m, s, d, k, n0,
                                         steps, N, tau, dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(),
                                         dcode.data_ptr(), dh.data_ptr(), stream)
            if rc:
                raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
            n0 += steps
            launches += 1
    torch.cuda.synchronize()
    x = dx.cpu().numpy()
    W = dW.cpu().numpy().transpose(0, 2, 1).copy()
    raw = dst[:, :words].cpu().numpy().copy().view(np.uint8).reshape(m, words * 8)
    code = dcode.cpu().numpy()
    hstep = dh.cpu().numpy()
    states = decode_states(raw, allr, d, k)
    if check:
        check_health(code, hstep, x, W, states, allr)
    prov = dict(backend="gpu", device=device_name(), backend_fallback_reason="", precision=PRECISION_LABELS[elt],
                seed=seed, tau=tau, k=k, chunk_size=m, gpu_chunks=1, gpu_chunk_size=m, launch_steps=L_,
                launches_per_chunk=launches)
    for r, _ in allr:
        prov.update(r.provenance())
    return PropagationResult(x=x, W=W, finals=finals_of(states, allr, d, k), provenance=prov, code=code,
                             hstep=hstep, states=raw if keep_states else None)


def propagate_ragged(sys, x0, Ns, *, k, tau=1, seed, reducers=(), W0=None, order=None, check=True,
                     keep_states=False, elt="f64"):
    """Chain ensembles with one step count per orbit (Ns[i], each a positive multiple of tau): every orbit is
    the same as a `propagate` call of its own length. `order` (a permutation) sets the launch order; by default
    the longest orbits go first. elt="f32tan": Float32 tangents (as in `propagate`)."""
    tau, k = int(tau), int(k)
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x0.shape
    d = sys.dimension
    Ns = np.asarray(Ns, dtype=np.int64)
    if Ns.shape != (m,) or np.any(Ns < 1) or np.any(Ns % tau):
        raise ValueError("Ns needs one positive multiple of tau per orbit")
    if s != sys.state_length or not (1 <= k <= d):
        raise ValueError("state length or k does not match the system")
    for r in reducers:
        r.check(d, k)
    lib, ctx = _context()
    sysd = sys.descriptor()
    chain, allr = make_chain(reducers)
    nbytes = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), ELEMENT_TYPES[elt], d, k)
    if nbytes < 0:
        raise L.HIPError(-nbytes, "no kernel instance for this system, k and reducer chain")
    if order is None:
        order = np.argsort(-Ns, kind="stable")
    return _propagate_orbit_major(lib, ctx, sys, sysd, chain, allr, x0, 0, k, tau, seed, W0, None, nbytes // 8,
                                  check, keep_states, Ns=Ns, order=order, elt=elt)


def propagate_mixed(lattices, *, K, eps, boundary="ring", k=16, tau=1, seed, reducers=None, chunk=0, check=True,
                    keep_states=False, _noalloc=False, elt="f64"):
    """Several coupled-chain lattices in one persistent launch per chunk (`cl_propagate_mixed`): `lattices` is a list of
    (N, x0, Ns), x0 of shape (m, 2N), Ns the per-orbit step counts (positive multiples of tau). All share K, eps and
    the boundary; the chain is QRSpectrum(HouseholderQR()) with k = 16. `chunk` is the step length of a chunk (0: one
    chunk). Seeded frames use orbit ids 1..m within each lattice, as `propagate_ragged` on that lattice alone, whose
    results this gives bit for bit. Returns one PropagationResult per lattice. elt="f32tan": Float32 tangents."""
    import torch
    from .systems import StandardMapChain
    from .reducers import QRSpectrum, HouseholderQR
    tau, k = int(tau), int(k)
    if reducers is None:
        reducers = (QRSpectrum(HouseholderQR()),)
    lib, ctx = _context()
    chain, allr = make_chain(reducers)
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    systems, bufs, lats, keep = [], [], (L.cl_lattice * max(len(lattices), 1))(), []
    for j, (N, x0, Ns) in enumerate(lattices):
        sys = StandardMapChain(N, K, eps, boundary)
        x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
        m, s = x0.shape
        d = sys.dimension
        Ns = np.ascontiguousarray(np.asarray(Ns, dtype=np.int64))
        if s != sys.state_length or Ns.shape != (m,) or np.any(Ns < 1) or np.any(Ns % tau):
            raise ValueError("each lattice needs x0 of shape (m, 2N) and one positive multiple of tau per orbit")
        sysd = sys.descriptor()
        nbytes = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), ELEMENT_TYPES[elt], d, k)
        if nbytes < 0:
            raise L.HIPError(-nbytes, "no kernel instance for this system, k and reducer chain")
        words = nbytes // 8
        dx = torch.from_numpy(x0.copy()).to(dev)
        dW = torch.empty((m, k, d), dtype=torch.float64, device=dev)
        rc = lib.cl_initial_deviations_device(ctx, int(seed) & (2 ** 64 - 1), 1, m, d, k, dW.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
        dst = torch.zeros((m, max(words, 1)), dtype=torch.int64, device=dev)
        dcode = torch.zeros(m, dtype=torch.int8, device=dev)
        dh = torch.zeros(m, dtype=torch.int64, device=dev)
        lats[j].N, lats[j].m, lats[j].Ns = int(N), m, Ns.ctypes.data
        lats[j].dx, lats[j].dW, lats[j].dstates = dx.data_ptr(), dW.data_ptr(), dst.data_ptr()
        lats[j].dcode, lats[j].dhstep = dcode.data_ptr(), dh.data_ptr()
        systems.append((sys, m, d, words))
        bufs.append((dx, dW, dst, dcode, dh))
        keep.append(Ns)
    sysd = StandardMapChain(max(8, int(lattices[0][0])), K, eps, boundary).descriptor()
    fn = lib.cl_propagate_mixed_noalloc if _noalloc else lib.cl_propagate_mixed
    rc = fn(ctx, ctypes.byref(sysd), ctypes.byref(chain), ELEMENT_TYPES[elt], len(lattices), lats, k, tau, int(chunk),
            stream)
    if rc:
        raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    torch.cuda.synchronize()
    out = []
    for (sys, m, d, words), (dx, dW, dst, dcode, dh), Ns in zip(systems, bufs, keep):
        x = dx.cpu().numpy()
        W = dW.cpu().numpy().transpose(0, 2, 1).copy()
        raw = dst[:, :words].cpu().numpy().copy().view(np.uint8).reshape(m, words * 8)
        code, hstep = dcode.cpu().numpy(), dh.cpu().numpy()
        states = decode_states(raw, allr, d, k)
        if check:
            check_health(code, hstep, x, W, states, allr)
        prov = dict(backend="gpu", device=device_name(), backend_fallback_reason="", precision=PRECISION_LABELS[elt],
                    seed=seed,
                    tau=tau, k=k, chunk_size=m, gpu_chunks=1, gpu_chunk_size=m, launch_steps=int(chunk),
                    launches_per_chunk=1, lattices=len(lattices))
        for r, _ in allr:
            prov.update(r.provenance())
        out.append(PropagationResult(x=x, W=W, finals=finals_of(states, allr, d, k), provenance=prov, code=code,
                                     hstep=hstep, states=raw if keep_states else None))
    return out


def device_name():
    lib = L.load()
    p = L.cl_probe_t()
    lib.cl_probe(L.CL_ELT_F64, ctypes.byref(p))
    return p.device.decode()


def initial_deviations(seed, orbit, d, k):
    """Seeded d x k orthonormal frame of orbit `orbit` (1-based), as the library forms it on the device."""
    lib = L.load()
    out = np.empty((1, k, d), dtype=np.float64)
    rc = lib.cl_initial_deviations(int(seed) & (2 ** 64 - 1), int(orbit), 1, d, k, out.ctypes.data)
    if rc:
        raise L.HIPError(rc, "no frame rule for this d, k")
    return out[0].T.copy()
