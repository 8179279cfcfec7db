"""`propagate_shadow`: the two-trajectory estimate X2 of the largest exponent (src/Lyapunov/two_trajectory.jl:190-205)
on the MI355X library.

A shadow orbit starts at distance d0 from the reference orbit along the seeded direction (the first column of the
seeded frame, or `U0` normalised) and is placed back at distance d0 along the current separation every tau steps;
X2 = (sum of log(d / d_prev)) / N. The argument meaning, validation and errors follow the reference: N a positive
multiple of tau, d0 finite and positive, every start's realized separation within [d0/2, 2 d0] (checked before the
run, `_check_starts`, :152-157), and a NumericalHealthError ("state" or "separation", `_shadow_error`, :87-92) for the
lowest failing orbit. The work runs on the GPU through the C ABI (`cl_shadow_device`); there is no CPU path.

Results: `x` the final reference states, `W` the final shadow states, `finals[0]` a dict of per-orbit arrays n, sum,
comp, X2, d_prev (`_final`, :85), and the raw end records (ShadowEnd{s, Float64}, 2s + 6 words each) in `states`.
"""
import ctypes
import math

import numpy as np

from . import _lib as L
from .systems import StandardMap, WrappedStandardMap
from .tangent import NumericalHealthError, PropagationResult

def check_d0(d0):
    """_check_d0 (two_trajectory.jl:22-28): a Python int or a float that is exactly a Float64, finite and positive."""
    if isinstance(d0, bool) or not isinstance(d0, (int, float, np.floating, np.integer)):
        raise ValueError(f"d0 of type {type(d0).__name__} would be narrowed to Float64")
    if isinstance(d0, np.floating) and np.finfo(type(d0)).bits > 64:
        raise ValueError(f"d0 of type {type(d0).__name__} would be narrowed to Float64")
    d = float(d0)
    if not (math.isfinite(d) and d > 0):
        raise ValueError(f"d0 must be finite and positive, got {d0}")
    return d


def unit_directions(U0, m):
    """U0 normalised per orbit, u / safe_norm(u) (_shadow_directions, two_trajectory.jl:159-163; safe_norm,
    src/Foundation/norm.jl:8-12: scale by the largest magnitude, NaN propagates)."""
    U = np.asarray(U0, dtype=np.float64)
    if U.shape != (m, 2):
        raise ValueError("U0 needs one vector per orbit")
    out = np.empty_like(U)
    for i in range(m):
        a, b = U[i]
        mx = max(abs(a), abs(b)) if not (math.isnan(a) or math.isnan(b)) else float("nan")
        if not (math.isfinite(mx) and mx > 0):
            nrm = mx
        else:
            nrm = mx * math.sqrt((a / mx) * (a / mx) + (b / mx) * (b / mx))
        with np.errstate(invalid="ignore", divide="ignore"):
            out[i] = (a / nrm, b / nrm)
    return out


def seeded_directions(seed, m):
    """First column of the seeded frame of orbits 1..m (initial_deviations(seed, j, Val(2), Val(1), T)[:, 1]), formed
    on the host by the library's frame rule (the device run forms the same frames with its frame kernel)."""
    lib = L.load()
    out = np.empty((m, 1, 2), dtype=np.float64)
    rc = lib.cl_initial_deviations(int(seed) & (2 ** 64 - 1), 1, m, 2, 1, out.ctypes.data)
    if rc:
        raise L.HIPError(rc, "no frame rule for d = 2, k = 1")
    return out[:, 0, :].copy()


def decode_ends(raw, s):
    """End records (m, 2s + 6) uint64 -> dict of x, y, sum, comp, dprev, code, step, value."""
    f = raw.view(np.float64)
    return dict(x=f[:, :s].copy(), y=f[:, s:2 * s].copy(), sum=f[:, 2 * s].copy(), comp=f[:, 2 * s + 1].copy(),
                dprev=f[:, 2 * s + 2].copy(), code=(raw[:, 2 * s + 3] & 0xff).astype(np.uint8).view(np.int8),
                step=raw[:, 2 * s + 4].view(np.int64).copy(), value=f[:, 2 * s + 5].copy())


def shadow_error(code, value, step, d0, orbit, backend="gpu"):
    """_shadow_error (two_trajectory.jl:87-92)."""
    v = float(value)
    if code == 1:
        return NumericalHealthError("state", v, math.inf, int(step), orbit, backend)
    thr = float(d0) if code == 3 else (0.0 if math.isfinite(v) else math.inf)
    return NumericalHealthError("separation", v, thr, int(step), orbit, backend)


def _validate(sys, x0, N, tau):
    if not isinstance(sys, (StandardMap, WrappedStandardMap)):
        raise ValueError("the two-trajectory kernel runs the standard map, lifted or wrapped")
    N, tau = int(N), int(tau)
    if tau < 1 or N < 1 or N % tau:
        raise ValueError(f"N must be a positive multiple of tau, got N = {N}, tau = {tau}")
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    if x0.ndim != 2:
        raise ValueError("x0 must have shape (orbits, state length)")
    if x0.shape[1] < 2:
        raise ValueError(f"state length {x0.shape[1]} is below the system dimension 2")
    if x0.shape[1] != sys.state_length:
        raise ValueError(f"state length {x0.shape[1]} does not match {sys!r}")
    return x0, N, tau


def propagate_shadow(sys, x0, N, *, d0, tau=1, seed, U0=None, launch_steps=None, check=True):
    """Two-trajectory estimate for every orbit of x0 (m, s): StandardMap (s = 2) or WrappedStandardMap (s = 4,
    torus_state form). `launch_steps` (a multiple of tau) splits the run into launches that carry the end records
    (Amendment D44 of the reference, two_trajectory.jl:124-143); the result is the same bit for bit. With
    check=False failing orbits are left in `code`/`hstep` instead of raising."""
    import torch
    from .tangent import _context
    x0, N, tau = _validate(sys, x0, N, tau)
    d0 = check_d0(d0)
    if launch_steps is not None and (int(launch_steps) < 1 or int(launch_steps) % tau):
        raise ValueError(f"forced steps per launch must be a positive multiple of tau, got L = {launch_steps}, tau = {tau}")
    m, s = x0.shape
    u = None if U0 is None else unit_directions(U0, m)
    lib, ctx = _context()
    sysd = sys.descriptor()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    if u is None:   # the seeded direction: the frame kernel's k = 1 column, [2][m]
        du = torch.empty((2, m), dtype=torch.float64, device=dev)
        rc = lib.cl_initial_deviations_device(ctx, int(seed) & (2 ** 64 - 1), 1, m, 2, 1, du.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    else:
        du = torch.from_numpy(np.ascontiguousarray(u.T)).to(dev)
    ew = 2 * s + 6
    de = torch.zeros((ew, m), dtype=torch.int64, device=dev)

    def launch(n0, steps):
        rc = lib.cl_shadow_device(ctx, ctypes.byref(sysd), m, s, n0, steps, N, tau, d0, dx.data_ptr(), du.data_ptr(),
                                  de.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())

    # _check_starts: every start's realized separation, before any step (a start-only launch: steps = 0)
    launch(0, 0)
    if check:
        codes = (de[2 * s + 3] & 0xff).cpu().numpy()
        bad = np.nonzero(codes)[0]
        if len(bad):
            i = int(bad[0])
            value = de[2 * s + 5, i].cpu().numpy().view(np.float64)
            raise shadow_error(int(codes[i]), float(value), 0, d0, i + 1)
    L_ = N if launch_steps is None else int(launch_steps)
    n0, launches = 0, 0
    while n0 < N:
        steps = min(L_, N - n0)
        launch(n0, steps)
        n0 += steps
        launches += 1
    torch.cuda.synchronize()
    raw = de.cpu().numpy().T.copy().view(np.uint64)
    e = decode_ends(raw, s)
    if check:
        bad = np.nonzero(e["code"] != 0)[0]
        if len(bad):
            i = int(bad[0])
            raise shadow_error(e["code"][i], e["value"][i], e["step"][i], d0, i + 1)
    total = e["sum"] + e["comp"]
    finals = [dict(n=np.full(m, N, dtype=np.int64), sum=e["sum"], comp=e["comp"], X2=total / N, d_prev=e["dprev"])]
    prov = dict(estimator="two-trajectory", state_mode="wrapped" if s == 4 else "lifted", k=0, d0=d0, tau=tau,
                seed=int(seed), precision="Float64", backend="gpu", launch_steps=L_, launches_per_chunk=launches,
                gpu_chunks=1, gpu_chunk_size=m)
    return PropagationResult(x=e["x"], W=e["y"], finals=finals, provenance=prov, code=e["code"], hstep=e["step"],
                             states=raw)
