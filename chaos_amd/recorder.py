"""The recorder's device path (src/Experiments/run_device.jl, src/Experiments/run_record.jl): grid records of an
ensemble of standard-map orbits.

Each orbit runs with the chain (estimator, RunMonitor) in segments that end at the grid steps; after each segment
the row of the grid step holds the estimator's exponents, the monitor's det and symplecticity maxima, the
orthogonality maximum, the running maxima of the sum and pairing residuals (k = 2), G_max, M_max and the point
(`_dev_row!`, run_device.jl:9-37). The segments are launches of the propagate kernels, so a row equals the final
values of the same estimator run directly to that step, bit for bit; the rows are written on the device by
`cl_record_rows`. An orbit whose segment fails records the 1-based failing row and no further rows.
"""
import ast
import ctypes
import math
from dataclasses import dataclass
from fractions import Fraction

import mpmath as mp
import numpy as np

from . import _lib as L
from .reducers import QRSpectrum, Benettin, CGS, MGS, RGS, HouseholderQR, AugmentedQR, RunMonitor
from .systems import StandardMap, WrappedStandardMap

# method codes with a device path (run_record.jl:47-53, run_device.jl:230); L6 (LAPACK QR) and L9 (SVD) run on the
# reference's CPU path only
QR_METHODS = {"L3": CGS, "L4": MGS, "L4b": RGS, "L5": HouseholderQR, "L5b": AugmentedQR}
VARIATIONAL = ("L1",) + tuple(QR_METHODS)   # the methods with a reducer chain on the tangent frame
METHODS = ("L1", "L2") + tuple(QR_METHODS)


def _to_host(t):
    """A device tensor's values in host memory, copied through a pinned buffer: one DMA transfer rather than the
    runtime's staging kernels for pageable memory (the row buffers reach hundreds of MB at 65536 orbits)."""
    import torch
    h = torch.empty(t.shape, dtype=t.dtype, pin_memory=True)
    h.copy_(t)
    return h.numpy()


def estimator(method, orthogonality_tol=8.0):
    """The estimator of a method code (_estimator, run_record.jl:49-53)."""
    if method == "L1":
        return Benettin()
    if method in QR_METHODS:
        return QRSpectrum(QR_METHODS[method](), orthogonality_tol)
    if method == "L2":
        raise ValueError("L2 (two trajectories) carries a shadow orbit, not a reducer chain")
    raise ValueError(f"method {method!r} has no device path (LAPACK QR L6 and SVD L9 do not)")


def vectors(method):
    """Deviation vectors of a method: 1 for Benettin and the two-trajectory estimator, 2 for the QR spectra."""
    return 1 if method in ("L1", "L2") else 2


def record_grid(N, tau, *, points_per_decade, first_steps, extra_fractions):
    """Steps at which an item of horizon N is recorded (record_grid, src/Experiments/run_grid.jl:18-38): every step
    1..first_steps, round(10^(j / points_per_decade)) evaluated in 256 bits and rounded half up, round(f N) for each
    extra fraction read as its exact decimal, and N; each moved up to the handoff tau ceil(n / tau); sorted, unique."""
    N, tau = int(N), int(tau)
    targets = set(range(1, min(int(first_steps), N) + 1))
    with mp.workprec(256):
        j = 0
        while True:
            v = mp.mpf(10) ** (mp.mpf(j) / int(points_per_decade))
            if v > N:
                break
            targets.add(int(mp.floor(v + mp.mpf(1) / 2)))
            j += 1
    for f in extra_fractions:
        targets.add(math.floor(Fraction(str(f)) * N + Fraction(1, 2)))
    targets.add(N)
    return sorted({tau * -(-t // tau) for t in targets})


def _eval_angle(text):
    """A decimal or closed-form angle text ("pi/4", "0.3 + pi/5", "2pi") in 256-bit arithmetic: numerals read at
    that precision, `pi`, + - * / and parentheses (_eval_exact, run_grid.jl:57-77)."""
    import re
    src = re.sub(r"(\d)pi", r"\1*pi", str(text))

    def ev(node):
        if isinstance(node, ast.Expression):
            return ev(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
            seg = ast.get_source_segment(src, node)
            return mp.mpf(seg)
        if isinstance(node, ast.Name) and node.id == "pi":
            return +mp.pi
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
            v = ev(node.operand)
            return -v if isinstance(node.op, ast.USub) else v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = ev(node.left), ev(node.right)
            return {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b}[type(node.op)]
        raise ValueError(f"unsupported angle text {text!r}")
    return ev(ast.parse(src, mode="eval"))


def deviation_frame(direction, k):
    """The 2 x k start frame of a direction (deviation_frame, run_grid.jl:90-104): first column (cos a, sin a) with
    a evaluated in 256 bits, second column (k = 2) its counterclockwise quarter turn (-sin a, cos a)."""
    if k not in (1, 2):
        raise ValueError(f"deviation_frame: k must be 1 or 2, got {k}")
    with mp.workprec(256):
        a = _eval_angle(direction)
        c, s = float(mp.cos(a)), float(mp.sin(a))
    return np.array([[c], [s]]) if k == 1 else np.array([[c, -s], [s, c]])


@dataclass
class Records:
    """Record array R[orbit, grid row, column] (columns: lambda_1..k, det, symp, orth, sum, pair, G_max, M_max, then
    the point), the 1-based failing row per orbit (0: none) and the final carried data."""
    method: str
    grid: np.ndarray
    R: np.ndarray
    fail: np.ndarray
    code: np.ndarray
    hstep: np.ndarray
    x: np.ndarray
    W: np.ndarray
    states: np.ndarray
    k: int
    s: int

    def unpack(self, j):
        """The record buffer of orbit j (_unpack! and _blank_failed!, run_device.jl:123-139, run_record.jl:251-261):
        rows before the failing row, NaN from it on; the G_max and M_max of the last good row (1.0 when none)."""
        k, s, G = self.k, self.s, len(self.grid)
        f = int(self.fail[j])
        last = G if f == 0 else f - 1
        Rj = self.R[j]
        nan = float("nan")
        out = dict(t=np.zeros(G, np.int64), lyap=np.full((G, k), nan), state=np.full((G, 2), nan),
                   det=np.full(G, nan), symp=np.full(G, nan), orth=np.full(G, nan), sum=np.full(G, nan),
                   pair=np.full(G, nan), gmax=1.0, mmax=1.0, failrow=f,
                   winding=np.zeros((G, 2), np.int64) if s == 4 else None)
        for g in range(last):
            out["t"][g] = self.grid[g]
            out["lyap"][g] = Rj[g, :k]
            out["det"][g], out["symp"][g], out["orth"][g] = Rj[g, k], Rj[g, k + 1], Rj[g, k + 2]
            out["sum"][g], out["pair"][g] = Rj[g, k + 3], Rj[g, k + 4]
            out["state"][g] = Rj[g, k + 7:k + 9]
            if s == 4:
                out["winding"][g] = Rj[g, k + 9:k + 11].astype(np.int64)
        if last >= 1:
            out["gmax"], out["mmax"] = float(Rj[last - 1, k + 5]), float(Rj[last - 1, k + 6])
        return out


def record(sys, x0, N, *, method, tau=1, grid=None, W0=None, direction="pi/4", orthogonality_tol=8.0, eps=None,
           launch_steps=None, timing=None):
    """Record an ensemble on the GPU: `sys` a StandardMap (lifted, x0 of shape (m, 2)) or WrappedStandardMap
    (x0 of shape (m, 4), torus_state form); `grid` the record steps (default: the study grid with 20 points per
    decade, every step to 100 and the fractions 0.1 and 0.5); the start frame `W0` (2 x k, or one per orbit) or the
    `direction` text of deviation_frame. L2 (two trajectories) takes the shadow distance `eps` (a decimal text or a
    float, read as the nearest Float64) and the first column of the frame as its direction. Launches are split at
    `launch_steps` (a multiple of tau) when given. A dict passed as `timing` receives the device time of the
    launches (HIP events on the launch stream) in "device_ms"."""
    import torch
    from .tangent import make_chain, _context
    if not isinstance(sys, (StandardMap, WrappedStandardMap)):
        raise ValueError("the recorder's device path runs the standard map, lifted or wrapped")
    N, tau = int(N), int(tau)
    k = vectors(method)
    if grid is None:
        grid = record_grid(N, tau, points_per_decade=20, first_steps=100, extra_fractions=(0.1, 0.5))
    grid = np.asarray(grid, dtype=np.int64)
    G = len(grid)
    if G == 0 or grid[-1] != N or np.any(grid % tau) or np.any(np.diff(grid) <= 0) or grid[0] <= 0:
        raise ValueError("grid must be increasing positive multiples of tau ending at N")
    if launch_steps is not None and (launch_steps < 1 or launch_steps % tau):
        raise ValueError("launch_steps must be a positive multiple of tau")
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x0.shape
    if s != sys.state_length:
        raise ValueError(f"state length {s} does not match {sys!r}")
    if W0 is None:
        W0 = deviation_frame(direction, k)
    W0 = np.asarray(W0, dtype=np.float64)
    if W0.shape == (2, k):
        W0 = np.broadcast_to(W0, (m, 2, k))
    if W0.shape != (m, 2, k):
        raise ValueError("W0 must be 2 x k or one 2 x k frame per orbit")
    if method == "L2":
        if eps is None:
            raise ValueError("L2 needs the shadow distance eps")
        return _record_shadow(sys, x0, N, tau, grid, np.ascontiguousarray(W0[:, :, 0]), float(str(eps)),
                              launch_steps, timing)
    chain, allr = make_chain((estimator(method, orthogonality_tol), RunMonitor(sys)))
    lib, ctx = _context()
    sysd = sys.descriptor()
    nbytes = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, 2, k)
    if nbytes < 0:
        raise L.HIPError(-nbytes, "no kernel instance for this method")
    words = nbytes // 8
    ncol = k + 7 + s
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    dW = torch.from_numpy(np.ascontiguousarray(W0.transpose(2, 1, 0).reshape(2 * k, m))).to(dev)
    dst = torch.zeros((words, m), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    dfail = torch.zeros(m, dtype=torch.int64, device=dev)
    dR = torch.zeros((m, G, ncol), dtype=torch.float64, device=dev)
    if timing is not None:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
    n = 0
    for g in range(G):
        target = int(grid[g])
        while n < target:
            steps = target - n if launch_steps is None else min(int(launch_steps), target - n)
            rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, m, s, 2, k, n,
                                         steps, N, tau, dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(),
                                         dcode.data_ptr(), dh.data_ptr(), stream)
            if rc:
                raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
            n += steps
        rc = lib.cl_record_rows(ctx, ctypes.byref(sysd), ctypes.byref(chain), m, s, k, g, G, dx.data_ptr(),
                                dst.data_ptr(), dcode.data_ptr(), dfail.data_ptr(), dR.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    if timing is not None:
        ev1.record()
    torch.cuda.synchronize()
    if timing is not None:
        timing["device_ms"] = ev0.elapsed_time(ev1)
    raw = dst.cpu().numpy().T.copy().view(np.uint8).reshape(m, words * 8)
    return Records(method=method, grid=grid, R=_to_host(dR), fail=dfail.cpu().numpy(), code=dcode.cpu().numpy(),
                   hstep=dh.cpu().numpy(), x=dx.cpu().numpy().T.copy(),
                   W=dW.cpu().numpy().reshape(k, 2, m).transpose(2, 1, 0).copy(), states=raw, k=k, s=s)


def _record_shadow(sys, x0, N, tau, grid, u0, d0, launch_steps, timing):
    """L2 (_shadow_batch!, run_device.jl:175-197; _shadow_record_kernel!, :74-98): the end records carried across
    segments ending at the grid steps (`cl_shadow_device`), the row of each grid step written on the device
    (`cl_shadow_record_rows`): total(sum) / step in column 0 and the point in columns 8..; no start check, a start
    out of range fails at row 1."""
    import torch
    from .tangent import _context
    from .two_trajectory import check_d0, decode_ends
    d0 = check_d0(d0)
    m, s = x0.shape
    G = len(grid)
    lib, ctx = _context()
    sysd = sys.descriptor()
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    du = torch.from_numpy(np.ascontiguousarray(u0.T)).to(dev)
    de = torch.zeros((2 * s + 6, m), dtype=torch.int64, device=dev)
    dfail = torch.zeros(m, dtype=torch.int64, device=dev)
    dR = torch.zeros((m, G, 1 + 7 + s), dtype=torch.float64, device=dev)
    if timing is not None:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
    n = 0
    for g in range(G):
        target = int(grid[g])
        while n < target:
            steps = target - n if launch_steps is None else min(int(launch_steps), target - n)
            rc = lib.cl_shadow_device(ctx, ctypes.byref(sysd), m, s, n, steps, N, tau, d0, dx.data_ptr(),
                                      du.data_ptr(), de.data_ptr(), stream)
            if rc:
                raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
            n += steps
        rc = lib.cl_shadow_record_rows(ctx, m, s, g, G, target, de.data_ptr(), dfail.data_ptr(), dR.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    if timing is not None:
        ev1.record()
    torch.cuda.synchronize()
    if timing is not None:
        timing["device_ms"] = ev0.elapsed_time(ev1)
    raw = de.cpu().numpy().T.copy().view(np.uint64)
    e = decode_ends(raw, s)
    return Records(method="L2", grid=grid, R=_to_host(dR), fail=dfail.cpu().numpy(), code=e["code"],
                   hstep=e["step"], x=e["x"], W=e["y"], states=raw, k=1, s=s)


# ---------------------------------------------------------------- the indicator recorder (device kinds)
# Columns of E24 per member (src/Experiments/run_indicator_chain.jl:16-17); RLI belongs to the pair kind.
IND_MEMBER_COLUMNS = {"SALI": ("SALI.sali", "SALI.L"), "FLI": ("FLI.fli", "FLI.fli1"), "RLI": ("RLI.DL", "RLI.RLI"),
                      "MEGNO": ("MEGNO.Y1", "MEGNO.Ybar1", "MEGNO.Ybar20_1", "MEGNO.Y2", "MEGNO.Ybar2",
                                "MEGNO.Ybar20_2")}


def ind_members(names, pair=False):
    """The members of a tangent (or, with `pair`, a pair) chain in order (chain_members, run_indicator_chain.jl:27-31):
    every member but the last non-replacing, the last replacing; SALI has one form (it always replaces); RLI runs on
    the pair kind only."""
    from .reducers import MEGNO, FLI, SALI, RLI
    names = tuple(names)
    if not 1 <= len(names) <= 4:
        raise ValueError("a tangent chain has 1 to 4 members")
    out = []
    for i, nm in enumerate(names):
        last = i == len(names) - 1
        if nm == "SALI":
            out.append(SALI())
        elif nm == "FLI":
            out.append(FLI(replace=last))
        elif nm == "MEGNO":
            out.append(MEGNO(replace=last))
        elif nm == "RLI" and pair:
            out.append(RLI(replace=last))
        else:
            raise ValueError(f"the device tangent kind runs FLI, MEGNO and SALI (and RLI on a pair), not {nm!r}")
    return tuple(out)


@dataclass
class IndRecords:
    """Record array R[orbit, grid row, column] of an indicator item (columns: the members' series columns, det_max,
    symp_max, M_max, the point), the 1-based failing row per orbit (0: none) and the final carried data; `reducers`
    is the chain, for the final tables of a window kind's reader (finals_of)."""
    kind: str
    columns: tuple
    grid: np.ndarray
    R: np.ndarray
    fail: np.ndarray
    code: np.ndarray
    hstep: np.ndarray
    x: np.ndarray
    W: np.ndarray
    states: np.ndarray
    reducers: tuple
    s: int

    def unpack(self, j):
        """The buffer of orbit j (_ind_unpack! and _blank_ind!, run_indicator_device.jl:111-125,
        run_indicator_record.jl:134-143): rows before the failing row, NaN from it on; M_max of the last good row
        (1.0 when none); t is the grid step of each recorded row and 0 after."""
        nser, G = len(self.columns), len(self.grid)
        f = int(self.fail[j])
        last = G if f == 0 else f - 1
        Rj = self.R[j]
        nan = float("nan")
        out = dict(t=np.zeros(G, np.int64), series=np.full((G, nser), nan), det=np.full(G, nan),
                   symp=np.full(G, nan), state=np.full((G, 2), nan), mmax=1.0, failrow=f,
                   winding=np.zeros((G, 2), np.int64) if self.s == 4 else None)
        for g in range(last):
            out["t"][g] = self.grid[g]
            out["series"][g] = Rj[g, :nser]
            out["det"][g], out["symp"][g] = Rj[g, nser], Rj[g, nser + 1]
            out["state"][g] = Rj[g, nser + 3:nser + 5]
            if self.s == 4:
                out["winding"][g] = Rj[g, nser + 5:nser + 7].astype(np.int64)
        if last >= 1:
            out["mmax"] = float(Rj[last - 1, nser + 2])
        return out


def record_indicators(sys, x0, N, *, members=("FLI", "MEGNO", "SALI"), reader=None, tau=1, grid=None, W0=None,
                      direction="pi/4", launch_steps=None, timing=None):
    """Record an indicator item on the GPU (ind_device_records, run_indicator_device.jl:164-197).

    The tangent kind (`reader` None): the chain is `members` (ind_members) then RunMonitor, on the frame `W0` (2 x 2,
    or one per orbit; default deviation_frame(direction, 2)); each row holds the members' series columns. The pair
    kind: `sys` a PairMap, x0 the pair states (pair_start), members with RLI (the study's spec RLI, MEGNO), W0 the
    block frame (4 x 2; default pair_frame of the direction); the point columns hold orbit A's state. A window
    kind (`reader` a WeightedBirkhoff, kind "birkhoff", or a WeightedPhaseAdvance, kind "phase_advance",
    run_indicator.jl:5-6): the chain is RunMonitor then the reader, stepped with the constant deviation column (1, 0)
    reset at every handoff; its rows hold the monitor and the point, and its tables come from the reader's final
    state (finals_of on `states`). `sys` is a StandardMap (x0 of shape (m, 2)) or a
    WrappedStandardMap (x0 of shape (m, 4)); `grid` the record steps (default: the study grid). Launches end at the
    grid steps and are split at `launch_steps` (a multiple of tau) when given."""
    import torch
    from .tangent import make_chain, _context
    from .systems import PairMap
    pair = isinstance(sys, PairMap)
    if not isinstance(sys, (StandardMap, WrappedStandardMap, PairMap)):
        raise ValueError("the indicator recorder's device path runs the standard map or its pair, lifted or wrapped")
    if pair and reader is not None:
        raise ValueError("a window kind runs on the map itself, not a pair")
    N, tau = int(N), int(tau)
    window = reader is not None
    if window:
        from .birkhoff import WeightedBirkhoff, WeightedPhaseAdvance
        if isinstance(reader, WeightedPhaseAdvance):
            window_kind = "phase_advance"
        elif isinstance(reader, WeightedBirkhoff):
            window_kind = "birkhoff"
        else:
            raise ValueError(f"a window kind's reader is a WeightedBirkhoff or a WeightedPhaseAdvance, got {reader!r}")
    k = 1 if window else 2
    if grid is None:
        grid = record_grid(N, tau, points_per_decade=20, first_steps=100, extra_fractions=(0.1, 0.5))
    grid = np.asarray(grid, dtype=np.int64)
    G = len(grid)
    if G == 0 or grid[-1] != N or np.any(grid % tau) or np.any(np.diff(grid) <= 0) or grid[0] <= 0:
        raise ValueError("grid must be increasing positive multiples of tau ending at N")
    if launch_steps is not None and (launch_steps < 1 or launch_steps % tau):
        raise ValueError("launch_steps must be a positive multiple of tau")
    x0 = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x0.shape
    if s != sys.state_length:
        raise ValueError(f"state length {s} does not match {sys!r}")
    if window:
        reds = (RunMonitor(sys), reader)
        reader.check_run(tau, N)
        W0 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
        columns = ()
    else:
        reds = ind_members(members, pair=pair) + (RunMonitor(sys),)
        columns = tuple(c for nm in members for c in IND_MEMBER_COLUMNS[nm])
        d = sys.dimension
        if W0 is None:
            from .systems import pair_frame
            W0 = pair_frame(deviation_frame(direction, 1)[:, 0]) if pair else deviation_frame(direction, 2)
        W0 = np.asarray(W0, dtype=np.float64)
        if W0.shape == (d, 2):
            W0 = np.broadcast_to(W0, (m, d, 2))
        if W0.shape != (m, d, 2):
            raise ValueError(f"W0 must be {d} x 2 or one {d} x 2 frame per orbit")
    d = W0.shape[1]
    ns = s // 2 if pair else s   # the point columns: orbit A's state on a pair
    chain, allr = make_chain(reds, reset_frame=window)
    lib, ctx = _context()
    sysd = sys.descriptor()
    nbytes = lib.cl_state_bytes(ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, d, k)
    if nbytes < 0:
        raise L.HIPError(-nbytes, "no kernel instance for this chain")
    words = nbytes // 8
    ncol = lib.cl_ind_row_width(ctypes.byref(chain), ns)
    assert ncol == len(columns) + 3 + ns
    dev = torch.device("cuda")
    stream = torch.cuda.current_stream().cuda_stream
    dx = torch.from_numpy(np.ascontiguousarray(x0.T)).to(dev)
    dW = torch.from_numpy(np.ascontiguousarray(W0.transpose(2, 1, 0).reshape(d * k, m))).to(dev)
    dst = torch.zeros((words, m), dtype=torch.int64, device=dev)
    dcode = torch.zeros(m, dtype=torch.int8, device=dev)
    dh = torch.zeros(m, dtype=torch.int64, device=dev)
    dfail = torch.zeros(m, dtype=torch.int64, device=dev)
    dR = torch.zeros((m, G, ncol), dtype=torch.float64, device=dev)
    if timing is not None:
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ev0.record()
    n = 0
    for g in range(G):
        target = int(grid[g])
        while n < target:
            steps = target - n if launch_steps is None else min(int(launch_steps), target - n)
            rc = lib.cl_propagate_device(ctx, ctypes.byref(sysd), ctypes.byref(chain), L.CL_ELT_F64, m, s, d, k, n,
                                         steps, N, tau, dx.data_ptr(), dW.data_ptr(), None, dst.data_ptr(),
                                         dcode.data_ptr(), dh.data_ptr(), stream)
            if rc:
                raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
            n += steps
        rc = lib.cl_ind_record_rows(ctx, ctypes.byref(sysd), ctypes.byref(chain), m, s, g, G, dx.data_ptr(),
                                    dst.data_ptr(), dcode.data_ptr(), dfail.data_ptr(), dR.data_ptr(), stream)
        if rc:
            raise L.HIPError(rc, lib.cl_last_error(ctx).decode())
    if timing is not None:
        ev1.record()
    torch.cuda.synchronize()
    if timing is not None:
        timing["device_ms"] = ev0.elapsed_time(ev1)
    raw = dst.cpu().numpy().T.copy().view(np.uint8).reshape(m, words * 8)
    return IndRecords(kind=window_kind if window else ("pair" if pair else "tangent"), columns=columns, grid=grid,
                      R=_to_host(dR), fail=dfail.cpu().numpy(), code=dcode.cpu().numpy(), hstep=dh.cpu().numpy(),
                      x=dx.cpu().numpy().T.copy(), W=dW.cpu().numpy().reshape(k, d, m).transpose(2, 1, 0).copy(),
                      states=raw, reducers=allr, s=ns)
