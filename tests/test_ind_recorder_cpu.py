"""The indicator recorder on the CPU: the oracle's restatement (or_ind_record) of _ind_series_loop! and
_ind_window_loop! (src/Experiments/run_indicator_record.jl:99-132) pinned against the host finals of the members
(reducers.py: fli.jl:91-94, megno.jl:97-111, sali.jl:62-66) on oracle states run directly to each grid step; the
row layout (run_indicator_chain.jl:16-17), the chain's replace forms (:27-31), unpacking and blanking
(run_indicator_device.jl:111-125, run_indicator_record.jl:134-143)."""
import ctypes
import math

import numpy as np
import pytest

from chaos_amd import (StandardMap, WrappedStandardMap, RunMonitor, WeightedBirkhoff, CosObservable, SinObservable,
                       torus_state)
from chaos_amd import _lib as L
from chaos_amd.recorder import IND_MEMBER_COLUMNS, IndRecords, ind_members, record_grid
from chaos_amd.reducers import MEGNO
from chaos_amd.tangent import make_chain, decode_states, finals_of
from tests import oracle_lib as O


def starts(m, wrapped, seed):
    rng = np.random.default_rng(seed)
    z = 2 * math.pi * rng.random((m, 2))
    z[1] = [np.nan, 0.5]               # fails in the first segment
    if wrapped:
        return np.array([torus_state(v) if np.all(np.isfinite(v)) else [v[0], v[1], 0.0, 0.0] for v in z])
    z[2] = [0.5, 1e306]                # x grows by about 1e306 a step: overflows after some 180 steps
    return z


def expected_row(members, reds, states, x, s):
    """The row from the host finals of the states (the members' series columns, det, symp, M_max, the point)."""
    fin = finals_of(states, reds, 2, 2)
    cols = []
    for nm, f in zip(members, fin):
        if nm == "FLI":
            cols += [f["fli"], f["fli1"]]
        elif nm == "MEGNO":
            for c in range(2):
                cols += [f["Y"][:, c], f["Ybar"][:, c], f["Ybar20"][:, c]]
        else:
            cols += [f["sali"], f["L"]]
    mon = fin[len(members)]
    cols += [mon["det_max"], mon["symp_max"], mon["M_max"]]
    cols += [x[:, j] for j in range(s)]
    return np.stack(cols, axis=1)


@pytest.mark.parametrize("members", [("FLI", "MEGNO", "SALI"), ("MEGNO", "FLI", "SALI"), ("MEGNO",), ("FLI",),
                                     ("SALI",)])
@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_oracle_series_rows_equal_finals(members, wrapped):
    m, N, tau = 64, 240, 2
    sysm = (WrappedStandardMap if wrapped else StandardMap)(6.0)
    x0 = starts(m, wrapped, 3)
    s = x0.shape[1]
    reds = ind_members(members) + (RunMonitor(sysm),)
    chain, allr = make_chain(reds)
    grid = record_grid(N, tau, points_per_decade=6, first_steps=10, extra_fractions=(0.5,))
    ncol = sum(len(IND_MEMBER_COLUMNS[n]) for n in members) + 3 + s
    W0 = O.frames(4, 0, m, 2, 2)
    rec = O.ind_record(sysm.descriptor(), chain, x0, W0, N, tau, 2, grid, ncol)
    full = O.run(sysm.descriptor(), chain, x0, W0, N, tau, 2)
    assert np.array_equal(rec["states"], full["states"]) and np.array_equal(rec["code"], full["code"])
    assert rec["fail"][1] == 1 and (wrapped or 1 < rec["fail"][2] < len(grid))
    for g, t in enumerate(grid):
        r = O.run(sysm.descriptor(), chain, x0, W0, int(t), tau, 2)
        ok = r["code"] == 0
        exp = expected_row(members, allr, decode_states(r["states"], allr, 2, 2), r["x"], s)
        assert np.array_equal(rec["R"][ok, g].view(np.uint64), exp[ok].view(np.uint64)), (members, g)
        # an orbit that has failed by this grid step has its failing row at or before it
        assert np.all((rec["fail"][~ok] >= 1) & (rec["fail"][~ok] <= g + 1))


def test_oracle_window_rows_and_final_states():
    """The window kind (RunMonitor, WeightedBirkhoff) with the constant deviation column: the rows hold the monitor
    and the point of a RunMonitor-only run to each grid step; the final states are the full chain's."""
    m, N, tau = 48, 400, 1
    sysm = WrappedStandardMap(2 * math.pi)
    rng = np.random.default_rng(8)
    x0 = np.array([torus_state(v) for v in 2 * math.pi * rng.random((m, 2))])
    wb = WeightedBirkhoff(window=(50, 200), observables=(CosObservable(1), SinObservable(2)), birkhoff_order=2)
    chain, allr = make_chain((RunMonitor(sysm), wb), reset_frame=True)
    mchain, _ = make_chain((RunMonitor(sysm),), reset_frame=True)
    grid = record_grid(N, tau, points_per_decade=5, first_steps=5, extra_fractions=(0.5,))
    W1 = np.broadcast_to(np.array([[1.0], [0.0]]), (m, 2, 1))
    rec = O.ind_record(sysm.descriptor(), chain, x0, W1, N, tau, 1, grid, 3 + 4)
    full = O.run(sysm.descriptor(), chain, x0, W1, N, tau, 1)
    assert np.array_equal(rec["states"], full["states"]) and np.all(rec["fail"] == 0)
    for g, t in enumerate(grid):
        r = O.run(sysm.descriptor(), mchain, x0, W1, int(t), tau, 1)
        mon = decode_states(r["states"], [(RunMonitor(sysm), 0)], 2, 1)[0]
        exp = np.column_stack([mon["det_max"], mon["symp_max"], mon["M_max"], r["x"]])
        assert np.array_equal(rec["R"][:, g], exp)


def test_chain_forms_and_row_width():
    reds = ind_members(("FLI", "MEGNO", "SALI"))
    assert [type(r).__name__ for r in reds] == ["FLI", "MEGNO", "SALI"]
    assert reds[0].replace is False and reds[1].replace is False
    assert ind_members(("SALI", "MEGNO"))[1].replace is True and ind_members(("MEGNO", "FLI"))[1].replace is True
    with pytest.raises(ValueError):
        ind_members(("RLI", "MEGNO"))
    with pytest.raises(ValueError):
        ind_members(())
    lib = L.load()
    sysm = StandardMap(1.0)
    for members, s in ((("FLI", "MEGNO", "SALI"), 2), (("FLI", "MEGNO", "SALI"), 4), (("MEGNO",), 2)):
        chain, _ = make_chain(ind_members(members) + (RunMonitor(sysm),))
        assert lib.cl_ind_row_width(ctypes.byref(chain), s) == sum(len(IND_MEMBER_COLUMNS[n]) for n in members) + 3 + s
    wb = WeightedBirkhoff(window=(16,), observables=(CosObservable(1),))
    chain, _ = make_chain((RunMonitor(sysm), wb), reset_frame=True)
    assert lib.cl_ind_row_width(ctypes.byref(chain), 4) == 7


def test_unpack_blanks_from_the_failing_row():
    G, nser, s = 5, 2, 4
    R = np.arange(3 * G * (nser + 3 + s), dtype=np.float64).reshape(3, G, nser + 3 + s)
    R[..., nser + 5:] = 7.0                      # winding columns
    rec = IndRecords(kind="tangent", columns=("SALI.sali", "SALI.L"), grid=np.array([1, 2, 4, 8, 10]), R=R,
                     fail=np.array([0, 3, 1]), code=np.zeros(3, np.int8), hstep=np.zeros(3, np.int64),
                     x=np.zeros((3, s)), W=np.zeros((3, 2, 2)), states=np.zeros((3, 8), np.uint8), reducers=(), s=s)
    u = rec.unpack(1)
    assert u["failrow"] == 3 and list(u["t"]) == [1, 2, 0, 0, 0]
    assert np.array_equal(u["series"][:2], R[1, :2, :2]) and np.all(np.isnan(u["series"][2:]))
    assert np.all(np.isnan(u["det"][2:])) and np.all(np.isnan(u["state"][2:]))
    assert u["mmax"] == R[1, 1, nser + 2] and np.all(u["winding"][:2] == 7)
    u0 = rec.unpack(2)
    assert u0["mmax"] == 1.0 and np.all(np.isnan(u0["series"]))
    assert np.array_equal(rec.unpack(0)["series"], R[0, :, :2])


def test_megno_ybar20_uses_the_literal_cube():
    """megno.jl:101 divides by tn^3, Julia's literal power (tn tn) tn. While tn^2 is exact (n < 2^26.5) that product
    is the correctly rounded cube; beyond, pow(tn, 3) differs from it in the last bit for many n. The host final
    takes the literal power."""
    ns = np.arange(10 ** 8, 10 ** 8 + 400, dtype=np.int64)
    tn = ns.astype(np.float64)
    lit = (tn * tn) * tn
    differ = ns[np.power(tn, 3) != lit]
    assert differ.size > 0
    n = int(differ[0])
    r = MEGNO()
    st = np.zeros(1, dtype=r.dtype(2, 2))
    st["n"], st["K"] = n, n
    st["c"][0, :, 4, 0] = 1.0                    # Q2 of both columns = 1
    f = r.final(st)
    t = float(n)
    assert f["Ybar20"][0, 0] == 1.0 * 1.0 / ((t * t) * t)


# ---------------------------------------------------------------- the pair kind (PairMap + RLI)
def pair_starts(base, m, seed, dx=1e-9):
    from chaos_amd import pair_start
    rng = np.random.default_rng(seed)
    u = np.array([math.cos(0.7), math.sin(0.7)])
    z = [pair_start(base, 2 * math.pi * rng.random(2), u, dx)[0] for _ in range(m)]
    return np.array(z), u


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_oracle_rli_sums_are_the_single_orbit_benettin_sums(wrapped):
    """rli.jl:14-18: RLI's log sums take Benettin's operations, so s_A and s_B are bitwise the single-orbit sums of
    orbits A and B with the start direction xi_0; DL and RLI follow from them (rli.jl:74, 116-119)."""
    from chaos_amd import PairMap, RLI, Benettin, MEGNO, pair_frame
    m, N, tau = 40, 300, 3
    base = (WrappedStandardMap if wrapped else StandardMap)(1.3)
    z, u = pair_starts(base, m, 12)
    s = base.state_length
    psys = PairMap(base)
    W = np.broadcast_to(pair_frame(u), (m, 4, 2))
    for reds in ((RLI(),), (RLI(replace=False), MEGNO())):
        chain, allr = make_chain(reds)
        r = O.run(psys.descriptor(), chain, z, W, N, tau, 2)
        assert np.all(r["code"] == 0)
        st = decode_states(r["states"], allr, 4, 2)[0]
        bchain, ball = make_chain((Benettin(),))
        u1 = np.broadcast_to(u.reshape(2, 1), (m, 2, 1))
        for half, key in ((slice(0, s), "sA"), (slice(s, 2 * s), "sB")):
            b = O.run(base.descriptor(), bchain, z[:, half], u1, N, tau, 1)
            bst = decode_states(b["states"], ball, 2, 1)[0]
            assert np.array_equal(st[key][:, 0], bst["sum"][:, 0]) and np.array_equal(st[key][:, 1], bst["sum"][:, 1])
            assert np.array_equal(r["x"][:, half], b["x"])
        f = RLI().final(st)
        assert np.all(f["handoffs"] == N // tau) and np.all(f["DL"] >= 0) and np.all(np.isfinite(f["RLI"]))
        # identical halves: DL and RLI exactly 0
    zz = np.concatenate([z[:, :s], z[:, :s]], axis=1)
    chain, allr = make_chain((RLI(),))
    r = O.run(psys.descriptor(), chain, zz, W, N, tau, 2)
    f = RLI().final(decode_states(r["states"], allr, 4, 2)[0])
    assert np.all(f["DL"] == 0.0) and np.all(f["RLI"] == 0.0)


def test_oracle_rli_pairing_and_norm_failures():
    """A frame with a nonzero off-block entry fails the first handoff as "pairing" (code 2, the column, the largest
    off-block magnitude); a zero column as "deviation matrix" (code 1) (rli.jl:84-95)."""
    from chaos_amd import PairMap, RLI, pair_frame
    base = StandardMap(1.0)
    z, u = pair_starts(base, 3, 4)
    W = np.broadcast_to(pair_frame(u), (3, 4, 2)).copy()
    W[0, 2, 0] = 1e-300          # block B row of column 1
    W[1, 1, 1] = -2.0            # block A row of column 2
    W[2, :, 1] = 0.0             # column 2 vanishes
    chain, allr = make_chain((RLI(),))
    r = O.run(PairMap(base).descriptor(), chain, z, W, 10, 1, 2)
    st = decode_states(r["states"], allr, 4, 2)[0]
    assert list(r["code"]) == [3, 3, 3] and list(r["hstep"]) == [1, 1, 1]
    assert list(st["code"]) == [2, 2, 1] and list(st["column"]) == [1, 2, 2]
    assert st["value"][1] > 0 and st["value"][2] == 0.0


def test_pair_start_separation_rule():
    from chaos_amd import pair_start, SeparationError, pair_distance
    base = WrappedStandardMap(1.0)
    z, dx, dr, mode = pair_start(base, (6.2830, 0.1), (1.0, 0.0), 1e-3)
    assert mode == "wrapped" and z[4] < 0.01 and z[6] == z[2] + 1   # B wrapped past 2 pi: its count steps
    assert abs(dr - 1e-3) < 1e-12 and pair_distance(base, z) == dr
    for bad in (0.0, -1e-9, float("nan"), float("inf")):
        with pytest.raises(SeparationError):
            pair_start(base, (1.0, 1.0), (1.0, 0.0), bad)
    with pytest.raises(SeparationError):   # a lifted gap of 2 pi reads as 0 (the minimal image)
        pair_start(StandardMap(1.0), (1.0, 1.0), (1.0, 0.0), 2 * math.pi)


@pytest.mark.parametrize("wrapped", [False, True], ids=["lifted", "wrapped"])
def test_oracle_pair_rows_equal_finals(wrapped):
    from chaos_amd import PairMap, pair_frame
    m, N, tau = 48, 200, 1
    base = (WrappedStandardMap if wrapped else StandardMap)(6.0)
    z, u = pair_starts(base, m, 9)
    psys = PairMap(base)
    members = ("RLI", "MEGNO")
    reds = ind_members(members, pair=True) + (RunMonitor(psys),)
    chain, allr = make_chain(reds)
    grid = record_grid(N, tau, points_per_decade=6, first_steps=10, extra_fractions=(0.5,))
    s = base.state_length
    W = np.broadcast_to(pair_frame(u), (m, 4, 2))
    rec = O.ind_record(psys.descriptor(), chain, z, W, N, tau, 2, grid, 2 + 6 + 3 + s)
    assert np.all(rec["fail"] == 0)
    for g, t in enumerate(grid):
        r = O.run(psys.descriptor(), chain, z, W, int(t), tau, 2)
        fin = finals_of(decode_states(r["states"], allr, 4, 2), allr, 4, 2)
        rl, mg, mon = fin
        cols = [rl["DL"], rl["RLI"]]
        for c in range(2):
            cols += [mg["Y"][:, c], mg["Ybar"][:, c], mg["Ybar20"][:, c]]
        cols += [mon["det_max"], mon["symp_max"], mon["M_max"]] + [r["x"][:, j] for j in range(s)]
        assert np.array_equal(rec["R"][:, g], np.stack(cols, axis=1))
    # the monitor reads orbit A: its maxima equal a RunMonitor run on orbit A alone
    mchain, mall = make_chain((RunMonitor(base),))
    b = O.run(base.descriptor(), mchain, z[:, :s], np.broadcast_to(u.reshape(2, 1), (m, 2, 1)), N, tau, 1)
    mon = decode_states(b["states"], mall, 2, 1)[0]
    assert np.array_equal(rec["R"][:, -1, 8], mon["det_max"]) and np.array_equal(rec["R"][:, -1, 10], mon["M_max"])
