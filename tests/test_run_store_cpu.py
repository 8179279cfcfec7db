"""The run store on the CPU: chunk files written and read back bit for bit, the HDF5.jl type and axis conventions
seen through the HDF5 library's own h5dump (an independent reader, as in verify/experiments-runner.jl:578-600), the
atomic write killed at each of its points (test/experiments/ex_store.jl:115-147), the store scan of a resume
(src/Experiments/run_resume.jl:24-131) and the orbit record of a device recording (run_record.jl:188-225)."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

from chaos_amd import h5
from chaos_amd import run_store as RS

try:
    h5.load()
except h5.H5Error as e:  # pragma: no cover - the image ships libhdf5 under /opt/conda/lib
    pytest.skip(f"no HDF5 library: {e}", allow_module_level=True)

H5DUMP = shutil.which("h5dump") or ("/opt/conda/bin/h5dump" if os.path.exists("/opt/conda/bin/h5dump") else None)


def _orbit(oid, n=7, k=2, wrapped=True, seed=0, ind=False):
    rng = np.random.default_rng(seed + oid)
    o = RS.OrbitRecord(
        id=oid, attrs={"system": "standard", "params": "0.97", "ic": f"0.{oid},0.5", "class_label": "regular",
                       "seed": str(1000 + oid), "item_hash": f"{oid:064x}", "design_class": "regular", "flags": [],
                       "steps_recorded": n},
        t=np.arange(1, n + 1, dtype=np.int64) * 10, state=rng.standard_normal((n, 2)),
        winding=rng.integers(-5, 5, (n, 2)).astype(np.int64) if wrapped else None,
        lyap={"L4": rng.standard_normal((n, k))},
        diagnostics={nm: np.abs(rng.standard_normal(n)) * 1e-15 for nm in ("det", "symp", "orth", "sum", "pair")},
        thresholds={nm: 1e-15 for nm in ("det", "symp", "orth", "sum", "pair")},
        first_exceedance={nm: 0 for nm in ("det", "symp", "orth", "sum", "pair")},
        cost={"L4": np.array([100.0, k, 0.0, 0.0, 4096.0])})
    if ind:
        o.indicators = {"st": RS.IndicatorTable(
            series=rng.standard_normal((n, 3)), columns=["mean", "var", "skew"], scalars=rng.standard_normal(2),
            scolumns=["lo", "hi"], range=np.array([-1.0, 2.0]), counts=np.arange(12, dtype=np.uint32).reshape(6, 2),
            pool=np.arange(12, dtype=np.int64).reshape(6, 2) * oid)}
    return o


def _attrs(chunk=1, h="h1", panel="p1"):
    return {"schema_version": 2, "study": "p1-le-benchmark", "chunk": chunk, "chunk_key": "L4:k1:regular",
            "chunk_hash": h, "panel_sha256": panel, "backend": "gpu", "backend_request": "auto", "device": "MI355X",
            "backend_fallback_reason": "", "precision": "Float64", "method": "L4", "horizon": 70, "tau": 1,
            "vectors": 2, "coordinate_mode": "wrapped", "status": "ok", "attempt": 1, "pilot": False,
            "first_attempt_reasons": [], "failure_reasons": [], "arms": ["main"], "dirty": False, "session": 1,
            "control_stop": False, "created_utc": "2026-12-04T12:00:00Z"}


def _store_bytes(path):
    """Every dataset's bytes and every attribute of a chunk file (the canonical content, I14)."""
    ch = RS.read_chunk(path)
    out = {"attrs": {k: v for k, v in ch["attrs"].items() if k != "created_utc"}}
    for o in ch["orbits"]:
        out[o.id] = (o.attrs, o.t.tobytes(), np.ascontiguousarray(o.state).tobytes(),
                     {m: np.ascontiguousarray(v).tobytes() for m, v in o.lyap.items()})
    return out


def test_write_read_bit_for_bit_with_special_values(tmp_path):
    orbs = [_orbit(i, ind=True) for i in (3, 1, 2)]
    m = orbs[0].lyap["L4"]
    m[0, 0], m[1, 0], m[2, 0], m[3, 0] = np.nan, np.inf, -0.0, 5e-324
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, _attrs(), orbs)
    back = RS.read_chunk(p)
    assert back["attrs"]["complete"] is True and not [f for f in os.listdir(tmp_path) if f.endswith(".tmp")]
    assert [o.id for o in back["orbits"]] == [1, 2, 3]  # id order
    by_id = {o.id: o for o in orbs}
    assert all(RS.same_orbit(by_id[o.id], o) for o in back["orbits"])
    b = back["orbits"][2].lyap["L4"]
    assert np.isnan(b[0, 0]) and b[1, 0] == np.inf and np.signbit(b[2, 0]) and b[3, 0] == 5e-324
    assert RS.attr_diff({**_attrs(), "complete": True}, back["attrs"]) == []
    # a planted defect in the buffers is seen by the comparison the write relies on
    other = _orbit(3, ind=True)
    other.lyap["L4"][4, 1] = np.nextafter(other.lyap["L4"][4, 1], np.inf)
    assert not RS.same_orbit(_orbit(3, ind=True), other)
    with pytest.raises(FileNotFoundError):
        RS.read_chunk(str(tmp_path / "chunk-00009.h5"))


def test_attribute_kinds_survive(tmp_path):
    a = {"b_true": True, "b_false": False, "i": -7, "f": -0.0, "s": "naïve text", "empty": "", "sl": ["x", "yz"],
         "sl_empty": [], "il": [1, 2, 3], "fl": [0.5, np.inf], "big": 2 ** 62}
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, {**_attrs(), **a}, [_orbit(1)])
    got = RS.read_chunk_attrs(p)
    assert got["b_true"] is True and got["b_false"] is False and got["complete"] is True
    assert type(got["i"]) is int and got["i"] == -7 and got["big"] == 2 ** 62
    assert type(got["f"]) is float and np.signbit(got["f"])
    assert got["s"] == "naïve text" and got["empty"] == "" and got["sl"] == ["x", "yz"] and got["sl_empty"] == []
    assert list(got["il"]) == [1, 2, 3] and np.asarray(got["il"]).dtype == np.int64
    assert list(got["fl"]) == [0.5, np.inf]
    assert RS.attr_diff({**_attrs(), **a, "complete": True}, got) == []
    assert RS.attr_diff({"complete": 1}, {"complete": True}) == ["complete"]  # a Bool is not an integer here


def test_layout_is_the_julia_shape(tmp_path):
    o = _orbit(5, n=9, k=2, ind=True)
    p = str(tmp_path / "chunk-00002.h5")
    RS.write_chunk(p, _attrs(chunk=2), [o])
    lay = RS.chunk_layout(p)
    b = "/orbits/5"
    assert lay[b + "/t"] == ("Int64", (9,))
    assert lay[b + "/state"] == ("Float64", (9, 2))
    assert lay[b + "/winding"] == ("Int64", (9, 2))
    assert lay[b + "/lyap/L4"] == ("Float64", (9, 2))
    assert lay[b + "/cost/L4"] == ("Float64", (5,))
    assert all(lay[f"{b}/diagnostics/{d}"] == ("Float64", (9,)) for d in ("det", "symp", "orth", "sum", "pair"))
    assert lay[b + "/indicators/st/counts"] == ("UInt32", (6, 2))
    assert lay["/pooled/st/counts"] == ("Int64", (6, 2))
    assert not any(k.startswith("/grid") for k in lay)
    q = str(tmp_path / "chunk-00003.h5")
    RS.write_chunk(q, _attrs(chunk=3), [_orbit(6, wrapped=False)])
    assert "/orbits/6/winding" not in RS.chunk_layout(q)  # lifted mode has no winding


@pytest.mark.skipif(H5DUMP is None, reason="h5dump not found")
def test_h5dump_reads_values_types_and_reversed_extents(tmp_path):
    """The HDF5 library's own reader: every value of lyap, state and t equal (the reference's spot check), the file
    extents the Julia shape reversed, and the attribute types HDF5.jl writes (B8 Bool, UTF-8 strings, I64, F64)."""
    o = _orbit(4, n=6, k=2)
    o.lyap["L4"][1, 1] = -0.0
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, _attrs(), [o])

    def dump(*args):
        return subprocess.run([H5DUMP, *args, p], capture_output=True, text=True, check=True).stdout

    for path, A in (("/orbits/4/lyap/L4", o.lyap["L4"]), ("/orbits/4/state", o.state), ("/orbits/4/t", o.t)):
        txt = dump("-d", path, "-m", "%.17g", "-y", "-w", "0")
        body = re.search(r"DATA \{([^}]*)\}", txt).group(1)
        toks = [t for t in re.split(r"[,\s]+", body) if t]
        vals = np.array([int(t) for t in toks] if A.dtype == np.int64 else [float(t) for t in toks])
        flat = np.asarray(A).T.ravel()  # Julia's vec(A): the file's C order over the reversed extents
        assert vals.shape == flat.shape and np.array_equal(vals, flat)
        ext = tuple(int(x) for x in re.search(r"DATASPACE\s+SIMPLE \{ \( ([\d, ]+) \)", txt).group(1).split(","))
        assert ext == tuple(reversed(A.shape))
    assert "-0" in dump("-d", "/orbits/4/lyap/L4", "-m", "%.17g", "-y", "-w", "0")
    hdr = dump("-A")  # headers and attribute values
    blocks = dict(re.findall(r'ATTRIBUTE "(\w+)" \{\s*DATATYPE\s+(H5T_\w+)', hdr))
    assert blocks["complete"] == "H5T_STD_B8LE" and blocks["pilot"] == "H5T_STD_B8LE"
    assert blocks["chunk"] == "H5T_STD_I64LE" and blocks["schema_version"] == "H5T_STD_I64LE"
    assert blocks["study"] == "H5T_STRING" and blocks["arms"] == "H5T_STRING"
    assert re.search(r'ATTRIBUTE "arms" \{\s*DATATYPE\s+H5T_STRING \{\s*STRSIZE H5T_VARIABLE;[^}]*CSET H5T_CSET_UTF8',
                     hdr)
    assert re.search(r'ATTRIBUTE "study" \{\s*DATATYPE\s+H5T_STRING \{\s*STRSIZE 15;[^}]*CSET H5T_CSET_UTF8', hdr)
    assert re.search(r'ATTRIBUTE "complete" \{\s*DATATYPE\s+H5T_STD_B8LE\s*DATASPACE\s+SCALAR\s*DATA \{\s*\(0\): 0x01',
                     hdr)


class _Kill(Exception):
    pass


@pytest.mark.parametrize("point", ["before_data", "during_data", "after_data", "before_rename", "after_rename"])
def test_atomic_write_killed_then_resumed(tmp_path, point):
    """No file with a final name is incomplete; a resume removes the temporaries, recomputes the chunk unless its
    file was already final, and the store then equals the uninterrupted one."""
    hashes = {1: "h1", 2: "h2", 3: "h3"}
    ref = tmp_path / "ref"
    st = tmp_path / "killed"
    data = {c: [_orbit(10 * c + j, ind=True) for j in range(3)] for c in hashes}
    for c in hashes:
        RS.write_chunk(RS.chunk_path(str(ref), c), _attrs(c, hashes[c]), data[c])

    def kill(p):
        if p == point:
            raise _Kill(p)

    RS.write_chunk(RS.chunk_path(str(st), 1), _attrs(1, "h1"), data[1])
    with pytest.raises(_Kill):
        RS.write_chunk(RS.chunk_path(str(st), 2), _attrs(2, "h2"), data[2], hook=kill)
    for f in RS.chunk_files(str(st)):
        assert RS.read_chunk_attrs(f)["complete"] is True
    assert os.path.isfile(RS.chunk_path(str(st), 2)) == (point == "after_rename")
    tmps = sorted(f for f in os.listdir(st) if f.endswith(".tmp"))
    assert tmps == ([] if point == "after_rename" else ["chunk-00002.h5.tmp"])
    pending, keep, removed, moved = RS.pending_chunks(str(st), hashes, "p1")
    assert removed == tmps and moved == [] and not [f for f in os.listdir(st) if f.endswith(".tmp")]
    assert pending == ([3] if point == "after_rename" else [2, 3])
    for c in pending:
        RS.write_chunk(RS.chunk_path(str(st), c), _attrs(c, hashes[c]), data[c])
    for c in hashes:
        assert _store_bytes(RS.chunk_path(str(st), c)) == _store_bytes(RS.chunk_path(str(ref), c))


def test_classify_store_and_stale_handling(tmp_path):
    s = str(tmp_path)
    hashes = {1: "h1", 2: "h2", 3: "h3", 4: "h4"}
    RS.write_chunk(RS.chunk_path(s, 1), _attrs(1, "h1"), [_orbit(1)])
    RS.write_chunk(RS.chunk_path(s, 2), _attrs(2, "old"), [_orbit(2)])                 # another chunk hash: stale
    RS.write_chunk(RS.chunk_path(s, 3), {**_attrs(3, "h3"), "control_stop": True}, [_orbit(3)])   # halted
    RS.write_chunk(RS.chunk_path(s, 4), {**_attrs(4, "h4"), "pilot": True}, [_orbit(4)])          # a pilot: stale
    RS.write_chunk(RS.chunk_path(s, 7), _attrs(7, "h7"), [_orbit(7)])                  # beyond the plan: stale
    with open(os.path.join(s, "notes.txt"), "w") as f:
        f.write("not a chunk")
    keep, stale, halted = RS.classify_store(s, hashes, "p1")
    assert keep == {1: RS.chunk_path(s, 1)} and halted == {3: RS.chunk_path(s, 3)}
    assert sorted(os.path.basename(f) for f in stale) == ["chunk-00002.h5", "chunk-00004.h5", "chunk-00007.h5"]
    k2, _, _ = RS.classify_store(s, hashes, "another panel")   # a different panel file: nothing is kept
    assert k2 == {}
    with pytest.raises(ValueError, match="control bound"):
        RS.refuse_halted(halted)
    os.rename(RS.chunk_path(s, 3), os.path.join(s, "held.h5"))
    with pytest.raises(ValueError, match="stale chunk files"):
        RS.pending_chunks(s, hashes, "p1")
    pending, keep, removed, moved = RS.pending_chunks(s, hashes, "p1", stale="move")
    assert pending == [2, 3, 4] and sorted(moved) == ["chunk-00002.h5", "chunk-00004.h5", "chunk-00007.h5"]
    assert sorted(os.listdir(os.path.join(s, "stale"))) == moved
    with open(RS.chunk_path(s, 2), "wb") as f:   # an unreadable file is stale, and a second move is numbered
        f.write(b"\x01\x02\x03")
    _, stale, _ = RS.classify_store(s, hashes, "p1")
    assert stale == [RS.chunk_path(s, 2)]
    assert RS.move_aside(s, stale[0], "stale").endswith("chunk-00002.h5.1")


def test_final_name_taken_moves_aside(tmp_path):
    p = RS.chunk_path(str(tmp_path), 1)
    RS.write_chunk(p, _attrs(1, "a"), [_orbit(1)])
    RS.write_chunk(p, _attrs(1, "b"), [_orbit(2)])
    assert RS.read_chunk_attrs(p)["chunk_hash"] == "b"
    assert RS.read_chunk_attrs(p + ".old1")["chunk_hash"] == "a"


def test_pooled_counts_shared_and_unshared_ranges(tmp_path):
    a, b = _orbit(1, ind=True), _orbit(2, ind=True)
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, _attrs(), [a, b])
    got = RS.read_chunk(p)["pooled"]["st"]
    assert got["orbits"] == 2 and np.array_equal(got["counts"], a.indicators["st"].pool + b.indicators["st"].pool)
    assert list(got["range"]) == [-1.0, 2.0]
    b.indicators["st"].range = np.array([-1.0, np.nextafter(2.0, 3.0)])
    q = str(tmp_path / "chunk-00002.h5")
    RS.write_chunk(q, _attrs(2), [a, b])
    got = RS.read_chunk(q)["pooled"]["st"]
    assert got["orbits"] == 2 and got["counts"] is None and got["range"] is None
    assert "/pooled/st/counts" not in RS.chunk_layout(q)


def test_provenance_groups(tmp_path):
    prov = {"MEGNO": {"replace": False, "k": 2, "spectrum_lo": "0.1234567890123456789012345"},
            "FLI": {"replace": True, "norm": "L2"}}
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, {**_attrs(), "provenance": prov, "kind": "tangent", "members": ["MEGNO", "FLI"]}, [_orbit(1)])
    back = RS.read_chunk(p)["attrs"]
    assert back["provenance"] == prov and back["members"] == ["MEGNO", "FLI"] and back["kind"] == "tangent"


def _fake_records(method="L4", k=2, s=4, G=6, m=3, fail=(0, 4, 1)):
    """A Records object as the device recorder returns it, with synthetic rows."""
    from chaos_amd.recorder import Records
    rng = np.random.default_rng(7)
    ncol = k + 7 + s
    R = rng.standard_normal((m, G, ncol)) * 1e-16
    R[:, :, k + 5] = 1.5 + np.arange(G)        # G_max
    R[:, :, k + 6] = 2.0                         # M_max
    R[:, :, k + 7:k + 9] = rng.random((m, G, 2))
    if s == 4:
        R[:, :, k + 9:k + 11] = rng.integers(-3, 3, (m, G, 2))
    grid = np.array([10, 20, 50, 100, 200, 500], dtype=np.int64)[:G]
    return Records(method=method, grid=grid, R=R, fail=np.array(fail, dtype=np.int64), code=np.zeros(m, np.int8),
                   hstep=np.zeros(m, np.int64), x=np.zeros((m, s)), W=np.zeros((m, 2, k)),
                   states=np.zeros((m, 8), np.uint8), k=k, s=s)


def test_orbit_record_of_a_device_recording(tmp_path):
    rec = _fake_records()
    thr = RS.record_thresholds(rec, 1, tau=1, det_residual_tol=64.0, pairing_residual_tol=64.0, orthogonality_tol=8.0)
    assert sorted(thr) == ["det", "orth", "pair", "sum", "symp"]
    # orbit 1 fails at row 4: the last good row is 3 (G_max 1.5 + 2, M_max 2)
    gmax, eps = 3.5, 2.0 ** -52
    assert thr["orth"] == ((8.0 * eps) * 2.0) * (gmax + 1.0 / gmax)   # MGS: g = G_w
    assert thr["det"] == (64.0 * eps) * 3.0 and thr["sum"] == thr["pair"]
    o = RS.orbit_record(rec, 1, orbit_id=42, attrs={"system": "standard"}, thresholds=thr)
    assert o.t.tolist() == [10, 20, 50, 0, 0, 0] and np.isnan(o.lyap["L4"][3:]).all() and np.isnan(o.state[3:]).all()
    assert o.winding.dtype == np.int64 and o.winding.shape == (6, 2)
    assert sorted(o.diagnostics) == ["det", "orth", "pair", "sum", "symp"]
    d = o.diagnostics["det"].copy()
    d[:3] = [0.0, 1.0, 0.0]
    o.diagnostics["det"] = d
    assert RS.first_exceedance(d, 0.5, o.t) == 20 and RS.first_exceedance(d, 2.0, o.t) == 0
    p = str(tmp_path / "chunk-00001.h5")
    RS.write_chunk(p, _attrs(), [o])
    assert RS.same_orbit(o, RS.read_chunk(p)["orbits"][0])
    hh = RS.record_thresholds(_fake_records(method="L5"), 0, tau=1, det_residual_tol=64.0, pairing_residual_tol=64.0,
                              orthogonality_tol=8.0)
    assert hh["orth"] == (8.0 * eps) * 2.0                             # Householder: g = 1
    assert RS.diagnostic_names("L1") == ["det", "symp"] and RS.diagnostic_names("L2") == []
    with pytest.raises(ValueError, match="no threshold"):
        RS.orbit_record(rec, 0, orbit_id=1, attrs={}, thresholds={"det": 1.0})
