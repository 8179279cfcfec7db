"""Device recordings through the run store: the chunk files hold the device recorder's rows bit for bit, and a
campaign killed during a chunk's write and resumed from the store (temporaries removed, the chunk recomputed on the
device) leaves the same store as an uninterrupted one (I14, src/Experiments/run_resume.jl:24-131,
test/experiments/ex_store.jl:115-147)."""
import os
import sys

import numpy as np
import pytest

from chaos_amd import StandardMap, WrappedStandardMap, torus_state
from chaos_amd import h5
from chaos_amd import run_store as RS
from chaos_amd.inputs import torus_uniform
from chaos_amd.recorder import record

pytestmark = pytest.mark.gpu

TOLS = dict(tau=1, det_residual_tol=64.0, pairing_residual_tol=64.0, orthogonality_tol=8.0)


def say(*a):
    print("[gpu run store]", *a, flush=True)
    sys.stdout.flush()


def _chunk(rec, rows, ids, method, wrapped, K):
    orbs = []
    for j, oid in zip(rows, ids):
        u = rec.unpack(j)
        attrs = {"system": "standard", "params": repr(K), "ic": f"orbit {oid}", "class_label": "none",
                 "seed": str(oid), "item_hash": f"{oid:064x}", "design_class": "none", "flags": [],
                 "steps_recorded": int(len(u["t"]))}
        orbs.append(RS.orbit_record(rec, j, orbit_id=oid, attrs=attrs, thresholds=RS.record_thresholds(rec, j, **TOLS)))
    return orbs


def _root(c, h):
    return {"schema_version": 2, "study": "device-store", "chunk": c, "chunk_hash": h, "panel_sha256": "panel",
            "backend": "gpu", "backend_request": "gpu", "device": "MI355X", "backend_fallback_reason": "",
            "precision": "Float64", "status": "ok", "attempt": 1, "pilot": False, "control_stop": False,
            "failure_reasons": [], "arms": ["main"]}


def _canon(path):
    ch = RS.read_chunk(path)
    return {o.id: (o.t.tobytes(), np.ascontiguousarray(o.state).tobytes(),
                   None if o.winding is None else np.ascontiguousarray(o.winding).tobytes(),
                   {m: np.ascontiguousarray(v).tobytes() for m, v in o.lyap.items()},
                   {m: np.ascontiguousarray(v).tobytes() for m, v in o.diagnostics.items()}) for o in ch["orbits"]}


class _Kill(Exception):
    pass


@pytest.mark.parametrize("wrapped,method", [(False, "L4"), (True, "L4"), (True, "L2")],
                         ids=["lifted-L4", "wrapped-L4", "wrapped-L2"])
def test_device_records_through_the_store_and_resume(tmp_path, wrapped, method):
    try:
        h5.load()
    except h5.H5Error as e:
        pytest.skip(f"no HDF5 library: {e}")
    K, m, N = 0.97, 96, 3000
    sysm = WrappedStandardMap(K) if wrapped else StandardMap(K)
    z = torus_uniform(11, m)
    x0 = np.array([torus_state(p) for p in z]) if wrapped else z
    opts = {"eps": "1e-8"} if method == "L2" else {}
    say("record", "wrapped" if wrapped else "lifted", m, "orbits", N, "steps", method)
    rec = record(sysm, x0, N, method=method, tau=1, **opts)
    ids = np.arange(1, m + 1)
    half = m // 2
    chunks = {1: (np.arange(half), ids[:half]), 2: (np.arange(half, m), ids[half:])}
    hashes = {1: "c1", 2: "c2"}
    ref = str(tmp_path / "ref")
    for c, (rows, oids) in chunks.items():
        RS.write_chunk(RS.chunk_path(ref, c), _root(c, hashes[c]), _chunk(rec, rows, oids, method, wrapped, K))
    # the files hold the device rows bit for bit
    for c, (rows, oids) in chunks.items():
        back = RS.read_chunk(RS.chunk_path(ref, c))["orbits"]
        for j, o in zip(rows, back):
            u = rec.unpack(j)
            assert o.t.tobytes() == u["t"].tobytes() and o.lyap[method].tobytes() == u["lyap"].tobytes()
            assert o.state.tobytes() == np.ascontiguousarray(u["state"]).tobytes()
            assert (o.winding is None) == (not wrapped)
            assert sorted(o.diagnostics) == sorted(RS.diagnostic_names(method))
            if wrapped:
                assert o.winding.tobytes() == u["winding"].tobytes()
    say("chunk files equal the device rows")
    # a campaign killed before chunk 2's rename, resumed: chunk 2 is recomputed on the device from its own starts
    st = str(tmp_path / "killed")
    RS.write_chunk(RS.chunk_path(st, 1), _root(1, "c1"), _chunk(rec, *chunks[1], method, wrapped, K))

    def kill(point):
        if point == "before_rename":
            raise _Kill(point)

    with pytest.raises(_Kill):
        RS.write_chunk(RS.chunk_path(st, 2), _root(2, "c2"), _chunk(rec, *chunks[2], method, wrapped, K), hook=kill)
    pending, keep, removed, moved = RS.pending_chunks(st, hashes, "panel")
    assert pending == [2] and removed == ["chunk-00002.h5.tmp"] and moved == []
    rows2, oids2 = chunks[2]
    rec2 = record(sysm, x0[rows2], N, method=method, tau=1, **opts)
    RS.write_chunk(RS.chunk_path(st, 2), _root(2, "c2"), _chunk(rec2, np.arange(len(rows2)), oids2, method, wrapped, K))
    for c in hashes:
        assert _canon(RS.chunk_path(st, c)) == _canon(RS.chunk_path(ref, c))
    assert sorted(os.listdir(st)) == ["chunk-00001.h5", "chunk-00002.h5"]
    say("resumed store equals the uninterrupted one")
