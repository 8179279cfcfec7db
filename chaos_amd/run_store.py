"""The run store (schema/run-store.md, version 2): one HDF5 file per chunk, written under a temporary name, read
back and compared bit for bit, marked complete, flushed to the disk and renamed, so that no file with a final name
is incomplete (src/Experiments/run_store.jl:143-181). Also the store scan a resume starts from
(src/Experiments/run_resume.jl:5-50, 124-131) and the orbit record of the device recorder's output
(src/Experiments/run_record.jl:188-225).

Layout (run_store.jl:87-106, run_indicator_store.jl:14-61): root attributes (with `complete`); `/provenance/<member>`
groups of attributes; `/orbits/<id>/` with the orbit attributes, `t`, `state`, `winding` (wrapped mode), `lyap/<method>`,
`diagnostics/<name>` (attributes `threshold`, `first_exceedance_step`), `cost/<method>` and
`indicators/<id>/{series, final, flags, scalars, counts}`; `/pooled/<id>` (attribute `orbits`, and `counts` with the
attribute `range` when the pooled orbits share one range). Arrays are held in their Julia shapes ((n, k) for an
exponent series); the file extents are reversed, as HDF5.jl writes them (chaos_amd/h5.py).
"""
import os
import re
import struct
from dataclasses import dataclass, field

import numpy as np

from . import h5

SCHEMA_VERSION = 2


@dataclass
class IndicatorTable:
    """The tables of one indicator spec on one orbit (run_indicator_store.jl:14-27); `pool` (the counts pooled over
    a chunk) is kept in memory only."""
    series: np.ndarray = None
    columns: list = field(default_factory=list)
    final: np.ndarray = None
    rows: list = field(default_factory=list)
    fcolumns: list = field(default_factory=list)
    flags: np.ndarray = None
    bits: list = field(default_factory=list)
    scalars: np.ndarray = None
    scolumns: list = field(default_factory=list)
    range: np.ndarray = None
    counts: np.ndarray = None
    pool: np.ndarray = None


@dataclass
class OrbitRecord:
    """One orbit's group (src/Experiments/run_record.jl:6-19): `t` (n,), `state` (n, 2), `winding` (n, 2) or None,
    `lyap` {method: (n, k)}, `diagnostics` {name: (n,)} with `thresholds` and `first_exceedance`, `cost`
    {method: (5,)}, `indicators` {spec id: IndicatorTable}."""
    id: int
    attrs: dict
    t: np.ndarray
    state: np.ndarray
    winding: np.ndarray = None
    lyap: dict = field(default_factory=dict)
    diagnostics: dict = field(default_factory=dict)
    thresholds: dict = field(default_factory=dict)
    first_exceedance: dict = field(default_factory=dict)
    cost: dict = field(default_factory=dict)
    failure: str = ""
    indicators: dict = field(default_factory=dict)


class StoreError(RuntimeError):
    pass


# ---------------------------------------------------------------- equality bit for bit
def _same_array(a, b):
    """Equal shapes, element types and bytes (NaN payloads, -0.0 and subnormals included; run_store.jl:110-113)."""
    if a is None or b is None:
        return a is None and b is None
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.ascontiguousarray(a).tobytes() == np.ascontiguousarray(b).tobytes()


def _same_value(a, b):
    """Attribute equality: the same kind (bool, integer, float, string, list) and the same bits."""
    if isinstance(a, (bool, np.bool_)) or isinstance(b, (bool, np.bool_)):
        return isinstance(a, (bool, np.bool_)) and isinstance(b, (bool, np.bool_)) and bool(a) == bool(b)
    if isinstance(a, (int, np.integer)) and isinstance(b, (int, np.integer)):
        return int(a) == int(b)
    if isinstance(a, (float, np.floating)) and isinstance(b, (float, np.floating)):
        return struct.pack("<d", float(a)) == struct.pack("<d", float(b))
    if isinstance(a, str) and isinstance(b, str):
        return a == b
    if isinstance(a, (list, tuple)) and isinstance(b, (list, tuple)) and all(isinstance(s, str) for s in [*a, *b]):
        return list(a) == list(b)
    if isinstance(a, (list, tuple, np.ndarray)) and isinstance(b, (list, tuple, np.ndarray)):
        x, y = np.asarray(a), np.asarray(b)
        return x.shape == y.shape and x.dtype.kind == y.dtype.kind and x.tobytes() == y.astype(x.dtype).tobytes()
    return False


def _same_attrs(a, b):
    return set(a) == set(b) and all(_same_value(a[k], b[k]) for k in a)


def attr_diff(want, got):
    """Names of the attributes whose read-back differs from what was written (run_store.jl:129)."""
    return sorted(k for k in set(want) | set(got) if not (k in want and k in got and _same_value(want[k], got[k])))


_IND_FIELDS = ("series", "columns", "final", "rows", "fcolumns", "flags", "bits", "scalars", "scolumns", "range",
               "counts")


def _same_indicators(a, b):
    if set(a) != set(b):
        return False
    for i in a:
        for f in _IND_FIELDS:
            x, y = getattr(a[i], f), getattr(b[i], f)
            if f in ("columns", "rows", "fcolumns", "bits", "scolumns"):
                if list(x or []) != list(y or []):
                    return False
            elif not _same_array(x, y):
                return False
    return True


def same_orbit(a, b):
    """The orbit records hold the same data bit for bit (run_store.jl:115-122)."""
    ok = a.id == b.id and _same_attrs(a.attrs, b.attrs) and _same_array(a.t, b.t) and _same_array(a.state, b.state)
    ok = ok and _same_array(a.winding, b.winding) and set(a.lyap) == set(b.lyap)
    ok = ok and all(_same_array(a.lyap[m], b.lyap[m]) for m in a.lyap) and _same_indicators(a.indicators, b.indicators)
    ok = ok and set(a.diagnostics) == set(b.diagnostics) and all(
        _same_array(a.diagnostics[m], b.diagnostics[m]) and _same_value(float(a.thresholds[m]), float(b.thresholds[m]))
        and int(a.first_exceedance[m]) == int(b.first_exceedance[m]) for m in a.diagnostics)
    return ok and set(a.cost) == set(b.cost) and all(_same_array(a.cost[m], b.cost[m]) for m in a.cost)


# ---------------------------------------------------------------- pooled counts
def pooled_of(orbits):
    """The pooled counts of a chunk (run_indicator_store.jl:42-53): per spec id with in-memory pools, the number of
    orbits, and, when all of them share one range bit for bit, the exact Int64 sum of their pools and that range."""
    out = {}
    ids = sorted({i for o in orbits for i, t in o.indicators.items() if t.pool is not None})
    for i in ids:
        ts = [o.indicators[i] for o in orbits if i in o.indicators and o.indicators[i].pool is not None]
        shared = all(_same_array(t.range, ts[0].range) for t in ts)
        counts = None
        if shared:
            counts = np.zeros(np.shape(ts[0].pool), dtype=np.int64)
            for t in ts:
                counts += np.asarray(t.pool, dtype=np.int64)
        out[i] = dict(orbits=len(ts), counts=counts, range=ts[0].range if shared else None)
    return out


def _same_pooled(a, b):
    return set(a) == set(b) and all(
        a[i]["orbits"] == b[i]["orbits"] and _same_array(a[i]["counts"], b[i]["counts"])
        and _same_array(None if a[i]["range"] is None else np.asarray(a[i]["range"], np.float64),
                        None if b[i]["range"] is None else np.asarray(b[i]["range"], np.float64)) for i in a)


# ---------------------------------------------------------------- writing
def _put_attrs(node, attrs):
    for k, v in attrs.items():
        node.set_attr(k, v)


def _put_ds(node, path, data, **attrs):
    with node.write_dataset(path, data) as d:
        _put_attrs(d, attrs)


def _put_orbit(f, o):
    """One orbit group (run_store.jl:87-106, run_indicator_store.jl:14-27)."""
    with f.create_group(f"orbits/{o.id}") as g:
        _put_attrs(g, o.attrs)
        _put_ds(g, "t", np.asarray(o.t, dtype=np.int64))
        _put_ds(g, "state", np.asarray(o.state, dtype=np.float64))
        if o.winding is not None:
            _put_ds(g, "winding", np.asarray(o.winding, dtype=np.int64))
        for m, v in o.lyap.items():
            _put_ds(g, "lyap/" + m, np.asarray(v, dtype=np.float64))
        for nm, v in o.diagnostics.items():
            _put_ds(g, "diagnostics/" + nm, np.asarray(v, dtype=np.float64), threshold=float(o.thresholds[nm]),
                    first_exceedance_step=int(o.first_exceedance[nm]))
        for m, c in o.cost.items():
            _put_ds(g, "cost/" + m, np.asarray(c, dtype=np.float64))
        for i, t in o.indicators.items():
            base = f"indicators/{i}/"
            if t.series is not None:
                _put_ds(g, base + "series", t.series, columns=list(t.columns))
            if t.final is not None:
                _put_ds(g, base + "final", t.final, columns=list(t.fcolumns), rows=list(t.rows))
                _put_ds(g, base + "flags", np.asarray(t.flags, dtype=np.int64), bits=list(t.bits))
            if t.scalars is not None:
                extra = {} if t.range is None else {"range": np.asarray(t.range, dtype=np.float64)}
                _put_ds(g, base + "scalars", t.scalars, columns=list(t.scolumns), **extra)
            if t.counts is not None:
                _put_ds(g, base + "counts", np.asarray(t.counts, dtype=np.uint32))


def _put_provenance(f, prov):
    for m, d in prov.items():
        with f.create_group(f"provenance/{m}") as g:
            _put_attrs(g, d)


def _put_pooled(f, orbits):
    for i, p in sorted(pooled_of(orbits).items()):
        with f.create_group(f"pooled/{i}") as g:
            g.set_attr("orbits", int(p["orbits"]))
            if p["counts"] is None:
                continue
            _put_ds(g, "counts", p["counts"])
            g.set_attr("range", np.asarray(p["range"], dtype=np.float64))


def _fsync(path):
    fd = os.open(path, os.O_RDWR)
    try:
        os.fsync(fd)
    finally:
        os.close(fd)


def _take_name(path):
    """A final name that is taken is moved to `<name>.old<k>` first (run_store.jl:131-138)."""
    if not os.path.lexists(path):
        return
    k = 1
    while os.path.lexists(f"{path}.old{k}"):
        k += 1
    os.rename(path, f"{path}.old{k}")


def _nohook(point):
    return None


def write_chunk(path, attrs, orbits, hook=_nohook, onread=_nohook):
    """Write a chunk file atomically (run_store.jl:140-181): under `path + ".tmp"` with `complete` false, the
    provenance, orbit and pooled groups; close; read back and compare bit for bit (root attributes included); set
    `complete`; flush to the disk; move a file already at `path` to `<path>.old<k>`; rename. `attrs` may hold
    "provenance" ({member: {key: value}}), written as groups. `hook(point)` is called at :before_data,
    :during_data (after half of the orbits), :after_data, :before_rename and :after_rename (as strings);
    `onread(kind)` before each read-back ("datasets", "attrs")."""
    tmp = path + ".tmp"
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    want = dict(attrs)
    want["complete"] = False
    prov = want.pop("provenance", {})
    with h5.open_file(tmp, "w") as f:
        _put_attrs(f, want)
        _put_provenance(f, prov)
        f.create_group("orbits").close()
        hook("before_data")
        half = -(-len(orbits) // 2)
        for j, o in enumerate(orbits, 1):
            _put_orbit(f, o)
            if j == half:
                hook("during_data")
        _put_pooled(f, orbits)
    onread("datasets")
    back = read_chunk(tmp)
    ordered = sorted(orbits, key=lambda o: int(o.id))  # the file lists the orbits in id order
    if not (len(back["orbits"]) == len(orbits) and all(same_orbit(a, b) for a, b in zip(ordered, back["orbits"]))):
        raise StoreError(f"write_chunk: {tmp} differs from the buffers after the write")
    if not _same_pooled(pooled_of(orbits), back["pooled"]):
        raise StoreError(f"write_chunk: pooled counts of {tmp} differ from the buffers after the write")
    bprov = back["attrs"].get("provenance", {})
    if set(prov) != set(bprov) or not all(_same_attrs(prov[m], bprov[m]) for m in prov):
        raise StoreError(f"write_chunk: provenance of {tmp} differs from the buffers after the write")
    bad = attr_diff(want, {k: v for k, v in back["attrs"].items() if k != "provenance"})
    if bad:
        raise StoreError(f"write_chunk: root attributes of {tmp} differ from the buffers after the write: "
                         + ", ".join(bad))
    hook("after_data")
    with h5.open_file(tmp, "r+") as f:
        f.del_attr("complete")
        f.set_attr("complete", True)
    onread("attrs")
    if read_chunk_attrs(tmp).get("complete") is not True:
        raise StoreError(f"write_chunk: {tmp} does not read back complete")
    _fsync(tmp)
    hook("before_rename")
    _take_name(path)
    os.rename(tmp, path)
    hook("after_rename")
    return path


# ---------------------------------------------------------------- reading
def _read_ds(g, name):
    if not g.has(name):
        return None, {}
    with g.dataset(name) as d:
        return d.read(), d.attrs()


def _read_indicators(g):
    out = {}
    if not g.has("indicators"):
        return out
    with g.group("indicators") as ig:
        for i in ig.keys():
            with ig.group(i) as s:
                series, sa = _read_ds(s, "series")
                final, fa = _read_ds(s, "final")
                flags, ga = _read_ds(s, "flags")
                scalars, ca = _read_ds(s, "scalars")
                counts, _ = _read_ds(s, "counts")
            out[i] = IndicatorTable(series=series, columns=sa.get("columns", []), final=final, rows=fa.get("rows", []),
                                    fcolumns=fa.get("columns", []), flags=flags, bits=ga.get("bits", []),
                                    scalars=scalars, scolumns=ca.get("columns", []), range=ca.get("range"),
                                    counts=counts)
    return out


def _read_group_arrays(g, name):
    if not g.has(name):
        return {}
    with g.group(name) as sub:
        out = {}
        for m in sub.keys():
            with sub.dataset(m) as d:
                out[m] = d.read()
        return out


def _read_orbit(g, oid):
    """One orbit group (run_store.jl:187-209)."""
    diag, thr, fe = {}, {}, {}
    if g.has("diagnostics"):
        with g.group("diagnostics") as dg:
            for nm in dg.keys():
                with dg.dataset(nm) as d:
                    diag[nm] = d.read()
                    thr[nm] = float(d.attr("threshold"))
                    fe[nm] = int(d.attr("first_exceedance_step"))
    wind = None
    if g.has("winding"):
        with g.dataset("winding") as d:
            wind = d.read()
    with g.dataset("t") as d:
        t = d.read()
    with g.dataset("state") as d:
        state = d.read()
    return OrbitRecord(id=oid, attrs=g.attrs(), t=t, state=state, winding=wind, lyap=_read_group_arrays(g, "lyap"),
                       diagnostics=diag, thresholds=thr, first_exceedance=fe, cost=_read_group_arrays(g, "cost"),
                       indicators=_read_indicators(g))


def _read_pooled(f):
    out = {}
    if not f.has("pooled"):
        return out
    with f.group("pooled") as pg:
        for i in pg.keys():
            with pg.group(i) as g:
                counts = None
                if g.has("counts"):
                    with g.dataset("counts") as d:
                        counts = d.read()
                names = g.attr_names()
                out[i] = dict(orbits=int(g.attr("orbits")), counts=counts,
                              range=g.attr("range") if "range" in names else None)
    return out


def read_chunk(path):
    """A chunk file as {"attrs", "orbits" (in id order), "pooled"} (run_store.jl:211-221); the provenance groups are
    the attribute "provenance"."""
    if not os.path.isfile(path):
        raise FileNotFoundError(f"no chunk file {path}")
    with h5.open_file(path, "r") as f:
        attrs = f.attrs()
        if f.has("provenance"):
            with f.group("provenance") as pg:
                prov = {}
                for m in pg.keys():
                    with pg.group(m) as g:
                        prov[m] = g.attrs()
            attrs["provenance"] = prov
        with f.group("orbits") as og:
            ids = sorted(int(k) for k in og.keys())
            orbits = []
            for i in ids:
                with og.group(str(i)) as g:
                    orbits.append(_read_orbit(g, i))
        return {"attrs": attrs, "orbits": orbits, "pooled": _read_pooled(f)}


def read_chunk_attrs(path):
    """The root attributes only (run_store.jl:224)."""
    with h5.open_file(path, "r") as f:
        return f.attrs()


def chunk_layout(path):
    """Dataset path => (Julia element type name, Julia size) of every dataset of a chunk file (run_store.jl:226-239)."""
    out = {}

    def walk(node, prefix):
        for k in node.keys():
            if node.is_dataset(k):
                with node.dataset(k) as d:
                    out[prefix + "/" + k] = d.layout()
            else:
                with node.group(k) as g:
                    walk(g, prefix + "/" + k)

    with h5.open_file(path, "r") as f:
        walk(f, "")
    return out


# ---------------------------------------------------------------- the store scan of a resume
_CHUNK_RE = re.compile(r"^chunk-(\d{5})\.h5$")


def chunk_path(store, ordinal):
    """`<store>/chunk-<NNNNN>.h5` (run_resume.jl:5)."""
    return os.path.join(store, f"chunk-{int(ordinal):05d}.h5")


def chunk_files(store):
    return sorted(os.path.join(store, f) for f in os.listdir(store) if _CHUNK_RE.match(f))


def classify_store(store, chunk_hashes, panel_sha256):
    """Files of the store by kind (run_resume.jl:24-44): `keep` {ordinal: path} (complete, chunk hash and panel hash
    equal to the plan's, not a pilot), `halted` (kept but written by a control-bound stop) and `stale` (everything
    else, including unreadable files and ordinals the plan does not have). `chunk_hashes` maps each ordinal of the
    plan (1..M) to its chunk hash."""
    keep, halted, stale = {}, {}, []
    for f in chunk_files(store):
        ordn = int(_CHUNK_RE.match(os.path.basename(f)).group(1))
        if ordn not in chunk_hashes:
            stale.append(f)
            continue
        try:
            a = read_chunk_attrs(f)
        except (h5.H5Error, OSError):
            a = {}
        ok = (a.get("complete") is True and a.get("chunk_hash", "") == chunk_hashes[ordn]
              and a.get("pilot", False) is False and a.get("panel_sha256", "") == panel_sha256)
        if ok:
            (halted if a.get("control_stop", False) is True else keep)[ordn] = f
        else:
            stale.append(f)
    return keep, stale, halted


def refuse_halted(halted):
    """A control-bound stop is not a result: a plain resume refuses (run_resume.jl:47-54)."""
    if halted:
        names = ", ".join(os.path.basename(halted[o]) for o in sorted(halted))
        raise ValueError(f"chunk files {names} were written by a campaign stopped at a control bound; resuming repeats "
                         "the same code: change the code version, or ask for the retry to recompute them")


def move_aside(store, path, dirname):
    """Move a file to `store/dirname/`, numbered when the name is taken; never deleted (run_resume.jl:7-19)."""
    d = os.path.join(store, dirname)
    os.makedirs(d, exist_ok=True)
    dest = os.path.join(d, os.path.basename(path))
    k = 0
    while os.path.lexists(dest):
        k += 1
        dest = os.path.join(d, os.path.basename(path) + f".{k}")
    os.rename(path, dest)
    return dest


def remove_temporaries(store):
    """List and remove the temporary files of a killed run (run_resume.jl:124-131); returns their names."""
    tmps = sorted(f for f in os.listdir(store) if f.endswith(".tmp"))
    for f in tmps:
        os.remove(os.path.join(store, f))
    return tmps


def pending_chunks(store, chunk_hashes, panel_sha256, stale="error"):
    """The ordinals a resume still has to compute: temporaries removed, halted files refused, stale files raised
    (`stale="error"`) or moved to `stale/` (`"move"`, run_resume.jl:106-121). Returns (pending ordinals, kept
    {ordinal: path}, removed temporaries, moved stale files)."""
    removed = remove_temporaries(store)
    keep, st, halted = classify_store(store, chunk_hashes, panel_sha256)
    refuse_halted(halted)
    moved = []
    if st:
        if stale != "move":
            shown = ", ".join(os.path.basename(f) for f in st[:20])
            raise ValueError(f"stale chunk files (chunk hash differs from the plan: code version or an input changed): "
                             f"{shown}; ask for stale='move' to move them to stale/ and recompute")
        for f in st:
            move_aside(store, f, "stale")
            moved.append(os.path.basename(f))
    pending = sorted(o for o in chunk_hashes if o not in keep)
    return pending, keep, removed, moved


# ---------------------------------------------------------------- the orbit record of the device recorder
def diagnostic_names(method):
    """Diagnostics kept for a method (run_record.jl:55-61)."""
    if method == "L2":
        return []
    if method == "L1":
        return ["det", "symp"]
    if method == "L9":
        return ["det", "symp", "sum", "pair"]
    return ["det", "symp", "orth", "sum", "pair"]


def first_exceedance(v, thr, t):
    """The recorded step of the first value above the threshold, 0 when none (run_record.jl:204-207)."""
    idx = np.flatnonzero(np.asarray(v) > thr)
    return 0 if idx.size == 0 else int(np.asarray(t)[idx[0]])


def record_thresholds(records, j, *, tau, det_residual_tol, pairing_residual_tol, orthogonality_tol):
    """The diagnostics' thresholds of orbit j (run_record.jl:189-202) from the G_max and M_max of its last good row:
    det and symp at M = [M_max 0; 0 0]; for k = 2 the sum and pairing threshold; for the QR family the
    orthogonality threshold tol eps k g, g = G_w = G_max + 1/G_max for CGS and MGS and 1 otherwise
    (src/Lyapunov/orthonormalize.jl:149-151)."""
    from .reducers import det_threshold, symp_threshold, sum_threshold, EPS64
    from .recorder import QR_METHODS
    from .reducers import CGS, MGS
    method = records.method
    if method == "L2":
        return {}
    u = records.unpack(j)
    M = np.array([[u["mmax"], 0.0], [0.0, 0.0]])
    out = {"det": det_threshold(det_residual_tol, M), "symp": symp_threshold(det_residual_tol, M)}
    if records.k == 2:
        s = float(sum_threshold(pairing_residual_tol, u["gmax"], tau, u["mmax"]))
        out["sum"] = s
        out["pair"] = s
        if method in QR_METHODS:
            g = u["gmax"] + 1.0 / u["gmax"] if QR_METHODS[method] in (CGS, MGS) else 1.0
            out["orth"] = ((float(orthogonality_tol) * EPS64) * 2.0) * g
    return out


def orbit_record(records, j, *, orbit_id, attrs, thresholds=None, cost=None):
    """The OrbitRecord of orbit j of a device recording (`recorder.Records`), as run_record.jl:213-225 assembles it
    from the record buffer: `t`, `state`, `winding` (wrapped mode), `lyap[method]`, the method's diagnostics with
    `thresholds` {name: value} and their first exceedance steps, `cost[method]` (default: horizon, vectors, 0, 0,
    0), the failure row blanked (t = 0, NaN values) as `Records.unpack` leaves it. `attrs` are the orbit attributes
    (system, params, ic, class_label, seed, item_hash, design_class, flags, steps_recorded)."""
    u = records.unpack(j)
    method = records.method
    names = diagnostic_names(method)
    thresholds = {} if thresholds is None else dict(thresholds)
    missing = [nm for nm in names if nm not in thresholds]
    if missing:
        raise ValueError(f"orbit_record: no threshold for {', '.join(missing)}")
    diag = {nm: np.array(u[nm], dtype=np.float64) for nm in names}
    fe = {nm: first_exceedance(diag[nm], thresholds[nm], u["t"]) for nm in names}
    if cost is None:
        cost = np.array([float(records.grid[-1]), float(records.k), 0.0, 0.0, 0.0])
    return OrbitRecord(id=int(orbit_id), attrs=dict(attrs), t=np.array(u["t"], dtype=np.int64),
                       state=np.array(u["state"], dtype=np.float64),
                       winding=None if u["winding"] is None else np.array(u["winding"], dtype=np.int64),
                       lyap={method: np.array(u["lyap"], dtype=np.float64)}, diagnostics=diag,
                       thresholds={nm: float(thresholds[nm]) for nm in names}, first_exceedance=fe,
                       cost={method: np.asarray(cost, dtype=np.float64)})
