"""ctypes binding of oracle/liboracle.so for the tests and the CPU-baseline leg of bench.py.

The oracle is the checker: a CPU restatement of the reference path (see oracle/oracle.cpp)."""
import ctypes
import os

import numpy as np

from chaos_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "liboracle.so")
MATH_LIBM, MATH_CL = 0, 1
_o = None


def load():
    global _o
    if _o is None:
        if not os.path.exists(PATH):
            raise ImportError(f"{PATH} missing: run `make -C oracle`")
        o = ctypes.CDLL(PATH)
        P, I64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        o.or_state_bytes.restype = I64
        o.or_state_bytes.argtypes = [ctypes.POINTER(L.cl_chain), I, I]
        o.or_propagate.restype = I
        o.or_propagate.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I, I64, I,
                                   P, P, P, P, P, P, I, I]
        o.or_propagate_elt.restype = I
        o.or_propagate_elt.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I, I64, I,
                                       P, P, P, P, P, P, I, I, I]
        o.or_propagate_quad_elt.restype = I
        o.or_propagate_quad_elt.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I, I64,
                                            I, P, P, P, P, P, P, I, I]
        o.or_propagate_quad.restype = I
        o.or_propagate_quad.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I, I64,
                                        I, P, P, P, P, P, P, I]
        o.or_propagate_linear.restype = I
        o.or_propagate_linear.argtypes = [P, ctypes.POINTER(L.cl_chain), I64, I, I64, I, P, P, P, P, P, I]
        o.or_propagate_linear_d.restype = I
        o.or_propagate_linear_d.argtypes = [P, I, ctypes.POINTER(L.cl_chain), I64, I, I64, I, P, P, P, P, P, I]
        o.or_plant_defect.restype = None
        o.or_plant_defect.argtypes = [I]
        o.or_initial_deviations.restype = I
        o.or_initial_deviations.argtypes = [ctypes.c_uint64, I64, I64, I, I, P]
        o.or_propagate_chain.restype = I
        o.or_propagate_chain.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, P, I64, I,
                                         P, P, P, P, P, I, I, I, I]
        o.or_same_orbit_quad.restype = I
        o.or_same_orbit_quad.argtypes = [ctypes.c_double, ctypes.POINTER(L.cl_chain), I64, I, I64, I, P, P, P, I]
        o.or_gauss_iterations.restype = None
        o.or_gauss_iterations.argtypes = [P, P]
        o.or_record.restype = I
        o.or_record.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I64, I, P, I, P, P,
                                P, P, P, P, P, I, I]
        o.or_ind_record.restype = I
        o.or_ind_record.argtypes = [ctypes.POINTER(L.cl_system), ctypes.POINTER(L.cl_chain), I64, I, I, I64, I, P, I,
                                    P, P, P, P, P, P, P, I, I]
        o.or_shadow.restype = I
        o.or_shadow.argtypes = [ctypes.POINTER(L.cl_system), P, I64, I, I64, I64, I64, I, ctypes.c_double, P, P, P, I, I]
        o.or_shadow_record.restype = I
        o.or_shadow_record.argtypes = [ctypes.POINTER(L.cl_system), I64, I, I64, I, ctypes.c_double, P, I, P, P, P, P,
                                       P, I, I]
        o.or_propagate_ws.restype = I
        o.or_propagate_ws.argtypes = [ctypes.POINTER(L.cl_system), P, ctypes.POINTER(L.cl_chain), I64, I, I, I, I64, I,
                                      P, P, P, P, P, P, I, I]
        o.or_wb_weights.restype = None
        o.or_wb_weights.argtypes = [I64, I, P, I]
        o.or_exp.restype = None
        o.or_exp.argtypes = [I64, P, P]
        o.or_atan2.restype = None
        o.or_atan2.argtypes = [I64, P, P, P]
        o.or_window_gw2.restype = None
        o.or_window_gw2.argtypes = [I64, P, P]
        o.or_csum.restype = None
        o.or_csum.argtypes = [I64, ctypes.c_double, P, P]
        o.or_csum_list.restype = None
        o.or_csum_list.argtypes = [I64, P, P]
        _o = o
    return _o


def _ptr(a):
    return None if a is None else a.ctypes.data


def run(sysd, chain, x0, W0, N, tau, k, W2=None, math=MATH_CL, nthreads=None, quad=False, elt=L.CL_ELT_F64):
    """Full N-step runs. x0 (m, s), W0 (m, d, k) as numpy; returns dict of x, W, W2, raw states, code, hstep."""
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m, s = x.shape
    d = 4 if sysd.kind in (L.CL_SYS_HENON_HEILES_GAUSS, L.CL_SYS_PAIR_STANDARD_MAP,
                           L.CL_SYS_PAIR_WRAPPED_STANDARD_MAP) else 2
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    W2a = None if W2 is None else np.array(np.asarray(W2, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    nb = o.or_state_bytes(ctypes.byref(chain), d, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    if quad and elt != L.CL_ELT_F64:
        rc = o.or_propagate_quad_elt(ctypes.byref(sysd), ctypes.byref(chain), m, s, d, k, N, tau, _ptr(x), _ptr(W),
                                     _ptr(W2a), _ptr(states), _ptr(code), _ptr(hstep), nt, elt)
    elif quad:
        rc = o.or_propagate_quad(ctypes.byref(sysd), ctypes.byref(chain), m, s, d, k, N, tau, _ptr(x), _ptr(W),
                                 _ptr(W2a), _ptr(states), _ptr(code), _ptr(hstep), nt)
    elif elt != L.CL_ELT_F64:
        rc = o.or_propagate_elt(ctypes.byref(sysd), ctypes.byref(chain), m, s, d, k, N, tau, _ptr(x), _ptr(W),
                                _ptr(W2a), _ptr(states), _ptr(code), _ptr(hstep), math, nt, elt)
    else:
        rc = o.or_propagate(ctypes.byref(sysd), ctypes.byref(chain), m, s, d, k, N, tau, _ptr(x), _ptr(W), _ptr(W2a),
                            _ptr(states), _ptr(code), _ptr(hstep), math, nt)
    assert rc == 0, rc
    out = dict(x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)
    if W2a is not None:
        out["W2"] = W2a.transpose(0, 2, 1).copy()
    return out


RED_ROWS, RED_SITES, RED_BLOCKED = 0, 1, 2


def run_chain(sysd, chain, x0, W0, N, tau, k, Ns=None, math=MATH_CL, mode=RED_SITES, nthreads=None, elt=1):
    """Chain runs from fresh states: x0 (m, 2N), W0 (m, 2N, k); Ns optional per-orbit step counts; elt 1 Float64,
    2 Float32 tangents."""
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m, d = x.shape
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    nb = 16 * k + 72 if chain.n else 0
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    Na = None if Ns is None else np.ascontiguousarray(np.asarray(Ns, dtype=np.int64))
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_propagate_chain(ctypes.byref(sysd), ctypes.byref(chain), m, k, _ptr(Na), int(N), tau, _ptr(x), _ptr(W),
                              _ptr(states), _ptr(code), _ptr(hstep), math, mode, nt, int(elt))
    assert rc == 0, rc
    return dict(x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)


def same_orbit_quad(K, chain, x0, W0, N, tau, k, nthreads=None):
    """lambda (m, k) of the standard map's QR spectrum with the FP64 orbit and __float128 tangents and QR."""
    o = load()
    x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m = x.shape[0]
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, 2, k).transpose(0, 2, 1), order="C", copy=True)
    lam = np.zeros((m, k))
    nt = nthreads or min(16, os.cpu_count() or 1)
    assert o.or_same_orbit_quad(float(K), ctypes.byref(chain), m, k, int(N), tau, _ptr(x), _ptr(W), _ptr(lam), nt) == 0
    return lam


def gauss_iterations():
    """(iterations, steps) of the Gauss stage solver summed over oracle runs since the last call."""
    it, st = ctypes.c_int64(0), ctypes.c_int64(0)
    load().or_gauss_iterations(ctypes.byref(it), ctypes.byref(st))
    return it.value, st.value


def run_linear(A, chain, x0, W0, N, tau, k, quad=False):
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m = x.shape[0]
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, 2, k).transpose(0, 2, 1), order="C", copy=True)
    Ac = np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.ravel())  # column-major
    nb = o.or_state_bytes(ctypes.byref(chain), 2, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    rc = o.or_propagate_linear(_ptr(Ac), ctypes.byref(chain), m, k, N, tau, _ptr(x), _ptr(W), _ptr(states), _ptr(code),
                               _ptr(hstep), int(quad))
    assert rc == 0, rc
    return dict(x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)


def run_linear_d(A, chain, x0, W0, N, tau, k, quad=False):
    """The linear map x' = A x in d = A.shape[0] dimensions (2..8): x0 (m, d), W0 (m, d, k). Returns x, W (m, d, k),
    raw states, code and hstep."""
    o = load()
    A = np.asarray(A, dtype=np.float64)
    d = A.shape[0]
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64).reshape(-1, d))
    m = x.shape[0]
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    Ac = np.ascontiguousarray(A.T.ravel())  # column-major
    nb = o.or_state_bytes(ctypes.byref(chain), d, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    rc = o.or_propagate_linear_d(_ptr(Ac), d, ctypes.byref(chain), m, k, N, tau, _ptr(x), _ptr(W), _ptr(states),
                                 _ptr(code), _ptr(hstep), int(quad))
    assert rc == 0, rc
    return dict(x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)


class planted:
    """Context manager: the oracle runs with a planted defect (oracle.cpp g_defect) inside the block."""
    GALI_RAW, OFLI_NOPROJ, MEGNO_Q2, GALI_LASTFACTOR, BLK_TNOTRANS, BLK_TSIGN = 1, 2, 3, 4, 5, 6

    def __init__(self, defect):
        self.defect = defect

    def __enter__(self):
        load().or_plant_defect(self.defect)
        return self

    def __exit__(self, *exc):
        load().or_plant_defect(0)
        return False


def frames(seed, orbit0, m, d, k):
    o = load()
    out = np.zeros((m, k, d), dtype=np.float64)
    assert o.or_initial_deviations(seed & (2 ** 64 - 1), orbit0, m, d, k, out.ctypes.data) == 0
    return out.transpose(0, 2, 1).copy()


def ind_record(sysd, chain, x0, W0, N, tau, k, grid, ncol, math=MATH_CL, nthreads=None):
    """The indicator recorder (oracle/oracle.cpp, or_ind_record): a series kind's chain (members, RunMonitor; k = 2)
    or a window kind's (RunMonitor, WeightedBirkhoff; k = 1, reset-frame flag); returns R (m, G, ncol), fail, x, W,
    raw states, code, hstep."""
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m, s = x.shape
    d = 4 if sysd.kind in (L.CL_SYS_PAIR_STANDARD_MAP, L.CL_SYS_PAIR_WRAPPED_STANDARD_MAP) else 2
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.int64))
    G = len(g)
    R = np.zeros((m, G, ncol))
    fail = np.zeros(m, dtype=np.int64)
    nb = o.or_state_bytes(ctypes.byref(chain), d, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_ind_record(ctypes.byref(sysd), ctypes.byref(chain), m, s, k, int(N), tau, _ptr(g), G, _ptr(x),
                         _ptr(W), _ptr(R), _ptr(fail), _ptr(code), _ptr(hstep), _ptr(states), math, nt)
    assert rc == 0, rc
    return dict(R=R, fail=fail, x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)


def record(sysd, chain, x0, W0, N, tau, k, grid, math=MATH_CL, nthreads=None):
    """The recorder (oracle/oracle.cpp, or_record): x0 (m, s), W0 (m, 2, k), grid ending at N; returns R (m, G,
    k + 7 + s), fail, x, W, raw states, code, hstep."""
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m, s = x.shape
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, 2, k).transpose(0, 2, 1), order="C", copy=True)
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.int64))
    G = len(g)
    R = np.zeros((m, G, k + 7 + s))
    fail = np.zeros(m, dtype=np.int64)
    nb = o.or_state_bytes(ctypes.byref(chain), 2, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_record(ctypes.byref(sysd), ctypes.byref(chain), m, s, k, int(N), tau, _ptr(g), G, _ptr(x), _ptr(W),
                     _ptr(R), _ptr(fail), _ptr(code), _ptr(hstep), _ptr(states), math, nt)
    assert rc == 0, rc
    return dict(R=R, fail=fail, x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep)


def shadow_ends(raw, s):
    """End records (m, 2s + 6) uint64 -> dict of x, y, sum, comp, dprev, code, step, value."""
    f = raw.view(np.float64)
    return dict(x=f[:, :s].copy(), y=f[:, s:2 * s].copy(), sum=f[:, 2 * s].copy(), comp=f[:, 2 * s + 1].copy(),
                dprev=f[:, 2 * s + 2].copy(), code=(raw[:, 2 * s + 3] & 0xff).astype(np.uint8).view(np.int8),
                step=raw[:, 2 * s + 4].view(np.int64).copy(), value=f[:, 2 * s + 5].copy(), raw=raw)


def shadow(sysd, x0, u0, N, tau, d0, *, n0=0, steps=None, ends=None, A=None, math=MATH_CL, nthreads=None):
    """The two-trajectory estimator (oracle/oracle.cpp, or_shadow): one launch of `steps` (default N - n0) from
    global step n0 over x0 (m, s) and unit directions u0 (m, 2), continuing `ends` (m, 2s + 6 words) when n0 > 0;
    with A (2 x 2) the linear map x' = A x. Returns the decoded end records."""
    o = load()
    x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x.shape
    u = np.ascontiguousarray(np.asarray(u0, dtype=np.float64))
    e = np.zeros((m, 2 * s + 6), dtype=np.uint64) if ends is None else np.array(ends, dtype=np.uint64, copy=True)
    Ac = None if A is None else np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.ravel())
    st = int(N) - n0 if steps is None else int(steps)
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_shadow(ctypes.byref(sysd) if sysd is not None else None, _ptr(Ac), m, s, n0, st, int(N), tau, float(d0),
                     _ptr(x), _ptr(u), _ptr(e), math, nt)
    assert rc == 0, rc
    return shadow_ends(e, s)


def shadow_record(sysd, x0, u0, N, tau, d0, grid, math=MATH_CL, nthreads=None):
    """The recorder's two-trajectory rows (or_shadow_record): R (m, G, 8 + s), fail and the end records."""
    o = load()
    x = np.ascontiguousarray(np.asarray(x0, dtype=np.float64))
    m, s = x.shape
    u = np.ascontiguousarray(np.asarray(u0, dtype=np.float64))
    g = np.ascontiguousarray(np.asarray(grid, dtype=np.int64))
    G = len(g)
    R = np.zeros((m, G, 8 + s))
    fail = np.zeros(m, dtype=np.int64)
    e = np.zeros((m, 2 * s + 6), dtype=np.uint64)
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_shadow_record(ctypes.byref(sysd), m, s, int(N), tau, float(d0), _ptr(g), G, _ptr(x), _ptr(u), _ptr(R),
                            _ptr(fail), _ptr(e), math, nt)
    assert rc == 0, rc
    return dict(R=R, fail=fail, ends=shadow_ends(e, s))


def run_ws(sysd, chain, nws, x0, W0, N, tau, k, A=None, math=MATH_CL, nthreads=None):
    """Full N-step runs of a chain with a workspace reducer (or_propagate_ws): x0 (m, s), W0 (m, d, k); with A (2 x 2)
    the linear map x' = A x. Returns x, W, raw states, code, hstep and the workspace rows ws (m, nws) UInt32."""
    o = load()
    x = np.ascontiguousarray(np.array(x0, dtype=np.float64))
    m, s = x.shape
    d = 2
    W = np.array(np.asarray(W0, dtype=np.float64).reshape(m, d, k).transpose(0, 2, 1), order="C", copy=True)
    nb = o.or_state_bytes(ctypes.byref(chain), d, k)
    states = np.zeros((m, max(nb, 1)), dtype=np.uint8)
    code = np.zeros(m, dtype=np.int8)
    hstep = np.zeros(m, dtype=np.int64)
    ws = np.full((m, nws), 0xDEADBEEF, dtype=np.uint32)
    Ac = None if A is None else np.ascontiguousarray(np.asarray(A, dtype=np.float64).T.ravel())
    nt = nthreads or min(16, os.cpu_count() or 1)
    rc = o.or_propagate_ws(None if A is not None else ctypes.byref(sysd), _ptr(Ac), ctypes.byref(chain), m, s, d, k,
                           int(N), tau, _ptr(x), _ptr(W), _ptr(states), _ptr(code), _ptr(hstep), _ptr(ws), math, nt)
    assert rc == 0, rc
    return dict(x=x, W=W.transpose(0, 2, 1).copy(), states=states[:, :nb], code=code, hstep=hstep, ws=ws)


def wb_weights(N, order, math=MATH_CL):
    w = np.empty(int(N))
    load().or_wb_weights(int(N), int(order), w.ctypes.data, math)
    return w


def exp(x):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float64))
    out = np.empty_like(x)
    load().or_exp(x.size, x.ctypes.data, out.ctypes.data)
    return out


def window_gw2(R):
    """window_gw of 2 x 2 upper-triangular factors R[n, 4] (column-major)."""
    R = np.ascontiguousarray(R, dtype=np.float64)
    out = np.empty(R.shape[0])
    load().or_window_gw2(R.shape[0], R.ctypes.data, out.ctypes.data)
    return out


def atan2(y, x):
    """The kernels' portable atan2 (cl_math.h) on the host."""
    y = np.ascontiguousarray(y, dtype=np.float64)
    x = np.ascontiguousarray(np.broadcast_to(x, y.shape), dtype=np.float64)
    out = np.empty_like(y)
    load().or_atan2(y.size, y.ctypes.data, x.ctypes.data, out.ctypes.data)
    return out

