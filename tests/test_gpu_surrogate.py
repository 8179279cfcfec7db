"""The surrogate triage pipeline on the GPU: the traced prefix, the continuation from the prefix state, and a
small end-to-end triage against full propagation."""
import sys

import numpy as np
import pytest

from chaos_amd import StandardMapK, MEGNO, FLI, QRSpectrum, MGS, propagate
from chaos_amd.tangent import make_chain
from tests import oracle_lib as O

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_prefix_states_and_trace():
    """The traced prefix pass equals the oracle's two-frame run bit for bit, and its trace at handoff n holds
    MEGNO's Y and FLI_B of the run stopped at n (as FP32)."""
    import torch
    from chaos_amd.surrogate import Ensemble, make_starts
    m = 2048
    x0 = make_starts(7, m)
    say("prefix", m)
    ens = Ensemble(x0, seed=7)
    tr = ens.prefix(256, 1000).cpu().numpy()
    ch, allr = make_chain((MEGNO(replace=False), FLI()), (QRSpectrum(MGS()),))
    W0 = O.frames(7, 1, m, 2, 2)
    for n in (1, 37, 256):
        orc = O.run(StandardMapK().descriptor(), ch, x0, W0, n, 1, 2, W2=W0, math=O.MATH_CL)
        from chaos_amd.tangent import decode_states, finals_of
        fin = finals_of(decode_states(orc["states"], allr, 2, 2), allr, 2, 2)
        Y = fin[0]["Y"][:, 0].astype(np.float32)
        fB = fin[1]["fli"].astype(np.float32)
        assert np.array_equal(tr[3 * (n - 1)], Y)
        assert np.array_equal(tr[3 * (n - 1) + 1], fB)
        if n == 256:
            raw = ens.st.cpu().numpy().T.copy().view(np.uint8).reshape(m, -1)
            assert np.array_equal(raw, orc["states"])
            assert np.array_equal(ens.x.cpu().numpy().T, orc["x"])
    # the log stretch of the first handoff: log of the first column's norm after one step
    orc1 = O.run(StandardMapK().descriptor(), make_chain(())[0], x0, W0, 1, 1, 2, math=O.MATH_CL)
    a1 = np.log(np.linalg.norm(orc1["W"][:, :, 0], axis=1))
    assert np.max(np.abs(tr[2] - a1)) < 1e-6


def test_continuation_equals_plain_run():
    """prefix (256 steps) + QR continuation to N equals one plain QRSpectrum(MGS()) run of N steps, bitwise."""
    import torch
    from chaos_amd.surrogate import Ensemble, make_starts, continue_qr, qr_lambda1
    m, N = 4096, 3000
    x0 = make_starts(9, m)
    say("continuation", m, N)
    ens = Ensemble(x0, seed=9)
    ens.prefix(256, N)
    w0 = ens.w_ind
    st = ens.st[w0:w0 + 13]
    continue_qr(ens.x, ens.W2, st, ens.code, ens.h, 256, N, N - 256)
    plain = propagate(StandardMapK(), x0, N, k=2, seed=9, reducers=(QRSpectrum(MGS()),), check=False, keep_states=True)
    raw = st.cpu().numpy().T.copy().view(np.uint8).reshape(m, -1)
    assert np.array_equal(raw, plain.states)
    assert np.array_equal(qr_lambda1(st).cpu().numpy(), plain.finals[0]["lambda"][:, 0])


def test_small_triage_gains_over_full_propagation():
    """Train on 32768 orbits (N = 4000), triage 8192 held-out orbits: most orbits skip the full run and the
    triaged labels agree with full propagation on at least 98% (the full-size workload reports its own figures)."""
    import torch
    from chaos_amd.surrogate import make_starts, run_full, fit, evaluate
    N, prefix = 4000, 256
    say("train set")
    tr_set = run_full(make_starts(1, 32768), 1, N, prefix)
    say("held-out set")
    ho = run_full(make_starts(2, 8192, orbit0=10 ** 7), 2, N, prefix)
    net, info = fit(tr_set, epochs=6)
    res = evaluate(net, ho, margin=2.0)
    say("triage", res)
    assert res["skipped_fraction"] > 0.5
    assert res["misclassified"] < 0.02
