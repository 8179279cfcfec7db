"""The autotuner on the GPU: every candidate setting of a two-frame map chain and of a small QR ensemble gives the
built-in rule's bits, the tuner's table covers the candidates and it leaves its choice on the context; a tau it may
choose is one the caller allowed."""
import sys

import numpy as np
import pytest

from chaos_amd import StandardMap, propagate, autotune
from chaos_amd.reducers import QRSpectrum, HouseholderQR, MGS, MEGNO, FLI, OFLI, SALI
from chaos_amd.inputs import grid_centres

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def test_every_setting_same_bits():
    say("autotune bits")
    red = (MEGNO(replace=False), FLI(replace=False), OFLI(replace=False), SALI())
    red2 = (QRSpectrum(HouseholderQR()),)
    x0 = grid_centres(48)
    autotune.set_tuning()
    ref = propagate(StandardMap(2.0), x0, 300, k=2, tau=1, seed=1, reducers=red, reducers2=red2, keep_states=True)
    try:
        for block in (64, 128, 256):
            for split in (1, 2):
                autotune.set_tuning(block, split, 0)
                g = propagate(StandardMap(2.0), x0, 300, k=2, tau=1, seed=1, reducers=red, reducers2=red2,
                              keep_states=True)
                assert np.array_equal(g.states, ref.states) and np.array_equal(g.W2, ref.W2), (block, split)
        x1 = grid_centres(40)
        autotune.set_tuning()
        r1 = propagate(StandardMap(6.0), x1, 500, k=2, tau=1, seed=2, reducers=(QRSpectrum(MGS()),), keep_states=True)
        for pipe in (1, 2):
            autotune.set_tuning(64, 0, pipe)
            g = propagate(StandardMap(6.0), x1, 500, k=2, tau=1, seed=2, reducers=(QRSpectrum(MGS()),),
                          keep_states=True)
            assert np.array_equal(g.states, r1.states), pipe
    finally:
        autotune.set_tuning()


def test_tune_picks_and_sets():
    say("autotune tune")
    x1 = grid_centres(64)
    out = autotune.tune(StandardMap(6.0), x1, k=2, tau=1, reducers=(QRSpectrum(MGS()),), trial_steps=400,
                        tau_allowed=(2, 4))
    try:
        assert len(out["table"]) == 3 * 2 * 3 + 1
        assert out["best"]["tau"] in (1, 2, 4)
        cur = autotune.get_tuning()
        assert (cur["block"], cur["split"], cur["pipe"]) == (out["best"]["block"], out["best"]["split"],
                                                              out["best"]["pipe"])
        assert out["gain"] >= 1.0 or out["best"]["ns_per_orbit_step"] <= out["default"]["ns_per_orbit_step"] * 1.5
    finally:
        autotune.set_tuning()
