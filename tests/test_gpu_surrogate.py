"""GPU: what the trained surrogate gains over the method without it.

The search for 3 dB couplers at 193 THz is run at reduced scale (768
training spectra from the batched FDTD, 4000 steps, 8192 screened designs,
32 checked by FDTD).  Without the surrogate a candidate is accepted at the rate of
random designs; with it, the designs the FDTD runs confirm among the
surrogate's picks must be at least 3x as frequent, and the network's
|S21| at the picks must be within 0.15 of the FDTD value on average.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_surrogate_picks_beat_random_designs(mi355x):
    from prismo_amd.surrogate import loop

    r = loop.run(n_train=768, n_screen=8192, top_k=32, epochs=200,
                 log=lambda m: None)
    assert r["accepted"] >= 4, r
    assert r["accept_rate_surrogate_picks"] >= 3 * max(r["accept_rate_random_designs"], 1 / 768), r
    assert r["s21_abs_err_mean"] <= 0.15, r
