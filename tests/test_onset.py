"""Onset driver: A depends on the Rayleigh number only through the buoyancy
force, linearly, so one resident system per m serves every Brent step
(CPU check of that premise on the assembled operators)."""
import dataclasses

import numpy as np

from conftest import params_dormy, params_jones


def test_A_is_affine_in_Ra_and_B_does_not_change():
    from kore_amd import assembly
    from kore_amd.onset import _on_pattern
    for mk, ras in ((params_dormy, (1.6e6, 3.2e6, 1.65e6)), (params_jones, (4.6e6, 9.2e6, 4.7e6))):
        par = dataclasses.replace(mk(1.0), N=24, lmax=32 if mk is params_dormy else 32)
        mats = [assembly.assemble(dataclasses.replace(par, Ra_gap=ra, Beyonce=None)) for ra in ras]
        (A1, B1, _), (A2, B2, _), (A3, B3, _) = mats
        assert abs(B1 - B2).max() == 0 and abs(B1 - B3).max() == 0
        pat = (abs(A1) + abs(A2) + abs(A3)).tocsr()
        a1, a2, a3 = (_on_pattern(pat, A) for A in (A1, A2, A3))
        pred = a1 + (ras[2] - ras[0]) * (a2 - a1) / (ras[1] - ras[0])
        assert np.abs(pred - a3).max() <= 1e-13 * np.abs(a3).max()
        # the change is confined to the buoyancy coupling: a small part of A
        changed = np.abs(a2 - a1) > 0
        assert 0 < changed.sum() < 0.2 * len(a1)
