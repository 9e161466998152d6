"""B_forced vectors against the ones the reference's own forcing branch
(bin/assemble.py:77-427) writes for the same parameters
(tests/golden/make_forcing_fixtures.py), and the A(w) identity of forced
problems."""
import glob
import json
import os

import numpy as np
import pytest
import scipy.sparse as ss

from conftest import GOLDEN

FIXTURES = sorted(glob.glob(os.path.join(GOLDEN, "forcing_*.npz")))


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p)[8:-4] for p in FIXTURES])
def test_forcing_vector_matches_reference(path):
    from kore_amd.forcing import forcing_vector
    from kore_amd.params import Params
    d = np.load(path)
    meta = json.loads(str(d["meta"]))
    ref = ss.csr_matrix((d["data"], d["indices"], d["indptr"]), shape=tuple(d["shape"])).toarray().ravel()
    par = Params(m=meta["m"], symm=meta["symm"], ricb=meta["ricb"], Ek=meta["Ek"], N=meta["N"], lmax=meta["lmax"],
                 bci=meta["bci"], bco=meta["bco"], forcing=meta["forcing"],
                 forcing_frequency=meta["forcing_frequency"], forcing_amplitude=meta["forcing_amplitude"],
                 forcing_amplitude_cmb=meta["forcing_amplitude_cmb"],
                 forcing_amplitude_icb=meta["forcing_amplitude_icb"])
    assert par.sizmat == meta["sizmat"]
    got = forcing_vector(par).toarray().ravel()
    assert np.array_equal(np.nonzero(got)[0], np.nonzero(ref)[0])
    assert np.abs(got - ref).max() <= 1e-15 * np.abs(ref).max()


@pytest.mark.parametrize("f", [1, 4, 5, 8])
def test_unbuildable_forcings_are_refused(f):
    from kore_amd.forcing import forcing_vector
    from kore_amd.params import Params
    with pytest.raises(NotImplementedError, match="never writes"):
        forcing_vector(Params(m=2, symm=1, forcing=f, N=16, lmax=17))


def test_tide_forcings_need_their_amplitude():
    from kore_amd.forcing import forcing_vector
    from kore_amd.params import Params
    with pytest.raises(ValueError, match="forcing_amplitude"):
        forcing_vector(Params(m=1, symm=-1, forcing=6, N=16, lmax=18))
