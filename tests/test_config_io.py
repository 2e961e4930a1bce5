"""The HDF5 reader (dfn_h5_*) on the reference's own input files, when the
reference tree is present: values against the script that wrote them
(wrapper_python/BP6QD_A.py) and against the workload builder that restates it."""
import os

import numpy as np
import pytest

from demystification_dfn_amd import _abi, problems
from demystification_dfn_amd.device import LIB_PATH

REF = "/root/reference/problems"
pytestmark = [pytest.mark.skipif(not os.path.isdir(REF), reason="reference input files absent"),
              pytest.mark.skipif(not os.path.exists(LIB_PATH), reason="device library not built")]


def test_bp6_config_matches_the_generator():
    from demystification_dfn_amd.config_io import read_config
    p = read_config(os.path.join(REF, "BP6QD_A_4000", "config"))
    assert p.name == "BP6QD_A_4000" and p.N == 4000 and p.nb_fault == 1
    assert p.fracture_mode == _abi.MODE_III and p.static_kernel == _abi.KERNEL_HMATRIX
    assert p.tol_solver == 1e-6 and p.meta["max_it"] == 30000
    ref = problems.single_fault(static_kernel=_abi.KERNEL_HMATRIX, tol=1e-6)
    np.testing.assert_allclose(p.node, ref.node, rtol=0, atol=1e-9)
    for nm in ("a", "b", "Dc", "f0", "V0", "sigmaN", "V", "permeability"):
        np.testing.assert_array_equal(p.arrays[nm], ref.arrays[nm], err_msg=nm)
    np.testing.assert_allclose(p.arrays["theta"], ref.arrays["theta"], rtol=1e-12)
    assert p.mu == 32.04e9 and p.cs == 3.464e3
    assert len(p.sources) == 1
    f, idx, Q, tb, te = p.sources[0]
    assert (f, Q, tb, te) == (1, 1.25e-6, 0.0, 100 * 86400.0)
    assert abs(idx - ref.sources[0][1]) <= 1


def test_network_config_layout():
    from demystification_dfn_amd.config_io import read_config
    p = read_config(os.path.join(REF, "network_rate_strengthening", "config"))
    assert p.nb_fault == 10 and p.fracture_mode == _abi.MODE_II
    assert p.N == int(np.sum(p.fault_nb_element + 1) - 1)
    fake = p.fake_indices()
    assert np.all(p.arrays["a"][fake] == 1.0) and np.all(p.arrays["V0"][fake] == 1.0)
    assert np.all(np.isfinite(p.node))
    for f, idx, Q, tb, te in p.sources:  # every source sits on an element of its fault
        assert 1 <= f <= p.nb_fault and 1 <= idx <= p.fault_nb_element[f - 1] and tb <= te


def test_missing_dataset_is_an_argument_error(tmp_path):
    from demystification_dfn_amd.config_io import H5File
    with H5File(os.path.join(REF, "BP6QD_A_4000", "config")) as f:
        assert not f.has("fault2/a")
        with pytest.raises(_abi.DfnError):
            f.f64("no/such/dataset")
    bad = tmp_path / "not_hdf5"
    bad.write_bytes(b"plain text, not an HDF5 file" * 8)
    with pytest.raises(_abi.DfnError):
        H5File(str(bad))
