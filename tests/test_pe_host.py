"""Host-side primitive-equation logic (no GPU): sigma-grid validation and
metadata against the oracle's restatement, state layout and rest state."""
import math

import numpy as np
import pytest

from oracle import pe


@pytest.mark.parametrize("bad", [(0.0,), (0.1, 0.5, 1.0), (0.0, 0.5, 0.999), (0.0, 0.6, 0.4, 1.0),
                                 (0.0, 0.5, 0.5, 1.0), (0.0, math.nan, 1.0), ("a", 1.0)])
def test_sigma_grid_rejects_invalid_interfaces(bad):
    from aeolus_amd.physics import SigmaGrid, SigmaGridError
    with pytest.raises(SigmaGridError):
        SigmaGrid(bad)


@pytest.mark.parametrize("nlev", [0, -1, 2.0, True])
def test_sigma_grid_uniform_rejects_bad_counts(nlev):
    from aeolus_amd.physics import SigmaGrid, SigmaGridError
    with pytest.raises(SigmaGridError):
        SigmaGrid.uniform(nlev)


@pytest.mark.parametrize("interfaces", [(0.0, 0.05, 0.15, 0.30, 0.50, 0.75, 1.0),
                                        tuple(k / 8 for k in range(9)), (0.0, 1.0)])
def test_sigma_grid_metadata_matches_oracle(interfaces):
    from aeolus_amd.physics import SigmaGrid
    g, o = SigmaGrid(interfaces), pe.Sigma(interfaces)
    assert g.nlev == o.nlev
    assert g.full_levels == o.full_levels and g.thickness == o.thickness
    assert g.interface_log_ratios == o.log_ratios and g.alpha == o.alpha
    assert g.alpha[0] == math.log(2.0) and g.interface_log_ratios[0] == math.inf
    assert np.array_equal(g.interfaces_array(), np.asarray(interfaces, dtype=np.float64))
    with pytest.raises(Exception):
        g.interfaces = (0.0, 1.0)            # frozen


def test_state_layout_and_rest_state():
    import torch
    from aeolus_amd.physics import (PrimitiveEquationsState, PrimitiveEquationsStateError,
                                    isothermal_rest_state)
    st = isothermal_rest_state(10, 3, temperature=250.0, surface_pressure=1.0e5, device="cpu")
    ref = pe.isothermal_rest_state(10, 3, 250.0, 1.0e5)
    assert st.nlev == 3 and tuple(st.coeffs.shape) == (10, 11, 11)
    assert np.array_equal(st.coeffs.numpy(), ref)
    assert st.ln_ps.shape == (11, 11) and st.temperature.shape == (3, 11, 11)
    z = torch.zeros((2, 11, 11), dtype=torch.complex128)
    st2 = PrimitiveEquationsState.from_fields(z, z + 1, z + 2, torch.ones((11, 11)))
    assert st2.nlev == 2 and float(st2.delta[1, 0, 0].real) == 1.0
    for shape in [(3, 11, 11), (5, 11, 11), (10, 11, 12), (11, 11)]:
        with pytest.raises(PrimitiveEquationsStateError):
            PrimitiveEquationsState(torch.zeros(shape, dtype=torch.complex128))
    for kw in (dict(temperature=0.0, surface_pressure=1e5),
               dict(temperature=250.0, surface_pressure=-1.0),
               dict(temperature=math.inf, surface_pressure=1e5)):
        with pytest.raises(ValueError):
            isothermal_rest_state(10, 3, **kw)
