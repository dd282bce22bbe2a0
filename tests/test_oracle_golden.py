"""The CPU oracle against the reference's own step/bc outputs.

The fixtures (tests/golden/step_bc_*.npz, written by make_golden.py) hold the
reference kernels' results on two tiny grids with every branch active: bed,
bank, hull blocks at the lid, a lone obstacle, solid cells in the inlet and
outlet rows, random wall friction and uref, forces in six cells, and random
previous-step macros so the wall function engages. The reference ran as
float32 numpy code with some float64 terms, so agreement is at float32
rounding level; the sensitivity tests show that a plausible mistake moves the
result far outside the tolerance.
"""
import numpy as np
import pytest

from tests.helpers import oracle_raw_steps, step_fixture

TOL_F = 1e-6      # populations, ~0.03..0.4 in size: a few float32 ulps
TOL_U = 1e-6


def _fluid(inp):
    fl = inp["solid"] == 0
    return fl, np.tile(fl, 19)


@pytest.mark.parametrize("case", ["a", "b"])
@pytest.mark.parametrize("precision", ["f32", "f64", "mixed"])
def test_oracle_matches_reference_step(case, precision):
    inp, ref = step_fixture(case)
    got = oracle_raw_steps(inp, 3, precision)
    fl, m19 = _fluid(inp)
    assert np.abs(got["fd_step1_prebc"][m19] - ref["fd_step1_prebc"][m19]).max() < TOL_F
    assert np.abs(got["fd_step1"][m19] - ref["fd_step1"][m19]).max() < TOL_F
    assert np.abs(got["f"][m19] - ref["f_step3"][m19]).max() < 2 * TOL_F
    for k in ("rho1", "ux1", "uy1", "uz1"):
        assert np.abs(got[k][fl] - ref[k][fl]).max() < TOL_U, k
    for k in ("rho", "ux", "uy", "uz"):
        assert np.abs(got[k][fl] - ref[k + "3"][fl]).max() < 2 * TOL_U, k
    # solid cells report rho = 1, u = 0 (lbm3d.py:44)
    assert np.all(got["rho"][~fl] == 1) and np.all(got["ux"][~fl] == 0)


@pytest.mark.parametrize("change", [
    dict(csmag2=0.0),                               # Smagorinsky term dropped
    dict(tau0=0.5010),                              # molecular viscosity off by 0.0004
    dict(Fx=None),                                  # forcing dropped
    dict(cfw=None),                                 # wall function dropped
    dict(uref=None),                                # reference speed ignored
])
def test_golden_comparison_detects_planted_mistakes(change):
    """Each planted change of the oracle's inputs (a dropped term) pushes the
    first-step result past the tolerance used above."""
    inp, ref = step_fixture("a")
    ov = {}
    for k, v in change.items():
        ov[k] = np.zeros_like(inp[k]) if v is None else v
    got = oracle_raw_steps(inp, 1, "f64", overrides=ov)
    fl, m19 = _fluid(inp)
    err = np.abs(got["fd_step1_prebc"][m19] - ref["fd_step1_prebc"][m19]).max()
    assert err > 20 * TOL_F
