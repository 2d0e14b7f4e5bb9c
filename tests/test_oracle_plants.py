"""The oracle's pins catch plausible restatement mistakes: each case makes the oracle
commit one planted mistake (or_test_plant, a test-only switch of the CPU oracle) and
runs the check that is meant to pin that function, which must then fail.

  1  u correction with the x-normal area xp*zp instead of yp*zp
  2  the ap = 0 -> 1e-10 guard of the correction dropped
  3  the pressure update reading the west neighbour's p'
  4  the x-boundary routine also zeroing the max-side ghost (the else-if of Q10 ignored)
  5  the lid written on one face more than the reference's i < cx
  6  Mode R energy convection with the east face velocity instead of the west
  7  Mode R energy spacing from the next cell instead of the previous
  8  solid relaxation weights swapped
  9  Mode C energy wall conductance with the Dirichlet point on the face
 10  Mode C outflow temperature ghost not copied
 11  Mode C momentum wall shear with the no-slip point on the face
 12  Mode C outflow face not copied after the velocity correction
"""
import ctypes as C

import pytest

import test_oracle_pins as pins
from hsolver_amd import abi

R, CFG = {"mode": abi.HS_MODE_REF}, {"mode": abi.HS_MODE_CONFIG}
PLANTS = [
    (1, pins.test_velocity_correction_and_pressure_update, R),
    (2, pins.test_velocity_correction_and_pressure_update, R),
    (3, pins.test_velocity_correction_and_pressure_update, CFG),
    (4, pins.test_mode_r_velocity_bcs_and_lid, {}),
    (5, pins.test_mode_r_velocity_bcs_and_lid, {}),
    (6, pins.test_mode_r_energy_point_update, {}),
    (7, pins.test_mode_r_energy_point_update, {}),
    (8, pins.test_solid_relaxation_point_update, {}),
    (9, pins.test_mode_c_global_energy_balance, {}),
    (10, pins.test_mode_c_boundary_values, {}),
    (11, pins.test_mode_c_developed_duct_flow_and_mass_balance, {}),
    (12, pins.test_correction_divergence_after_exact_mode_c_solve, {}),
]


@pytest.mark.parametrize("n,check,kw", PLANTS, ids=[f"plant{n}-{c.__name__[5:]}" for n, c, _ in PLANTS])
def test_planted_mistake_is_caught(oracle_lib, n, check, kw):
    oracle_lib.or_test_plant.argtypes = [C.c_int32]
    oracle_lib.or_test_plant(n)
    try:
        with pytest.raises(AssertionError):
            check(oracle_lib, **kw)
    finally:
        oracle_lib.or_test_plant(0)
