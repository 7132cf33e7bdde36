"""The closed-form pins of tests/test_oracle_pins.py must catch plausible
mistakes in the parts they pin.  Each case plants one mistake in a copy of
oracle/spark_oracle.c (a single textual edit, on the CPU), builds it with
the oracle's own flags into a temporary directory, routes pyoracle to that
build and requires at least one pin to fail.  The unmodified source passes
every pin (test_oracle_pins.py), so a failure here means the planted
mistake was caught.
"""
import shutil
import subprocess
from pathlib import Path

import pytest

import pyoracle
import test_oracle_pins as pins

ORACLE = Path(__file__).resolve().parent.parent / "oracle"
CFLAGS = ["-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-shared"]

VISC = "double cvisc = P->cvisc * fmax2(-(pn[i1] - pn[i1 - 1]), 0.);"
VISC_MASS = "F[HY_MASS] = F[HY_MASS] + cvisc * (pd[i1 - 1] - pd[i1]);"
GRAV_E = "Sg[HY_ENER] = Us[HY_XMOM] * P->grav[0] + Us[HY_YMOM] * P->grav[1] + Us[HY_ZMOM] * P->grav[2];"
GRAV_M = "Sg[HY_ZMOM] = Us[HY_MASS] * P->grav[2];"
SCAL_L = "if (F[HY_MASS] > 0.) F[HY_NUM_FLUX + s] = left[HY_NUM_VARS + s] * F[HY_MASS];"
SCAL_R = "else F[HY_NUM_FLUX + s] = right[HY_NUM_VARS + s] * F[HY_MASS];"

MUTANTS = {
    "viscosity_off": [(VISC, VISC.replace("P->cvisc *", "0.0 *"))],
    "viscosity_on_expansion": [(VISC, VISC.replace("-(pn[i1] - pn[i1 - 1])", "(pn[i1] - pn[i1 - 1])"))],
    "viscosity_mass_sign": [(VISC_MASS, VISC_MASS.replace("+ cvisc", "- cvisc"))],
    "gravity_dropped": [("if (P->use_gravity) {", "if (0) {")],
    "gravity_energy_sign": [(GRAV_E, GRAV_E.replace("= Us[HY_XMOM]", "= -Us[HY_XMOM]")
                             .replace("+ Us[HY_YMOM]", "- Us[HY_YMOM]").replace("+ Us[HY_ZMOM]", "- Us[HY_ZMOM]"))],
    "gravity_energy_from_stage0": [(GRAV_E, GRAV_E.replace("Us[", "U0["))],
    "gravity_z_momentum_dropped": [(GRAV_M, "Sg[HY_ZMOM] = 0.0 * P->grav[2];")],
    "scalars_downwinded": [(SCAL_L, SCAL_L.replace("left[", "right[")), (SCAL_R, SCAL_R.replace("right[", "left["))],
}

CHECKS = [
    ("viscosity compressive", lambda: pins.check_viscosity(True)),
    ("viscosity expanding", lambda: pins.check_viscosity(False)),
    ("scalars u > 0", lambda: pins.check_scalar_upwinding(5.0)),
    ("scalars u < 0", lambda: pins.check_scalar_upwinding(-5.0)),
    ("gravity 1D", lambda: pins.check_gravity(1, (-1.0, 0.5, 0.25), (0.3, -0.2, 0.1))),
    ("gravity 3D", lambda: pins.check_gravity(3, (0.0, 0.0, -1.0), (0.1, -0.05, 0.4))),
]


def build_mutant(tmp, edits):
    src = (ORACLE / "spark_oracle.c").read_text()
    for old, new in edits:
        assert src.count(old) == 1, f"mutation site not unique in spark_oracle.c: {old}"
        src = src.replace(old, new)
    (tmp / "spark_oracle.c").write_text(src)
    shutil.copy(ORACLE / "spark_oracle.h", tmp / "spark_oracle.h")
    out = tmp / "libspark_oracle_mutant.so"
    subprocess.run(["gcc", *CFLAGS, "-o", str(out), str(tmp / "spark_oracle.c"), "-lm"], check=True)
    return out


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_planted_mistake_fails_a_pin(name, tmp_path):
    lib = build_mutant(tmp_path, MUTANTS[name])
    caught = []
    with pyoracle.using(lib):
        for label, check in CHECKS:
            try:
                check()
            except AssertionError:
                caught.append(label)
    assert caught, f"mutant {name} passed every pin"


def test_unmodified_copy_passes_every_pin(tmp_path):
    """The build recipe itself changes nothing: a copy with no edit passes."""
    lib = build_mutant(tmp_path, [])
    with pyoracle.using(lib):
        for _, check in CHECKS:
            check()
