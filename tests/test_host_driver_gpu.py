"""The C++ drop-in classes (saras_amd/host: parser, grid, the field classes, poisson / multigrid_d3) driven
like the reference's own programs: the restated known-answer tests and the TEST_POISSON r^2/6 run
of tests/testPoisson.sh, on one GPU through the C ABI."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "saras_amd", "_build", "saras_mg_driver")


def run(*args):
    assert os.path.exists(DRIVER), "C++ driver not built"
    r = subprocess.run([DRIVER, *map(str, args)], capture_output=True, text=True, timeout=120)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.parametrize("ind,domain", [(5, "PPP"), (4, "PPN"), (5, "NNN")])
def test_reference_kats_pass(ind, domain):
    rc, out = run("kat", ind, domain)
    print(out)
    assert rc == 0 and out.count("[PASSED]") == 3, out


def test_mgtest_r2_case():
    rc, out = run("mgtest", 5, 4, 25)
    print(out[-2000:])
    assert rc == 0
    res = [float(x) for x in re.findall(r"Residual after V Cycle \d+ is (\S+)", out)]
    assert len(res) == 25 and res[-1] < res[0] * 1e-6
    dev = float(re.search(r"Maximum absolute deviation from analytic solution is: (\S+)", out).group(1))
    assert dev < 1e-8


@pytest.mark.parametrize("ind", [4, 5])
def test_field_classes_pass_reference_unit_tests(ind):
    """derivative / vfield / sfield / plainsf / plainvf of saras_amd/host/field.h over
    include/saras_field.h: differ_test, field_test and nlin_test of src/tests/units and the ghost
    values of hydro_d3::testPeriodic, each printed [PASSED] like the reference's unit tests."""
    rc, out = run("fieldtest", ind)
    print(out)
    assert rc == 0 and out.count("[PASSED]") == 4, out
