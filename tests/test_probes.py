"""Probe lists and ProbesData.dat lines (saras_amd.probes): parser::parseProbes / testProbes
(parser.cc:424-570), probes::placeProbes and probeData's format (probes.cc:61-143).  CPU only."""
import numpy as np
import pytest

from saras_amd import probes


def test_parse_the_reference_example():
    """input/parameters.yaml:140-145 of the reference: one probe, a 3 x 3 x 3 block and a 3 x 1 x 3
    plane; 1:62:3 is linspace(1, 62, 3) rounded half away from zero: 1, 32, 62."""
    text = "[5, 5, 5] [1:62:3, 1:62:3, 1:62:3] [1:62:3, 7, 1:62:3]\n"   # the folded YAML scalar
    pl = probes.parse_probes(text)
    assert len(pl) == 1 + 27 + 9
    assert pl[0] == (5, 5, 5)
    block = pl[1:28]
    assert block[0] == (1, 1, 1) and block[1] == (1, 1, 32) and block[2] == (1, 1, 62) and block[3] == (1, 32, 1)
    assert sorted({p[0] for p in block}) == [1, 32, 62]
    assert all(p[1] == 7 for p in pl[28:])


def test_parse_ranges_like_the_reference():
    # equal ends or one probe give the start alone; a missing count reads the end as the count
    assert probes.parse_probes("[2:2:5, 3:9:1, 0:4:5]") == [(2, 3, k) for k in (0, 1, 2, 3, 4)]
    assert probes.parse_probes("[ 1 ,  2 , 3 ]  [4,5,6]") == [(1, 2, 3), (4, 5, 6)]
    # 0:10:4 -> 0 + round(i * 10/3): 0, 3, 7 (6.67), 10
    assert [p[2] for p in probes.parse_probes("[0, 0, 0:10:4]")] == [0, 3, 7, 10]
    with pytest.raises(ValueError, match="does not match dimensionality"):
        probes.parse_probes("[1, 2]")


def test_check_and_place():
    pl = [(0, 0, 0), (32, 32, 33), (31, 31, 32)]
    probes.check_probes(pl, (5, 5, 5))
    with pytest.raises(ValueError, match="The Z index"):
        probes.check_probes([(1, 1, 34)], (5, 5, 5))
    with pytest.raises(ValueError, match="The X index"):
        probes.check_probes([(-1, 1, 1)], (5, 5, 5))
    # one sub-domain keeps i, j below N - 1 = 32 (subarrayEnds), any k
    assert probes.place_probes(pl, (33, 33, 33)) == [(0, 0, 0), (31, 31, 32)]


def test_line_format():
    """std::fixed time setw(6) setprecision(4), the x index setw(4), values setw(16) setprecision(8)."""
    line = probes.format_line(0.25, (5, 17, 120), np.array([1.0, -2.5e-3, 12345.678901234]))
    assert line == "0.2500\t   5\t17\t120\t      1.00000000\t     -0.00250000\t  12345.67890123"


_CXX = r"""
#include <fstream>
#include <iomanip>
#include <iostream>
int main() {
    double t[3] = {0.0, 0.25, 123.456789};
    int x[3] = {5, 17, 1024}, y[3] = {0, 33, 7}, z[3] = {120, 1, 65};
    double v[3][3] = {{1.0, -2.5e-3, 12345.678901234}, {-0.0, 3.14159265358979, -1e-9}, {42.0, 1e6, -7.125}};
    for (int i = 0; i < 3; i++) {
        std::cout << std::fixed << std::setw(6) << std::setprecision(4) << t[i] << "\t";
        std::cout << std::setw(4) << std::setprecision(0) << x[i] << "\t" << y[i] << "\t" << z[i] << "\t";
        std::cout << std::setw(16) << std::setprecision(8) << v[i][0];
        for (int j = 1; j < 3; j++) std::cout << "\t" << std::setw(16) << std::setprecision(8) << v[i][j];
        std::cout << std::endl;
    }
    return 0;
}
"""


def test_line_format_matches_iostream(tmp_path):
    """The stream statements of probes::probeData (probes.cc:116-135) compiled here print the same
    lines as format_line.  Note the reference's setw(4) applies to the first index only; the y and
    z indices follow a tab unpadded."""
    import subprocess
    src, exe = tmp_path / "p.cc", tmp_path / "p"
    src.write_text(_CXX)
    subprocess.run(["g++", "-O0", "-o", str(exe), str(src)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines()
    cases = [(0.0, (5, 0, 120), [1.0, -2.5e-3, 12345.678901234]),
             (0.25, (17, 33, 1), [-0.0, 3.14159265358979, -1e-9]),
             (123.456789, (1024, 7, 65), [42.0, 1e6, -7.125])]
    assert got == [probes.format_line(t, p, v) for t, p, v in cases]
