"""Each pin of tests/pins.py fails on a one-line mistake planted in a copy of the oracle.

The oracle's sources are copied to a temporary folder, one line is changed, the copy is compiled
with the checker's flags and loaded beside the real oracle; the pin that guards that line must
raise.  The real oracle (oracle/) is never modified.  CPU only (gcc, about a second per build)."""
import os
import shutil
import subprocess

import pytest

import pins
from oracle import pyoracle as po

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")

# (file, anchor, old, new, pin): `old` is replaced at its first occurrence after `anchor`
MUTATIONS = [
    ("field_oracle.c", "double or_ts_nusselt", "i, j, k - 1)]) / 2.0)", "i, j, k)]) / 2.0)",
     pins.nusselt_face_pair),
    ("field_oracle.c", "double or_ts_nusselt", "met(t, 2, 0, 1, k)", "met(t, 2, 1, 1, k)",
     pins.nusselt_tanh_metric),
    ("field_oracle.c", "double or_ts_nusselt", "T[FI(p, i, j, k)] * dv", "T[FI(p, i + 1, j, k)] * dv",
     pins.nusselt_xmode),
    ("field_oracle.c", "default: t->nu", "t->kappa = 1.0 / sqrt(p->Pr * p->Ra)", "t->kappa = 1.0 / sqrt(p->Ra)",
     pins.nusselt_xmode),
    ("mg_oracle.c", "int or_mg_solve", "double mean = sum / (double)L->size;",
     "double mean = sum / (double)(L->nx * L->ny * L->nz);", pins.reference_mean_over_padded),
    ("mg_oracle.c", "int or_mg_solve", "if (fabs(prev - r) < p->mgTolerance) break;",
     "if (r < p->mgTolerance) break;", pins.reference_stop_rule),
    ("mg_oracle.c", "int or_mg_solve", "double prev = 0.0;", "double prev = 1.0;",
     pins.reference_restype0_one_cycle),
    ("mg_oracle.c", "static void consistent_pads", "a[IX(L, -1, j, k)] = a[IX(L, 1, j, k)];",
     "a[IX(L, -1, j, k)] = a[IX(L, 0, j, k)];", pins.pad_mode1_weighted_transpose),
    ("mg_oracle.c", "static void consistent_pads", "m->p.zPer ? Mz - 1 : 1", "m->p.zPer ? Mz : 1",
     pins.pad_mode1_cosine_modes),
    ("mg_oracle.c", "static void consistent_pads", "a[IX(L, i, My + 1, k)] = a[IX(L, i, My - 1, k)];",
     "a[IX(L, i, My + 1, k)] = a[IX(L, i, My, k)];", pins.pad_mode1_cosine_modes),
    ("field_oracle.c", "void or_fd_diff", "H[FI(f, i, j, k)] += t->deriv[FI(f, i, j, k)];",
     "H[FI(f, i, j, k)] = t->deriv[FI(f, i, j, k)];", pins.fd_diff_quadratic),
    ("field_oracle.c", "static void d2(", "out[c] = xx * tmp + 0.5 * x2 * sec;", "out[c] = xx * tmp + x2 * sec;",
     pins.fd_diff_quadratic),
    ("field_oracle.c", "void or_fd_nlin", "off[1][d] = -1;", "off[1][d] = 1;", pins.fd_nlin_linear),
    ("field_oracle.c", "static int slices", "off[2][c] = 1;", "off[2][c] = -1;", pins.fd_nlin_linear),
    ("field_oracle.c", "static void d1(", "out[c] = met(t, d, f->s[d], 1, r) * v;", "out[c] = met(t, d, 1, 1, r) * v;",
     pins.fd_nlin_linear),
    ("field_oracle.c", "static void gradient(", "met(t, c, 0, 1, r)", "met(t, c, 1, 1, r)",
     pins.fd_gradient_divergence_linear),
    ("field_oracle.c", "void or_fd_divergence", "V[1][FI(&t->sV[1], i, j - 1, k)]", "V[1][FI(&t->sV[1], i, j + 1, k)]",
     pins.fd_gradient_divergence_linear),
    ("field_oracle.c", "void or_fd_bc", "int kind = kinds[w] == 1 ? (node ? 2 : 3) : (node ? 4 : 5);",
     "int kind = kinds[w] == 1 ? (node ? 3 : 2) : (node ? 4 : 5);", pins.fd_bc_rules),
]


def _mutated_library(tmp_path, fname, anchor, old, new):
    for f in ("mg_oracle.c", "field_oracle.c", "oracle.h"):
        shutil.copy(os.path.join(ORACLE, f), tmp_path / f)
    src = (tmp_path / fname).read_text()
    a = src.index(anchor)
    b = src.index(old, a)
    (tmp_path / fname).write_text(src[:b] + new + src[b + len(old):])
    out = tmp_path / "liboracle_mutant.so"
    subprocess.run(["gcc", "-std=c11", "-O2", "-fno-fast-math", "-ffp-contract=off", "-fPIC", "-shared", "-o",
                    str(out), str(tmp_path / "mg_oracle.c"), str(tmp_path / "field_oracle.c"), "-lm"],
                   check=True, capture_output=True)
    return po.load(str(out))


@pytest.mark.parametrize("fname,anchor,old,new,pin", MUTATIONS, ids=[f"{m[4].__name__}-{i}" for i, m in enumerate(MUTATIONS)])
def test_pin_catches_planted_mistake(tmp_path, fname, anchor, old, new, pin):
    pin()                                                        # the real oracle passes
    lib = _mutated_library(tmp_path, fname, anchor, old, new)
    with pytest.raises(AssertionError):
        pin(library=lib)
