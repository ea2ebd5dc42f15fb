"""The C-ABI library: loads, exports every declared entry point, and rejects bad input
before it touches a device (no GPU needed)."""
import os
import re

import numpy as np
import pytest

from mainline_amd import engine as E
from tests._util import ROOT, golden

HEADER = os.path.join(ROOT, "include", "qwalk_amd.h")


def declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\w[\w\s\*]*?\b(qw_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 10
    L = E.lib()
    for n in names:
        assert hasattr(L, n), n


def test_missing_deck_is_an_input_error():
    with pytest.raises(E.QwError) as ex:
        E.Engine("/nonexistent/deck")
    assert ex.value.status == -2


def test_determinant_input_checks(tmp_path):
    """STATES must hold ndet x Ne orbitals, a CSF a coefficient and weights, and CSF excludes
    DETWT, and a centers file must exist (bad input); several determinants with a mixed
    precision are valid QWalk the engine does not run (unsupported)."""
    g, deck = golden("h_atom")
    d = os.path.dirname(deck)
    base = open(os.path.join(d, "qw.slater")).read()
    for f in ["qw.sys", "qw.basis", "qw.orb"]:
        (tmp_path / f).write_text(open(os.path.join(d, f)).read())
    (tmp_path / "x.wfeval").write_text("INCLUDE qw.sys\nTRIALFUNC { INCLUDE qw.slater }\n")
    for txt, status in [(base.replace("DETWT { 1.0 }", "DETWT { 1.0 0.5 }"), -2),
                        (base.replace("DETWT { 1.0 }", "CSF { 1.0 }"), -2),
                        (base.replace("DETWT { 1.0 }", "DETWT { 1.0 } CSF { 1.0 1.0 }"), -2),
                        (re.sub(r"CENTERS\s*\{[^}]*\}", "CENTERS { READ qw.centers }", base, flags=re.I), -2)]:
        (tmp_path / "qw.slater").write_text(txt)
        with pytest.raises(E.QwError) as ex:
            E.Engine(str(tmp_path / "x.wfeval"))
        assert ex.value.status == status, ex.value
    g, deck = golden("n2_ci")
    with pytest.raises(E.QwError) as ex:
        E.Engine(deck, "mixed_fp32")
    assert ex.value.status == -3
    # BLAS_MO orbitals and basis centers read from a file are read (on a machine without a GPU
    # engine creation then stops at the device)
    for name in ("n2_blas", "n2_cen"):
        g, deck = golden(name)
        try:
            E.Engine(deck, "fp64").close()
        except E.QwError as ex:
            assert ex.status == -4, (name, ex)


def test_method_read_applies_reference_defaults(tmp_path):
    p = tmp_path / "m.dmc"
    p.write_text("METHOD { DMC TIMESTEP 0.01 }\n")
    m = E.read_method(str(p))
    assert (m.method, m.nblock, m.nconfig, m.nstep, m.feedback_interval) == (1, 100, 2048, 100, 5)
    assert (m.max_weight, m.branch_start_cutoff, m.feedback) == (7.0, 10.0, 1.0)
    p.write_text("METHOD { VMC NCONFIG 256 TIMESTEP 0.5 NBLOCK 50 NSTEP 20 }\n")
    m = E.read_method(str(p))
    assert (m.method, m.nconfig, m.nblock, m.nstep, m.ndecorr, m.timestep) == (0, 256, 50, 20, 2, 0.5)


def test_method_read_tmoves_keywords(tmp_path):
    """TMOVES (one draw per walker) and TMOVESSC (per electron); TMOVES wins when both are
    given, as doTmove tests it first (src/method/Dmc_method.cpp:70-73,869)."""
    p = tmp_path / "m.dmc"
    for text, want in [("", 0), ("TMOVES", 1), ("TMOVESSC", 2), ("TMOVESSC TMOVES", 1)]:
        p.write_text("METHOD { DMC TIMESTEP 0.01 %s }\n" % text)
        assert E.read_method(str(p)).tmoves == want


def test_csf_fixture_is_the_same_expansion():
    """The reference evaluates the CSF form of the N2 CI expansion exactly as its DETWT form."""
    a, _ = golden("n2_ci")
    b, _ = golden("n2_csf")
    for k in ["logpsi", "sign", "grad", "lap", "kinetic", "vloc", "vnonloc"]:
        assert np.array_equal(a[k], b[k]), k


def test_every_method_section_in_order(tmp_path):
    """Several METHOD sections run one after another (src/program_utils/ooqmc.cpp:148-170):
    each is read with its own defaults, LABEL (default vmc / dmc) and START_FEEDBACK
    (default 1, src/method/Dmc_method.cpp:94-98)."""
    p = tmp_path / "two.run"
    p.write_text("METHOD { VMC NSTEP 7 }\nMETHOD { DMC TIMESTEP 0.02 LABEL mydmc START_FEEDBACK 0.5 }\n"
                 "METHOD { DMC TIMESTEP 0.01 }\n")
    ms = E.read_methods(str(p))
    assert [(m.method, lab) for m, lab in ms] == [(0, "vmc"), (1, "mydmc"), (1, "dmc")]
    assert ms[0][0].nstep == 7 and ms[0][0].timestep < 0  # no TIMESTEP: tuned by the driver
    assert (ms[1][0].timestep, ms[1][0].start_feedback, ms[2][0].start_feedback) == (0.02, 0.5, 1.0)
    assert E.read_method(str(p)).nstep == 7
    p.write_text("METHOD { DMC TIMESTEP 0.02 FEEDBACK_INTERVAL 0 }\n")
    with pytest.raises(E.QwError):
        E.read_methods(str(p))


def test_driver_rejects_unknown_precision(tmp_path):
    """A misspelt --precision is an error, not a silent choice of another mode."""
    import subprocess
    driver = os.path.join(ROOT, "mainline_amd", "qwalk_amd")
    if not os.path.exists(driver):
        pytest.skip("driver not built")
    out = subprocess.run([driver, str(tmp_path / "none.run"), "--precision", "fp32"], capture_output=True,
                         text=True, timeout=60)
    assert out.returncode == 2
    assert "unknown precision" in out.stderr
