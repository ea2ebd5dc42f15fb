"""Walker files (.config): the library's reader and writer against the reference's own
read_configurations / write_configurations (src/program_utils/qmc_io.h:371-503), run through
oracle/_ref/qw_configs (our driver around the reference's objects, built from its sources), and
against files the reference program itself wrote (tests/golden/configs, written by
tests/golden/ref_configs.py with oracle/_ref/qwalk).  Host code only: no device calls."""
import os
import subprocess

import numpy as np
import pytest

from mainline_amd import engine as E
from tests._util import GOLDEN, ROOT

REF_CONFIGS = os.path.join(ROOT, "oracle", "_ref", "qw_configs")
needs_ref = pytest.mark.skipif(not os.access(REF_CONFIGS, os.X_OK),
                               reason="oracle/_ref/qw_configs not built (oracle/Makefile.ref)")
FIXTURES = {"h4_vmc": (4, False), "n2_dmc": (10, True)}


def _ref_read(path, ne):
    """The reference's read_configurations<Dmc_point> of a file: (pos, weights, signs)."""
    out = subprocess.run([REF_CONFIGS, "read", path, str(ne)], capture_output=True, text=True, timeout=60, check=True)
    lines = [ln for ln in out.stdout.splitlines() if ln.startswith(("nwalkers", "walker")) or ln[:1] in "-0123456789"]
    n = int(lines[0].split()[1])
    pos, w, sg, k = np.zeros((n, ne, 3)), np.zeros(n), np.zeros(n, dtype=int), 1
    for i in range(n):
        f = lines[k].split()
        assert f[0] == "walker" and int(f[1]) == i
        w[i], sg[i] = float(f[3]), int(f[5])
        for e in range(ne):
            pos[i, e] = [float(x) for x in lines[k + 1 + e].split()]
        k += 1 + ne
    return pos, w, sg


def _ref_write(path, pos, weights, dmc):
    """The reference's write_configurations of walkers handed over at %.17g."""
    nw, ne = pos.shape[0], pos.shape[1]
    txt = ["%d %d" % (nw, ne)]
    for i in range(nw):
        txt.append("%.17g " % weights[i] + " ".join("%.17g" % x for x in pos[i].ravel()))
    subprocess.run([REF_CONFIGS, "write", path, "1" if dmc else "0"], input="\n".join(txt) + "\n", text=True,
                   capture_output=True, timeout=60, check=True)


def _walkers(seed, nw, ne):
    rng = np.random.default_rng(seed)
    # magnitudes over several decades and both signs, so the 15-digit text is exercised
    pos = rng.normal(size=(nw, ne, 3)) * 10.0 ** rng.integers(-6, 3, size=(nw, ne, 3))
    w = rng.uniform(0.05, 3.0, size=nw)
    return pos, w


@pytest.mark.parametrize("name", sorted(FIXTURES))
def test_reads_files_the_reference_program_wrote(name):
    ne, dmc = FIXTURES[name]
    path = os.path.join(GOLDEN, "configs", name + ".config")
    pos, w, sg = E.read_configs(path, ne)
    text = open(path).read()
    assert pos.shape[0] == text.split().count("walker")
    assert np.all(np.isfinite(pos))
    if dmc:
        assert np.all(w > 0) and not np.all(w == 1.0)
    else:
        assert np.all(w == 1.0)
    assert np.all(sg == 1)
    if os.access(REF_CONFIGS, os.X_OK):
        rp, rw, rs = _ref_read(path, ne)
        # both parse the same decimal text: equal bit for bit
        assert np.array_equal(pos, rp) and np.array_equal(w, rw) and np.array_equal(sg, rs)


@needs_ref
@pytest.mark.parametrize("dmc", [False, True])
def test_written_file_is_the_reference_writers_byte_for_byte(dmc, tmp_path):
    pos, w = _walkers(7 + dmc, 5, 6)
    ours, theirs = str(tmp_path / "ours.config"), str(tmp_path / "ref.config")
    E.write_configs(ours, pos, w if dmc else None)
    _ref_write(theirs, pos, w, dmc)
    assert open(ours, "rb").read() == open(theirs, "rb").read()


@needs_ref
def test_reference_reads_back_what_the_library_writes(tmp_path):
    pos, w = _walkers(11, 9, 10)
    path = str(tmp_path / "ours.config")
    E.write_configs(path, pos, w)
    rp, rw, rs = _ref_read(path, 10)
    # the file holds 15 significant digits (the reference's stream precision)
    r15 = np.vectorize(lambda x: float("%.15g" % x))
    assert np.array_equal(rp, r15(pos)) and np.array_equal(rw, r15(w)) and np.all(rs == 1)
    # and the library reads its own file to the same values
    p2, w2, _ = E.read_configs(path, 10)
    assert np.array_equal(p2, rp) and np.array_equal(w2, rw)


def test_old_file_becomes_the_backup(tmp_path):
    pos, w = _walkers(3, 2, 2)
    path = str(tmp_path / "x.config")
    E.write_configs(path, pos)
    first = open(path).read()
    E.write_configs(path, pos + 1.0, w)
    assert open(path + ".backup").read() == first
    assert "DMC {" in open(path).read() and "DMC {" not in first


def test_reader_rejects_other_electron_counts_and_empty_files(tmp_path):
    path = os.path.join(GOLDEN, "configs", "h4_vmc.config")
    with pytest.raises(E.QwError, match="electrons"):
        E.read_configs(path, 5)
    empty = tmp_path / "empty.config"
    empty.write_text("\n")
    with pytest.raises(E.QwError, match="no walkers"):
        E.read_configs(str(empty), 4)
    with pytest.raises(E.QwError, match="cannot open"):
        E.read_configs(str(tmp_path / "missing.config"), 4)


def test_method_walker_file_names(tmp_path):
    deck = tmp_path / "run.in"
    deck.write_text("METHOD { VMC NSTEP 5 }\nMETHOD { DMC TIMESTEP 0.02 READCONFIG a.config STORECONFIG b.config }\n")
    d = str(deck)
    assert E.method_files(d, 0) == (d + ".config", d + ".config")
    assert E.method_files(d, 1) == ("a.config", "b.config")
    with pytest.raises(E.QwError):
        E.method_files(d, 2)
