"""Serving loop (qw_server_*): concurrent jobs of different sizes on one device, admission by
HBM budget, and each job's result the same alone and in a mix."""
import numpy as np
import pytest

from mainline_amd import engine as E
from tests._util import synth_deck

pytestmark = pytest.mark.gpu


def _jobs():
    h4, _ = synth_deck("h4")
    pp16, _ = synth_deck("pp16")
    ae32, _ = synth_deck("ae32")
    return [
        (h4, E.method("VMC", 256, nblock=3, nstep=5, timestep=0.5), 256, "fp64", 11),
        (pp16, E.method("DMC", 128, nblock=2, nstep=5, timestep=0.01), 128, "mixed_fp32", 12),
        (ae32, E.method("DMC", 64, nblock=2, nstep=3, timestep=0.01), 64, "fp64", 13),
        (h4, E.method("DMC", 512, nblock=2, nstep=10, timestep=0.02), 512, "mixed_fp32", 14),
        (pp16, E.method("VMC", 96, nblock=2, nstep=4, timestep=0.3), 96, "fp64", 15),
        (ae32, E.method("VMC", 32, nblock=2, nstep=2, timestep=0.3), 32, "mixed_bf16x6", 16),
    ]


def test_each_job_is_the_same_alone_and_in_the_mix():
    jobs = _jobs()
    alone = []
    for deck, m, nw, prec, seed in jobs:
        s = E.Server(max_active=1)
        jid = s.submit(deck, m, nw, precision=prec, seed=seed, equil_steps=3)
        eb, st = s.wait(jid)
        s.close()
        alone.append(eb)
        print("alone", prec, nw, eb[:, 0], "latency %.3f s" % st[3], flush=True)
    s = E.Server(max_active=6)
    ids = [s.submit(deck, m, nw, precision=prec, seed=seed, equil_steps=3) for deck, m, nw, prec, seed in jobs]
    res = [s.wait(j) for j in ids]
    s.close()
    admitted = [st[1] for _, st in res]
    finished = [st[2] for _, st in res]
    # several jobs were on the device at once
    overlap = sum(1 for i in range(len(res)) for j in range(len(res))
                  if i < j and admitted[j] < finished[i] and admitted[i] < finished[j])
    print("overlapping job pairs", overlap, flush=True)
    assert overlap >= 1
    for a, (eb, st) in zip(alone, res):
        assert np.array_equal(a, eb)
        assert st[4] > 0 and st[6] > 0


def test_admission_respects_the_hbm_budget():
    deck, _ = synth_deck("pp16")
    m = E.method("VMC", 256, nblock=1, nstep=3, timestep=0.3)
    probe = E.Engine(deck, "fp64")
    b = ctypes_bytes(probe, 256)
    probe.close()
    s = E.Server(hbm_budget=int(b * 1.5), max_active=8)  # room for one job at a time
    ids = [s.submit(deck, m, 256, seed=20 + i) for i in range(3)]
    res = [s.wait(j)[1] for j in ids]
    s.close()
    for prev, nxt in zip(res, res[1:]):
        assert nxt[1] >= prev[2] - 1e-6  # admitted only after the previous one finished
    big = E.Server(hbm_budget=int(b // 2))
    jid = big.submit(deck, m, 256)
    with pytest.raises(E.QwError) as ex:
        big.wait(jid)
    big.close()
    assert ex.value.status == -5


def ctypes_bytes(eng, nw):
    import ctypes
    v = ctypes.c_ulonglong()
    assert E.lib().qw_engine_walker_bytes(eng.h, nw, ctypes.byref(v)) == 0
    return v.value


def test_serving_workload_sizes_alone_and_in_the_mix(tmp_path):
    """Jobs at the serving workload's sizes (8 to 512 electrons, 1024 walkers, VMC and DMC,
    fp64 and mixed_fp32): the batched jobs share launches step by step (the 192-electron and
    512-electron ones on the block-per-walker path; the 512-electron mixed-precision one, which
    no batch holds, on its own engine), and every job's block records equal the same job run
    alone, bit for bit."""
    from mainline_amd import synth
    decks = {}
    for ne in (8, 48, 96, 192, 512):
        m = synth.z4_molecule(ne // 4, 20261300 + ne, name="srv%d" % ne)
        decks[ne] = m.write(str(tmp_path / m.name))
    jobs = [
        (8, E.method("VMC", 1024, nblock=2, nstep=4, timestep=0.3, ndecorr=1), "fp64", 31),
        (48, E.method("DMC", 1024, nblock=2, nstep=3, timestep=0.01), "mixed_fp32", 32),
        (96, E.method("VMC", 1024, nblock=1, nstep=3, timestep=0.3, ndecorr=1), "mixed_fp32", 33),
        (192, E.method("DMC", 1024, nblock=1, nstep=2, timestep=0.01), "mixed_fp32", 34),
        (512, E.method("DMC", 1024, nblock=1, nstep=2, timestep=0.01), "fp64", 35),
        (512, E.method("VMC", 1024, nblock=1, nstep=1, timestep=0.3, ndecorr=1), "mixed_fp32", 36),
        (48, E.method("DMC", 1024, nblock=2, nstep=6, timestep=0.01), "fp64", 37),
    ]
    alone = []
    for ne, m, prec, seed in jobs:
        s = E.Server(max_active=1)
        jid = s.submit(decks[ne], m, 1024, precision=prec, seed=seed, equil_steps=2)
        eb, st = s.wait(jid)
        s.close()
        alone.append(eb)
        print("alone", ne, prec, eb[:, 0], "latency %.3f s" % st[3], flush=True)
        assert np.all(np.isfinite(eb[:, 0]))
    s = E.Server(max_active=16, hbm_budget=96 << 30)
    ids = [s.submit(decks[ne], m, 1024, precision=prec, seed=seed, equil_steps=2) for ne, m, prec, seed in jobs]
    res = [s.wait(j) for j in ids]
    s.close()
    for (ne, m, prec, seed), a, (eb, st) in zip(jobs, alone, res):
        print("mix", ne, prec, eb[:, 0], "latency %.3f s" % st[3], flush=True)
        assert np.array_equal(a, eb), (ne, prec)


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_cloned_engine_equals_a_fresh_engine(prec):
    """qw_engine_clone (the server's per-deck template) gives the engine qw_engine_create gives:
    the same DMC block records bit for bit, also when the clone changes precision and seed."""
    deck, mol = synth_deck("pp16")
    src = E.Engine(deck, "fp64", seed=1)
    print("clone", prec, flush=True)
    a = E.Engine(deck, prec, seed=77)
    b = src.clone(precision=prec, seed=77)
    m = E.method("DMC", 64, nblock=2, nstep=4, timestep=0.01)
    pos = a.random_walkers(64, seed=3)
    a.set_walkers(pos)
    b.set_walkers(pos)
    ra, rb = a.run(m), b.run(m)
    assert np.array_equal(ra["eblock"], rb["eblock"])
    pa, wa = a.get_walkers(64)
    pb, wb = b.get_walkers(64)
    assert np.array_equal(pa, pb) and np.array_equal(wa, wb)
    for e in (a, b, src):
        e.close()
