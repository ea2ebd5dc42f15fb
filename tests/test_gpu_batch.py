"""Batches of systems (qw_batch_*): the eight molecules of the batching workload (8 to 192
electrons, the 192-electron one on the block-per-walker path) run DMC in shared launches.
Each molecule's block energies, walkers and weights in the batch equal its own engine run
alone bit for bit (fp64 and mixed_fp32), and in FP64 the molecules of up to 64 electrons per
spin follow the oracle's trajectory."""
import sys

import numpy as np
import pytest

from mainline_amd import engine as E
from mainline_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

_DECKS = {}


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _molecules(tmp_path_factory):
    if not _DECKS:
        d = tmp_path_factory.mktemp("batch")
        for m, nw in synth.batch_molecules():
            _DECKS[m.name] = (m.write(str(d / m.name)), m, nw)
    return _DECKS


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_batch_equals_each_engine_alone(prec, tmp_path_factory):
    mols = _molecules(tmp_path_factory)
    m = E.method("DMC", 0, nblock=2, nstep=6, timestep=0.01)
    solo, batched, starts = [], [], []
    for name, (deck, mol, nw) in mols.items():
        nw = max(16, nw // 16)  # a sixteenth of the workload's walkers keeps the test short
        a = E.Engine(deck, prec, seed=1701)
        pos = synth.nonsingular_configs(mol, nw, lambda p: a.eval(p)["logpsi"], seed=61)
        starts.append((name, mol, nw, pos))
        a.set_walkers(pos)
        b = E.Engine(deck, prec, seed=1701)
        b.set_walkers(pos)
        solo.append(a)
        batched.append(b)
    _say("batch", prec, [(n, mol.nelec, nw) for n, mol, nw, _ in starts])
    ref = [e.run(m)["eblock"] for e in solo]
    B = E.Batch(batched)
    res = B.run(m)
    times, buckets = B.kernel_times()
    _say("batch kernels", times, "buckets", buckets)
    for k, (name, mol, nw, _) in enumerate(starts):
        pa, wa = solo[k].get_walkers(nw)
        pb, wb = batched[k].get_walkers(nw)
        assert np.array_equal(pa, pb), (name, np.max(np.abs(pa - pb)))
        assert np.array_equal(wa, wb), name
        assert np.array_equal(ref[k], res["eblock"][k]), (name, ref[k][:, 0], res["eblock"][k][:, 0])
    # one sweep launch per spin channel and step (per recompute interval of 32 moves on the
    # mixed path: the 192-electron molecule has 96 electrons of a spin)
    nseg = 1 if prec == "fp64" else 3
    assert times["sweep"][1] == 2 * nseg * m.nstep * m.nblock
    # a second run continues every engine (DMC state kept) and still equals the engines alone
    ref2 = [e.run(m)["eblock"] for e in solo]
    res2 = B.run(m)
    for k in range(len(starts)):
        assert np.array_equal(ref2[k], res2["eblock"][k]), starts[k][0]
    B.close()


def test_batch_fp64_follows_the_oracle(tmp_path_factory):
    mols = _molecules(tmp_path_factory)
    names = [n for n, (_, mol, _) in mols.items() if mol.nelec <= 64]
    m = E.method("DMC", 0, nblock=1, nstep=3, timestep=0.02)
    engs, starts = [], []
    for name in names:
        deck, mol, _ = mols[name]
        nw = 16
        e = E.Engine(deck, "fp64", seed=4242)
        pos = synth.nonsingular_configs(mol, nw, lambda p: e.eval(p)["logpsi"], seed=62)
        e.set_walkers(pos)
        engs.append(e)
        starts.append((name, deck, mol, nw, pos))
    res = E.Batch(engs).run(m)
    for k, (name, deck, mol, nw, pos) in enumerate(starts):
        _say("batch vs oracle", name, mol.nelec)
        gp, gw = engs[k].get_walkers(nw)
        o = O.Oracle(deck).run(pos, "DMC", nblock=1, nstep=3, timestep=0.02, seed=4242, step0=0)
        assert np.max(np.abs(gp - o["pos"])) < 1e-8, name
        assert np.max(np.abs(gw - o["weights"])) < 1e-8, name
        assert np.max(np.abs(res["eblock"][k][:, :4] - o["eblock"][:, :4])) < 1e-7 * max(1, abs(o["eblock"][0, 0]))


def test_batch_refuses_what_it_does_not_run(tmp_path_factory):
    mols = _molecules(tmp_path_factory)
    deck, mol, _ = next(iter(mols.values()))
    a = E.Engine(deck, "fp64")
    b = E.Engine(deck, "mixed_fp32")
    with pytest.raises(E.QwError) as ex:
        E.Batch([a, b])
    assert ex.value.status == -1
    with pytest.raises(E.QwError):
        E.Batch([a, a])


@pytest.mark.parametrize("prec", ["fp64", "mixed_fp32"])
def test_step_mode_members_join_and_leave(prec, tmp_path_factory):
    """Continuous batching: engines join an open batch at different steps with a VMC
    equilibration and then VMC or DMC, share each step's launches, leave when done, and each
    one's last run equals qw_run of the same runs on an engine alone, bit for bit."""
    mols = _molecules(tmp_path_factory)
    picks = [n for n, (_, mol, _) in mols.items() if mol.nelec in (16, 64, 192)]
    eq = E.method("VMC", 0, nblock=1, nstep=2, timestep=0.3, ndecorr=1)
    plans = [[eq, E.method("DMC", 0, nblock=2, nstep=3, timestep=0.01)],
             [eq, E.method("VMC", 0, nblock=2, nstep=2, timestep=0.3, ndecorr=2)],
             [E.method("DMC", 0, nblock=1, nstep=7, timestep=0.01)]]
    alone, inbatch = [], []
    for k, name in enumerate(picks):
        deck, mol, nw = mols[name]
        nw = max(16, nw // 16)
        a = E.Engine(deck, prec, seed=90 + k)
        pos = synth.nonsingular_configs(mol, nw, lambda p: a.eval(p)["logpsi"], seed=63)
        a.set_walkers(pos)
        for m in plans[k]:
            r = a.run(m)
        alone.append(r["eblock"])
        b = E.Engine(deck, prec, seed=90 + k)
        b.set_walkers(pos)
        inbatch.append(b)
    B = E.Batch(precision=prec)
    ids = {B.join(inbatch[0], plans[0]): 0}
    out, steps = {}, 0
    while len(out) < len(picks):
        if steps == 2:
            ids[B.join(inbatch[1], plans[1])] = 1
        if steps == 3:
            ids[B.join(inbatch[2], plans[2])] = 2
        for mid in B.step():
            out[ids[mid]] = B.leave(mid)["eblock"]
        steps += 1
        assert steps < 100
    _say("step mode", prec, "steps", steps)
    for k in range(len(picks)):
        assert np.array_equal(out[k], alone[k]), (picks[k], out[k][:, 0], alone[k][:, 0])


def test_mixed_batch_recomputes_each_system_at_its_own_interval(tmp_path_factory):
    """Mixed precision: engines with different FP64 recompute intervals share a batch.  The
    launch interval is the common divisor of the intervals that fall inside a sweep (here 64,
    from the 192-electron molecule's explicit 64; the smaller molecules recompute once per
    sweep), each system is recomputed at its own interval, and every engine equals its run
    alone bit for bit.  An open batch refuses an engine whose interval is not a multiple of
    its launch interval."""
    mols = _molecules(tmp_path_factory)
    picks = [n for n, (_, mol, _) in mols.items() if mol.nelec in (16, 64, 192)]
    m = E.method("DMC", 0, nblock=1, nstep=3, timestep=0.01)
    solo, batched, starts = [], [], []
    for name in picks:
        deck, mol, nw = mols[name]
        nw = max(16, nw // 16)
        rf = 64 if mol.nelec == 192 else 0
        a = E.Engine(deck, "mixed_fp32", seed=1702, refresh_interval=rf)
        pos = synth.nonsingular_configs(mol, nw, lambda p: a.eval(p)["logpsi"], seed=64)
        a.set_walkers(pos)
        b = E.Engine(deck, "mixed_fp32", seed=1702, refresh_interval=rf)
        b.set_walkers(pos)
        solo.append(a)
        batched.append(b)
        starts.append((name, mol, nw))
    _say("own intervals", [(n, mol.nelec, nw) for n, mol, nw in starts])
    ref = [e.run(m)["eblock"] for e in solo]
    B = E.Batch(batched)
    res = B.run(m)
    times, _ = B.kernel_times()
    for k, (name, mol, nw) in enumerate(starts):
        pa, wa = solo[k].get_walkers(nw)
        pb, wb = batched[k].get_walkers(nw)
        assert np.array_equal(pa, pb), name
        assert np.array_equal(wa, wb), name
        assert np.array_equal(ref[k], res["eblock"][k]), name
    # 96 electrons of a spin in launches of 64 moves: two sweep launches per spin channel
    assert times["sweep"][1] == 2 * 2 * m.nstep * m.nblock
    B.close()
    deck, mol, _ = mols[picks[-1]]
    odd = E.Engine(deck, "mixed_fp32", refresh_interval=48)
    odd.set_walkers(synth.random_configs(mol, 16, seed=3))
    Bo = E.Batch(precision="mixed_fp32")
    with pytest.raises(E.QwError):
        Bo.join(odd, [m])
    Bo.close()
