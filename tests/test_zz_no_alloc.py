"""The allocation-free fixed-configuration evaluation (qw_eval_fixed_static) against
qw_eval_fixed on the same engine: two calls in a row, the second with fewer configurations,
bit for bit (the same kernels run on the same inputs; only where the per-configuration state
lives differs)."""
import numpy as np
import pytest

from mainline_amd import engine as E
from tests._util import golden

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

CASES = [("ae32", "fp64", 0), ("ae32", "mixed_fp32", 0), ("pp16", "fp64", 1), ("n2_ci", "fp64", 0),
         ("h4", "fp64", 0)]


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to("cuda:0")


def _static(eng, pos, rot, pp_mode, pp_rand):
    K, ne = pos.shape[0], eng.ne
    d = dict(logpsi=torch.zeros(K, dtype=torch.float64, device="cuda:0"),
             sign=torch.zeros(K, dtype=torch.float64, device="cuda:0"),
             grad=torch.zeros((K, ne, 3), dtype=torch.float64, device="cuda:0"),
             lap=torch.zeros((K, ne), dtype=torch.float64, device="cuda:0"),
             eparts=torch.zeros((K, 3), dtype=torch.float64, device="cuda:0"))
    dp = _dev(pos)
    dr = None if rot is None else _dev(rot)
    dq = None if pp_rand is None else _dev(pp_rand)
    torch.cuda.synchronize()
    eng.eval_static(K, dp.data_ptr(), None if dr is None else dr.data_ptr(), pp_mode,
                    None if dq is None else dq.data_ptr(),
                    [d[k].data_ptr() for k in ("logpsi", "sign", "grad", "lap", "eparts")])
    torch.cuda.synchronize()
    return {k: v.cpu().numpy() for k, v in d.items()}


@pytest.mark.parametrize("name,prec,pp_mode", CASES)
def test_static_matches_eval_fixed(name, prec, pp_mode):
    g, deck = golden(name)
    eng = E.Engine(deck, prec, device=0, seed=11)
    rng = np.random.default_rng(5)
    # the fixture's configurations, repeated with small displacements
    pos_all = np.concatenate([g["pos"] + 0.01 * i * rng.normal(size=g["pos"].shape) for i in range(4)])
    kmax = eng.eval_static_max()
    assert kmax >= 2
    for K in (min(kmax, pos_all.shape[0]), max(1, min(kmax, pos_all.shape[0]) // 3)):
        pos = np.ascontiguousarray(pos_all[:K])
        rot = None
        pr = None
        if pp_mode == 1:
            q = rng.normal(size=4)
            q /= np.linalg.norm(q)
            a, b, c, d = q
            rot = np.array([[a*a+b*b-c*c-d*d, 2*(b*c-a*d), 2*(b*d+a*c)],
                            [2*(b*c+a*d), a*a-b*b+c*c-d*d, 2*(c*d-a*b)],
                            [2*(b*d-a*c), 2*(c*d+a*b), a*a-b*b-c*c+d*d]])
            pr = rng.random((K, eng.natoms, eng.ne))
        ref = eng.eval(pos, rot=rot, pp_mode=pp_mode, pp_rand=pr)
        got = _static(eng, pos, rot, pp_mode, pr)
        for k in ("logpsi", "sign", "grad", "lap", "eparts"):
            assert np.array_equal(got[k], ref[k]), (name, prec, K, k, np.max(np.abs(got[k] - ref[k])))


def test_static_refuses_too_many_configurations():
    g, deck = golden("ae32")
    eng = E.Engine(deck, "fp64", device=0, seed=11)
    kmax = eng.eval_static_max()
    K = kmax + 1
    buf = torch.zeros(K * eng.ne * 3, dtype=torch.float64, device="cuda:0")
    with pytest.raises(E.QwError) as ex:
        eng.eval_static(K, buf.data_ptr(), None, 0, None, [buf.data_ptr()] * 5)
    assert ex.value.status == -1 and "fit" in str(ex.value)


def test_static_allocates_no_device_memory():
    g, deck = golden("pp16")
    eng = E.Engine(deck, "fp64", device=0, seed=11)
    pos = np.ascontiguousarray(g["pos"][:4])
    _static(eng, pos, None, 0, None)  # first call: buffers of this test are made here
    K, ne = pos.shape[0], eng.ne
    outs = [torch.zeros(n, dtype=torch.float64, device="cuda:0") for n in (K, K, K * ne * 3, K * ne, K * 3)]
    dp = _dev(pos)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info(0)[0]
    for _ in range(3):
        eng.eval_static(K, dp.data_ptr(), None, 0, None, [o.data_ptr() for o in outs])
    torch.cuda.synchronize()
    assert torch.cuda.mem_get_info(0)[0] == free0
