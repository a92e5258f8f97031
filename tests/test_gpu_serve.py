"""The draw-request server: a served request equals the one-shot sampler on
the same factors and uniforms (bitwise), a request's result is the same
served alone and in a mix, admission respects the HBM budget, and cancelled
requests leave the others unchanged."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MIX = [
    {"dims": [3000, 2000, 1500], "R": 16, "J": 4096, "seed": 11},
    {"dims": [70000, 512, 1024, 900], "R": 32, "J": 20000, "seed": 12},
    {"dims": [1024, 1100, 1300, 1500, 1700], "R": 8, "J": 1024, "seed": 13},
    {"dims": [200000, 3000, 5000], "R": 64, "J": 8192, "seed": 14},
    {"dims": [50000, 60000, 70000], "R": 32, "J": 65536, "seed": 15},
    {"dims": [5000, 4000, 3000, 2000], "R": 16, "J": 3000, "seed": 16},
]


@pytest.fixture(scope="module")
def ftl():
    import fast_tensor_leverage_amd as ftl
    return ftl


def _run_alone(ftl, q, precision):
    srv = ftl.Server(8 << 30, 4)
    rid = srv.submit(q["dims"], q["R"], q["J"], None, q["seed"], precision)
    return srv.wait(rid)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_served_request_equals_sampler(ftl, precision):
    print(f"[gpu] serve vs sampler precision={precision}", flush=True)
    q = MIX[1]
    srv = ftl.Server(8 << 30, 4)
    rid = srv.submit(q["dims"], q["R"], q["J"], None, q["seed"], precision)
    while rid not in srv.progress():
        pass
    U = [srv.factor(rid, k) for k in range(len(q["dims"]))]
    got = srv.result(rid)
    for u in U:  # N(0,1) entries rounded to FP32
        assert np.array_equal(u, u.astype(np.float32).astype(np.float64))
        assert abs(u.mean()) < 0.05 and abs(u.std() - 1.0) < 0.05
    s = ftl.Sampler(U, q["J"], q["R"], "efficient", precision=precision)
    s.set_seed(q["seed"], 0)
    ref = np.zeros((len(U), q["J"]), dtype=np.uint64)
    s.KRPDrawSamples(len(U) + 1, ref)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_alone_equals_mixed(ftl, precision):
    print(f"[gpu] alone vs mixed precision={precision}", flush=True)
    alone = {i: _run_alone(ftl, q, precision) for i, q in enumerate(MIX)}
    srv = ftl.Server(64 << 30, 4)
    ids = {}
    got = {}
    # staggered submission: later requests arrive while earlier ones run
    for i, q in enumerate(MIX):
        ids[srv.submit(q["dims"], q["R"], q["J"], None, q["seed"], precision)] = i
        for rid in srv.progress():
            got[ids[rid]] = srv.result(rid)
    while len(got) < len(MIX):
        for rid in srv.progress():
            got[ids[rid]] = srv.result(rid)
    st = srv.stats()
    assert st["completed"] == len(MIX) and st["bytes_in_use"] == 0
    for i in range(len(MIX)):
        np.testing.assert_array_equal(got[i], alone[i], err_msg=f"request {i}")


def test_admission_by_budget(ftl):
    """With a budget for about two requests the arena never exceeds it and
    every request completes."""
    q = {"dims": [100000, 100000, 100000], "R": 32, "J": 4096}
    srv = ftl.Server(1 << 30, 8)
    rid0 = srv.submit(q["dims"], q["R"], q["J"], None, 1)
    st = srv.stats()
    one = st["bytes_in_use"]
    assert st["active"] == 1 and one > 0
    del srv
    budget = int(2.5 * one)
    srv = ftl.Server(budget, 8)
    ids = [srv.submit(q["dims"], q["R"], q["J"], None, 100 + i) for i in range(6)]
    st = srv.stats()
    assert st["active"] == 2 and st["waiting"] == 4
    done = set()
    while len(done) < len(ids):
        for rid in srv.progress():
            srv.result(rid)
            done.add(rid)
        assert srv.stats()["peak_bytes"] <= budget
    assert done == set(ids)
    with pytest.raises(MemoryError):
        srv.submit([10 ** 7, 10 ** 7, 10 ** 7], 64, 1024, None, 5)
    del rid0


def test_cancel_leaves_others_unchanged(ftl):
    alone = {i: _run_alone(ftl, q, "fp32") for i, q in enumerate(MIX[:4])}
    srv = ftl.Server(64 << 30, 2)
    ids = {srv.submit(q["dims"], q["R"], q["J"], None, q["seed"], "fp32"): i for i, q in enumerate(MIX[:4])}
    victim = {v: k for k, v in ids.items()}
    srv.cancel(victim[3])  # still waiting (two slots)
    srv.cancel(victim[0])  # running
    got = {}
    while len(got) < 2:
        for rid in srv.progress():
            got[ids[rid]] = srv.result(rid)
    assert set(got) == {1, 2}
    for i in got:
        np.testing.assert_array_equal(got[i], alone[i])
    st = srv.stats()
    assert st["cancelled"] == 2 and st["active"] == 0


def test_bad_requests(ftl):
    srv = ftl.Server(1 << 30, 2)
    with pytest.raises(ValueError):
        srv.submit([100, 100], 0, 10)
    with pytest.raises(ValueError):
        srv.submit([100, 100], 8, 10, precision="fp16")
    with pytest.raises(ValueError):
        srv.result(12345)


def _one_shot(ftl, U, J, R, j, seed, precision):
    s = ftl.Sampler(U, J, R, "efficient", precision=precision)
    s.set_seed(seed, 0)
    out = np.zeros((len(U), J), dtype=np.uint64)
    s.KRPDrawSamples(len(U) + 1 if j is None else j, out)
    s.close()
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_registered_factor_set_equals_sampler(ftl, precision):
    """Draws on a registered (resident) factor set equal the one-shot sampler
    on the same factors and seed, for J below J_max and an excluded mode; an
    update of one factor rebuilds its tree for the next draw."""
    print(f"[gpu] registered factor set precision={precision}", flush=True)
    rng = np.random.default_rng(21)
    dims, R, Jmax = [30000, 4000, 9000], 16, 8192
    U = [rng.standard_normal((d, R)) for d in dims]
    srv = ftl.Server(8 << 30, 4)
    fs = srv.add_factors(U, Jmax, precision)
    for J, j, seed in ((Jmax, None, 5), (3001, 1, 6), (1, None, 7)):
        got = srv.wait(srv.submit_on(fs, J, j, seed))
        exp = _one_shot(ftl, U, J, R, j, seed, precision)
        if j is not None:
            exp[j] = 0
        np.testing.assert_array_equal(got, exp, err_msg=f"J={J} j={j}")
    U1 = rng.standard_normal((dims[1], R)) * rng.random((dims[1], 1))
    srv.update_factor(fs, 1, U1)
    got = srv.wait(srv.submit_on(fs, 2048, None, 9))
    np.testing.assert_array_equal(got, _one_shot(ftl, [U[0], U1, U[2]], 2048, R, None, 9, precision))
    srv.remove_factors(fs)
    assert srv.stats()["bytes_in_use"] == 0


def test_factor_sets_and_generated_requests_mixed(ftl):
    """Factor-set draws (one of them registered from a Sampler handle) and
    seed-generated requests submitted together, several per packed launch:
    each result equals the same request served alone."""
    print("[gpu] factor sets and generated requests mixed", flush=True)
    rng = np.random.default_rng(22)
    sets = []
    for dims, R in (([5000, 6000, 7000], 32), ([2000, 3000, 4000, 5000], 8), ([100000, 900], 32)):
        sets.append(([rng.standard_normal((d, R)) for d in dims], R))
    srv = ftl.Server(32 << 30, 8)
    fss = [srv.add_factors(U, 4096, "fp32") for U, _ in sets[:2]]
    smp = ftl.Sampler(sets[2][0], 4096, sets[2][1], "efficient", precision="fp32")
    fss.append(srv.add_sampler(smp))
    jobs = {}
    for rep in range(2):
        for i, fs in enumerate(fss):
            jobs[srv.submit_on(fs, 4096 - 1000 * rep, None, 100 + 10 * i + rep)] = ("fs", i, rep)
        for i, q in enumerate(MIX[:4]):
            jobs[srv.submit(q["dims"], q["R"], q["J"], None, q["seed"] + rep, "fp32")] = ("gen", i, rep)
    got = {}
    while len(got) < len(jobs):
        srv.wait_for(0, 10.0)
        for rid in srv.progress():
            got[rid] = srv.result(rid)
    for rid, (kind, i, rep) in jobs.items():
        if kind == "fs":
            U, R = sets[i]
            exp = _one_shot(ftl, U, 4096 - 1000 * rep, R, None, 100 + 10 * i + rep, "fp32")
        else:
            q = MIX[i]
            exp = _run_alone(ftl, {**q, "seed": q["seed"] + rep}, "fp32")
        np.testing.assert_array_equal(got[rid], exp, err_msg=f"{kind} {i} rep {rep}")
    for fs in fss:
        srv.remove_factors(fs)
    assert srv.stats()["bytes_in_use"] == 0


def test_device_factors_are_read_before_the_call_returns(ftl):
    """add_factors / update_factor with torch device tensors: the tensors are
    overwritten and freed right after the calls, and later draws still equal
    the one-shot sampler on the original factors."""
    import torch
    print("[gpu] device factors read before return", flush=True)
    rng = np.random.default_rng(23)
    dims, R = [20000, 3000, 7000], 16
    U = [rng.standard_normal((d, R)) for d in dims]
    dev = torch.device("cuda", 0)
    srv = ftl.Server(8 << 30, 4)
    T = [torch.from_numpy(u).to(dev) for u in U]
    fs = srv.add_factors(T, 4096, "fp32")
    for t in T:
        t.fill_(float("nan"))
    del T
    junk = [torch.full((d, R), 7.0, dtype=torch.float64, device=dev) for d in dims]  # reuse the freed blocks
    got = srv.wait(srv.submit_on(fs, 4096, None, 3))
    np.testing.assert_array_equal(got, _one_shot(ftl, U, 4096, R, None, 3, "fp32"))
    U1 = rng.standard_normal((dims[1], R))
    t1 = torch.from_numpy(U1).to(dev)
    srv.update_factor(fs, 1, t1)
    t1.fill_(float("nan"))
    del t1, junk
    got = srv.wait(srv.submit_on(fs, 4096, None, 4))
    np.testing.assert_array_equal(got, _one_shot(ftl, [U[0], U1, U[2]], 4096, R, None, 4, "fp32"))
    srv.remove_factors(fs)


def test_factor_set_errors(ftl):
    srv = ftl.Server(1 << 30, 2)
    U = [np.ones((100, 4)), np.ones((50, 4))]
    fs = srv.add_factors(U, 100)
    with pytest.raises(ValueError):
        srv.submit_on(fs, 101)          # above J_max
    with pytest.raises(ValueError):
        srv.submit_on(fs + 7, 10)       # unknown set
    with pytest.raises(ValueError):
        srv.update_factor(fs, 2, np.ones((50, 4)))   # mode out of range
    with pytest.raises(MemoryError):
        srv.add_factors([np.ones((10 ** 7, 8))] * 3, 1000)  # over the budget
    srv.remove_factors(fs)
