"""Device parity of the configuration-pair costs and distances (mrcc_config_cost).

Bit-exact against the oracle (or_config_cost) for every metric and reduction, over tile edges
(row counts around the 64-row tiles), robot layouts, broadcast / per-row / consecutive-row
(path) operands, strided rows, large dimensions (several tiles of fewer rows), non-finite values;
the reference's own tests of batch_config_cost / batch_config_dist (tests/test_config.py) through
the package API; device tensors in and out; argument errors.
"""
import ctypes

import numpy as np
import pytest

from helpers import O, pkg

pytestmark = pytest.mark.gpu

L = pkg.lib
CODES = {"max": 0, "euclidean": 1, "absmax": 2}
REDS = {"max": 0, "sum": 1, "max_only": 2}
DIMS = [[2, 2], [7, 7], [3, 3, 3], [2, 5], [14]]
NUM_POINTS = [1, 10, 100, 1000, 5000]


def _slices(dims):
    return [(sum(dims[:i]), sum(dims[: i + 1])) for i in range(len(dims))]


def _dev(P, Q, slices, metric, red, w=0.01, ps=None, qs=None, n=None, D=None):
    """Raw C-ABI call on device copies; returns numpy."""
    import torch
    Pd = torch.from_numpy(np.ascontiguousarray(P, dtype=np.float64)).cuda()
    Qd = torch.from_numpy(np.ascontiguousarray(Q, dtype=np.float64)).cuda()
    n = Q.shape[0] if n is None else n
    D = Q.shape[1] if D is None else D
    ps = (0 if P.ndim == 1 else P.shape[1]) if ps is None else ps
    qs = Q.shape[1] if qs is None else qs
    sl = np.ascontiguousarray(np.asarray(slices, dtype=np.int32).reshape(-1))
    out = torch.full((max(n, 1),), -7.0, dtype=torch.float64, device="cuda")
    rc = L().mrcc_config_cost(ctypes.c_void_p(Pd.data_ptr()), ps, ctypes.c_void_p(Qd.data_ptr()), qs, n, D,
                              sl.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)), len(sl) // 2, CODES[metric],
                              REDS[red], float(w), ctypes.c_void_p(out.data_ptr()), None)
    torch.cuda.synchronize()
    return rc, out.cpu().numpy()[:n]


def _same(a, b):
    return np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("metric", ["max", "euclidean", "absmax"])
@pytest.mark.parametrize("red", ["max", "sum", "max_only"])
@pytest.mark.parametrize("dims", [[7, 7, 7, 7], [3, 3, 3], [14], [1, 20, 2], [7] * 8])
def test_matches_oracle(metric, red, dims):
    rng = np.random.default_rng(len(dims) * 31 + CODES[metric] * 7 + REDS[red])
    D = sum(dims)
    for n in (1, 63, 64, 65, 1000, 20011):
        Q = rng.normal(size=(n, D)) * np.exp2(rng.integers(-12, 12, size=(n, D)))
        for P in (rng.normal(size=D), rng.normal(size=(n, D))):
            print(f"[case] {metric} {red} dims={dims} n={n} P={P.shape}", flush=True)
            rc, got = _dev(P, Q, _slices(dims), metric, red)
            assert rc == 0
            exp = O.config_cost(P, Q, _slices(dims), metric, red)
            assert _same(got, exp), (n, P.shape)


@pytest.mark.parametrize("D", [1, 63, 64, 65, 200, 1024])
def test_large_and_odd_dimensions(D):
    """One slice per robot of up to 64 coordinates; dimensions whose tiles hold fewer than 64 rows."""
    rng = np.random.default_rng(D)
    n = 3001
    Q, P = rng.normal(size=(n, D)), rng.normal(size=(n, D))
    cuts = list(range(0, D, 64)) + [D]
    sl = [(cuts[i], cuts[i + 1]) for i in range(len(cuts) - 1)]
    for metric in ("max", "euclidean", "absmax"):
        for red in ("max", "sum", "max_only"):
            print(f"[case] D={D} {metric} {red}", flush=True)
            rc, got = _dev(P, Q, sl, metric, red)
            assert rc == 0 and _same(got, O.config_cost(P, Q, sl, metric, red))


def test_strided_rows_and_path_form():
    """Rows of a wider array (row stride > D), and consecutive rows of one array (the path form
    of batch_config_cost: arr[1:] - arr[:-1], configuration.py:462-468)."""
    rng = np.random.default_rng(3)
    n, D, W = 5000, 28, 33
    A = rng.normal(size=(n, W))
    sl = _slices([7, 7, 7, 7])
    # Q = A[:, :28] with row stride 33, P broadcast
    P = rng.normal(size=D)
    rc, got = _dev(P, A, sl, "euclidean", "max", qs=W, D=D)
    assert rc == 0 and _same(got, O.config_cost(P, np.ascontiguousarray(A[:, :D]), sl, "euclidean", "max"))
    # path form through the package: device tensor of waypoints
    import torch
    Wp = rng.normal(size=(n, D))
    got = pkg.batch_config_cost(torch.from_numpy(Wp).cuda(), None, "euclidean", "max", tmp_agent_slice=np.array(sl))
    assert got.is_cuda and got.shape[0] == n - 1
    exp = O.config_cost(Wp[1:], Wp[:-1], sl, "euclidean", "max")
    assert _same(got.cpu().numpy(), exp)
    got = pkg.batch_config_cost(list(Wp), None, "max", "sum", tmp_agent_slice=np.array(sl))
    assert _same(got, O.config_cost(Wp[1:], Wp[:-1], sl, "max", "sum"))


def test_non_finite_values():
    rng = np.random.default_rng(4)
    n, D = 700, 12
    Q = rng.normal(size=(n, D))
    P = rng.normal(size=(n, D))
    pos = rng.integers(0, D, size=n // 3)
    Q[np.arange(n // 3), pos] = rng.choice([np.nan, np.inf, -np.inf], size=n // 3)
    Q[n // 3: n // 3 + 20, 0] = np.nan  # NaN in the first coordinate
    sl = _slices([4, 4, 4])
    for metric in ("max", "euclidean", "absmax"):
        for red in ("max", "sum", "max_only"):
            rc, got = _dev(P, Q, sl, metric, red)
            assert rc == 0 and _same(got, O.config_cost(P, Q, sl, metric, red)), (metric, red)
    rc, got = _dev(P, Q, [(0, D)], "absmax", "max_only")
    assert _same(got, O.config_cost(P, Q, [(0, D)], "absmax", "max_only"))


def test_argument_errors():
    Q = np.zeros((4, 6))
    for sl, metric in (([(0, 7)], "max"), ([(3, 3)], "max"), ([(4, 2)], "euclidean"), ([], "max"),
                       ([(0, 1)] * 65, "max")):
        rc, _ = _dev(np.zeros(6), Q, sl, metric, "max")
        assert rc == -1, (sl, metric)
    rc, got = _dev(np.ones(6), Q, [(3, 3)], "euclidean", "sum")  # an empty slice is a zero norm
    assert rc == 0 and np.all(got == 0)
    rc, _ = _dev(np.zeros(6), Q, [(0, 6)], "max", "max", D=1025)
    assert rc == -1
    rc, _ = _dev(np.zeros(6), Q, [(0, 6)], "max", "max", n=0)
    assert rc == 0


# ---------------------------------------------------------------------------------------------
# tests/test_config.py through the package API
def _cfg(dims, pt):
    return pkg.NpConfiguration(pt, _slices(dims))


@pytest.mark.parametrize("reduction", ["max", "sum"])
@pytest.mark.parametrize("dims", DIMS)
def test_reference_cost_properties(reduction, dims):
    rng = np.random.default_rng(0)
    D = sum(dims)
    for num_points in NUM_POINTS:
        assert np.all(pkg.batch_config_cost(_cfg(dims, np.zeros(D)), np.zeros((num_points, D)), reduction=reduction) == 0)
        pt = rng.random(D)
        pts = np.tile(pt, (num_points, 1))
        assert np.all(pkg.batch_config_cost(_cfg(dims, pt), pts, reduction=reduction) == 0)
        c, pts = _cfg(dims, rng.random(D)), rng.random((num_points, D))
        a = pkg.batch_config_cost(c, pts, reduction=reduction)
        b = pkg.batch_config_cost(c.from_flat(c.state() + 1), pts + 1, reduction=reduction)
        assert np.allclose(a, b)
        assert _same(a, O.config_cost(c.state(), pts, _slices(dims), "max", reduction))


@pytest.mark.parametrize("dims", [[2], [6], [14]])
def test_reference_cost_single_agent(dims):
    rng = np.random.default_rng(1)
    D = sum(dims)
    for num_points in NUM_POINTS:
        c, pts = _cfg(dims, rng.random(D)), rng.random((num_points, D))
        assert np.allclose(pkg.batch_config_cost(c, pts, reduction="max", w=0.0),
                           pkg.batch_config_cost(c, pts, reduction="sum"))


@pytest.mark.parametrize("dims", DIMS)
def test_reference_manual_comparisons(dims):
    """test_config.py:110-145: equality with the per-robot np.linalg.norm / np.max summed in Python."""
    rng = np.random.default_rng(2)
    D = sum(dims)
    for _ in range(20):
        c, pts = _cfg(dims, rng.random(D)), rng.random((1, D))
        sl = c._array_slice
        e = pkg.batch_config_cost(c, pts, reduction="sum", metric="euclidean")[0]
        m = pkg.batch_config_cost(c, pts, reduction="sum", metric="max")[0]
        me = mm = 0
        for s, t in sl:
            me += np.linalg.norm(pts[0][s:t] - c.state()[s:t])
            mm += np.max(np.abs(pts[0][s:t] - c.state()[s:t]))
        assert e == me and m == mm


@pytest.mark.parametrize("metric", ["max", "euclidean", "sum_euclidean", "max_euclidean"])
@pytest.mark.parametrize("dims", DIMS)
def test_reference_dist_properties(metric, dims):
    rng = np.random.default_rng(3)
    D = sum(dims)
    for num_points in NUM_POINTS:
        assert np.all(pkg.batch_config_dist(_cfg(dims, np.zeros(D)), np.zeros((num_points, D)), metric) == 0)
        pt = rng.random(D)
        assert np.all(pkg.batch_config_dist(_cfg(dims, pt), np.tile(pt, (num_points, 1)), metric) == 0)
        c, pts = _cfg(dims, rng.random(D)), rng.random((num_points, D))
        d0 = pkg.batch_config_dist(c, pts, metric)
        d1 = pkg.batch_config_dist(c.from_flat(c.state() + 1), pts + 1, metric)
        assert np.allclose(d0, d1)
        # the forms of _batch_dist (configuration.py:303-329) against the oracle
        if metric == "euclidean":
            exp = O.config_cost(c.state(), pts, [(0, D)], "euclidean", "max_only")
        elif metric == "max":
            exp = O.config_cost(c.state(), pts, [(0, D)], "absmax", "max_only")
        else:
            exp = O.config_cost(c.state(), pts, _slices(dims), "euclidean",
                                "sum" if metric == "sum_euclidean" else "max_only")
        assert _same(d0, exp)


def test_single_pair_and_list_forms():
    rng = np.random.default_rng(5)
    dims = [7, 7, 7]
    a, b = _cfg(dims, rng.random(21)), _cfg(dims, rng.random(21))
    assert pkg.config_dist(a, b, "max") == np.max(np.abs(a.state() - b.state()))
    assert pkg.config_cost(a, b, "euclidean", "max") == O.config_cost(a.state(), b.state()[None], _slices(dims),
                                                                       "euclidean", "max")[0]
    starts = [_cfg(dims, rng.random(21)) for _ in range(9)]
    ends = [_cfg(dims, rng.random(21)) for _ in range(9)]
    got = pkg.batch_config_cost(starts, ends, "euclidean", "sum")
    exp = O.config_cost(np.array([s.state() for s in starts]), np.array([e.state() for e in ends]),
                        _slices(dims), "euclidean", "sum")
    assert _same(got, exp)
    states = [pkg.State(s, None) for s in starts]
    got = pkg.batch_config_cost(states, [pkg.State(e, None) for e in ends], "max", "max")
    assert _same(got, O.config_cost(np.array([s.state() for s in starts]), np.array([e.state() for e in ends]),
                                    _slices(dims), "max", "max"))
    with pytest.raises(ValueError):
        pkg.batch_config_cost(starts[0], np.zeros((2, 21)), "max", "mean")


def test_environment_methods():
    """abstract_env.py:238-253: the environment's cost uses its cost_metric / cost_reduction."""
    from helpers import envs
    envs()
    env = pkg.get_env_by_name("mrcc.quad_arms")
    rng = np.random.default_rng(6)
    q = [env.start_pos.from_flat(rng.uniform(-1, 1, env.start_pos.state().shape[0])) for _ in range(5)]
    c = env.config_cost(q[0], q[1])
    sl = q[0]._array_slice
    assert c == O.config_cost(q[0].state(), q[1].state()[None], sl, env.cost_metric,
                              env.cost_reduction)[0]
    got = env.batch_config_cost(q[:4], q[1:])
    assert _same(got, O.config_cost(np.array([x.state() for x in q[:4]]), np.array([x.state() for x in q[1:]]),
                                    sl, env.cost_metric, env.cost_reduction))
