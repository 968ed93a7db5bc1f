"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from the oracle):
the oracle must keep reproducing them (CPU) and the engine must match them (GPU)."""
import ctypes
import glob
import os

import numpy as np
import pytest

from helpers import O, backend, gpu_configs, gpu_edges, pkg, scenes

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIXTURES = sorted(glob.glob(os.path.join(HERE, "c*.npz")))


def load(path):
    z = np.load(path, allow_pickle=False)
    d = {k: z[k] for k in z.files}
    kind = str(d["kind"])
    desc = scenes.SceneDesc(kind=kind, limits=d["limits"])
    if kind == "disks":
        desc.robot_dim = int(d["robot_dim"])
        for k in ("robot_radius", "sphere_center", "sphere_radius", "rect_center", "rect_size"):
            setattr(desc, k, d[k])
    elif kind == "planar":
        for k in ("robot_half", "obstacle_pose", "obstacle_half"):
            setattr(desc, k, d[k])
    else:
        desc.n_arms, desc.dof = int(d["n_arms"]), int(d["dof"])
        for k in ("base", "dh", "link_sphere_count", "sphere_local", "boxes"):
            setattr(desc, k, d[k])
        if "mode_n_active" in d:
            desc.modes = {k[5:]: d[k] for k in d if k.startswith("mode_")}
    return desc, d


def test_bfs_fixture():
    z = np.load(os.path.join(HERE, "bfs.npz"))
    off = z["offsets"]
    for n in range(1, 65):
        exp = tuple(int(v) for v in z["flat"][off[n - 1]:off[n]])
        assert O.bfs_indices(n) == exp == pkg.generate_binary_search_indices(n)


def test_sincos_fixture():
    z = np.load(os.path.join(HERE, "sincos.npz"))
    for x, s, c in zip(z["x"], z["s"], z["c"]):
        assert O.sincos(x) == (s, c)


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(path):
    desc, d = load(path)
    o = O.Scene(desc)
    assert np.array_equal(o.check_configs(d["configs"], modes=d.get("config_modes")), d["config_free"])
    assert np.array_equal(o.check_edges(d["q1"], d["q2"], float(d["resolution"]), modes=d.get("edge_modes")),
                          d["edge_free"])
    if "win_N" in d:
        assert np.array_equal(o.check_edges(d["q1"], d["q2"], 0.01, N=d["win_N"], N_start=d["win_start"],
                                            N_max=d["win_max"]), d["win_free"])
    if "fk" in d:
        assert np.array_equal(np.array([o.arm_fk(q) for q in d["fk_q"]]), d["fk"])
        assert np.array_equal(np.array([o.clearance_cost(q, 0.02) for q in d["clr_q"]]), d["clr_cost"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_engine_matches_fixture(path):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    print(f"[case] golden fixture {os.path.basename(path)}", flush=True)
    desc, d = load(path)
    sc = backend().Scene(desc)
    assert np.array_equal(gpu_configs(sc, d["configs"], modes=d.get("config_modes")), d["config_free"])
    assert np.array_equal(gpu_edges(sc, d["q1"], d["q2"], float(d["resolution"]), modes=d.get("edge_modes")),
                          d["edge_free"])
    if "win_N" in d:
        for incl, key in ((False, "win_free"), (True, "win_free_incl")):
            got = gpu_edges(sc, d["q1"], d["q2"], 0.01, N=d["win_N"], N_start=d["win_start"], N_max=d["win_max"],
                            incl=incl)
            assert np.array_equal(got, d[key])
    if "fk" in d:
        q = torch.from_numpy(np.ascontiguousarray(d["fk_q"])).cuda()
        out = torch.empty((q.shape[0], desc.dim, 12), dtype=torch.float64, device="cuda")
        pkg._lib.check(pkg.lib().mrcc_arm_fk(sc.handle, ctypes.c_void_p(q.data_ptr()), q.shape[0], desc.dim, 1,
                                             ctypes.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy(), d["fk"].reshape(out.shape))
