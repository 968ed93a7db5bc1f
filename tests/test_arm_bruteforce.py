"""The arm scenes' composite predicate (which pairs, the mask, the active arms of a mode, held boxes)
checked against an independent numpy / scipy brute force (tests/arm_bruteforce.py), on CPU for the
oracle and on the GPU for the engine.

Scenes are tiny (2-3 arms x 3-4 links x 1-3 spheres, a few boxes, modes with fewer active arms and
held boxes) so that every rule decides some states: for each rule the tests require states whose
only contacts are of that rule (a checker that drops or misstates the rule flips them)."""
import math

import numpy as np
import pytest

import arm_bruteforce as BF
from helpers import O, backend, gpu_configs, gpu_edges, scenes

MIN_DECISIVE = 8  # states per rule whose only contacts come from that rule


def _cases(seed, kind, n_cfg=260):
    """(desc, list of (q, mode, mask, expected collides, deciding category)).  kind: "dh" or
    "chain" (3 arms x 4 links, 3 boxes), "dh-close" / "chain-close" (3 arms x 3 links close
    together, no boxes, larger held boxes: inter-arm and held-box contacts decide more states),
    "dh-reach" (2 arms, held boxes far out on their end-effectors: held boxes meet each other),
    "dh-rest" / "chain-rest" (3 arms x 3 links and 3 resting boxes among them), "chain-prism" (3 arms
    x 4 links, 40% of the joints prismatic; one state in ten has a prismatic coordinate beyond its
    travel, which counts as a collision)."""
    rng = np.random.default_rng(seed)
    chain = kind.startswith("chain")
    if kind.endswith("close"):
        kw = dict(n_arms=3, dof=3, n_boxes=0, spread=0.2, held_half=(0.05, 0.16))
    elif kind.endswith("reach"):
        kw = dict(n_arms=2, dof=3, n_boxes=0, spread=0.25, held_half=(0.06, 0.18), rel_offset=0.35,
                  sphere_r=(0.01, 0.03))
    elif kind.endswith("rest"):
        kw = dict(n_arms=3, dof=3, n_boxes=0, spread=0.25, held_half=(0.04, 0.12), n_rest=3)
    elif kind.endswith("prism"):
        kw = dict(n_arms=3, dof=4, n_boxes=3, prism=0.4)
    else:
        kw = dict(n_arms=3, dof=4, n_boxes=3)
    desc = BF.tiny_arm_scene(rng, scenes, chain=chain, **kw)
    prism = kind.endswith("prism")
    nm = len(desc.modes["n_active"])
    masks = [0xFFFFFFFF, 0b1, 0b10, 0b101, 0b110]
    out = []
    reach = kind.endswith("reach")
    for t in range(n_cfg):
        q = rng.uniform(desc.limits[0], desc.limits[1]) * rng.choice([0.3, 0.6, 1.0])
        if prism and t % 10 == 9:  # a prismatic coordinate beyond its travel
            k = int(rng.choice(np.nonzero(desc.joint_type == 1)[0]))
            q[k] = rng.choice([-1, 1]) * rng.uniform(0.3001, 0.5)
        for _ in range(200 if reach else 0):  # end-effectors close to each other (held boxes meet)
            F = BF.link_frames(desc, q)
            if np.linalg.norm(F[desc.dof - 1][:3, 3] - F[2 * desc.dof - 1][:3, 3]) < 0.3:
                break
            q = rng.uniform(-math.pi, math.pi, desc.dim)
        mode = int(rng.integers(0, nm))
        if BF.out_of_range(desc, q, mode):
            out.append((q, mode, masks[0], True, "travel"))
            continue
        pairs = BF.pair_slacks(desc, q, mode)
        for mask in masks[:1 + (t % len(masks))]:
            exp = BF.decide(pairs, mask)
            if exp is None:
                continue
            holder_only = (not exp) and any(c == "held_sphere_holder" and s < -1e-9 and
                                            all((mask >> int(x)) & 1 for x in rob) for c, rob, s in pairs)
            cat = BF.deciding_category(pairs, mask) if exp else ("holder_only" if holder_only else None)
            out.append((q, mode, mask, exp, cat))
    return desc, out


SEEDS = [(11, "dh"), (12, "dh"), (13, "chain"), (14, "chain"), (15, "dh-close"), (16, "chain-close"),
         (17, "dh-reach"), (18, "dh-reach"), (19, "dh-rest"), (20, "chain-rest"), (21, "chain-prism"),
         (22, "chain-prism")]


@pytest.fixture(scope="module", params=SEEDS, ids=[f"seed{s}-{c}" for s, c in SEEDS])
def case_set(request):
    seed, chain = request.param
    return _cases(seed, chain)


def test_bruteforce_box_overlap_closed_forms():
    I = np.eye(4)
    h = np.array([0.5, 0.5, 0.5])
    # face-to-face touching boxes overlap (closed), slack 0
    assert abs(BF.box_box_slack(I, h, BF.trans(1.0, 0, 0), h, 1e-9)) < 1e-12
    assert BF.box_box_slack(I, h, BF.trans(1.1, 0, 0), h, 1e-9) > 0
    assert BF.box_box_slack(I, h, BF.trans(0.9, 0, 0), h, 1e-9) < 0
    # a cube rotated 45 degrees about z: its edge reaches sqrt(2)/2 along x
    R = BF.rot_z(math.pi / 4)
    assert BF.box_box_slack(I, h, BF.trans(0.5 + 0.70, 0, 0) @ R, h, 1e-9) < 0
    assert BF.box_box_slack(I, h, BF.trans(0.5 + 0.72, 0, 0) @ R, h, 1e-9) > 0
    # edge-edge configuration that every face axis misses: two long bars crossed at 90 degrees
    # over each other, rotated 45 degrees about x, separated along z by a little more than the
    # sum of their half-diagonals in z
    bar = np.array([1.0, 0.1, 0.1])
    A = BF.rot_x(math.pi / 4)
    B = BF.trans(0, 0, 2 * 0.1 * math.sqrt(2) + 0.01) @ BF.rot_z(math.pi / 2) @ BF.rot_x(math.pi / 4)
    assert BF.box_box_slack(A, bar, B, bar, 1e-9) > 0
    B2 = BF.trans(0, 0, 2 * 0.1 * math.sqrt(2) - 0.01) @ BF.rot_z(math.pi / 2) @ BF.rot_x(math.pi / 4)
    assert BF.box_box_slack(A, bar, B2, bar, 1e-9) < 0


def test_bruteforce_kinematics_vs_oracle_fk():
    """The 4x4-product kinematics agree with the oracle's transforms (to rounding)."""
    for seed, kind in SEEDS:
        desc, _ = _cases(seed, kind, n_cfg=0)
        o = O.Scene(desc)
        rng = np.random.default_rng(seed)
        for _ in range(20):
            q = rng.uniform(-3, 3, desc.dim)
            assert np.max(np.abs(BF.link_frames(desc, q) - o.fk_matrices(q))) < 1e-13


def test_oracle_matches_bruteforce(case_set):
    desc, cases = case_set
    o = O.Scene(desc)
    bad = [(i, m, mk) for i, (q, m, mk, exp, _) in enumerate(cases) if o.config_free(q, m, mk) == exp]
    assert not bad, f"{len(bad)} of {len(cases)} states disagree, first {bad[:5]}"


def test_every_rule_decides_states():
    """Across the scene set, each rule is the only source of contact in some states, and held
    boxes touching only their holding link are free in some states."""
    counts = {}
    for seed, kind in SEEDS:
        _, cases = _cases(seed, kind)
        for *_, exp, cat in cases:
            if cat is not None:
                counts[cat] = counts.get(cat, 0) + 1
    for cat in BF.CATS + ("holder_only", "travel"):
        assert counts.get(cat, 0) >= MIN_DECISIVE, (cat, counts)
    # free states exist (a checker answering "collides" everywhere fails)
    assert sum(1 for s, c in SEEDS for *_, exp, _ in _cases(s, c)[1] if not exp) > 100


def test_oracle_mutants_fail():
    """Plausible mistakes in the composite rule, applied to the brute force, disagree with the
    oracle on some state: held box excluded against its whole arm, mask ignored, self pairs of
    adjacent links included, inactive arms present."""
    for which in ("whole_arm", "no_mask", "gap1", "all_active"):
        diff = 0
        for seed, kind in (SEEDS[0], SEEDS[4]):
            desc, cases = _cases(seed, kind)
            o = O.Scene(desc)
            for q, m, mk, exp, _ in cases:
                pairs = BF.pair_slacks(desc, q, m, mutant=which)
                counted = [s for c, rob, s in pairs
                           if c != "held_sphere_holder" and not (which == "whole_arm" and c == "held_sphere_own")
                           and (which == "no_mask" or all((mk >> int(x)) & 1 for x in rob))]
                if any(s < -1e-9 for s in counted) != (not o.config_free(q, m, mk)):
                    diff += 1
        assert diff > 0, which


def test_prismatic_joints_decide_states():
    """Reading the prismatic joints of the chain-prism scenes as revolute (T <- T Rz(q) C) changes
    the oracle-checked verdict of some states: the joint type is exercised."""
    import copy
    diff = 0
    for seed, kind in SEEDS:
        if not kind.endswith("prism"):
            continue
        desc, cases = _cases(seed, kind)
        rev = copy.copy(desc)
        rev.joint_type = None
        o = O.Scene(desc)
        for q, m, mk, exp, cat in cases:
            if cat == "travel":
                continue
            r = BF.decide(BF.pair_slacks(rev, q, m), mk)
            assert o.config_free(q, m, mk) == (not exp)
            diff += r is not None and r != exp
    assert diff >= 20, diff


# ---------------------------------------------------------------------------------------------
def _held_forearm_scene():
    """One 3-link planar arm (DH, all alpha 0) with a held box at the end-effector that can swing
    back onto link 0 or 1 when the arm folds."""
    dh = np.array([[[0.0, 0.5, 0.0], [0.0, 0.5, 0.0], [0.0, 0.25, 0.0]]])
    # one sphere per link at the link's origin-side end (link frame: previous origin at (-a, 0, 0))
    sph = np.array([[-0.25, 0.0, 0.0, 0.05], [-0.25, 0.0, 0.0, 0.05], [-0.125, 0.0, 0.0, 0.05]])
    desc = scenes.SceneDesc(kind="arms", limits=np.array([[-3.1] * 3, [3.1] * 3]), n_arms=1, dof=3,
                            base=np.array([[1.0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]]), dh=dh,
                            link_sphere_count=np.array([1, 1, 1], dtype=np.int32), sphere_local=sph,
                            boxes=np.zeros((0, 6)))
    rel = np.zeros((2, 1, 12))
    rel[:, 0, [0, 4, 8]] = 1.0
    # box centred at the end-effector origin; it always overlaps the holding link's sphere
    # (0.125 from the origin, radius 0.05) and reaches link 0 when the arm folds back
    desc.modes = {"n_active": np.array([1, 1], dtype=np.int32), "has_grasp": np.array([[0], [1]], dtype=np.uint8),
                  "grasp_rel": rel, "grasp_half": np.array([[[0.1] * 3], [[0.1] * 3]])}
    return desc


def test_held_box_vs_own_forearm_oracle():
    desc = _held_forearm_scene()
    o = O.Scene(desc)
    rng = np.random.default_rng(5)
    n_own = n_holder = 0
    for _ in range(3000):
        q = rng.uniform(-3.1, 3.1, 3)
        pairs = BF.pair_slacks(desc, q, 1)
        exp = BF.decide(pairs)
        if exp is None:
            continue
        cat = BF.deciding_category(pairs)
        n_own += cat == "held_sphere_own"
        n_holder += any(c == "held_sphere_holder" and s < -1e-9 for c, _, s in pairs) and not exp
        assert o.config_free(q, 1) == (not exp)
        assert o.config_free(q, 0) == (not BF.decide(BF.pair_slacks(desc, q, 0)))
    assert n_own >= 20 and n_holder >= 20, (n_own, n_holder)


# ---------------------------------------------------------------------------------------------
def _need_gpu():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")


@pytest.mark.gpu
@pytest.mark.parametrize("broadphase", [1, 0])
def test_engine_matches_bruteforce(case_set, broadphase):
    _need_gpu()
    desc, cases = case_set
    print(f"[case] arm brute force, {len(cases)} states, broadphase {broadphase}", flush=True)
    sc = backend().Scene(desc)
    sc.set_broadphase(broadphase)
    Q = np.array([c[0] for c in cases])
    modes = np.array([c[1] for c in cases], dtype=np.int32)
    exp = np.array([c[3] for c in cases])
    masks = sorted(set(c[2] for c in cases))
    for mk in masks:
        sel = np.array([c[2] == mk for c in cases])
        got = gpu_configs(sc, Q[sel], modes=modes[sel], mask=mk)
        assert np.array_equal(got, ~exp[sel]), f"mask {mk:#x}: {np.sum(got != ~exp[sel])} mismatches"


@pytest.mark.gpu
def test_engine_held_box_vs_own_forearm():
    _need_gpu()
    print("[case] held box vs own forearm", flush=True)
    desc = _held_forearm_scene()
    sc = backend().Scene(desc)
    rng = np.random.default_rng(5)
    Q = rng.uniform(-3.1, 3.1, (3000, 3))
    exp = [BF.decide(BF.pair_slacks(desc, q, 1)) for q in Q]
    keep = np.array([e is not None for e in exp])
    got = gpu_configs(sc, Q[keep], modes=np.ones(int(keep.sum()), dtype=np.int32))
    assert np.array_equal(got, ~np.array([e for e in exp if e is not None]))
    # edges in mode 1 between states on both sides: engine vs oracle (the oracle is pinned above)
    o = O.Scene(desc)
    a, b = Q[:400], Q[400:800]
    m = np.ones(400, dtype=np.int32)
    assert np.array_equal(gpu_edges(sc, a, b, 0.05, modes=m), o.check_edges(a, b, 0.05, modes=m))
