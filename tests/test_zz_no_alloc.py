"""mrcc_check_edges_noalloc (the batched check on the library's per-GPU workspace, no device
allocation) against mrcc_check_edges (same batch, allocated workspace): identical verdicts on two
calls in a row, the second with a smaller batch, for the disk, planar-box and arm scenes (the arm
batch with vertex indices and modes); invalid edges counted; oversize batches rejected."""
import ctypes

import numpy as np
import pytest

from helpers import backend, gpu_edges, scenes

pytestmark = pytest.mark.gpu


def _torch():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _noalloc(sc, q1, q2, res, **kw):
    torch = _torch()
    a = torch.from_numpy(np.ascontiguousarray(q1)).cuda()
    b = torch.from_numpy(np.ascontiguousarray(q2)).cuda()
    out = torch.full((len(q1),), 7, dtype=torch.uint8, device="cuda")
    iv = lambda v: None if v is None else torch.from_numpy(np.ascontiguousarray(v, dtype=np.int32)).cuda()
    # the batch holds raw pointers: keep the tensors alive until the kernels have run
    Ns, md = iv(kw.get("N_start")), iv(kw.get("modes"))
    batch = sc.edge_batch(a, b, out, res, N_start=Ns, modes=md)
    cnt = torch.full((1,), -1, dtype=torch.int64, device="cuda")
    sc.check_edges_noalloc_device(batch, cnt)
    torch.cuda.synchronize()
    o = out.cpu().numpy()
    assert np.all((o == 0) | (o == 1))
    return o.astype(bool), int(cnt.item())


@pytest.mark.parametrize("kind", ["disks", "planar", "arms", "mixed"])
def test_noalloc_matches_allocating_entry_point(kind):
    _torch()
    print(f"[case] no-alloc vs allocating, {kind}", flush=True)
    rng = np.random.default_rng(21)
    modes = None
    if kind == "disks":
        desc, res = scenes.planar_disks(rng=rng), 0.01
        q1, q2 = scenes.random_pair_edges(rng, desc, 3000)
    elif kind == "planar":
        desc, res = scenes.planar_boxes(rng=rng), 0.01
        q1, q2 = scenes.lognormal_edges(rng, desc, 20000, res)
    elif kind == "arms":
        desc, res = scenes.quad_arms(rng=rng), 0.05
        q1, q2 = scenes.lognormal_edges(rng, desc, 30000, res)
    else:
        desc, res = scenes.mixed_arms(rng=rng, n_modes=64), 0.05
        q1 = np.clip(rng.normal(0, 0.4, (20000, desc.dim)), -2.9, 2.9)
        q2 = np.clip(q1 + rng.normal(0, 0.2, q1.shape), -2.9, 2.9)
        modes = rng.integers(0, 64, len(q1)).astype(np.int32)
    sc = backend().Scene(desc)
    for n in (len(q1), len(q1) // 3):  # two calls in a row, the second smaller
        ref = gpu_edges(sc, q1[:n], q2[:n], res, modes=None if modes is None else modes[:n])
        got, bad = _noalloc(sc, q1[:n], q2[:n], res, modes=None if modes is None else modes[:n])
        assert bad == 0
        assert np.array_equal(got, ref), f"{kind} n={n}: {np.sum(got != ref)} mismatches"
        assert 0 < ref.sum() < n or kind == "planar"


def test_noalloc_invalid_edges_counted():
    torch = _torch()
    print("[case] no-alloc invalid edges", flush=True)
    rng = np.random.default_rng(22)
    desc = scenes.mixed_arms(rng=rng, n_modes=8)
    sc = backend().Scene(desc)
    q1 = np.clip(rng.normal(0, 0.3, (500, desc.dim)), -2.9, 2.9)
    q2 = np.clip(q1 + rng.normal(0, 0.1, q1.shape), -2.9, 2.9)
    N = np.maximum(2, (np.max(np.abs(q2 - q1), axis=1) / 0.05).astype(np.int64) + 1)
    Ns = np.zeros(500, dtype=np.int32)
    Ns[:7] = N[:7] + 1          # N_start > N: the reference asserts
    md = rng.integers(0, 8, 500).astype(np.int32)
    md[7:10] = [8, -1, 100]     # unknown modes
    got, bad = _noalloc(sc, q1, q2, 0.05, N_start=Ns, modes=md)
    assert bad == 10
    assert not got[:10].any()
    ref = gpu_edges(sc, q1[10:], q2[10:], 0.05, modes=md[10:])
    assert np.array_equal(got[10:], ref)


def test_noalloc_rejects_oversize_batch():
    torch = _torch()
    print("[case] no-alloc oversize", flush=True)
    pkg = backend()
    desc = scenes.quad_arms(rng=np.random.default_rng(3))
    sc = pkg.Scene(desc)
    a = torch.zeros((4, desc.dim), dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.uint8, device="cuda")
    batch = sc.edge_batch(a, a, out, 0.05)
    batch.n_edges = 100000  # (24 + 8 * 28) * 100000 bytes > 16 MiB: rejected before any device work
    with pytest.raises(pkg._lib.MrccError) as e:
        sc.check_edges_noalloc_device(batch)
    assert e.value.code == pkg._lib.MRCC_ERR_INVALID
