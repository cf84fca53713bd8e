"""GPU tests of the learned surrogate and of the culled sphere-union fold, through the C ABI.

* The BF16-MFMA forward pass and the backward kernels against the numpy oracle (mlp_oracle), which
  emulates the device's BF16 operand rounding; the oracle's own backward is pinned against finite
  differences on the CPU (test_mlp_oracle.py).
* Training on the device lowers the loss and fits the body.
* The culled fold is bitwise the plain fold: phi and every column equal the oracle's.
* The surrogate update leaves every CutCellData column bitwise equal to the exact path, whatever
  the network (trained or untrained); phi differs only away from the cut cells, never in sign.
* What the trained network gains over the path without it: far fewer exact evaluations than an
  untrained network (and than the culled-exact path, which evaluates every node), with far-field
  phi close to the exact one.
"""
import numpy as np
import pytest

import mlp_oracle as mo
import oracle as orc
from cutcell import (CartesianIsoGrid, MarchingCubesCutCell, SDFSphereUnion, Surrogate,
                     allocate_cache, update_cache)
from cutcell.workloads import csg_spheres

pytestmark = pytest.mark.gpu

COLS = ("kind", "ambiguous", "volume_fraction", "centroid", "face_fraction",
        "face_centroid_local", "interface_centroid")


def test_forward_matches_the_bf16_oracle():
    rng = np.random.default_rng(0)
    net = Surrogate(nfreq=64, depth=4, sigma=2.0, seed=7)
    p, B = net.params()
    p[-1] = 0.25  # a non-zero output bias
    net.set_params(p)
    x = rng.uniform(-1, 1, (1000, 3)).astype(np.float32)  # not a multiple of the 64-point tile
    y = net.eval(x)
    ye, _ = mo.forward(p, B, x, 64, 256, 4, emulate=True)
    yd, _ = mo.forward(p.astype(np.float64), B, x, 64, 256, 4, emulate=False)
    scale = float(np.sqrt(np.mean(yd ** 2)))
    err_e = float(np.max(np.abs(y - ye)))
    err_d = float(np.max(np.abs(y - yd)))
    print(f"rms(y)={scale:.3e} max|gpu-bf16 oracle|={err_e:.3e} max|gpu-fp64|={err_d:.3e}")
    assert scale > 1e-2
    # the emulating oracle explains the device to a few BF16 rounding flips; a layout error
    # (transposed operand, wrong lane mapping) would be O(scale)
    assert err_e <= 2e-2 * scale
    assert err_e < err_d


@pytest.mark.parametrize("n", [512, 64])  # 64: one forward tile, one or two point chunks per dW tile
def test_loss_and_gradients_match_the_bf16_oracle(n):
    rng = np.random.default_rng(1)
    net = Surrogate(nfreq=64, depth=4, sigma=2.0, seed=3)
    p, B = net.params()
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    y = (np.linalg.norm(x, axis=1) - 0.5).astype(np.float32)
    loss, g = net.loss_grad(x, y)
    lo, go = mo.loss_grad(p, B, x, y, 64, 256, 4, emulate=True)
    print(f"loss gpu={loss:.6e} oracle={lo:.6e}")
    assert abs(loss - lo) <= 1e-3 * abs(lo)
    # block by block: every weight matrix, the biases, w_out, b_out
    Ks = [128, 256, 256, 256]
    off, blocks = 0, []
    for K in Ks:
        blocks.append((f"W{len(blocks)}", off, off + K * 256))
        off += K * 256
    for l in range(4):
        blocks.append((f"b{l}", off, off + 256))
        off += 256
    blocks.append(("wout", off, off + 256))
    blocks.append(("bout", off + 256, off + 257))
    for name, a, b in blocks:
        rel = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-30)
        print(f"{name}: |g|={np.linalg.norm(go[a:b]):.3e} rel err={rel:.2e}")
        assert rel <= 2e-2, name


def test_training_reduces_the_loss_and_fits_the_body():
    body = csg_spheres(5, count=16, lo=-0.5, hi=0.5, rmin=0.1, rmax=0.25)
    net = Surrogate(nfreq=64, depth=4, sigma=4.0, seed=11)  # a smooth body: lower frequencies
    hist = net.train(body, nsamples=1 << 20, band=0.05, epochs=12, batch=16384, lr=5e-3, seed=2)
    print("loss per epoch", hist)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.05 * hist[0]
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, (20000, 3))
    exact = np.array([orc.sdf_value(body, 3, q) for q in x])
    err = np.sqrt(np.mean((net.eval(x.astype(np.float32)) - exact) ** 2))
    print(f"held-out rms error {err:.3e}")
    assert err < 0.02


def test_training_is_deterministic():
    # no atomics anywhere in the step: the weight gradients are summed from per-chunk partial
    # tiles and the other gradients from per-workgroup sums, in fixed orders, and the samples and
    # the epochs' orders come from counter-based generators -- two runs agree to the last bit
    body = csg_spheres(5, count=16, lo=-0.5, hi=0.5, rmin=0.1, rmax=0.25)
    runs = []
    for _ in range(2):
        net = Surrogate(nfreq=64, depth=4, sigma=4.0, seed=11)
        hist = net.train(body, nsamples=1 << 16, band=0.05, epochs=3, batch=4096, lr=5e-3, seed=2)
        runs.append((hist, net.params()[0]))
        net.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1].view(np.uint32), runs[1][1].view(np.uint32))
    assert runs[0][0][-1] < runs[0][0][0]


def union_case(n=(70, 45, 33), count=64, seed=4):
    body = csg_spheres(seed, count=count, lo=-0.6, hi=0.6, rmin=0.06, rmax=0.2)
    g = CartesianIsoGrid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), n=n)
    return body, g


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_culled_fold_is_bitwise_the_plain_fold(dtype):
    body, g = union_case()
    plain = allocate_cache(g, MarchingCubesCutCell, dtype)
    culled = allocate_cache(g, MarchingCubesCutCell, dtype)
    culled.set_phi_source("culled")
    update_cache(plain, body, g)
    update_cache(culled, body, g)
    st = culled.phi_stats()
    print(st, "mean kept", st["kept_sum"] / st["bricks"])
    assert st["nodes"] == np.prod(np.array(g.n) + 1) == st["exact"]
    assert st["kept_sum"] < st["bricks"] * 64  # the bound does prune
    assert np.array_equal(plain.phi.view(np.uint8), culled.phi.view(np.uint8))
    for name in COLS:
        assert np.array_equal(plain.read(name).view(np.uint8), culled.read(name).view(np.uint8)), name
    if dtype == np.float64:
        ref = orc.sample_nodes(body, 3, g.x0, g.d, g.n)
        assert np.array_equal(culled.phi.view(np.uint64), np.asarray(ref).view(np.uint64))


def test_screened_folds_far_from_the_origin_and_with_degenerate_spheres():
    # the plain and culled folds screen primitives in single precision with margins that grow with
    # the coordinates' magnitude: a body around (1000, -500, 250), spheres of radius 0 and one of
    # negative radius (never screened), and a tiny smoothing; phi and every column bitwise the oracle
    rng = np.random.default_rng(17)
    c0 = np.array([1000.0, -500.0, 250.0])
    cen = c0 + rng.uniform(-0.5, 0.5, (40, 3))
    rad = rng.uniform(0.02, 0.2, 40)
    rad[3] = 0.0
    rad[7] = -0.01
    body = SDFSphereUnion(cen, rad, 1e-3)
    g = CartesianIsoGrid(tuple(c0 - 0.7), tuple(c0 + 0.7), n=(36, 40, 33))
    ref = orc.sample_nodes(body, 3, g.x0, g.d, g.n)
    cols = orc.mc_cells(ref, g.x0, g.d, g.n)
    for source in ("naive", "culled"):
        c = allocate_cache(g, MarchingCubesCutCell)
        c.set_phi_source(source)
        update_cache(c, body, g)
        assert np.array_equal(c.phi.view(np.uint64), np.asarray(ref).view(np.uint64)), source
        for name in COLS:
            got = np.ascontiguousarray(c.read(name)).view(np.uint8).ravel()
            want = np.ascontiguousarray(cols[name]).view(np.uint8).ravel()
            assert np.array_equal(got, want), (source, name)


def cut_corner_nodes(kind, n):
    """Node linear indices of every corner of every cut cell."""
    n0, n1, n2 = n
    cells = np.flatnonzero(kind == 2)
    i = cells % n0
    j = (cells // n0) % n1
    k = cells // (n0 * n1)
    m0, m1 = n0 + 1, n1 + 1
    out = []
    for dk in (0, 1):
        for dj in (0, 1):
            for di in (0, 1):
                out.append((i + di) + m0 * ((j + dj) + m1 * (k + dk)))
    return np.unique(np.concatenate(out))


def check_surrogate_against_plain(body, g, net, coarse=4):
    plain = allocate_cache(g, MarchingCubesCutCell)
    update_cache(plain, body, g)
    sur = allocate_cache(g, MarchingCubesCutCell)
    sur.set_phi_source("surrogate", net, coarse=coarse, band_cells=3.0)
    update_cache(sur, body, g)
    st = sur.phi_stats()
    for name in COLS:
        assert np.array_equal(plain.read(name).view(np.uint8), sur.read(name).view(np.uint8)), name
    pe, ps = plain.phi, sur.phi
    kind = plain.read("kind")
    corners = cut_corner_nodes(kind, g.n)
    assert np.array_equal(pe[corners].view(np.uint64), ps[corners].view(np.uint64))
    assert np.array_equal(pe > 0, ps > 0)  # every node's sign is the exact one
    diff = pe.view(np.uint64) != ps.view(np.uint64)
    h = max(g.d)
    assert np.all(np.abs(pe[diff]) > np.sqrt(3.0) * h)  # kept network values sit off the surface
    assert np.all(np.abs(ps[diff]) >= 3.0 * h)          # and outside the band
    return st, pe, ps, diff


def test_surrogate_update_is_bitwise_the_exact_path_even_untrained():
    body, g = union_case()
    net = Surrogate(nfreq=64, depth=4, sigma=8.0, seed=1)  # untrained: a poor predictor
    st, pe, ps, diff = check_surrogate_against_plain(body, g, net)
    print(st)
    assert st["nodes"] == pe.size
    assert st["exact"] + int(diff.sum()) <= st["nodes"]


def test_what_the_trained_network_gains():
    """Same body and grid, surrogate path with an untrained and with a trained network: kind and
    moments are bitwise the exact path's either way; the trained one needs far fewer exact
    evaluations (the culled-exact path needs one per node) and keeps far-field phi close."""
    body = csg_spheres(8, count=128, lo=-0.6, hi=0.6, rmin=0.05, rmax=0.15)
    g = CartesianIsoGrid((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), n=(96, 96, 96))
    untrained = Surrogate(nfreq=64, depth=4, sigma=8.0, seed=2)
    trained = Surrogate(nfreq=64, depth=4, sigma=8.0, seed=2)
    hist = trained.train(body, nsamples=1 << 21, band=0.05, epochs=16, batch=16384, lr=5e-3, seed=3)
    print("training loss", hist)
    su, peu, psu, du = check_surrogate_against_plain(body, g, untrained)
    st, pet, pst, dt = check_surrogate_against_plain(body, g, trained)
    fu = su["exact"] / su["nodes"]
    ft = st["exact"] / st["nodes"]
    err_t = float(np.sqrt(np.mean((pst[dt] - pet[dt]) ** 2))) if dt.any() else 0.0
    err_u = float(np.sqrt(np.mean((psu[du] - peu[du]) ** 2))) if du.any() else 0.0
    print(f"exact fraction: culled-exact 1.0, untrained {fu:.3f}, trained {ft:.3f}; "
          f"sign contradictions untrained {su['sign_fail']} trained {st['sign_fail']}; "
          f"far-field rms error untrained {err_u:.3e} trained {err_t:.3e} (h={g.d[0]:.3e})")
    assert ft < 0.5 * fu
    assert st["sign_fail"] < 0.05 * max(su["sign_fail"], 1)
    assert err_t < 0.25 * err_u and err_t < 4 * g.d[0]


def _plain_vs(source, body, g, dtype=np.float64, net=None):
    plain = allocate_cache(g, MarchingCubesCutCell, dtype)
    other = allocate_cache(g, MarchingCubesCutCell, dtype)
    if source == "surrogate":
        other.set_phi_source("surrogate", net, coarse=4)
    else:
        other.set_phi_source(source)
    update_cache(plain, body, g)
    update_cache(other, body, g)
    return plain, other


def test_culled_fold_with_more_kept_primitives_than_lds_holds():
    # 700 nearly coincident spheres: no bound separates them, every brick near them keeps all 700
    rng = np.random.default_rng(12)
    cen = 0.1 + 1e-3 * rng.standard_normal((700, 3))
    rad = 0.3 + 1e-3 * rng.standard_normal(700)
    body = SDFSphereUnion(cen, rad, 0.02)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(24, 20, 18))
    plain, culled = _plain_vs("culled", body, g)
    st = culled.phi_stats()
    assert st["kept_sum"] > 512 * 4  # several bricks over the 512-primitive LDS stage
    assert np.array_equal(plain.phi.view(np.uint64), culled.phi.view(np.uint64))
    for name in COLS:
        assert np.array_equal(plain.read(name).view(np.uint8), culled.read(name).view(np.uint8)), name


@pytest.mark.parametrize("n", [(1, 1, 1), (5, 3, 2), (17, 31, 15), (33, 16, 47)])
def test_culled_and_surrogate_on_ragged_grids(n):
    body = csg_spheres(6, count=40, lo=-0.6, hi=0.6, rmin=0.1, rmax=0.3)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=n)
    plain, culled = _plain_vs("culled", body, g)
    assert np.array_equal(plain.phi.view(np.uint64), culled.phi.view(np.uint64))
    net = Surrogate(seed=3)
    st, pe, ps, diff = check_surrogate_against_plain(body, g, net, coarse=4)
    assert st["nodes"] == int(np.prod(np.array(n) + 1))


def test_single_primitive_union_and_moved_grid():
    body = SDFSphereUnion(np.array([[0.05, -0.1, 0.2]]), np.array([0.45]), 0.02)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(40, 36, 44))
    for shift in (0.0, 0.37 * g.d[0], -2.5 * g.d[0]):
        gm = g.moved((shift, 0.5 * shift, -shift))
        plain, culled = _plain_vs("culled", body, gm)
        assert np.array_equal(plain.phi.view(np.uint64), culled.phi.view(np.uint64))
        for name in COLS:
            assert np.array_equal(plain.read(name).view(np.uint8), culled.read(name).view(np.uint8)), name


def test_exact_zero_nodes_under_every_phi_route():
    # a dyadic grid and a sphere centred on a node with a radius of four cells: six nodes read phi
    # exactly 0, which the cells nudge to +eps (outside, marching_cubes.jl:225-243); the brick pass's
    # sign bitmap must classify them the same way for the z-march that reads it
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(32, 32, 32))
    body = SDFSphereUnion(np.array([[0.0, 0.0, 0.0], [0.75, 0.75, 0.75]]), np.array([0.25, 0.1]), 0.02)
    plain = update_cache(allocate_cache(g, MarchingCubesCutCell), body, g)
    assert int(np.count_nonzero(plain.phi == 0.0)) == 6
    net = Surrogate(seed=3)  # untrained: the exact fallback decides the nodes near the surface
    for source in ("culled", "surrogate"):
        for tune in ((1, 2, False), (1, 16, False), (2, 16, False)):
            c = allocate_cache(g, MarchingCubesCutCell)
            c.set_tuning(*tune)
            c.set_phi_source(source, net if source == "surrogate" else None, coarse=4)
            update_cache(c, body, g)
            for name in COLS:
                assert np.array_equal(plain.read(name).view(np.uint8), c.read(name).view(np.uint8)), \
                    (source, tune, name)


def test_surrogate_on_a_float32_cache():
    body, g = union_case(n=(48, 40, 36))
    net = Surrogate(seed=5)
    net.train(body, nsamples=1 << 18, epochs=4, batch=16384)
    plain, sur = _plain_vs("surrogate", body, g, np.float32, net)
    for name in COLS:
        assert np.array_equal(plain.read(name).view(np.uint8), sur.read(name).view(np.uint8)), name
    pe, ps = plain.phi, sur.phi
    assert np.array_equal(pe > 0, ps > 0)
    corners = cut_corner_nodes(plain.read("kind"), g.n)
    assert np.array_equal(pe[corners].view(np.uint32), ps[corners].view(np.uint32))
    # the phi pass and the z-march reading its sign bitmap, in Float32 too
    sur.set_tuning(1, 4, False)
    update_cache(sur, body, g)
    for name in COLS:
        assert np.array_equal(plain.read(name).view(np.uint8), sur.read(name).view(np.uint8)), name


def test_phi_source_argument_errors():
    from cutcell import ArgumentError, MarchingSquaresCutCell, PLICCutCell, Unsupported
    g3 = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(8, 8, 8))
    c = allocate_cache(g3, MarchingCubesCutCell)
    with pytest.raises(ArgumentError):
        c.set_phi_source("surrogate", None)
    net = Surrogate(seed=1)
    with pytest.raises(ArgumentError):
        c.set_phi_source("surrogate", net, coarse=0)
    with pytest.raises(ArgumentError):
        c.set_phi_source("surrogate", net, band_cells=-1.0)
    p = allocate_cache(g3, PLICCutCell)
    with pytest.raises(Unsupported):
        p.set_phi_source("culled")
    g2 = CartesianIsoGrid((-1.0, -1.0), (1.0, 1.0), n=(8, 8))
    m = allocate_cache(g2, MarchingSquaresCutCell)
    with pytest.raises(Unsupported):
        m.set_phi_source("culled")


def test_fused_phi_statistics_equal_the_brick_pass():
    # the fused stream pass counts every node once (the owned ones of its work items) and the cull
    # pass counts the certificates: the same counters as the separate brick pass, phi the same bits
    body = csg_spheres(9, count=48, lo=-0.6, hi=0.6, rmin=0.08, rmax=0.25)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(45, 38, 70))
    net = Surrogate(seed=4)
    net.train(body, nsamples=1 << 16, epochs=3, batch=4096, seed=1)
    for source in ("culled", "surrogate"):
        out = []
        for tune in ((1, 8, 0), (2, 8, 0), (2, 16, 1)):
            c = allocate_cache(g, MarchingCubesCutCell)
            c.set_tuning(*tune)
            c.set_phi_source(source, net if source == "surrogate" else None, coarse=4)
            update_cache(c, body, g)
            out.append((c.phi_stats(), c.phi.copy()))
        for st, phi in out[1:]:
            assert st == out[0][0], (source, st, out[0][0])
            assert np.array_equal(phi.view(np.uint64), out[0][1].view(np.uint64)), source
        assert out[0][0]["nodes"] == int(np.prod(np.array(g.n) + 1))


@pytest.mark.parametrize("tune", [(0, 16, 0), (1, 1, 0), (1, 3, 0), (1, 16, 1), (2, 64, 0), (2, 100, 0)])
def test_culled_and_surrogate_under_every_stream_shape(tune):
    """Every launch shape of the stream pass reads what the brick pass wrote (phi, or for the z-march
    reading a field the sign bitmap beside it), update after update."""
    body = csg_spheres(9, count=48, lo=-0.6, hi=0.6, rmin=0.08, rmax=0.25)
    g = CartesianIsoGrid((-1.0,) * 3, (1.0,) * 3, n=(40, 36, 150))  # several 64-layer slabs
    plain = allocate_cache(g, MarchingCubesCutCell)
    update_cache(plain, body, g)
    net = Surrogate(seed=4)
    for source in ("culled", "surrogate"):
        c = allocate_cache(g, MarchingCubesCutCell)
        c.set_tuning(*tune)
        c.set_phi_source(source, net if source == "surrogate" else None, coarse=4)
        for step in range(2):  # the second update reuses the streams and events
            update_cache(c, body, g)
        for name in COLS:
            assert np.array_equal(plain.read(name).view(np.uint8), c.read(name).view(np.uint8)), (source, name)
        if source == "culled":
            assert np.array_equal(plain.phi.view(np.uint64), c.phi.view(np.uint64))
