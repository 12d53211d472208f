"""The benchmark workloads at their full sizes.

* The 1M-cell channel (100^3 jittered, randomly renumbered hex box; CG +
  DILU to rtol 1e-10 for p): one SIMPLE iteration on the device against the
  oracle's (C++, OpenMP) within 1e-8 relative L2, the channel tolerance of
  the smaller parity tests; and three iterations with FP32 inner solves in
  FP64 refinement against FP64 within 1e-8.
* The 256 batched cavities (4.3 M cells): three cases -- the smallest, the
  median and the largest -- are bit-identical to the same cases run alone,
  after two SIMPLE iterations of the whole batch.
"""
import numpy as np
import pytest

import bench
from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def channel_mesh(x):
    return x.mesh.box3d(100, 100, 100, jitter=0.1, seed_jitter=bench.SEED_JITTER,
                        seed_perm=bench.SEED_PERM)


def _device_run(x, mesh, model, config, iters):
    H = x.hip
    sim = H.Simple(H.DeviceMesh(x.api._context(0), mesh), x.api.simple_setup(model, config))
    sim.set_U(model.momentum.U)
    sim.set_p(model.momentum.p)
    out = sim.run(iters)
    return out, sim.get("U"), sim.get("p")


def test_channel_1m_matches_oracle(x, A, O, channel_mesh):
    api = x.api
    model, config = bench.channel(api, channel_mesh, 1)
    out, U, p = _device_run(x, channel_mesh, model, config, 1)
    ref = O.OracleMesh(channel_mesh, A).simple_run(api.simple_setup(model, config),
                                                   model.momentum.U, model.momentum.p, 1)
    eU, ep = rel(U, ref["U"]), rel(p, ref["p"])
    print(f"1M channel, 1 iteration: U {eU:.2e} p {ep:.2e}; p CG iterations "
          f"{out['krylov_iters'][3, 0]} (oracle {ref['krylov_iters'][3, 0]})")
    assert eU < 1e-8 and ep < 1e-8
    assert abs(int(out["krylov_iters"][3, 0]) - int(ref["krylov_iters"][3, 0])) <= 10


def test_channel_1m_mixed_precision_within_1e8(x, channel_mesh):
    api = x.api
    res = {}
    for prec in ("fp64", "fp32-ir", "fp32-precon"):
        model, config = bench.channel(api, channel_mesh, 3, precision=prec)
        res[prec] = _device_run(x, channel_mesh, model, config, 3)
    _, U64, p64 = res["fp64"]
    for prec in ("fp32-ir", "fp32-precon"):
        _, U32, p32 = res[prec]
        eU, ep = rel(U32, U64), rel(p32, p64)
        print(f"1M channel, 3 iterations, {prec} vs FP64: U {eU:.2e} p {ep:.2e}")
        assert eU < 1e-8 and ep < 1e-8


def test_batch_256_cases_alone_equal_mix(x):
    api, H = x.api, x.hip
    sizes, Re = bench.batch_cases(256)
    meshes = [x.mesh.cavity2d(int(m)) for m in sizes]
    model, config = bench.cavity(api, meshes[0], 2)
    setup = api.simple_setup(model, config, pref=0.0)
    ctx = api._context(0)
    dm = H.DeviceMesh.batch(ctx, meshes)
    sim = H.Simple(dm, setup)
    sim.set_case_nu(1.0 / Re)
    out = sim.run_batch(2)
    assert out["done"] == 2
    U, p = sim.get("U"), sim.get("p")
    off = np.cumsum([0] + [m.n_cells for m in meshes])
    order = np.argsort(sizes, kind="stable")
    for k in (order[0], order[len(order) // 2], order[-1]):
        d1 = H.DeviceMesh.batch(ctx, [meshes[k]])
        s1 = H.Simple(d1, setup)
        s1.set_case_nu([1.0 / Re[k]])
        o1 = s1.run_batch(2)
        sl = slice(off[k], off[k + 1])
        assert np.array_equal(U[:, sl], s1.get("U")), f"case {k} U"
        assert np.array_equal(p[sl], s1.get("p")), f"case {k} p"
        assert np.array_equal(out["residuals"][k], o1["residuals"][0])
