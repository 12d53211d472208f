"""Mesh ingest on the device: the CSR pattern, slot maps and DILU levels
built by kernels equal the host build element for element, satisfy the
definitions (test_ingest.check_connectivity) and match the oracle's CSR; a
device mesh created from them runs as before; and the build time of a
250k-cell step mesh is reported both ways."""
import numpy as np
import pytest

from test_ingest import check_connectivity

pytestmark = pytest.mark.gpu

MESHES = ["laplace3", "cavity40", "box12", "cavity_hex", "step", "trig40", "chain5", "box20"]


@pytest.fixture(scope="module")
def ctx(x):
    return x.hip.Context(0)


def _same(Kd, Kh):
    assert Kd.levels == Kh.levels and Kd.nnz == Kh.nnz
    for name in Kh.NAMES:
        assert np.array_equal(Kd[name], Kh[name]), name


@pytest.mark.parametrize("name", MESHES)
def test_device_connectivity_equals_host_and_oracle(x, A, O, ctx, meshes, name):
    mesh = meshes(name)
    Kd = x.hip.Connectivity(ctx, mesh, on_device=True)
    Kh = x.hip.Connectivity(ctx, mesh, on_device=False)
    _same(Kd, Kh)
    orp, ocv = O.OracleMesh(mesh, A).csr()
    check_connectivity(Kd, mesh, orp, ocv)
    dm = x.hip.DeviceMesh(ctx, mesh)  # built through the device form
    assert dm.levels == Kd.levels and dm.nnz == Kd.nnz
    rp, cv = dm.csr_pattern()
    assert np.array_equal(rp, Kd["rowptr"]) and np.array_equal(cv, Kd["colval"])


def test_device_ingest_250k_step_and_1m_box(x, ctx, meshes):
    """Large meshes: a 250k-cell 2D step (about 1000 DILU levels, so about as
    many propagation rounds) and the 1M-cell jittered, renumbered box."""
    import bench
    M = x.mesh
    step = M.backward_facing_step_2d(0.45, 1.0, 1.0, 4.0, nx_in=234, nx_out=937, ny_low=105,
                                     ny_up=129)
    assert step.n_cells > 245000
    box = M.box3d(100, 100, 100, jitter=0.1, seed_jitter=bench.SEED_JITTER,
                  seed_perm=bench.SEED_PERM)
    for name, mesh in (("step 250k", step), ("box 1M", box)):
        Kd = x.hip.Connectivity(ctx, mesh, on_device=True)
        Kd = x.hip.Connectivity(ctx, mesh, on_device=True)  # second build: warm
        Kh = x.hip.Connectivity(ctx, mesh, on_device=False)
        _same(Kd, Kh)
        check_connectivity(Kd, mesh)
        print(f"{name}: {mesh.n_cells} cells, levels {Kd.levels}: device kernels "
              f"{Kd.device_ms:.2f} ms, device build {Kd.wall_ms:.1f} ms (transfers included), "
              f"host build {Kh.wall_ms:.1f} ms")


def test_mesh_creation_same_with_host_ingest(x, A, ctx, meshes, monkeypatch):
    """A SIMPLE run on a mesh built with the device ingest equals the run on
    the same mesh built with the host loops, bit for bit."""
    from conftest import cavity_case
    mesh = meshes("cavity40")
    runs = []
    for host in (False, True):
        if host:
            monkeypatch.setenv("XCAL_HOST_INGEST", "1")
        dm = x.hip.DeviceMesh(ctx, mesh)
        monkeypatch.delenv("XCAL_HOST_INGEST", raising=False)
        model, config = cavity_case(x.api, mesh, iterations=20)
        sim = x.hip.Simple(dm, x.api.simple_setup(model, config))
        sim.set_U(model.momentum.U)
        sim.set_p(model.momentum.p)
        sim.run(20)
        runs.append((sim.get("U"), sim.get("p")))
    assert np.array_equal(runs[0][0], runs[1][0]) and np.array_equal(runs[0][1], runs[1][1])
