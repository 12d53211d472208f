"""Pinning the oracle's k-omega functions (RANS_kOmega.jl:175-235,
RANS_functions.jl, RANS_tensor_algebra.jl:68) by what the model and the law
of the wall fix, on the CPU.  Each test is chosen so that a plausible slip --
alpha1, beta1, beta* or a sigma swapped, a wall-function branch inverted, a
wrong constant or exponent -- fails it:

* production 2 S:S of simple shear, pure rotation and pure strain, and of the
  Green-Gauss gradient of a linear field;
* the log-law intersection y+_lam of kappa 0.41, E 9.8;
* the wall viscosity reproducing the wall shear u*^2 of the log law, and the
  wall production equal to the log-layer equilibrium u*^3/(kappa y);
* the omega wall values (Wilcox's sublayer 6 nu/(beta1 y^2), log layer
  u*/(sqrt(C_mu) kappa y)) as identity rows;
* bound!'s clipping, nut = k/omega, and which constant feeds which term;
* one coupled step: omega, U and p do not depend on beta* or sigma_k, and k
  does not depend on alpha1, beta1 or sigma_omega (the k equation uses the
  old omega, RANS_kOmega.jl:198-206)."""
import numpy as np
import pytest
from scipy.optimize import brentq

from conftest import komega_bfs_case

NU = 1e-5
KAPPA, E, CMU = 0.41, 9.8, 0.09


def patch_faces(mesh, name):
    k = mesh.patch_index(name)
    f0, f1 = mesh.patch_range[2 * k] - 1, mesh.patch_range[2 * k + 1]
    return np.arange(f0, f1), mesh.boundary_cellsID[f0:f1] - 1


def bcs(A, mesh, **kinds):
    return [(kinds.get(name, A.BC_ZEROGRADIENT), i, 0.0)
            for i, name in enumerate(mesh.patch_names)]


def test_production_closed_forms(A, O, meshes):
    """2 S:S: simple shear dU/dy = g gives g^2; pure rotation gives 0; pure
    strain diag(a, -a, 0) gives 4 a^2.  Pw = alpha1 2S:S, Pk = nut 2S:S."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    n = m.n_cells
    g, a, alpha1 = 3.0, 0.7, 0.52
    nut = np.linspace(1e-4, 2e-3, n)
    for grad, expect in (({1: g}, g * g), ({1: g, 3: -g}, 0.0), ({0: a, 4: -a}, 4 * a * a)):
        gU = [np.zeros(n) for _ in range(9)]
        for idx, v in grad.items():  # gU[3*j + q] = dU_j/dx_q
            gU[idx][:] = v
        Pk, Pw = om.kw_production(gU, nut, alpha1)
        assert np.allclose(Pw, alpha1 * expect, rtol=1e-14, atol=1e-300)
        assert np.allclose(Pk, nut * expect, rtol=1e-14, atol=1e-300)


def test_production_of_green_gauss_linear_field(A, O, meshes):
    """U = (g y, 0, 0) on a uniform mesh: Green-Gauss of the exact face values
    is the exact gradient in every cell, so 2 S:S = g^2 everywhere."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    g = 2.5
    yf = m.face_centre.reshape(-1, 3)[:, 1]
    t = om.green_gauss_tensor(g * yf, np.zeros(m.n_faces), np.zeros(m.n_faces))
    Pk, Pw = om.kw_production(t, np.ones(m.n_cells), 0.52)
    assert np.allclose(Pk, g * g, rtol=1e-12)


def test_yplus_laminar_is_the_log_law_intersection(O):
    """y+_lam solves y = ln(E y)/kappa (viscous sublayer u+ = y+ meets the
    log law u+ = ln(E y+)/kappa): about 11.53 for kappa 0.41, E 9.8."""
    c = O.OracleMesh.kw_wall_constants(NU)
    y_star = brentq(lambda y: y - np.log(E * y) / KAPPA, 5.0, 30.0, xtol=1e-14)
    assert abs(c["yplus_lam"] - y_star) < 1e-5
    assert 11.4 < c["yplus_lam"] < 11.6
    assert c["cmu25"] == pytest.approx(CMU ** 0.25, rel=1e-15)


def _wall_case(A, O, meshes):
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    faces, cells = patch_faces(m, "bottom")
    y = m.face_delta[faces]
    # k per bottom cell: half in the viscous sublayer (y+ ~ 1.7), half in the
    # log layer (y+ ~ 170)
    k = np.full(m.n_cells, 1e-2)
    k[cells[::2]] = 1e-6
    ustar = CMU ** 0.25 * np.sqrt(k[cells])
    yplus = ustar * y / NU
    return m, om, faces, cells, y, k, ustar, yplus


def test_nut_wall_reproduces_log_law_shear(A, O, meshes):
    """correct_eddy_viscosity!: above y+_lam, (nu + nut_w) U_P / y = u*^2 for
    the log-law velocity U_P = u*/kappa ln(E y+); below it nut_w = 0.
    Interior faces carry the interpolated k/omega."""
    m, om, faces, cells, y, k, ustar, yplus = _wall_case(A, O, meshes)
    lam = O.OracleMesh.kw_wall_constants(NU)["yplus_lam"]
    omega = np.linspace(50.0, 80.0, m.n_cells)
    bc = bcs(A, m, bottom=A.BC_NUTWALL)
    nut, nutf = om.kw_nut(bc, NU, k, omega)
    assert np.array_equal(nut, k / omega)
    log = yplus > lam
    assert log.any() and (~log).any()
    UP = ustar / KAPPA * np.log(E * yplus)
    tau = (NU + nutf[faces]) * UP / y
    assert np.allclose(tau[log], ustar[log] ** 2, rtol=1e-12)
    assert np.all(nutf[faces][~log] == 0.0)
    own = m.face_owner.reshape(-1, 2) - 1
    w = m.face_weight
    i = np.arange(m.n_bfaces, m.n_faces)
    assert np.allclose(nutf[i], w[i] * nut[own[i, 0]] + (1 - w[i]) * nut[own[i, 1]],
                       rtol=1e-14)


def test_wall_production_is_log_layer_equilibrium(A, O, meshes):
    """correct_production!: with the wall-parallel cell velocity of the log
    law, Pk = u*^3 / (kappa y) (production equals dissipation in the log
    layer); the wall-normal velocity component does not contribute; below
    y+_lam, Pk = 0.  Cells of other patches keep their value."""
    m, om, faces, cells, y, k, ustar, yplus = _wall_case(A, O, meshes)
    lam = O.OracleMesh.kw_wall_constants(NU)["yplus_lam"]
    U = np.zeros((3, m.n_cells))
    U[0, cells] = ustar / KAPPA * np.log(E * yplus)  # parallel to the bottom wall
    U[1, cells] = 0.3                                # normal to it
    Pk0 = np.full(m.n_cells, 7.0)
    Pk = om.kw_wall_production(bcs(A, m, bottom=A.BC_KWALL), NU, U, k, Pk0)
    log = yplus > lam
    assert np.allclose(Pk[cells][log], ustar[log] ** 3 / (KAPPA * y[log]), rtol=1e-12)
    assert np.all(Pk[cells][~log] == 0.0)
    other = np.setdiff1d(np.arange(m.n_cells), cells)
    assert np.all(Pk[other] == 7.0)


def test_omega_wall_rows_are_identity_with_wall_value(A, O, meshes):
    """constrain!: each wall cell's row becomes (0 ... 1 ... 0) with b the
    sublayer value 6 nu / (beta1 y^2), beta1 = 0.075 of the wall function,
    below y+_lam and the log-law value with omega sqrt(C_mu) kappa y = u*
    above it; other rows are untouched."""
    m, om, faces, cells, y, k, ustar, yplus = _wall_case(A, O, meshes)
    lam = O.OracleMesh.kw_wall_constants(NU)["yplus_lam"]
    rp, cv = om.csr()
    rng = np.random.default_rng(1)
    nz = rng.uniform(1.0, 2.0, om.nnz)
    b = rng.uniform(1.0, 2.0, m.n_cells)
    A2, b2 = om.kw_constrain(bcs(A, m, bottom=A.BC_OMEGAWALL), NU, k, nz, b)
    log = yplus > lam
    for c, wv, lg, ys, us in zip(cells, b2[cells], log, y, ustar):
        row = slice(rp[c], rp[c + 1])
        assert np.array_equal(A2[row], (cv[row] == c).astype(float))
        if lg:
            assert wv * np.sqrt(CMU) * KAPPA * ys == pytest.approx(us, rel=1e-12)
        else:
            assert wv == pytest.approx(6 * NU / (0.075 * ys * ys), rel=1e-14)
    other = np.setdiff1d(np.arange(m.n_cells), cells)
    for c in other:
        assert np.array_equal(A2[rp[c]:rp[c + 1]], nz[rp[c]:rp[c + 1]])
    assert np.array_equal(b2[other], b[other])


def test_destruction_and_diffusivity_take_their_own_constants(A, O, meshes):
    """Dw = beta1 omega and Dk = beta* omega (implicit destruction
    beta omega^2, beta* k omega), mueff_omega = nu + sigma_omega nut_f and
    mueff_k = nu + sigma_k nut_f; four distinct constants."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    omega = np.linspace(1.0, 2.0, m.n_cells)
    nutf = np.linspace(1e-3, 3e-3, m.n_faces)
    Dw, Dk, mw, mk = om.kw_coeffs(omega, nutf, NU, 0.072, 0.09, 0.5, 0.3)
    assert np.allclose(Dw / omega, 0.072, rtol=1e-14)
    assert np.allclose(Dk / omega, 0.09, rtol=1e-14)
    assert np.allclose((mw - NU) / nutf, 0.5, rtol=1e-10)
    assert np.allclose((mk - NU) / nutf, 0.3, rtol=1e-10)


def test_bound_clips_to_the_neighbour_average(A, O, meshes):
    """bound!: a non-negative value v stays max(v, eps); a negative value
    becomes the average over its internal-face neighbours of max(v_n, eps)
    (eps when that is smaller)."""
    m = meshes("cavity16")
    om = O.OracleMesh(m, A)
    eps = np.finfo(np.float64).eps
    rng = np.random.default_rng(3)
    v = rng.uniform(0.5, 1.5, m.n_cells)
    v[rng.choice(m.n_cells, 20, replace=False)] *= -1.0
    v[0] = 0.0
    out = om.bound(v)
    rngs = m.cell_faces_range.reshape(-1, 2) - 1
    nb = m.cell_neighbours - 1
    for i in range(m.n_cells):
        if v[i] >= 0:
            assert out[i] == max(v[i], eps)
        else:
            nbr = nb[rngs[i, 0]:rngs[i, 1] + 1]
            assert out[i] == pytest.approx(max(np.maximum(v[nbr], eps).mean(), eps), rel=1e-15)


def _run(A, O, mesh, api, coeffs, iterations):
    """The reference example's k-omega case from a uniform state, with some
    model coefficients replaced, through the oracle's SIMPLE loop."""
    model, config = komega_bfs_case(api, mesh, iterations, "DILU")
    t = model.turbulence
    for kk, vv in coeffs.items():
        setattr(t, kk, vv)
    for f, v in ((model.momentum.U, [1.5, 0.0, 0.0]), (t.k, 1.0), (t.omega, 1000.0),
                 (t.nut, 1e-3)):
        api.initialise(f, v)
    setup = api.simple_setup(model, config)
    ts = api.turbulence_setup(model, config)
    return O.OracleMesh(mesh, A).simple_run(setup, model.momentum.U, model.momentum.p,
                                            iterations, turb=ts, k=t.k, omega=t.omega, nut=t.nut)


def test_komega_step_takes_each_constant_where_the_model_puts_it(x, A, O, meshes):
    """One SIMPLE k-omega iteration on the reference example's step mesh:
    changing beta* leaves omega, U and p unchanged and changes k; changing
    alpha1 and beta1 leaves k unchanged and changes omega (the k equation
    takes the old omega)."""
    api, mesh = x.api, meshes("bfs10")
    b, a1, a2 = (_run(A, O, mesh, api, c, 1)
                 for c in ({}, dict(beta_star=0.11), dict(alpha1=0.6, beta1=0.09)))
    for f in ("omega", "U", "p"):
        assert np.array_equal(a1[f], b[f]), f
    assert not np.array_equal(a1["k"], b["k"])
    assert np.array_equal(a2["k"], b["k"])
    assert not np.array_equal(a2["omega"], b["omega"])


def test_komega_diffusivities_take_their_own_sigma(x, A, O, meshes):
    """The face eddy viscosity starts at zero, so the sigmas act from the
    second iteration on.  After two iterations, changing sigma_k leaves omega,
    U and p unchanged and changes k; changing sigma_omega leaves k, U and p
    unchanged and changes omega."""
    api, mesh = x.api, meshes("bfs10")
    b, a1, a2 = (_run(A, O, mesh, api, c, 2)
                 for c in ({}, dict(sigma_k=0.8), dict(sigma_omega=0.9)))
    for f in ("omega", "U", "p"):
        assert np.array_equal(a1[f], b[f]), f
    assert not np.array_equal(a1["k"], b["k"])
    for f in ("k", "U", "p"):
        assert np.array_equal(a2[f], b[f]), f
    assert not np.array_equal(a2["omega"], b["omega"])
