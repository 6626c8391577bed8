"""The exterior field as a matrix: the P x N radiation operator assembled by the
collocation kernel (field points as rows off the surface) and applied to all
solutions with one product.

Checks:
  * the operator's entries against the oracle's rows at the same points
    (assembled as CHIEF-position rows): max |dF| <= 2e-5 max |F|;
  * the field from the operator against the oracle's field (genExtPressure,
    FP64) for the same surface solutions, for 1 and for 20 solutions, rigid
    and impedance boundary conditions: <= 2e-5 of the largest value;
  * the operator path against the direct field kernel: <= 2e-5;
  * applying an operator assembled at another wavenumber is refused.
"""
import numpy as np
import pytest

import bem_gpu_clean_cpp_amd as bem
import bemcases as C
import oracle_py as O

pytestmark = pytest.mark.gpu

RHO, SPEED = float(np.float32(1.21)), float(np.float32(343.0))


@pytest.fixture(autouse=True, scope="module")
def _setup():
    bem.quad_init(6)
    O.quad_init(6)
    bem.set_medium(RHO, SPEED)
    yield
    bem.set_medium(1.2041, 343.21)


def _mesh(impedance):
    m = bem.Mesh().icosphere(3, 0.1)
    if impedance:
        m.set_bc(complex(0.3, 0.1), 1.0, complex(0.2, -0.4))
    else:
        m.set_bc(0, 1, 0)
    m.set_chief(C.ball_points(31, 8, 0.05))
    return m


@pytest.mark.parametrize("impedance", [False, True])
@pytest.mark.parametrize("R", [1, 20])
def test_field_operator_matches_oracle(R, impedance):
    m = _mesh(impedance)
    k = 20.0
    dirs = C.sphere_points(17, R, 1.0)
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, dirs, 1.0)
    s.solve(bem.SOLVE_GMRES, refine_tol=1e-10)
    x = s.solution()
    xs = C.sphere_points(5, 96, 0.5)
    s.field_matrix(xs)
    out = s.field_apply()
    direct = s.field(xs)
    p, e, bc = C.mesh_arrays(m)
    inc = np.exp(-1j * k * xs.astype(np.float64) @ np.asarray(dirs, np.float64).T)
    fo = O.field(k, p.astype(np.float64), p.shape[0], e, bc, xs, x, inc=inc, dens=RHO, speed=SPEED)
    scale = np.max(np.abs(fo))
    assert out.shape == (96, R)
    assert np.max(np.abs(out - fo)) / scale < 2e-5
    assert np.max(np.abs(out - direct)) / scale < 2e-5


def test_field_operator_entries_match_oracle_rows():
    m = _mesh(True)
    k = 12.0
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    xs = C.sphere_points(9, 40, 0.3)
    s.field_matrix(xs)
    F = s.get_field_matrix()
    p, e, bc = C.mesh_arrays(m)
    pts = np.vstack([p, xs]).astype(np.float64)
    Ao, _ = O.assemble(k, pts, p.shape[0], e, bc, dens=RHO, speed=SPEED)
    Fo = Ao[p.shape[0]:]
    assert F.shape == Fo.shape
    assert np.max(np.abs(F - Fo)) / np.max(np.abs(Fo)) < 2e-5


def test_field_operator_stale_wavenumber_refused():
    m = _mesh(False)
    s = bem.Solver(m)
    s.assemble(10.0, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    s.solve()
    s.field_matrix(C.sphere_points(1, 8, 0.5))
    s.assemble(11.0, bem.PREC_FP32)
    s.solve()
    with pytest.raises(bem.BemError):
        s.field_apply()


def test_plain_config_field_matches_sphere_series():
    """The small plain workload: rigid sphere r = 0.1 m (icosphere s3), 171 Hz,
    8 CHIEF points, unit plane wave along +z, QR least squares; the field at 64
    seeded points on r = 1 m against the exterior series.  The scattered part
    (total minus incident) carries the comparison: <= 2 % of its size."""
    a, k = 0.1, float(np.float32(2 * np.pi * 171.0 / SPEED))
    m = bem.Mesh().icosphere(3, a)
    m.set_bc(0, 1, 0)
    m.set_chief(C.ball_points(101, 8, a / 2))
    s = bem.Solver(m)
    s.assemble(k, bem.PREC_FP32)
    s.set_rhs(bem.RHS_PLANE, [[0, 0, 1]], 1.0)
    st = s.solve(bem.SOLVE_QR, refine_tol=1e-10, max_refine=10)
    assert st["converged"] == 1
    xs = C.sphere_points(101, 64, 1.0)
    s.field_matrix(xs)
    p = s.field_apply()[:, 0]
    inc = np.exp(-1j * k * xs[:, 2].astype(np.float64))
    ser = C.rigid_sphere_field(a, k, xs)
    scat, scat_ser = p - inc, ser - inc
    err = np.linalg.norm(scat - scat_ser) / np.linalg.norm(scat_ser)
    print("scattered-field error vs series %.3g" % err)
    assert err < 0.02
