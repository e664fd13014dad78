"""GPU tests of the drop-in API around the operator: contrast edits in place,
the flat-vector forms, the host-driven BiCGStab on a foreign operator, the
solve judge and the stated ``persistent_nbytes`` meaning.

Oracles: the dense operator ``A = I - B chi`` assembled cell pair by cell pair
(src/em3d/dense.py:35-209 restated in oracle/vie.py), the oracle's FFT
operator (src/em3d/operator.py:353-409), its BiCGStab trajectory
(src/em3d/solvers/bicgstab.py:12-211) and true residual
(src/em3d/experiments/stationary_validation.py:172-180).
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def em():
    import em3d_amd
    return em3d_amd


def _be(em, prec):
    return em.Backend.hip(em.Precision.DOUBLE if prec == "double" else em.Precision.SINGLE)


def _rand(shape, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _host(t):
    return t.detach().cpu().numpy()


def _sphere_problem(em, be, N, k, eps=2.0, radius=0.3):
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    chi = em.geometry.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=eps,
                                             geometry=em.geometry.Ellipsoid(center=(0, 0, 0), radii=(radius,) * 3))
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return em.Problem(grid=grid, eps_tensor=chi, wave=wave, k0=k, volume=1.0)


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
def test_in_place_contrast_edit_reaches_every_path(em, prec, tol):
    """Editing problem.eps_tensor in place (iso -> diag -> full) changes what
    matvec, rmatvec, the device solve and the complex64 companion apply."""
    from oracle import vie
    N, k = (8, 6, 4), 3.0
    _say(f"[contrast edit] N={N} prec={prec}")
    be = _be(em, prec)
    prob = _sphere_problem(em, be, N, k)
    op = em.Operator(prob)
    Kh = vie.kernel_hat(N, (1, 1, 1), k)
    u = _rand((3,) + N, 3)
    ud = be.array(u.astype(be.complex_dtype))
    assert op.contrast_kind == "isotropic"
    comp = op.single_precision_companion() if prec == "double" else None
    edits = [lambda c: c[1, 1].mul_(1.5 - 0.2j), lambda c: c[0, 2].add_(0.1j)]
    for kind, edit in zip(("diagonal", "full"), edits):
        edit(prob.eps_tensor)
        chi = _host(prob.eps_tensor).astype(np.complex128)
        assert op.contrast_kind == kind
        assert _rel(_host(op.matvec(ud)), vie.em_matvec(chi, u, Kh)) < tol
        assert _rel(_host(op.rmatvec(ud)), vie.em_rmatvec(chi, u, Khat=Kh)) < tol
        if comp is not None:
            comp2 = op.single_precision_companion()
            assert comp2 is not comp
            comp = comp2
            y32 = _host(comp2.matvec(ud.to(comp2.backend.precision.torch_complex_dtype)))
            assert _rel(y32, vie.em_matvec(chi, u, Kh)) < 1e-6
        res = em.BiCGStab(em.SolverConfig(max_iter=100, rtol=1e-8 if prec == "double" else 1e-5)).solve(op, ud)
        assert res.status == "converged"
        true = _rel(vie.em_matvec(chi, _host(res.u).astype(np.complex128), Kh), u)
        assert true < (1e-7 if prec == "double" else 1e-4)


def test_in_place_eta_edit(em):
    from oracle import vie
    N, L, k = (6, 5, 4), (1.0, 0.9, 0.8), 4.0
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    eta = em.acoustics.eta_sphere(grid, center=(0, 0, 0), radius=0.3, eta_inside=2.0 + 0.1j)
    prob = em.acoustics.make_acoustic_problem(grid, eta, k0=k)
    op = em.acoustics.AcousticOperator(prob)
    u = _rand(N, 4)
    ud = be.array(u)
    op.matvec(ud)
    prob.eta[1:3].mul_(1.7 + 0.3j)
    chi = _host(prob.eta) - 1.0
    assert _rel(_host(op.matvec(ud)), vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k))) < 1e-12
    assert _rel(_host(op.rmatvec(ud)),
                vie.acoustic_rmatvec(chi, u, vie.acoustic_kernel_hat(N, L, k, adjoint=True))) < 1e-12


def test_flat_forms_match_dense_operator(em):
    """matvec_flat / rmatvec_flat / iteration_matvec_flat against the dense
    A = I - B chi and its adjoint (cell-major, component-minor order)."""
    from oracle import vie
    N, L, k = (4, 3, 2), (1.0, 0.8, 0.6), 2.1
    _say("[flat forms] dense 72 x 72")
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    chi = _rand((3, 3) + N, 7) * 0.3
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=be.array(_rand((3,) + N, 1)), k0=k, volume=1.0)
    op = em.Operator(prob)
    A = vie.dense_operator_matrix(N, L, (0, 0, 0), k, chi)
    v = _rand(3 * int(np.prod(N)), 8)
    vd = be.array(v)
    assert op.shape == A.shape
    assert _rel(_host(op.matvec_flat(vd)), A @ v) < 1e-12
    assert _rel(_host(op.rmatvec_flat(vd)), A.conj().T @ v) < 1e-12
    mu = 1.3 - 0.4j
    assert _rel(_host(op.iteration_matvec_flat(vd, mu=mu)), v - (A @ v) / mu) < 1e-12
    with pytest.raises(ValueError, match="mu"):
        op.iteration_matvec_flat(vd, mu=0.0)
    with pytest.raises(ValueError):
        op.matvec_flat(vd[:-3])


class _Wrapped:
    """A foreign operator: only the reference's duck-typed protocol
    (``backend`` and ``matvec``), so BiCGStab takes the host-driven loop."""

    def __init__(self, op):
        self._op = op
        self.backend = op.backend
        self.calls = 0

    def matvec(self, x):
        self.calls += 1
        return self._op.matvec(x)


def test_solve_host_on_foreign_operator_matches_oracle(em):
    """Config-1 sphere at 16^3 through the host-driven loop: iterations,
    status, action counts and history of the oracle's BiCGStab."""
    from oracle import inputs, solvers
    N, L, k = (16, 16, 16), (1.0, 1.0, 1.0), 4.0
    _say("[solve_host] 16^3 sphere, wrapped operator")
    mask = inputs.ellipsoid_center_mask(N, L, (0, 0, 0), ell_center=(0, 0, 0), radii=(0.25,) * 3)
    chi = inputs.contrast_from_weights(mask.astype(float), background_eps_r=1.0, feature_eps_r=2.0)
    wave = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    cfg = dict(max_iter=200, rtol=1e-8)
    ref = solvers.bicgstab(solvers.EMOperator(N, L, k, chi), wave, solvers.Config(**cfg))
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=be.array(wave), k0=k, volume=1.0)
    w = _Wrapped(em.Operator(prob))
    res = em.BiCGStab(em.SolverConfig(**cfg)).solve(w, prob.wave)
    assert res.status == ref.status == "converged"
    assert res.iterations == ref.iterations and res.matvec_count == ref.matvec_count == w.calls
    assert res.residual_action_counts == ref.residual_action_counts
    np.testing.assert_allclose(res.residual_history, ref.residual_history, rtol=1e-9)
    assert _rel(_host(res.u), ref.u) < 1e-10


def test_judge_matches_oracle_and_qualifies(em):
    """true_relative_residual against the oracle's on the same solution; the
    qualification rule on a converged solve, a solve stopped early and an
    explicit qualification rtol."""
    from oracle import solvers, vie
    N, k = (12, 12, 12), 6.0
    _say("[judge] 12^3 sphere")
    be = _be(em, "double")
    prob = _sphere_problem(em, be, N, k, eps=3.0)
    op = em.Operator(prob)
    cfg = em.SolverConfig(max_iter=200, rtol=1e-8)
    j = em.judge_solve(em.BiCGStab(cfg), op, prob.wave)
    chi = _host(prob.eps_tensor).astype(np.complex128)
    ref_op = solvers.EMOperator(N, (1, 1, 1), k, chi)
    want = solvers.true_relative_residual(ref_op, _host(j.result.u), _host(prob.wave))
    assert abs(j.true_residual - want) <= 1e-6 * want
    assert j.qualified and j.true_residual < 1e-8 and j.tolerance == 1e-8
    assert j.result.true_final_residual == j.true_residual and j.elapsed_seconds > 0
    short = em.judge_solve(em.BiCGStab(em.SolverConfig(max_iter=2, rtol=1e-8)), op, prob.wave)
    assert short.result.status == "max_iter" and not short.qualified
    loose = em.judge_solve(em.BiCGStab(cfg), op, prob.wave, qualification_rtol=1e-12)
    assert loose.tolerance == 1e-12 and not loose.qualified
    mixed = em.judge_solve(em.MixedBiCGStab(cfg), op, prob.wave)
    assert mixed.qualified and mixed.true_residual < 1e-8
    u = _rand((3,) + N, 2)
    assert abs(em.true_relative_residual(op, be.array(u), prob.wave)
               - solvers.true_relative_residual(ref_op, u, _host(prob.wave))) < 1e-12
    Kh = vie.kernel_hat(N, (1, 1, 1), k)
    del Kh


def test_persistent_nbytes_is_the_octant(em):
    """persistent_nbytes counts what the kernel object holds: the octant
    spectrum, about 1/8 of the reference's dense (3, 3, 2N, 2N, 2N) tensor,
    which kernel_hat still materialises on request."""
    be = _be(em, "double")
    grid = em.Grid(N=(8, 6, 4), L=(1, 1, 1), center=(0, 0, 0), backend=be)
    pk = em.PreparedEMKernel.build(grid, k=2.0)
    t = pk.khat_octant
    assert pk.persistent_nbytes == t.numel() * t.element_size()
    dense = pk.kernel_hat
    assert tuple(dense.shape) == (3, 3, 16, 12, 8)
    assert pk.dense_kernel_hat_nbytes == dense.numel() * dense.element_size()
    assert pk.persistent_nbytes < dense.numel() * dense.element_size() / 4


def test_operator_and_far_field_share_the_contrast_cache(em):
    """The operators and the far field cache packed contrasts per tensor and
    packing rule: interleaving an acoustic solve, its far field, an in-place
    eta edit and both again gives the oracle's values each time."""
    from oracle import farfield as ff
    from oracle import vie
    N, L, k = (8, 8, 8), (1.0, 1.0, 1.0), 5.0
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    eta = em.acoustics.eta_sphere(grid, center=(0, 0, 0), radius=0.3, eta_inside=2.0 + 0.1j)
    prob = em.acoustics.make_acoustic_problem(grid, eta, k0=k)
    op = em.acoustics.AcousticOperator(prob)
    dirs = np.array([[1.0, 0, 0], [0, 0.6, 0.8]])
    u = _rand(N, 3)
    for edit in (False, True):
        if edit:
            prob.eta.mul_(1.1)
        chi = _host(prob.eta) - 1.0
        y = _host(op.matvec(be.array(u)))
        amp = em.acoustics.farfield.farfield_amplitude(be.array(u), prob, dirs)
        y2 = _host(op.matvec(be.array(u)))
        assert _rel(y, vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k))) < 1e-12
        assert np.array_equal(y, y2)
        ref = k ** 2 / (4 * np.pi) * ff.scatter_integral(u, chi, N, L, (0, 0, 0), k, dirs)
        assert _rel(amp, ref) < 1e-12
