"""GPU parity: the HIP operator and solvers against the oracle (runs on an MI355X).

Tolerances (the precision contract of the build): FP64 matvec/rmatvec within
1e-12 relative L2 of the complex128 oracle; FP32 (complex64 kernels, FP64
kernel build) within 1e-6; solver trajectories follow the oracle's to the
same iteration count and status.  Large sizes are checked through properties
that do not need the oracle: adjoint identity, linearity, zero contrast,
FP32 vs FP64 agreement.
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


def _problem(em, be, N, L, k, chi_np):
    grid = em.Grid(N=N, L=L, center=(0.0, 0.0, 0.0), backend=be)
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return em.Problem(grid=grid, eps_tensor=be.array(chi_np.astype(be.complex_dtype)), wave=wave, k0=k,
                      volume=grid.dv * int(np.prod(N)))


def _chi(kind, N, seed=3):
    rng = np.random.default_rng(seed)
    if kind == "full":
        return (rng.standard_normal((3, 3) + N) + 1j * rng.standard_normal((3, 3) + N)) * 0.3
    from oracle import inputs
    iso = inputs.ellipsis_refraction(N, (1, 1, 1), (0, 0, 0), eps_real=2.5, eps_imag=0.1, ell_center=(0, 0, 0),
                                     radius=(0.3, 0.2, 0.15) if min(N) >= 8 else (0.45, 0.45, 0.45))
    if kind == "iso":
        return iso
    d = iso.copy()
    d[1, 1] *= (2.0 + 0.2j) / (1.5 + 0.1j)
    d[2, 2] *= 1.0 / (1.5 + 0.1j)
    return d


EM_CASES = [((4, 4, 4), 1.0, "full"), ((4, 3, 2), 2.1, "full"), ((5, 6, 7), 3.0, "diag"),
            ((8, 8, 8), 4.0, "iso"), ((16, 16, 16), 4.0, "diag"), ((12, 10, 14), 8.0, "full"),
            # power-of-two fast kernels with a full contrast (the adjoint holds all components)
            ((32, 32, 32), 6.0, "full"),
            # doubled axes 3 * 2^k and 5 * 2^k: mixed-radix register kernels in every pass
            ((24, 40, 48), 6.0, "diag"), ((48, 24, 40), 5.0, "full"), ((40, 48, 24), 4.0, "iso")]


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("N,k,kind", EM_CASES)
def test_em_matvec_rmatvec_vs_oracle(em, prec, tol, N, k, kind):
    from oracle import vie
    _say(f"[em parity] N={N} k={k} chi={kind} prec={prec}")
    be = _be(em, prec)
    L = (1.0, 1.0, 1.0)
    chi = _chi(kind, N)
    op = em.Operator(_problem(em, be, N, L, k, chi))
    Kh = vie.kernel_hat(N, L, k)
    u = _rand((3,) + N, 42)
    y = _host(op.matvec(be.array(u.astype(be.complex_dtype))))
    ya = _host(op.rmatvec(be.array(u.astype(be.complex_dtype))))
    assert _rel(y, vie.em_matvec(chi, u, Kh)) < tol
    assert _rel(ya, vie.em_rmatvec(chi, u, Khat=Kh)) < tol


def test_kernel_build_matches_oracle_spectrum(em):
    """Materialised spectrum vs FFT of the oracle's spatial kernel with the
    index-N planes zeroed (the only difference from the reference's storage)."""
    from oracle import vie
    N, L, k = (6, 5, 4), (1.0, 0.9, 0.8), 3.3
    _say(f"[kernel build] N={N}")
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    pk = em.PreparedEMKernel.build(grid, k=k)
    K = vie.kernel_tensor(N, L, k)
    K[:, :, N[0], :, :] = 0
    K[:, :, :, N[1], :] = 0
    K[:, :, :, :, N[2]] = 0
    ref = np.fft.fftn(K, axes=(-3, -2, -1))
    got = _host(pk.kernel_hat)
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
def test_acoustic_vs_oracle(em, prec, tol):
    from oracle import vie
    be = _be(em, prec)
    for N, L, k in [((3, 3, 2), (1.0, 0.8, 0.6), 1.2), ((16, 12, 10), (1.0, 1.0, 1.0), 10.0)]:
        _say(f"[acoustic parity] N={N} prec={prec}")
        grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
        eta = em.acoustics.eta_sphere(grid, center=(0, 0, 0), radius=0.3, eta_inside=2.0 + 0.2j)
        prob = em.acoustics.make_acoustic_problem(grid, eta, k0=k)
        op = em.acoustics.AcousticOperator(prob)
        chi = _host(prob.chi).astype(np.complex128)
        u = _rand(N, 9)
        y = _host(op.matvec(be.array(u.astype(be.complex_dtype))))
        ya = _host(op.rmatvec(be.array(u.astype(be.complex_dtype))))
        assert _rel(y, vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k))) < tol
        assert _rel(ya, vie.acoustic_rmatvec(chi, u, vie.acoustic_kernel_hat(N, L, k, adjoint=True))) < tol


def test_batched_rhs_equals_single_calls(em):
    N, k = (16, 16, 16), 8.0
    _say(f"[batch] N={N}")
    be = _be(em, "single")
    chi = _chi("diag", N)
    op = em.Operator(_problem(em, be, N, (1, 1, 1), k, chi))
    U = be.array(_rand((5, 3) + N, 11).astype(np.complex64))
    Y = op.matvec(U)
    for b in range(5):
        assert _rel(_host(Y[b]), _host(op.matvec(U[b]))) < 1e-6


def test_streaming_batch_equals_cached_single_calls(em):
    """A batch of 16 fields at 64^3 has a 384 MiB x-pass working set, so its
    plan runs the passes with non-temporal row accesses; one field (24 MiB)
    runs the plain kernels.  Each field of the batch equals its own call,
    forward and adjoint."""
    N, k = (64, 64, 64), 8.0
    _say(f"[batch nt] N={N} x 16")
    be = _be(em, "single")
    chi = _chi("diag", N)
    op = em.Operator(_problem(em, be, N, (1, 1, 1), k, chi))
    U = be.array(_rand((16, 3) + N, 12).astype(np.complex64))
    Y, YA = op.matvec(U), op.rmatvec(U)
    for b in (0, 7, 15):
        assert _rel(_host(Y[b]), _host(op.matvec(U[b]))) < 1e-6
        assert _rel(_host(YA[b]), _host(op.rmatvec(U[b]))) < 1e-6


@pytest.mark.parametrize("N", [(64, 64, 64), (128, 128, 128)])
def test_full_size_properties(em, N):
    """Adjoint identity, linearity, zero-contrast identity and FP32/FP64
    agreement at the benchmark sizes (no oracle needed)."""
    import torch
    _say(f"[properties] N={N}")
    k = 8.0
    chi = _chi("diag", N)
    be64, be32 = _be(em, "double"), _be(em, "single")
    op64 = em.Operator(_problem(em, be64, N, (1, 1, 1), k, chi))
    op32 = em.Operator(_problem(em, be32, N, (1, 1, 1), k, chi))
    u = be64.array(_rand((3,) + N, 1))
    v = be64.array(_rand((3,) + N, 2))
    au, ahv = op64.matvec(u), op64.rmatvec(v)
    lhs = complex(torch.vdot(au.flatten(), v.flatten()))
    rhs = complex(torch.vdot(u.flatten(), ahv.flatten()))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    a = 0.7 - 0.3j
    lin = op64.matvec(a * u + v) - (a * au + op64.matvec(v))
    assert float(torch.linalg.vector_norm(lin) / torch.linalg.vector_norm(au)) < 1e-13
    y32 = op32.matvec(u.to(torch.complex64)).to(torch.complex128)
    assert float(torch.linalg.vector_norm(y32 - au) / torch.linalg.vector_norm(au)) < 1e-6
    zero = em.Operator(_problem(em, be64, N, (1, 1, 1), k, np.zeros((3, 3) + N, complex)),
                       prepared_kernel=op64.prepared_kernel)
    assert torch.equal(zero.matvec(u), u)


def test_bicgstab_trajectory_matches_oracle(em):
    """Config-1 sphere at 16^3: same iterations, status and residual history."""
    from oracle import inputs, solvers
    N, L, k = (16, 16, 16), (1.0, 1.0, 1.0), 4.0
    _say("[bicgstab] 16^3 sphere")
    mask = inputs.ellipsoid_center_mask(N, L, (0, 0, 0), ell_center=(0, 0, 0), radii=(0.25,) * 3)
    chi = inputs.contrast_from_weights(mask.astype(float), background_eps_r=1.0, feature_eps_r=2.0)
    wave = inputs.flat_wave_vec(N, L, (0, 0, 0), k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    ref = solvers.bicgstab(solvers.EMOperator(N, L, k, chi), wave, solvers.Config(max_iter=200, rtol=1e-8))
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=be.array(wave), k0=k, volume=grid.dv * 4096)
    op = em.Operator(prob)
    res = em.BiCGStab(em.SolverConfig(max_iter=200, rtol=1e-8)).solve(op, prob.wave)
    assert res.status == ref.status == "converged"
    assert res.iterations == ref.iterations and res.matvec_count == ref.matvec_count
    np.testing.assert_allclose(res.residual_history, ref.residual_history, rtol=1e-6)
    assert _rel(_host(res.u), ref.u) < 1e-9


@pytest.mark.parametrize("name", ["SIM", "TwoStep"])
def test_sim_twostep_match_oracle(em, name):
    from oracle import inputs, solvers
    N, L, k = (6, 6, 6), (1.0, 1.0, 1.0), 1.0
    _say(f"[{name}] 6^3")
    chi = inputs.cylinder_refraction(N, L, (0, 0, 0), eps_real=1.3, eps_imag=0.1, radius=0.3)
    rhs = _rand((3,) + N, 5)
    cfg = dict(max_iter=60, rtol=1e-9, mu=1.0 + 0j, radius=0.5)
    ref = getattr(solvers, name.lower())(solvers.EMOperator(N, L, k, chi), rhs, solvers.Config(**cfg))
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=be.array(rhs), k0=k, volume=1.0)
    res = getattr(em, name)(em.SolverConfig(**cfg)).solve(em.Operator(prob), prob.wave)
    assert res.status == ref.status and res.iterations == ref.iterations
    assert res.matvec_count == ref.matvec_count and res.rmatvec_count == ref.rmatvec_count
    np.testing.assert_allclose(res.residual_history, ref.residual_history, rtol=1e-7)
    assert _rel(_host(res.u), ref.u) < 1e-9


def test_forced_statuses_on_gpu(em):
    N = (4, 4, 4)
    be = _be(em, "double")
    grid = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=be)
    chi = _chi("full", N)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=be.array(_rand((3,) + N, 1)), k0=1.0, volume=1.0)
    op = em.Operator(prob)
    r = em.BiCGStab(em.SolverConfig()).solve(op, be.zeros((3,) + N, "complex"))
    assert r.status == "converged" and r.matvec_count == 0
    bad = _rand((3,) + N, 2)
    bad[0, 0, 0, 0] = np.nan
    assert em.BiCGStab(em.SolverConfig()).solve(op, be.array(bad)).status == "nonfinite"
    r = em.BiCGStab(em.SolverConfig(max_iter=100, max_operator_actions=2)).solve(op, prob.wave)
    assert r.status == "max_operator_actions" and r.operator_action_count <= 2
    pk = em.PreparedEMKernel.build(grid, k=1.0, include_adjoint=False)
    with pytest.raises(RuntimeError, match="adjoint action is unavailable"):
        em.Operator(prob, prepared_kernel=pk).rmatvec(prob.wave)
    with pytest.raises(ValueError, match="wave number"):
        em.Operator(prob, prepared_kernel=em.PreparedEMKernel.build(grid, k=2.0))


def _config2(em, be, N, k):
    """Anisotropic absorbing ellipsoid: radii (0.3, 0.2, 0.15), chi = diag(1.5+0.1i, 2.0+0.2i, 1.0)."""
    return _problem(em, be, N, (1.0, 1.0, 1.0), k, _chi("diag", N))


def test_mixed_refinement_reaches_fp64_tolerance(em):
    """complex64 inner BiCGStab + FP64 refinement: true residual (recomputed
    here with the FP64 operator) below 1e-8 and the solution within 1e-6 of
    the oracle's complex128 BiCGStab solution."""
    from oracle import solvers
    N, k = (16, 16, 16), 8.0
    _say("[mixed] 16^3 config-2 ellipsoid k0=8")
    be = _be(em, "double")
    prob = _config2(em, be, N, k)
    op = em.Operator(prob)
    res = em.MixedBiCGStab(em.SolverConfig(max_iter=200, rtol=1e-8)).solve(op, prob.wave)
    assert res.status == "converged", res.extras
    f = prob.wave
    true = float(np.linalg.norm(_host(op.matvec(res.u) - f)) / np.linalg.norm(_host(f)))
    assert true < 1e-8 and abs(true - res.true_final_residual) <= 1e-12 + 1e-6 * true
    assert res.extras["fp32_matvecs"] > res.extras["fp64_matvecs"] >= 2
    chi = _chi("diag", N)
    ref = solvers.bicgstab(solvers.EMOperator(N, (1, 1, 1), k, chi), _host(f), solvers.Config(max_iter=200, rtol=1e-10))
    assert ref.status == "converged"
    assert _rel(_host(res.u), ref.u) < 1e-6


def test_mixed_refinement_fallback_and_errors(em):
    """A one-iteration inner solve cannot halve the residual: the FP64
    BiCGStab fallback takes over and the result still meets the tolerance."""
    N, k = (8, 8, 8), 4.0
    _say("[mixed] 8^3 forced fallback")
    be = _be(em, "double")
    prob = _config2(em, be, N, k)
    op = em.Operator(prob)
    solver = em.MixedBiCGStab(em.SolverConfig(max_iter=200, rtol=1e-9), inner_max_iter=1, stagnation=1e-3)
    res = solver.solve(op, prob.wave)
    assert res.status == "converged" and res.extras["fallback"] == "converged"
    true = float(np.linalg.norm(_host(op.matvec(res.u) - prob.wave)) / np.linalg.norm(_host(prob.wave)))
    assert true < 1e-9
    z = em.MixedBiCGStab(em.SolverConfig()).solve(op, be.zeros((3,) + N, "complex"))
    assert z.status == "converged" and z.matvec_count == 0
    be32 = _be(em, "single")
    with pytest.raises(ValueError, match="FP64"):
        em.MixedBiCGStab(em.SolverConfig()).solve(em.Operator(_config2(em, be32, N, k)), prob.wave)


def test_batched_bicgstab_matches_single_solves(em):
    """One device-resident run over a batch of RHS (incident directions x
    polarisations, a zero RHS and a NaN RHS) gives each RHS exactly the
    iterations, status, history and solution of its own single solve."""
    import torch
    from em3d_amd.sources import incident_set, plane_waves
    N, k = (16, 16, 16), 6.0
    _say("[batched bicgstab] 16^3 sphere eps 4, 6 RHS")
    for prec in ("double", "single"):
        be = _be(em, prec)
        grid = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=be)
        chi = em.geometry.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=4.0,
                                                 geometry=em.geometry.Ellipsoid(center=(0, 0, 0), radii=(0.3,) * 3))
        o, a = incident_set(2, seed=4)
        waves = plane_waves(grid, k=k, orients=o, amplitudes=a)
        zero = torch.zeros_like(waves[0])
        bad = waves[1].clone()
        bad[0, 1, 2, 3] = float("nan")
        rhs = torch.stack([waves[0], zero, waves[1], bad, waves[2], waves[3]])
        prob = em.Problem(grid=grid, eps_tensor=chi, wave=waves[0], k0=k, volume=1.0)
        op = em.Operator(prob)
        cfg = em.SolverConfig(max_iter=200, rtol=1e-8 if prec == "double" else 1e-5)
        batch = em.BiCGStab(cfg).solve_batch(op, rhs)
        assert [r.status for r in batch][1] == "converged" and batch[1].matvec_count == 0
        assert batch[3].status == "nonfinite"
        for b in range(rhs.shape[0]):
            single = em.BiCGStab(cfg).solve(op, rhs[b].contiguous())
            got = batch[b]
            assert (got.status, got.iterations, got.matvec_count) == (single.status, single.iterations,
                                                                        single.matvec_count), b
            assert got.residual_action_counts == single.residual_action_counts
            np.testing.assert_allclose(got.residual_history, single.residual_history, rtol=1e-12, atol=0)
            if single.status != "nonfinite":
                assert torch.equal(got.u, single.u), b


@pytest.mark.parametrize("prec,tol", [("double", 1e-13), ("single", 2e-6)])
def test_plane_waves_match_oracle(em, prec, tol):
    """Device plane waves (em3d_plane_waves) against the oracle's flat_wave_vec
    (src/em3d/wave.py:9-37) and against the product's host flat_wave_vec.  The
    device phase is FP64; the FP32 oracle forms it in float32, hence 2e-6."""
    from em3d_amd.sources import incident_set, plane_waves
    from oracle import inputs
    be = _be(em, prec)
    N, L, c = (12, 10, 8), (1.0, 0.8, 1.3), (0.1, -0.2, 0.05)
    grid = em.Grid(N=N, L=L, center=c, backend=be)
    o, a = incident_set(3, seed=7)
    a = a * (0.5 - 0.25j)
    dev = plane_waves(grid, k=9.0, orients=o, amplitudes=a, phi0=0.3, sign=-1)
    for b in range(o.shape[0]):
        ref = inputs.flat_wave_vec(N, L, c, k=9.0, orient=tuple(o[b]), amplitude=tuple(a[b]), phi0=0.3, sign=-1,
                                   precision=prec)
        assert _rel(_host(dev[b]), ref.astype(np.complex128)) < tol
        host = em.flat_wave_vec(grid, k=9.0, orient=tuple(o[b]), amplitude=tuple(a[b]), phi0=0.3, sign=-1)
        assert _rel(_host(dev[b]), _host(host)) < tol


@pytest.mark.parametrize("prec", ["double", "single"])
def test_device_materials_match_oracle(em, prec):
    """Device ellipsoid contrast (em3d_ellipsoid_contrast) and inclusion eta
    (em3d_eta_inclusions) equal the oracle's cell-centre sampling bit for bit
    (src/em3d/geometry.py:62-93, 233-252; acoustics/materials.py:62-83), cells
    on the surface included, and agree with the product's host builders."""
    import torch
    from em3d_amd.acoustics.materials import eta_inclusions, random_spheres
    from oracle import inputs
    be = _be(em, prec)
    N, L, c = (16, 16, 16), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0)
    grid = em.Grid(N=N, L=L, center=c, backend=be)
    cdt = np.complex128 if prec == "double" else np.complex64
    geo = em.geometry
    eps1 = np.diag([2.5 + 0.1j, 3.0 + 0.2j, 2.0])
    eps2 = np.array([[2.0, 0.3j, 0], [-0.3j, 2.0, 0], [0, 0, 1.5]])
    # centre on a cell centre, radius 2 cells: the axis cells at distance 2 lie on the surface
    c0 = float(inputs.axis(16, 1.0, 0.0, prec)[7])
    e1 = geo.Ellipsoid(center=(c0, c0, c0), radii=(0.125,) * 3)
    e2 = geo.Ellipsoid(center=(0.1, 0, 0), radii=(0.2, 0.15, 0.1))
    m1 = inputs.ellipsoid_center_mask(N, L, c, ell_center=e1.center, radii=e1.radii, precision=prec)
    m2 = inputs.ellipsoid_center_mask(N, L, c, ell_center=e2.center, radii=e2.radii, precision=prec)
    assert int(m1.sum()) == 33
    ref1 = inputs.contrast_from_weights(m1.astype(float), background_eps_r=1.0, feature_eps_r=eps1, precision=prec)
    got1 = _host(geo.ellipsoids_contrast(grid, [(e1, eps1)]))
    assert np.array_equal(got1, ref1)
    assert np.array_equal(_host(geo.ellipsoids_contrast(grid, [(e1, eps1)], kind="diag")),
                          np.stack([ref1[i, i] for i in range(3)]))
    assert torch.equal(geo.ellipsoids_contrast(grid, [(e1, eps1)]),
                       geo.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=eps1, geometry=e1))
    two = _host(geo.ellipsoids_contrast(grid, [(e1, eps1), (e2, eps2)], background_eps_r=1.2))
    for a in range(3):
        for b in range(3):
            want = np.full(N, (1.2 - 1.0) if a == b else 0.0, dtype=cdt)
            want = np.where(m1, cdt(eps1[a, b] - (a == b)), want)
            want = np.where(m2, cdt(eps2[a, b] - (a == b)), want)
            assert np.array_equal(two[a, b], want), (a, b)
    spheres = random_spheres(8, seed=3, radius=(0.1, 0.2))
    spheres.append(((c0, c0, c0), (0.125,) * 3, 2.0 + 0.1j))
    eta = _host(eta_inclusions(grid, spheres))
    want = np.ones(N, dtype=cdt)
    for cc, r, v in spheres:
        m = inputs.eta_ellipsoid(N, L, c, ell_center=cc, radii=r, eta_inside=1.0, eta_outside=0.0,
                                 precision=prec) != 0
        want = np.where(m, cdt(v), want)
    assert np.array_equal(eta, want)


def test_autotuner_keeps_results(em):
    """The autotuner picks valid variants and the matvec it leaves behind
    agrees with the default kernels (single precision, 1e-6)."""
    import torch
    from em3d_amd.operator import HipPlan
    N = (64, 64, 64)
    _say("[autotune] 64^3 single")
    be = _be(em, "single")
    prob = _config2(em, be, N, 8.0)
    op = em.Operator(prob)
    plan = HipPlan.get(N, 3, 1, em.Precision.SINGLE)
    plan.set_variants([0, 0, 0])
    ref = op.matvec(prob.wave)
    counts = plan.variant_counts()
    out = op.autotune(reps=2)
    v = out["variants"]
    assert all(-1 <= v[k] < counts[k] for k in range(3)) and out["ms"] > 0
    got = op.matvec(prob.wave)
    assert float(torch.linalg.vector_norm(got - ref) / torch.linalg.vector_norm(ref)) < 1e-6
    with pytest.raises(Exception, match="out of range"):
        plan.set_variants([counts[0], 0, 0])
    plan.set_variants([0, 0, 0])


def test_packed_group_graph_replays_match_eager(em):
    """Grids 16^3 / 24^3 / 32^3 / 40^3 in one captured graph: every replay,
    with eager work and synchronisations in between and new inputs each time,
    equals the eager matvecs."""
    import torch
    from em3d_amd.packed import PackedOperators
    _say("[packed] 16/24/32/40 graph replays")
    be = _be(em, "single")
    ops = []
    for i, n in enumerate((16, 24, 32, 40)):
        prob = _config2(em, be, (n, n, n), 4.0 + i)
        ops.append(em.Operator(prob))
    group = PackedOperators(ops, autotune=False)
    rng = np.random.default_rng(11)
    for rep in range(3):
        fields = [be.array(_rand((3,) + op._N, 100 + 10 * rep + j).astype(np.complex64)) for j, op in enumerate(ops)]
        got = group.matvec(fields)
        torch.cuda.synchronize()
        noise = ops[rep].matvec(be.array(_rand((3,) + ops[rep]._N, 7).astype(np.complex64)))   # eager work between
        torch.cuda.synchronize()
        for j, op in enumerate(ops):
            assert torch.equal(got[j], op.matvec(fields[j])), (rep, j)
        del noise
    with pytest.raises(ValueError):
        group.matvec(fields[:2])


def test_gamma0_coarse_operator_and_sim(em):
    """The coarse operator assembled by the batched HIP matvec equals the dense
    oracle on the resampled contrast (config-1 sphere 16^3 -> 4^3), and SIM
    with the resulting (mu, R) runs on the fine operator."""
    from em3d_amd import gamma0
    from oracle import inputs, vie
    N, L, k = (16, 16, 16), (1.0, 1.0, 1.0), 4.0
    _say("[gamma0] 16^3 sphere, coarse 4^3")
    mask = inputs.ellipsoid_center_mask(N, L, (0, 0, 0), ell_center=(0, 0, 0), radii=(0.25,) * 3)
    chi = inputs.contrast_from_weights(mask.astype(float), background_eps_r=1.0, feature_eps_r=2.0)
    be = _be(em, "double")
    grid = em.Grid(N=N, L=L, center=(0, 0, 0), backend=be)
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi), wave=wave, k0=k, volume=1.0)
    A = gamma0.coarse_operator_matrix(prob, (4, 4, 4))
    Nc = (4, 4, 4)
    fine_ax = [inputs.axis(N[a], L[a], 0.0, "double") for a in range(3)]
    coarse_ax = [inputs.axis(Nc[a], L[a], 0.0, "double") for a in range(3)]
    idx = [gamma0._nearest(fine_ax[a], coarse_ax[a]) for a in range(3)]
    chi_c = chi[:, :, idx[0]][:, :, :, idx[1]][:, :, :, :, idx[2]]
    D = vie.dense_operator_matrix(Nc, L, (0, 0, 0), k, chi_c)
    n = 64
    perm = np.array([(j % 3) * n + j // 3 for j in range(3 * n)])    # oracle: cell-major, component-minor
    assert np.abs(A[np.ix_(perm, perm)] - D).max() <= 1e-12 * np.abs(D).max()
    a = gamma0.estimate_from_problem(prob, (4, 4, 4))
    assert a.matrix_shape == (192, 192) and abs(a.mu) > a.radius
    res = em.SIM(em.SolverConfig(max_iter=300, rtol=1e-6, mu=a.mu, radius=a.radius)).solve(em.Operator(prob), wave)
    _say(f"[gamma0] mu={a.mu:.4f} R={a.radius:.4f} SIM {res.status} {res.iterations} it")
    assert res.status in ("converged", "max_iter", "divergence_guard") and res.residual_history[-1] < 1.0


def test_reference_format_records(em):
    from em3d_amd import reports
    be = _be(em, "double")
    prob = _config2(em, be, (8, 8, 8), 4.0)
    op = em.Operator(prob)
    rec = reports.matvec_record("ellipsoid_N8", op, prob.wave, repeats=3, operator_build_sec=0.1)
    assert list(rec) == ["case_name", "N", "dof", "operator_build_sec", "matvec_avg_sec", "matvec_min_sec",
                         "matvec_max_sec", "repeats"]
    assert rec["N"] == "8x8x8" and rec["dof"] == 1536 and rec["matvec_min_sec"] <= rec["matvec_avg_sec"]
    res, run = reports.timed_solve(em.BiCGStab(em.SolverConfig(rtol=1e-8)), op, prob.wave, case_name="ellipsoid_N8",
                                   solver_name="BiCGStab")
    assert run.iterations == res.iterations and run.final_residual == res.residual_history[-1]


def test_same_shape_operators_share_a_workspace(em):
    """Operators of one shape share one FFT workspace per stream; their
    interleaved matvecs, and an adjoint between them, equal each operator's
    own results."""
    import torch
    _say("[workspace] two same-shape operators, interleaved")
    be = _be(em, "single")
    N, L = (24, 20, 16), (1.0, 0.8, 0.7)
    ops = [em.Operator(_problem(em, be, N, L, k, _chi(kind, N, seed=s)))
           for k, kind, s in ((3.0, "diag", 11), (5.0, "full", 12))]
    u = be.array(_rand((3,) + N, 7).astype(be.complex_dtype))
    ref = [op.matvec(u).clone() for op in ops]
    refa = ops[0].rmatvec(u).clone()
    for _ in range(2):
        y1 = ops[1].matvec(u)
        ya = ops[0].rmatvec(u)
        y0 = ops[0].matvec(u)
        assert torch.equal(y0, ref[0]) and torch.equal(y1, ref[1]) and torch.equal(ya, refa)
