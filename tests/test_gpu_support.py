"""GPU tests of the restriction to the contrast's support (em3d_amd.support,
em3d_amd.SupportRestricted).

Oracles: the oracle's FFT operators (src/em3d/operator.py:353-398 and
acoustics/operator.py:40-67, restated in oracle/vie.py) on the full grid and
on the box grid (same k0, same cell size, contrast cropped), and the oracle's
BiCGStab (src/em3d/solvers/bicgstab.py:12-211) on the full grid with its true
residual (src/em3d/experiments/stationary_validation.py:172-180).  Bounds,
crops and scatters are integer/byte work and are compared exactly.
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


N0 = (32, 24, 20)


def _em_problem(em, be, N, k, kind="diag", center=(0.1, -0.05, 0.0), radii=(0.2, 0.15, 0.2)):
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    eps = {"iso": 2.5 + 0.1j,
           "diag": np.diag([2.5 + 0.1j, 3.0 + 0.2j, 2.0]),
           "full": np.array([[2.5 + 0.1j, 0.2, 0.0], [0.2, 3.0 + 0.2j, 0.1j], [0.0, 0.1j, 2.0]])}[kind]
    chi = em.geometry.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=eps,
                                             geometry=em.geometry.Ellipsoid(center=center, radii=radii))
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    return em.Problem(grid=grid, eps_tensor=chi, wave=wave, k0=k, volume=1.0)


def _np_bounds(mask):
    out = []
    for a in range(3):
        other = tuple(b for b in range(3) if b != a)
        idx = np.nonzero(mask.any(axis=other))[0]
        out.append((int(idx[0]), int(idx[-1])))
    return tuple(out)


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("kind", ["iso", "diag", "full"])
def test_support_bounds_exact(em, prec, kind):
    """em3d_support_bounds equals the bounds of the non-zero cells of the
    contrast, for every packing and for a contrast that is zero everywhere."""
    from em3d_amd.operator import pack_contrast
    from em3d_amd.support import support_bounds
    _say(f"[support bounds] kind={kind} prec={prec}")
    be = _be(em, prec)
    prob = _em_problem(em, be, N0, 4.0, kind)
    ck, chi = pack_contrast(prob.eps_tensor)
    host = _host(prob.eps_tensor)
    want = _np_bounds((host != 0).any(axis=(0, 1)))
    assert support_bounds(N0, ck, chi, be.precision) == want
    # one non-zero cell at a corner of the grid
    import torch
    z = torch.zeros_like(chi)
    assert support_bounds(N0, ck, z, be.precision) is None
    z[..., N0[0] - 1, 0, N0[2] - 1] = 1.0
    assert support_bounds(N0, ck, z, be.precision) == ((N0[0] - 1, N0[0] - 1), (0, 0), (N0[2] - 1, N0[2] - 1))


def _np_box(full, box, ncc):
    """NumPy reference of the crop: planes in groups of ncc, the box's axis
    order, components following the axes (ncc 3) or both indices (ncc 9)."""
    planes = full.shape[0]
    sl = tuple(slice(box.lo[a], box.lo[a] + box.n[a]) for a in range(3))
    out = np.empty((planes,) + box.shape, dtype=full.dtype)
    perm = box.perm
    for p in range(planes):
        g, q = divmod(p, ncc)
        src = q if ncc == 1 else (perm[q] if ncc == 3 else 3 * perm[q // 3] + perm[q % 3])
        out[p] = np.transpose(full[g * ncc + src][sl], perm)
    return out


@pytest.mark.parametrize("prec", ["double", "single"])
@pytest.mark.parametrize("perm,ncc", [((0, 1, 2), 1), ((1, 0, 2), 3), ((2, 1, 0), 9), ((1, 2, 0), 3)])
def test_gather_scatter_exact(em, prec, perm, ncc):
    """The crop and the two scatters equal NumPy slicing and transposes bit
    for bit, in every box axis order and plane grouping."""
    from em3d_amd.support import SupportBox, gather, scatter
    _say(f"[support gather/scatter] prec={prec} perm={perm} ncc={ncc}")
    be = _be(em, prec)
    box = SupportBox(N=N0, lo=(5, 3, 2), n=(16, 8, 16), bounds=((5, 20), (3, 10), (2, 17)), perm=perm)
    planes = 2 * ncc
    c = be.complex_dtype
    full = _rand((planes,) + N0, 1).astype(c)
    f = _rand((planes,) + N0, 2).astype(c)
    y = _rand((planes,) + N0, 3).astype(c)
    g = _host(gather(be.array(full), box, planes, ncc))
    assert np.array_equal(g, _np_box(full, box, ncc))
    # scatter is the inverse of gather inside the box
    bf = _rand((planes,) + box.shape, 4).astype(c)
    back = _host(scatter(be.array(bf), box, planes, ncc))
    assert np.array_equal(_np_box(back, box, ncc), bf)
    inside = np.zeros(N0, bool)
    inside[tuple(slice(box.lo[a], box.lo[a] + box.n[a]) for a in range(3))] = True
    assert not np.any(back[:, ~inside])
    yd = be.array(y)
    out = _host(scatter(be.array(bf), box, planes, ncc, be.array(f), yd, out=yd))   # in place over y
    assert np.array_equal(out[:, ~inside], (f - y)[:, ~inside])
    assert np.array_equal(_np_box(out, box, ncc), bf)
    with pytest.raises(RuntimeError, match="inside the grid"):
        scatter(be.array(bf), SupportBox(N=N0, lo=(20, 3, 2), n=(16, 8, 16), bounds=box.bounds, perm=perm), planes,
                ncc)


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("kind", ["iso", "diag", "full"])
def test_box_operator_is_the_full_operator_on_the_box(em, prec, tol, kind):
    """(A u) on the box cells equals A_box (u on the box) for any u, and A_box
    is the oracle's operator of the box grid: same k0, same cell size, the
    contrast cropped.  Its adjoint is the oracle's adjoint of that grid."""
    from oracle import vie
    from em3d_amd.support import support_operator
    k = 5.0
    _say(f"[support operator] N={N0} kind={kind} prec={prec}")
    be = _be(em, prec)
    prob = _em_problem(em, be, N0, k, kind)
    op = em.Operator(prob)
    sub = support_operator(op)
    assert sub is not None and sub.box.n == (16, 8, 8) and sub.box.perm == (0, 1, 2), sub.box
    _check_box_operator(em, op, sub, prob, k, be, tol)


def _check_box_operator(em, op, sub, prob, k, be, tol):
    """A_box (in its own axis order, mapped back to the grid) against the
    oracle's operator of the box grid in the grid's order, and against the
    full operator on the box; the box adjoint against the oracle's."""
    from oracle import vie
    N = tuple(prob.grid.N)
    lo, n = sub.box.lo, sub.box.n
    sl = (slice(None),) + tuple(slice(lo[a], lo[a] + n[a]) for a in range(3))
    u = _rand((3,) + N, 7)
    ud = be.array(u.astype(be.complex_dtype))
    full = _host(op.matvec(ud))[sl]
    boxed = _host(sub.extend(sub.matvec(sub.restrict(ud))))[sl]
    chi = _host(prob.eps_tensor).astype(np.complex128)
    Lb = tuple(n[a] * (1.0 / N[a]) for a in range(3))
    Kb = vie.kernel_hat(n, Lb, k)
    chib = chi[(slice(None),) + sl]
    want = vie.em_matvec(chib, u[sl], Kb)
    assert _rel(boxed, want) < tol
    assert _rel(full, want) < tol
    assert _rel(full, vie.em_matvec(chi, u, vie.kernel_hat(N, (1, 1, 1), k))[sl]) < tol
    got = _host(sub.extend(sub.rmatvec(sub.restrict(ud))))[sl]
    assert _rel(got, vie.em_rmatvec(chib, u[sl], Khat=Kb)) < tol


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("kind", ["diag", "full"])
def test_permuted_box_operator(em, prec, tol, kind):
    """A box whose x side is not a power of two takes a power-of-two side as
    its first axis (its x pass): the permuted box operator, with components
    and contrast entries permuted, is still the full operator on the box."""
    from em3d_amd.support import support_operator
    N, k = (48, 32, 24), 5.0
    _say(f"[support permuted operator] N={N} kind={kind} prec={prec}")
    be = _be(em, prec)
    prob = _em_problem(em, be, N, k, kind, center=(0.05, 0.0, 0.02), radii=(0.22, 0.2, 0.15))
    op = em.Operator(prob)
    sub = support_operator(op)
    assert sub is not None and sub.box.perm != (0, 1, 2), sub.box
    assert sub.box.shape[0] & (sub.box.shape[0] - 1) == 0, sub.box
    _check_box_operator(em, op, sub, prob, k, be, tol)


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
def test_acoustic_box_operator(em, prec, tol):
    from oracle import vie
    from em3d_amd.support import support_operator
    N, k = (32, 32, 24), 8.0
    _say(f"[support acoustic operator] N={N} prec={prec}")
    be = _be(em, prec)
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    eta = em.acoustics.materials.eta_inclusions(grid, [((-0.1, 0.05, 0.0), (0.08, 0.08, 0.08), 2.0 + 0.1j),
                                                       ((0.1, 0.0, 0.05), (0.06, 0.06, 0.06), 1.5 + 0.2j)])
    prob = em.acoustics.make_acoustic_problem(grid, eta, k0=k)
    op = em.acoustics.AcousticOperator(prob)
    sub = support_operator(op)
    assert sub is not None and sub.box.fraction < 0.1, sub.box
    lo, n = sub.box.lo, sub.box.n
    sl = tuple(slice(lo[a], lo[a] + n[a]) for a in range(3))
    u = _rand(N, 8)
    ud = be.array(u.astype(be.complex_dtype))
    chi = _host(eta).astype(np.complex128) - 1.0
    Lb = tuple(n[a] * (1.0 / N[a]) for a in range(3))
    want = vie.acoustic_matvec(chi[sl], u[sl], vie.acoustic_kernel_hat(n, Lb, k))
    assert _rel(_host(sub.extend(sub.matvec(sub.restrict(ud))))[sl], want) < tol
    assert _rel(_host(op.matvec(ud))[sl], want) < tol


def _oracle_op(chi, Kh, acoustic=False):
    from oracle import vie

    class Op:
        def matvec(self, u):
            return vie.acoustic_matvec(chi, u, Kh) if acoustic else vie.em_matvec(chi, u, Kh)
    return Op()


def _boxed(shape, lo, n, seed, lead=()):
    """Random values inside the box [lo, lo + n) of the last three axes, zero outside."""
    x = np.zeros(lead + shape, complex)
    sl = tuple(slice(lo[a], lo[a] + n[a]) for a in range(3))
    x[(Ellipsis,) + sl] = _rand(lead + tuple(n), seed)
    return x


# (N, lo, n): boxes inside, filling, on a face and along an edge of grids whose
# doubled axes take the register kernels (64, 48, 40 / 96, 80, 48 points)
WINDOW_CASES = [((32, 24, 20), (5, 3, 2), (20, 15, 12)), ((32, 24, 20), (0, 0, 0), (32, 24, 20)),
                ((32, 24, 20), (31, 0, 7), (1, 24, 1)), ((48, 40, 24), (10, 5, 3), (30, 30, 15))]


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("kind", ["diag", "full"])
@pytest.mark.parametrize("N,lo,n", WINDOW_CASES)
def test_window_matvec(em, prec, tol, kind, N, lo, n):
    """em3d_matvec_window with chi zero outside the box (u is not): the same
    bits as em3d_matvec, whose forward passes transform exact zeros there, and
    the oracle's matvec (operator.py:353-363)."""
    from oracle import vie
    _say(f"[window matvec] N={N} lo={lo} n={n} kind={kind} prec={prec}")
    be = _be(em, prec)
    L, k = (1.0, 0.8, 1.2), 3.0
    chi = _boxed(N, lo, n, 1, (3, 3)) * 0.3
    if kind == "diag":
        chi *= np.eye(3)[:, :, None, None, None]
    grid = em.Grid(N=N, L=L, center=(0.0, 0.0, 0.0), backend=be)
    prob = em.Problem(grid=grid, eps_tensor=be.array(chi.astype(be.complex_dtype)),
                      wave=be.array(_rand((3,) + N, 2).astype(be.complex_dtype)), k0=k, volume=1.0)
    op = em.Operator(prob)
    u = _rand((3,) + N, 3)
    ud = be.array(u.astype(be.complex_dtype))
    yw = _host(op.matvec_window(ud, lo, n))
    assert np.array_equal(yw, _host(op.matvec(ud)))
    assert _rel(yw, vie.em_matvec(chi, u, vie.kernel_hat(N, L, k))) < tol


def test_window_matvec_full_size_and_acoustic(em):
    """At 128^3 (the bench's ellipsoid, its support box) and for the scalar
    operator: the windowed matvec gives the same bits as the whole-grid one;
    a box outside the grid is refused."""
    import em3d_amd._lib as lib
    from em3d_amd.support import support_operator
    _say("[window matvec] 128^3 ellipsoid box, FP64 and complex64; acoustic 40^3")
    for prec in ("double", "single"):
        be = _be(em, prec)
        prob = _em_problem(em, be, (128, 128, 128), 8.0, "diag", center=(0.0, 0.0, 0.0), radii=(0.3, 0.2, 0.15))
        op = em.Operator(prob)
        box = support_operator(op).box
        u = be.array(_rand((3, 128, 128, 128), 4).astype(be.complex_dtype))
        assert np.array_equal(_host(op.matvec_window(u, box.lo, box.n)), _host(op.matvec(u)))
        with pytest.raises(lib.HipLibraryError):
            op.matvec_window(u, (100, 0, 0), (40, 8, 8))
    N, lo, n = (40, 40, 40), (6, 11, 2), (25, 20, 30)
    be = _be(em, "double")
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    eta = 1.0 + _boxed(N, lo, n, 5) * 0.4
    prob = em.acoustics.AcousticProblem(grid=grid, eta=be.array(eta), wave=be.array(_rand(N, 6)), k0=6.0,
                                        volume=1.0)
    op = em.acoustics.AcousticOperator(prob)
    u = be.array(_rand((2,) + N, 7))
    assert np.array_equal(_host(op.matvec_window(u, lo, n)), _host(op.matvec(u)))


def test_restricted_solves_match_the_oracle_full_solve(em):
    """SupportRestricted(BiCGStab) and SupportRestricted(MixedBiCGStab) reach
    the requested true relative residual of the FULL FP64 operator and agree
    with the oracle's full-grid BiCGStab solution."""
    from oracle import solvers as osolv
    from oracle import vie
    N, k = N0, 6.0
    _say(f"[support solve] N={N} k={k}")
    be = _be(em, "double")
    prob = _em_problem(em, be, N, k, "diag")
    op = em.Operator(prob)
    f = prob.wave
    chi = _host(prob.eps_tensor)
    oop = _oracle_op(chi, vie.kernel_hat(N, (1, 1, 1), k))
    ref = osolv.bicgstab(oop, _host(f), osolv.Config(max_iter=400, rtol=1e-11))
    assert ref.status == "converged"
    for solver, rtol, sol_tol in ((em.BiCGStab(em.SolverConfig(max_iter=400, rtol=1e-10)), 1e-10, 1e-8),
                                  (em.MixedBiCGStab(em.SolverConfig(max_iter=400, rtol=1e-8)), 1e-8, 1e-6)):
        s = em.SupportRestricted(solver)
        j = em.judge_solve(s, op, f)
        res = j.result
        _say(f"  {type(solver).__name__}: {res.status} it={res.iterations} true={j.true_residual:.3e} "
             f"box={res.extras['support']}")
        assert res.status == "converged" and j.qualified
        assert j.true_residual < rtol
        assert osolv.true_relative_residual(oop, _host(res.u), _host(f)) < rtol * (1 + 1e-6)
        assert _rel(_host(res.u), ref.u) < sol_tol
        assert res.residual_history[-1] < rtol


@pytest.mark.parametrize("name", ["SIM", "TwoStep"])
def test_restricted_host_solvers(em, name):
    """SupportRestricted around the host-driven solvers (SIM with the
    operator's (mu, R) enclosure, TwoStep with its rmatvec on the box): the
    full operator's true residual meets the tolerance and the field agrees
    with the full-grid BiCGStab solution."""
    N, k = (24, 20, 16), 1.0
    _say(f"[support {name}] N={N}")
    be = _be(em, "double")
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    chi = em.geometry.sample_contrast_tensor(grid, background_eps_r=1.0, feature_eps_r=1.3 + 0.1j,
                                             geometry=em.geometry.Ellipsoid(center=(0.05, 0.0, -0.1),
                                                                            radii=(0.2, 0.15, 0.25)))
    wave = em.flat_wave_vec(grid, k=k, orient=(0, 0, 1), amplitude=(1, 0, 0))
    prob = em.Problem(grid=grid, eps_tensor=chi, wave=wave, k0=k, volume=1.0)
    op = em.Operator(prob)
    cfg = em.SolverConfig(max_iter=400, rtol=1e-7, mu=1.0 + 0j, radius=0.5)
    j = em.judge_solve(em.SupportRestricted(getattr(em, name)(cfg)), op, wave)
    res = j.result
    _say(f"  {res.status} it={res.iterations} true={j.true_residual:.3e} box={res.extras['support']}")
    assert res.extras["support"]["fraction"] < 0.5
    assert res.status == "converged" and j.qualified and j.true_residual < 1e-7
    ref = em.BiCGStab(em.SolverConfig(max_iter=400, rtol=1e-11)).solve(op, wave)
    assert _rel(_host(res.u), _host(ref.u)) < 1e-5


def test_restricted_acoustic_solve(em):
    from oracle import solvers as osolv
    from oracle import vie
    N, k = (32, 32, 24), 8.0
    _say(f"[support acoustic solve] N={N}")
    be = _be(em, "double")
    grid = em.Grid(N=N, L=(1.0, 1.0, 1.0), center=(0.0, 0.0, 0.0), backend=be)
    eta = em.acoustics.materials.eta_inclusions(grid, [((-0.1, 0.05, 0.0), (0.08, 0.08, 0.08), 2.0 + 0.1j),
                                                       ((0.1, 0.0, 0.05), (0.06, 0.06, 0.06), 1.5 + 0.2j)])
    prob = em.acoustics.make_acoustic_problem(grid, eta, k0=k)
    op = em.acoustics.AcousticOperator(prob)
    f = prob.wave
    chi = _host(eta) - 1.0
    oop = _oracle_op(chi, vie.acoustic_kernel_hat(N, (1, 1, 1), k), acoustic=True)
    ref = osolv.bicgstab(oop, _host(f), osolv.Config(max_iter=400, rtol=1e-11))
    j = em.judge_solve(em.SupportRestricted(em.BiCGStab(em.SolverConfig(max_iter=400, rtol=1e-10))), op, f)
    assert j.qualified and j.true_residual < 1e-10
    assert _rel(_host(j.result.u), ref.u) < 1e-8


def test_restriction_follows_contrast_edits_and_falls_back(em):
    """An in-place edit that moves the support re-plans the box; a contrast
    that fills most of the grid runs the inner solver on the full operator,
    with the inner solver's own result."""
    from em3d_amd.support import support_operator
    N, k = N0, 4.0
    _say(f"[support edits] N={N}")
    be = _be(em, "double")
    prob = _em_problem(em, be, N, k, "full")
    op = em.Operator(prob)
    cfg = em.SolverConfig(max_iter=400, rtol=1e-10)
    box0 = support_operator(op).box
    prob.eps_tensor[0, 0, 1, 1, 1] = 0.5            # a cell far from the ellipsoid, in place
    box1 = support_operator(op, max_fraction=1.0).box   # (32, 16, 16): more than half the grid
    assert box1 != box0 and box1.lo[0] <= 1 and box1.lo[2] <= 1
    j = em.judge_solve(em.SupportRestricted(em.BiCGStab(cfg), max_fraction=1.0), op, prob.wave)
    assert j.qualified and j.true_residual < 1e-10
    plain = em.BiCGStab(cfg).solve(op, prob.wave)
    fb = em.SupportRestricted(em.BiCGStab(cfg), max_fraction=1e-3).solve(op, prob.wave)
    assert "support" not in (fb.extras or {})
    assert fb.iterations == plain.iterations and np.array_equal(_host(fb.u), _host(plain.u))


def test_restricted_batch_equals_restricted_singles(em):
    """solve_batch on the box: every right-hand side equals its own restricted
    single solve (plane waves of one amplitude share the rescaled tolerance)."""
    import torch
    N, k = N0, 5.0
    _say(f"[support batch] N={N}")
    be = _be(em, "double")
    prob = _em_problem(em, be, N, k, "diag")
    op = em.Operator(prob)
    grid = prob.grid
    waves = [em.flat_wave_vec(grid, k=k, orient=o, amplitude=a)
             for o, a in (((0, 0, 1), (1, 0, 0)), ((1, 0, 0), (0, 1, 0)), ((0, 1, 0), (0, 0, 1)))]
    rhs = torch.stack(waves)
    s = em.SupportRestricted(em.BiCGStab(em.SolverConfig(max_iter=400, rtol=1e-9)))
    batch = s.solve_batch(op, rhs)
    for b, w in enumerate(waves):
        one = s.solve(op, w)
        assert batch[b].iterations == one.iterations
        assert _rel(_host(batch[b].u), _host(one.u)) < 1e-13
        assert em.true_relative_residual(op, batch[b].u, w) < 1e-9
