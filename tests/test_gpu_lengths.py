"""GPU oracle parity for every FFT length and kernel path the workloads run.

The operator picks its pass kernels by the doubled length of each axis
(2^k, 3 2^k, 5 2^k register kernels; 1024-thread row-pair tiles for 512-point
scalar x rows) and by the pass the axis lands in (z forward/inverse, strided
y, x with the spectral product).  Thin grids put one long axis in each pass
position, so every length the 64^3 / 96^3 / 128^3 / 256^3 / 80^3 workloads
run is compared with the oracle's complex128 FFT operator
(src/em3d/operator.py:353-398, src/em3d/acoustics/operator.py:40-67) at the
reference's own tolerances (tests/test_operator_vs_dense.py:81-116 uses 1e-10;
here FP64 1e-12, FP32 1e-6 relative L2).

Batched fields, packed-group members and every autotuner variant are compared
with the oracle too, not only with other HIP results.
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


def _em_problem(em, be, N, L, k, chi):
    grid = em.Grid(N=N, L=L, center=(0.0, 0.0, 0.0), backend=be)
    wave = be.array(_rand((3,) + N, 1).astype(be.complex_dtype))
    return em.Problem(grid=grid, eps_tensor=be.array(chi.astype(be.complex_dtype)), wave=wave, k0=k,
                      volume=grid.dv * int(np.prod(N)))


# (N, L, k): the long axis is doubled to 256, 512, 192 or 160 points and sits
# in the x pass (axis 0), the strided y passes (axis 1) or the z passes (axis 2)
THIN = [
    ((128, 2, 8), (1.0, 0.8, 0.6), 2.0), ((2, 128, 8), (0.6, 1.0, 0.8), 3.0), ((2, 2, 128), (0.5, 0.7, 1.0), 4.0),
    ((256, 2, 8), (1.0, 0.1, 0.1), 6.0), ((2, 256, 8), (0.2, 1.0, 0.3), 5.0), ((4, 2, 256), (0.2, 0.2, 1.0), 3.0),
    ((96, 3, 16), (1.0, 0.3, 0.5), 3.0), ((2, 96, 8), (0.4, 1.0, 0.5), 4.0), ((2, 8, 96), (0.5, 0.7, 1.0), 6.0),
    ((80, 2, 8), (1.0, 0.4, 0.3), 5.0), ((2, 80, 8), (0.3, 1.0, 0.6), 3.0), ((4, 4, 80), (0.5, 0.5, 1.2), 3.0),
]


class _Tiles:
    """The plan's x and z passes on their default tiles (variant 0) for the
    block, instead of the small-grid tiles a thin grid gets by default; the
    plan (shared by every operator of the shape) is restored after."""

    def __init__(self, em, N, ncomp, be, tiles):
        from em3d_amd.operator import HipPlan
        self.plan = HipPlan.get(N, ncomp, 1, be.precision, be.device_index)
        self.tiles = tiles

    def __enter__(self):
        self.old = self.plan.variants()
        if self.tiles == "default":
            self.plan.set_variants([self.old[0], 0, 0])
        return self

    def __exit__(self, *exc):
        self.plan.set_variants(self.old)
        return False


@pytest.mark.parametrize("tiles", ["auto", "default"])
@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("N,L,k", THIN)
def test_em_lengths_vs_oracle(em, N, L, k, prec, tol, tiles):
    """auto: the plan's own variants (small-grid tiles on these thin grids);
    default: the tiles large grids run."""
    from oracle import vie
    _say(f"[em length] N={N} (doubled {tuple(2 * n for n in N)}) prec={prec} tiles={tiles}")
    be = _be(em, prec)
    chi = _rand((3, 3) + N, 8) * 0.3
    op = em.Operator(_em_problem(em, be, N, L, k, chi))
    Kh = vie.kernel_hat(N, L, k)
    u = _rand((3,) + N, 42)
    ud = be.array(u.astype(be.complex_dtype))
    with _Tiles(em, N, 3, be, tiles):
        assert _rel(_host(op.matvec(ud)), vie.em_matvec(chi, u, Kh)) < tol
        assert _rel(_host(op.rmatvec(ud)), vie.em_rmatvec(chi, u, Khat=Kh)) < tol


@pytest.mark.parametrize("tiles", ["auto", "default"])
@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
@pytest.mark.parametrize("N,L,k", THIN)
def test_acoustic_lengths_vs_oracle(em, N, L, k, prec, tol, tiles):
    """Scalar passes: the (256, 2, 8) case runs the 512-point x pass with
    1024-thread blocks and both rows of a ky pair in one schedule."""
    from oracle import vie
    _say(f"[acoustic length] N={N} prec={prec} tiles={tiles}")
    be = _be(em, prec)
    grid = em.Grid(N=N, L=L, center=(0.0, 0.0, 0.0), backend=be)
    eta = 1.0 + _rand(N, 10) * 0.3
    prob = em.acoustics.AcousticProblem(grid=grid, eta=be.array(eta.astype(be.complex_dtype)),
                                        wave=be.array(_rand(N, 2).astype(be.complex_dtype)), k0=k, volume=1.0)
    op = em.acoustics.AcousticOperator(prob)
    chi = _host(prob.chi).astype(np.complex128)
    u = _rand(N, 11)
    ud = be.array(u.astype(be.complex_dtype))
    with _Tiles(em, N, 1, be, tiles):
        assert _rel(_host(op.matvec(ud)), vie.acoustic_matvec(chi, u, vie.acoustic_kernel_hat(N, L, k))) < tol
        assert _rel(_host(op.rmatvec(ud)),
                    vie.acoustic_rmatvec(chi, u, vie.acoustic_kernel_hat(N, L, k, adjoint=True))) < tol


@pytest.mark.parametrize("prec,tol", [("double", 1e-12), ("single", 1e-6)])
def test_batched_members_vs_oracle(em, prec, tol):
    """Every field of a batched call (doubled axes 64 / 96 / 160) against the oracle."""
    from oracle import vie
    N, L, k = (32, 48, 80), (0.8, 1.0, 1.2), 7.0
    _say(f"[batch members] N={N} x 3 prec={prec}")
    be = _be(em, prec)
    chi = np.zeros((3, 3) + N, complex)
    d = _rand((3,) + N, 4) * 0.4
    for a in range(3):
        chi[a, a] = d[a]
    op = em.Operator(_em_problem(em, be, N, L, k, chi))
    Kh = vie.kernel_hat(N, L, k)
    U = _rand((3, 3) + N, 13)
    Ud = be.array(U.astype(be.complex_dtype))
    Y, YA = _host(op.matvec(Ud)), _host(op.rmatvec(Ud))
    for b in range(3):
        assert _rel(Y[b], vie.em_matvec(chi, U[b], Kh)) < tol, b
        assert _rel(YA[b], vie.em_rmatvec(chi, U[b], Khat=Kh)) < tol, b


def test_packed_members_vs_oracle(em):
    """The packed 32^3 / 48^3 / 64^3 / 80^3 group of the uneven-batch workload
    (one random ellipsoid each, autotuned kernels, one graph launch): each
    member against the oracle, FP32 1e-6."""
    from em3d_amd.packed import PackedOperators
    from oracle import inputs, vie
    _say("[packed members] 32/48/64/80 autotuned graph vs oracle")
    be = _be(em, "single")
    rng = np.random.default_rng(40)
    ops, chis, ks = [], [], []
    for i, n in enumerate((32, 48, 64, 80)):
        N = (n, n, n)
        radii = tuple(rng.uniform(0.15, 0.4, 3))
        eps = np.diag(rng.uniform(1.5, 4.0, 3) + 1j * rng.uniform(0.0, 0.5, 3))
        mask = inputs.ellipsoid_center_mask(N, (1, 1, 1), (0, 0, 0), ell_center=(0, 0, 0), radii=radii)
        chi = inputs.contrast_from_weights(mask.astype(float), background_eps_r=1.0, feature_eps_r=eps)
        ops.append(em.Operator(_em_problem(em, be, N, (1.0, 1.0, 1.0), 8.0, chi)))
        chis.append(chi)
        ks.append(8.0)
    group = PackedOperators(ops, autotune=True)
    _say(f"[packed members] variants {[t['variants'] for t in group.tuned]}")
    fields = [_rand((3,) + op._N, 60 + j) for j, op in enumerate(ops)]
    for rep in range(2):
        got = group.matvec([be.array(f.astype(np.complex64)) for f in fields])
        for j, op in enumerate(ops):
            ref = vie.em_matvec(chis[j], fields[j], vie.kernel_hat(op._N, (1.0, 1.0, 1.0), ks[j]))
            assert _rel(_host(got[j]), ref) < 1e-6, (rep, j)


# (N, pass kind, precision): x pass at 256 / 192 / 128, strided passes at
# 256 / 192, z passes at 256 / 192, the FP64 x and strided passes at 256 -- every variant
# the autotuner may pick, and the generic kernels (-1)
VARIANT_CASES = [((128, 2, 8), 1, "single"), ((96, 2, 8), 1, "single"), ((64, 2, 8), 1, "single"),
                 ((2, 128, 8), 0, "single"), ((2, 96, 8), 0, "single"), ((2, 2, 128), 2, "single"),
                 ((2, 2, 96), 2, "single"), ((128, 2, 8), 1, "double"), ((2, 128, 8), 0, "double")]


@pytest.mark.parametrize("N,kind,precision", VARIANT_CASES)
def test_every_variant_vs_oracle_on_gpu(em, N, kind, precision):
    from em3d_amd.operator import HipPlan
    from oracle import vie
    L, k = (1.0, 0.6, 0.8), 3.0
    be = _be(em, precision)
    chi = _rand((3, 3) + N, 21) * 0.3
    op = em.Operator(_em_problem(em, be, N, L, k, chi))
    u = _rand((3,) + N, 22)
    ud = be.array(u.astype(be.complex_dtype))
    Kh = vie.kernel_hat(N, L, k)
    ref, refa = vie.em_matvec(chi, u, Kh), vie.em_rmatvec(chi, u, Khat=Kh)
    plan = HipPlan.get(N, 3, 1, be.precision)
    tol = 1e-6 if precision == "single" else 1e-12
    n = plan.variant_counts()[kind]
    assert n >= 2
    try:
        for v in range(-1, n):
            _say(f"[variant] N={N} kind={kind} precision={precision} variant={v}")
            vs = [0, 0, 0]
            vs[kind] = v
            plan.set_variants(vs)
            assert _rel(_host(op.matvec(ud)), ref) < tol, v
            assert _rel(_host(op.rmatvec(ud)), refa) < tol, v
    finally:
        plan.set_variants([0, 0, 0])


def test_acoustic_256_full_size_properties(em):
    """Acoustic 256^3 (eta = 1 plus 32 seeded inclusions, k0 = 10): adjoint
    identity, linearity, FP32 against FP64 and eta = 1 as the identity.  The
    FP32 operator takes the FP64 eta cast to complex64: eta built on FP32
    axes tests its inclusions in FP32 and differs at boundary cells."""
    import torch
    from em3d_amd.acoustics.materials import eta_inclusions, random_spheres
    N = (256, 256, 256)
    _say("[acoustic 256 properties]")
    be64, be32 = _be(em, "double"), _be(em, "single")
    grid = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=be64)
    eta = eta_inclusions(grid, random_spheres(32, seed=3))
    prob64 = em.acoustics.make_acoustic_problem(grid, eta, k0=10.0)
    op64 = em.acoustics.AcousticOperator(prob64)
    grid32 = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=be32)
    prob32 = em.acoustics.AcousticProblem(grid=grid32, eta=eta.to(torch.complex64),
                                          wave=prob64.wave.to(torch.complex64), k0=10.0, volume=1.0)
    op32 = em.acoustics.AcousticOperator(prob32)
    g = torch.Generator(device="cuda").manual_seed(5)
    u = torch.randn(N, dtype=torch.complex128, device="cuda", generator=g)
    v = torch.randn(N, dtype=torch.complex128, device="cuda", generator=g)
    au, ahv = op64.matvec(u), op64.rmatvec(v)
    lhs = complex(torch.vdot(au.flatten(), v.flatten()))
    rhs = complex(torch.vdot(u.flatten(), ahv.flatten()))
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    a = 0.3 + 0.8j
    lin = op64.matvec(a * u + v) - (a * au + op64.matvec(v))
    assert float(torch.linalg.vector_norm(lin) / torch.linalg.vector_norm(au)) < 1e-13
    y32 = op32.matvec(u.to(torch.complex64)).to(torch.complex128)
    assert float(torch.linalg.vector_norm(y32 - au) / torch.linalg.vector_norm(au)) < 1e-6
    ya32 = op32.rmatvec(v.to(torch.complex64)).to(torch.complex128)
    assert float(torch.linalg.vector_norm(ya32 - ahv) / torch.linalg.vector_norm(ahv)) < 1e-6
    grid = prob64.grid
    one = em.acoustics.AcousticProblem(grid=grid, eta=torch.ones(N, dtype=torch.complex128, device="cuda"),
                                       wave=prob64.wave, k0=10.0, volume=1.0)
    assert torch.equal(em.acoustics.AcousticOperator(one).matvec(u), u)


def test_96_batch_full_size_properties(em):
    """The 96^3 sphere (eps 4, k0 8) with 8 of its incident fields in one
    batched call: each member equals its single call, satisfies the adjoint
    identity, and the FP32 batch agrees with the FP64 batch (same contrast
    and fields, cast)."""
    import torch
    from em3d_amd.sources import incident_set, plane_waves
    N, k = (96, 96, 96), 8.0
    _say("[96^3 batch properties] 8 RHS")
    res = {}
    g64 = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=_be(em, "double"))
    chi64 = em.geometry.sample_contrast_tensor(g64, background_eps_r=1.0, feature_eps_r=4.0,
                                               geometry=em.geometry.Ellipsoid(center=(0, 0, 0), radii=(0.3,) * 3))
    o, a = incident_set(4, seed=4)
    waves64 = plane_waves(g64, k=k, orients=o, amplitudes=a)
    for prec in ("double", "single"):
        be = _be(em, prec)
        grid = em.Grid(N=N, L=(1, 1, 1), center=(0, 0, 0), backend=be)
        cdt = be.precision.torch_complex_dtype
        chi, waves = chi64.to(cdt), waves64.to(cdt)
        op = em.Operator(em.Problem(grid=grid, eps_tensor=chi, wave=waves[0], k0=k, volume=1.0))
        Y, YA = op.matvec(waves), op.rmatvec(waves)
        for b in (0, 3, 7):
            s = op.matvec(waves[b].contiguous())
            tol = 1e-13 if prec == "double" else 1e-6
            assert float(torch.linalg.vector_norm(Y[b] - s) / torch.linalg.vector_norm(s)) < tol, b
            lhs = complex(torch.vdot(Y[b].flatten(), waves[(b + 1) % 8].flatten()))
            rhs = complex(torch.vdot(waves[b].flatten(), YA[(b + 1) % 8].flatten()))
            assert abs(lhs - rhs) / abs(lhs) < (1e-12 if prec == "double" else 1e-5), b
        res[prec] = (Y.to(torch.complex128), YA.to(torch.complex128))
    for i in range(2):
        d = torch.linalg.vector_norm(res["single"][i] - res["double"][i]) / torch.linalg.vector_norm(res["double"][i])
        assert float(d) < 1e-6
