"""Multigrid solver: the CPU statement (oracle/mg.py) is pinned to the
reference's converged JPCG answer and to dense least squares; the device MG
is then checked against both."""
import numpy as np
import pytest

import oracle as O
import oracle.mg as M
import rheon_amd as R
from rheon_amd.scenes import GridSpec, courant_dt, fourier_modes, taylor_green_noise

from helpers import config1, ogrid, rel_err


def test_mg_statement_operator_matches_reference_apply():
    spec = GridSpec((8, 6, 4), (0.3, 0.2, 0.5))
    L = M.Level(spec.counts, spec.spacing, 1.7)
    x = np.random.default_rng(1).standard_normal(spec.cell_len)
    got = L.apply(x.reshape(4, 6, 8)).ravel()
    want = O.apply_full(ogrid(spec), 1.7, x)
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))


def test_mg_statement_levels_and_transfers():
    lv = M.build_levels((64, 32, 16), (1 / 64, 1 / 32, 1 / 16), 1.0)
    assert [L.n for L in lv] == [(64, 32, 16), (32, 16, 8), (16, 8, 4), (8, 4, 2)]
    lv = M.build_levels((128,) * 3, (1 / 128,) * 3, 1.0, levels=4)
    assert [L.n for L in lv][-1] == (16, 16, 16)
    # prolongation reproduces constants and linear fields away from walls
    xc = np.ones((4, 4, 4))
    assert np.allclose(M.prolong(xc, (8, 8, 8)), 1.0)
    # restriction preserves the total (integrated residuals add)
    r = np.random.default_rng(2).standard_normal((8, 8, 8))
    assert abs(M.restrict(r).sum() - r.sum()) < 1e-12


def test_mg_statement_chebyshev_weights():
    """n sweeps with the Chebyshev weights apply prod_k (1 - tau_k lam) to
    an eigencomponent lam of D^-1 K: on [lo, 2] its largest magnitude is
    1 / T_n((2 + lo) / (2 - lo)) (T_n = cosh(n acosh .)), reached with
    alternating signs at the n + 1 mapped extrema of T_n."""
    for n in range(1, 7):
        tau = M.sweep_weights(n)
        lo = 0.3 if n <= 2 else 0.25
        poly = lambda lam: np.prod([1.0 - t * lam for t in tau], axis=0)
        bound = 1.0 / np.cosh(n * np.arccosh((2 + lo) / (2 - lo)))
        lam = np.linspace(lo, 2.0, 20001)
        assert abs(np.max(np.abs(poly(lam))) - bound) < 1e-9
        ext = 0.5 * (2 + lo) + 0.5 * (2 - lo) * np.cos(np.pi * np.arange(n + 1) / n)
        vals = poly(ext)
        assert np.allclose(np.abs(vals), bound, rtol=1e-12)
        assert np.all(np.sign(vals[1:]) != np.sign(vals[:-1]))
        assert poly(np.array([0.0]))[0] == 1.0  # the constant (coarse) mode is left to the coarse grid
    assert M.sweep_weights(3, "jacobi") == [6.0 / 7.0] * 3


def test_mg_statement_chebyshev_beats_jacobi():
    """Same sweep count per cycle, fewer cycles: Chebyshev V(2,2) vs Jacobi V(2,2)."""
    spec, vel, dt = config1(32, seed=3)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    args = (spec.counts, spec.spacing, 1.0, rhs, dt, 1e-10, 1e-14, 1.0, 60)
    _, _, cj, _, _ = M.solve(*args, pre=2, post=2, smoother="jacobi")
    _, _, cc, _, _ = M.solve(*args, pre=2, post=2)
    assert cc < cj


def test_mg_statement_coarse_sweeps():
    """Coarse-level sweep counts: 0 means the finest level's (same history),
    and more coarse smoothing gives a more accurate coarse-grid correction
    (fewer cycles on a 64^3 Fourier field, 12 -> 10)."""
    spec = GridSpec.unit_box(64)
    vel = fourier_modes(spec, seed=2026, modes=32, kmax=16, amplitude=1.0, dtype=np.float32)
    dt = courant_dt(spec, vel, 0.5)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    args = (spec.counts, spec.spacing, 1.0, rhs, dt, 1e-10, 1e-14, 1.0, 60)
    h0, h1 = [], []
    _, _, c0, _, _ = M.solve(*args, dtype=np.float32, history=h0)
    _, _, c1, _, _ = M.solve(*args, dtype=np.float32, coarse_pre=3, coarse_post=1, history=h1)
    assert c0 == c1 and h0 == h1
    _, _, c2, _, _ = M.solve(*args, dtype=np.float32, coarse_pre=6, coarse_post=3)
    assert c2 <= c0 - 2


def test_mg_statement_converges_to_reference_pressure():
    spec, vel, dt = config1(32, seed=3)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(1e-12, 1e-14, 1.0, 5000), threads=4)
    st, p, cycles, l2, rmax = M.solve(spec.counts, spec.spacing, 1.0, rhs, dt, 1e-12, 1e-14, 1.0, 60)
    assert st == "ok" and cycles < 25
    assert p[0] == 0.0
    assert rel_err(p, ref.pressure) < 1e-9


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [R.F32, R.F64])
def test_device_mg_reaches_reference_pressure(precision):
    spec, vel, dt = config1(64, seed=4, dtype=np.float32)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(1e-12, 1e-14, 1.0, 5000), threads=8)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", precision)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1.0, 100))
    assert rep.ok, R.STATUS_NAMES.get(rep.status)
    assert rep.iterations < 25
    p = ws.get_pressure()
    assert p[0] == 0.0
    assert rel_err(p, ref.pressure) < 1e-8
    # the reported residual is that of the returned (gauged) pressure
    res = rhs - O.apply_full(g, 1.0, p)
    assert abs(np.sqrt(np.sum(res * res)) - rep.true_residual_l2) <= 1e-6 * rep.true_residual_l2


@pytest.mark.gpu
@pytest.mark.parametrize("counts", [(40, 24, 16), (60, 60, 60), (33, 17, 7), (8, 8, 8)])
def test_device_mg_uneven_hierarchies(counts):
    """Dense coarsest + one-workgroup tail (40x24x16), CG coarsest (60^3 ->
    15^3), no coarsening at all (odd axes), tiny grids."""
    spec = GridSpec.unit_box(*counts)
    vel = taylor_green_noise(spec, 7, 0.1, np.float32)
    dt = 0.25 * min(spec.spacing)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(1e-12, 1e-14, 1.0, 20000), threads=8)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1.0, 400))
    assert rep.ok, R.STATUS_NAMES.get(rep.status)
    assert rel_err(ws.get_pressure(), ref.pressure) < 1e-7


@pytest.mark.gpu
@pytest.mark.parametrize("coarse", [(0, 0), (6, 3)])
def test_device_mg_cycles_track_the_cpu_statement(coarse):
    spec, vel, dt = config1(32, seed=5)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    hist = []
    M.solve(spec.counts, spec.spacing, 1.0, rhs, dt, 0.0, 0.0, 0.0, 6, history=hist,
            coarse_pre=coarse[0], coarse_post=coarse[1])
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F64, mg_coarse_sweeps=coarse)
    for k in range(1, 6):
        rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(0.0, 0.0, 0.0, k))
        assert rep.status == R.ITERATION_LIMIT
        # residual after k cycles agrees with the statement (coarsest solved
        # by CG on the device vs least squares on the CPU)
        assert abs(rep.true_residual_l2 - hist[k]) <= 1e-6 * hist[k], (k, rep.true_residual_l2, hist[k])


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [R.F32, R.F64])
def test_device_mg_no_cycle_leaves_zero_pressure(precision):
    """A solve that ends before its first cycle (zero right-hand side, or no
    iterations allowed) returns p = 0 exactly, whatever the output buffer
    held (the FP32 four-cell path does not read p = 0 and fills it only in
    that case)."""
    spec, vel, dt = config1(64, seed=4, dtype=np.float32)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", precision)
    assert ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-8, 1e-14, 1.0, 100)).ok
    assert np.any(ws.get_pressure() != 0.0)
    rep = ws.solve_rhs(1.0, np.zeros_like(rhs), dt, R.PressureSettings(1e-8, 1e-14, 1.0, 100))
    assert rep.ok and rep.iterations == 0
    assert np.all(ws.get_pressure().view(np.uint64) == 0)
    assert ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-8, 1e-14, 1.0, 100)).ok
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-8, 1e-14, 1.0, 0))
    assert rep.status == R.ITERATION_LIMIT
    assert np.all(ws.get_pressure().view(np.uint64) == 0)


@pytest.mark.gpu
def test_device_mg_four_levels_config2_shape():
    """128^3 FP32 V-cycles, 4 levels (128 -> 16), to rel 1e-5."""
    spec, vel, dt = config1(128, seed=6, dtype=np.float32)
    rhs = R.build_rhs(spec, vel, dt)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32, mg_levels=4)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-5, 1e-14, 1.0, 100))
    assert rep.ok and rep.iterations <= 12
    out = R.correct_velocity(spec, 1.0, ws.get_pressure(), dt, vel)
    assert R.divergence_max(spec, out) < 1e-1


@pytest.mark.gpu
def test_device_mg_fourier_projection_divergence():
    spec = GridSpec.unit_box(64)
    vel = fourier_modes(spec, 9, dtype=np.float32)
    dt = 0.5 / 64
    rhs = R.build_rhs(spec, vel, dt)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1.0, 100))
    assert rep.ok
    out = R.correct_velocity(spec, 1.0, ws.get_pressure(), dt, vel)
    # f32 face rounding floor ~ |u| eps32 / h
    assert R.divergence_max(spec, out) < 64 * 64 * 6e-8


def wall_blocks(spec):
    """Solid blocks against the walls and one in the interior: rows next to a
    wall or a block are not plain box rows, their neighbours are plain rows
    with non-plain neighbour rows, and one block sits in a single z-plane."""
    nx, ny, nz = spec.counts
    s = np.zeros((nz, ny, nx), np.uint8)
    s[:, :3, :4] = 1             # along the x- / y- edge, every plane
    s[nz // 2, ny // 2, nx - 4:] = 1   # one row's end against the x+ wall
    s[1:3, ny - 2:, 8:12] = 1    # against the y+ wall near the z- wall
    s[nz // 3, ny // 3:ny // 3 + 2, nx // 2:nx // 2 + 4] = 1
    return s.ravel()


@pytest.mark.gpu
@pytest.mark.parametrize("counts,sphere", [((64, 32, 48), False), ((64, 32, 48), True), ((40, 24, 16), False),
                                           ((40, 24, 16), True), ((256, 256, 128), False),
                                           ((256, 256, 128), True), ((64, 32, 48), "walls"),
                                           ((40, 24, 16), "walls")])
def test_device_mg_quad_kernels_match_scalar(counts, sphere):
    """The four-cells-per-lane FP32 kernels give the scalar kernels' values
    bit for bit (same per-cell arithmetic).  256x256x128: its first coarse
    level has 2^20 cells, so it also runs the four-cell kernels on a level
    with face fractions (with the sphere).  On obstacle grids the four-cell
    kernels take the box form on rows flagged plain (no mask load) and the
    mask form elsewhere; "walls" puts solids against the walls so that
    both forms, and rows whose neighbour rows differ, occur."""
    import rheon_amd as R
    from rheon_amd.scenes import centred_sphere, close_obstacle_faces
    spec = GridSpec((counts), tuple(1.0 / n for n in counts))
    vel = fourier_modes(spec, 9, modes=8, kmax=6, dtype=np.float32)
    solid = wall_blocks(spec) if sphere == "walls" else (centred_sphere(spec, 0.25) if sphere else None)
    if solid is not None:
        vel = close_obstacle_faces(spec, solid, vel)
    dt = courant_dt(spec, vel)
    rhs = O.build_rhs(ogrid(spec), vel, dt)
    out = []
    for variant in (0, R.VARIANT_SCALAR_MG):
        ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32, kernel_variant=variant)
        if solid is not None:
            ws.set_solid(solid)
        rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1e-3, 60))
        assert rep.ok, R.STATUS_NAMES.get(rep.status)
        out.append((rep.iterations, ws.get_pressure()))
        ws.close()
    assert out[0][0] == out[1][0]
    assert np.array_equal(out[0][1].view(np.uint64), out[1][1].view(np.uint64))


@pytest.mark.gpu
@pytest.mark.parametrize("counts,sphere", [((44, 28, 60), False), ((44, 28, 60), True), ((33, 17, 7), False)])
def test_device_mg_padded_coarsening(counts, sphere):
    """Odd axes on a level too large for the dense coarsest inverse are
    coarsened with ceil(n/2) (the missing children act as solid cells); the
    solve still reaches the reference method's converged pressure."""
    from rheon_amd.scenes import centred_sphere, close_obstacle_faces, isotropic_spec
    spec = isotropic_spec(counts)
    vel = fourier_modes(spec, 13, modes=8, kmax=6, dtype=np.float32)
    solid = centred_sphere(spec, 0.15) if sphere else None
    if solid is not None:
        vel = close_obstacle_faces(spec, solid, vel)
    dt = courant_dt(spec, vel)
    g = ogrid(spec)
    rhs = O.build_rhs(g, vel, dt)
    ws = R.PressureWorkspace(spec, "mixed-mg-dc", R.F32)
    if solid is not None:
        ws.set_solid(solid)
    rep = ws.solve_rhs(1.0, rhs, dt, R.PressureSettings(1e-10, 1e-14, 1e-3, 40))
    assert rep.ok, R.STATUS_NAMES.get(rep.status)
    assert rep.iterations <= 30, rep.iterations
    p = ws.get_pressure()
    if solid is None:
        ref = O.solve_rhs(g, 1.0, rhs, dt, O.Settings(1e-12, 1e-16, 1e-3, 4000), "rbsgs-pcg")
    else:
        ob = O.obstacle_mask(g, solid)
        ref = O.solve_rhs_masked(g, 1.0, ob, rhs, dt, O.Settings(1e-12, 1e-16, 1e-3, 4000), "rbsgs-pcg")
    assert ref.status == O.OK
    assert rel_err(p, ref.pressure) < 1e-7
