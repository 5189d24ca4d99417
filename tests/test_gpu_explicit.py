"""GPU parity of the explicit path against the oracle: updatePadding, FTCS
fast path, the expression interpreter for every operator, reductions."""
import numpy as np
import pytest

import opflow_amd as of
import oracle
from helpers import build_field, random_storage, rel_err, SEED

pytestmark = pytest.mark.gpu
DIRC, NEUM, PER, SYMM, ASYMM = 1, 2, 3, 5, 6


def _mesh(ctx, nodes, ranges=None):
    b = of.MeshBuilder(ctx).newMesh(*nodes)
    for d in range(len(nodes)):
        lo, hi = (ranges[d] if ranges else (0.0, 1.0))
        b.setMeshOfDim(d, lo, hi)
    return b.build(), oracle.Mesh(nodes, ranges)


def test_ghost_fill_random_configs(ctx):
    rng = np.random.default_rng(11)
    types = [DIRC, NEUM, SYMM, ASYMM]
    for trial in range(40):
        dim = int(rng.integers(1, 4))
        nodes = [int(rng.integers(5, 12)) for _ in range(dim)]
        loc = [int(rng.integers(0, 2)) for _ in range(dim)]
        bc, bcv, ext = [], [], []
        for q in range(dim):
            bc.append((PER, PER) if rng.random() < 0.3 else (int(rng.choice(types)), int(rng.choice(types))))
            bcv.append((float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1))))
            e = int(rng.integers(1, 3))
            ext.append((e, e))
        ranges = [(0.0, float(rng.uniform(0.5, 3.0))) for _ in range(dim)]
        mesh, om = _mesh(ctx, nodes, ranges)
        u = build_field(ctx, mesh, dim, loc, bc, bcv, ext=ext, padding=int(rng.integers(0, 3)))
        raw = random_storage(u, SEED + trial)
        u.upload(raw)
        u.updatePadding()
        got = u.numpy()
        ref = oracle.update_padding(u.desc, om, raw.copy())
        print(f"trial {trial} dim {dim} loc {loc} bc {bc} ext {ext}", flush=True)
        np.testing.assert_array_equal(got, ref)
        u.close()


@pytest.mark.parametrize("n,steps", [(32, 50), (256, 1000), (256, 13), (37, 11)])
def test_ftcs2d_dirichlet_bitexact(ctx, n, steps):
    """config 1 shape: 256^2 centre field, Dirichlet 0, 4 bumps, dt = 0.2 h^2, 1000 steps."""
    mesh, om = _mesh(ctx, [n + 1, n + 1])
    u = build_field(ctx, mesh, 2, 1, [(DIRC, DIRC), (DIRC, DIRC)])
    init = oracle.update_padding(u.desc, om, of.gen_bumps(u, SEED + 1, 4))
    u.upload(init)
    h = 1.0 / n
    coef = 0.2 * h * h * 1.0
    u.ftcs(coef, steps)
    got = u.numpy()
    ref = oracle.ftcs(u.desc, om, init, coef, steps)
    assert rel_err(got, ref) <= 1e-12
    np.testing.assert_array_equal(got, ref)  # same operation order, no FMA: identical bits
    u.close()


@pytest.mark.parametrize("n,steps", [(40, 30), (6, 9), (45, 17)])
def test_ftcs2d_nonpow2_and_mixed_bc(ctx, n, steps):
    # h = 1.3/n is not a power of two (the division path); a periodic y side, a
    # 6-cell grid narrower than the blocked kernel's tile plus halo, remainders
    mesh, om = _mesh(ctx, [n + 1, n + 1], [(0.0, 1.3), (0.0, 0.7)])
    u = build_field(ctx, mesh, 2, 1, [(NEUM, DIRC), (PER, PER)], [(0.3, -0.5), (0, 0)])
    init = oracle.update_padding(u.desc, om, random_storage(u, SEED + 2))
    u.upload(init)
    coef = 0.2 * (1.3 / n) ** 2
    u.ftcs(coef, steps)
    ref = oracle.ftcs(u.desc, om, init, coef, steps)
    np.testing.assert_array_equal(u.numpy(), ref)
    u.close()


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_ftcs3d_periodic(ctx, dtype):
    n, steps = 48, 12
    mesh, om = _mesh(ctx, [n + 1] * 3)
    u = build_field(ctx, mesh, 3, 1, [(PER, PER)] * 3, dtype=dtype)
    raw = random_storage(u, SEED + 3)
    init = oracle.update_padding(u.desc, om, raw)
    u.upload(init)
    h = 1.0 / n
    u.ftcs(0.2 * h * h, steps)
    ref = oracle.ftcs(u.desc, om, init, 0.2 * h * h, steps)
    got = u.numpy()
    if dtype == "f64":
        np.testing.assert_array_equal(got, ref)
    else:
        assert rel_err(got, ref) < 1e-5
    u.close()


@pytest.mark.parametrize("steps", [13, 14], ids=["remainder-1", "remainder-2"])
def test_ftcs3d_periodic_ragged_blocked(ctx, steps):
    """Temporal blocking on a grid that no tile divides, with a remainder of one
    step (a trailing single step) or two (a trailing two-step launch): equal bits
    to the oracle and to one launch per step."""
    n = [40, 36, 20]
    mesh, om = _mesh(ctx, [m + 1 for m in n], [(0.0, m / 32.0) for m in n])
    fields = []
    for tblock in (0, 1):
        u = build_field(ctx, mesh, 3, 1, [(PER, PER)] * 3)
        init = oracle.update_padding(u.desc, om, random_storage(u, SEED + 9))
        u.upload(init)
        nl, k = u.ftcs_plan(steps, tblock)
        if tblock == 1:
            assert (nl, k) == (steps, 1)
        else:
            assert k > 1 and nl == -(-steps // k)  # K-step launches, then one launch for the rest
        u.ftcs_blocked(0.2 / 32**2, steps, tblock)
        fields.append(u.numpy())
        u.close()
    ref = oracle.ftcs(u.desc, om, init, 0.2 / 32**2, steps)
    np.testing.assert_array_equal(fields[0], ref)
    np.testing.assert_array_equal(fields[1], ref)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_ftcs3d_periodic_full_size_blocked(ctx, dtype):
    """512^3 periodic (the benchmark shape): blocked and one-launch-per-step
    results are identical, and a Fourier mode decays by lambda^N."""
    n, steps = 512, 6
    mesh, om = _mesh(ctx, [n + 1] * 3)
    h = 1.0 / n
    coef = 0.2 * h * h
    xc = (np.arange(n) + 0.5) / n
    modes = (1, 2, 3)
    init = None
    got = []
    for tblock in (0, 1):
        u = build_field(ctx, mesh, 3, 1, [(PER, PER)] * 3, dtype=dtype)
        if init is None:
            init = np.zeros(u.shape)
            init[1:-1, 1:-1, 1:-1] = (np.cos(2 * np.pi * modes[2] * xc)[:, None, None] *
                                      np.cos(2 * np.pi * modes[1] * xc)[None, :, None] *
                                      np.cos(2 * np.pi * modes[0] * xc)[None, None, :])
            init = oracle.update_padding(u.desc, om, init)
        u.upload(init)
        u.ftcs_blocked(coef, steps, tblock)
        got.append(u.numpy())
        u.close()
    np.testing.assert_array_equal(got[0], got[1])
    lam = 1 - 4 * coef / h**2 * sum(np.sin(np.pi * m * h) ** 2 for m in modes)
    tol = 1e-12 if dtype == "f64" else 2e-6
    assert rel_err(got[0], lam**steps * init) < tol


def test_ftcs2d_blocked_equals_single_large(ctx):
    """2048^2 Dirichlet/Neumann, 19 steps: 8-step blocked launches + 3 single steps == 19 single steps."""
    n, steps = 2048, 19
    mesh, om = _mesh(ctx, [n + 1, n + 1])
    got = []
    for tblock in (0, 1):
        u = build_field(ctx, mesh, 2, 1, [(DIRC, NEUM), (NEUM, DIRC)], [(0.5, -1.0), (0.25, 0.0)])
        init = oracle.update_padding(u.desc, om, random_storage(u, SEED + 12))
        u.upload(init)
        u.ftcs_blocked(0.2 / n**2, steps, tblock)
        got.append(u.numpy())
        u.close()
    np.testing.assert_array_equal(got[0], got[1])


def test_ftcs3d_dirichlet_eigenmode_full_size(ctx):
    """Size-independent property at 256^3 (Dirichlet): the discrete eigenmode decays by lambda^N."""
    n, steps = 256, 10
    mesh, om = _mesh(ctx, [n + 1] * 3)
    u = build_field(ctx, mesh, 3, 1, [(DIRC, DIRC)] * 3)
    xc = (np.arange(n) + 0.5) / n
    s = np.sin(np.pi * xc)
    init = np.zeros(u.shape)
    init[1:-1, 1:-1, 1:-1] = s[:, None, None] * s[None, :, None] * s[None, None, :]
    init = oracle.update_padding(u.desc, om, init)
    u.upload(init)
    h = 1.0 / n
    u.ftcs(0.125 * h * h, steps)
    lam = 1 - 4 * 0.125 * 3 * np.sin(np.pi * h / 2) ** 2
    got = u.numpy()[1:-1, 1:-1, 1:-1]
    assert rel_err(got, lam ** steps * init[1:-1, 1:-1, 1:-1]) < 1e-12
    u.close()


def _sweep_case(ctx, expr_fn, fields_spec, nodes, out_spec, seed, op=0, ranges=None):
    mesh, om = _mesh(ctx, nodes, ranges)
    dim = len(nodes)
    fs, arrs = [], []
    for k, (loc, bc, bcv, ext) in enumerate(fields_spec):
        f = build_field(ctx, mesh, dim, loc, bc, bcv, ext=ext, name=f"f{k}")
        a = oracle.update_padding(f.desc, om, random_storage(f, seed + k, 0.5, 1.5))
        f.upload(a)
        fs.append(f)
        arrs.append(a)
    if out_spec is None:
        out, out_arr = fs[0], arrs[0]
    else:
        loc, bc, bcv, ext = out_spec
        out = build_field(ctx, mesh, dim, loc, bc, bcv, ext=ext, name="out")
        out_arr = oracle.update_padding(out.desc, om, random_storage(out, seed + 99, 0.5, 1.5))
        out.upload(out_arr)
    expr = expr_fn(*fs)
    nodes_, flds, root = expr.flatten()
    slot_arrays = [(arrs[next(i for i, g in enumerate(fs) if g is f)], f.desc) for f in flds]
    ref = oracle.assign(nodes_, root, slot_arrays, om, out_arr, out.desc, op,
                        update_padding=lambda r: oracle.update_padding(out.desc, om, r))
    out.assign(expr, op)
    got = out.numpy()
    return got, ref


C_DIRC2 = (1, [(DIRC, DIRC), (DIRC, DIRC)], [(0.1, -0.2), (0.3, 0.4)], (1, 1))
C_NEUM2 = (1, [(NEUM, NEUM), (NEUM, NEUM)], [(0.5, 0.0), (0.0, -0.25)], (1, 1))
N_DIRC2 = (0, [(DIRC, DIRC), (DIRC, DIRC)], [(0.0, 1.0), (0.0, 0.0)], (1, 1))
C_PER3 = (1, [(PER, PER)] * 2, None, (3, 3))

OPS = {
    "d2": lambda u: u + 0.001 * (of.d2x(of.D2SecondOrderCentered, u) + of.d2y(of.D2SecondOrderCentered, u)),
    "d1c": lambda u: of.dx(of.D1FirstOrderCentered, u) * 2.0 - of.dy(of.D1FirstOrderCentered, u),
    "up": lambda u: of.dx(of.D1FirstOrderBiasedUpwind, u) + of.dy(of.D1FirstOrderBiasedUpwind, u),
    "down": lambda u: u - 0.01 * of.dx(of.D1FirstOrderBiasedDownwind, u) / u,
    "intp": lambda u: of.d1IntpCornerToCenter(0, of.d1IntpCenterToCorner(0, u)) * -1.0,
    "neg": lambda u: -(u * u) / (u + 3.0),
    "pow2div": lambda u: u / 0.25 - u / -1024.0 + u / 3.0,
}


@pytest.fixture(params=["interp", "jit", "jit-tile"])
def engine(ctx, request):
    """Run a sweep test with the stack-program interpreter and with the kernel compiled
    from the expression at run time, in its direct form and in its LDS-staged form;
    all must equal the oracle bit for bit."""
    ctx.set_jit({"interp": 0, "jit": 4, "jit-tile": 3}[request.param])
    yield request.param
    ctx.set_jit(1)


@pytest.mark.parametrize("name", list(OPS))
@pytest.mark.parametrize("spec", [C_DIRC2, C_NEUM2, N_DIRC2], ids=["cdirc", "cneum", "ndirc"])
def test_sweep_operators_bitexact(ctx, engine, name, spec):
    got, ref = _sweep_case(ctx, OPS[name], [spec], [17, 13], None, SEED + 5, ranges=[(0.0, 1.0), (0.0, 2.5)])
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["up", "down"])
def test_sweep_weno_periodic(ctx, engine, kind):
    k = of.D1WENO53Upwind if kind == "up" else of.D1WENO53Downwind
    got, ref = _sweep_case(ctx, lambda u: u - 0.001 * (of.dx(k, u) + of.dy(k, u)), [C_PER3], [33, 21], None,
                           SEED + 6)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("op", [0, 2])
def test_sweep_3d_periodic_ghosts(ctx, engine, op):
    """Every dim periodic, the whole range assigned, 3 and 2 ghost layers: the whole
    storage (ghost faces, edges and corners) equals the oracle's sweep followed by
    updatePadding bit for bit, in place and compound into another field."""
    u_spec = (1, [(PER, PER)] * 3, None, (3, 3))
    w_spec = (1, [(PER, PER)] * 3, None, (2, 2))
    W, D2 = of.D1WENO53Upwind, of.D2SecondOrderCentered

    def fn(u, w):
        return u - 0.001 * (of.dx(W, u) + of.dz(W, u)) + 0.01 * of.d2y(D2, w) * w
    got, ref = _sweep_case(ctx, fn, [u_spec, w_spec], [13, 9, 21], None if op == 0 else w_spec, SEED + 31, op=op)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("nodes", [[25, 25], [33, 17]], ids=["h1/24", "pow2"])
def test_sweep_lid_driven_terms(ctx, engine, nodes):
    """Convective terms of examples/LidDriven/LidDriven2D.cpp:30-41 on a MAC grid; on a mesh
    of power-of-two spacings (1/32, 1/16) the compiled sweep divides by multiples of h as
    multiplications by their reciprocals, which must give the same bits."""
    u_spec = ([0, 1], [(DIRC, DIRC), (DIRC, DIRC)], [(0, 0), (0, 1.0)], (1, 1))
    v_spec = ([1, 0], [(DIRC, DIRC), (DIRC, DIRC)], [(0, 0), (0, 0)], (1, 1))
    p_spec = ([1, 1], [(NEUM, NEUM), (NEUM, NEUM)], [(0, 0), (0, 0)], (1, 1))
    I_c2n, I_n2c = of.d1IntpCenterToCorner, of.d1IntpCornerToCenter
    D1 = of.D1FirstOrderCentered

    def conv_u(u, v, p):
        conv_xx = of.dx(D1, I_n2c(0, u) * I_n2c(0, u))
        conv_xy = of.dy(D1, I_c2n(1, u) * I_c2n(0, v))
        return u - 0.001 * (conv_xx + conv_xy) + 0.0001 * (of.d2x(of.D2SecondOrderCentered, u) +
                                                            of.d2y(of.D2SecondOrderCentered, u)) \
            - 0.001 * of.dx(D1, p)

    got, ref = _sweep_case(ctx, conv_u, [u_spec, v_spec, p_spec], nodes, None, SEED + 7)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("op", [1, 2, 3, 4])
def test_sweep_compound_ops(ctx, engine, op):
    got, ref = _sweep_case(ctx, lambda u, w: w * 0.5 + u, [C_DIRC2, C_NEUM2], [13, 13], C_DIRC2, SEED + 8, op=op)
    np.testing.assert_array_equal(got, ref)


def test_sweep_3d_f64_and_range_error(ctx, engine):
    spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
    lap3 = (lambda u: u + 0.01 * (of.d2x(of.D2SecondOrderCentered, u) + of.d2y(of.D2SecondOrderCentered, u) +
                                  of.d2z(of.D2SecondOrderCentered, u)))
    got, ref = _sweep_case(ctx, lap3, [spec], [9, 11, 10], None, SEED + 9)
    np.testing.assert_array_equal(got, ref)
    # WENO needs 3 ghost layers: with ext 1 the library must refuse (OpFlow would read out of bounds)
    mesh, om = _mesh(ctx, [17, 17])
    u = build_field(ctx, mesh, 2, 1, [(DIRC, DIRC), (DIRC, DIRC)])
    with pytest.raises(of._lib.OpfError):
        u.assign(of.dx(of.D1WENO53Upwind, u))


@pytest.mark.parametrize("op", [0, 1])
def test_sweep_3d_zmarch_bitexact(ctx, engine, op):
    """A 3-D sweep on a ragged 20 x 12 x 40 box with WENO in z (ext 3, periodic),
    two fields, a z spacing different from x and y, in place and compound: equal to
    the oracle bit for bit on both engines."""
    u_spec = (1, [(DIRC, NEUM), (NEUM, DIRC), (PER, PER)], [(0.2, 0.1), (0.3, -0.1), (0, 0)], (3, 3))
    w_spec = (1, [(DIRC, DIRC)] * 3, [(0.5, -0.5), (0.0, 1.0), (0.25, 0.75)], (1, 1))
    D2, W, DN = of.D2SecondOrderCentered, of.D1WENO53Upwind, of.D1FirstOrderBiasedDownwind

    def fn(u, w):
        return (u - 0.001 * of.dz(W, u) + 0.01 * (of.d2x(D2, u) + of.d2z(D2, u))
                - 0.02 * w * of.dz(DN, w) + of.dy(DN, u) * w)
    got, ref = _sweep_case(ctx, fn, [u_spec, w_spec], [21, 13, 41], None if op == 0 else w_spec, SEED + 12, op=op,
                           ranges=[(0.0, 1.0), (0.0, 1.0), (0.0, 2.5)])
    np.testing.assert_array_equal(got, ref)


def test_sweep_pow2_spacing_bitexact(ctx, engine):
    """Uniform power-of-two spacings in every dim (1/16, 1/4, 1/128): divisions by h and by
    its multiples (D2, centred D1 at centres and corners, upwind, the linear interpolations,
    WENO's h) on both engines equal the oracle bit for bit, FP64 in place and into a field."""
    u_spec = (1, [(DIRC, NEUM), (NEUM, DIRC), (PER, PER)], [(0.2, 0.1), (0.3, -0.1), (0, 0)], (3, 3))
    w_spec = ([0, 1, 1], [(DIRC, DIRC)] * 3, [(0.5, -0.5), (0.0, 1.0), (0.25, 0.75)], (1, 1))
    D2, W, UP, D1 = of.D2SecondOrderCentered, of.D1WENO53Upwind, of.D1FirstOrderBiasedUpwind, of.D1FirstOrderCentered

    def fn(u, w):
        return (u + 0.01 * (of.d2x(D2, u) + of.d2y(D2, u) + of.d2z(D2, u)) - 0.001 * of.dz(W, u)
                - 0.02 * of.dx(D1, w) * of.dy(UP, u)
                + of.d1IntpCornerToCenter(0, w) * of.dy(D1, of.d1IntpCenterToCorner(1, u)))
    for out in (None, u_spec):
        got, ref = _sweep_case(ctx, fn, [u_spec, w_spec], [17, 9, 33], out, SEED + 14,
                               ranges=[(0.0, 1.0), (0.0, 2.0), (0.0, 0.25)])
        np.testing.assert_array_equal(got, ref)


def test_sweep_zextent_two_cells_per_thread(ctx, engine):
    """A z extent that is even but not a multiple of 4 (18 cells): the compiled kernel computes
    two cells along z per thread (plane-stride load offsets, lo + OP stores); bit-exact with the
    oracle on both engines, in place and into another field."""
    u_spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
    w_spec = (1, [(DIRC, DIRC)] * 3, [(0.5, -0.5), (0.0, 1.0), (0.25, 0.75)], (1, 1))
    D2, UP = of.D2SecondOrderCentered, of.D1FirstOrderBiasedUpwind

    def fn(u, w):
        return u + 0.01 * (of.d2x(D2, u) + of.d2z(D2, u)) - 0.02 * w * of.dz(UP, u)
    for out in (None, w_spec):
        got, ref = _sweep_case(ctx, fn, [u_spec, w_spec], [15, 11, 19], out, SEED + 13,
                               ranges=[(0.0, 1.0), (0.0, 1.0), (0.0, 1.7)])
        np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("mode", [0, 2], ids=["interp", "jit"])
def test_sweep_new_constant_reuses_program_and_kernel(ctx, mode):
    """Scalar constants are launch arguments: the same expression with a different dt (as an
    adaptive time step does) runs the program and kernel of the first sweep, with no new
    compile, kernel load or program upload, and every result equals the oracle's."""
    ctx.set_jit(mode)
    try:
        mesh, om = _mesh(ctx, [23, 17, 13])
        spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
        u = build_field(ctx, mesh, 3, spec[0], spec[1], spec[2], ext=spec[3])
        a = oracle.update_padding(u.desc, om, random_storage(u, SEED + 14, 0.5, 1.5))
        D2 = of.D2SecondOrderCentered
        before = None
        for k, dt in enumerate((1e-3, 2.5e-3, 7e-4, 1.9e-3, -3e-3)):
            u.upload(a)
            # a structure no other test sweeps (the 0.375 factor), so its first sweep builds it here
            expr = u + dt * (0.375 * of.d2x(D2, u) + of.d2y(D2, u) + of.d2z(D2, u))
            u.assign(expr)
            nodes, flds, root = expr.flatten()
            ref = oracle.assign(nodes, root, [(a, u.desc)], om, a, u.desc, 0,
                                update_padding=lambda r: oracle.update_padding(u.desc, om, r))
            np.testing.assert_array_equal(u.numpy(), ref)
            if k == 0:
                before = ctx.cache_stats()
        after = ctx.cache_stats()
        print("cache counters after the first sweep", before, "after five", after, flush=True)
        assert after["programs_uploaded"] == before["programs_uploaded"]
        assert after["jit_compiles"] == before["jit_compiles"]
        assert after["jit_loaded"] == before["jit_loaded"]
        u.close()
    finally:
        ctx.set_jit(1)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_sweep_engines_agree_at_size(ctx, dtype):
    """A 3-D sweep of 96^3 cells (above the size from which the default engine compiles
    the expression): the compiled kernel and the interpreter give the same bits."""
    n = 96
    mesh, om = _mesh(ctx, [n + 1] * 3)
    spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
    outs = []
    for mode in (1, 0):
        ctx.set_jit(mode)
        u = build_field(ctx, mesh, 3, spec[0], spec[1], spec[2], ext=spec[3], dtype=dtype)
        a = oracle.update_padding(u.desc, om, random_storage(u, SEED + 11, 0.5, 1.5)).astype(u.dtype)
        u.upload(a)
        D2, D1 = of.D2SecondOrderCentered, of.D1FirstOrderCentered
        u.assign(u + 1e-5 * (of.d2x(D2, u) + of.d2y(D2, u) + of.d2z(D2, u)) - 1e-3 * u * of.dz(of.D1FirstOrderBiasedUpwind, u))
        outs.append(u.numpy())
        u.close()
    ctx.set_jit(1)
    np.testing.assert_array_equal(outs[0], outs[1])


def test_reduce_matches_oracle(ctx):
    mesh, om = _mesh(ctx, [101, 77])
    u = build_field(ctx, mesh, 2, 1, [(NEUM, NEUM), (NEUM, NEUM)])
    a = oracle.update_padding(u.desc, om, random_storage(u, SEED + 10))
    u.upload(a)
    rng = u.assignableRange
    lo, hi = [r[0] for r in rng], [r[1] for r in rng]
    for red in ("sum", "max", "min", "sumsq", "maxabs"):
        g = of.rangeReduce(u, red)
        r = oracle.reduce(u.desc, a, red, lo, hi)
        assert g == pytest.approx(r, rel=1e-13, abs=1e-13), red
    # deterministic: same bits on repeat
    assert of.rangeReduce(u, "sum") == of.rangeReduce(u, "sum")
    # known sum: ones
    u.assign(1.0)
    assert of.rangeReduce(u, "sum") == 100 * 76


def test_context_on_caller_stream(ctx):
    """opf_ctx_set_stream: the same FTCS steps and expression sweep enqueued on a
    caller-owned HIP stream give the same bits as on the context's own stream."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")  # the runtime the library itself links
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    raw = C.c_void_p()
    assert hip.hipStreamCreate(C.byref(raw)) == 0
    mesh, om = _mesh(ctx, [65, 49, 33])
    spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
    D2 = of.D2SecondOrderCentered
    outs = []

    class _S:
        cuda_stream = raw.value

        @staticmethod
        def synchronize():
            assert hip.hipStreamSynchronize(raw) == 0
    s = _S
    for caller in (False, True):
        u = build_field(ctx, mesh, 3, spec[0], spec[1], spec[2], ext=spec[3])
        u.upload(oracle.update_padding(u.desc, om, random_storage(u, SEED + 13, 0.5, 1.5)))
        ctx.sync()
        try:
            if caller:
                ctx.set_stream(s.cuda_stream)
                assert ctx.stream == s.cuda_stream
            u.ftcs(1e-4, 7)
            u.assign(u + 1e-5 * (of.d2x(D2, u) + of.d2z(D2, u)) * u)
            if caller:
                s.synchronize()
        finally:
            ctx.set_stream(None)
        outs.append(u.numpy())
        u.close()
    assert hip.hipStreamDestroy(raw) == 0
    np.testing.assert_array_equal(outs[0], outs[1])


LIMITERS = [of.D1Central, of.D1QUICK, of.D1CUI, of.D1Fromm, of.D1LinearUpwind, of.D1Minmod, of.D1Superbee,
            of.D1MUSCL, of.D1Harmonic, of.D1Albada, of.D1UMIST, of.D1SPL13, of.D1SPLMax12, of.D1SPLMax13,
            of.D1KorenLimitedCUI, of.D1SMART, of.D1BSOU]


def _fluxlim_case(ctx, fn, specs, nodes, out_spec, seed):
    """Like _sweep_case, with the fields after the first velocities of both signs (zero at some cells)."""
    mesh, om = _mesh(ctx, nodes)
    dim = len(nodes)
    fs, arrs = [], []
    for k, (loc, bc, bcv, ext) in enumerate(specs):
        f = build_field(ctx, mesh, dim, loc, bc, bcv, ext=ext, name=f"f{k}")
        lo, hi = (0.5, 1.5) if k == 0 else (-1.0, 1.0)
        raw = random_storage(f, seed + k, lo, hi)
        if k >= 1:
            raw.reshape(-1)[::11] = 0.0  # u = 0 takes the downwind branch (u > 0. fails)
        a = oracle.update_padding(f.desc, om, raw)
        f.upload(a)
        fs.append(f)
        arrs.append(a)
    loc, bc, bcv, ext = out_spec
    out = build_field(ctx, mesh, dim, loc, bc, bcv, ext=ext, name="out")
    out_arr = oracle.update_padding(out.desc, om, random_storage(out, seed + 99, 0.5, 1.5))
    out.upload(out_arr)
    expr = fn(*fs)
    nodes_, flds, root = expr.flatten()
    slot_arrays = [(arrs[next(i for i, g in enumerate(fs) if g is f)], f.desc) for f in flds]
    ref = oracle.assign(nodes_, root, slot_arrays, om, out_arr, out.desc, 0,
                        update_padding=lambda r: oracle.update_padding(out.desc, om, r))
    out.assign(expr)
    return out.numpy(), ref


@pytest.mark.parametrize("k", range(len(LIMITERS)))
def test_sweep_fluxlim_cen2cor_bitexact(ctx, engine, k):
    """Conservative advection step e - dt * (dx(u * intp(u, e)) + dy(...)) with the limiter's
    Cen2Cor interpolation (D1FluxLimiter.hpp:43-158) in x and y, u of both signs; periodic,
    two ghost layers; equal to the oracle bit for bit on both engines."""
    K, D1 = LIMITERS[k], of.D1FirstOrderCentered
    e_spec = (1, [(PER, PER), (PER, PER)], None, (2, 2))
    ux_spec = ([0, 1], [(PER, PER), (PER, PER)], None, (2, 2))  # face velocities (MAC)
    uy_spec = ([1, 0], [(PER, PER), (PER, PER)], None, (2, 2))

    def fn(e, ux, uy):
        fx = of.d1IntpCenterToCorner(0, ux, e, kernel=K)
        fy = of.d1IntpCenterToCorner(1, uy, e, kernel=K)
        return e - 0.001 * (of.dx(D1, ux * fx) + of.dy(D1, uy * fy))
    got, ref = _fluxlim_case(ctx, fn, [e_spec, ux_spec, uy_spec], [41, 29], e_spec, SEED + 21 + k)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("k", [1, 5, 6, 9, 12, 16])
def test_sweep_fluxlim_cor2cen_3d_bitexact(ctx, engine, k):
    """Cor2Cen interpolation of a corner-located field to centres (the branch reads i-1..i+2),
    in z on a 3-D grid (the compiled sweep's z ring), Dirichlet / Neumann ghosts."""
    K = LIMITERS[k]
    e_spec = ([1, 1, 0], [(DIRC, NEUM), (NEUM, DIRC), (DIRC, DIRC)], [(0.2, 0.1), (0.3, -0.1), (0.5, -0.5)], (2, 2))
    u_spec = ([1, 1, 1], [(NEUM, NEUM)] * 3, [(0.0, 0.0)] * 3, (2, 2))
    o_spec = ([1, 1, 1], [(NEUM, NEUM)] * 3, [(0.0, 0.0)] * 3, (1, 1))

    def fn(e, u):
        return of.d1IntpCornerToCenter(2, u, e, kernel=K) * 2.0 - u
    got, ref = _fluxlim_case(ctx, fn, [e_spec, u_spec], [13, 11, 21], o_spec, SEED + 41 + k)
    np.testing.assert_array_equal(got, ref)


def test_sweep_conditional_minmax(ctx, engine):
    """conditional / max / min / comparisons (Conditional.hpp, MinMax.hpp, Compare.hpp) in a
    sweep, as the level-set example writes its upwind terms."""
    spec_e = (1, [(NEUM, NEUM), (NEUM, NEUM)], [(0.0, 0.0), (0.0, 0.0)], (3, 3))
    spec_u = (1, [(NEUM, NEUM), (NEUM, NEUM)], [(0.0, 0.0), (0.0, 0.0)], (3, 3))
    DU, DD = of.D1WENO53Upwind, of.D1WENO53Downwind

    def fn(p, u):
        return (p - 0.01 * u * of.conditional(u > 0.0, of.dx(DD, p), of.dx(DU, p))
                + of.max(-of.min(0.0, of.dy(DU, p)), of.max(of.dy(DD, p), 0.0)) * 1e-3
                + of.conditional(p >= u, p, u) * of.conditional(u <= 0.25, 1.0, -1.0) * 1e-4
                + of.conditional(u < p, 1.0, 0.0) * 1e-5)
    got, ref = _fluxlim_case(ctx, fn, [spec_e, spec_u], [37, 23], spec_e, SEED + 60)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("n,steps,bc", [
    ((40, 36, 20), 13, [(DIRC, NEUM), (NEUM, DIRC), (DIRC, DIRC)]),
    ((70, 33, 67), 14, [(NEUM, NEUM), (PER, PER), (DIRC, NEUM)]),
    ((65, 97, 9), 8, [(PER, PER), (DIRC, DIRC), (NEUM, DIRC)]),
    ((6, 5, 7), 7, [(DIRC, DIRC), (DIRC, NEUM), (PER, PER)]),
    ((32, 32, 32), 11, [(DIRC, NEUM), (NEUM, NEUM), (DIRC, DIRC)]),
    ((64, 32, 16), 9, [(NEUM, DIRC), (DIRC, DIRC), (NEUM, NEUM)]),
], ids=["dirc-neum", "per-y", "per-x", "tiny", "pow2-equal", "pow2-unequal"])
def test_ftcs3d_nonperiodic_blocked_bitexact(ctx, n, steps, bc):
    """3-D FTCS with Dirichlet / Neumann sides runs 3 steps per launch, rebuilding the ghosts of
    the intermediate levels inside the launch; equal to the oracle's step-by-step
    FieldAssigner + updatePadding sequence bit for bit (ragged tiles, non-zero boundary values,
    a z extent below the chunk, a grid smaller than the halo)."""
    mesh, om = _mesh(ctx, [k + 1 for k in n], [(0.0, 1.0), (0.0, 1.0), (0.0, 1.0)])
    bcv = [(0.4 if t0 == DIRC else -0.3 if t0 == NEUM else 0.0, -0.2 if t1 == DIRC else 0.6 if t1 == NEUM else 0.0)
           for (t0, t1) in bc]
    u = build_field(ctx, mesh, 3, 1, bc, bcv)
    init = oracle.update_padding(u.desc, om, random_storage(u, SEED + 70 + steps))
    u.upload(init)
    launches, kmax = u.ftcs_plan(steps)
    assert kmax == 3
    coef = 0.1 / max(n) ** 2
    u.ftcs(coef, steps)
    ref = oracle.ftcs(u.desc, om, init, coef, steps)
    np.testing.assert_array_equal(u.numpy(), ref)
    # and equal to single steps on the device
    u.upload(init)
    u.ftcs_blocked(coef, steps, 1)
    np.testing.assert_array_equal(u.numpy(), ref)
    u.close()


# <cmath> functions of expressions (Arithmetic/AMDS.hpp:39-90): the argument mapped into each
# function's domain from the test fields' values in [0.5, 1.5]
FN1_ARGS = {"acos": lambda u: u - 1.0, "asin": lambda u: u - 1.0, "atanh": lambda u: u - 1.0,
            "acosh": lambda u: u + 1.0, "log1p": lambda u: u - 1.0, "expm1": lambda u: u - 1.0,
            "round": lambda u: u * 4.0 - 4.0, "nearbyint": lambda u: u * 4.0 - 4.0, "rint": lambda u: u * 4.0 - 4.0,
            "ceil": lambda u: u * 4.0 - 4.0, "floor": lambda u: u * 4.0 - 4.0, "trunc": lambda u: u * 4.0 - 4.0,
            "logb": lambda u: u * 1e3 - 900.0, "tan": lambda u: u - 1.0}
FN2_ARGS = {"fmod": lambda u, w: (u * 7.0, w), "remainder": lambda u, w: (u * 7.0, w),
            "copysign": lambda u, w: (u, w - 1.0), "atan2": lambda u, w: (u - 1.0, w - 1.0)}
SPEC_C2 = (1, [(DIRC, NEUM), (PER, PER)], [(0.2, 0.1), (0, 0)], (2, 2))


def _ulp_close(got, ref, name, exact):
    if exact:
        np.testing.assert_array_equal(got, ref, err_msg=name)
    else:  # the device's function and the C library's round differently in the last bits
        np.testing.assert_allclose(got, ref, rtol=1e-14, atol=1e-15, err_msg=name)


@pytest.mark.parametrize("name", oracle.expr.FN1_NAMES[:32])
def test_sweep_math_fn1(ctx, engine, name):
    """Each unary <cmath> function of an expression on all engines against the oracle (the C
    library's function through Python's math module): bit for bit where the C standard fixes
    the result (sqrt, rounding functions, logb), within a few ulp otherwise."""
    f = getattr(of, name)
    arg = FN1_ARGS.get(name, lambda u: u)
    got, ref = _sweep_case(ctx, lambda u, w: f(arg(u)) * 0.5 + w, [SPEC_C2, SPEC_C2], [21, 14], SPEC_C2,
                           SEED + 60 + oracle.expr.FN1_NAMES.index(name))
    _ulp_close(got, ref, name, name in oracle.expr.FN1_EXACT)


@pytest.mark.parametrize("name", oracle.expr.FN2_NAMES)
def test_sweep_math_fn2(ctx, engine, name):
    f = getattr(of, name)
    args = FN2_ARGS.get(name, lambda u, w: (u, w))
    got, ref = _sweep_case(ctx, lambda u, w: f(*args(u, w)) - u, [SPEC_C2, SPEC_C2], [17, 12], SPEC_C2,
                           SEED + 100 + oracle.expr.FN2_NAMES.index(name))
    _ulp_close(got, ref, name, name in oracle.expr.FN2_EXACT)


def test_sweep_smooth_delta_heaviside(ctx, engine):
    """Math::smoothDelta / smoothHeviside (Math/Function/Numeric.hpp:86-96) of an expression,
    the functors of the level-set examples' UniOpAdaptor, eps = 0.3 against values in
    [-0.5, 0.5] (both branches and the cut-offs)."""
    got, ref = _sweep_case(ctx, lambda u, w: of.smoothDelta(0.3, u - 1.0) + of.smoothHeviside(0.3, w - 1.0) * 2.0,
                           [SPEC_C2, SPEC_C2], [23, 15], SPEC_C2, SEED + 120)
    _ulp_close(got, ref, "smooth", False)


@pytest.mark.parametrize("dim", [2, 3])
def test_sweep_conv_bitexact(ctx, engine, dim):
    """conv(expr, kernel) (Convolution.hpp:52-68): the level-set example's 3 x 3 kernel
    (16/24 centre, 1/24 around) in 2-D and an asymmetric 3 x 3 x 3 kernel in 3-D, over an
    expression of two fields, equal to the oracle bit for bit (x-fastest accumulation)."""
    if dim == 2:
        c_, o_ = 16. / 24., 1. / 24.
        ker = np.array([[o_, o_, o_], [o_, c_, o_], [o_, o_, o_]])
        spec, nodes = SPEC_C2, [19, 13]
    else:
        ker = np.arange(27, dtype=np.float64).reshape(3, 3, 3) / 17.0 - 0.5
        spec = (1, [(DIRC, NEUM), (PER, PER), (NEUM, DIRC)], [(0.2, 0.1), (0, 0), (-0.3, 0.7)], (1, 1))
        nodes = [11, 9, 13]
    got, ref = _sweep_case(ctx, lambda u, w: 0.25 * of.conv(u * w - 1.0, ker) + u, [spec, spec], nodes,
                           spec, SEED + 130 + dim)
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize("kind", ["ftcs", "weno"])
def test_sweep_f32_error_against_fp64_oracle(ctx, engine, kind):
    """FP32 sweeps (both engines) against the FP64 oracle on the same input (the FP32
    field holds the FP64 input rounded to float): the FP32 path computes in float from
    float spacings, so the bound is a small multiple of float's epsilon (6e-8) relative
    to the largest value: 1e-6."""
    n = [37, 29, 24]
    mesh, om = _mesh(ctx, [k + 1 for k in n], ranges=[(0.0, 1.0), (0.0, 0.75), (0.0, 0.5)])
    W, D2 = of.D1WENO53Upwind, of.D2SecondOrderCentered
    if kind == "ftcs":
        fn = lambda u: u + 1e-5 * (of.d2x(D2, u) + of.d2y(D2, u) + of.d2z(D2, u))
    else:
        fn = lambda u: u - 1e-3 * (of.dx(W, u) + of.dy(W, u) + of.dz(W, u))
    u64 = build_field(ctx, mesh, 3, 1, [(PER, PER)] * 3, ext=(3, 3), dtype="f64", name="u")
    a = oracle.update_padding(u64.desc, om, random_storage(u64, SEED + 51, 0.5, 1.5))
    a32 = a.astype(np.float32).astype(np.float64)  # the FP32 field's input, exactly
    expr = fn(u64)
    nodes_, flds, root = expr.flatten()
    ref = oracle.assign(nodes_, root, [(a32, u64.desc)], om, a32.copy(), u64.desc, 0,
                        update_padding=lambda r: oracle.update_padding(u64.desc, om, r))
    u32 = build_field(ctx, mesh, 3, 1, [(PER, PER)] * 3, ext=(3, 3), dtype="f32", name="u")
    u32.upload(a32.astype(np.float32))
    u32.assign(fn(u32))
    got = u32.numpy().astype(np.float64)
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    print(kind, "max rel err", err)
    assert err <= 1e-6, err
    u64.close()
    u32.close()
