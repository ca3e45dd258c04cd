"""Learned warm start: the network's CPU oracle pinned by mathematics, the
device's BF16-MFMA forward / backward / Adam kernels against it, the
feature rows against their definition, and what the trained network gains
over the warm starts without it.

Tolerances.  The device accumulates matrix products in FP32 in its own
order; the oracle emulates the BF16 rounding points exactly and sums in
float64.  Outputs therefore agree to FP32 rounding except where an FP32
difference moves a hidden activation across a BF16 rounding boundary
(one BF16 step, 2^-8 relative, in a small fraction of activations): the
tests bound the median error tightly and the worst case loosely."""
import numpy as np
import pytest

from kizuriphysics3d_amd import device as D
from kizuriphysics3d_amd import scenes as S
from oracle import mlp as M


# ---------------------------------------------------------------------------
# CPU: the oracle against mathematics
# ---------------------------------------------------------------------------

def _random_problem(n, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, scale, M.N_PARAMS).astype(np.float32)
    mean = rng.normal(0.0, 1.0, M.N_IN).astype(np.float32)
    istd = rng.uniform(0.5, 2.0, M.N_IN).astype(np.float32)
    s = np.zeros((n, 20), np.float32)
    s[:, :M.N_IN] = rng.normal(0.0, 1.0, (n, M.N_IN))
    s[:, 16:19] = rng.normal(0.0, 1.0, (n, 3))
    return p, mean, istd, s


def test_bf16_rounding_is_nearest_even():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)  # BF16 spacing at 1
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 3.0e38],
                 np.float32)
    got = M.bf16(x)
    assert got[0] == one
    assert got[1] == one                      # tie -> even (mantissa 0)
    assert got[2] == one + 2 * ulp            # tie -> even (mantissa 2)
    assert got[3] == one + ulp                # above the tie -> up
    assert got[4] == np.float32(-2.5)         # exactly representable
    assert np.isfinite(got[5])


def test_oracle_gradient_matches_finite_differences():
    p, mean, istd, s = _random_problem(6, 1)
    p = p.astype(np.float64)
    _, g = M.loss_and_grad(p, mean, istd, s, emulate=False)
    h = 1e-6
    fd = np.zeros_like(g)
    for k in range(M.N_PARAMS):
        q = p.copy()
        q[k] += h
        a = M.mean_loss(q, mean, istd, s)
        q[k] -= 2 * h
        b = M.mean_loss(q, mean, istd, s)
        fd[k] = (a - b) / (2 * h)
    err = np.abs(fd - g)
    assert err.max() <= 1e-6 * max(1.0, np.abs(g).max()), (err.argmax(), err.max())
    # every parameter block receives gradient (no dropped term)
    for lo, hi in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                   (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        assert np.abs(g[lo:hi]).max() > 0


def test_emulated_network_is_close_to_exact():
    p, mean, istd, s = _random_problem(64, 2)
    ye, _ = M.forward(p, mean, istd, s[:, :16], emulate=True)
    yx, _ = M.forward(p.astype(np.float64), mean, istd, s[:, :16], emulate=False)
    assert np.abs(ye - yx).max() <= 0.03 * np.abs(yx).max()
    _, ge = M.loss_and_grad(p, mean, istd, s, emulate=True)
    _, gx = M.loss_and_grad(p.astype(np.float64), mean, istd, s, emulate=False)
    assert np.linalg.norm(ge - gx) <= 0.03 * np.linalg.norm(gx)


def test_adam_first_step_closed_form():
    # t = 1 from zero moments: m^ = g and v^ = g^2, so w1 = w0 - lr * g / (|g| + eps).
    rng = np.random.default_rng(3)
    w0 = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    w1, m, v = M.adam(w0, np.zeros(100), np.zeros(100), g, 1, 1e-3)
    expect = w0.astype(np.float64) - 1e-3 * g / (np.abs(g) + 1e-8)
    assert np.allclose(w1, expect, rtol=0, atol=1e-6)
    # (1 - beta2 in FP32 is 0.00099998713: relative 1.3e-5 off 0.001)
    assert np.allclose(m, 0.1 * g, rtol=1e-6) and np.allclose(v, 0.001 * g * g, rtol=1e-4)
    # second step with the same gradient moves by the same amount again
    w2, _, _ = M.adam(w1, m, v, g, 2, 1e-3)
    assert np.allclose(w2 - w1, w1 - w0, rtol=1e-3, atol=1e-7)


def test_features_definition_on_a_resting_box():
    # Unit-mass box resting on a static floor, contact point straight below
    # the centre: rA parallel to n, so the normal projected inertia is 0 and
    # the normal relative velocity is the body's vertical velocity.
    ra = np.array([[0.0, -0.5, 0.0]])
    rb = np.array([[0.0, 0.5, 0.0]])
    n = np.array([[0.0, 1.0, 0.0]])
    t1 = np.array([[1.0, 0.0, 0.0]])
    t2 = np.array([[0.0, 0.0, -1.0]])
    I = np.eye(3)[None] * 6.0  # inverse inertia of a unit cube of 1 kg
    Z = np.zeros((1, 3, 3))
    f = M.contact_features(ra, rb, n, t1, t2, np.array([-0.01]), np.array([1.0]), np.array([0.0]), I, Z,
                           np.array([[0.0, -0.16, 0.0]]), np.zeros((1, 3)), np.zeros((1, 3)), np.zeros((1, 3)),
                           np.array([1.0]))
    assert f[0, 0] == pytest.approx(-0.16)
    assert f[0, 1] == 0 and f[0, 2] == 0
    assert f[0, 6] == pytest.approx(0.01)
    assert f[0, 9] == 0.0                       # r x n = 0
    assert f[0, 11] == pytest.approx(6 * 0.25)  # |r x t|^2 = 0.25 for both tangents
    assert f[0, 14] == 0 and f[0, 15] == 0      # no load features


def test_body_loads_of_a_resting_stack():
    # Two boxes stacked on a static floor (body 0), at rest: the floor holds
    # both weights over 4 points, the boxes press on each other with the top
    # box's weight over 4 points.  Each box's load is the sum of its points'
    # normal impulses over its own weight impulse: the top box 1, the bottom
    # box (m_bot + m_top) / m_bot from the floor + m_top / m_bot from the top
    # box, the floor 0.
    g, dt = 9.81, 1 / 60
    m_bot, m_top = 2.0, 0.5
    inv_mass = np.array([0.0, 1 / m_bot, 1 / m_top])
    dynamic = np.array([False, True, True])
    a = np.array([1] * 4 + [2] * 4)   # bottom on floor, top on bottom
    b = np.array([0] * 4 + [1] * 4)
    lam = np.array([(m_bot + m_top) * g * dt / 4] * 4 + [m_top * g * dt / 4] * 4)
    la, lb = M.body_loads(a, b, lam, inv_mass, dynamic, g * dt)
    assert la[:4] == pytest.approx([(m_bot + m_top) / m_bot + m_top / m_bot] * 4)
    assert la[4:] == pytest.approx([1.0] * 4)
    assert lb[:4] == pytest.approx([0.0] * 4)
    assert lb[4:] == pytest.approx([(m_bot + m_top) / m_bot + m_top / m_bot] * 4)
    # negative cached impulses count as 0 (Prepare clamps them)
    la2, _ = M.body_loads(a, b, np.where(np.arange(8) == 0, -5.0, lam), inv_mass, dynamic, g * dt)
    assert la2[0] == pytest.approx(la[0] - lam[0] / m_bot / (g * dt))


# ---------------------------------------------------------------------------
# GPU: device kernels against the oracle
# ---------------------------------------------------------------------------

def _ctx():
    return D.DeviceContext(S.world_settings())


@pytest.mark.gpu
def test_mlp_forward_matches_oracle():
    p, mean, istd, s = _random_problem(1000, 11)
    ctx = _ctx()
    ctx.mlp_set_params(p)
    ctx.mlp_set_normalization(mean, istd)
    assert np.array_equal(ctx.mlp_params(), p)
    y = ctx.mlp_forward(s[:, :16])
    assert np.all(y[:, 3] == 0)
    yo, _ = M.forward(p, mean, istd, s[:, :16], emulate=True)
    err = np.abs(y[:, :3] - yo)
    scale = np.abs(yo).max()
    print(f"forward: max |y| {scale:.3f}, median err {np.median(err):.3e}, max err {err.max():.3e}", flush=True)
    assert np.median(err) <= 1e-5 * scale
    assert np.quantile(err, 0.99) <= 1e-3 * scale
    assert err.max() <= 2e-2 * scale


@pytest.mark.gpu
def test_mlp_gradients_match_oracle():
    p, mean, istd, s = _random_problem(1000, 12, scale=0.2)
    ctx = _ctx()
    ctx.mlp_set_params(p)
    ctx.mlp_set_normalization(mean, istd)
    g, loss = ctx.mlp_gradients(s)
    lo, go = M.loss_and_grad(p, mean, istd, s, emulate=True)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    print(f"gradients: |g| {np.linalg.norm(go):.4e}, relative L2 error {rel:.3e}, loss {loss:.6e} vs {lo:.6e}",
          flush=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert rel <= 1e-3
    for lo_, hi_ in ((M.OFF_W1, M.OFF_B1), (M.OFF_B1, M.OFF_W2), (M.OFF_W2, M.OFF_B2), (M.OFF_B2, M.OFF_W3),
                     (M.OFF_W3, M.OFF_B3), (M.OFF_B3, M.N_PARAMS)):
        blk = np.linalg.norm(g[lo_:hi_] - go[lo_:hi_]) / max(np.linalg.norm(go[lo_:hi_]), 1e-30)
        assert blk <= 5e-3, (lo_, blk)
    # planted mistake in the oracle: a wrong target is seen at once
    s2 = s.copy()
    s2[:, 17] += 1.0
    _, gw = M.loss_and_grad(p, mean, istd, s2, emulate=True)
    assert np.linalg.norm(g - gw) / np.linalg.norm(gw) > 0.05


@pytest.mark.gpu
def test_mlp_gradients_with_target_clip():
    """kzg_mlp_set_target_clip: the device gradient equals the oracle's on
    targets clamped to [-clip, clip], and differs from the unclamped one."""
    p, mean, istd, s = _random_problem(1000, 13, scale=0.2)
    s[:, 16:19] *= 5.0
    clip = 2.0
    assert (np.abs(s[:, 16:19]) > clip).mean() > 0.3
    ctx = _ctx()
    ctx.mlp_set_params(p)
    ctx.mlp_set_normalization(mean, istd)
    ctx.mlp_set_target_clip(clip)
    g, loss = ctx.mlp_gradients(s)
    sc = s.copy()
    sc[:, 16:19] = np.clip(sc[:, 16:19], -clip, clip)
    lo, go = M.loss_and_grad(p, mean, istd, sc, emulate=True)
    assert loss == pytest.approx(lo, rel=1e-4)
    assert np.linalg.norm(g - go) / np.linalg.norm(go) <= 1e-3
    _, gu = M.loss_and_grad(p, mean, istd, s, emulate=True)
    assert np.linalg.norm(g - gu) / np.linalg.norm(gu) > 0.05
    ctx.mlp_set_target_clip(0.0)
    g0, _ = ctx.mlp_gradients(s)
    assert np.linalg.norm(g0 - gu) / np.linalg.norm(gu) <= 1e-3


def _small_pile(n=1024, seed=0x4B5A0510, y_range=(1.0, 16.0)):
    from kizuriphysics3d_amd.world import PhysicsWorld
    return PhysicsWorld(S.pile(n, seed=seed, width=12.0, y_range=y_range))


@pytest.mark.gpu
@pytest.mark.parametrize("load", [False, True])
def test_feature_rows_and_targets_match_definition(load):
    w = _small_pile()
    ctx = w.device_context()
    for _ in range(150):  # the bodies have landed: most contacts persist from the last step
        w.step_fixed()
    mass = w.mass()
    w.update_broadphase()
    pairs = w.pairs()
    ctx.mlp_collect(1 << 16)
    ctx.set_learned_features(D.FEATURES_LOAD if load else 0)
    ctx.narrowphase(pairs)
    ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE)
    cached = ctx.constraints()[1]  # the cache's impulses, before Prepare clamps them
    ctx.integrate_velocities()
    st = ctx.states()  # velocities after integration, the ones Prepare reads
    ctx.solve()
    rows = ctx.mlp_samples()
    hdr, pts = ctx.constraints()
    npts = hdr[:, 2].astype(int)
    live = np.arange(8)[None, :] < npts[:, None]
    P = pts[live].astype(np.float64)
    a = np.repeat(hdr[:, 0], npts)
    b = np.repeat(hdr[:, 1], npts)
    assert len(rows) == len(P) > 1000
    def quat_mat(q):
        x, y, z, w_ = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
        return np.stack([
            np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w_), 2 * (x * z + y * w_)], -1),
            np.stack([2 * (x * y + z * w_), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w_)], -1),
            np.stack([2 * (x * z - y * w_), 2 * (y * z + x * w_), 1 - 2 * (x * x + y * y)], -1)], 1)
    R = quat_mat(st["rotation"].astype(np.float64))
    Il = mass[:, 1:10].astype(np.float64).reshape(-1, 3, 3).transpose(0, 2, 1)  # column-major -> matrix
    Iw = R @ Il @ R.transpose(0, 2, 1)
    dyn = mass[:, 0] > 0
    Iw[~dyn] = 0
    f = M.contact_features(P[:, 6:9], P[:, 9:12], P[:, 18:21], P[:, 21:24], P[:, 24:27], P[:, 33],
                           mass[a, 0], mass[b, 0], Iw[a], Iw[b],
                           st["linear_velocity"][a], st["angular_velocity"][a],
                           st["linear_velocity"][b], st["angular_velocity"][b],
                           rows[:, 13])
    gdt = np.float32(9.81) * np.float32(1 / 60)
    if load:
        la, lb = M.body_loads(a, b, cached[live][:, 30].astype(np.float64), mass[:, 0].astype(np.float64), dyn,
                              float(gdt))
        f[:, 14], f[:, 15] = la, lb
        assert la.max() > 2.0  # bodies low in the pile carry several weights
    for k in range(16):
        tol = 1e-4 * max(1.0, np.abs(f[:, k]).max())
        err = np.abs(rows[:, k] - f[:, k]).max()
        assert err <= tol, (k, err, tol)
    assert set(np.unique(rows[:, 13])) <= {0.0, 1.0}
    assert rows[:, 13].mean() > 0.5  # a settling pile: most points come from the cache
    assert np.all(rows[:, 19] == 0)
    # targets: (final impulses - the cache's) in units of the scale s =
    # normal mass * |g| dt; where the cache had no match its impulses are zero
    s = (P[:, 27].astype(np.float32) * gdt).astype(np.float32)
    miss = rows[:, 13] == 0
    for k, col in ((16, 30), (17, 31), (18, 32)):
        expect = np.where(s > 0, (P[:, col].astype(np.float32) / np.where(s > 0, s, 1)).astype(np.float32), 0)
        assert np.array_equal(rows[miss, k], expect[miss]), k
    hit = ~miss & (s > 0)
    assert np.all(np.isfinite(rows[:, 16:19])) and hit.sum() > 0
    print(f"median |target| on hits {np.median(np.abs(rows[hit, 16])):.3f}, "
          f"median |impulse|/s {np.median(np.abs(P[hit, 30] / s[hit])):.3f}", flush=True)


@pytest.mark.gpu
def test_mlp_train_step_is_adam_on_the_device_gradient():
    w = _small_pile(512, 0x4B5A0511)
    ctx = w.device_context()
    ctx.mlp_init(0x4B5A0005)
    ctx.mlp_collect(1 << 16)
    for _ in range(30):
        w.step_fixed()
    ctx.mlp_fit_normalization()
    written, held = ctx.mlp_num_samples()
    assert held > 2000
    rows = ctx.mlp_samples(0, held)
    mean, istd = ctx.mlp_normalization()
    fm = rows[:, :16].astype(np.float64)
    assert np.allclose(mean, fm.mean(0), rtol=1e-4, atol=1e-5)
    p0 = ctx.mlp_params()
    batch, seed, lr = 1024, 77, 1e-3
    idx = (M.sample_hash(seed, np.arange(batch)) % np.uint64(held)).astype(np.int64)
    g, _ = ctx.mlp_gradients(rows[idx])
    losses = ctx.mlp_train(1, batch=batch, lr=lr, seed=seed)
    p1 = ctx.mlp_params()
    w1, _, _ = M.adam(p0, np.zeros_like(p0), np.zeros_like(p0), g, 1, lr)
    ulps = np.abs(p1.view(np.int32).astype(np.int64) - w1.view(np.int32).astype(np.int64))
    assert ulps.max() <= 1, ulps.max()
    lo, go = M.loss_and_grad(p0, mean, istd, rows[idx], emulate=True)
    assert losses[0] == pytest.approx(lo / (3 * batch), rel=1e-3)
    assert np.linalg.norm(g - go) <= 1e-3 * np.linalg.norm(go)


@pytest.mark.gpu
def test_learned_warm_start_gains_over_cold_start():
    """Train on a settling pile, then solve the same state with every warm
    start and compare the PGS residual after k sweeps.  The figures for all
    modes are printed; the assertions are what the trained network gains
    over the methods without it (cold start, cache alone)."""
    from kizuriphysics3d_amd import learned as Lw
    w = _small_pile(1024, 0x4B5A0512, y_range=(1.0, 40.0))
    losses = Lw.train_on_world(w, collect_steps=300, train_steps=1500, batch=4096, lr=1e-3)
    print(f"training loss: first {losses[:10].mean():.4f}, last {losses[-100:].mean():.4f}", flush=True)
    assert losses[-100:].mean() < 0.9 * losses[:10].mean()
    res = Lw.compare_warm_starts(w, penetration_iterations=(4, 8))
    for name, r in res.items():
        print(f"{name:13s} residual k=1..8 {np.array2string(r['residual'], precision=4)} "
              f"penetration {r['penetration']}", flush=True)
    cold, cache = res["cold"]["residual"], res["cache"]["residual"]
    learned, miss = res["learned"]["residual"], res["learned_miss"]["residual"]
    assert len(cold) == len(cache) == len(learned) == 8
    # After 8 sweeps both warm starts beat the cold start, and the network's
    # correction on top of the cache is at least as good as the cache alone.
    assert cache[7] < cold[7]
    assert learned[7] < cold[7] and miss[7] < cold[7]
    assert learned[7] <= cache[7] and miss[7] <= cache[7]


@pytest.mark.gpu
def test_no_collect_flag_appends_no_rows():
    """KZG_WARM_NO_COLLECT: a solve built with it appends no training rows;
    the next one without it does."""
    w = _small_pile(256)
    ctx = w.device_context()
    for _ in range(60):
        w.step_fixed()
    ctx.mlp_collect(1 << 16)
    w.update_broadphase()
    pairs = w.pairs()
    ctx.narrowphase(pairs)
    ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE | D.WARM_NO_COLLECT)
    ctx.integrate_velocities()
    ctx.solve()
    assert ctx.mlp_num_samples() == (0, 0)
    ctx.narrowphase(pairs)
    ctx.build_constraints(D.WARM_CACHE | D.WARM_NO_STORE)
    ctx.integrate_velocities()
    ctx.solve()
    written, held = ctx.mlp_num_samples()
    assert written == held == ctx.sizes()[2] > 0
