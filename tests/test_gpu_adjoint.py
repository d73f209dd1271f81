"""Derivatives of recorded substeps: replay, forward mode, adjoint (GPU).

* The replay of a recording reproduces the stepped state bit for bit (the replay re-evaluates
  every continuous value of the substep with the recorded discrete structure).
* The adjoint is the transpose of the forward-mode (dual-number) run of the same substeps:
  <gE, J d0> = <J^T gE, d0> to 1e-9 relative to the sum of the terms' magnitudes.
* The adjoint matches central finite differences (step 1e-6) of the CPU oracle stepped from the
  same state: 1e-4 relative, for single-body initial-velocity directions, friction and
  restitution, wherever the loss is smooth over the step.  Where it is not (a contact that
  switches, an impulse sitting on its bound), the adjoint is the derivative along the branch
  the forward pass took and must equal one of the two one-sided slopes; a direction matching
  neither has a jump inside the step and is counted, with a small cap, never for the material
  parameters.  A reduced form of the 50,000-sphere differentiation case checks every direction.
"""
import numpy as np
import pytest

from oracle_binding import oracle_sim
from scenes import apply_scene, make_scene


def hip_sim(capacity):
    from peclet.dem import Simulation

    return Simulation(capacity)


def state(sim):
    return np.concatenate([sim.get_positions(), sim.get_velocities(),
                           sim.get_angular_velocities(), sim.get_quaternions()], axis=1)


def recorded(kind, n, pre, steps, every, seed=0):
    scene = make_scene(kind, n, seed)
    h = apply_scene(hip_sim(n), scene)
    for _ in range(pre):
        h.step(scene["dt"])
    h.tape_begin(every)
    for _ in range(steps):
        h.step(scene["dt"])
    h.tape_end()
    assert h.tape_length() == steps
    return scene, h


@pytest.mark.gpu
@pytest.mark.parametrize("kind,every", [("bed", 4), ("bed_periodic", 3), ("bed_periodic", 16)])
def test_tape_replay_is_the_stepped_state(hip_lib, kind, every):
    print(kind, every, flush=True)
    scene, h = recorded(kind, 300, 4, 10, every)
    assert h.get_stats()["num_contacts"] > 0
    rep = h.tape_replay()
    assert np.array_equal(rep, state(h))


@pytest.mark.gpu
@pytest.mark.parametrize("kind,every", [("bed", 4), ("bed_periodic", 3)])
def test_segment_replayed_once_or_twice_gives_the_same_adjoint(hip_lib, kind, every):
    # each substep of a checkpoint segment replayed once into its own workspace (default), or
    # the segment replayed for its states and each substep again before its reverse.  The
    # replays are bit-equal either way; the reverse sums scattered adjoints with atomic adds, so
    # two adjoints agree to rounding (the same bound as two runs of one form)
    print(kind, every, flush=True)
    scene, h = recorded(kind, 300, 4, 10, every)
    gE = np.random.default_rng(3).normal(size=(h.num_particles(), 13))
    g_once, p_once = h.tape_adjoint(gE)
    g_again, _ = h.tape_adjoint(gE)
    h.set_tuning("adjoint_keep_segment", 0)
    g_twice, p_twice = h.tape_adjoint(gE)
    scale = np.abs(g_twice).max()
    spread = np.abs(g_once - g_again).max()
    assert np.abs(g_once - g_twice).max() <= max(1e-12 * scale, 4 * spread)
    np.testing.assert_allclose(p_once, p_twice, rtol=1e-10)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["bed", "bed_periodic"])
def test_adjoint_is_the_transpose_of_the_tangent(hip_lib, kind):
    print(kind, flush=True)
    scene, h = recorded(kind, 300, 4, 10, 4)
    n = h.num_particles()
    rng = np.random.default_rng(1)
    d0 = np.zeros((n, 13))
    d0[:, 0:3] = rng.normal(size=(n, 3)) * 1e-5
    d0[:, 3:6] = rng.normal(size=(n, 3)) * 1e-2
    d0[:, 6:9] = rng.normal(size=(n, 3))
    dmu, de = 0.3, -0.2
    val, tan = h.tape_tangent(d0, dmu, de)
    assert np.array_equal(val, state(h))
    gE = rng.normal(size=(n, 13))
    g0, gp = h.tape_adjoint(gE)
    lhs = np.sum(gE * tan)
    terms = np.concatenate([(g0 * d0).ravel(), [gp[0] * dmu, gp[1] * de]])
    rhs = terms.sum()
    scale = np.abs(gE * tan).sum() + np.abs(terms).sum()
    assert abs(lhs - rhs) <= 1e-9 * scale, (lhs, rhs, scale)
    assert abs(gp[0]) > 0.0 and abs(gp[1]) > 0.0


def fd_directions(n, rng):
    """Single-particle velocity directions (both components of a few bodies) and two random
    directions over every body."""
    dirs = []
    for i in (0, n // 4, n // 2, 3 * n // 4, n - 1):
        for c in (0, 2):
            d = np.zeros((n, 3))
            d[i, c] = 1.0
            dirs.append((f"v[{i},{c}]", d))
    for k in range(2):
        d = rng.normal(size=(n, 3))
        dirs.append((f"random{k}", d / np.linalg.norm(d)))
    return dirs


@pytest.mark.gpu
@pytest.mark.parametrize("kind,friction", [("bed", 0.5), ("bed_periodic", 0.5), ("bed", 0.0)])
def test_adjoint_matches_central_differences_of_the_oracle(hip_lib, kind, friction):
    # Contact dynamics is piecewise smooth: jumps (the restitution threshold, contact switching)
    # and kinks where a constraint sits on its bound at the base point (a resting contact left at
    # C = 0 by the position projection, an impulse at 0) make a central difference the mean of
    # two one-sided slopes there.  A difference is trusted only where the loss is smooth over
    # the step: both one-sided slopes at 1e-7 and the central differences at 1e-6 and 1e-7 agree.
    n, pre, steps = 200, 4, 8
    print(kind, friction, n, pre, steps, flush=True)
    scene = make_scene(kind, n, 0)
    e0, et0, _ = scene["material"]
    scene["material"] = (e0, et0, friction)
    h = apply_scene(hip_sim(n), scene)
    for _ in range(pre):
        h.step(scene["dt"])
    h.tape_begin(4)
    for _ in range(steps):
        h.step(scene["dt"])
    h.tape_end()
    rng = np.random.default_rng(2)
    gE = np.zeros((n, 13))
    gE[:, 0:3] = rng.normal(size=(n, 3))
    gE[:, 3:6] = rng.normal(size=(n, 3)) * 1e-3
    g0, gp = h.tape_adjoint(gE)

    def loss(dv=None, dmu=0.0, de=0.0):
        r = apply_scene(oracle_sim(n), scene)
        for _ in range(pre):
            r.step(scene["dt"])
        if dv is not None:
            r.set_velocities(r.get_velocities() + dv)
        r.set_material_params(e0 + de, et0, friction + dmu)
        for _ in range(steps):
            r.step(scene["dt"])
        return float(np.sum(gE * state(r)))

    f0 = loss()

    # all-body random directions are reported but not asserted: over 12 substeps of a 200-body
    # bed some contact switches inside any step that moves every body
    checks = [(name, (lambda s, d=d: loss(dv=s * d)), float(np.sum(g0[:, 3:6] * d)))
              for name, d in fd_directions(n, rng)]
    report_only = {name for name, _, _ in checks if name.startswith("random")}
    checks.append(("restitution", lambda s: loss(de=s), gp[1]))
    if friction > 0.0:
        checks.append(("friction", lambda s: loss(dmu=s), gp[0]))
    jumps = []
    for name, f, adj in checks:
        p6, m6, p7, m7 = f(1e-6), f(-1e-6), f(1e-7), f(-1e-7)
        fd6, fd7 = (p6 - m6) / 2e-6, (p7 - m7) / 2e-7
        right, left = (p7 - f0) / 1e-7, (f0 - m7) / 1e-7
        size = max(abs(fd6), abs(fd7), 1e-12)
        ok_fd = abs(fd6 - fd7) <= 1e-3 * size and abs(right - left) <= 2e-3 * size
        print(f"  {name:12s} fd {fd6:.9e} {fd7:.9e} one-sided {right:.6e} {left:.6e} "
              f"adjoint {adj:.9e} smooth {ok_fd}", flush=True)
        if name in report_only:
            continue
        if ok_fd:
            assert abs(fd6 - adj) <= 1e-4 * abs(adj) + 1e-10, (name, fd6, adj)
            continue
        # a kink at the base point: the adjoint is the derivative along the branch the forward
        # pass took, one of the two one-sided slopes; neither matching means a jump inside the
        # step
        if min(abs(adj - right), abs(adj - left)) > 1e-3 * max(abs(right), abs(left)):
            jumps.append(name)
    print(f"  directions with a jump inside the step: {jumps}", flush=True)
    assert len(jumps) <= 2, jumps
    assert "restitution" not in jumps and "friction" not in jumps, jumps


@pytest.mark.gpu
def test_torch_rollout_gradients(hip_lib):
    import torch

    from peclet.dem.adjoint import bed_height, packing_fraction_below, rollout

    n = 200
    scene = make_scene("bed_periodic", n, 0)
    h = apply_scene(hip_sim(n), scene)
    for _ in range(4):
        h.step(scene["dt"])
    v0 = torch.tensor(h.get_velocities(), dtype=torch.float64, requires_grad=True)
    mu = torch.tensor(scene["material"][2], dtype=torch.float64, requires_grad=True)
    e = torch.tensor(scene["material"][0], dtype=torch.float64, requires_grad=True)
    end = rollout(h, v0, mu, e, scene["dt"], 8, checkpoint_every=4,
                  restitution_tangent=scene["material"][1])
    rad = h.get_scales() * 1e-3
    L = scene["domain"][0]
    loss = bed_height(end[:, 0:3]) + packing_fraction_below(end[:, 0:3], rad, L * L)
    loss.backward()
    # the same gradient through the C-ABI directly
    gE = np.zeros((n, 13))
    gE[:, 2] = 1.0 / n
    H = 2.0 * end[:, 2].detach().numpy().mean()
    inside = end[:, 2].detach().numpy() < 0.6 * H
    vol = (4.0 / 3.0) * np.pi * rad ** 3
    phi = (inside * vol).sum() / (L * L * 0.6 * H)
    gE[:, 2] += -phi / H * (2.0 / n)
    g0, gp = h.tape_adjoint(gE)
    # two adjoint runs: their atomic accumulations may be summed in different orders
    scale = np.abs(g0[:, 3:6]).max()
    assert np.abs(v0.grad.numpy() - g0[:, 3:6]).max() <= 1e-9 * scale
    assert np.isclose(mu.grad.item(), gp[0], rtol=1e-9)
    assert np.isclose(e.grad.item(), gp[1], rtol=1e-9)
    assert np.abs(v0.grad.numpy()).max() > 0.0


def settling_bed_64(seed=5):
    """A reduced form of the 50,000-sphere differentiation case: 64 spheres r ~ U[0.9, 1.1] mm
    seeded at solid fraction 0.294 in a 3:3:10 periodic-x/y box over a floor, v ~ N(0, 0.01 m/s),
    mu 0.5, e 0.3, dt 1e-5 s."""
    from peclet.dem.workloads import sphere_bed

    n = 64
    mean_r3 = (1.1e-3 ** 4 - 0.9e-3 ** 4) / (4.0 * 0.2e-3)
    vol = n * 4.0 / 3.0 * np.pi * mean_r3 / 0.294
    a = (vol * 0.3) ** (1.0 / 3.0)
    # seeded in a column 1.3 x the box height (the few spheres of a small box jam sooner at the
    # floor and top faces than the 50,000 of the full case)
    bed = sphere_bed(n, 0.9e-3, 1.1e-3, (a, a, a / 0.3), seed, column=1.3 * a / 0.3)
    bed["vel"] = np.random.default_rng(seed).normal(0.0, 0.01, (n, 3))
    return bed


def loss_gradient(end, radius, area):
    """Loss = mean bed height + solid fraction below 0.6 H (H = 2 x the mean centre height) of an
    end state (n, 13), and its gradient with respect to the end state."""
    n = len(end)
    z = end[:, 2]
    H = 2.0 * z.mean()
    inside = z < 0.6 * H
    phi = (inside * 4.0 / 3.0 * np.pi * radius ** 3).sum() / (area * 0.6 * H)
    g = np.zeros((n, 13))
    g[:, 2] = 1.0 / n - phi / H * (2.0 / n)
    return z.mean() + phi, g


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_settling_bed_gradients_every_direction(hip_lib):
    """Every initial-velocity component, friction and restitution of the reduced differentiation
    case against central differences of the oracle, over 16 recorded substeps (two checkpoint
    segments) after 3000 substeps of settling.  The step is 1e-6 for the velocities and 1e-4 for
    the two material parameters, whose derivatives are small (|dL/de| ~ 2e-8): at 1e-6 their
    differences sit at the loss's rounding floor, at 1e-4 they agree with the adjoint to ~1e-5.
    The tolerance is 1e-4 relative plus that floor, 8 eps |L| / h.  Where the loss is not
    differentiable over the step (its one-sided slopes differ) the adjoint must equal one of them
    (the branch the forward pass took); at most 10 of the 194 directions may match neither, and
    the material parameters must match outright."""
    from peclet.dem.workloads import setup_gravity_bed

    bed = settling_bed_64()
    n, dt, pre, steps = len(bed["pos"]), 1e-5, 3000, 16
    area = bed["domain"][0] * bed["domain"][1]

    def prepared(sim):
        setup_gravity_bed(sim, bed, dt, 0.5, 0.3)
        sim.set_velocities(bed["vel"])
        for _ in range(pre):
            sim.step(dt)
        return sim

    h = prepared(hip_sim(n))
    assert h.get_stats()["active_manifolds"] > 0
    h.tape_begin(8)
    for _ in range(steps):
        h.step(dt)
    h.tape_end()
    _, gE = loss_gradient(state(h), bed["radius"], area)
    g0, gp = h.tape_adjoint(gE)

    def loss(dv=None, dmu=0.0, de=0.0):
        r = prepared(oracle_sim(n))
        if dv is not None:
            r.set_velocities(r.get_velocities() + dv)
        r.set_material_params(0.3 + de, 0.0, 0.5 + dmu)
        for _ in range(steps):
            r.step(dt)
        return loss_gradient(state(r), bed["radius"], area)[0]

    f0 = loss()
    checks = []
    for i in range(n):
        for c in range(3):
            d = np.zeros((n, 3))
            d[i, c] = 1.0
            checks.append((f"v[{i},{c}]", (lambda s, d=d: loss(dv=s * d)), g0[i, 3 + c]))
    checks.append(("restitution", lambda s: loss(de=s), gp[1]))
    checks.append(("friction", lambda s: loss(dmu=s), gp[0]))
    kinks, jumps = [], []
    worst = 0.0
    scale = max(abs(a) for _, _, a in checks)
    for name, f, adj in checks:
        step = 1e-4 if name in ("restitution", "friction") else 1e-6
        p6, m6 = f(step), f(-step)
        fd = (p6 - m6) / (2.0 * step)
        right, left = (p6 - f0) / step, (f0 - m6) / step
        err = abs(fd - adj)
        floor = 8.0 * np.finfo(float).eps * abs(f0) / step  # rounding of the loss itself
        tol = 1e-4 * abs(adj) + max(1e-8 * scale, floor)
        if err <= tol:
            worst = max(worst, err / max(abs(adj), 1e-30))
            continue
        # not differentiable at the base point: the adjoint follows the branch the forward pass
        # took, one of the one-sided slopes; matching neither means a jump inside the step
        side = 1e-3 * max(abs(right), abs(left)) + max(1e-8 * scale, 2.0 * floor)
        if min(abs(adj - right), abs(adj - left)) <= side:
            kinks.append(name)
        else:
            jumps.append(name)
        print(f"  {name}: fd {fd:.6e}, one-sided {right:.6e} {left:.6e}, adjoint {adj:.6e}",
              flush=True)
    print(f"  {len(checks) - len(kinks) - len(jumps)} of {len(checks)} directions within 1e-4 of "
          f"the central difference (worst relative {worst:.2e}); at a kink, on the forward "
          f"branch: {kinks}; a jump inside the step: {jumps}", flush=True)
    assert len(jumps) <= 10  # at most 5 % of the directions
    assert len(kinks) + len(jumps) <= 20
    for name in ("restitution", "friction"):
        assert name not in jumps and name not in kinks, name
    assert abs(gp[0]) > 0.0 and abs(gp[1]) > 0.0


@pytest.mark.gpu
def test_recording_guards(hip_lib):
    """A rollout's backward refuses once a later recording replaced its tape; setting masses
    while recording fails the next substep; state arrays of the wrong size are refused (Python
    shape check and the C-ABI's element count)."""
    import ctypes as C

    import torch

    from peclet.dem.adjoint import rollout

    n = 100
    scene = make_scene("bed_periodic", n, 0)
    h = apply_scene(hip_sim(n), scene)
    h.step(scene["dt"])
    v0 = torch.tensor(h.get_velocities(), dtype=torch.float64, requires_grad=True)
    g0 = h.tape_generation()
    first = rollout(h, v0, 0.5, 0.3, scene["dt"], 4, checkpoint_every=2)
    assert h.tape_generation() == g0 + 1
    second = rollout(h, v0, 0.5, 0.3, scene["dt"], 4, checkpoint_every=2)
    assert h.tape_generation() == g0 + 2
    with pytest.raises(RuntimeError, match="replaced"):
        first[:, 2].sum().backward()
    second[:, 2].sum().backward()  # the current recording still differentiates
    assert torch.all(torch.isfinite(v0.grad))
    with pytest.raises(ValueError, match="shape"):
        h.tape_adjoint(np.zeros((n, 6)))
    g = np.zeros((n, 13))
    out = np.zeros((n, 13))
    gp = np.zeros(2)
    dp = C.POINTER(C.c_double)
    rc = h._b.tape_adjoint(h._h, g.ctypes.data_as(dp), out.ctypes.data_as(dp),
                           gp.ctypes.data_as(dp), 13 * n - 1)
    assert rc == -2
    h.tape_begin(4)
    h.step(scene["dt"])
    h.set_inv_mass(np.full(n, 2.0))
    with pytest.raises(ValueError, match="while recording"):
        h.step(scene["dt"])
    h.tape_end()
