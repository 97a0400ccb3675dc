"""Hand-written adjoint of the substep (GPU), pinned by float64 central
finite differences of the same differentiable forward, and the
differentiable forward pinned against the plain substep."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene(fx):
    from flowx_amd.config import converging_update

    cfg = converging_update(fx.lattice_config((-0.5, -0.5, 0.0), (0.5, 0.5, 1.0), 0.1, (10, 10, 6), iterations=4))
    cfg.extra.update(colliders=[("box", (-0.2, -0.2, 0.0), (0.2, 0.2, 0.2))], collider_resolution=10,
                     surface_resolution=24, fps=24.0)
    return cfg


@pytest.mark.parametrize("with_collider", [False, True])
def test_adjoint_forward_matches_plain_substep(fx, with_collider):
    import torch

    from flowx_amd.adjoint import DiffSolver

    def scene():
        c = _scene(fx)
        if not with_collider:
            c.extra["colliders"] = []
        return c

    cfg = scene()
    plain = fx.Solver(cfg, keep_scratch=False)
    plain.reset()
    dt = plain.cfg.substep_dt(1 / 24.0)[1]
    # the seeded lattice (large first corrections) and a developed state
    for label, warm in (("seeded", 0), ("developed", 6)):
        plain.substeps_n(dt, warm) if warm else None
        x0, v0 = plain.pos.clone(), plain.vel.clone()
        plain.substep(dt)
        for dtype in (torch.float32, torch.float64):
            d = DiffSolver(scene(), dtype=dtype)
            x1, v1 = d.forward_raw(x0.to(dtype).contiguous(), v0.to(dtype).contiguous(), -9.81, (0, 0, 0))
            torch.cuda.synchronize()
            e = (x1[:, :3].double() - plain.pos[:, :3].double()).abs().max(1).values.cpu().numpy()
            print(f"{label} collider={with_collider} {dtype}: |dx| max {e.max():.3e} p99 "
                  f"{np.percentile(e, 99):.3e} median {np.median(e):.3e}")
            assert not d.overflowed()
            # one substep chains 4 Jacobi iterations: float32 rounding of
            # ~1e-7 per operation reaches ~1e-5 h on a developed state
            bar = 1e-5 if label == "seeded" else 1e-4
            assert e.max() < bar * cfg.smoothing_radius
        plain.load_state(x0.cpu().numpy(), v0.cpu().numpy())
        plain.substep(dt)


def _loss(fx, d, x0, v0, g, tr, steps, target, com_t, decisions=None, mode=0):
    from flowx_amd.adjoint import loss_fn, rollout

    x, v = rollout(d, x0, v0, g, tr, steps, decisions, mode)
    return loss_fn(d, x, target, com_t)


def test_adjoint_matches_finite_differences_fp64(fx):
    import torch

    from flowx_amd.adjoint import DiffSolver, gaussian_target

    torch.manual_seed(0)
    cfg = _scene(fx)
    d = DiffSolver(cfg, dtype=torch.float64)
    x0 = torch.as_tensor(fx.seed_positions(cfg), dtype=torch.float64).cuda()
    x0[:, 3] = 0
    rng = np.random.default_rng(1)
    v0 = torch.as_tensor(rng.normal(0, 0.05, (len(x0), 4)), dtype=torch.float64).cuda()
    v0[:, 3] = 0
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
    tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    target = gaussian_target(d, (0.1, -0.1, 0.2), sigma=0.2)
    com_t = torch.tensor([0.05, 0.0, 0.15], dtype=torch.float64, device="cuda")
    steps = 4
    v = v0.clone().requires_grad_(True)
    decisions = []
    L = _loss(fx, d, x0, v, g, tr, steps, target, com_t, decisions, 1)
    L.backward()
    gv, gg, gt = v.grad.clone(), g.grad.item(), tr.grad.cpu().numpy()

    def L_at(dv=None, dg=0.0, dtr=None):
        with torch.no_grad():
            vv = v0 + (dv if dv is not None else 0)
            tt = tr.detach() + (torch.as_tensor(dtr, dtype=torch.float64, device="cuda") if dtr is not None else 0)
            return _loss(fx, d, x0, vv, g.detach() + dg, tt, steps, target, com_t, decisions, 2).item()

    eps = 1e-6
    worst = 0.0
    for k in range(8):
        dirn = torch.as_tensor(rng.normal(0, 1, v0.shape), dtype=torch.float64).cuda()
        dirn[:, 3] = 0
        dirn /= dirn.norm()
        fd = (L_at(dv=eps * dirn) - L_at(dv=-eps * dirn)) / (2 * eps)
        an = float((gv * dirn).sum())
        rel = abs(fd - an) / max(abs(fd), 1e-12)
        worst = max(worst, rel)
        print(f"v0 direction {k}: fd {fd:.9e} adjoint {an:.9e} rel {rel:.2e}")
    fdg = (L_at(dg=eps) - L_at(dg=-eps)) / (2 * eps)
    print(f"gravity: fd {fdg:.9e} adjoint {gg:.9e}")
    worst = max(worst, abs(fdg - gg) / max(abs(fdg), 1e-12))
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps
        fdt = (L_at(dtr=e) - L_at(dtr=-e)) / (2 * eps)
        print(f"translation[{a}]: fd {fdt:.9e} adjoint {gt[a]:.9e}")
        if abs(fdt) > 1e-8:
            worst = max(worst, abs(fdt - gt[a]) / abs(fdt))
        else:
            assert abs(gt[a]) < 1e-6
    assert worst < 1e-3


def _cluster_scene(fx, lattice, cluster, seed=3):
    """A lattice plus `cluster` particles packed into a sphere of radius
    0.4 h: the cluster's neighbour lists are far longer than the average."""
    import torch

    from flowx_amd.adjoint import DiffSolver
    from flowx_amd.config import converging_update

    cfg = converging_update(fx.lattice_config((-0.5, -0.5, 0.0), (0.5, 0.5, 1.0), 0.1, lattice, iterations=4))
    cfg.extra.update(colliders=[], surface_resolution=16, fps=24.0)
    base = fx.seed_positions(cfg) if lattice[0] else np.zeros((0, 4), np.float32)
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(cluster, 3))
    u *= (0.04 * rng.uniform(0, 1, (cluster, 1)) ** (1 / 3)) / np.linalg.norm(u, axis=1, keepdims=True)
    extra = np.concatenate([u + np.array([0.0, 0.0, 0.3]), np.ones((cluster, 1))], 1).astype(np.float32)
    pos = np.concatenate([base, extra])
    cfg.particle_count = len(pos)
    d = DiffSolver(cfg, dtype=torch.float64)
    x0 = torch.as_tensor(pos, dtype=torch.float64).cuda()
    x0[:, 3] = 0
    return cfg, d, x0


def test_adjoint_long_lists_in_dense_cluster(fx):
    """Lists longer than the average capacity (192) in a dense cluster: the
    adjoint still matches float64 finite differences, and replaying the
    record reproduces the recorded substep exactly."""
    import torch

    from flowx_amd.adjoint import gaussian_target, loss_fn, rollout

    cfg, d, x0 = _cluster_scene(fx, (20, 20, 8), 250)
    n = len(x0)
    rng = np.random.default_rng(4)
    v0 = torch.as_tensor(rng.normal(0, 0.05, (n, 4)), dtype=torch.float64).cuda()
    v0[:, 3] = 0
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
    tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    target = gaussian_target(d, (0.0, 0.0, 0.25), sigma=0.2)
    com_t = torch.tensor([0.0, 0.0, 0.2], dtype=torch.float64, device="cuda")
    for steps in (1, 2):
        v = v0.clone().requires_grad_(True)
        decisions = []
        x, _ = rollout(d, x0, v, g, tr, steps, decisions, 1)
        L = loss_fn(d, x, target, com_t)
        L.backward()
        offs = decisions[0][2 * n: 3 * n + 1].cpu().numpy().astype(np.int64)
        longest = int(np.diff(offs).max())
        print(f"steps {steps}: longest list {longest}, mean {offs[-1] / n:.1f}")
        assert longest > 192 and offs[-1] < 192 * n
        with torch.no_grad():
            xr, _ = rollout(d, x0, v0, g.detach(), tr.detach(), steps, decisions, 2)
        assert torch.equal(xr, x.detach())
        gv = v.grad.clone()
        eps = 1e-7
        for k in range(3):
            dirn = torch.as_tensor(rng.normal(0, 1, v0.shape), dtype=torch.float64).cuda()
            dirn[:, 3] = 0
            dirn /= dirn.norm()
            with torch.no_grad():
                lp = loss_fn(d, rollout(d, x0, v0 + eps * dirn, g.detach(), tr.detach(), steps, decisions, 2)[0],
                             target, com_t).item()
                lm = loss_fn(d, rollout(d, x0, v0 - eps * dirn, g.detach(), tr.detach(), steps, decisions, 2)[0],
                             target, com_t).item()
            fd = (lp - lm) / (2 * eps)
            an = float((gv * dirn).sum())
            rel = abs(fd - an) / max(abs(fd), 1e-12)
            print(f"  direction {k}: fd {fd:.9e} adjoint {an:.9e} rel {rel:.2e}")
            assert rel < 1e-4


def test_adjoint_recorded_lists_are_sorted_stencil_pairs(fx):
    """The recorded neighbour lists of a float64 substep over a lattice plus
    a 300-particle cluster (lists past the 256 entries one wave sorts in
    registers, so the queued block sort runs too): each list is exactly the
    set of j != i within h (1 + skin) of i's predicted position whose cell is
    within one of i's cell on every axis, sorted by index.  Brute force over
    all pairs on the host: at zero velocity the predicted position is
    x + dt (g dt), and the host repeats the device's float64 and float32
    operations one for one, so the sets compare exactly."""
    import torch

    from flowx_amd.adjoint import rollout

    cfg, d, x0 = _cluster_scene(fx, (20, 20, 8), 300)
    n = len(x0)
    decisions = []
    with torch.no_grad():
        rollout(d, x0, torch.zeros_like(x0), torch.tensor(-9.81, dtype=torch.float64, device="cuda"),
                torch.zeros(3, dtype=torch.float64, device="cuda"), 1, decisions, 1)
    rec = decisions[0].cpu().numpy()
    offs = rec[2 * n: 3 * n + 1].astype(np.int64)
    lists = rec[3 * n + 1:]
    p = d.params
    dt = float(p.dt)
    P = x0[:, :3].cpu().numpy().copy()
    P[:, 2] = P[:, 2] + dt * (0.0 + -9.81 * dt)
    lo = np.array([p.lo[0], p.lo[1], p.lo[2]], np.float32)
    cells = np.array([p.cells[0], p.cells[1], p.cells[2]])
    cell = np.floor((P.astype(np.float32) - lo) / np.float32(p.cell_size)).astype(np.int64)
    cell = np.clip(cell, 0, cells - 1)
    R = float(p.h) * (1.0 + d.list_skin)
    r2 = R * R
    assert offs[0] == 0 and np.all(np.diff(offs) >= 0) and offs[-1] < 192 * n
    longest = int(np.diff(offs).max())
    assert longest > 256, longest
    for a in range(0, n, 500):
        b = min(n, a + 500)
        dd = P[a:b, None, :] - P[None, :, :]
        dist2 = dd[..., 0] * dd[..., 0] + dd[..., 1] * dd[..., 1] + dd[..., 2] * dd[..., 2]
        near = np.all(np.abs(cell[a:b, None, :] - cell[None, :, :]) <= 1, axis=2)
        want = (dist2 < r2) & near
        want[np.arange(b - a), np.arange(a, b)] = False
        for i in range(a, b):
            got = lists[offs[i]:offs[i + 1]]
            assert np.array_equal(got, np.flatnonzero(want[i - a])), i


def test_rollout_reuses_preallocated_records(fx):
    """A recording rollout into records allocated ahead equals one that
    appends new records: same positions, same records."""
    import torch

    from flowx_amd.adjoint import rollout

    cfg, d, x0 = _cluster_scene(fx, (12, 12, 6), 50)
    v0 = torch.zeros_like(x0)
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda")
    tr = torch.zeros(3, dtype=torch.float64, device="cuda")
    appended = []
    with torch.no_grad():
        xa, va = rollout(d, x0, v0, g, tr, 3, appended, 1)
        ahead = [d.new_decisions().fill_(-7) for _ in range(3)]
        xb, vb = rollout(d, x0, v0, g, tr, 3, ahead, 1)
    assert len(ahead) == 3 and torch.equal(xa, xb) and torch.equal(va, vb)
    n = len(x0)
    for ra, rb in zip(appended, ahead):
        used = 3 * n + 1 + int(ra[3 * n].item())  # branches, offsets, then the lists
        assert torch.equal(ra[:used], rb[:used])


def test_adjoint_list_overflow_raises(fx):
    """More pairs than n * 192 in total: the capped lists are flagged and the
    rollout raises instead of returning a wrong gradient."""
    import torch

    from flowx_amd.adjoint import rollout

    cfg, d, x0 = _cluster_scene(fx, (0, 0, 0), 400)
    v0 = torch.zeros_like(x0)
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda")
    tr = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="overflow"):
        rollout(d, x0, v0, g, tr, 1, [], 1)
    assert not d.overflowed()  # reading the flag cleared it


def test_nonfinite_state_and_adjoint_raise(fx):
    """A rollout whose state is not finite raises, and so does the backward
    of one whose output adjoint is not finite (flagged on the device by the
    substep that produced it, reported by the rollout's first substep)."""
    import torch

    from flowx_amd.adjoint import NonFiniteError, Substep, rollout

    cfg = _scene(fx)
    from flowx_amd.adjoint import DiffSolver

    d = DiffSolver(cfg, dtype=torch.float64)
    x0 = torch.as_tensor(fx.seed_positions(cfg), dtype=torch.float64).cuda()
    x0[:, 3] = 0
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
    tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    bad = torch.zeros_like(x0)
    bad[5, 0] = float("nan")
    with pytest.raises(NonFiniteError):
        rollout(d, x0, bad, g, tr, 2, [], 1)
    v = torch.zeros_like(x0).requires_grad_(True)
    x, _ = rollout(d, x0, v, g, tr, 3, [], 1)
    seed_adj = torch.zeros_like(x)
    seed_adj[len(x) // 2 + 5, 1] = float("inf")  # a particle in the middle (a wall clamp would zero it)
    with pytest.raises(NonFiniteError):
        x.backward(seed_adj)
    # the device flag was cleared by the report: a finite backward passes
    v = torch.zeros_like(x0).requires_grad_(True)
    x, _ = rollout(d, x0, v, g, tr, 3, [], 1)
    x[:, :3].sum().backward()
    assert torch.isfinite(v.grad).all()


def _config5(fx, dtype):
    import torch

    from flowx_amd.adjoint import DiffSolver, gaussian_target

    d = DiffSolver(fx.adjoint_config(), dtype=dtype)
    x0 = torch.as_tensor(fx.seed_positions(d.cfg), dtype=dtype).cuda()
    x0[:, 3] = 0
    rng = np.random.default_rng(0)  # the target blob: a seeded point in the lower half
    centre = (float(rng.uniform(-0.5, 0.5)), float(rng.uniform(-0.5, 0.5)), float(rng.uniform(0.1, 0.9)))
    target = gaussian_target(d, centre, sigma=0.2)
    com_t = torch.tensor(centre, dtype=dtype, device="cuda")
    return d, x0, target, com_t


def test_config5_forward_matches_plain_substep(fx):
    """The 200k-particle adjoint scene (h = 0.05, 80x80x31, 0.4 m box): the
    differentiable forward's first substep equals the plain substep within
    1e-5 h (float32 and float64)."""
    import torch

    cfg = fx.adjoint_config()
    plain = fx.Solver(cfg, keep_scratch=False)
    plain.reset()
    dt = cfg.substep_dt(1 / 24.0)[1]
    x0, v0 = plain.pos.clone(), plain.vel.clone()
    plain.substep(dt)
    ref = plain.pos[:, :3].double()
    for dtype in (torch.float32, torch.float64):
        d, _, _, _ = _config5(fx, dtype)
        x1, _ = d.forward_raw(x0.to(dtype).contiguous(), v0.to(dtype).contiguous(), -9.81, (0, 0, 0))
        e = float((x1[:, :3].double() - ref).abs().max())
        print(f"200k {dtype}: |dx| max {e:.3e} ({e / cfg.smoothing_radius:.2e} h)")
        assert not d.overflowed()
        assert e < 1e-5 * cfg.smoothing_radius


@pytest.mark.parametrize("steps,eps,eps_t", [(4, 1e-4, 1e-8), (8, 1e-7, 1e-10)])
def test_config5_adjoint_matches_finite_differences(fx, steps, eps, eps_t):
    """Float64 central differences against the adjoint on the 200k scene,
    along 32 seeded initial-velocity directions and for gravity (eps, the
    configuration's 1e-4 at 4 substeps) and the three collider-translation
    axes (eps_t).  Bar: 1e-3 relative; a translation component is compared
    relative to the translation gradient's norm (at 8 substeps its z
    component is 1/700 of the norm, below what a difference quotient
    resolves).  Why these horizons: the substep map amplifies perturbations
    about tenfold per substep once the dam collapses (|dL/dv0| = 0.027 at 4
    substeps, 3.9 at 8, 1.3e8 at 16 in float32), so a perturbation of eps
    leaves the linear range after a few substeps - at 8 substeps the
    translation quotient converges to the adjoint only below eps_t = 1e-9
    (1e-9: 2.3e-2 on z; 1e-10: 1.0e-3) - and the loss is no longer a useful
    function of v0 past ~10 substeps.  The perturbed runs replay the base
    run's branches and neighbour lists."""
    import torch

    from flowx_amd.adjoint import loss_fn, rollout

    d, x0, target, com_t = _config5(fx, torch.float64)
    v0 = torch.zeros_like(x0)
    g = torch.tensor(-9.81, dtype=torch.float64, device="cuda", requires_grad=True)
    tr = torch.zeros(3, dtype=torch.float64, device="cuda", requires_grad=True)
    v = v0.clone().requires_grad_(True)
    dec = []
    x, _ = rollout(d, x0, v, g, tr, steps, dec, 1)
    loss_fn(d, x, target, com_t).backward()
    gv, gg, gt = v.grad.clone(), g.grad.item(), tr.grad.cpu().numpy()

    def L_at(dv=None, dg=0.0, dtr=None):
        with torch.no_grad():
            vv = v0 + (dv if dv is not None else 0)
            tt = tr.detach() + (torch.as_tensor(dtr, dtype=torch.float64, device="cuda") if dtr is not None else 0)
            xx, _ = rollout(d, x0, vv, g.detach() + dg, tt, steps, dec, 2)
            return loss_fn(d, xx, target, com_t).item()

    rng = np.random.default_rng(7)
    errs = []
    for _ in range(32):
        dn = torch.as_tensor(rng.normal(0, 1, v0.shape), dtype=torch.float64).cuda()
        dn[:, 3] = 0
        dn /= dn.norm()
        fd = (L_at(dv=eps * dn) - L_at(dv=-eps * dn)) / (2 * eps)
        errs.append(abs(fd - float((gv * dn).sum())) / abs(fd))
    fdg = (L_at(dg=eps) - L_at(dg=-eps)) / (2 * eps)
    e_g = abs(fdg - gg) / abs(fdg)
    fdt = []
    for a in range(3):
        e = np.zeros(3)
        e[a] = eps_t
        fdt.append((L_at(dtr=e) - L_at(dtr=-e)) / (2 * eps_t))
    fdt = np.array(fdt)
    e_t = float(np.abs(fdt - gt).max() / np.linalg.norm(fdt))
    print(f"{steps} substeps: |dL/dv0| {float(gv.norm()):.3e}; 32 directions max {max(errs):.2e} median "
          f"{np.median(errs):.2e}; gravity {e_g:.2e}; translation fd {fdt} adjoint {gt} -> {e_t:.2e} of the norm")
    assert max(errs) <= 1e-3 and e_g <= 1e-3 and e_t <= 1e-3
