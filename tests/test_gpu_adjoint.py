"""The adjoint of one vpm(WhatToDo=2) velocity evaluation (vpmx_adjoint_probe_energy) on a
vortex ring: J = 1/2 sum_k |u(p_k)|^2 at probe points, and its gradient in the particle
strengths and positions.

  * J itself: the oracle's velocity grid (the reference's FP64 sequence) interpolated at the
    probes with the M4' weights equals the library's J to 1e-12.
  * dJ/dq: J is a quadratic form in the strengths, so J(q + v) - J(q - v) = 2 v . dJ/dq holds
    exactly for any v (up to rounding); a transposition mistake anywhere in the chain (a
    dropped boundary term, a wrong sign, a transposed stencil) breaks it.
  * dJ/dx: central finite differences in FP64 (step 1e-7, far below a cell) along the
    gradient itself and, for the particles with the largest gradients, per coordinate.
"""
import numpy as np
import pytest

from vpm_free_amd import cases

pytestmark = pytest.mark.gpu

H = 1.0 / 8  # a 33^3 global grid on [-2, 2]^3
NPROBE = 64


def ring_case():
    xp, qp = cases.single_ring(h=H, sigma=0.25)
    rng = np.random.default_rng(11)
    probes = rng.uniform(-1.4, 1.4, size=(NPROBE, 3))
    return xp, qp, probes


def m4(x):
    a = np.abs(x)
    return np.where(a < 1.0, 1.0 - 2.5 * a * a + 1.5 * a ** 3, np.where(a <= 2.0, 0.5 * (2.0 - a) ** 2 * (1.0 - a), 0.0))


def interp(vel, x0, h, pts):
    """vel: (3, nz, ny, nx) on nodes x0 + i h; M4' weights over the 4 x 4 x 4 nodes."""
    out = np.zeros((pts.shape[0], 3))
    n = vel.shape[1:][::-1]
    for p, x in enumerate(pts):
        t = (x - x0) / h
        c = np.floor(t).astype(int)
        for k in range(c[2] - 1, c[2] + 3):
            for j in range(c[1] - 1, c[1] + 3):
                for i in range(c[0] - 1, c[0] + 3):
                    if 0 <= i < n[0] and 0 <= j < n[1] and 0 <= k < n[2]:
                        out[p] += m4(t[0] - i) * m4(t[1] - j) * m4(t[2] - k) * vel[:, k, j, i]
    return out


@pytest.fixture(scope="module")
def ctx(vpmlib):
    V = vpmlib
    xp, qp, probes = ring_case()
    c = V.VPM(V.Params.make(H))
    c.set_precision(0)
    up, gp = np.zeros((xp.shape[0], 3)), np.zeros((xp.shape[0], 3))
    c.vpm(xp, qp, up, gp, xp.shape[0], 3, 0, 0)
    yield c, xp, qp, probes
    c.close()


def test_adjoint_energy_matches_oracle(ctx, oracle):
    c, xp, qp, probes = ctx
    print("case: adjoint J vs oracle, ring h=1/8,", xp.shape[0], "particles", flush=True)
    J, dq, dx = c.adjoint_probe_energy(xp, qp, probes)
    p = oracle.Params.make(H)
    g = oracle.define_sizes(p, xp)
    _, _, f = oracle.vpm2(p, g, xp, qp, fields=True)
    u = interp(f["vel"], np.array(g.xbound[:3]), H, probes)
    J_ref = 0.5 * float(np.sum(u * u))
    assert abs(J - J_ref) <= 1e-12 * J_ref, (J, J_ref)
    assert np.all(np.isfinite(dq)) and np.all(np.isfinite(dx))


def test_adjoint_strength_gradient_exact(ctx):
    c, xp, qp, probes = ctx
    print("case: adjoint dJ/dq, quadratic identity", flush=True)
    J, dq, _ = c.adjoint_probe_energy(xp, qp, probes)
    rng = np.random.default_rng(5)
    scale = np.abs(qp[:, :3]).max()
    for trial in range(3):
        v = rng.standard_normal((xp.shape[0], 3)) * scale
        qa, qb = qp.copy(), qp.copy()
        qa[:, :3] += v
        qb[:, :3] -= v
        Ja = c.adjoint_probe_energy(xp, qa, probes)[0]
        Jb = c.adjoint_probe_energy(xp, qb, probes)[0]
        fd = (Ja - Jb) / 2.0
        an = float(np.sum(v * dq))
        print("  trial", trial, "fd", fd, "adjoint", an, flush=True)
        assert abs(fd - an) <= 1e-9 * (abs(Ja) + abs(Jb)), (trial, fd, an)


def test_adjoint_position_gradient_fd(ctx):
    c, xp, qp, probes = ctx
    print("case: adjoint dJ/dx, central differences", flush=True)
    J, _, dx = c.adjoint_probe_energy(xp, qp, probes)
    eps = 1e-7
    v = dx / np.linalg.norm(dx)
    Ja = c.adjoint_probe_energy(xp + eps * v, qp, probes)[0]
    Jb = c.adjoint_probe_energy(xp - eps * v, qp, probes)[0]
    fd = (Ja - Jb) / (2 * eps)
    an = float(np.linalg.norm(dx))
    print("  along the gradient: fd", fd, "adjoint", an, flush=True)
    assert abs(fd - an) <= 1e-6 * an, (fd, an)
    top = np.argsort(-np.abs(dx).max(axis=1))[:4]
    for p in top:
        for d in range(3):
            xa, xb = xp.copy(), xp.copy()
            xa[p, d] += eps
            xb[p, d] -= eps
            fd = (c.adjoint_probe_energy(xa, qp, probes)[0] - c.adjoint_probe_energy(xb, qp, probes)[0]) / (2 * eps)
            assert abs(fd - dx[p, d]) <= 1e-5 * np.abs(dx[p]).max(), (p, d, fd, dx[p, d])


def test_adjoint_spectral_vs_direct_transposed_sums(ctx):
    """The transposed boundary sums by spectral correlation on the forward's plan give the
    gradient of the direct O(faces^2) sums to rounding."""
    c, xp, qp, probes = ctx
    print("case: adjoint, spectral vs direct transposed boundary sums", flush=True)
    J1, dq1, dx1 = c.adjoint_probe_energy(xp, qp, probes)
    c.set_option("adjoint_boundsum", 1)
    try:
        J2, dq2, dx2 = c.adjoint_probe_energy(xp, qp, probes)
    finally:
        c.set_option("adjoint_boundsum", 2)
    assert J1 == J2
    e = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in ((dq1, dq2), (dx1, dx2))]
    print("  rel L2 dq %.3e dx %.3e" % tuple(e), flush=True)
    assert max(e) <= 1e-12, e


def test_adjoint_m2p_gather_vs_scatter(ctx):
    """The transposed interpolation as a sorted gather projection of the particles' adjoints
    equals its FP64-atomic scatter form to rounding (one roll-out step exercises it)."""
    c, xp, qp, probes = ctx
    print("case: roll-out, M2P^T gather vs scatter", flush=True)
    J1, dq1, dx1 = c.adjoint_rollout(xp, qp, probes, 1, 0.05)
    c.set_option("adjoint_m2p_scatter", 1)
    try:
        J2, dq2, dx2 = c.adjoint_rollout(xp, qp, probes, 1, 0.05)
    finally:
        c.set_option("adjoint_m2p_scatter", 0)
    assert J1 == J2
    e = [float(np.linalg.norm(a - b) / np.linalg.norm(b)) for a, b in ((dq1, dq2), (dx1, dx2))]
    print("  rel L2 dq %.3e dx %.3e" % tuple(e), flush=True)
    assert max(e) <= 1e-12, e


def test_adjoint_autograd_function(ctx):
    import torch
    from vpm_free_amd.autograd import probe_energy
    c, xp, qp, probes = ctx
    print("case: adjoint through torch.autograd", flush=True)
    X = torch.tensor(xp, requires_grad=True)
    Q = torch.tensor(qp, requires_grad=True)
    J = probe_energy(c, X, Q, torch.tensor(probes))
    J.backward()
    J2, dq, dx = c.adjoint_probe_energy(xp, qp, probes)
    assert float(J.detach()) == J2
    # the scatters' FP64 atomics leave the gradients' last bits to the order of the adds
    assert np.linalg.norm(X.grad.numpy() - dx) <= 1e-12 * np.linalg.norm(dx)
    assert np.linalg.norm(Q.grad.numpy()[:, :3] - dq) <= 1e-12 * np.linalg.norm(dq)
    assert not Q.grad.numpy()[:, 3].any()


def test_adjoint_rejects_several_blocks(vpmlib):
    V = vpmlib
    xp, qp, probes = ring_case()
    with V.VPM(V.Params.make(H, blocks=(2, 1, 1))) as c:
        up, gp = np.zeros((xp.shape[0], 3)), np.zeros((xp.shape[0], 3))
        c.vpm(xp, qp, up, gp, xp.shape[0], 3, 0, 0)
        with pytest.raises(Exception):
            c.adjoint_probe_energy(xp, qp, probes)


# ---------------------------------------------------------------- roll-out
NSTEPS, DT = 4, 0.05


def oracle_rollout_energy(oracle, xp, qp, probes):
    p = oracle.Params.make(H)
    g = oracle.define_sizes(p, xp)
    for k in range(NSTEPS):
        xp, qp = oracle.step_rk2(p, g, xp, qp, DT, ntime=k + 1)
    _, _, f = oracle.vpm2(p, g, xp, qp, fields=True)
    u = interp(f["vel"], np.array(g.xbound[:3]), H, probes)
    return 0.5 * float(np.sum(u * u)), xp.shape[0]


def test_rollout_energy_matches_oracle(ctx, oracle):
    c, xp, qp, probes = ctx
    print("case: roll-out J vs oracle,", NSTEPS, "steps", flush=True)
    J, dq, dx = c.adjoint_rollout(xp, qp, probes, NSTEPS, DT, ckpt=2)
    J_ref, n_ref = oracle_rollout_energy(oracle, xp, qp, probes)
    print("  J", J, "oracle", J_ref, "particles after", n_ref, flush=True)
    assert abs(J - J_ref) <= 1e-10 * J_ref, (J, J_ref)
    assert np.all(np.isfinite(dq)) and np.all(np.isfinite(dx))


def test_rollout_zero_steps_is_probe_energy(ctx):
    c, xp, qp, probes = ctx
    print("case: roll-out of 0 steps", flush=True)
    J0, dq0, dx0 = c.adjoint_probe_energy(xp, qp, probes)
    J1, dq1, dx1 = c.adjoint_rollout(xp, qp, probes, 0, DT)
    assert J0 == J1
    for a, b in ((dq0, dq1), (dx0, dx1)):
        assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(a)


def test_rollout_checkpoint_spacing(ctx):
    c, xp, qp, probes = ctx
    print("case: roll-out checkpoints every 1, 2, 3 steps and every step kept", flush=True)
    res = [c.adjoint_rollout(xp, qp, probes, NSTEPS, DT, ckpt=k) for k in (1, 2, 3, 0)]
    for J, dq, dx in res[1:]:
        assert J == res[0][0]
        assert np.linalg.norm(dq - res[0][1]) <= 1e-12 * np.linalg.norm(res[0][1])
        assert np.linalg.norm(dx - res[0][2]) <= 1e-12 * np.linalg.norm(res[0][2])


def test_rollout_gradient_fd(ctx):
    """Central differences in FP64: strengths along a random direction, each particle's
    perturbation 1e-4 of its own strength; positions along the gradient (step 1e-7).  The
    remesh threshold mask is piecewise constant and the gradient holds it fixed, so the steps
    must not switch nodes on: a perturbation of fixed size would give the zero-strength
    particles (the box corners) strengths that the remesh then keeps."""
    c, xp, qp, probes = ctx
    print("case: roll-out gradient vs central differences", flush=True)
    J, dq, dx = c.adjoint_rollout(xp, qp, probes, NSTEPS, DT)
    rng = np.random.default_rng(9)
    v = rng.standard_normal((xp.shape[0], 3)) * qp[:, :3] * 1e-4
    qa, qb = qp.copy(), qp.copy()
    qa[:, :3] += v
    qb[:, :3] -= v
    fd = (c.adjoint_rollout(xp, qa, probes, NSTEPS, DT)[0] - c.adjoint_rollout(xp, qb, probes, NSTEPS, DT)[0]) / 2
    an = float(np.sum(v * dq))
    print("  strengths: fd", fd, "adjoint", an, flush=True)
    assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)
    eps = 1e-7
    d = dx / np.linalg.norm(dx)
    fd = (c.adjoint_rollout(xp + eps * d, qp, probes, NSTEPS, DT)[0] -
          c.adjoint_rollout(xp - eps * d, qp, probes, NSTEPS, DT)[0]) / (2 * eps)
    an = float(np.linalg.norm(dx))
    print("  positions: fd", fd, "adjoint", an, flush=True)
    assert abs(fd - an) <= 1e-6 * an, (fd, an)


def test_rollout_autograd_function(ctx):
    import torch
    from vpm_free_amd.autograd import rollout_energy
    c, xp, qp, probes = ctx
    print("case: roll-out through torch.autograd", flush=True)
    X = torch.tensor(xp, requires_grad=True)
    Q = torch.tensor(qp, requires_grad=True)
    J = rollout_energy(c, X, Q, torch.tensor(probes), NSTEPS, DT)
    (2.0 * J).backward()
    _, dq, dx = c.adjoint_rollout(xp, qp, probes, NSTEPS, DT)
    assert np.linalg.norm(X.grad.numpy() - 2 * dx) <= 1e-12 * np.linalg.norm(dx)
    assert np.linalg.norm(Q.grad.numpy()[:, :3] - 2 * dq) <= 1e-12 * np.linalg.norm(dq)
