"""Material gradient dJ/d(lam, mu, rho) of a receiver misfit.

The CPU reference gradient is reverse-mode differentiation (torch autograd,
float64) of a vectorised restatement of the forward step (oracle/torch_ref);
that restatement is pinned to the FP64 oracle's traces, and its gradient to
central finite differences.  The GPU gradient comes from the hand-written
adjoint kernels (awp_adjoint_step) with checkpointed recomputation; it is
checked against the CPU reference gradient field by field, against central
finite differences of the GPU misfit along seeded random directions, and
for independence of the checkpoint interval.
"""
import numpy as np
import pytest
import torch

from awp_amd import configs

NT = 60


def _case():
    return configs.adjoint_pair(nx=20, ny=18, nz=16, nt=NT, DH=100.0, DT=0.005, nsrc=2,
                                nrec_side=3, f0=4.0)


def _arrays(med):
    return (torch.from_numpy(med.lam.copy()), torch.from_numpy(med.mu.copy()),
            torch.from_numpy(med.d1.copy()))


def _reference(bg, true, src, rec):
    from oracle.torch_ref import TorchForward
    obs = TorchForward(true, src, rec, NT).run(*_arrays(true)).detach()
    tf = TorchForward(bg, src, rec, NT)
    lam, mu, rho = (a.requires_grad_(True) for a in _arrays(bg))
    d = tf.run(lam, mu, rho)
    J = 0.5 * ((d - obs) ** 2).sum()
    J.backward()
    return tf, obs, float(J), lam.grad, mu.grad, rho.grad


@pytest.fixture(scope="module")
def case():
    bg, true, src, rec = _case()
    return (bg, true, src, rec) + _reference(bg, true, src, rec)


def test_torch_reference_matches_oracle(oracle_lib, case):
    from oracle.oracle import OracleSolver
    bg, true, src, rec, tf = case[:5]
    o = OracleSolver(bg, src, rec, nt_record=NT, dtype=np.float64)
    o.run(NT)
    d = tf.run(*_arrays(bg)).detach().numpy()
    assert np.abs(o.traces()).max() > 0
    assert np.linalg.norm(d - o.traces()) <= 1e-12 * np.linalg.norm(o.traces())


def test_reference_gradient_matches_finite_differences(case):
    bg, true, src, rec, tf, obs, J0, gl, gm, gr = case
    rng = np.random.default_rng(5)
    lam0, mu0, rho0 = _arrays(bg)
    for name, base, g in (("lam", lam0, gl), ("mu", mu0, gm), ("rho", rho0, gr)):
        dirn = torch.from_numpy(rng.standard_normal(base.shape)) * base
        eps = 1e-4

        def J(a):
            args = {"lam": lam0, "mu": mu0, "rho": rho0}
            args[name] = a
            d = tf.run(args["lam"], args["mu"], args["rho"]).detach()
            return float(0.5 * ((d - obs) ** 2).sum())
        fd = (J(base + eps * dirn) - J(base - eps * dirn)) / (2 * eps)
        ad = float((g * dirn).sum())
        assert abs(fd - ad) <= 1e-5 * abs(ad), (name, fd, ad)


@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return True


def _gpu_problem(case, K, resident_steps=0):
    from awp_amd.adjoint import AdjointProblem
    bg, true, src, rec, tf, obs = case[:6]
    return AdjointProblem(bg, src, rec, NT, obs.numpy().astype(np.float32), checkpoint_every=K,
                          resident_steps=resident_steps)


def _rel(a, b):
    return float(torch.linalg.norm((a - b).flatten()) / torch.linalg.norm(b.flatten()))


@pytest.mark.gpu
def test_gpu_gradient_matches_reference(gpu, case):
    from awp_amd.adjoint import misfit
    bg = case[0]
    J0, gl, gm, gr = case[6:]
    P = _gpu_problem(case, 7)
    lam, mu, rho = (a.clone().requires_grad_(True) for a in _arrays(bg))
    J = misfit(lam, mu, rho, P)
    J.backward()
    errs = {n: _rel(a.grad, b) for n, a, b in (("lam", lam, gl), ("mu", mu, gm), ("rho", rho, gr))}
    print(f"GPU adjoint vs torch-AD FP64 reference: J {float(J):.6e} vs {J0:.6e}; "
          f"gradient rel-L2 {errs}")
    assert abs(float(J) / J0 - 1) < 1e-4
    for n, e in errs.items():
        assert e < 2e-3, (n, e)


@pytest.mark.gpu
def test_gpu_gradient_matches_finite_differences(gpu, case):
    """Directional derivatives along 5 seeded random directions (relative
    size 1 of each array's magnitude, all three arrays at once) against
    central differences of the GPU misfit."""
    bg = case[0]
    P = _gpu_problem(case, 10)
    lam0, mu0, rho0 = _arrays(bg)
    P.set_medium(lam0, mu0, rho0)
    P.forward()
    glam, gmu, grho = (g.cpu() for g in P.backward())
    rng = np.random.default_rng(9)
    worst = 0.0
    for t in range(5):
        dl, dm, dr = (torch.from_numpy(rng.standard_normal(a.shape)) * a for a in (lam0, mu0, rho0))
        eps = 2e-3

        def J(s):
            P.set_medium(lam0 + s * dl, mu0 + s * dm, rho0 + s * dr)
            return P.forward()
        fd = (J(eps) - J(-eps)) / (2 * eps)
        ad = float((glam * dl).sum() + (gmu * dm).sum() + (grho * dr).sum())
        err = abs(fd - ad) / abs(ad)
        worst = max(worst, err)
        print(f"direction {t}: adjoint {ad:.6e} finite difference {fd:.6e} rel {err:.2e}")
    assert worst < 1e-3


@pytest.mark.gpu
def test_gpu_gradient_independent_of_checkpoint_interval(gpu, case):
    bg = case[0]
    out = []
    for K in (1, 13, NT):
        P = _gpu_problem(case, K)
        P.set_medium(*_arrays(bg))
        P.forward()
        out.append([g.cpu() for g in P.backward()])
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            assert _rel(a, b) < 1e-6


@pytest.mark.gpu
def test_gpu_out_of_place_recompute_matches_in_place(gpu, case):
    """The segment recompute (awp_dvelcx_oop / awp_dstrqc_oop into the ring
    of states) reproduces the in-place forward (dvelcx_H_opt, dstrqc_H) bit
    for bit: velocities and stresses of every step, memory variables after
    the segment; one segment from the start, one from a mid-run checkpoint."""
    from awp_amd.adjoint import STRESS, VEL
    from awp_amd.solver import FIELDS
    K = 13
    P = _gpu_problem(case, K)
    P.set_medium(*_arrays(case[0]))
    P.forward()
    s = P.solver
    R = ("r1", "r2", "r3", "r4", "r5", "r6")
    for n0 in (0, 26):
        for x, t in zip(FIELDS, P.checkpoints[n0]):
            s.f[x].copy_(t)
        s.cur_step = n0
        ref = []
        for _ in range(K):
            s.step()
            ref.append({x: s.f[x].clone() for x in VEL + STRESS})
        ref_r = {x: s.f[x].clone() for x in R}
        for x, t in zip(FIELDS, P.checkpoints[n0]):
            (P.ring[0][x] if x in P.ring[0] else s.f[x]).copy_(t)
        s.cur_step = n0
        for n in range(K):
            P._recompute_step(P.ring[n], P.ring[n + 1])
        torch.cuda.synchronize()
        for n in range(K):
            for x in VEL + STRESS:
                assert torch.equal(P.ring[n + 1][x], ref[n][x]), (n0, n, x)
        for x in R:
            assert torch.equal(s.f[x], ref_r[x]), (n0, x)


@pytest.mark.gpu
def test_gpu_resident_tail_gives_the_same_gradient(gpu, case):
    """States of the last segments kept from the forward pass (no recompute
    for them) against recomputing every segment: the same states, so the
    same misfit, traces and gradients bit for bit; a partial tail (2 of 6
    segments) and a fully resident run."""
    bg = case[0]
    out = []
    for resident in (0, 25, None):
        P = _gpu_problem(case, 10, resident_steps=resident)
        print(f"resident_steps={resident}: first resident step {P.n_tail}", flush=True)
        P.set_medium(*_arrays(bg))
        J = P.forward()
        tr = P.traces()
        out.append((J, tr, [g.cpu() for g in P.backward()], P.n_tail))
    assert out[0][3] > NT and out[1][3] == 40 and out[2][3] == 0
    for J, tr, g, _ in out[1:]:
        assert J == out[0][0]
        assert np.array_equal(tr, out[0][1])
        for a, b in zip(g, out[0][2]):
            assert torch.equal(a, b)


@pytest.mark.gpu
@pytest.mark.parametrize("jgroup", [0, 3, 64])
def test_gpu_block_order_gives_the_same_gradient(gpu, case, jgroup):
    """The adjoint kernels' block order (awp_set_tuning("adj_jgroup"): plane
    by plane, groups of 3 j tiles over all planes with a partial last group,
    one group of every tile; the default is groups of 8) only changes which
    blocks run together, not what any cell computes: misfit and gradients bit
    for bit as the default."""
    from awp_amd import _lib
    lib = _lib.load()
    bg = case[0]
    out = []
    for g in (None, jgroup):
        if g is not None:
            assert lib.awp_set_tuning(b"adj_jgroup", g) == 0
        try:
            P = _gpu_problem(case, 10)
            P.set_medium(*_arrays(bg))
            J = P.forward()
            out.append((J, [x.cpu() for x in P.backward()]))
        finally:
            lib.awp_set_tuning(b"adj_jgroup", 8)
    assert out[1][0] == out[0][0]
    for a, b in zip(out[1][1], out[0][1]):
        assert torch.equal(a, b)
