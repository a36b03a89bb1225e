"""GPU: the 24-step coupled rollout (CG2D + momentum correction + DST3, checkpoint every 6 steps): reverse-mode
gradients of sum(rA drF hFacC T_final) with respect to T0, the rhs b and the initial velocities u0, v0 against
central differences (best relative error of an h sweep below 1e-6), and the checkpointed gradients equal to the
stored-everything gradients bitwise (the recomputed forward is deterministic)."""

import numpy as np
import pytest

from tests.conftest import to_dev

pytestmark = pytest.mark.gpu


def _inputs(dev):
    from mit_jax_amd import rollout, synthetic as S
    case = rollout.RolloutCase.synthetic(dev)
    g = S.make_grid(0)
    u, v, _ = S.flow(g, 2)
    tr = S.tracer(g, 3)
    b = S.rhs(g, 1) * 1e-3  # a modest surface-pressure forcing: the corrected velocities stay within CFL 0.3
    return case, g, (to_dev(tr, dev), to_dev(u, dev), to_dev(v, dev), to_dev(b, dev))


def _grads(case, X, nsteps, every):
    import torch
    from mit_jax_amd import rollout
    leaves = [x.clone().requires_grad_(True) for x in X]
    T, _, _ = rollout.rollout(case, *leaves, nsteps=nsteps, checkpoint_every=every)
    J = rollout.loss(case, T)
    return float(J.detach()), torch.autograd.grad(J, leaves)


def test_checkpointed_equals_stored(dev):
    case, g, X = _inputs(dev)
    J1, g1 = _grads(case, X, 8, 4)
    J2, g2 = _grads(case, X, 8, 0)
    assert J1 == J2
    for a, b in zip(g1, g2):
        assert np.array_equal(a.cpu().numpy(), b.cpu().numpy())


def test_rollout_gradients_vs_finite_differences(dev):
    import torch
    from mit_jax_amd import rollout
    case, g, X = _inputs(dev)
    J0, grads = _grads(case, X, 24, 6)
    print(f"J = {J0:.15e}", flush=True)
    for a in grads:
        assert bool(torch.isfinite(a).all())
    rng = np.random.default_rng(21)
    L = g.L

    def T_final(Y):
        with torch.no_grad():
            return rollout.rollout(case, *Y, nsteps=24, checkpoint_every=0)[0]

    # a random direction on a wet window of the interior (tile 4, levels 0-2, 3x3 points)
    t, j, i = 4, L.jj(41), L.ii(41)
    for k, name, rhs in ((0, "T0", (1e-1, 1e-2, 1e-3)), (1, "u0", (1e-2, 1e-3, 1e-4)), (2, "v0", (1e-2, 1e-3, 1e-4)),
                         (3, "b", (1e-1, 1e-2, 1e-3))):
        d = np.zeros(tuple(X[k].shape))
        if k == 3:
            d[t, j:j + 3, i:i + 3] = rng.normal(size=(3, 3)) * g.maskC[t, 0, j:j + 3, i:i + 3]
        else:
            d[t, 0:3, j:j + 3, i:i + 3] = rng.normal(size=(3, 3, 3)) * g.maskC[t, 0:3, j:j + 3, i:i + 3]
        scale = float(X[k][tuple(slice(None) for _ in range(X[k].dim()))].abs().max())
        dd = to_dev(d, dev)
        ad = float((grads[k] * dd).sum())
        errs = []
        for r in rhs:
            h = r * scale
            Yp = [x + h * dd if n == k else x for n, x in enumerate(X)]
            Ym = [x - h * dd if n == k else x for n, x in enumerate(X)]
            # difference the final states before the weighted sum: J itself (~1e19) would drown the FD in rounding
            fd = float(rollout.loss(case, T_final(Yp) - T_final(Ym))) / (2 * h)
            errs.append(abs(fd - ad) / abs(ad))
            print(f"{name} h={h:.2e}: ad {ad:.12e} fd {fd:.12e} rel {errs[-1]:.2e}", flush=True)
        assert min(errs) < 1e-6, (name, errs)
