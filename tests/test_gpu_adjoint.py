"""The adjoint of the step (pg_adjoint_forward / pg_adjoint_backward).

Gates: gradients of L = sum(w * Sigma_final) with respect to the planet
mass, the viscosity alpha and the initial surface density match central
finite differences of the forward run with the same (frozen) step sizes;
FP64 within 1e-5 relative, the FP32 adjoint within 1e-3 of the FP64 one;
and the autograd wrapper returns the same numbers.  The configurations
switch the planet and the viscosity on and off so that a wrong pass shows
up in the case that isolates it.
"""
import numpy as np
import pytest

from penguin_amd import Disk, config_c5

pytestmark = pytest.mark.gpu


def _params(prec, nr=64, nphi=96, **kw):
    return config_c5(precision=prec, xres=nr, yres=nphi).replace(noise_rho=0.0, **kw)


def _grid(p):
    from penguin_amd.adjoint import radial_centres
    R = radial_centres(p)
    phi = (np.arange(int(p.yres)) + 0.5) * 2 * np.pi / int(p.yres)
    return np.meshgrid(R, phi)


def _smooth_state(d, p, sigma0=None):
    """Start state with resolved azimuthal structure (a two-armed spiral in
    Sigma, smooth radial and azimuthal velocity perturbations), so that the
    limiters sit away from their switching points and finite differences
    measure the derivative the adjoint computes."""
    R, PHI = _grid(p)
    g, _ = d.get_state()
    cs = p.sc_h / np.sqrt(R)
    g[0] = g[0] * (1.0 + 0.05 * np.cos(2 * PHI - 4 * np.log(R))) if sigma0 is None else sigma0
    g[2] = g[2] + 0.02 * cs * np.sin(PHI + 2 * np.log(R))
    g[3] = g[3] + 0.02 * cs * np.cos(3 * PHI - np.log(R))
    d.set_state(g)
    return g[0].copy()


def _weights(p, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.0, 1.0, (int(p.yres), int(p.xres)))


def _adjoint(p, nsteps, ckpt, w, sigma0=None):
    d = Disk(p)
    s0 = _smooth_state(d, p, sigma0)
    dts = d.adjoint_forward(nsteps, ckpt)
    g, _ = d.get_state()
    L = float(np.sum(w * g[0]))
    dL = np.zeros((5, d.ny, d.nx))
    dL[0] = w
    dq, dal, d0 = d.adjoint_backward(dL)
    d.close()
    return L, dts, dq, dal, d0, s0


def _frozen(p, sigma0, dts, w):
    d = Disk(p)
    _smooth_state(d, p, sigma0)
    t = 0.0
    for dt in dts:
        d.solve(t, float(dt))
        t += float(dt)
    g, _ = d.get_state()
    d.close()
    return float(np.sum(w * g[0]))


def _fd_checks(p, nsteps, ckpt, tol):
    w = _weights(p)
    L, dts, dq, dal, d0, s0 = _adjoint(p, nsteps, ckpt, w)
    assert abs(_frozen(p, s0, dts, w) - L) <= 1e-12 * abs(L) + 1e-14, "frozen-dt replay differs from the forward"
    out = {}
    if p.n_planet and p.planet_mass > 0:
        e = 1e-6 * p.planet_mass
        fd = (_frozen(p.replace(planet_mass=p.planet_mass + e), s0, dts, w) -
              _frozen(p.replace(planet_mass=p.planet_mass - e), s0, dts, w)) / (2 * e)
        out["q"] = (dq, fd)
    if p.visc_flag:
        e = 1e-6 * p.ss_alpha
        fd = (_frozen(p.replace(ss_alpha=p.ss_alpha + e), s0, dts, w) -
              _frozen(p.replace(ss_alpha=p.ss_alpha - e), s0, dts, w)) / (2 * e)
        out["alpha"] = (dal, fd)
    R, PHI = _grid(p)
    delta = s0 * (0.5 + 0.5 * np.sin(PHI + 3 * R))
    e = 1e-6
    fd = (_frozen(p, s0 + e * delta, dts, w) - _frozen(p, s0 - e * delta, dts, w)) / (2 * e)
    out["sigma0"] = (float(np.sum(d0[0] * delta)), fd)
    for k, (a, f) in out.items():
        rel = abs(a - f) / max(abs(f), 1e-300)
        print(f"{k}: adjoint {a:.10e} fd {f:.10e} rel {rel:.2e}", flush=True)
        assert rel <= tol, (k, a, f, rel)
    return out


@pytest.mark.parametrize("case", ["sweeps", "planet", "viscosity", "full"])
def test_adjoint_one_step_fp64(case):
    kw = {"sweeps": dict(n_planet=0, visc_flag=0), "planet": dict(visc_flag=0),
          "viscosity": dict(n_planet=0), "full": {}}[case]
    _fd_checks(_params(8, **kw), 1, 1, 1e-5)


def test_adjoint_segments_fp64():
    # 7 steps in checkpoint segments of 3: recompute of full and partial segments
    _fd_checks(_params(8), 7, 3, 1e-5)


def test_adjoint_fp32_matches_fp64():
    p8, p4 = _params(8), _params(4)
    w = _weights(p8)
    _, _, dq8, da8, d08, s0 = _adjoint(p8, 10, 4, w)
    _, _, dq4, da4, d04, _ = _adjoint(p4, 10, 4, w, sigma0=s0)
    rq = abs(dq4 - dq8) / abs(dq8)
    ra = abs(da4 - da8) / abs(da8)
    rs = float(np.linalg.norm(d04[0] - d08[0]) / np.linalg.norm(d08[0]))
    print(f"fp32 vs fp64: q {rq:.2e} alpha {ra:.2e} sigma0 {rs:.2e}", flush=True)
    assert rq <= 1e-3 and ra <= 1e-3 and rs <= 1e-3


def test_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from penguin_amd.adjoint import evolve

    p = _params(8)
    w = _weights(p)
    L, dts, dq, dal, d0, s0 = _adjoint(p, 5, 2, w)
    sig = torch.tensor(s0, requires_grad=True)
    q = torch.tensor(p.planet_mass, dtype=torch.float64, requires_grad=True)
    al = torch.tensor(p.ss_alpha, dtype=torch.float64, requires_grad=True)
    out = evolve(sig, q, al, p, 5, 2)
    loss = torch.sum(torch.from_numpy(w) * out)
    loss.backward()
    assert abs(float(loss) - L) <= 1e-12 * abs(L)
    assert abs(float(q.grad) - dq) <= 1e-12 * abs(dq)
    assert abs(float(al.grad) - dal) <= 1e-12 * abs(dal)
    assert np.allclose(sig.grad.numpy(), d0[0], rtol=0, atol=1e-14 * np.abs(d0[0]).max())
