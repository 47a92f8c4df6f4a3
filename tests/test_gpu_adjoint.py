"""The adjoint of the step (pg_adjoint_forward / pg_adjoint_backward).

Gates: gradients of L = sum(w * Sigma_final) with respect to the planet
mass, the viscosity alpha and the initial surface density match central
finite differences of the forward run with the same (frozen) step sizes;
FP64 within 1e-5 relative, the FP32 adjoint within 1e-3 of the FP64 one;
and the autograd wrapper returns the same numbers.  The configurations
switch the planet and the viscosity on and off so that a wrong pass shows
up in the case that isolates it.  The state is smooth and resolved and
p_alpha = 0.5 (distinct radial ghost velocities), so that no limiter or
flattening switch sits within a finite-difference step of its threshold.
"""
import numpy as np
import pytest

from penguin_amd import Disk, config_c5

pytestmark = pytest.mark.gpu


def _params(prec, nr=64, nphi=96, **kw):
    return config_c5(precision=prec, xres=nr, yres=nphi).replace(noise_rho=0.0, p_alpha=0.5, **kw)


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


def _adjoint(p, nsteps, ckpt, w, sigma0=None, smooth=True):
    d = Disk(p)
    if smooth:
        s0 = _smooth_state(d, p, sigma0)
    else:   # only Sigma replaced, as DiskEvolve does
        g, _ = d.get_state()
        g[0] = sigma0
        d.set_state(g)
        s0 = g[0].copy()
    dts = d.adjoint_forward(nsteps, ckpt)
    g, _ = d.get_state()
    L = float(np.sum(w * g[0]))
    dL = np.zeros((5, d.ny, d.nx))
    dL[0] = w
    dq, dal, d0 = d.adjoint_backward(dL)
    d.close()
    return L, dts, dq, dal, d0, s0


def _smooth_weights(p):
    """Weights of a smooth loss (a function of radius, like the gradient of
    the gap-profile loss)."""
    R, PHI = _grid(p)
    return np.sin(3 * np.log(R)) / int(p.yres)


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


def _fd_checks(p, nsteps, ckpt, tol, w=None, p_fd=None):
    """Adjoint of p against central differences of frozen-dt forward runs of
    p_fd (default p; an FP64 twin to check an FP32 adjoint)."""
    w = _weights(p) if w is None else w
    pf = p if p_fd is None else p_fd
    L, dts, dq, dal, d0, s0 = _adjoint(p, nsteps, ckpt, w)
    if p_fd is None:
        assert abs(_frozen(p, s0, dts, w) - L) <= 1e-12 * abs(L) + 1e-14, "frozen-dt replay differs from the forward"
    out = {}
    if p.n_planet and p.planet_mass > 0:
        e = 1e-4 * p.planet_mass
        fd = (_frozen(pf.replace(planet_mass=p.planet_mass + e), s0, dts, w) -
              _frozen(pf.replace(planet_mass=p.planet_mass - e), s0, dts, w)) / (2 * e)
        out["q"] = (dq, fd)
    if p.visc_flag:
        e = 1e-4 * p.ss_alpha
        fd = (_frozen(pf.replace(ss_alpha=p.ss_alpha + e), s0, dts, w) -
              _frozen(pf.replace(ss_alpha=p.ss_alpha - e), s0, dts, w)) / (2 * e)
        out["alpha"] = (dal, fd)
    R, PHI = _grid(p)
    delta = s0 * (0.5 + 0.5 * np.sin(PHI + 3 * R))
    e = 1e-6
    fd = (_frozen(pf, s0 + e * delta, dts, w) - _frozen(pf, s0 - e * delta, dts, w)) / (2 * e)
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


def test_adjoint_fp32_against_fp64_fd():
    # the FP32 adjoint (forward, checkpoints and backward in FP32) against
    # FP64 central differences with the FP32 run's step sizes, for a smooth
    # loss; the bar is 1e-3 relative
    p4 = _params(4)
    _fd_checks(p4, 20, 5, 1e-3, w=_smooth_weights(p4), p_fd=_params(8))


def test_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from penguin_amd.adjoint import evolve

    p = _params(8)
    w = _weights(p)
    d = Disk(p)
    s_in = _smooth_state(d, p)
    d.close()
    L, dts, dq, dal, d0, s0 = _adjoint(p, 5, 2, w, sigma0=s_in, smooth=False)
    sig = torch.tensor(s0, requires_grad=True)
    q = torch.tensor(p.planet_mass, dtype=torch.float64, requires_grad=True)
    al = torch.tensor(p.ss_alpha, dtype=torch.float64, requires_grad=True)
    out = evolve(sig, q, al, p, 5, 2)
    loss = torch.sum(torch.from_numpy(w) * out)
    loss.backward()
    assert abs(float(loss.detach()) - L) <= 1e-12 * abs(L)
    assert abs(float(q.grad) - dq) <= 1e-12 * abs(dq)
    assert abs(float(al.grad) - dal) <= 1e-12 * abs(dal)
    assert np.allclose(sig.grad.numpy(), d0[0], rtol=0, atol=1e-14 * np.abs(d0[0]).max())


@pytest.mark.parametrize("prec", [8, 4])
@pytest.mark.parametrize("visc", [1, 0])
def test_adjoint_tape_equals_checkpoints(prec, visc):
    """ckpt_every 0 (the forward keeps every step's intermediates, the
    backward recomputes nothing): the device loop's step sizes and final
    state bit for bit (the same kernels on the same inputs, the step chosen
    on the device), and the checkpointed sweep's gradients (whose segments
    restart the viscous forces and the CFL rate from their checkpoints: equal
    to rounding)."""
    p = _params(prec, visc_flag=visc)
    w = _weights(p)
    d = Disk(p)
    _smooth_state(d, p)
    dts_run = d.run(7, use_graph=True)
    g_run, _ = d.get_state()
    d.close()
    L0, dts0, dq0, dal0, d00, _ = _adjoint(p, 7, 3, w)
    d = Disk(p)
    _smooth_state(d, p)
    dts1 = d.adjoint_forward(7, 0)
    g1, _ = d.get_state()
    assert np.array_equal(dts1, dts_run)
    assert np.array_equal(g1, g_run)
    dL = np.zeros((5, d.ny, d.nx))
    dL[0] = w
    dq1, dal1, d01 = d.adjoint_backward(dL)
    g2, _ = d.get_state()
    assert np.array_equal(g2, g_run), "the backward sweep must leave the forward's final state"
    dq2, dal2, d02 = d.adjoint_backward(dL)   # a second sweep over the same tape
    d.run(1, use_graph=True)                  # the device loop continues from the tape's end
    d.close()
    assert dq2 == dq1 and dal2 == dal1 and np.array_equal(d02, d01)
    tol = 1e-10 if prec == 8 else 1e-4
    print("tape == checkpoints bitwise:", dq1 == dq0 and dal1 == dal0 and np.array_equal(d01, d00), flush=True)
    assert np.allclose(dts1, dts0, rtol=tol, atol=0)
    if p.n_planet:
        assert abs(dq1 - dq0) <= tol * abs(dq0)
    if p.visc_flag:
        assert abs(dal1 - dal0) <= tol * abs(dal0)
    assert np.abs(d01 - d00).max() <= tol * np.abs(d00).max()


def test_adjoint_tape_fd_fp64():
    _fd_checks(_params(8), 7, 0, 1e-5)


def test_adjoint_tape_refuses_the_time_average():
    from penguin_amd.disk import PenguinError
    p = _params(8)
    d = Disk(p)
    d.average_begin(1.0)
    with pytest.raises(PenguinError):
        d.adjoint_forward(3, 0)
    d.average_end()
    d.close()


def test_adjoint_error_paths():
    """Refused without touching the state: dust or wave killing, a negative
    checkpoint interval, a backward sweep before any forward run."""
    from penguin_amd import config_c2
    from penguin_amd.disk import PenguinError
    d = Disk(config_c2(precision=4, xres=64, yres=96))   # gas + dust, wave killing
    with pytest.raises(PenguinError):
        d.adjoint_forward(2, 1)
    d.close()
    p = _params(8)
    d = Disk(p)
    with pytest.raises(PenguinError):
        d.adjoint_backward(np.zeros((5, d.ny, d.nx)))
    with pytest.raises(PenguinError):
        d.adjoint_forward(2, -1)
    g0, _ = d.get_state()
    dts = d.adjoint_forward(2, 0)   # still usable after the refusals
    assert dts.shape == (2,) and np.all(dts > 0)
    d.close()
