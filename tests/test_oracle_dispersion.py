"""Plane waves through the CPU oracle against the staggered-grid dispersion relation.

In a homogeneous elastic medium with the CPML switched off, a wave that varies
along one axis only splits into independent (stress, velocity) pairs: the P
pair (v along the axis and its normal stress, modulus M = lambda + 2 mu) and
two S pairs (a transverse v and the shear stress coupling it, modulus mu).
Each pair is the 1-D staggered leapfrog

    sigma^{n+1} = sigma^n + dt M D v^n,   v^{n+1} = v^n + dt / rho D' sigma^{n+1},

where D, D' are the reference's D+ / D- stencils (sum_c C_c (f[i+c+1] - f[i-c]) / d
and sum_c C_c (f[i+c] - f[i-c-1]) / d).  For a Fourier mode e^{ikj} both
stencils have the symbol i K(k), K(k) = (2/d) sum_m C_m sin((m + 1/2) k), once
sigma sits half a cell after v (sigma from D+ v) or before it (sigma from
D- v).  One step multiplies (S, V) by G = [[1, i a], [i b, 1 - a b]] with
a = dt M K, b = dt K / rho; det G = 1, so G^N = (sin(N th) G - sin((N-1) th) I)
/ sin(th), with the staggered-grid dispersion relation sin(th/2) = c dt K / 2
(c^2 = M / rho) giving the phase th = omega dt per step.  The other normal
stresses follow the P pair as slaves: sigma_yy^N = sigma_yy^0 +
lambda / M (sigma_xx^N - sigma_xx^0) for a wave along x.

The tests seed each pair with a smooth, localised packet along the wave
axis, uniform along the others, run the FP64 oracle N steps, and compare
with the closed form evaluated by FFT: 1e-10 relative, on the rows farther
from the grid edges than the stencils can carry the edges' own treatment in
N steps (2 * ord cells per step).  This pins, over many steps, every
stencil's coefficient order, its D+/D- direction and axis, the spacing, dt,
the moduli each stress uses and the leapfrog order (stress first, then
velocity: wgsl2d :829-1072, wgsl3d :1340-1773, sim_taichi_elastic.py:138-309),
for every FD order; planted mistakes in the closed form (coefficients
reversed, a D+/D- swap, lambda for M, a halved dt) miss the oracle by far more.
"""
import numpy as np
import pytest

import oracle
from taichi_sim_amd.sim_support.simulator import fd_coefficients
from test_oracle_closed_form import Setup, _profiles

TOL = 1e-10
RHO, CP, CS = 8.0, 6.0, 3.5
D = 0.125                       # spacing; 1/d exact in float32 (the Taichi variant scales by f32(1/d))
DT = 2.0 ** -7                  # cp dt / d = 0.375


def _symbol(C, k, d):
    m = np.arange(len(C))[:, None]
    return 2.0 / d * np.sum(C[:, None] * np.sin((m + 0.5) * k[None, :]), axis=0)


def _propagate(sig0, v0, s, M, n_steps, C, d=D, dt=DT, rho=RHO):
    """Closed-form (sigma, v) after n_steps of one pair; s = +1 when sigma is
    half a cell after v (sigma from D+ v), -1 when before (sigma from D- v)."""
    n = sig0.size
    k = 2.0 * np.pi * np.fft.fftfreq(n)
    K = _symbol(np.asarray(C, np.float64), k, d)
    S = np.fft.fft(sig0) * np.exp(-0.5j * s * k)
    V = np.fft.fft(v0).astype(complex)
    a, b = dt * M * K, dt * K / rho
    c = np.sqrt(M / rho)
    th = 2.0 * np.arcsin(c * dt * K / 2.0)            # dispersion relation
    sth = np.sin(th)
    small = np.abs(sth) < 1e-300
    safe = np.where(small, 1.0, sth)
    u1 = np.where(small, n_steps, np.sin(n_steps * th) / safe)
    u2 = np.where(small, n_steps - 1, np.sin((n_steps - 1) * th) / safe)
    SN = u1 * (S + 1j * a * V) - u2 * S
    VN = u1 * (1j * b * S + (1.0 - a * b) * V) - u2 * V
    return np.fft.ifft(SN * np.exp(0.5j * s * k)).real, np.fft.ifft(VN).real


def _packet(n, centre, width, k0, phase):
    j = np.arange(n, dtype=np.float64)
    return np.exp(-((j - centre) / width) ** 2) * np.cos(k0 * (j - centre) + phase)


def _moduli():
    lam = np.float64(RHO) * (np.float64(CP) ** 2 - 2.0 * np.float64(CS) ** 2)
    mu = np.float64(RHO) * np.float64(CS) ** 2
    return lam, mu, lam + 2.0 * mu


# pairs per wave axis: (stress, velocity, s, modulus) -- s from the stencil
# the stress takes of that velocity along the axis
PAIRS2D = {
    0: [("sxx", "vx", +1, "P"), ("sxy", "vy", -1, "S")],
    1: [("syy", "vy", -1, "P"), ("sxy", "vx", +1, "S")],
}
SLAVES2D = {0: [("syy", "sxx")], 1: [("sxx", "syy")]}
PAIRS3D = {
    0: [("sxx", "vx", +1, "P"), ("sxy", "vy", -1, "S"), ("sxz", "vz", -1, "S")],
    1: [("syy", "vy", -1, "P"), ("sxy", "vx", +1, "S"), ("syz", "vz", +1, "S")],
    2: [("szz", "vz", -1, "P"), ("sxz", "vx", +1, "S"), ("syz", "vy", +1, "S")],
}
SLAVES3D = {0: [("syy", "sxx"), ("szz", "sxx")], 1: [("sxx", "syy"), ("szz", "syy")],
            2: [("sxx", "szz"), ("syy", "szz")]}


def _run(ndim, axis, acc, n_steps, n_wave, variant="wgsl"):
    """Oracle run of a wave along `axis`; returns (oracle, initial 1-D packets,
    index of the compared line)."""
    ord_ = len(fd_coefficients(acc))
    n_flat = 2 * (2 * ord_ * n_steps) + 8
    shape = [n_flat] * ndim
    shape[axis] = n_wave
    rng = np.random.default_rng(7)
    mat = np.ones(shape, np.float32)
    sim = Setup(tuple(shape), (D,) * ndim, DT, acc, mat * RHO, mat * CP, mat * CS,
                _profiles(shape, acc, rng, inert=True), visco=False, n_steps=n_steps)
    o = oracle.Oracle(sim, variant=variant, precision="f64", visco=False)
    pairs = (PAIRS2D if ndim == 2 else PAIRS3D)[axis]
    slaves = (SLAVES2D if ndim == 2 else SLAVES3D)[axis]
    init = {}
    centre = n_wave / 2.0
    for i, (sig, vel, _, _) in enumerate(pairs):
        init[sig] = _packet(n_wave, centre + 3 * i, 5.0, 2 * np.pi / 11, 0.4 + i)
        init[vel] = 0.7 * _packet(n_wave, centre - 2 * i, 4.0, 2 * np.pi / 9, 1.1 * i)
    for sl, _ in slaves:
        init[sl] = 0.3 * _packet(n_wave, centre + 1, 6.0, 2 * np.pi / 13, 0.2)
    bshape = [1] * ndim
    bshape[axis] = n_wave
    for name, line in init.items():
        o.f[name][...] = np.broadcast_to(line.reshape(bshape), shape)
    o.run(n_steps)
    mid = [n_flat // 2] * ndim
    return o, init, mid, pairs, slaves


def _line(field, mid, axis):
    idx = list(mid)
    idx[axis] = slice(None)
    return field[tuple(idx)]


def _check(o, init, mid, axis, pairs, slaves, n_steps, C):
    lam, mu, lp2m = _moduli()
    M = {"P": lp2m, "S": mu}
    for sig, vel, s, kind in pairs:
        want_s, want_v = _propagate(init[sig], init[vel], s, M[kind], n_steps, C)
        got_s, got_v = _line(o.f[sig], mid, axis), _line(o.f[vel], mid, axis)
        assert np.abs(want_s).max() > 0.1 and np.abs(got_s - init[sig]).max() > 1e-3
        assert oracle.rel_l2(got_s, want_s) < TOL, (sig, vel, oracle.rel_l2(got_s, want_s))
        assert oracle.rel_l2(got_v, want_v) < TOL, (sig, vel, oracle.rel_l2(got_v, want_v))
    for sl, master in slaves:
        want = init[sl] + lam / lp2m * (_line(o.f[master], mid, axis) - init[master])
        assert oracle.rel_l2(_line(o.f[sl], mid, axis), want) < TOL, sl


@pytest.mark.parametrize("variant", ["wgsl", "taichi"])
@pytest.mark.parametrize("acc", [2, 3, 4, 5, 6])
@pytest.mark.parametrize("axis", [0, 1])
def test_2d_plane_waves_follow_the_dispersion_relation(variant, acc, axis):
    n_steps = 24
    o, init, mid, pairs, slaves = _run(2, axis, acc, n_steps, 192, variant)
    _check(o, init, mid, axis, pairs, slaves, n_steps, fd_coefficients(acc))


@pytest.mark.parametrize("acc", [2, 4])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_3d_plane_waves_follow_the_dispersion_relation(acc, axis):
    n_steps = 12 if acc == 2 else 8
    o, init, mid, pairs, slaves = _run(3, axis, acc, n_steps, 128)
    _check(o, init, mid, axis, pairs, slaves, n_steps, fd_coefficients(acc))


def test_planted_mistakes_in_the_closed_form_are_caught():
    n_steps, acc, axis = 24, 4, 0
    o, init, mid, pairs, slaves = _run(2, axis, acc, n_steps, 192)
    C = fd_coefficients(acc)
    lam, mu, lp2m = _moduli()
    sig, vel, s, _ = pairs[0]
    got = _line(o.f[sig], mid, axis)
    assert oracle.rel_l2(got, _propagate(init[sig], init[vel], s, lp2m, n_steps, C)[0]) < TOL
    wrong = {
        "coefficients reversed": _propagate(init[sig], init[vel], s, lp2m, n_steps, C[::-1]),
        "D+/D- swap": _propagate(init[sig], init[vel], -s, lp2m, n_steps, C),
        "lambda for lambda + 2 mu": _propagate(init[sig], init[vel], s, lam, n_steps, C),
        "dt / 2": _propagate(init[sig], init[vel], s, lp2m, n_steps, C, dt=DT / 2),
        "one step short": _propagate(init[sig], init[vel], s, lp2m, n_steps - 1, C),
    }
    for name, (ws, _) in wrong.items():
        assert oracle.rel_l2(got, ws) > 1e-4, name


def test_dispersion_relation_reduces_to_the_continuum_for_long_waves():
    # sin(th/2) = c dt K / 2 with K -> k / d as k -> 0: the phase velocity of
    # long waves tends to c, and the FD coefficients' Taylor condition
    # sum_m C_m (2m + 1) = 1 is what makes it so.
    for acc in range(2, 7):
        C = np.asarray(fd_coefficients(acc), np.float64)
        k = np.array([1e-3])
        K = _symbol(C, k, D)[0]
        assert abs(K * D / k[0] - 1.0) < 1e-6, acc
        th = 2.0 * np.arcsin(CP * DT * K / 2.0)
        assert abs(th / DT / (k[0] / D) / CP - 1.0) < 1e-6
