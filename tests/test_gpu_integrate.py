"""Batched device-resident integrator (crs_rk4, integrate.py) on the GPU.

The workloads' synthetic mechanisms are extremely stiff (species time
scales down to 1e-57 s: random thermochemistry makes some reverse rates
enormous), so no explicit step of physical size stays finite on them;
mech9ns is generated non-stiff (Gibbs energies within a few kJ/mol,
activation energies of 22-28 kcal/mol: every eigenvalue of the Jacobian
below ~2e4 1/s at 1200-1500 K) to test the integration itself:

* on mech9ns, over 0.2 ms of physical time (temperatures move by up to
  ~100 K, trace species by orders of magnitude), the GPU's RK4 converges to
  an independent DOP853 solution of the oracle's dydt (scipy, rtol 1e-13)
  at fourth order (error ratio 16 per halving of the step), agrees with a
  host RK4 on the oracle's dydt to 1e-10, and conserves every element to
  1e-12;

and on the stiff mechanisms the integrator's own arithmetic, which is what
it adds to crs_dydt (pinned against the oracle elsewhere), at steps small
enough to stay finite:

* RK4 on the GPU is bit-identical to the textbook RK4 combination done on
  the host with the stage derivatives the GPU's crs_dydt returns for the
  host's own stage states (the energy-and-stage kernels are not
  contracted, so their derivatives are crs_dydt's bits), for 1, 2 and 3
  steps;
* the harness' captured HIP graph, replayed several times from the same
  initial state with eager work and synchronisations in between, gives the
  eager result every time;
* ignition bookkeeping (the reference's IGN block) flags a state the step
  its temperature first reaches T0 + dT_ign, and no other;
* the step count follows the reference's while (t < tend) loop (floating
  accumulation included).
"""

import numpy as np
import pytest
import torch

from create_rate_subs_amd import integrate, states

from _helpers import hip_lib, mechanism, oracle, to_dev, to_host

pytestmark = pytest.mark.gpu


# near-equilibrium states' fastest species move by up to ~1e-3 of their
# value per step at these sizes, and every stage stays finite
H_OF = {'mech13co': 1.0e-39, 'mech53': 1.0e-39, 'mech21pc': 1.0e-30}
H = H_OF['mech53']


def _state(name, n, seed):
    specs, _r, _f = mechanism(name)
    nsp = len(specs)
    T, P, Y = states.equilibrium_states(specs, n, seed, perturb=1e-3)
    y = np.empty((nsp, n))
    y[0] = T
    y[1:] = Y[:nsp - 1]
    return P, y


def _rk4_host(lib, Pd, y, h, steps):
    """Textbook RK4 on the host, stage derivatives from the GPU's dydt."""
    def f(v):
        return to_host(lib.dydt(Pd, to_dev(v)[0]))
    y = y.copy()
    for _ in range(steps):
        k1 = f(y)
        k2 = f(y + 0.5 * h * k1)
        k3 = f(y + 0.5 * h * k2)
        k4 = f(y + h * k3)
        y = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
    return y


@pytest.mark.parametrize('name', ['mech13co', 'mech53', 'mech21pc'])
def test_rk4_is_the_textbook_combination(name):
    P, y0 = _state(name, 256, 61)
    h, steps = H_OF[name], 3
    lib = hip_lib(name)
    Pd, yd = to_dev(P, y0)
    ref = _rk4_host(lib, Pd, y0, h, steps)
    lib.rk4(Pd, yd, 0.0, h, steps)
    got = to_host(yd)
    moved = (got != y0).mean()
    print(name, 'entries changed by the steps', moved)
    assert np.isfinite(got).all()
    assert moved > 0.01
    np.testing.assert_array_equal(got, ref)


@pytest.mark.parametrize('steps', [1, 2])
def test_short_runs_are_the_textbook_combination(steps):
    """One and two steps (the step count's edge cases)."""
    name = 'mech53'
    P, y0 = _state(name, 300, 64)
    lib = hip_lib(name)
    Pd, yd = to_dev(P, y0)
    ref = _rk4_host(lib, Pd, y0, H, steps)
    lib.rk4(Pd, yd, 0.0, H, steps)
    np.testing.assert_array_equal(to_host(yd), ref)


def test_graph_replays_match_eager():
    name = 'mech53'
    P, y0 = _state(name, 1000, 62)
    lib = hip_lib(name)
    Pd, y0d = to_dev(P, y0)
    h, steps = H, 3
    eager = y0d.clone()
    lib.rk4(Pd, eager, 0.0, h, steps)
    y = y0d.clone()
    work = torch.empty((3 * lib.nsp, 1000), dtype=torch.float64,
                       device='cuda')
    lib.rk4(Pd, y, 0.0, h, 1, work=work)      # set-up outside the capture
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        lib.rk4(Pd, y, 0.0, h, steps, work=work)
    for rep in range(3):
        y.copy_(y0d)
        other = lib.eval_wdot(Pd, Pd, torch.ones((lib.nsp, 1000),
                                                 dtype=torch.float64,
                                                 device='cuda'))
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, eager), rep
        del other


@pytest.mark.parametrize('t0', [0.0, -5.0 * H])
def test_ignition_time_is_the_first_step_reaching_threshold(t0):
    """The reference's IGN block (create_rate_subs.py:267-276) records t
    before its t += h: a state over the threshold after the first step gets
    t_ign = t0, and keeps it on later steps, negative t0 included."""
    name = 'mech53'
    P, y0 = _state(name, 64, 63)
    lib = hip_lib(name)
    Pd, yd = to_dev(P, y0)
    h = H
    # threshold already met for the first half, unreachable for the rest
    T0 = torch.tensor(np.where(np.arange(64) < 32, y0[0] - 1000.0,
                               y0[0] + 1.0e6), device='cuda')
    t_ign = torch.full((64,), float('nan'), dtype=torch.float64,
                       device='cuda')
    lib.rk4(Pd, yd, t0, h, 3, T0=T0, t_ign=t_ign, dT_ign=400.0)
    ti = to_host(t_ign)
    assert (ti[:32] == t0).all()
    assert np.isnan(ti[32:]).all()


def test_harness_runs_the_reference_loop_on_device():
    name = 'mech53'
    P, y0 = _state(name, 4096, 64)
    lib = hip_lib(name)
    Pd, yd = to_dev(P, y0)
    # the reference's defaults: t accumulates 1e-8 ten times to just below
    # 1e-7, so its loop runs 11 steps (its timer divides by (tend-t0)/h = 10)
    assert integrate.loop_steps(0.0, 1.0e-7, 1.0e-8) == 11
    h = H
    out = integrate.integrate(lib, Pd, yd, 0.0, 5.0 * H, h)
    steps = integrate.loop_steps(0.0, 5.0 * H, h)
    assert out['steps'] == steps
    ref = to_dev(y0)[0]
    lib.rk4(Pd, ref, 0.0, h, steps)
    assert torch.equal(yd, ref)
    print('steps', steps, 'ms/step', out['ms_per_step'])


def _nonstiff_states(n, seed):
    from create_rate_subs_amd import states as st
    specs, _r, _f = mechanism('mech9ns')
    nsp = len(specs)
    _T, _P, Y = st.random_states(nsp, n, seed)
    rng = np.random.default_rng(seed)
    y = np.empty((nsp, n))
    y[0] = rng.uniform(1200.0, 1500.0, n)
    y[1:] = Y[:nsp - 1]
    return np.full(n, 101325.0), y


def _element_totals(y):
    """Moles of each element per kg of mixture from the explicit mass
    fractions (the bath gas, last and inert, carries only N)."""
    specs, _r, _f = mechanism('mech9ns')
    elems = sorted({e for sp in specs[:-1] for e, _n in sp.elem})
    A = np.array([[dict(sp.elem).get(e, 0) / sp.mw for sp in specs[:-1]]
                  for e in elems])
    return A @ y[1:]


def test_rk4_converges_at_fourth_order_on_a_nonstiff_mechanism():
    from scipy.integrate import solve_ivp
    name, n, tend = 'mech9ns', 16, 2.0e-4
    P, y0 = _nonstiff_states(n, 71)
    orc = oracle(name)
    ref = np.empty_like(y0)
    for i in range(n):
        def f(_t, v, i=i):
            return orc.dydt(P[i:i + 1], v[:, None].copy())[:, 0]
        sol = solve_ivp(f, (0.0, tend), y0[:, i], method='DOP853',
                        rtol=1e-13, atol=1e-22)
        assert sol.success
        ref[:, i] = sol.y[:, -1]
    assert np.abs(ref[0] - y0[0]).max() > 10.0       # T moves
    lib = hip_lib(name)
    Pd = to_dev(P)[0]
    scale = np.maximum(np.abs(ref), 1e-8)
    errs = []
    for steps in (10, 20, 40):
        yd = to_dev(y0)[0]
        lib.rk4(Pd, yd, 0.0, tend / steps, steps)
        got = to_host(yd)
        assert np.isfinite(got).all()
        errs.append((np.abs(got - ref) / scale).max())
        # the same RK4 on the host with the oracle's dydt
        h, y = tend / steps, y0.copy()
        for _ in range(steps):
            k1 = orc.dydt(P, y)
            k2 = orc.dydt(P, y + 0.5 * h * k1)
            k3 = orc.dydt(P, y + 0.5 * h * k2)
            k4 = orc.dydt(P, y + h * k3)
            y = y + (h / 6.0) * (k1 + 2.0 * k2 + 2.0 * k3 + k4)
        host = (np.abs(got - y) / scale).max()
        e0, e1 = _element_totals(y0), _element_totals(got)
        cons = (np.abs(e1 - e0) / np.abs(e0).max(axis=0)).max()
        print('steps', steps, 'err vs DOP853 %.3e' % errs[-1],
              'vs host RK4 %.2e' % host, 'element drift %.2e' % cons)
        assert host <= 1e-10, host
        assert cons <= 1e-12, cons
    ratios = [errs[k] / errs[k + 1] for k in range(len(errs) - 1)]
    print('error ratios per halving', ratios)
    assert all(12.0 <= r <= 24.0 for r in ratios), ratios
    assert errs[-1] <= 2e-7, errs
