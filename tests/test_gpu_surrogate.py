"""Surrogate kernels and pipeline on the GPU.

Every test runs on the 13-species network and on the 53-species one of the
learned-surrogate workload (55 -> 256 x 3 -> 53, GELU, on the 53-species /
325-reaction mechanism):

* the BF16-MFMA forward product and the full backward pass against the
  NumPy reference (oracle/mlp_ref.py) with the same BF16 roundings;
* training by our kernels lowers the loss;
* gated evaluation: every state the gate sends back is bit-identical to the
  exact FP64 kernel, out-of-box and invalid states are always sent back;
* checkpoint round trip, and refusal of another mechanism's checkpoint.
"""

import numpy as np
import pytest
import torch

from create_rate_subs_amd import states
from create_rate_subs_amd.surrogate import Surrogate, load_mlp, _p, _s
from oracle import mlp_ref

from _helpers import hip_lib, mechanism

pytestmark = pytest.mark.gpu

MECHS = ['mech13', 'mech53']


def _net(name='mech13', seed=3):
    specs, reacs, _ = mechanism(name)
    return Surrogate(hip_lib(name), specs, reacs, hidden=256, seed=seed)


def _host_params(net):
    W, b = [], []
    for li in range(len(net.shapes)):
        W.append(net.w(li).float().cpu().numpy().astype(np.float64))
        b.append(net.b(li).cpu().numpy().astype(np.float64))
    return W, b


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('M', [1, 100, 300])
def test_forward_matches_reference(M, name):
    net = _net(name)
    assert net.nin == net.nsp + 2 and net.nout == net.nsp
    rng = np.random.default_rng(M)
    X = mlp_ref.bf16(rng.normal(size=(M, net.kin)))
    X[:, net.nin:] = 0.0
    Xd = torch.tensor(X, dtype=torch.bfloat16, device='cuda')
    # perturb biases so they matter
    net.p.add_(torch.randn_like(net.p) * 0.01)
    lib = load_mlp()
    lib('crs_mlp_to_bf16', _p(net.p), _p(net.pb), net.nparam, _s())
    Yh = net.forward(Xd).cpu().numpy()[:, :net.nout].astype(np.float64)
    W, b = _host_params(net)
    Yt = np.zeros((M, net.kout))
    ref, _dW, _db = mlp_ref.forward_backward(X, W, b, Yt, net.nout)
    err = np.abs(Yh - ref[:, :net.nout]).max() / np.abs(ref).max()
    assert err < 5e-3, err


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('M', [64, 300])
def test_gradients_match_cpu_oracle(M, name):
    net = _net(name, seed=5)
    rng = np.random.default_rng(7 + M)
    X = mlp_ref.bf16(rng.normal(size=(M, net.kin)))
    X[:, net.nin:] = 0.0
    Ytgt = rng.normal(size=(M, net.nout)).astype(np.float32)
    Xd = torch.tensor(X, dtype=torch.bfloat16, device='cuda')
    Yd = torch.tensor(Ytgt, device='cuda')
    Yh = net.forward(Xd, save=True)
    net.backward(Yh, Yd)
    torch.cuda.synchronize()
    W, b = _host_params(net)
    Yt = np.zeros((M, net.kout))
    Yt[:, :net.nout] = Ytgt
    _y, dW, db = mlp_ref.forward_backward(X, W, b, Yt, net.nout)
    for li in range(len(net.shapes)):
        n_out, n_in = net.shapes[li]
        gw = net.gw(li).view(n_out, n_in).cpu().numpy().astype(np.float64)
        gb = net.gb(li).cpu().numpy().astype(np.float64)
        ew = np.linalg.norm(gw - dW[li]) / max(np.linalg.norm(dW[li]), 1e-30)
        eb = np.linalg.norm(gb - db[li]) / max(np.linalg.norm(db[li]), 1e-30)
        print('layer', li, 'rel dW', ew, 'rel db', eb)
        assert ew < 2e-2 and eb < 2e-2, (li, ew, eb)


def _labelled(name, n, seed):
    specs, _r, _f = mechanism(name)
    T, P, Y = states.random_states(len(specs), n, seed)
    Td, Pd, Yd = (torch.from_numpy(a).cuda() for a in (T, P, Y))
    W = hip_lib(name).eval_wdot(Td, Pd, Yd)
    return Td, Pd, Yd, W


@pytest.mark.parametrize('name', MECHS)
def test_captured_training_step_matches_eager(name):
    """The training step's forward and backward pass replayed from a
    captured HIP graph, with new batch indices each time and eager work
    between the replays: the gradients equal an eager pass over the same
    batch up to the order of the FP32 atomic sums of the split reduction."""
    net = _net(name, seed=31)
    Td, Pd, Yd, W = _labelled(name, 1 << 14, 83)
    net.fit_normalisation(Td, Pd, Yd, W)
    X = torch.empty((Td.numel(), net.kin), dtype=torch.bfloat16,
                    device='cuda')
    net.features(Td, Pd, Yd, out=X)
    Yt = net.targets(W)
    sel = torch.zeros(4096, dtype=torch.int64, device='cuda')
    loss = torch.zeros(1, dtype=torch.float32, device='cuda')
    net.grad_step(X, Yt, sel, loss)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        net.grad_step(X, Yt, sel, loss)
    gen = torch.Generator(device='cuda')
    gen.manual_seed(5)
    for rep in range(3):
        idx = torch.randperm(Td.numel(), device='cuda', generator=gen)[:4096]
        sel.copy_(idx)
        loss.zero_()
        g.replay()
        torch.cuda.synchronize()
        g_graph, l_graph = net.g.clone(), loss.item()
        other = hip_lib(name).eval_wdot(Td, Pd, Yd)      # unrelated work
        loss.zero_()
        net.grad_step(X, Yt, idx, loss)
        torch.cuda.synchronize()
        scale = net.g.abs().max().item()
        diff = (g_graph - net.g).abs().max().item()
        print(rep, 'max |dg|', diff, 'of', scale)
        assert diff <= 1e-5 * scale
        assert abs(l_graph - loss.item()) <= 1e-5 * abs(loss.item())
        del other


def test_training_lowers_loss():
    net = _net(seed=11)
    Td, Pd, Yd, W = _labelled('mech13', 1 << 15, 50)
    hist = net.train(Td, Pd, Yd, W, epochs=6, batch=4096, lr=2e-3)
    print('mse by epoch', hist)
    assert hist[-1] < 0.25 * hist[0]


@pytest.mark.parametrize('name', MECHS)
def test_gate_fallback_is_exact_and_catches_invalid_states(name):
    net = _net(name, seed=13)
    Td, Pd, Yd, W = _labelled(name, 1 << 14, 51)
    net.train(Td, Pd, Yd, W, epochs=2, batch=4096)
    net.gate.rmax = 1e-3
    T2, P2, Y2, Wex = _labelled(name, 5000, 52)
    T2[7] = 5000.0                    # outside the training box
    Y2[0, 9] = float('nan')           # invalid composition
    Y2[1, 11] = -0.1                  # invalid composition
    lib = hip_lib(name)
    Wex = lib.eval_wdot(T2, P2, Y2)
    Wout = torch.empty_like(Wex)
    m = net.evaluate(T2, P2, Y2, Wout)
    flags = net._bufs['flags'][:5000].cpu().numpy()
    assert flags[7] and flags[9] and flags[11]
    assert m == int(flags.sum())
    fb = np.nonzero(flags)[0]
    a = Wout.cpu().numpy()[:, fb]
    b = Wex.cpu().numpy()[:, fb]
    assert np.array_equal(a, b, equal_nan=True)
    audit = net.audit(T2, P2, Y2, Wout, frac=0.05)
    print('fallback', m, 'audit', audit)
    assert audit['states'] == 250


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('mode', ['fp64', 'mixed'])
def test_gated_c_abi_matches_its_parts(name, mode):
    """crs_mlp_eval_gated (inference, gate and the mechanism library's
    crs_eval_wdot_subset_dev on one stream, the count never read on the
    host): states it sends back carry the exact kernel's rates bit for bit,
    the others the fused surrogate's, and the device count is the number of
    flags.  Also replayed from a captured HIP graph, with other work between
    replays, against the eager result."""
    net = _net(name, seed=23)
    Td, Pd, Yd, W = _labelled(name, 1 << 14, 81)
    net.train(Td, Pd, Yd, W, epochs=1, batch=4096)
    T2, P2, Y2, _w = _labelled(name, 3000, 82)
    T2[5] = 9000.0                          # outside the training box
    # admit about half of the states: rmax at the median residual
    Wsur = torch.empty_like(_w)
    net.predict(T2, P2, Y2, Wsur)
    net.gate.rmax = float(net.residuals(Wsur).median().item())
    Wex = hip_lib(name).eval_wdot(T2, P2, Y2, mode)
    out = torch.full_like(Wex, float('nan'))
    cnt = net.evaluate(T2, P2, Y2, out, mode=mode, read_count=False)
    flags = net._bufs['flags'][:3000].bool()
    m = int(cnt.item())
    print(name, mode, 'fallback', m, 'of 3000')
    assert m == int(flags.sum().item()) and 0 < m < 3000
    assert bool(flags[5])
    assert torch.equal(out[:, flags], Wex[:, flags])
    assert torch.equal(out[:, ~flags], Wsur[:, ~flags])
    # graph capture of the same call: no host read, no memset node
    g_out = torch.full_like(Wex, float('nan'))
    net.evaluate(T2, P2, Y2, g_out, mode=mode, read_count=False)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        net.evaluate(T2, P2, Y2, g_out, mode=mode, read_count=False)
    for rep in range(3):
        g_out.fill_(float('nan'))
        other = hip_lib(name).eval_wdot(T2, P2, Y2, 'fp64')
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(g_out, out), rep
        assert int(net._bufs['fb_count'][0].item()) == m
        del other


def _fit_random(name, n, seed):
    """Network with normalisation fitted to n labelled states, no training."""
    net = _net(name, seed=seed)
    Td, Pd, Yd, W = _labelled(name, n, seed)
    net.fit_normalisation(Td, Pd, Yd, W)
    net.p.add_(torch.randn_like(net.p) * 0.02)
    lib = load_mlp()
    lib('crs_mlp_to_bf16', _p(net.p), _p(net.pb), net.nparam, _s())
    net.step += 1
    return net, Td, Pd, Yd


@pytest.mark.parametrize('name', MECHS)
@pytest.mark.parametrize('n', [5, 1000, 4096])
def test_fused_inference_matches_layered(n, name):
    """One-kernel inference (activations in registers, permuted tiles)
    against the layer-by-layer GEMM path, in the network's output space."""
    net, Td, Pd, Yd = _fit_random(name, n, 60 + n)
    Wf = torch.empty((net.nsp, n), dtype=torch.float64, device='cuda')
    Wl = torch.empty_like(Wf)
    net.predict(Td, Pd, Yd, Wf, fused=True)
    net.predict(Td, Pd, Yd, Wl, fused=False)
    s = net.scale[:, None]
    yf = (torch.asinh(Wf / s) - net.ymu[:, None].double()) / net.ysd[:, None].double()
    yl = (torch.asinh(Wl / s) - net.ymu[:, None].double()) / net.ysd[:, None].double()
    # the fused decode covers |arcsinh| <= 132 (FP32 exponential); beyond it
    # returns inf, which the gate sends to the exact kernel
    big = torch.asinh(Wl / s).abs() > 130.0
    ok = torch.isfinite(yf) & torch.isfinite(yl)
    assert bool((ok | big).all())
    d = (yf - yl).abs()[ok]
    print('n', n, 'max', d.max().item(), 'mean', d.mean().item(),
          'beyond range', int(big.sum().item()))
    # bf16 activations may round differently (summation order differs)
    assert d.max().item() < 3e-2 and d.mean().item() < 2e-3


@pytest.mark.parametrize('name', MECHS)
def test_fused_gate_flags_match_gate_kernel(name):
    net, Td, Pd, Yd = _fit_random(name, 3000, 77)
    Td[10] = 1.0e4                    # outside the training box
    Yd[2, 20] = -0.5                  # invalid composition
    Yd[4, 30] = float('nan')
    for rmax in (1e300, 1e-1, 1e-3):
        net.gate.rmax = rmax
        W1 = torch.empty((net.nsp, 3000), dtype=torch.float64, device='cuda')
        f1, i1, c1 = net.infer(Td, Pd, Yd, W1)
        f1 = f1.clone()
        i1 = i1[:int(c1.item())].clone()
        f2, i2, c2 = net.gate_states(Td, Pd, Yd, W1)
        if rmax > 1e100:
            # box, composition and finiteness only: identical decisions
            assert torch.equal(f1, f2)
            assert torch.equal(i1.sort().values,
                               i2[:int(c2.item())].sort().values)
        else:
            # balance sums are added in another order: the threshold may
            # fall differently for a state right at it
            assert int((f1 != f2).sum().item()) <= 3
        assert f1[10] and f1[20] and f1[30]
        print('rmax', rmax, 'fallback', int(c1.item()))


@pytest.mark.parametrize('name', MECHS)
def test_gate_that_admits_nothing_skips_the_network(name):
    """rmax < 0, what calibration returns when no residual threshold meets
    the target: the gated call (the C ABI's fused path and the Python path)
    gives the exact FP64 rates for every state, with every state on the
    fallback list."""
    n = 3000
    net, Td, Pd, Yd = _fit_random(name, n, 79)
    assert net.fused_ok
    net.gate.rmax = -1.0
    exact = net.mech.eval_wdot(Td, Pd, Yd, 'fp64')
    for fused in (True, False):
        W = torch.full((net.nsp, n), float('nan'), dtype=torch.float64,
                       device='cuda')
        assert net.evaluate(Td, Pd, Yd, W, fused=fused) == n
        assert torch.equal(W, exact)
        if fused:
            idx = net._bufs['fb_idx'][:n]
            flags = net._bufs['flags'][:n]
            assert torch.equal(idx.long().cpu(), torch.arange(n))
            assert bool((flags == 1).all())
    # a gate that admits states still runs the network
    net.gate.rmax = 1e300
    W = torch.empty_like(exact)
    assert net.evaluate(Td, Pd, Yd, W) < n


def test_trained_network_gain():
    """What training adds, measured on held-out states.

    Against the same network untrained, the arcsinh-space error must drop
    at least tenfold and the per-state normwise error must drop too.  The
    gated evaluation then keeps every state it does not trust exact
    (bit-identical to the FP64 kernel); the admitted fraction at a
    normwise target of 0.3 and both paths' times are printed (the honest
    gain over exact evaluation depends on that fraction).
    """
    from create_rate_subs_amd import metrics
    name = 'mech13'
    Td, Pd, Yd, W = _labelled(name, 1 << 16, 70)
    Tv, Pv, Yv, Wv = _labelled(name, 1 << 13, 71)
    untrained = _net(name, seed=21)
    untrained.fit_normalisation(Td, Pd, Yd, W)
    trained = _net(name, seed=21)
    trained.train(Td, Pd, Yd, W, epochs=12, batch=4096, lr=2e-3)

    def mse(net):
        Yh = net.forward(net.features(Tv, Pv, Yv))[:, :net.nout]
        return ((Yh - net.targets(Wv)) ** 2).mean().item()

    def normwise_median(net):
        out = torch.empty_like(Wv)
        net.predict(Tv, Pv, Yv, out)
        return metrics.normwise_err(out, Wv).median().item()

    m0, m1 = mse(untrained), mse(trained)
    n0, n1 = normwise_median(untrained), normwise_median(trained)
    print('arcsinh mse untrained', m0, 'trained', m1,
          '| normwise median untrained', n0, 'trained', n1)
    assert m1 < 0.1 * m0
    assert n1 < n0
    S = trained.gross_scale(Tv, Pv, Yv)
    _r, admitted, curve = trained.calibrate_gate(Tv, Pv, Yv, Wv, S, 0.3,
                                                 metric='normwise')
    out = torch.empty_like(Wv)
    m = trained.evaluate(Tv, Pv, Yv, out)
    flags = trained._bufs['flags'][:Tv.numel()].bool()
    assert m == int(flags.sum().item())
    assert torch.equal(out[:, flags], Wv[:, flags])
    ok = ~flags
    if ok.any():
        err = metrics.normwise_err(out[:, ok], Wv[:, ok])
        print('admitted', ok.double().mean().item(), 'max normwise',
              err.max().item())
    print('calibration curve (normwise)', curve)


@pytest.mark.parametrize('name', MECHS)
def test_checkpoint_roundtrip(tmp_path, name):
    net = _net(name, seed=17)
    Td, Pd, Yd, W = _labelled(name, 1 << 13, 53)
    net.train(Td, Pd, Yd, W, epochs=1, batch=4096)
    net.gate.rmax = 0.123
    net.save(str(tmp_path))
    other = _net(name, seed=99)
    other.load(str(tmp_path))
    a = net.predict(Td, Pd, Yd, torch.empty_like(W))
    b = other.predict(Td, Pd, Yd, torch.empty_like(W))
    assert torch.equal(a, b)
    assert other.gate.rmax == 0.123 and other.step == net.step


def test_checkpoint_of_another_mechanism_is_refused(tmp_path):
    """Same species count and network shape, different mechanism: mech13co
    is mech13 with CO moved last, so only the hash tells them apart."""
    import json
    import os
    net = _net('mech13', seed=17)
    net.save(str(tmp_path))
    other = _net('mech13co', seed=5)
    assert other.shapes == net.shapes
    before = other.p.clone()
    with pytest.raises(ValueError, match='another mechanism'):
        other.load(str(tmp_path))
    assert torch.equal(other.p, before)
    # an unexpected array name or shape in the manifest is refused too
    man_path = os.path.join(str(tmp_path), 'manifest.json')
    man = json.load(open(man_path))
    man['files']['p']['shape'] = [3]
    json.dump(man, open(man_path, 'w'))
    same = _net('mech13', seed=1)
    with pytest.raises(ValueError, match='array p'):
        same.load(str(tmp_path))
    man['files']['p']['shape'] = [net.nparam]
    man['files']['extra'] = dict(file='p.bin', dtype='float32', shape=[1])
    json.dump(man, open(man_path, 'w'))
    with pytest.raises(ValueError, match='differ'):
        same.load(str(tmp_path))
