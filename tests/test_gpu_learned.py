"""Learned absorbing boundary on the device (needs a GPU): the BF16 MFMA
inference kernel and the FP32 training kernels against a plain PyTorch FP32
reference of the same network (CPU), band extraction against the host, and
the network attached to a grid for one step."""
import numpy as np
import pytest

import jobs
from biotfd import learned, lib

pytestmark = pytest.mark.gpu

FIELDS = ("U", "W", "QX", "QZ", "TXX", "TZZ", "TXZ", "P")


def _weights(seed, last=1.0):
    w = learned.init_weights(seed)
    for s in range(3):
        w[s, learned.NW1 + learned.NW2:] *= last
    return w


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("side,L", [(0, 130), (1, 52), (2, 209)])
def test_inference_kernel_matches_fp32_reference(side, L):
    """BF16 inputs, weights and activations with FP32 accumulation against the
    FP32 network: the predicted change of the strip agrees to 2e-2 relative
    L2 (BF16 keeps 8 significant bits; three layers)."""
    rng = np.random.default_rng(side)
    w = _weights(10 + side, last=10.0)
    scales = rng.uniform(0.5, 2.0, (3, 8)).astype(np.float32) * np.float32(10.0) ** rng.integers(-3, 3, (3, 1))
    bands = rng.standard_normal((4, L, learned.BAND)).astype(np.float32)
    bands *= np.repeat(scales[side], learned.CELLS)[None, None, :]
    net = learned.Network(w, scales)
    got = net.infer(side, bands)
    want = learned.torch_forward(w[side], scales[side], bands)
    prev = bands[3].reshape(L, 8, learned.CELLS)[:, :, :8]
    err = _rel_l2(got - prev, want - prev)
    print(f"side {side} L={L}: relative L2 of the change {err:.2e}", flush=True)
    assert np.abs(want - prev).max() > 0
    assert err < 2e-2
    # zero bands give no change at all (no bias, leaky ReLU)
    np.testing.assert_array_equal(net.infer(side, np.zeros_like(bands)), 0.0)
    net.close()


def _random_state(sim, seed):
    rng = np.random.default_rng(seed)
    for f in FIELDS:
        sim.set_field(f, rng.standard_normal((sim.nz, sim.nx)).astype(np.float32))


def _host_fields(sim):
    return {f: sim.field(f) for f in FIELDS}


@pytest.mark.parametrize("residual,inner", [(False, 8), (True, 3)])
def test_band_extract_matches_host_and_training_kernels_match_torch(residual, inner):
    import torch
    job = jobs.tiny_layered(steps=30, tpml=8)
    tr = learned.Trainer(job.mx, job.mzp, max_groups=4, batch=2, residual=residual, inner=inner)
    host = []   # [group][level] bands on the host
    for g in range(4):
        sim = lib.Simulation(job)
        _random_state(sim, 100 + g)
        sim.run(g)            # groups at different steps
        levels = []
        for lev in range(5):
            if lev:
                sim.run(1)
            tr.capture(sim, g, lev)
            levels.append(learned.band_host(_host_fields(sim), job.mx, job.mzp, job.tpml))
        host.append(np.stack(levels))
        sim.close()
    host = np.stack(host)                                   # [4][5][band]
    scales = tr.compute_scales(4)
    lz, lx = job.mzp + 1, job.mx + 2 * learned.S
    offs = [0, lz * learned.BAND, 2 * lz * learned.BAND]
    lens = [lz, lz, lx]
    for s in range(3):
        b = host[:, :4, offs[s]:offs[s] + lens[s] * learned.BAND].reshape(-1, 8, learned.CELLS)
        rms = np.sqrt((b.astype(np.float64) ** 2).mean(axis=(0, 2)))
        np.testing.assert_allclose(scales[s], rms, rtol=1e-5)
    w0 = _weights(7, last=3.0)
    tr.set_weights(w0)
    groups = np.array([1, 3], np.int32)
    for s in range(3):
        # torch FP32 loss and gradients of the same minibatch
        inv = 1.0 / scales[s]
        xs, ys = [], []
        for g in groups:
            lv = host[g, :, offs[s]:offs[s] + lens[s] * learned.BAND].reshape(5, lens[s], 8, learned.CELLS)
            x = (lv[:4] * inv[None, None, :, None]).transpose(1, 0, 2, 3).reshape(lens[s], -1)
            tgt = lv[4, :, :, :8] if residual else lv[4, :, :, :8] - lv[3, :, :, :8]
            y = (tgt * inv[None, :, None]).reshape(lens[s], -1)
            xs.append(x.T)
            ys.append(y.T)
        x = torch.tensor(np.stack(xs), dtype=torch.float32)
        y = torch.tensor(np.stack(ys), dtype=torch.float32)
        ws = [torch.tensor(np.ascontiguousarray(w.transpose(2, 1, 0)), requires_grad=True)
              for w in learned.split_weights(w0[s])]
        h = x
        for k, w in enumerate(ws):
            h = torch.nn.functional.conv1d(h, w, padding=2)
            if k < 2:
                h = torch.nn.functional.leaky_relu(h, 0.1)
        mask = torch.tensor((np.arange(64) % 8) >= 8 - inner)[None, :, None].expand_as(h)
        loss = ((h - y) ** 2)[mask].mean()
        loss.backward()
        got = tr.step(s, groups, 0.0)
        assert abs(got - loss.item()) <= 1e-4 * loss.item(), (s, got, loss.item())
        # the first Adam step moves each weight by lr * g / (|g| + 1e-8):
        # its sign is the gradient's
        lr = 1e-3
        tr.step(s, groups, lr)
        w1 = tr.get_weights()[s]
        gt = np.concatenate([w.grad.numpy().transpose(2, 1, 0).reshape(-1) for w in ws])
        moved = (w0[s] - w1) / lr
        big = np.abs(gt) > max(1e-3 * np.abs(gt).max(), 1e-4)
        assert big.sum() > 100
        agree = np.mean(np.sign(moved[big]) == np.sign(gt[big]))
        print(f"side {s}: loss {got:.6e} (torch {loss.item():.6e}), gradient signs agree on "
              f"{agree:.5f} of {big.sum()} weights", flush=True)
        assert agree > 0.999
        np.testing.assert_allclose(np.abs(moved[big]), 1.0, rtol=1e-3)
        tr.set_weights(w0)
    tr.close()


@pytest.mark.parametrize("residual,gain,inner,damping",
                         [(False, 1.0, 8, False), (True, 0.5, 8, False), (True, 2.0, 2, False),
                          (True, 3.0, 8, True)])
def test_attached_network_replaces_only_the_strip(residual, gain, inner, damping):
    """One step from a random state: every cell outside the 8-cell strip is
    bit-identical to the plain 8-cell PML run; the strip cells are the
    network's prediction from the band of the start level (direct form), or
    the plain run's strip plus the network's correction (residual form), or
    that sum limited to the interval between 0 and the plain run's value
    (damping form)."""
    job = jobs.tiny_layered(steps=10, tpml=8)
    plain, run = lib.Simulation(job), lib.Simulation(job)
    _random_state(plain, 5)
    _random_state(run, 5)
    band0 = learned.band_host(_host_fields(run), job.mx, job.mzp, job.tpml)
    w = _weights(3, last=10.0)
    scales = np.full((3, 8), 0.7, np.float32)
    net = learned.Network(w, scales, residual=residual, gain=gain, inner=inner, damping=damping)
    net.attach(run)
    plain.run(1)
    run.run(1)
    lz, lx = job.mzp + 1, job.mx + 2 * learned.S
    offs, lens = [0, lz * learned.BAND, 2 * lz * learned.BAND], [lz, lz, lx]
    pmll = 1 - job.tpml
    for f, name in enumerate(FIELDS):
        a, b = run.field(name), plain.field(name)
        strip = np.zeros(a.shape, bool)
        for s in range(3):
            bands = np.zeros((4, lens[s], learned.BAND), np.float32)
            bands[3] = band0[offs[s]:offs[s] + lens[s] * learned.BAND].reshape(lens[s], -1)
            pred = learned.torch_forward(w[s], scales[s], bands, gain)[:, f, :]     # [L][8]
            prev = bands[3].reshape(lens[s], 8, -1)[:, f, :8]
            c = np.arange(8)
            if s == 0:
                ii, ll = (1 - 8 + c)[None, :], np.arange(lz)[:, None]
            elif s == 1:
                ii, ll = (job.mx + 8 - c)[None, :], np.arange(lz)[:, None]
            else:
                ii, ll = (1 - 8 + np.arange(lx))[:, None], (job.mzp + 8 - c)[None, :]
            ii, ll = np.broadcast_to(ii, pred.shape), np.broadcast_to(ll, pred.shape)
            strip[ll, ii - pmll] = True
            got = a[ll, ii - pmll]
            cpml = b[ll, ii - pmll]
            want = cpml + (pred - prev) if residual else pred
            if damping:
                want = np.where(cpml >= 0, np.clip(want, 0, np.maximum(cpml, 0)),
                                np.clip(want, np.minimum(cpml, 0), 0))
                # the limit binds on a good share of the cells at this gain
                assert np.mean(np.abs(want - (cpml + pred - prev)) > 0) > 0.1
            outer = np.arange(8) < 8 - inner       # strip cells the network leaves alone
            want[:, outer] = b[ll, ii - pmll][:, outer]
            scale = np.abs(pred - prev).max() + 1e-30
            assert np.abs(got - want).max() <= 2e-2 * scale + 1e-6 * np.abs(want).max(), (name, s)
        np.testing.assert_array_equal(a[~strip], b[~strip], err_msg=name)
    net.close()


def test_strip_copy_and_residual_capture():
    """bq_strip_copy puts a wide-PML run's values into an 8-cell grid's strip
    (bit for bit, same physical cells); bq_lb_trainer_capture_residual records
    band(wide) - band(8-cell grid), whose strip cells are the residual
    trainer's target (checked through the loss of a zero network)."""
    job_w = jobs.tiny_layered(steps=30, tpml=14)
    job_t = jobs.tiny_layered(steps=30, tpml=8)
    wide, trunc = lib.Simulation(job_w), lib.Simulation(job_t)
    _random_state(wide, 11)   # every strip cell active from the start
    wide.run(6)
    trunc.run(6)
    learned.strip_copy(trunc, wide)
    fw, ft = _host_fields(wide), _host_fields(trunc)
    bw = learned.band_host(fw, job_w.mx, job_w.mzp, job_w.tpml)
    bt = learned.band_host(ft, job_t.mx, job_t.mzp, job_t.tpml)
    lz, lx = job_t.mzp + 1, job_t.mx + 2 * learned.S
    views = [bw[:2 * lz * 96].reshape(2 * lz, 8, 12), bt[:2 * lz * 96].reshape(2 * lz, 8, 12),
             bw[2 * lz * 96:].reshape(lx, 8, 12), bt[2 * lz * 96:].reshape(lx, 8, 12)]
    np.testing.assert_array_equal(views[0][:, :, :8], views[1][:, :, :8])
    np.testing.assert_array_equal(views[2][:, :, :8], views[3][:, :, :8])
    assert np.abs(views[0][:, :, :8]).max() > 0
    wide.run(1)
    trunc.run(1)
    tr = learned.Trainer(job_t.mx, job_t.mzp, max_groups=2, batch=2, residual=True)
    for g in range(2):
        for lev in range(4):
            tr.capture(wide, g, lev)
        tr.capture_residual(wide, trunc, g)
    scales = tr.compute_scales(2)
    tr.set_weights(np.zeros((3, learned.NSIDE), np.float32))
    diff = (learned.band_host(_host_fields(wide), job_w.mx, job_w.mzp, job_w.tpml)
            - learned.band_host(_host_fields(trunc), job_t.mx, job_t.mzp, job_t.tpml))
    offs, lens = [0, lz * 96, 2 * lz * 96], [lz, lz, lx]
    for s in range(3):
        d = diff[offs[s]:offs[s] + lens[s] * 96].reshape(lens[s], 8, 12)[:, :, :8] / scales[s][None, :, None]
        want = float((d.astype(np.float64) ** 2).mean())
        got = tr.step(s, np.array([0, 1], np.int32), 0.0)
        assert want > 0 and abs(got - want) <= 1e-5 * want, (s, got, want)


def test_inner_profile_strip_tracks_the_wide_run():
    """An 8-cell grid with the inner 8 nodes of a 40-cell PML's coefficients
    (learned.inner_profile_pml), started from the wide run's fields at the same
    cells, advances one step exactly as the wide run does at every cell at
    least 3 from its outer edge (the others miss the wide run's outer cells)."""
    wide_job = jobs.tiny_layered(steps=10, tpml=40)
    job8 = jobs.tiny_layered(steps=10, tpml=8)
    wide = lib.Simulation(wide_job)
    _random_state(wide, 21)
    trunc = lib.Simulation(job8, pml_arrays=learned.inner_profile_pml(wide, 8))
    off = 40 - 8
    for f in FIELDS:
        trunc.set_field(f, np.ascontiguousarray(wide.field(f)[:trunc.nz, off:off + trunc.nx]))
    wide.run(1)
    trunc.run(1)
    for f in FIELDS:
        a = trunc.field(f)[:trunc.nz - 3, 3:trunc.nx - 3]
        b = wide.field(f)[:trunc.nz - 3, off + 3:off + trunc.nx - 3]
        np.testing.assert_array_equal(a, b, err_msg=f)
    wide.close()
    trunc.close()
