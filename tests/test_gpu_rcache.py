"""Radiance cache on the GPU against the CPU oracle, and what it gains.

- Forward outputs and the gradients of every weight matrix and of the hash
  table agree with the oracle's bf16-operand restatement (which the CPU suite
  pins against finite differences) to within accumulation-order differences.
- The first Adam step moves every parameter by lr * g / (|g| + eps).
- The learning-rate decay scales step t by 1 / (1 + (t - 1) / D)^P, and the
  growing parameter average (decay min(emaDecay, 1 - G / t)) predicts with
  exactly the parameters it should: the latest ones while G / t >= 1, then the
  FP32 average of the steps so far.
- On a small scene, after the same number of frames, the trained cache's
  per-voxel irradiance is closer to a 1024-frame FP64-accumulated reference
  than the plain running mean is.
"""
from __future__ import annotations

import numpy as np
import pytest

from doonengine_amd import scenes
from doonengine_amd._hip import DeviceBuffer
from doonengine_amd.volume import RadianceCache, Volume, rc_pack_weights, rc_unpack_weights
from oracle import radiance_cache as orc

pytestmark = pytest.mark.gpu

LOG2T = 14


def make_samples(rng, n, extent=128.0):
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(0.5, extent - 0.5, (n, 3))
    nrm = rng.normal(size=(n, 3))
    s[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    s[:, 6:9] = rng.uniform(0.1, 0.9, (n, 3))
    s[:, 9:12] = rng.uniform(0.0, 1.0, (n, 3))
    s[:, 12] = 1.0
    return s


@pytest.fixture(scope="module")
def small_volume(gpu):
    sc = scenes.config_small()
    vol = Volume(sc.map_size, min_chunks=64)
    sc.apply(vol)
    vol.request_all()
    vol.sync()
    vol.synchronize()
    yield vol, sc
    vol.close()


def rel_fro(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def test_gradients_match_oracle(small_volume):
    vol, _ = small_volume
    cfg = orc.Config(log2_table=LOG2T, extent=128.0)
    rng = np.random.default_rng(3)
    mats, table = orc.init_params(cfg, seed=5)
    table = (rng.normal(0, 0.05, table.shape)).astype(np.float32)      # larger than init so gradients flow
    B = 512
    smp = make_samples(rng, B)
    rc = RadianceCache(vol, log2Table=LOG2T, extent=128.0)
    try:
        rc.set_params(rc_pack_weights(mats), table.reshape(-1))
        buf = DeviceBuffer.from_array(smp)
        out = DeviceBuffer(B * 16)
        rc.predict_samples(buf, B, out)
        rc.train_samples(buf, B, step=False)
        vol.synchronize()
        y = out.download(np.float32, (B, 4))
        _, _, g, tg = rc.get_params()
        ref = orc.forward_backward(cfg, mats, table, smp[:, 0:3], smp[:, 3:6], smp[:, 6:9], smp[:, 9:12],
                                   bf16_ops=True)
        ey = np.abs(y[:, :3] - ref["y"]).max() / np.abs(ref["y"]).max()
        print(f"[rc] forward max rel err {ey:.2e}", flush=True)
        assert ey < 2e-3
        assert np.all(y[:, 3] == 1.0)
        for i, (gm, rm) in enumerate(zip(rc_unpack_weights(g), ref["dW"])):
            e = rel_fro(gm, rm)
            print(f"[rc] dW{i} rel err {e:.2e}", flush=True)
            assert e < 5e-3, (i, e)
        # padding of the weight layout receives no gradient
        pad = g.copy()
        for off, rows, ld, used in ((0, 64, 48, 38), (11264, 16, 64, 3)):
            blk = pad[off:off + rows * ld].reshape(rows, ld)
            if used == 38:
                assert np.all(blk[:, 38:] == 0.0)
            else:
                assert np.all(blk[3:, :] == 0.0)
        et = rel_fro(tg.reshape(table.shape), ref["dtable"])
        print(f"[rc] dtable rel err {et:.2e}", flush=True)
        assert et < 5e-3
    finally:
        rc.close()


def test_steps_drawn_together_equal_steps_drawn_alone(small_volume):
    """DNX_rc_train_frame draws the training samples of several steps in one
    launch, batch k with the seed of the frame counter plus k: the last step's
    samples of a 3-step frame equal, bit for bit, those of the third of three
    1-step frames, and differ from the first step's; a 17-step frame (two
    launches) ends on the samples of the 17th of 17 1-step frames."""
    vol, _ = small_volume
    vol.mark_visible(1 << 20)
    vol.sync()
    vol.synchronize()
    a = RadianceCache(vol, log2Table=LOG2T, extent=128.0, samplesPerFrame=4096, seed=5)
    b = RadianceCache(vol, log2Table=LOG2T, extent=128.0, samplesPerFrame=4096, seed=5)
    try:
        a.train_frame(10.0, 3)
        sa = a.last_samples()
        b.train_frame(10.0, 1)
        first = b.last_samples()
        b.train_frame(10.0, 1)
        b.train_frame(10.0, 1)
        sb = b.last_samples()
        assert sa.shape == (4096, 16)
        assert (sa[:, 12] == 1.0).sum() > 2000                       # rows holding a drawn voxel
        assert np.array_equal(sa.view(np.uint32), sb.view(np.uint32))
        assert not np.array_equal(sa.view(np.uint32), first.view(np.uint32))
        # more steps than one launch draws (16): a launch of 16 batches, then one of 1
        a.train_frame(11.0, 17)
        for _ in range(17):
            b.train_frame(11.0, 1)
        assert np.array_equal(a.last_samples().view(np.uint32), b.last_samples().view(np.uint32))
    finally:
        a.close()
        b.close()


def test_adam_first_step(small_volume):
    vol, _ = small_volume
    cfg = orc.Config(log2_table=LOG2T, extent=128.0)
    rng = np.random.default_rng(11)
    mats, table = orc.init_params(cfg, seed=2)
    B = 256
    smp = make_samples(rng, B)
    rc = RadianceCache(vol, log2Table=LOG2T, extent=128.0)
    try:
        w0 = rc_pack_weights(mats)
        t0 = table.reshape(-1).copy()
        buf = DeviceBuffer.from_array(smp)
        rc.set_params(w0, t0)
        rc.train_samples(buf, B, step=False)
        _, _, g, tg = rc.get_params()
        rc.set_params(w0, t0)
        rc.train_samples(buf, B, step=True)
        w1, t1, g1, tg1 = rc.get_params()
        assert np.all(g1 == 0.0) and np.all(tg1 == 0.0)            # gradients cleared by the step
        for p0, p1, gg in ((w0, w1, g), (t0, t1, tg)):
            m = np.abs(gg) > 1e-6
            want = p0[m] - cfg.lr * gg[m] / (np.abs(gg[m]) + cfg.eps)
            assert np.allclose(p1[m], want, rtol=1e-5, atol=1e-6)
            assert np.all(p1[~(np.abs(gg) > 0)] == p0[~(np.abs(gg) > 0)])
    finally:
        rc.close()


def test_lr_decay_and_growing_average(small_volume):
    vol, _ = small_volume
    cfg = orc.Config(log2_table=LOG2T, extent=128.0)
    rng = np.random.default_rng(23)
    mats, table = orc.init_params(cfg, seed=4)
    B = 256
    smp = make_samples(rng, B)
    lr, D, G, ema = 0.01, 4.0, 8.0, 0.95
    # beta1 = beta2 = 0: step t moves a parameter by lr_t * g / (|g| + eps), so
    # the typical move of the weights is lr_t itself
    rc = RadianceCache(vol, log2Table=LOG2T, extent=128.0, lr=lr, beta1=0.0, beta2=0.0, lrDecaySteps=D,
                       lrDecayPower=1.0, emaGrowth=G, emaDecay=ema)
    ref = RadianceCache(vol, log2Table=LOG2T, extent=128.0)
    try:
        buf = DeviceBuffer.from_array(smp)
        rc.set_params(rc_pack_weights(mats), table.reshape(-1))
        params = [rc.get_params()[:2]]
        for _ in range(9):
            rc.train_samples(buf, B, step=True)
            params.append(rc.get_params()[:2])
        for t in (1, 2, 5, 9):
            d = np.abs(params[t][0] - params[t - 1][0])
            moved = d > 0
            assert moved.sum() > 1000
            assert abs(float(np.median(d[moved])) / (lr / (1.0 + (t - 1) / D)) - 1.0) < 1e-3, t
        # the average as the Adam kernel keeps it, in FP32
        ew, et = params[1]
        for t in range(2, 10):
            dk = np.float32(min(float(np.float32(ema)), max(0.0, 1.0 - G / t)))
            ew = dk * ew + (np.float32(1.0) - dk) * params[t][0]
            et = dk * et + (np.float32(1.0) - dk) * params[t][1]
        assert not np.array_equal(ew, params[9][0])         # step 9 is the first real average
        out_rc, out_ref = DeviceBuffer(B * 16), DeviceBuffer(B * 16)
        rc.predict_samples(buf, B, out_rc)
        ref.set_params(ew, et)
        ref.predict_samples(buf, B, out_ref)
        assert np.array_equal(out_rc.download(np.float32, (B, 4)), out_ref.download(np.float32, (B, 4)))
    finally:
        rc.close()
        ref.close()


def test_prediction_exact_term_matches_lighting_pass(gpu):
    # With no diffuse bounces the pass's first-frame estimate is exactly its
    # ambient + direct term, and an untrained cache (zero output layer)
    # predicts exactly its query rows' ambient + direct term: the two must
    # agree bit for bit on every voxel.
    sc = scenes.config_small(seed=5, chunks=(6, 4, 6))
    vol = Volume(sc.map_size, min_chunks=256)
    try:
        sc.apply(vol)
        vol.v.diffuseBounceLimit = 0
        vol.v.numDirectRays = 16
        vol.request_all()
        vol.sync()
        vol.mark_visible(1 << 20)
        vol.sync()
        vol.synchronize()
        n = vol.fetch_requests().size
        slots = n * 32
        dbg, idx = DeviceBuffer(slots * 24), DeviceBuffer(slots * 4)
        vol.lib.DNX_set_lighting_debug_output(vol.ptr, dbg.ptr, idx.ptr, n)
        vol.update_lighting(1, 1000, 123.456)
        vol.synchronize()
        diff = dbg.download(np.float32, (slots, 6))[:, :3]
        vi = idx.download(np.uint32, (slots,))
        vol.lib.DNX_set_lighting_debug_output(vol.ptr, None, None, 0)
        rc = RadianceCache(vol)
        try:
            pred = DeviceBuffer(slots * 16)
            rc.predict_requests(pred, n)
            vol.synchronize()
            p = pred.download(np.float32, (slots, 4))
        finally:
            rc.close()
        valid = vi != 0xFFFFFFFF
        assert valid.sum() > 1000
        assert np.any(diff[valid] > 0)
        assert np.array_equal(p[valid, :3], diff[valid])
    finally:
        vol.close()


def test_warm_cache_beats_cold_accumulation(gpu):
    """What the trained network gains: after a lighting-history reset (an edit,
    or a region seen for the first time) the engine's running mean restarts
    from one sample, while a cache trained on the scene predicts at once.
    Scored against a FP64 reference of 16384 independent samples per voxel.
    Also: training lowers the cache's own error, and its exact direct term
    plus learned indirect term beats the exact term alone."""
    sc = scenes.config_small()
    sc.params = dict(sc.params, numDirectRays=8)

    def load():
        v = Volume(sc.map_size, min_chunks=64)
        sc.apply(v)
        v.request_all()
        v.sync()
        v.mark_visible(1 << 20)
        v.sync()
        v.synchronize()
        return v

    vol = load()
    cold = None
    rc = RadianceCache(vol, samplesPerFrame=8192, seed=3)
    try:
        n = vol.fetch_requests().size
        slots = n * 32
        acc = DeviceBuffer(slots * 32)
        acc.zero()
        rc.reference(acc, n, 16384, seed=11)
        vol.synchronize()
        ref = acc.download(np.float64, (slots, 4))
        valid = ref[:, 3] > 0
        assert valid.sum() > 1000
        refm = np.clip(ref[valid, :3] / ref[valid, 3:4], 0.0, 1.0)
        pred = DeviceBuffer(slots * 16)

        def cache_err():
            rc.predict_requests(pred, n)
            vol.synchronize()
            p = pred.download(np.float32, (slots, 4))
            assert np.all(p[valid, 3] == 1.0)
            return orc.relative_rmse(np.clip(p[valid, :3], 0, 1), refm), p

        # the exact term alone: a cache whose network predicts 0 (fresh, zero output layer)
        e_exact, _ = cache_err()
        errs = {}
        for f in range(96):
            vol.mark_visible(1 << 20)
            vol.sync()
            rc.train_frame(123.456 + f / 60.0, 1)
            if f + 1 in (8, 96):
                errs[f + 1], _ = cache_err()
        cold = load()
        dbg, idx = DeviceBuffer(slots * 24), DeviceBuffer(slots * 4)
        cold.lib.DNX_set_lighting_debug_output(cold.ptr, dbg.ptr, idx.ptr, n)
        cold_err = {}
        for k in range(1, 5):
            cold.mark_visible(1 << 20)
            cold.sync()
            cold.update_lighting(1, 1000, 300.0 + k / 60.0)
            cold.synchronize()
            cold_err[k] = orc.relative_rmse(np.clip(dbg.download(np.float32, (slots, 6))[valid, :3], 0, 1), refm)
        cold.lib.DNX_set_lighting_debug_output(cold.ptr, None, None, 0)
        print(f"[rc] exact term alone {e_exact:.4f}; cache after 8 frames {errs[8]:.4f}, after 96 {errs[96]:.4f}; "
              f"cold running mean after 1-4 frames " + " ".join(f"{cold_err[k]:.4f}" for k in range(1, 5)), flush=True)
        assert errs[96] < errs[8]
        assert errs[96] < e_exact
        assert errs[96] < cold_err[2]
    finally:
        rc.close()
        vol.close()
        if cold is not None:
            cold.close()
