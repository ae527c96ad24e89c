"""DNX_rc_predict_samples_noalloc against DNX_rc_predict_samples through the
C-ABI: the allocation-free variant keeps its encoded rows in a fixed
per-device array instead of a buffer grown with hipMalloc, with the same
kernels, so its predictions must equal the existing entry point's bit for bit.
Two calls in a row (the second smaller, so the array holds rows of the first
call past the second one's end), and the size checks that run on the host."""
from __future__ import annotations

import numpy as np
import pytest

from doonengine_amd import scenes
from doonengine_amd._hip import DeviceBuffer
from doonengine_amd.volume import RadianceCache, Volume, rc_pack_weights
from oracle import radiance_cache as orc

pytestmark = pytest.mark.gpu

MAX_SAMPLES = 262144                     # DNX_RC_NOALLOC_MAX_SAMPLES


def samples(rng, n, extent):
    s = np.zeros((n, 16), np.float32)
    s[:, 0:3] = rng.uniform(0.5, extent - 0.5, (n, 3))
    nrm = rng.normal(size=(n, 3))
    s[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    s[:, 6:9] = rng.uniform(0.1, 0.9, (n, 3))
    s[:, 12] = 1.0
    return s


def test_noalloc_prediction_equals_allocating_prediction(gpu):
    sc = scenes.config_small(seed=31)
    extent = 8.0 * max(sc.map_size)
    rng = np.random.default_rng(77)
    with Volume(sc.map_size, min_chunks=64) as vol:
        sc.apply(vol)
        rc = RadianceCache(vol, log2Table=14, extent=extent)
        try:
            cfg = orc.Config(log2_table=14, extent=extent)
            mats, table = orc.init_params(cfg, seed=5)
            table = rng.normal(0, 0.1, table.shape).astype(np.float32)
            rc.set_params(rc_pack_weights(mats), table.reshape(-1))
            for call, n in enumerate((MAX_SAMPLES, 4096)):
                print(f"[no-alloc] call {call}: {n} samples", flush=True)
                smp = DeviceBuffer.from_array(samples(rng, n, extent))
                ref = DeviceBuffer(n * 16)
                got = DeviceBuffer(n * 16)
                rc.predict_samples(smp, n, ref)
                assert rc.predict_samples_noalloc(smp, n, got)
                vol.synchronize()
                a = ref.download(np.float32, (n, 4))
                b = got.download(np.float32, (n, 4))
                assert np.abs(a[:, :3]).max() > 0.0
                np.testing.assert_array_equal(b.view(np.uint32), a.view(np.uint32), err_msg=f"call {call}")
            # rejected on the host, before any launch
            smp = DeviceBuffer(16 * 16 * 4)
            out = DeviceBuffer(16 * 16)
            assert not rc.predict_samples_noalloc(smp, MAX_SAMPLES + 16, out)
            assert not rc.predict_samples_noalloc(smp, 24, out)
            assert not rc.predict_samples_noalloc(smp, 0, out)
            vol.synchronize()
        finally:
            rc.close()
