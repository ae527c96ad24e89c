"""DN_load_volume / DN_save_volume (voxel.c:520-654) against the oracle's codec
and the reference's own saved file, and the reference's demo scene as a
real-world parity fixture: its specular (mirror and glossy), emissive and
transparent refracting materials go through the draw and lighting passes,
checked against the FP32 oracle exactly as the synthetic scenes are."""
from __future__ import annotations

import types
from pathlib import Path

import numpy as np
import pytest

import oracle
from doonengine_amd.volume import Volume
from oracle import voxvol
from test_gpu_parity import draw_gpu, oracle_scene, run_frame_and_compare

pytestmark = pytest.mark.gpu

DEMO = Path(__file__).resolve().parent / "golden" / "demo.voxvol"


def assert_matches_file(vol: Volume, vf: voxvol.VolumeFile):
    mp, ch, _ = vol.host_state()
    assert vol.map_size == tuple(vf.map_size)
    assert len(ch) == len(vf.chunks)
    expect_flag = np.zeros(len(mp), np.uint8)
    for i, c in enumerate(vf.chunks):
        np.testing.assert_array_equal(ch["pos"][i], c.pos)
        if c.voxels is None:
            continue
        t = c.pos[0] + vf.map_size[0] * (c.pos[1] + vf.map_size[1] * c.pos[2])
        expect_flag[t] = 1
        assert mp["chunkIndex"][t] == i
        assert ch["numVoxels"][i] == c.num_voxels and ch["updated"][i] == 0
        np.testing.assert_array_equal(ch["voxels"][i], voxvol.chunk_array(c.voxels), err_msg=f"slot {i}")
    np.testing.assert_array_equal(mp["flag"], expect_flag)
    assert bytes(np.ascontiguousarray(vol.materials).tobytes()) == vf.materials


def test_load_demo_matches_oracle(gpu):
    vf = voxvol.decode(DEMO.read_bytes())
    with Volume.load(DEMO) as vol:
        assert_matches_file(vol, vf)
        v = vol.v
        tail = (v.camPos.x, v.camPos.y, v.camPos.z, v.camOrient.x, v.camOrient.y, v.camOrient.z, v.camFOV,
                v.camViewMode, v.sunDir.x, v.sunDir.y, v.sunDir.z, v.sunStrength.x, v.sunStrength.y,
                v.sunStrength.z, v.ambientLightStrength.x, v.ambientLightStrength.y, v.ambientLightStrength.z,
                v.diffuseBounceLimit, v.specBounceLimit, v.shadowSoftness, v.skyGradientBot.x, v.skyGradientBot.y,
                v.skyGradientBot.z, v.skyGradientTop.x, v.skyGradientTop.y, v.skyGradientTop.z)
        np.testing.assert_array_equal(np.array(tail, np.float64), np.array(vf.tail, np.float64))


def test_save_reproduces_reference_file(gpu, tmp_path):
    with Volume.load(DEMO) as vol:
        vol.save(tmp_path / "again.voxvol")
    assert (tmp_path / "again.voxvol").read_bytes() == DEMO.read_bytes()


def test_save_load_round_trip_after_edits(gpu, tmp_path):
    with Volume.load(DEMO) as vol:
        vol.shape_sphere(3, (200, 10, 10), (40.0, 12.0, 40.0), 9.0)
        vol.shape_sphere(255, (0, 0, 0), (20.0, 6.0, 20.0), 7.5, flip_normals=True)
        vol.save(tmp_path / "edited.voxvol")
        mp, ch, _ = vol.host_state()
    data = (tmp_path / "edited.voxvol").read_bytes()
    vf = voxvol.decode(data)
    assert voxvol.encode(vf) == data
    with Volume.load(tmp_path / "edited.voxvol") as back:
        assert_matches_file(back, vf)
        mp2, ch2, _ = back.host_state()
    live = mp["flag"] != 0
    np.testing.assert_array_equal(mp["flag"], mp2["flag"])
    for t in np.nonzero(live)[0]:
        a, b = ch[mp["chunkIndex"][t]], ch2[mp2["chunkIndex"][t]]
        solid = (a["voxels"][:, 0] >> 24) != 255
        np.testing.assert_array_equal(a["voxels"][solid], b["voxels"][solid])
        assert a["numVoxels"] == b["numVoxels"]


def test_malformed_files_are_rejected(gpu, tmp_path):
    data = DEMO.read_bytes()
    for name, blob in (("truncated", data[: len(data) // 2]), ("tail cut", data[:-4]),
                       ("no header", data[:10]), ("zero slots", data[:12] + bytes(8) + data[20:])):
        (tmp_path / name).write_bytes(blob)
        with pytest.raises(RuntimeError):
            Volume.load(tmp_path / name)
    bad = bytearray(data)
    bad[20:22] = (60000).to_bytes(2, "little")              # first record longer than the file allows
    (tmp_path / "bad").write_bytes(bytes(bad))
    with pytest.raises(RuntimeError):
        Volume.load(tmp_path / "bad")


def demo_scene(vol, w=192, h=128):
    return types.SimpleNamespace(map_size=vol.map_size, width=w, height=h)


def test_demo_scene_draw_parity(gpu):
    o = oracle.load("f32")
    with Volume.load(DEMO) as vol:
        sc = demo_scene(vol)
        vol.request_all()
        vol.sync()
        vol.synchronize()
        s, (m, c, v) = oracle_scene(o, vol, sc)
        rgba, px, view, proj = draw_gpu(vol, sc)
        from doonengine_amd.volume import draw_matrices, mat_to_np
        iv, ivp = draw_matrices(view, proj)
        orgba, opx, _ = o.draw(s, mat_to_np(iv), mat_to_np(ivp), sc.width, sc.height)
        assert (px["voxelIndex"] != 0xFFFFFFFF).mean() > 0.15   # the file's camera sees ~21% geometry
        for f in ("voxelIndex", "mapSteps", "voxelSteps"):
            assert np.array_equal(px[f], opx[f]), f
        np.testing.assert_allclose(rgba[..., :3], orgba[..., :3], rtol=2e-6, atol=1e-7)


def test_demo_scene_lighting_parity(gpu):
    o = oracle.load("f32")
    with Volume.load(DEMO) as vol:
        sc = demo_scene(vol)
        vol.request_all()
        vol.sync()
        vol.synchronize()
        lit = 0
        for f in range(3):
            _, n = run_frame_and_compare(vol, sc, o, f, 50.0 + f / 60.0, sc.width, sc.height)
            lit += n
        assert lit > 2000


def test_demo_scene_direct_rays_through_glass(gpu):
    """The demo scene with 16 direct rays: its glass (opacity 0.5) makes some
    rays stop inside a transparent voxel, where the shader carries lastVoxID
    from one ray to the next; those voxels must fall back from the direct terms
    traced ahead to tracing in sequence, and every voxel must still match."""
    o = oracle.load("f32")
    with Volume.load(DEMO) as vol:
        vol.v.numDirectRays = 16
        sc = demo_scene(vol)
        vol.request_all()
        vol.sync()
        vol.synchronize()
        for f in range(2):
            run_frame_and_compare(vol, sc, o, f, 60.0 + f / 60.0, sc.width, sc.height)
