"""The committed config-1 fixture (tests/golden/room_c1.npz, written by
tests/golden/make_golden.py): the closed room with two boxes, 12,312
triangles in 8 objects, one ceiling rect light, 256 x 256 queries.

On the CPU the oracle must reproduce it: the scene generator and the oracle's
hierarchy by their digests, every query of the image by the digests of its
outputs, and the sampled pixels value for value. On the GPU the library's
whole path (host builder, camera pass, cone kernel, sky) must reproduce the
same bits through the C ABI."""
import numpy as np
import pytest

from continuum_amd.scenes import meshes_to_arrays, room_scene
from tests.golden.make_golden import OUT, compute, digest, record


@pytest.fixture(scope="module")
def golden():
    z = np.load(OUT, allow_pickle=False)
    return {k: z[k] for k in z.files}


def _same(a, b, what):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype == np.float32:
        a, b = a.view(np.uint32), b.view(np.uint32)
    bad = np.nonzero((a != b).reshape(len(a), -1).any(axis=1))[0]
    assert bad.size == 0, f"{what}: {bad.size} of {len(a)} differ, first at {bad[:5].tolist()}"


def test_the_fixture_names_config_1(golden):
    assert int(golden["width"]) * int(golden["height"]) == 65_536
    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    assert len(counts) == 8 and int(np.sum(counts)) == 12_312
    assert str(golden["scene_digest"]) == digest(tris9, counts), "the room generator changed"
    _same(golden["camera"], sc.camera, "camera")
    _same(golden["lights"], sc.lights, "lights")
    assert int(golden["capped_count"]) == 0 and int(golden["overflows"]) == 0


def test_the_oracle_reproduces_the_fixture(golden):
    rec = record(compute())
    for k in ("hierarchy_digest", "hits_digest", "vis_digest", "full_digest", "touched_digest", "sky_digest"):
        assert str(rec[k]) == str(golden[k]), f"{k} moved"
    for k in ("sample", "hits", "receivers", "vis", "full", "touched", "sky"):
        _same(rec[k], golden[k], k)
    # The image is lit and shadowed: most queries see the light, some none of it.
    assert (golden["vis"] > 0).mean() > 0.3 and (golden["vis"] == 0).mean() > 0.02


@pytest.mark.gpu
def test_the_library_reproduces_the_fixture(device, golden):
    import torch

    from continuum_amd import _lib
    from continuum_amd.kernel import Flat, Scene

    sc = room_scene()
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    assert digest(flat.nodes, flat.items, flat.blas, flat.blas_items) == str(golden["hierarchy_digest"])
    scene = Scene(device, flat)
    p = _lib.default_options(hierarchy=2)
    hits, vis, full, touched = scene.render_view(sc.camera, sc.width, sc.height, sc.lights, p)
    idx = golden["sample"]
    _same(hits[idx], golden["hits"], "hits")
    _same(vis[idx], golden["vis"], "vis")
    _same(full[idx], golden["full"], "full")
    _same(touched[idx], golden["touched"], "touched")
    assert digest(hits) == str(golden["hits_digest"])
    assert digest(vis) == str(golden["vis_digest"])
    assert digest(full) == str(golden["full_digest"])
    assert digest(touched) == str(golden["touched_digest"])

    dev = torch.device("cuda:0")
    n = sc.width * sc.height
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    dhits = torch.zeros((n, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((n, 8), dtype=torch.float32, device=dev)
    scene.primary_device(cam, sc.width, sc.height, lights, p, dhits, recv)
    torch.cuda.synchronize()
    r = recv.cpu().numpy()
    _same(r[idx], golden["receivers"], "receivers")
    sky = scene.sky(r, p)[0]
    _same(sky[idx], golden["sky"], "sky")
    assert digest(sky) == str(golden["sky_digest"])
