"""GPU: cnq_sky_irradiance_device_noalloc, the sky launch that allocates
nothing (its face values go to the library's static per-GPU array), called
through the C ABI, against cnq_sky_irradiance_device with the caller's own
face buffer: sums and full masks bit for bit, on two calls in a row, the
second with fewer receivers than the first; sizes past its maximum are
refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd._lib import CnqError, check, lib
from continuum_amd.kernel import Flat, Scene
from continuum_amd.scenes import meshes_to_arrays, rock_field

pytestmark = pytest.mark.gpu


def _noalloc(scene, recv, n, opts, sky, out, full, stream):
    check(lib().cnq_sky_irradiance_device_noalloc(scene.handle, recv.data_ptr(), n, C.byref(opts), C.byref(sky),
                                                  out.data_ptr(), full.data_ptr() if full is not None else None,
                                                  stream.cuda_stream))


def test_the_allocation_free_sky_equals_the_sky_launch(device):
    import torch

    sc = rock_field(width=96, height=64)  # open to the sky
    tris9, counts = meshes_to_arrays(sc.meshes)
    scene = Scene(device, Flat.build_arrays(tris9, counts))
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    sky = _lib.default_sky_options()
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    px = sc.width * sc.height
    cam = torch.from_numpy(sc.camera).to(dev)
    lights = torch.from_numpy(np.ascontiguousarray(sc.lights)).to(dev)
    hits = torch.zeros((px, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((px, 8), dtype=torch.float32, device=dev)
    scene.primary_device(cam, sc.width, sc.height, lights, opts, hits, recv)
    torch.cuda.synchronize()
    assert px > 4000
    for n in (px, 1000):  # the second call smaller than the first
        r = recv[:n].contiguous()
        want = torch.zeros(n, dtype=torch.float32, device=dev)
        want_faces = torch.zeros(6 * n, dtype=torch.float32, device=dev)
        want_full = torch.zeros(6 * n, dtype=torch.int64, device=dev)
        scene.sky_device(r, opts, want, want_faces, want_full, sky=sky, stream=stream)
        got = torch.full((n,), -1.0, dtype=torch.float32, device=dev)
        got_full = torch.zeros(6 * n, dtype=torch.int64, device=dev)
        _noalloc(scene, r, n, opts, sky, got, got_full, stream)
        torch.cuda.synchronize()
        a, b = got.cpu().numpy(), want.cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), n
        assert np.array_equal(got_full.cpu().numpy(), want_full.cpu().numpy()), n
        assert (b > 0).sum() > n // 2
    # without the full masks too
    got = torch.zeros(1000, dtype=torch.float32, device=dev)
    _noalloc(scene, recv[:1000].contiguous(), 1000, opts, sky, got, None, stream)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), a.view(np.uint32))
    # past the maximum: refused on the host
    with pytest.raises(CnqError):
        _noalloc(scene, recv, (1 << 20) + 1, opts, sky, got, None, stream)
