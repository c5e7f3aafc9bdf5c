"""GPU: camera receivers with interpolated shading normals (the reference's
`shading_normal`, mask_cones.wgsl :2047-2058, as `per_pixel_hit_color`
:2280 shades with it) against the oracle, bit for bit: the receivers the
camera pass writes, then every query's masks and irradiance (capped walks
excepted, as everywhere), on the rock field with smooth corner normals and
on an instanced world of rotated, scaled boulders."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, Scene
from continuum_amd.scenes import camera_uniform, meshes_to_arrays, rock_field
from continuum_amd.surfaces import smooth_normals
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _per_mesh_normals(tris12, ranges):
    """Each mesh's smooth corner normals (mesh.rs:330-376), in the scene's
    triangle order."""
    t9 = tris12.reshape(-1, 3, 4)[:, :, :3].reshape(-1, 9)
    out = np.zeros((len(t9), 9), np.float32)
    for first, count, _, _ in ranges:
        if count:
            out[first:first + count] = smooth_normals(t9[first:first + count])
    return out


def _camera_pass(scene, cam, w, h, lights, opts):
    import torch

    dev = torch.device("cuda:0")
    hits = torch.zeros((w * h, 4), dtype=torch.int32, device=dev)
    recv = torch.zeros((w * h, 8), dtype=torch.float32, device=dev)
    scene.primary_device(torch.from_numpy(cam).to(dev), w, h, torch.from_numpy(np.ascontiguousarray(lights)).to(dev),
                         opts, hits, recv)
    torch.cuda.synchronize()
    return hits.cpu().numpy().view(np.uint32), recv.cpu().numpy()


def _check(scene, oflat, normals, cam, w, h, lights, opts, o):
    hits, recv = _camera_pass(scene, cam, w, h, lights, opts)
    rh, rrecv, _ = O.primary(oflat, cam, w, h, lights, normals=normals)
    assert np.array_equal(hits, rh)
    assert np.array_equal(_u32(recv), _u32(rrecv)), int((_u32(recv) != _u32(rrecv)).any(axis=1).sum())
    g = scene.query(recv, lights, opts)
    r = O.irradiance(oflat, rrecv, lights, o)
    ok = r[4] == 0
    bad = (_u32(g[0]) != _u32(r[0])) | (g[1] != r[1]) | (g[2] != r[2])
    assert not (bad & ok).any(), int((bad & ok).sum())
    return recv, g


def test_rock_field_with_smooth_normals(device):
    sc = rock_field(width=160, height=90)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat = Flat.build_arrays(tris9, counts)
    scene = Scene(device, flat)
    normals = _per_mesh_normals(flat.tris, flat.ranges)
    opts = _lib.default_options(hierarchy=2, prefilter=sc.prefilter)
    o = O.opts_array(hierarchy=2, prefilter=sc.prefilter)
    flat_recv, flat_q = _check(scene, flat, None, sc.camera, 160, 90, sc.lights, opts, o)
    scene.set_normals(normals)
    recv, q = _check(scene, flat, normals, sc.camera, 160, 90, sc.lights, opts, o)
    lit = recv[:, 7] >= 0
    # smooth normals differ from the faces' on the rocks, and change the light
    assert (np.abs(recv[lit, 4:7] - flat_recv[lit, 4:7]).max(axis=1) > 1e-3).mean() > 0.2
    assert not np.array_equal(_u32(q[0]), _u32(flat_q[0]))
    # the shading points themselves stay where the face normal puts them
    assert np.array_equal(_u32(recv[:, 0:3]), _u32(flat_recv[:, 0:3]))
    scene.set_normals(None)  # back to face normals
    _, again = _camera_pass(scene, sc.camera, 160, 90, sc.lights, opts)
    assert np.array_equal(_u32(again), _u32(flat_recv))


def test_instanced_boulders_with_smooth_normals(device):
    from test_gpu_instanced import _world

    tris9, counts, pl, lights = _world()
    f = Flat.build_instanced(tris9, counts, pl, object_prefilter=True)
    o_flat = O.OracleInstancedFlat(tris9, counts, pl, object_prefilter=True)
    scene = Scene(device, f)
    firsts = np.concatenate([[0], np.cumsum(counts.astype(np.int64))[:-1]]).astype(np.int64)
    normals = _per_mesh_normals(f.tris, [(int(a), int(b), 0, 0) for a, b in zip(firsts, counts)])
    scene.set_normals(normals)
    cam = camera_uniform([-24.0, -24.0, 2.5], [5.0, 5.0, 0.5], 55.0, 96, 64)
    opts = _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    o = O.opts_array(hierarchy=2, prefilter=1.0, object_cov=f.object_cov)
    recv, _ = _check(scene, o_flat, normals, cam, 96, 64, lights, opts, o)
    assert (recv[:, 7] >= 0).sum() > 3000
