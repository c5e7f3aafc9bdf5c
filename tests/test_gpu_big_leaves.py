"""GPU: triangle-hierarchy leaves larger than the builder makes. The ABI takes
a caller's own hierarchy (the reference's `flatten` arrays), whose leaves
may hold any number of triangles; a walk tests a leaf's triangles with the
whole packet and each lane then draws the ones it passed, in runs of 32.
The room's hierarchies with subtrees of 33-128 triangles collapsed into
single leaves must give the oracle's camera hits, masks and irradiance bit
for bit, with and without the prefilter's splats."""
import numpy as np
import pytest

from continuum_amd import _lib
from continuum_amd.kernel import Flat, Scene
from continuum_amd.scenes import meshes_to_arrays, room_scene
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _collapsed(flat, lo=33, hi=128):
    """The hierarchy with every subtree of lo..hi items (contiguous in
    blas_items) turned into one leaf over them; its descendants stay in the
    arrays, unreachable."""
    blas = flat.blas.copy()
    n = len(blas)
    first = np.zeros(n, np.int64)
    total = np.zeros(n, np.int64)
    for k in range(n - 1, -1, -1):  # children follow their parents
        a, b = int(blas[k, 8]), int(blas[k, 9])
        if b > 0:
            first[k], total[k] = a, b
        else:
            c0, c1 = a, a + 1
            if total[c0] >= 0 and total[c1] >= 0 and first[c0] + total[c0] == first[c1]:
                first[k], total[k] = first[c0], total[c0] + total[c1]
            else:
                first[k], total[k] = first[c0], -1  # not contiguous: never collapsed
    made = 0
    for k in range(n):
        if int(blas[k, 9]) == 0 and lo <= total[k] <= hi:
            blas[k, 8], blas[k, 9] = first[k], total[k]
            made += 1
    out = Flat(flat.tris, flat.ranges, blas, flat.blas_items, flat.blas_cov, flat.nodes, flat.items,
               flat.xforms, flat.object_cov)
    return out, made


@pytest.mark.parametrize("prefilter", [0.0, 1.0])
def test_leaves_of_more_than_32_triangles_match_the_oracle(device, prefilter):
    sc = room_scene(width=96, height=96)
    tris9, counts = meshes_to_arrays(sc.meshes)
    flat, made = _collapsed(Flat.build_arrays(tris9, counts))
    big = flat.blas[:, 9][(flat.blas[:, 9] > 32) & (flat.blas[:, 9] < 0x80000000)]
    assert made > 0 and big.size > 0 and int(big.max()) > 32, (made, big[:8])
    scene = Scene(device, flat)
    p = _lib.default_options(hierarchy=2, prefilter=prefilter)
    o = O.opts_array(hierarchy=2, prefilter=prefilter)
    hits, vis, full, touched = scene.render_view(sc.camera, sc.width, sc.height, sc.lights, p)
    rh, rrecv, _ = O.primary(flat, sc.camera, sc.width, sc.height, sc.lights)
    rv, rf, rt, _, capped, _ = O.irradiance(flat, rrecv, sc.lights, o)
    assert np.array_equal(hits, rh), "camera hits differ"
    ok = capped == 0
    assert np.array_equal(full[ok], rf[ok]) and np.array_equal(touched[ok], rt[ok]), "masks differ"
    assert np.array_equal(vis[ok].view(np.uint32), rv[ok].view(np.uint32)), "irradiance differs"
    assert int((touched != 0).sum()) > 0


def test_a_bake_over_leaves_of_more_than_32_triangles_matches_the_oracle(device):
    """The view cache's bake queues each lane's draws across a mesh's leaves
    only where the scene's leaves hold at most 32 triangles; with larger ones
    it draws leaf by leaf. The boulder world's hierarchies collapsed as
    above bake the oracle's pages bit for bit."""
    from continuum_amd.kernel import ViewCache
    from continuum_amd.scenes import vcache_world
    from tests.test_gpu_vcache import _keys, _same_state

    tris9, counts, pl, surf, lights, _ = vcache_world()
    flat, made = _collapsed(Flat.build_instanced(tris9, counts, pl, object_prefilter=True))
    assert made > 0 and int(flat.blas[:, 9].max()) > 32
    scene = Scene(device, flat)
    kw = dict(pages=1024, append_pages=1024, candidates=1024, base_texel=0.25, light_ratio=1, visible_frames=1)
    vc = ViewCache(scene, surf, lights, _lib.default_options(hierarchy=2, prefilter=1.0, object_cov=flat.object_cov),
                   **kw)
    ov = O.OracleViewCache(flat, surf, lights, O.opts_array(hierarchy=2, prefilter=1.0, object_cov=flat.object_cov),
                           **kw)
    keys = _keys(vc, pl)
    sa, sb = vc.request(keys), ov.request(keys)
    assert np.array_equal(sa, sb)
    assert vc.frame() == ov.frame()
    _same_state(vc, ov, "bake over large leaves")
