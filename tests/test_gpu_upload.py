"""Object-space upload on the GPU (aur_geometry_upload_objects): transformed
and padded geometry bit-exact with the oracle, for ascending offsets (binary
search) and unsorted ones (the reference's linear scan); the tree built from
it equals the tree of the oracle-transformed geometry uploaded directly."""
from __future__ import annotations

import numpy as np
import pytest

from aurora_amd import native, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shuffle", [False, True])
def test_objects_bitexact(gpu_ctx, oracle, shuffle):
    v, n, ids, xf = scenes.instanced_spheres(seed=11, n_objects=16, subdivisions=2)
    assert v.shape[0] % 2 == 1
    if shuffle:
        xf = xf[np.random.default_rng(0).permutation(xf.shape[0])]
    gpu_ctx.upload_objects(v, n, xf, shader_ids=ids)
    gpu_ctx.rebuild_nmh()
    perm = gpu_ctx.read_permutation()
    geo = gpu_ctx.read_geometry()
    ov, on = oracle.transform_objects(v, n, xf)
    assert perm.shape[0] == v.shape[0] + 1
    assert np.array_equal(geo["vertices"].view(np.uint32), ov[perm].view(np.uint32))
    assert np.array_equal(geo["normals"].view(np.uint32), on[perm].view(np.uint32))
    # the same tree as uploading the oracle-transformed geometry directly
    ids_p = np.concatenate([ids, ids[-1:]])
    gpu_ctx.upload_geometry(ov, on, None, ids_p)
    gpu_ctx.rebuild_nmh()
    assert np.array_equal(gpu_ctx.read_permutation(), perm)
    assert np.array_equal(oracle.nmh_build(ov), perm)


def test_objects_reject_overflow(gpu_ctx):
    v, n, ids, xf = scenes.instanced_spheres(seed=1, n_objects=2, subdivisions=1)
    xf = xf.copy()
    xf[1]["world"][0, 0] = 3e38  # x * 3e38 overflows for |x| > ~1.1
    xf[1]["world"][1, 0] = 3e38
    with pytest.raises(native.AuroraError):
        gpu_ctx.upload_objects(v, n, xf, shader_ids=ids)
