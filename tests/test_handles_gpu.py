"""Object lifetime across the C ABI: objects created from a context may be
destroyed after the context (a garbage collector at interpreter exit destroys
them in any order); the context stays alive until its last object is gone."""
import gc

import pytest

pytestmark = pytest.mark.gpu


def test_objects_outlive_their_context(ctx):
    from untangling26_amd.lbvh import LBVH, TREE_EDGE, TREE_FACE
    from untangling26_amd.runtime import Context

    c = Context(0)
    trees = [LBVH(c, TREE_FACE, 64), LBVH(c, TREE_EDGE, 96)]
    c.close()  # the caller's reference goes first
    del trees
    gc.collect()
    # the session context is unaffected
    ctx.synchronize()
