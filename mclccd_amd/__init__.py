"""MI355X-native continuous collision detection: the ``mcl::BVHTree<double,DIM>``
path of mattoverby/mclccd (swept-box BVH self-traversal, ACCD vertex-face and
edge-edge narrow phase, discrete tri-tri / edge-edge tests) as gfx950 HIP
kernels behind the C ABI of include/mclg.h.

Importing the package does not load the native library; ``mclccd_amd.tree``
(and everything that runs on the GPU) does, and raises if it is not built.
"""

__all__ = ["BVHTree", "Options"]


def __getattr__(name):
    if name in ("BVHTree", "Options"):
        from . import tree

        return getattr(tree, name)
    raise AttributeError(name)
