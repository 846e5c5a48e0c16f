"""MI355X-native RGBDFusion dense path (voxel-hash TSDF fusion, ray casting,
marching cubes, dense ICP + photometric tracking) behind a C ABI
(include/yas.h, libyas.so).

Submodules: ``native`` (ctypes loader of libyas.so), ``rgbdfusion``
(reference-shaped classes over torch device tensors), ``synth`` (synthetic
RGB-D sequences), ``build`` (in-tree build of the native libraries).
"""
__all__ = ["native", "rgbdfusion", "synth", "build"]
