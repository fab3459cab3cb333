"""final184_amd — MI355X-native voxel global illumination (voxelize -> mips -> cone trace).

Submodules:
  vgi       ctypes binding of libvgi.so (the C ABI declared in include/vgi.h)
  scenegen  seeded synthetic workloads (libvgi_gen.so)
  build     explicit hipcc / g++ build recipes
  serving   continuous-batching request loop over resident voxel volumes
"""
__version__ = "0.1.0"
