"""HIP backend: ctypes bindings over libsmoke_hip.so (gfx950 kernels)."""
