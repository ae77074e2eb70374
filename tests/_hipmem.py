"""Device buffers for GPU tests from the HIP runtime the engine library
itself links (libamdhip64.so.7, already loaded with the library), through
ctypes.  Tests do not use torch for device memory: its wheel carries a HIP
runtime of its own, and two runtimes in one process are avoided."""
import ctypes as C

import numpy as np

import turboransac_amd as tr

tr.load_library()  # loads /opt/rocm's HIP runtime first
_hip = C.CDLL("libamdhip64.so.7")
_hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
_hip.hipFree.argtypes = [C.c_void_p]
_hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
_hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
_H2D, _D2H = 1, 2


class DeviceArray:
    """nbytes of device memory, freed on close()."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        rc = _hip.hipMalloc(C.byref(p), C.c_size_t(max(1, int(nbytes))))
        if rc != 0:
            raise RuntimeError(f"hipMalloc({nbytes}) failed: {rc}")
        self.ptr = p.value
        self.nbytes = int(nbytes)

    @classmethod
    def from_numpy(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.nbytes)
        rc = _hip.hipMemcpy(C.c_void_p(d.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, _H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to device failed: {rc}")
        return d

    def zero(self):
        if _hip.hipMemset(C.c_void_p(self.ptr), 0, self.nbytes) != 0:
            raise RuntimeError("hipMemset failed")
        return self

    def to_numpy(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype=dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("read past the buffer")
        rc = _hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, _D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy to host failed: {rc}")
        return out

    def close(self):
        if self.ptr:
            _hip.hipFree(C.c_void_p(self.ptr))
            self.ptr = 0
