"""Neural radiance cache over the C ABI of include/krr_nrc.h.

Device buffers are passed as integer device pointers (kiraray_amd.hip
buffers, or torch tensors' data_ptr() when torch shares the runtime).
"""
import ctypes as C
import json as _json

import numpy as np

from ._abi import NRC_PARAMS, load_library


class NeuralRadianceCache:
    """5 x 64 ReLU MLP on BF16 matrix cores with FP32 accumulation."""

    PARAMS = NRC_PARAMS

    def __init__(self, config=None):
        from . import KrrError, _check  # noqa: F401
        self._lib = load_library()
        h = C.c_void_p()
        cfg = b"" if config is None else (_json.dumps(config).encode() if isinstance(config, dict) else config.encode())
        _check(self._lib.krr_nrc_create(cfg, C.byref(h)))
        self._h = h

    def _chk(self, rc):
        from . import _check
        return _check(rc, self._lib)

    def get_weights(self):
        w = np.zeros(self.PARAMS, np.float32)
        self._chk(self._lib.krr_nrc_get_weights(self._h, w.ctypes.data_as(C.c_void_p)))
        return w

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.PARAMS
        self._chk(self._lib.krr_nrc_set_weights(self._h, w.ctypes.data_as(C.c_void_p)))

    def encode(self, vertices_ptr, n, out_ptr, stream=0):
        self._chk(self._lib.krr_nrc_encode(self._h, C.c_void_p(vertices_ptr), int(n), C.c_void_p(out_ptr), C.c_void_p(stream)))

    def infer(self, vertices_ptr, n, radiance_ptr, stream=0):
        self._chk(self._lib.krr_nrc_infer(self._h, C.c_void_p(vertices_ptr), int(n), C.c_void_p(radiance_ptr), C.c_void_p(stream)))

    def gradient(self, vertices_ptr, targets_ptr, n, grad_ptr, stream=0):
        loss = C.c_float()
        self._chk(self._lib.krr_nrc_gradient(self._h, C.c_void_p(vertices_ptr), C.c_void_p(targets_ptr), int(n), C.c_void_p(grad_ptr),
                                             C.byref(loss), C.c_void_p(stream)))
        return loss.value

    def train_step(self, vertices_ptr, targets_ptr, n, stream=0):
        loss = C.c_float()
        self._chk(self._lib.krr_nrc_train_step(self._h, C.c_void_p(vertices_ptr), C.c_void_p(targets_ptr), int(n), C.byref(loss),
                                               C.c_void_p(stream)))
        return loss.value

    @property
    def steps(self):
        return int(self._lib.krr_nrc_steps(self._h))

    def close(self):
        if getattr(self, "_h", None):
            self._lib.krr_nrc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
