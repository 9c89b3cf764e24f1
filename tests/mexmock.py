"""Python side of the mock MEX harness (tests/mex): drives the engine's MEX
gateways (mrilab_amd/mex/DoXScanAtY.cpp, built against the mock runtime
into tests/mex/_build/libmex_<entry>.so) the way MATLAB does: the seven
scan structs as globals, mexFunction called with no arguments, DoExtPlugin
and DoGPUFetch evaluated by a handler. Test infrastructure only."""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "mex", "_build")
CLS = {np.dtype(np.float64): 6, np.dtype(np.float32): 7, np.dtype(np.int32): 12}
DTYPE = {6: np.float64, 7: np.float32, 12: np.int32}
CHAR = 4
EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_char_p)


def path(entry: str) -> str:
    return os.path.join(BUILD, f"libmex_{entry}.so")


class MockMex:
    def __init__(self, entry: str):
        from mrilab_amd import engine
        engine.load()  # the HIP runtime and the engine library first
        self.lib = C.CDLL(path(entry))
        L = self.lib
        L.mock_set.argtypes = [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.POINTER(C.c_int64), C.c_void_p]
        L.mock_get.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int64)]
        L.mock_get.restype = C.c_void_p
        L.mock_remove.argtypes = [C.c_char_p, C.c_char_p]
        L.mock_set_eval.argtypes = [EVAL_FN]
        L.mock_error.restype = C.c_char_p
        L.mock_error_id.restype = C.c_char_p
        L.mock_reset()
        self._eval = None

    def set(self, var: str, field: str, value) -> None:
        """var.field = value: a numpy array keeps its dtype and shape
        (column-major), a str becomes a char array, a Python float a double."""
        if isinstance(value, str):
            a = np.frombuffer(value.encode("utf-16-le"), dtype=np.uint16)
            dims = (C.c_int64 * 2)(1, a.size)
            self.lib.mock_set(var.encode(), field.encode(), CHAR, 2, dims, a.ctypes.data)
            return
        a = np.asarray(value)
        if a.dtype not in CLS:
            a = a.astype(np.float64)
        shape = a.shape if a.ndim >= 2 else ((1, a.size) if a.ndim == 1 else (1, 1))
        a = np.asfortranarray(a)
        dims = (C.c_int64 * len(shape))(*shape)
        self.lib.mock_set(var.encode(), field.encode(), CLS[a.dtype], len(shape), dims, a.ctypes.data)

    def get(self, var: str, field: str, shape=None):
        """A copy of var.field (column-major, reshaped to ``shape`` if given)."""
        cls, n = C.c_int(), C.c_int64()
        p = self.lib.mock_get(var.encode(), field.encode(), C.byref(cls), C.byref(n))
        if not p:
            return None
        dt = np.dtype(DTYPE[cls.value])
        a = np.frombuffer((C.c_char * (n.value * dt.itemsize)).from_address(p), dtype=dt).copy()
        return a.reshape(shape, order="F") if shape is not None else a

    def address(self, var: str, field: str) -> int:
        return int(self.lib.mock_get(var.encode(), field.encode(), None, None) or 0)

    def remove(self, var: str, field: str) -> None:
        self.lib.mock_remove(var.encode(), field.encode())

    def load(self, sc) -> None:
        """The globals from a workloads.Scan (its dicts, as DoDataTypeConv.m
        leaves them; VCtl strings as MATLAB chars)."""
        for var, d in zip(("VObj", "VMag", "VCoi", "VCtl", "VSeq", "VVar", "VSig"), sc.structs):
            for k, v in d.items():
                if isinstance(v, (np.ndarray, np.generic, str, float, int)):
                    if isinstance(v, int) and not isinstance(v, bool):
                        v = float(v)
                    self.set(var, k, v)

    def on_eval(self, handler) -> None:
        """handler(command: str) -> int for mexEvalString."""
        def _h(cmd):
            try:
                return int(handler(cmd.decode()))
            except Exception:  # noqa: BLE001 - a MATLAB error in the plugin
                return 1
        self._eval = EVAL_FN(_h)
        self.lib.mock_set_eval(self._eval)

    def run(self):
        """mexFunction(); returns None or (error id, message)."""
        if self.lib.mock_run() == 0:
            return None
        return self.lib.mock_error_id().decode(), self.lib.mock_error().decode()
