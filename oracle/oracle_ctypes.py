"""ctypes binding of the FP64 CPU oracle (oracle/tis_oracle.h).

Test infrastructure: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the CPU baseline.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

OE_E2_HB, OE_E3_HB, OE_E2_STACK, OE_E3_STACK, OE_WCA, OE_COUL_SR, OE_COUL_LR, OE_BOND, OE_ANGLE, OE_RG = range(10)
OE_N = 10
OE_NAMES = ["e2_hb", "e3_hb", "e2_stack", "e3_stack", "e_wca", "e_coul_sr", "e_coul_lr",
            "e_bond", "e_angle", "rg"]

_lib = None


def load(path=LIB_PATH):
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} not built (make -C oracle)")
    lib = C.CDLL(path)
    vp, dp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_int)
    lib.oracle_create.restype = vp
    lib.oracle_create.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_int]
    lib.oracle_destroy.argtypes = [vp]
    lib.oracle_destroy.restype = None
    lib.oracle_set_state.argtypes = [vp, dp, dp, dp, C.c_longlong, C.c_int]
    lib.oracle_get_state.argtypes = [vp, dp, dp, dp]
    lib.oracle_set_stream.argtypes = [vp, C.POINTER(C.c_int32), C.c_longlong]
    lib.oracle_set_stream_glibc.argtypes = [vp, C.c_uint, C.c_longlong]
    lib.oracle_stream_glibc_more.argtypes = [vp, C.c_longlong]
    lib.oracle_stream_cursor.argtypes = [vp]
    lib.oracle_stream_cursor.restype = C.c_longlong
    lib.oracle_eval.argtypes = [vp, C.c_uint, dp, dp]
    lib.oracle_step.argtypes = [vp, C.c_longlong, C.c_double]
    lib.oracle_pairs.argtypes = [vp, C.c_int, ip, C.c_longlong]
    lib.oracle_pairs.restype = C.c_longlong
    lib.oracle_hb_status.argtypes = [vp, ip]
    lib.oracle_rebuilds.argtypes = [vp]
    lib.oracle_rebuilds.restype = C.c_longlong
    lib.oracle_progress.argtypes = [vp]
    lib.oracle_progress.restype = C.c_double
    lib.oracle_klok.argtypes = [vp]
    lib.oracle_klok.restype = C.c_longlong
    lib.oracle_stack_energies.argtypes = [vp, dp]
    lib.oracle_hb_outcome.argtypes = [vp, ip]
    lib.oracle_hb_iterations.argtypes = [vp]
    lib.oracle_hb_iterations.restype = C.c_longlong
    _lib = lib
    return lib


def _d(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One replica on the CPU: lists, forces, steps in FP64."""

    def __init__(self, handle, r_sr=15.0, r_lr=53.0, skin=3.0, lr_every=10):
        self.lib = load()
        self.handle = handle   # keeps the tis_system arrays alive
        self.n = handle.d["n_particles"]
        self.n_hb = handle.d["n_hb"]
        self.n_st3 = handle.d["n_st3"]
        self.ctx = self.lib.oracle_create(C.cast(handle.ptr(), C.c_void_p), r_sr, r_lr, skin, lr_every)
        if not self.ctx:
            raise RuntimeError("oracle_create failed")

    def __del__(self):
        if getattr(self, "ctx", None):
            self.lib.oracle_destroy(self.ctx)
            self.ctx = None

    def set_state(self, pos, old=None, vel=None, klok=1, brute=False):
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
        old = None if old is None else np.ascontiguousarray(old, dtype=np.float64).reshape(-1)
        vel = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(-1)
        rc = self.lib.oracle_set_state(self.ctx, _d(pos), None if old is None else _d(old),
                                       None if vel is None else _d(vel), klok, int(brute))
        assert rc == 0

    def get_state(self):
        out = [np.zeros(3 * self.n) for _ in range(3)]
        self.lib.oracle_get_state(self.ctx, *[_d(a) for a in out])
        return [a.reshape(-1, 3) for a in out]

    def set_stream(self, raw):
        self._stream = np.ascontiguousarray(raw, dtype=np.int32)
        self.lib.oracle_set_stream(self.ctx, self._stream.ctypes.data_as(C.POINTER(C.c_int32)),
                                   len(self._stream))

    def set_stream_glibc(self, seed, n):
        assert self.lib.oracle_set_stream_glibc(self.ctx, seed, n) == 0

    def stream_glibc_more(self, n):
        """Keeps the unconsumed values and appends the next n of the glibc sequence."""
        assert self.lib.oracle_stream_glibc_more(self.ctx, n) == 0

    def cursor(self):
        return self.lib.oracle_stream_cursor(self.ctx)

    def eval(self, mask):
        f = np.zeros(3 * self.n)
        e = np.zeros(OE_N)
        rc = self.lib.oracle_eval(self.ctx, mask, _d(f), _d(e))
        return rc, f.reshape(-1, 3), dict(zip(OE_NAMES, e))

    def step(self, n, noise_scale=1.0):
        return self.lib.oracle_step(self.ctx, n, noise_scale)

    def pairs(self, which):
        n = self.lib.oracle_pairs(self.ctx, which, None, 0)
        out = np.zeros(2 * n, dtype=np.int32)
        self.lib.oracle_pairs(self.ctx, which, out.ctypes.data_as(C.POINTER(C.c_int)), n)
        return out.reshape(-1, 2)

    def hb_status(self):
        out = np.zeros(self.n_hb, dtype=np.int32)
        self.lib.oracle_hb_status(self.ctx, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def hb_outcome(self):
        out = np.zeros(self.n_hb, dtype=np.int32)
        self.lib.oracle_hb_outcome(self.ctx, out.ctypes.data_as(C.POINTER(C.c_int)))
        return out

    def stack_energies(self):
        out = np.zeros(self.n_st3)
        self.lib.oracle_stack_energies(self.ctx, _d(out))
        return out

    def rebuilds(self):
        return self.lib.oracle_rebuilds(self.ctx)

    def progress(self):
        return self.lib.oracle_progress(self.ctx)

    def klok(self):
        return self.lib.oracle_klok(self.ctx)

    def hb_iterations(self):
        return self.lib.oracle_hb_iterations(self.ctx)


def glibc_stream(seed, n):
    """n raw values of glibc rand() after srand(seed), from the C library itself."""
    libc = C.CDLL("libc.so.6")
    libc.srand(C.c_uint(seed))
    return np.array([libc.rand() for _ in range(n)], dtype=np.int32)
