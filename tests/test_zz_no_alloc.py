"""tis_force_match_noalloc: force matching with caller-owned device buffers and
no allocation, its working sums in the library's fixed device array.  Checked
against tis_force_match (host buffers, working sums in memory the context
allocated) on two calls in a row, the second on a smaller set of snapshots, so
that a sum left over from the first call would show.

Device buffers come from the HIP runtime the library itself uses (the same
libamdhip64.so.7 handle), through ctypes."""
import ctypes as C

import numpy as np
import pytest

import force_match as FM
import tis_ctypes as T

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.live = []

    def alloc(self, n):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), max(1, n) * 8) == 0
        self.live.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        p = self.alloc(a.size)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.size * 8, H2D) == 0
        return p

    def get(self, p, n):
        out = np.empty(n)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, n * 8, D2H) == 0
        return out

    def free_all(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def _close(a, b, tol=1e-10):
    """Equal up to the summation order of FP64 atomics (norm-wise)."""
    a, b = np.asarray(a, dtype=float).ravel(), np.asarray(b, dtype=float).ravel()
    return np.linalg.norm(a - b) <= tol * max(np.linalg.norm(b), 1e-300)


def _compare(lib, hip, h, snaps, prec):
    d = h.d
    n = len(snaps) * d["n_particles"]
    tp = FM.Problem(lib, h, snaps, T.PREC_FP64)
    target = tp.run(snaps, FM.perturbed_params(d), np.zeros((n, 3)), grads=False)["forces"]
    tp.close()
    pb = FM.Problem(lib, h, snaps, prec)
    ref = pb.run(snaps, FM.base_params(d), target)
    sizes = {"forces": 3 * n, "pos": 3 * n, "hb_E": d["n_hb"], "st3_E": d["n_st3"], "st2_E": d["n_st2"],
             "ion_R": 3, "loss": 1}
    dev = {k: hip.alloc(v) for k, v in sizes.items()}
    tgt = hip.put(target)
    rc = lib.tis_force_match_noalloc(pb.eng.ctx, tgt, dev["loss"], dev["forces"], dev["pos"], dev["hb_E"],
                                     dev["st3_E"], dev["st2_E"], dev["ion_R"])
    T.check(lib, rc, "tis_force_match_noalloc")
    got = {k: hip.get(dev[k], v) for k, v in sizes.items()}
    print(len(snaps), "snapshots: loss", ref["loss"], got["loss"][0])
    assert got["loss"][0] == pytest.approx(ref["loss"], rel=1e-12)
    for k in ("forces", "pos", "hb_E", "st3_E", "st2_E", "ion_R"):
        print(k, np.linalg.norm(got[k] - np.asarray(ref[k]).ravel()), np.linalg.norm(ref[k]))
        assert np.abs(ref[k]).max() > 0 or k == "ion_R", k
        assert _close(got[k], ref[k]), k
    pb.close()
    hip.free_all()


@pytest.mark.parametrize("prec", [T.PREC_FP64, T.PREC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_matches_force_match_on_two_calls(lib, cfg1, prec):
    hip = Hip()
    h = T.SystemHandle(cfg1[0])
    d = h.d
    _compare(lib, hip, h, FM.snapshots(d, 4, sigma=0.5, seed=227), prec)   # first call
    _compare(lib, hip, h, FM.snapshots(d, 2, sigma=0.5, seed=375), prec)   # second, smaller
