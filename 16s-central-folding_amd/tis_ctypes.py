"""ctypes binding of the C ABI in include/tis_gpu.h.

Thin plumbing for tests, smoke() and bench.py: it loads the in-tree
libtis_mi355x.so, mirrors the tis_system / tis_options / tis_energies structs
and converts systems to and from numpy dictionaries.  No physics lives here.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# the engine library; TIS_LIB names another build of it
LIB_PATH = os.environ.get("TIS_LIB", os.path.join(HERE, "libtis_mi355x.so"))
DATA_16SCD = os.path.join(HERE, "data", "16SCD")

NTYPES = 9
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class TisSystem(C.Structure):
    _fields_ = [
        ("n_particles", C.c_int), ("n_beads", C.c_int), ("n_units", C.c_int),
        ("n_bonds", C.c_int), ("n_angles", C.c_int),
        ("n_st3", C.c_int), ("n_st2", C.c_int),
        ("n_hb", C.c_int), ("n_hb_sites", C.c_int), ("n_hb_incid", C.c_int),
        ("n_mg", C.c_int), ("n_cl", C.c_int), ("n_k", C.c_int),
        ("box", C.c_double * 3), ("T", C.c_double), ("lB", C.c_double), ("h", C.c_double),
        ("recentre", C.c_int),
        ("type_mass", C.c_double * NTYPES), ("type_R", C.c_double * NTYPES),
        ("type_eps", C.c_double * NTYPES), ("type_q", C.c_double * NTYPES),
        ("type", _ip), ("cls", _ip),
        ("pos", _dp), ("pos_old", _dp), ("vel", _dp),
        ("unit_bead0", _ip), ("unit_bead1", _ip), ("unit_kind", C.POINTER(C.c_char)),
        ("bond_ij", _ip), ("bond_kr", _dp), ("angle_ijk", _ip), ("angle_kt", _dp),
        ("st3_beads", _ip), ("st3_geo", _dp), ("st3_E", _dp), ("st3_face", _ip), ("st3_disable", _ip),
        ("st2_beads", _ip), ("st2_r0", _dp), ("st2_E", _dp), ("st_excess", _ip),
        ("hb_beads", _ip), ("hb_geo", _dp), ("hb_E", _dp), ("hb_dode", _ip),
        ("hb_site_off", _ip), ("hb_site_idx", _ip),
        ("site_hb_off", _ip), ("site_hb_idx", _ip), ("site_valence", _ip),
    ]


class BuildArgs(C.Structure):
    _fields_ = [("ss_D", C.c_double), ("hs_D", C.c_double), ("st_D", C.c_double),
                ("T", C.c_double), ("c_mg", C.c_double), ("c_k", C.c_double),
                ("seed", C.c_int), ("box", C.c_double)]


class ChainFiles(C.Structure):
    _fields_ = [("pdb", C.c_char_p), ("bonds", C.c_char_p), ("stacks", C.c_char_p), ("maxikey", C.c_char_p),
                ("double_face_res", _ip), ("n_double_face", C.c_int), ("recentre_bead", C.c_int)]


class SyntheticArgs(C.Structure):
    _fields_ = [("n_nt", C.c_int), ("stack_frac", C.c_double), ("n_hb", C.c_int),
                ("radius", C.c_double), ("build", BuildArgs)]


class Options(C.Structure):
    _fields_ = [("precision", C.c_int), ("rng", C.c_int), ("seed", C.c_uint64),
                ("noise_scale", C.c_double), ("skin", C.c_double), ("r_sr", C.c_double),
                ("r_lr", C.c_double), ("lr_every", C.c_int), ("use_graph", C.c_int),
                ("graph_steps", C.c_int), ("exchange_every", C.c_longlong), ("pair_lanes", C.c_int),
                ("pair_block", C.c_int), ("list_waves", C.c_int), ("hb_groups", C.c_int),
                ("hb_table_lag", C.c_int), ("persistent", C.c_int)]


class Tuning(C.Structure):
    _fields_ = [("ms_per_step_start", C.c_double), ("ms_per_step_best", C.c_double), ("trials", C.c_int)]


class Energies(C.Structure):
    _fields_ = [("e2_hb", C.c_double), ("e3_hb", C.c_double), ("e2_stack", C.c_double),
                ("e3_stack", C.c_double), ("e_wca", C.c_double), ("e_coul_sr", C.c_double),
                ("e_coul_lr", C.c_double), ("e_bond", C.c_double), ("e_angle", C.c_double),
                ("rg", C.c_double), ("klok", C.c_longlong), ("n_rebuilds", C.c_longlong),
                ("progress", C.c_double), ("n_sr_pairs", C.c_longlong),
                ("n_lr_pairs", C.c_longlong), ("n_hb_active", C.c_longlong), ("hb_rounds", C.c_longlong)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Observables(C.Structure):
    _fields_ = [("n_samples", C.c_longlong), ("cut", C.c_double), ("rg", C.c_double), ("rg2", C.c_double),
                ("asphericity", C.c_double), ("shape", C.c_double), ("e2_hb", C.c_double),
                ("e3_hb", C.c_double), ("e2_stack", C.c_double), ("e3_stack", C.c_double),
                ("e_wca", C.c_double), ("e_coul_sr", C.c_double), ("e_coul_lr", C.c_double),
                ("e_bond", C.c_double), ("e_angle", C.c_double), ("bound_mg", C.c_double),
                ("bound_cl", C.c_double), ("bound_k", C.c_double)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


R_KT = 7.2684   # P-ion contact distance of the reference's analysis (dcd_traj_SaltCount_lB.py:9)


def contact_concentration(mean_count, cut=R_KT):
    """Mean ion count within `cut` of a phosphate as a local concentration in
    mol/L (analysis/salt_nt/integrate_num_ions.py:51-55)."""
    return np.asarray(mean_count) / (4.0 / 3.0 * np.pi * cut ** 3 * 6.02214 / 1e4)


PREC_FP64, PREC_MIXED = 0, 1
RNG_PHILOX, RNG_STREAM = 0, 1
TERM_STACK, TERM_HB, TERM_BACKBONE, TERM_SR, TERM_LR, TERM_ALL = 1, 2, 4, 8, 16, 31

# (field, dtype, length-expression) for every array of tis_system.
_ARRAYS = [
    ("type", np.int32, "N"), ("cls", np.int32, "N"),
    ("pos", np.float64, "3N"), ("pos_old", np.float64, "3N"), ("vel", np.float64, "3N"),
    ("unit_bead0", np.int32, "U"), ("unit_bead1", np.int32, "U"), ("unit_kind", np.uint8, "U"),
    ("bond_ij", np.int32, "2B"), ("bond_kr", np.float64, "2B"),
    ("angle_ijk", np.int32, "3A"), ("angle_kt", np.float64, "2A"),
    ("st3_beads", np.int32, "6S3"), ("st3_geo", np.float64, "6S3"), ("st3_E", np.float64, "S3"),
    ("st3_face", np.int32, "2S3"), ("st3_disable", np.int32, "2S3"),
    ("st2_beads", np.int32, "7S2"), ("st2_r0", np.float64, "S2"), ("st2_E", np.float64, "S2"),
    ("st_excess", np.int32, "U"),
    ("hb_beads", np.int32, "6H"), ("hb_geo", np.float64, "6H"), ("hb_E", np.float64, "H"),
    ("hb_dode", np.int32, "H"), ("hb_site_off", np.int32, "H+1"), ("hb_site_idx", np.int32, "I"),
    ("site_hb_off", np.int32, "S+1"), ("site_hb_idx", np.int32, "I"), ("site_valence", np.int32, "S"),
]
_SCALARS = ["n_particles", "n_beads", "n_units", "n_bonds", "n_angles", "n_st3", "n_st2",
            "n_hb", "n_hb_sites", "n_hb_incid", "n_mg", "n_cl", "n_k", "T", "lB", "h", "recentre"]


def _length(expr, d):
    env = {"N": d["n_particles"], "U": d["n_units"], "B": d["n_bonds"], "A": d["n_angles"],
           "S3": d["n_st3"], "S2": d["n_st2"], "H": d["n_hb"], "I": d["n_hb_incid"],
           "S": d["n_hb_sites"]}
    for key in ("S3", "S2"):
        if expr.endswith(key) and expr != key:
            return int(expr[:-len(key)]) * env[key]
    if expr in env:
        return env[expr]
    if expr.endswith("+1"):
        return env[expr[:-2]] + 1
    return int(expr[:-1]) * env[expr[-1]]


def load_library(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run make -C 16s-central-folding_amd first")
    lib = C.CDLL(path)
    lib.tis_last_error.restype = C.c_char_p
    lib.tis_build_16scd.argtypes = [C.c_char_p, C.POINTER(BuildArgs), C.POINTER(C.POINTER(TisSystem)),
                                    C.POINTER(C.c_longlong)]
    lib.tis_build_chain.argtypes = [C.POINTER(ChainFiles), C.POINTER(BuildArgs), C.POINTER(C.POINTER(TisSystem)),
                                    C.POINTER(C.c_longlong)]
    lib.tis_build_synthetic.argtypes = [C.c_char_p, C.POINTER(SyntheticArgs), C.POINTER(C.POINTER(TisSystem))]
    lib.tis_system_free.argtypes = [C.POINTER(TisSystem)]
    lib.tis_system_free.restype = None
    lib.tis_glibc_rand_fill.argtypes = [C.c_uint, C.c_longlong, C.POINTER(C.c_int32), C.c_longlong]
    if hasattr(lib, "tis_create"):
        lib.tis_default_options.argtypes = [C.POINTER(Options)]
        lib.tis_default_options.restype = None
        lib.tis_create.argtypes = [C.POINTER(TisSystem), C.c_int, C.POINTER(Options), C.POINTER(C.c_void_p)]
        lib.tis_destroy.argtypes = [C.c_void_p]
        lib.tis_destroy.restype = None
        lib.tis_set_state.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, C.c_longlong]
        lib.tis_get_state.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp]
        lib.tis_set_step_counter.argtypes = [C.c_void_p, C.c_int, C.c_longlong]
        lib.tis_set_rng_epoch.argtypes = [C.c_void_p, C.c_int, C.c_uint]
        lib.tis_prepare_graph.argtypes = [C.c_void_p, C.c_int]
        lib.tis_trace_hb.argtypes = [C.c_void_p, C.c_int]
        lib.tis_get_hb_trace.argtypes = [C.c_void_p, C.POINTER(C.c_ulonglong), C.c_longlong]
        lib.tis_get_launch_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                              C.POINTER(C.c_longlong)]
        lib.tis_step.argtypes = [C.c_void_p, C.c_longlong]
        lib.tis_set_persistent.argtypes = [C.c_void_p, C.c_int]
        lib.tis_get_persistent_counts.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_longlong),
                                                  C.POINTER(C.c_longlong), C.c_char_p, C.c_int]
        lib.tis_get_energies.argtypes = [C.c_void_p, C.c_int, C.POINTER(Energies)]
        lib.tis_get_stack_energies.argtypes = [C.c_void_p, C.c_int, _dp, C.c_int]
        lib.tis_eval_forces.argtypes = [C.c_void_p, C.c_int, C.c_uint, _dp, C.POINTER(Energies)]
        lib.tis_get_pairs.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, C.c_longlong,
                                      C.POINTER(C.c_longlong)]
        lib.tis_get_hb_status.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        lib.tis_get_hb_outcome.argtypes = [C.c_void_p, C.c_int, _ip, C.c_int]
        lib.tis_set_uniform_stream.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int32), C.c_longlong]
        lib.tis_get_stream_cursor.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_longlong)]
        lib.tis_kernel_times.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        lib.tis_set_profiling.argtypes = [C.c_void_p, C.c_int]
        lib.tis_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        lib.tis_philox4x32_10.restype = None
        lib.tis_exchange_now.argtypes = [C.c_void_p]
        lib.tis_get_exchange_stats.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _dp,
                                               C.c_int]
        lib.tis_get_walkers.argtypes = [C.c_void_p, _ip, C.c_int]
        lib.tis_force_match.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.tis_force_match_noalloc.argtypes = [C.c_void_p] + [C.c_void_p] * 8
        lib.tis_set_force_field.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        lib.tis_autotune.argtypes = [C.POINTER(TisSystem), C.c_int, C.POINTER(Options), C.c_int, C.POINTER(Tuning)]
        lib.tis_observe.argtypes = [C.c_void_p, C.c_longlong, C.c_double]
        lib.tis_observe_now.argtypes = [C.c_void_p]
        lib.tis_get_observables.argtypes = [C.c_void_p, C.c_int, C.POINTER(Observables), _dp, C.c_int, _dp,
                                            C.c_int, _dp, C.c_int]
    return lib


class TisError(RuntimeError):
    def __init__(self, lib, code, what):
        msg = lib.tis_last_error().decode(errors="replace")
        super().__init__(f"{what} failed with {code}: {msg}")
        self.code = code


def check(lib, code, what):
    if code != 0:
        raise TisError(lib, code, what)


def system_to_dict(sp):
    """Deep-copies a tis_system (pointer or struct) into a dict of numpy arrays."""
    s = sp.contents if hasattr(sp, "contents") else sp
    d = {k: getattr(s, k) for k in _SCALARS}
    d["box"] = np.array(list(s.box))
    for k in ("type_mass", "type_R", "type_eps", "type_q"):
        d[k] = np.array(list(getattr(s, k)))
    for name, dt, expr in _ARRAYS:
        n = _length(expr, d)
        ptr = getattr(s, name)
        if n == 0:
            d[name] = np.zeros(0, dtype=dt)
            continue
        raw = C.cast(ptr, C.POINTER(C.c_ubyte * (n * np.dtype(dt).itemsize))).contents
        d[name] = np.frombuffer(bytes(raw), dtype=dt).copy()
    return d


class SystemHandle:
    """A tis_system struct whose arrays are kept alive by numpy buffers."""

    def __init__(self, d):
        self.d = {k: (np.ascontiguousarray(v) if isinstance(v, np.ndarray) else v) for k, v in d.items()}
        s = TisSystem()
        for k in _SCALARS:
            setattr(s, k, self.d[k])
        s.box = (C.c_double * 3)(*[float(v) for v in self.d["box"]])
        for k in ("type_mass", "type_R", "type_eps", "type_q"):
            setattr(s, k, (C.c_double * NTYPES)(*[float(v) for v in self.d[k]]))
        for name, dt, expr in _ARRAYS:
            arr = np.ascontiguousarray(self.d[name], dtype=dt)
            self.d[name] = arr
            ctype = C.c_char if name == "unit_kind" else (C.c_int if dt == np.int32 else C.c_double)
            setattr(s, name, arr.ctypes.data_as(C.POINTER(ctype)))
        self.struct = s

    def ptr(self):
        return C.pointer(self.struct)


def build_16scd(lib, T=0.615, c_mg=0.005, c_k=0.050, seed=1, box=350.0,
                ss_D=0.850, hs_D=3.975, st_D=6.5, data_dir=DATA_16SCD):
    """Runs the host set-up (tis_build_16scd) and returns (dict, rand draws used)."""
    args = BuildArgs(ss_D, hs_D, st_D, T, c_mg, c_k, seed, box)
    out = C.POINTER(TisSystem)()
    draws = C.c_longlong(0)
    check(lib, lib.tis_build_16scd(data_dir.encode(), C.byref(args), C.byref(out), C.byref(draws)),
          "tis_build_16scd")
    try:
        d = system_to_dict(out)
    finally:
        lib.tis_system_free(out)
    return d, draws.value


def build_chain(lib, pdb, bonds, stacks, maxikey, double_face_res=(), recentre_bead=500, T=0.615, c_mg=0.005,
                c_k=0.050, seed=1, box=350.0, ss_D=0.850, hs_D=3.975, st_D=6.5):
    """Runs tis_build_chain on files in the reference's formats; returns (dict, rand draws used)."""
    faces = np.ascontiguousarray(double_face_res, dtype=np.int32)
    files = ChainFiles(pdb.encode(), bonds.encode(), stacks.encode(), maxikey.encode(),
                       faces.ctypes.data_as(_ip) if len(faces) else None, len(faces), recentre_bead)
    args = BuildArgs(ss_D, hs_D, st_D, T, c_mg, c_k, seed, box)
    out = C.POINTER(TisSystem)()
    draws = C.c_longlong(0)
    check(lib, lib.tis_build_chain(C.byref(files), C.byref(args), C.byref(out), C.byref(draws)), "tis_build_chain")
    try:
        d = system_to_dict(out)
    finally:
        lib.tis_system_free(out)
    return d, draws.value


def build_synthetic(lib, n_nt=1500, stack_frac=0.6, n_hb=400, radius=75.0, T=0.596, c_mg=0.010,
                    c_k=0.100, seed=1, box=600.0, ss_D=0.850, hs_D=3.975, st_D=6.5, data_dir=DATA_16SCD):
    """Runs tis_build_synthetic (a random-sequence self-avoiding chain) and returns the system dict."""
    args = SyntheticArgs(n_nt, stack_frac, n_hb, radius, BuildArgs(ss_D, hs_D, st_D, T, c_mg, c_k, seed, box))
    out = C.POINTER(TisSystem)()
    check(lib, lib.tis_build_synthetic(data_dir.encode(), C.byref(args), C.byref(out)), "tis_build_synthetic")
    try:
        return system_to_dict(out)
    finally:
        lib.tis_system_free(out)


def glibc_stream(lib, seed, n, skip=0):
    out = np.zeros(n, dtype=np.int32)
    check(lib, lib.tis_glibc_rand_fill(seed, skip, out.ctypes.data_as(C.POINTER(C.c_int32)), n),
          "tis_glibc_rand_fill")
    return out


def default_options(lib):
    o = Options()
    lib.tis_default_options(C.byref(o))
    return o


LAUNCH_FIELDS = ("graph_steps", "pair_lanes", "pair_block", "list_waves", "hb_groups")


def autotune(lib, handles, trial_steps=400, **options):
    """tis_autotune over the replicas of `handles` with the given options
    (precision, rng, seed, ...); returns ({launch field: value}, Tuning dict)."""
    arr = (TisSystem * len(handles))(*[h.struct for h in handles])
    o = default_options(lib)
    for k, v in options.items():
        setattr(o, k, v)
    rep = Tuning()
    check(lib, lib.tis_autotune(arr, len(handles), C.byref(o), trial_steps, C.byref(rep)), "tis_autotune")
    return ({k: getattr(o, k) for k in LAUNCH_FIELDS},
            {name: getattr(rep, name) for name, _ in Tuning._fields_})


class Engine:
    """A tis_ctx over one or more replicas (SystemHandle objects)."""

    def __init__(self, lib, handles, precision=PREC_FP64, rng=RNG_PHILOX, seed=1, noise_scale=1.0,
                 use_graph=1, graph_steps=50, **overrides):
        self.lib = lib
        self.handles = list(handles)
        arr = (TisSystem * len(self.handles))(*[h.struct for h in self.handles])
        self._arr = arr
        o = default_options(lib)
        o.precision, o.rng, o.seed, o.noise_scale = precision, rng, seed, noise_scale
        o.use_graph, o.graph_steps = use_graph, graph_steps
        for k, v in overrides.items():
            setattr(o, k, v)
        self.opt = o
        self.ctx = C.c_void_p()
        check(lib, lib.tis_create(arr, len(self.handles), C.byref(o), C.byref(self.ctx)), "tis_create")

    def close(self):
        if self.ctx:
            self.lib.tis_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self, rep=0):
        return self.handles[rep].d["n_particles"]

    def eval(self, mask, rep=0):
        f = np.zeros(3 * self.n(rep))
        e = Energies()
        check(self.lib, self.lib.tis_eval_forces(self.ctx, rep, mask, f.ctypes.data_as(_dp), C.byref(e)),
              "tis_eval_forces")
        return f.reshape(-1, 3), e.as_dict()

    def step(self, n):
        check(self.lib, self.lib.tis_step(self.ctx, n), "tis_step")

    def get_state(self, rep=0):
        out = [np.zeros(3 * self.n(rep)) for _ in range(3)]
        check(self.lib, self.lib.tis_get_state(self.ctx, rep, *[a.ctypes.data_as(_dp) for a in out]),
              "tis_get_state")
        return [a.reshape(-1, 3) for a in out]

    def set_state(self, pos, old, vel, klok=1, rep=0):
        a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1) for x in (pos, old, vel)]
        check(self.lib, self.lib.tis_set_state(self.ctx, rep, *[x.ctypes.data_as(_dp) for x in a], klok),
              "tis_set_state")

    def set_step_counter(self, klok, rep=0):
        check(self.lib, self.lib.tis_set_step_counter(self.ctx, rep, klok), "tis_set_step_counter")

    def prepare_graph(self, graph_steps):
        check(self.lib, self.lib.tis_prepare_graph(self.ctx, graph_steps), "tis_prepare_graph")

    def launch_counts(self):
        v = [C.c_longlong(0) for _ in range(3)]
        check(self.lib, self.lib.tis_get_launch_counts(self.ctx, *[C.byref(x) for x in v]), "tis_get_launch_counts")
        return {"graph_launches": v[0].value, "graph_steps": v[1].value, "eager_steps": v[2].value}

    # layout of tis_get_hb_trace: (step slot, replica, group, word)
    HB_TRACE_STEPS, HB_TRACE_GROUPS, HB_TRACE_WORDS = 64, 16, 10

    def trace_hb(self, on):
        check(self.lib, self.lib.tis_trace_hb(self.ctx, 1 if on else 0), "tis_trace_hb")

    def hb_trace(self):
        """The resolution's phase ring as a uint64 array (64, n_rep, 16, 10)."""
        shape = (self.HB_TRACE_STEPS, len(self.handles), self.HB_TRACE_GROUPS, self.HB_TRACE_WORDS)
        out = np.zeros(shape, dtype=np.uint64)
        check(self.lib, self.lib.tis_get_hb_trace(self.ctx, out.ctypes.data_as(C.POINTER(C.c_ulonglong)), out.size),
              "tis_get_hb_trace")
        return out

    def set_persistent(self, on):
        check(self.lib, self.lib.tis_set_persistent(self.ctx, 1 if on else 0), "tis_set_persistent")

    def persistent_counts(self):
        act, nl, ns = C.c_int(0), C.c_longlong(0), C.c_longlong(0)
        why = C.create_string_buffer(256)
        check(self.lib, self.lib.tis_get_persistent_counts(self.ctx, C.byref(act), C.byref(nl), C.byref(ns), why, 256),
              "tis_get_persistent_counts")
        return {"active": bool(act.value), "launches": nl.value, "steps": ns.value, "why": why.value.decode()}

    def set_rng_epoch(self, epoch, rep=-1):
        check(self.lib, self.lib.tis_set_rng_epoch(self.ctx, rep, epoch), "tis_set_rng_epoch")

    def energies(self, rep=0):
        e = Energies()
        check(self.lib, self.lib.tis_get_energies(self.ctx, rep, C.byref(e)), "tis_get_energies")
        return e.as_dict()

    def stack_energies(self, rep=0):
        n = self.handles[rep].d["n_st3"]
        out = np.zeros(max(n, 1))
        check(self.lib, self.lib.tis_get_stack_energies(self.ctx, rep, out.ctypes.data_as(_dp), n),
              "tis_get_stack_energies")
        return out[:n]

    def pairs(self, which, rep=0):
        n = C.c_longlong(0)
        check(self.lib, self.lib.tis_get_pairs(self.ctx, rep, which, None, 0, C.byref(n)), "tis_get_pairs")
        out = np.zeros(2 * n.value, dtype=np.int32)
        check(self.lib, self.lib.tis_get_pairs(self.ctx, rep, which, out.ctypes.data_as(_ip), n.value,
                                               C.byref(n)), "tis_get_pairs")
        return out.reshape(-1, 2)

    def hb_status(self, rep=0):
        n = self.handles[rep].d["n_hb"]
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(self.lib, self.lib.tis_get_hb_status(self.ctx, rep, out.ctypes.data_as(_ip), n), "tis_get_hb_status")
        return out[:n]

    def hb_outcome(self, rep=0):
        n = self.handles[rep].d["n_hb"]
        out = np.zeros(max(n, 1), dtype=np.int32)
        check(self.lib, self.lib.tis_get_hb_outcome(self.ctx, rep, out.ctypes.data_as(_ip), n), "tis_get_hb_outcome")
        return out[:n]

    def set_stream(self, raw, rep=0):
        self._stream = np.ascontiguousarray(raw, dtype=np.int32)
        check(self.lib, self.lib.tis_set_uniform_stream(
            self.ctx, rep, self._stream.ctypes.data_as(C.POINTER(C.c_int32)), len(self._stream)),
            "tis_set_uniform_stream")

    def exchange_now(self):
        check(self.lib, self.lib.tis_exchange_now(self.ctx), "tis_exchange_now")

    def exchange_stats(self):
        n = max(0, len(self.handles) - 1)
        att = np.zeros(max(n, 1), dtype=np.int64)
        acc = np.zeros(max(n, 1), dtype=np.int64)
        d = np.zeros(max(n, 1))
        check(self.lib, self.lib.tis_get_exchange_stats(
            self.ctx, att.ctypes.data_as(C.POINTER(C.c_longlong)), acc.ctypes.data_as(C.POINTER(C.c_longlong)),
            d.ctypes.data_as(_dp), n), "tis_get_exchange_stats")
        return att[:n], acc[:n], d[:n]

    def walkers(self):
        w = np.zeros(len(self.handles), dtype=np.int32)
        check(self.lib, self.lib.tis_get_walkers(self.ctx, w.ctypes.data_as(_ip), len(self.handles)),
              "tis_get_walkers")
        return w

    def n_total(self):
        return sum(h.d["n_particles"] for h in self.handles)

    def force_match(self, target, want_grads=True):
        """Loss, forces and gradients of tis_force_match (target: (N_total, 3))."""
        d0 = self.handles[0].d
        n = self.n_total()
        tgt = np.ascontiguousarray(target, dtype=np.float64).reshape(-1)
        loss = np.zeros(1)
        f = np.zeros(3 * n)
        out = {"forces": f}
        ptrs = [None] * 5
        if want_grads:
            out.update(pos=np.zeros(3 * n), hb_E=np.zeros(max(1, d0["n_hb"])), st3_E=np.zeros(max(1, d0["n_st3"])),
                       st2_E=np.zeros(max(1, d0["n_st2"])), ion_R=np.zeros(3))
            ptrs = [out[k].ctypes.data_as(_dp) for k in ("pos", "hb_E", "st3_E", "st2_E", "ion_R")]
        check(self.lib, self.lib.tis_force_match(self.ctx, tgt.ctypes.data_as(_dp), loss.ctypes.data_as(_dp),
                                                 f.ctypes.data_as(_dp), *ptrs), "tis_force_match")
        out["loss"] = float(loss[0])
        out["forces"] = f.reshape(-1, 3)
        if want_grads:
            out["pos"] = out["pos"].reshape(-1, 3)
            out["hb_E"] = out["hb_E"][:d0["n_hb"]]
            out["st3_E"] = out["st3_E"][:d0["n_st3"]]
            out["st2_E"] = out["st2_E"][:d0["n_st2"]]
        return out

    def set_force_field(self, hb_E=None, st3_E=None, st2_E=None, ion_R=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (hb_E, st3_E, st2_E, ion_R)]
        ptrs = [None if a is None else a.ctypes.data_as(_dp) for a in arrs]
        check(self.lib, self.lib.tis_set_force_field(self.ctx, *ptrs), "tis_set_force_field")

    def observe(self, every, cut=R_KT):
        """Clear the observable accumulators; sample every `every` steps (0: only observe_now)."""
        check(self.lib, self.lib.tis_observe(self.ctx, every, cut), "tis_observe")

    def observe_now(self):
        check(self.lib, self.lib.tis_observe_now(self.ctx), "tis_observe_now")

    def observables(self, rep=0):
        """Means over the samples: a dict of the scalars plus per_bead (n_beads, 4):
        Mg, Cl, K counts within the cut of each P bead and P(Mg count >= 1);
        hb_kept (n_hb); st3_mean (n_st3)."""
        d = self.handles[rep].d
        o = Observables()
        pb = np.zeros(4 * d["n_beads"])
        hb = np.zeros(max(1, d["n_hb"]))
        st = np.zeros(max(1, d["n_st3"]))
        check(self.lib, self.lib.tis_get_observables(
            self.ctx, rep, C.byref(o), pb.ctypes.data_as(_dp), d["n_beads"], hb.ctypes.data_as(_dp), d["n_hb"],
            st.ctypes.data_as(_dp), d["n_st3"]), "tis_get_observables")
        out = o.as_dict()
        out.update(per_bead=pb.reshape(-1, 4), hb_kept=hb[:d["n_hb"]], st3_mean=st[:d["n_st3"]])
        return out

    def set_profiling(self, on):
        check(self.lib, self.lib.tis_set_profiling(self.ctx, int(on)), "tis_set_profiling")

    def kernel_times(self):
        buf = C.create_string_buffer(8192)
        check(self.lib, self.lib.tis_kernel_times(self.ctx, buf, len(buf)), "tis_kernel_times")
        out = {}
        for line in buf.value.decode().splitlines():
            name, n, ms = line.split()
            out[name] = (int(n), float(ms))
        return out

    def cursor(self, rep=0):
        c = C.c_longlong(0)
        check(self.lib, self.lib.tis_get_stream_cursor(self.ctx, rep, C.byref(c)), "tis_get_stream_cursor")
        return c.value
