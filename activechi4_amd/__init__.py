"""activechi4_amd: MI355X-native active Kob-Andersen MD + dynamic-heterogeneity
analysis behind a C ABI (include/ac4.h, libac4.so).

This module is a thin ctypes binding used by the tests and bench.py; the
product path is libac4.so (hand-written HIP kernels for gfx950) and the C
drop-in executables built from activechi4_amd/host/.  There is no CPU
fallback: every compute call goes through libac4.so and fails loudly when the
library or a GPU is missing.
"""
import ctypes
import os

import numpy as np

__all__ = ["lib", "load", "MD", "Analysis", "ReplicaParams", "AC4Error", "FP64", "FP32",
           "OBS_MSD", "OBS_OVERLAP", "OBS_S4", "OBS_FS", "NQ", "LIB_PATH"]

HERE = os.path.dirname(os.path.abspath(__file__))
# AC4_LIB names another build of the library (the sanitizer build of the host C, say)
LIB_PATH = os.environ.get("AC4_LIB") or os.path.join(HERE, "libac4.so")
FP64, FP32 = 0, 1
OBS_MSD, OBS_OVERLAP, OBS_S4, OBS_FS = 1, 2, 4, 8
NQ = 162

_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)
_U64 = ctypes.POINTER(ctypes.c_uint64)
_U32 = ctypes.POINTER(ctypes.c_uint32)
_VP = ctypes.c_void_p


class ReplicaParams(ctypes.Structure):
    """ac4_replica_params of include/ac4.h."""
    _fields_ = [("n", ctypes.c_int), ("density", ctypes.c_double), ("T", ctypes.c_double),
                ("f0", ctypes.c_double), ("n_act", ctypes.c_int), ("tau_per", ctypes.c_int),
                ("seed", ctypes.c_uint64)]


class AC4Error(RuntimeError):
    pass


class Job(ctypes.Structure):
    """ac4_job of include/ac4.h."""
    _fields_ = [("n", ctypes.c_int), ("T", ctypes.c_double), ("f0", ctypes.c_double), ("tau_p", ctypes.c_double),
                ("steps", ctypes.c_long), ("seed", ctypes.c_uint64), ("t_arrive", ctypes.c_double)]


_SIGS = {
    "ac4_last_error": (ctypes.c_char_p, []),
    "ac4_device_count": (ctypes.c_int, []),
    "ac4_md_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ReplicaParams), ctypes.c_int,
                                     ctypes.POINTER(_VP)]),
    "ac4_md_destroy": (ctypes.c_int, [_VP]),
    "ac4_md_total_particles": (ctypes.c_int, [_VP]),
    "ac4_md_set_state": (ctypes.c_int, [_VP, ctypes.c_int] + [_D] * 9 + [_I] * 5),
    "ac4_md_get_state": (ctypes.c_int, [_VP, ctypes.c_int] + [_D] * 9 + [_I] * 3),
    "ac4_md_get_forces": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D]),
    "ac4_md_set_chain": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_md_get_chain": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_md_update_cell_list": (ctypes.c_int, [_VP, ctypes.c_int, _I, _I]),
    "ac4_md_calculate_forces": (ctypes.c_int, [_VP, _D]),
    "ac4_md_step": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_step_graph": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_set_rebuild_criterion": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_relax": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "ac4_md_shuffle": (ctypes.c_int, [_VP]),
    "ac4_md_shuffle_uniforms": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_md_stats": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_md_rebuilds": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_sync": (ctypes.c_int, [_VP]),
    "ac4_md_kernel_times": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_philox4x32": (ctypes.c_int, [_U32, _U32, _U32]),
    "ac4_ana_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, _I, ctypes.c_int,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                      ctypes.POINTER(_VP)]),
    "ac4_ana_destroy": (ctypes.c_int, [_VP]),
    "ac4_ana_gap": (ctypes.c_int, [_VP]),
    "ac4_ana_push_frame": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D, _I]),
    "ac4_ana_push_frame_device": (ctypes.c_int, [_VP, ctypes.c_int, _VP, _VP, _VP, _VP]),
    "ac4_ana_push_md": (ctypes.c_int, [_VP, ctypes.c_int, _VP, ctypes.c_int]),
    "ac4_ana_finalize": (ctypes.c_int, [_VP]),
    "ac4_ana_nslots": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_ana_get_msd": (ctypes.c_int, [_VP, _D, _D, _D, _I, _D, _I]),
    "ac4_ana_get_overlap": (ctypes.c_int, [_VP, _D, _D, _I, _D, _I]),
    "ac4_ana_get_phase": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D, _D, _I, _D, _D]),
    "ac4_ana_npairs": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_ana_get_pairs": (ctypes.c_int, [_VP, ctypes.c_int, _D, _I]),
    "ac4_ana_write": (ctypes.c_int, [_VP, ctypes.c_char_p]),
    "ac4_schedule": (ctypes.c_int, [ctypes.c_uint64, ctypes.c_int, ctypes.c_double, _U64, ctypes.c_int]),
    "ac4_origin_gap": (ctypes.c_int, [_I, ctypes.c_int]),
    "ac4_set_tuning": (ctypes.c_int, [ctypes.c_int, ctypes.c_int]),
    "ac4_frames_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.POINTER(_VP)]),
    "ac4_frames_destroy": (ctypes.c_int, [_VP]),
    "ac4_frames_push": (ctypes.c_int, [_VP, _D, _D, _D, ctypes.c_longlong]),
    "ac4_frames_push_md": (ctypes.c_int, [_VP, _VP, ctypes.c_int, ctypes.c_longlong]),
    "ac4_frames_count": (ctypes.c_int, [_VP]),
    "ac4_frames_find": (ctypes.c_int, [_VP, ctypes.c_longlong]),
    "ac4_frames_overlap": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I]),
    "ac4_md_fix_drift": (ctypes.c_int, [_VP, ctypes.c_int, _D]),
    "ac4_md_format_frame": (ctypes.c_longlong, [_VP, ctypes.c_int, ctypes.c_char_p, ctypes.c_longlong]),
    "ac4_frames_block": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, ctypes.c_char_p]),
    "ac4_frames_overlap_cage": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                               ctypes.c_char_p, _I, ctypes.c_int, _I]),
    "ac4_frames_guu": (ctypes.c_int, [_VP, _I, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_char_p,
                                      _D, _I, _I, _I]),
    "ac4_shell_directions": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, _D]),
    "ac4_frames_s4_shells": (ctypes.c_int, [_VP, ctypes.c_int, _I, _I, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _D, _D,
                                            _D]),
    "ac4_frames_s4_shells_noalloc": (ctypes.c_int, [_VP, ctypes.c_int, _I, _I, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_uint64, _D,
                                                    _D, _D]),
    "ac4_md_set_nve": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "ac4_chain_init": (None, [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _D]),
    "ac4_autotune": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ReplicaParams), ctypes.c_int, ctypes.c_int,
                                    ctypes.c_int, _D, _D, ctypes.c_int, _I]),
    "ac4_set_skin": (ctypes.c_int, [ctypes.c_double]),
    "ac4_group_plan": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ReplicaParams), ctypes.c_int, _I]),
    "ac4_md_step_graph_multi": (ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, ctypes.c_int]),
    "ac4_format_frame": (ctypes.c_longlong, [ctypes.c_int, _D, _D, _D, _I, ctypes.c_char_p, ctypes.c_longlong]),
    "ac4_parse_frame": (ctypes.c_longlong, [ctypes.c_char_p, ctypes.c_longlong, ctypes.c_int, _D, _D, _D, _I]),
    "ac4_textframes_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_longlong, ctypes.POINTER(_VP)]),
    "ac4_textframes_buffer": (ctypes.c_void_p, [_VP, ctypes.POINTER(ctypes.c_longlong)]),
    "ac4_textframes_parse": (ctypes.c_int, [_VP, ctypes.c_longlong, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                            ctypes.POINTER(ctypes.c_longlong)]),
    "ac4_textframes_copy": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D, _I]),
    "ac4_textframes_destroy": (ctypes.c_int, [_VP]),
    "ac4_md_checkpoint_size": (ctypes.c_longlong, [_VP, ctypes.c_int]),
    "ac4_md_checkpoint": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong]),
    "ac4_md_restore": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_void_p, ctypes.c_longlong]),
    "ac4_md_create_slots": (ctypes.c_int, [ctypes.c_int, _I, ctypes.c_int, ctypes.POINTER(_VP)]),
    "ac4_md_assign": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.POINTER(ReplicaParams)]),
    "ac4_md_release": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_slot_n": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_slot_offset": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_md_load_slot": (ctypes.c_int, [_VP, ctypes.c_int] + [_D] * 9 + [_I] * 5),
    "ac4_md_relax_slot": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double]),
    "ac4_md_export_all": (ctypes.c_int, [_VP, _D, _D, _D]),
    "ac4_md_overlap_enable": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                             ctypes.c_double]),
    "ac4_md_overlap_sample": (ctypes.c_int, [_VP]),
    "ac4_md_overlap_get": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _I]),
    "ac4_md_overlap_chi4": (ctypes.c_int, [_VP, ctypes.c_int, _I, _D]),
    "ac4_md_overlap_reset": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_server_create": (ctypes.c_int, [ctypes.c_int, _I, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(_VP)]),
    "ac4_server_destroy": (ctypes.c_int, [_VP]),
    "ac4_server_plan": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I,
                                       ctypes.c_int]),
    "ac4_server_create_budget": (ctypes.c_int, [ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_double, ctypes.POINTER(_VP)]),
    "ac4_server_enable_overlap": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "ac4_server_slots": (ctypes.c_int, [_VP, _I]),
    "ac4_server_job_error": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_server_job_overlap": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _I]),
    "ac4_md_slot_bytes": (ctypes.c_longlong, [ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "ac4_md_errors": (ctypes.c_int, [_VP, _I]),
    "ac4_md_raise_error": (ctypes.c_int, [_VP, ctypes.c_int, ctypes.c_int]),
    "ac4_server_handle": (_VP, [_VP]),
    "ac4_server_slot_handle": (_VP, [_VP, ctypes.c_int, _I]),
    "ac4_server_groups": (ctypes.c_int, [_VP]),
    "ac4_server_create_groups": (ctypes.c_int, [ctypes.c_int, _I, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                ctypes.c_int, ctypes.POINTER(_VP)]),
    "ac4_server_job_slot": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_server_submit": (ctypes.c_int, [_VP, ctypes.POINTER(Job)]),
    "ac4_server_tick": (ctypes.c_int, [_VP, ctypes.c_double]),
    "ac4_server_pending": (ctypes.c_int, [_VP]),
    "ac4_server_preempt": (ctypes.c_int, [_VP, ctypes.c_int]),
    "ac4_server_job": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, ctypes.POINTER(ctypes.c_long), _I, _I]),
    "ac4_server_result": (ctypes.c_int, [_VP, ctypes.c_int, _D, _D, _D]),
    "ac4_server_particle_steps": (ctypes.c_longlong, [_VP]),
    "ac4_init_lattice": (ctypes.c_int, [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_uint64]
                         + [_D] * 9 + [_I] * 5),
}

lib = None


def load(path=LIB_PATH):
    """Load libac4.so (raises when it is missing: there is no fallback)."""
    global lib
    if lib is not None:
        return lib
    if not os.path.exists(path):
        raise AC4Error("libac4.so not built at %s (run __graft_entry__.build())" % path)
    L = ctypes.CDLL(path)
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    lib = L
    return L


def _check(rc):
    if rc != 0:
        raise AC4Error("libac4 error %d: %s" % (rc, lib.ac4_last_error().decode(errors="replace")))
    return rc


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _ip(a):
    return None if a is None else a.ctypes.data_as(_I)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _len(n, **arrays):
    """Every array must hold n entries: the C side reads n of each."""
    for name, a in arrays.items():
        if a is not None and len(a) != n:
            raise ValueError("%s holds %d entries, %d expected" % (name, len(a), n))


def require_gpu():
    load()
    if lib.ac4_device_count() < 1:
        raise AC4Error("no HIP device visible: libac4 has no CPU path")


class MD:
    """R replicas of the active KA system on one GPU (ac4_md_* of include/ac4.h)."""

    def __init__(self, params, precision=FP64):
        load()
        params = list(params)
        arr = (ReplicaParams * len(params))(*params)
        h = _VP()
        _check(lib.ac4_md_create(len(params), arr, precision, ctypes.byref(h)))
        self.h = h
        self.params = params
        self.R = len(params)

    def close(self):
        if self.h:
            lib.ac4_md_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def n(self, r=0):
        return self.params[r].n

    def set_state(self, r, x, y, z, xu, yu, zu, px, py, pz, typ, act_tag=None, kx=None, ky=None, kz=None):
        a = [_f64(v) for v in (x, y, z, xu, yu, zu, px, py, pz)]
        t = _i32(typ)
        at = None if act_tag is None else _i32(act_tag)
        ks = [None if k is None else _i32(k) for k in (kx, ky, kz)]
        _len(self.n(r), x=a[0], y=a[1], z=a[2], xu=a[3], yu=a[4], zu=a[5], px=a[6], py=a[7], pz=a[8], type=t,
             act_tag=at)
        if all(k is not None for k in ks) and self.nact(r) > 0:
            _len(self.nact(r), kx=ks[0], ky=ks[1], kz=ks[2])
        _check(lib.ac4_md_set_state(self.h, r, *[_dp(v) for v in a], _ip(t), _ip(at), *[_ip(k) for k in ks]))

    def get_state(self, r=0):
        n = self.n(r)
        out = [np.empty(n) for _ in range(9)]
        nact = self.nact(r)
        ks = [np.empty(max(nact, 1), dtype=np.int32) for _ in range(3)]
        _check(lib.ac4_md_get_state(self.h, r, *[_dp(v) for v in out], *[_ip(k) for k in ks]))
        keys = ["x", "y", "z", "xu", "yu", "zu", "px", "py", "pz"]
        d = dict(zip(keys, out))
        d["kx"], d["ky"], d["kz"] = [k[:nact] for k in ks]
        return d

    def nact(self, r=0):
        p = self.params[r]
        return int(p.n * 0.1) if p.n_act < 0 else p.n_act

    def forces(self, r=0):
        n = self.n(r)
        f = [np.empty(n) for _ in range(3)]
        _check(lib.ac4_md_get_forces(self.h, r, *[_dp(v) for v in f]))
        return f

    def set_chain(self, r, chain9):
        c = _f64(chain9)
        _check(lib.ac4_md_set_chain(self.h, r, _dp(c)))

    def get_chain(self, r=0):
        c = np.empty(9)
        _check(lib.ac4_md_get_chain(self.h, r, _dp(c)))
        return c

    def update_cell_list(self, r=0):
        n = self.n(r)
        cell = np.empty(n, dtype=np.int32)
        cnt = np.empty(n, dtype=np.int32)
        _check(lib.ac4_md_update_cell_list(self.h, r, _ip(cell), _ip(cnt)))
        return cell, cnt

    def calculate_forces(self):
        out = np.empty(2 * self.R)
        _check(lib.ac4_md_calculate_forces(self.h, _dp(out)))
        return out.reshape(self.R, 2)

    def step(self, nsteps, graph=False):
        _check((lib.ac4_md_step_graph if graph else lib.ac4_md_step)(self.h, nsteps))

    def export_xu(self):
        """x_u of every replica in reference order (ac4_md_export_all): a list
        of (3, N_r) arrays."""
        nt = lib.ac4_md_total_particles(self.h)
        a = [np.empty(nt) for _ in range(3)]
        _check(lib.ac4_md_export_all(self.h, *[_dp(v) for v in a]))
        out = []
        for r in range(self.R):
            o, n = lib.ac4_md_slot_offset(self.h, r), self.n(r)
            out.append(np.stack([v[o:o + n] for v in a]))
        return out

    def relax(self, nsteps, f_start=1.5, f_end=1.0):
        """Velocity-rescale ramp from f_start*T to f_end*T (ac4_md_relax)."""
        _check(lib.ac4_md_relax(self.h, nsteps, f_start, f_end))

    def init_chain(self, r, seed=None):
        """InitialiseNoseHoover: eta = 0, p_eta ~ N(0, T*Q) (MDcode/time_evolve.c:476-493)
        through ac4_chain_init.  The draw depends on `seed` only (default: 7919*r)."""
        p = self.params[r]
        self.set_chain(r, chain_init(p.n, p.T, 7919 * r if seed is None else seed))

    def set_nve(self, r, on=True):
        """Plain velocity Verlet for replica r (ac4_md_set_nve); Heff = KE + PE."""
        _check(lib.ac4_md_set_nve(self.h, r, 1 if on else 0))

    def format_frame(self, r=0):
        """data_coor.out text of replica r's current frame, formatted on the
        device (ac4_md_format_frame), as bytes."""
        n = self.n(r)
        if getattr(self, "_fmt_buf", None) is None or len(self._fmt_buf) < 80 * n + 1:
            self._fmt_buf = ctypes.create_string_buffer(80 * n + 1)
        buf = self._fmt_buf
        k = lib.ac4_md_format_frame(self.h, r, buf, len(buf))
        if k < 0:
            _check(int(k))
        return ctypes.string_at(buf, k)

    def fix_drift(self, r=0):
        """Remove the mean momentum of replica r (ac4_md_fix_drift); returns it."""
        m = np.zeros(3)
        _check(lib.ac4_md_fix_drift(self.h, r, _dp(m)))
        return m

    def checkpoint(self, r=0):
        """Exact checkpoint of replica r as bytes (ac4_md_checkpoint)."""
        size = lib.ac4_md_checkpoint_size(self.h, r)
        if size < 0:
            _check(int(size))
        buf = ctypes.create_string_buffer(size)
        _check(lib.ac4_md_checkpoint(self.h, r, buf, size))
        return buf.raw

    def restore(self, r, blob):
        """Loads an exact checkpoint into replica r (ac4_md_restore)."""
        buf = ctypes.create_string_buffer(blob, len(blob))
        _check(lib.ac4_md_restore(self.h, r, buf, len(blob)))

    def set_rebuild_criterion(self, crit):
        _check(lib.ac4_md_set_rebuild_criterion(self.h, crit))

    def shuffle(self):
        _check(lib.ac4_md_shuffle(self.h))

    def shuffle_uniforms(self, r=0):
        u = np.empty(3 * max(self.nact(r), 1))
        _check(lib.ac4_md_shuffle_uniforms(self.h, r, _dp(u)))
        return u[: 3 * self.nact(r)].reshape(3, -1)

    def stats(self, r=0):
        o = np.empty(8)
        _check(lib.ac4_md_stats(self.h, r, _dp(o)))
        return dict(zip(["pe", "virial", "ke", "T", "P", "drift", "Heff", "steps"], o))

    def rebuilds(self, r=0):
        v = lib.ac4_md_rebuilds(self.h, r)
        if v < 0:
            _check(v)
        return v

    def sync(self):
        _check(lib.ac4_md_sync(self.h))

    def overlap_enable(self, sample_every, origin_every, n_origin, n_lags, a=0.3):
        self.ov = (sample_every, n_lags)
        _check(lib.ac4_md_overlap_enable(self.h, sample_every, origin_every, n_origin, n_lags, a))

    def overlap_sample(self):
        _check(lib.ac4_md_overlap_sample(self.h))

    def overlap(self, r):
        nl = self.ov[1] + 1
        q, q2 = np.empty(nl), np.empty(nl)
        c = np.empty(nl, dtype=np.int32)
        _check(lib.ac4_md_overlap_get(self.h, r, _dp(q), _dp(q2), _ip(c)))
        return q, q2, c

    def chi4(self, reps):
        """Pooled chi4(l), <Q>(l) and origin counts over the replicas `reps`
        (ac4_md_overlap_chi4, FP64 on the device), l = 0..n_lags."""
        nl = self.ov[1] + 1
        lst = np.ascontiguousarray(reps, dtype=np.int32)
        o = np.empty(3 * nl)
        _check(lib.ac4_md_overlap_chi4(self.h, len(lst), _ip(lst), _dp(o)))
        o = o.reshape(nl, 3)
        return o[:, 0].copy(), o[:, 1].copy(), o[:, 2].astype(np.int64)

    def errors(self):
        e = np.empty(self.R, dtype=np.int32)
        _check(lib.ac4_md_errors(self.h, _ip(e)))
        return e

    def overlap_reset(self, r):
        _check(lib.ac4_md_overlap_reset(self.h, r))

    def kernel_times(self, nsteps):
        """Average ms per step of {first_half, rebuild, force, shuffle,
        force_end} (HIP events; force is k_force alone)."""
        o = np.empty(5)
        _check(lib.ac4_md_kernel_times(self.h, nsteps, _dp(o)))
        return dict(zip(["first_half", "rebuild", "force", "shuffle", "force_end"], o))


class Analysis:
    """Streaming analysis handle (ac4_ana_* of include/ac4.h)."""

    def __init__(self, n, L, fidx, gap=0, factor=1.4, a=0.3, obs=15, type_flag=0):
        load()
        self.fidx = _i32(fidx)
        h = _VP()
        _check(lib.ac4_ana_create(n, L, len(self.fidx), _ip(self.fidx), gap, factor, a, obs, type_flag,
                                  ctypes.byref(h)))
        self.h = h
        self.n = n
        self.obs = obs

    def close(self):
        if self.h:
            lib.ac4_ana_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def gap(self):
        return lib.ac4_ana_gap(self.h)

    def push(self, pos, x, y, z, t):
        x, y, z, t = _f64(x), _f64(y), _f64(z), _i32(t)
        _len(self.n, x=x, y=y, z=z, type=t)
        _check(lib.ac4_ana_push_frame(self.h, pos, _dp(x), _dp(y), _dp(z), _ip(t)))

    def push_md(self, pos, md, r=0):
        _check(lib.ac4_ana_push_md(self.h, pos, md.h, r))

    def finalize(self):
        _check(lib.ac4_ana_finalize(self.h))

    def nslots(self, obs):
        return lib.ac4_ana_nslots(self.h, obs)

    def msd(self):
        ns = self.nslots(OBS_MSD)
        m0, m1, m2, dt = (np.empty(ns) for _ in range(4))
        cn = np.empty(ns, dtype=np.int32)
        noi = ctypes.c_int()
        _check(lib.ac4_ana_get_msd(self.h, _dp(m0), _dp(m1), _dp(m2), _ip(cn), _dp(dt), ctypes.byref(noi)))
        return dict(m0=m0, m1=m1, m2=m2, counters=cn, delta_t=dt, noi=noi.value)

    def overlap(self):
        ns = self.nslots(OBS_OVERLAP)
        c, x4, dt = (np.empty(ns) for _ in range(3))
        cn = np.empty(ns, dtype=np.int32)
        noi = ctypes.c_int()
        _check(lib.ac4_ana_get_overlap(self.h, _dp(c), _dp(x4), _ip(cn), _dp(dt), ctypes.byref(noi)))
        return dict(c=c, x4=x4, counters=cn, delta_t=dt, noi=noi.value)

    def phase(self, obs):
        ns = self.nslots(obs)
        cs, ss, cs2, ss2 = (np.empty((ns, NQ)) for _ in range(4))
        cn = np.empty(ns, dtype=np.int32)
        dt = np.empty(ns)
        qv = np.empty(NQ)
        _check(lib.ac4_ana_get_phase(self.h, obs, _dp(cs), _dp(ss), _dp(cs2), _dp(ss2), _ip(cn), _dp(dt), _dp(qv)))
        return dict(cs=cs, ss=ss, cs2=cs2, ss2=ss2, counters=cn, delta_t=dt, qval=qv)

    def pairs(self, obs):
        np_ = lib.ac4_ana_npairs(self.h, obs)
        if obs == OBS_MSD:
            v = np.empty((max(np_, 1), 3))
            _check(lib.ac4_ana_get_pairs(self.h, obs, _dp(v), None))
            return v[:np_]
        if obs == OBS_OVERLAP:
            iv = np.empty(max(np_, 1), dtype=np.int32)
            _check(lib.ac4_ana_get_pairs(self.h, obs, None, _ip(iv)))
            return iv[:np_]
        v = np.empty((max(np_, 1), 2 * NQ))
        iv = np.empty(max(np_, 1), dtype=np.int32)
        _check(lib.ac4_ana_get_pairs(self.h, obs, _dp(v), _ip(iv)))
        return (v[:np_], iv[:np_]) if obs == OBS_S4 else v[:np_]

    def write(self, sub_dir):
        _check(lib.ac4_ana_write(self.h, sub_dir.encode()))


class Server:
    """Request loop over device-resident slots (ac4_server_* of include/ac4.h)."""

    def __init__(self, slot_caps=None, precision=FP32, chunk=100, relax=300, budget=None, n_range=(500, 20000),
                 overlap=None, groups=0):
        """Slots of the capacities slot_caps, or planned from an HBM byte
        budget for requests with N in n_range (ac4_server_create_budget).
        overlap = (origin_every, n_origin, n_lags, a): every job's Q(t)
        accumulated on the device from its frames.  groups: launch groups of
        the slots (ac4_server_create_groups; 0 = the server's choice)."""
        load()
        h = _VP()
        if budget is not None:
            oe, no, nl, a = overlap if overlap else (0, 0, 0, 0.3)
            _check(lib.ac4_server_create_budget(float(budget), int(n_range[0]), int(n_range[1]), precision, chunk,
                                                relax, oe, no, nl, a, ctypes.byref(h)))
        else:
            caps = _i32(slot_caps)
            _check(lib.ac4_server_create_groups(len(caps), _ip(caps), precision, chunk, relax, int(groups),
                                                ctypes.byref(h)))
        self.h = h
        self.n_of = {}
        self.n_lags = None
        if overlap:
            self.n_lags = overlap[2]
            if budget is None:
                _check(lib.ac4_server_enable_overlap(self.h, *overlap))

    def slots(self):
        ns = lib.ac4_server_slots(self.h, None)
        caps = np.empty(ns, dtype=np.int32)
        lib.ac4_server_slots(self.h, _ip(caps))
        return caps

    def groups(self):
        return lib.ac4_server_groups(self.h)

    def raise_error(self, jid, bits=2):
        """Sets error bits on the replica of running job jid, as a kernel
        would on an overflow (ac4_md_raise_error); the next tick fails it."""
        slot = lib.ac4_server_job_slot(self.h, jid)
        if slot < 0:
            raise AC4Error("job %d is not running" % jid)
        local = np.zeros(1, dtype=np.int32)
        md = lib.ac4_server_slot_handle(self.h, slot, _ip(local))
        if not md:
            raise AC4Error(lib.ac4_last_error().decode(errors="replace"))
        _check(lib.ac4_md_raise_error(md, int(local[0]), bits))

    def job_error(self, jid):
        rc = lib.ac4_server_job_error(self.h, jid)
        if rc < 0:
            _check(rc)
        return rc

    def job_overlap(self, jid):
        nl = self.n_lags + 1
        q, q2 = np.empty(nl), np.empty(nl)
        c = np.empty(nl, dtype=np.int32)
        _check(lib.ac4_server_job_overlap(self.h, jid, _dp(q), _dp(q2), _ip(c)))
        return q, q2, c

    def close(self):
        if self.h:
            lib.ac4_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, n, T, f0, tau_p, steps, seed, t_arrive=0.0):
        j = Job(n, T, f0, tau_p, steps, seed, t_arrive)
        rc = lib.ac4_server_submit(self.h, ctypes.byref(j))
        if rc < 0:
            _check(rc)
        self.n_of[rc] = n
        return rc

    def tick(self, now):
        rc = lib.ac4_server_tick(self.h, now)
        if rc < 0:
            _check(rc)
        return rc

    def pending(self):
        return lib.ac4_server_pending(self.h)

    def preempt(self, jid):
        """Checkpoint a running job to host memory and free its slot (ac4_server_preempt)."""
        _check(lib.ac4_server_preempt(self.h, jid))

    def job(self, jid):
        ts, td = ctypes.c_double(), ctypes.c_double()
        sd = ctypes.c_long()
        fr, st = ctypes.c_int(), ctypes.c_int()
        _check(lib.ac4_server_job(self.h, jid, ctypes.byref(ts), ctypes.byref(td), ctypes.byref(sd),
                                  ctypes.byref(fr), ctypes.byref(st)))
        return dict(t_start=ts.value, t_done=td.value, steps_done=sd.value, frames=fr.value, state=st.value)

    def result(self, jid):
        n = self.n_of[jid]
        a = [np.empty(n) for _ in range(3)]
        _check(lib.ac4_server_result(self.h, jid, *[_dp(v) for v in a]))
        return a

    def particle_steps(self):
        return lib.ac4_server_particle_steps(self.h)


class Frames:
    """Device-resident trajectory of one replica (ac4_frames_* of include/ac4.h)."""

    def __init__(self, n, L, cap):
        load()
        h = _VP()
        _check(lib.ac4_frames_create(n, L, cap, ctypes.byref(h)))
        self.h = h
        self.n = n
        self.L = L

    def close(self):
        if self.h:
            lib.ac4_frames_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def push(self, xu, yu, zu, step):
        a = [_f64(v) for v in (xu, yu, zu)]
        _len(self.n, xu=a[0], yu=a[1], zu=a[2])
        _check(lib.ac4_frames_push(self.h, *[_dp(v) for v in a], int(step)))

    def push_md(self, md, r, step):
        _check(lib.ac4_frames_push_md(self.h, md.h, r, int(step)))

    def count(self):
        return lib.ac4_frames_count(self.h)

    def find(self, step):
        return lib.ac4_frames_find(self.h, int(step))

    def overlap(self, i0, i1, a=0.3):
        c = ctypes.c_int()
        _check(lib.ac4_frames_overlap(self.h, i0, i1, a, ctypes.byref(c)))
        return c.value

    def s4_shells(self, orig, lagf, q0, dq, nshell, ndir, seed, a=0.3, phases=False, noalloc=False):
        """Returns (s4_ref, s4_dir[, phases[npairs][ndir][nshell][2]]);
        noalloc=True calls ac4_frames_s4_shells_noalloc."""
        o, lf = _i32(orig), _i32(lagf)
        s4r, s4d = np.empty(nshell), np.empty(nshell)
        ph = np.empty((len(o), ndir, nshell, 2)) if (phases or noalloc) else None
        fn = lib.ac4_frames_s4_shells_noalloc if noalloc else lib.ac4_frames_s4_shells
        _check(fn(self.h, len(o), _ip(o), _ip(lf), a, q0, dq, nshell, ndir, int(seed) & (2**64 - 1), _dp(s4r),
                  _dp(s4d), _dp(ph) if ph is not None else None))
        return (s4r, s4d, ph) if phases else (s4r, s4d)

    def block(self, gap, cell_cut, sub_dir, factor=1.4, a=0.3, vh_time=1500.0):
        """Block MSD, overlap and van Hove files for one block edge under
        sub_dir/BlockAnalysis/ (ac4_frames_block)."""
        _check(lib.ac4_frames_block(self.h, int(gap), factor, a, cell_cut, vh_time, os.fsencode(sub_dir)))

    def overlap_cage(self, gap, path=None, factor=1.4, a=0.3, cell_cut=2.85, cap=1 << 20):
        """Cage-relative overlap (ac4_frames_overlap_cage); returns the count of
        every (origin, lag) pair in walk order."""
        pc = np.zeros(cap, dtype=np.int32)
        npairs = ctypes.c_int()
        _check(lib.ac4_frames_overlap_cage(self.h, int(gap), factor, a, cell_cut,
                                           os.fsencode(path) if path else None, _ip(pc), cap, ctypes.byref(npairs)))
        return pc[:min(npairs.value, cap)]

    def guu(self, act_tag, gap, tau, path=None, factor=1.4):
        """g_uu(r) (ac4_frames_guu); returns (hist[3][maxbin], frame_num, n_origins)."""
        at = _i32(act_tag)
        _len(self.n, act_tag=at)
        mb = int(0.5 * self.L / 0.005)
        h = np.zeros(3 * max(mb, 1))
        out = [ctypes.c_int() for _ in range(3)]
        _check(lib.ac4_frames_guu(self.h, _ip(at), int(gap), factor, tau, os.fsencode(path) if path else None,
                                  _dp(h), *[ctypes.byref(o) for o in out]))
        return h[:3 * out[0].value].reshape(3, out[0].value), out[1].value, out[2].value


def shell_directions(ndir, seed):
    """ndir unit vectors from seed (ac4_shell_directions)."""
    load()
    n3 = np.empty((ndir, 3))
    _check(lib.ac4_shell_directions(ndir, int(seed) & (2**64 - 1), _dp(n3)))
    return n3


def chi4_peak(md, reps, dt_lag, first=1):
    """The chi4(t) peak of the replicas `reps` of an MD handle with overlap
    accumulators: chi4(l) pooled over the replicas and their origins
    (MD.chi4, FP64 on the device), its maximum over lags >= first, and
    jackknife standard errors over the replicas (each replica left out in
    turn; replicas are independent runs).  Returns a dict with the curve, the
    peak height, time (lag * dt_lag) and lag, and their errors."""
    chi, q, c = md.chi4(reps)
    lags = np.arange(len(chi))
    ok = (c > 0) & (lags >= first)
    if not ok.any():
        return None
    k = int(np.flatnonzero(ok)[np.argmax(chi[ok])])
    out = {"chi4": chi, "Q": q, "origins": c, "peak": float(chi[k]), "peak_lag": k, "peak_time": k * dt_lag,
           "Q_at_peak": float(q[k])}
    m = len(reps)
    if m >= 3:
        hs, ts = [], []
        for r in range(m):
            sub = [x for i, x in enumerate(reps) if i != r]
            cj, _, cc = md.chi4(sub)
            okj = (cc > 0) & (lags >= first)
            kj = int(np.flatnonzero(okj)[np.argmax(cj[okj])])
            hs.append(cj[kj])
            ts.append(kj * dt_lag)
        hs, ts = np.array(hs), np.array(ts)
        f = (m - 1) / m
        out["peak_err"] = float(np.sqrt(f * ((hs - hs.mean()) ** 2).sum()))
        out["peak_time_err"] = float(np.sqrt(f * ((ts - ts.mean()) ** 2).sum()))
    return out


def server_plan(budget, n_min, n_max, n_origin=0, n_lags=0):
    """Slot capacities ac4_server_plan derives from an HBM byte budget."""
    load()
    caps = np.empty(65536, dtype=np.int32)
    ns = lib.ac4_server_plan(float(budget), n_min, n_max, n_origin, n_lags, _ip(caps), len(caps))
    if ns < 0:
        _check(ns)
    return caps[:ns].copy()


def slot_bytes(cap_n, n_origin=0, n_lags=0):
    load()
    b = lib.ac4_md_slot_bytes(cap_n, n_origin, n_lags)
    if b < 0:
        _check(int(b))
    return int(b)


def chain_init(n, T, seed):
    """Nose-Hoover chain {eta[3], p_eta[3], Q[3]} drawn from seed (ac4_chain_init)."""
    load()
    c = np.empty(9)
    lib.ac4_chain_init(int(n), float(T), int(seed) & (2**64 - 1), _dp(c))
    return c


def autotune(params, precision=FP32, relax=300, steps=200):
    """Times every kernel shape on this replica set and keeps the fastest for
    handles created afterwards (ac4_autotune).  Returns (tpp, nlist_mode,
    skin, groups) and the table of rows (tpp, nlist_mode, skin, groups,
    particle-steps/s); groups = launch groups the set is best split into."""
    load()
    arr = (ReplicaParams * len(params))(*params)
    best = np.zeros(4)
    table = np.zeros((40, 5))
    nc = ctypes.c_int(0)
    _check(lib.ac4_autotune(len(params), arr, precision, relax, steps, _dp(best), _dp(table), 40,
                            ctypes.byref(nc)))
    return (int(best[0]), int(best[1]), float(best[2]), int(best[3])), table[: nc.value]


def group_plan(params, groups):
    """Launch group of every replica (ac4_group_plan: balanced particle counts)."""
    load()
    arr = (ReplicaParams * len(params))(*params)
    g = np.empty(len(params), dtype=np.int32)
    _check(lib.ac4_group_plan(len(params), arr, int(groups), _ip(g)))
    return g


def step_multi(mds, nsteps):
    """nsteps graph steps of several handles together (ac4_md_step_graph_multi)."""
    load()
    arr = (_VP * len(mds))(*[m.h for m in mds])
    _check(lib.ac4_md_step_graph_multi(arr, len(mds), nsteps))


def format_frame(x, y, z, typ):
    """data_coor.out text of one frame (ac4_format_frame), as bytes."""
    load()
    x, y, z, t = _f64(x), _f64(y), _f64(z), _i32(typ)
    _len(len(x), y=y, z=z, type=t)
    buf = ctypes.create_string_buffer(80 * len(x) + 1)
    k = lib.ac4_format_frame(len(x), _dp(x), _dp(y), _dp(z), _ip(t), buf, 80 * len(x) + 1)
    if k < 0:
        _check(int(k))
    return buf.raw[:k]


def parse_frame(text, n):
    """Inverse of format_frame (ac4_parse_frame): (x, y, z, type, bytes consumed)."""
    load()
    x, y, z = np.empty(n), np.empty(n), np.empty(n)
    t = np.empty(n, dtype=np.int32)
    k = lib.ac4_parse_frame(text, len(text), n, _dp(x), _dp(y), _dp(z), _ip(t))
    if k < 0:
        _check(int(k))
    return x, y, z, t, k


class TextFrames:
    """data_coor.out text parsed on the GPU (ac4_textframes_* of include/ac4.h)."""

    def __init__(self, n, chunk_bytes=0):
        load()
        h = _VP()
        _check(lib.ac4_textframes_create(n, chunk_bytes, ctypes.byref(h)))
        self.h, self.n = h, n
        cap = ctypes.c_longlong()
        self.buf = lib.ac4_textframes_buffer(h, ctypes.byref(cap))
        self.cap = cap.value

    def close(self):
        if self.h:
            lib.ac4_textframes_destroy(self.h)
            self.h = None

    def parse(self, text, max_frames=1 << 30):
        """Parses the whole frames at the start of `text` (bytes, at most the
        chunk size): returns (frames, bytes used)."""
        if len(text) > self.cap:
            raise ValueError("text longer than the chunk (%d bytes)" % self.cap)
        ctypes.memmove(self.buf, text, len(text))
        nf, used = ctypes.c_int(), ctypes.c_longlong()
        _check(lib.ac4_textframes_parse(self.h, len(text), max_frames, ctypes.byref(nf), ctypes.byref(used)))
        return nf.value, used.value

    def frame(self, k):
        """Frame k of the last parse as host arrays (x, y, z, type)."""
        x, y, z = np.empty(self.n), np.empty(self.n), np.empty(self.n)
        t = np.empty(self.n, dtype=np.int32)
        _check(lib.ac4_textframes_copy(self.h, k, _dp(x), _dp(y), _dp(z), _ip(t)))
        return x, y, z, t


def set_skin(skin=0.35):
    """Verlet skin for handles created afterwards (ac4_set_skin)."""
    load()
    _check(lib.ac4_set_skin(skin))


def set_tuning(tpp=0, nlist_mode=3):
    """Kernel shape for handles created afterwards (ac4_set_tuning): force
    team width tpp (0 = automatic) and dynamics list builder nlist_mode."""
    load()
    _check(lib.ac4_set_tuning(tpp, nlist_mode))


def schedule(run_length, n_origins=200, factor=1.4):
    """prepareSavingArray (MDcode/main.c:62-119) through libac4 (host logic)."""
    load()
    cap = n_origins * 64 + 64
    out = np.zeros(cap, dtype=np.uint64)
    n = lib.ac4_schedule(run_length, n_origins, factor, out.ctypes.data_as(_U64), cap)
    if n < 0:
        _check(n)
    return out[:n].copy(), out


def init_lattice(n, T, seed, n_act=-1):
    """Seeded lattice state (ac4_init_lattice): dict of SPHERE fields + k's."""
    load()
    nact = int(n * 0.1) if n_act < 0 else n_act
    d = {k: np.empty(n) for k in ("x", "y", "z", "xu", "yu", "zu", "px", "py", "pz")}
    d["type"] = np.empty(n, dtype=np.int32)
    d["act_tag"] = np.empty(n, dtype=np.int32)
    for k in ("kx", "ky", "kz"):
        d[k] = np.empty(max(nact, 1), dtype=np.int32)
    _check(lib.ac4_init_lattice(n, T, n_act, seed, *[_dp(d[k]) for k in ("x", "y", "z", "xu", "yu", "zu",
                                                                         "px", "py", "pz")],
                                _ip(d["type"]), _ip(d["act_tag"]), _ip(d["kx"]), _ip(d["ky"]), _ip(d["kz"])))
    for k in ("kx", "ky", "kz"):
        d[k] = d[k][:nact]
    return d


def origin_gap(fidx):
    load()
    f = _i32(fidx)
    g = lib.ac4_origin_gap(_ip(f), len(f))
    if g < 0:
        _check(g)
    return g


def philox(ctr4, key2):
    load()
    c = np.ascontiguousarray(ctr4, dtype=np.uint32)
    k = np.ascontiguousarray(key2, dtype=np.uint32)
    o = np.empty(4, dtype=np.uint32)
    _check(lib.ac4_philox4x32(c.ctypes.data_as(_U32), k.ctypes.data_as(_U32), o.ctypes.data_as(_U32)))
    return o
