"""ctypes binding of ``librepeat_hip.so`` (declarations in ``include/repeat_hip.h``).

The library is the only compute path: if it is missing or cannot be loaded,
every entry point raises ``RepeatHipError``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librepeat_hip.so")

RP_OK = 0
RP_EINVAL = -1
RP_ENOMEM = -2
RP_EHIP = -3
RP_ESINGULAR = -4
RP_EBUDGET = -5
RP_ESTATE = -6
RP_EEMPTY = -7
RP_EAGAIN = -8

RP_F64, RP_F32 = 0, 1
RP_ORDER_CUBE, RP_ORDER_FLAT = 0, 1
RP_PREC_FP64, RP_PREC_OZAKI, RP_PREC_OZAKI_REFINED, RP_PREC_FP64_FACT = 0, 1, 2, 3

EXPORTED = (
    "rp_ctx_create", "rp_ctx_destroy", "rp_last_error", "rp_ctx_hbm_in_use",
    "rp_fit_create", "rp_fit_destroy", "rp_fit_info_get", "rp_fit_grid", "rp_fit_ewald",
    "rp_fit_accumulate", "rp_fit_normal_equations", "rp_fit_solve", "rp_fit_phi_matvec",
    "rp_fit_phi", "rp_fit_timings", "rp_ctx_profile", "rp_ctx_profile_kinds", "rp_ctx_kernel_stats", "rp_fit_refit",
    "rp_fit_reset", "rp_server_create", "rp_server_destroy", "rp_server_error", "rp_server_now",
    "rp_request_bytes", "rp_submit", "rp_poll", "rp_server_stats_get",
    "rp_cube_read_header", "rp_cube_read", "rp_parse_values", "rp_fit_qeq", "rp_fit_refine_stats",
    "rp_chunk_plan", "rp_fit_phi_noalloc", "rp_fit_solve_report",
)
RP_SOLVER_CHOLESKY, RP_SOLVER_LU, RP_SOLVER_LSTSQ = 0, 1, 2
RP_REFINE_TARGET_E = 1e-6
RP_SENSITIVITY_LIMIT_E = 1e-5
KERNEL_KINDS = {"kspace": 0, "gram": 1, "filter": 2, "solve": 3, "phase": 4, "slice": 5, "reduce": 6, "refine": 7}


class RepeatHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"librepeat_hip error {code}: {msg}")
        self.code = code


class SingularKKTError(RepeatHipError):
    pass


class RpSystem(ctypes.Structure):
    _fields_ = [("n_grid", ctypes.c_int32 * 3), ("axis", ctypes.c_double * 9), ("n_atoms", ctypes.c_int32),
                ("atomic_number", ctypes.POINTER(ctypes.c_int32)), ("atom_xyz", ctypes.POINTER(ctypes.c_double)),
                ("esp", ctypes.c_void_p), ("esp_dtype", ctypes.c_int32), ("esp_order", ctypes.c_int32)]


class RpOptions(ctypes.Structure):
    _fields_ = [("fit_type", ctypes.c_int32), ("precision", ctypes.c_int32), ("vdw_factor", ctypes.c_double),
                ("vdw_max", ctypes.c_double), ("r_cutoff", ctypes.c_double), ("refine_steps", ctypes.c_int32),
                ("flags", ctypes.c_int32)]


class RpFitInfo(ctypes.Structure):
    _fields_ = [("n_total", ctypes.c_int64), ("n_keep", ctypes.c_int64), ("n_eff", ctypes.c_int64),
                ("n_atoms", ctypes.c_int32), ("n_k", ctypes.c_int32), ("n_shift", ctypes.c_int32),
                ("n_shift_pruned", ctypes.c_int32), ("volume", ctypes.c_double), ("alpha", ctypes.c_double),
                ("sqrt_alpha", ctypes.c_double), ("v_mean", ctypes.c_double)]


class RpConstraints(ctypes.Structure):
    _fields_ = [("q_tot", ctypes.c_double), ("n_reduced", ctypes.c_int32),
                ("red_ptr", ctypes.POINTER(ctypes.c_int32)), ("red_src", ctypes.POINTER(ctypes.c_int32)),
                ("red_mult", ctypes.POINTER(ctypes.c_double)), ("weights", ctypes.POINTER(ctypes.c_double)),
                ("diag_add", ctypes.POINTER(ctypes.c_double)), ("rhs_add", ctypes.POINTER(ctypes.c_double)),
                ("expand_src", ctypes.POINTER(ctypes.c_int32))]


class RpRequest(ctypes.Structure):
    _fields_ = [("system", RpSystem), ("options", RpOptions), ("constraints", RpConstraints)]


class RpResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("slot", ctypes.c_int32), ("n_atoms", ctypes.c_int32),
                ("reserved_", ctypes.c_int32), ("n_keep", ctypes.c_int64), ("t_submit", ctypes.c_double),
                ("t_admit", ctypes.c_double), ("t_done", ctypes.c_double), ("device_ms", ctypes.c_double * 3),
                ("bytes_bound", ctypes.c_uint64), ("error", ctypes.c_char * 256)]


class RpServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_int64), ("completed", ctypes.c_int64), ("failed", ctypes.c_int64),
                ("queued", ctypes.c_int32), ("running", ctypes.c_int32), ("max_running", ctypes.c_int32),
                ("n_slots", ctypes.c_int32), ("budget", ctypes.c_uint64), ("reserved", ctypes.c_uint64),
                ("peak_reserved", ctypes.c_uint64)]


class RpCubeHeader(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int32), ("n_grid", ctypes.c_int32 * 3), ("origin", ctypes.c_double * 3),
                ("axis", ctypes.c_double * 9), ("values_offset", ctypes.c_int64)]


class RpSolveReport(ctypes.Structure):
    _fields_ = [("solver", ctypes.c_int32), ("refine_steps", ctypes.c_int32), ("last_correction", ctypes.c_double),
                ("error_bound", ctypes.c_double), ("within_target", ctypes.c_int32), ("well_conditioned", ctypes.c_int32),
                ("sensitivity", ctypes.c_double)]


class RpChunkPlanInfo(ctypes.Structure):
    _fields_ = [("rows", ctypes.c_int64), ("n_chunks", ctypes.c_int64), ("np", ctypes.c_int32), ("kp", ctypes.c_int32),
                ("splits", ctypes.c_int32), ("n_tiles", ctypes.c_int32), ("slices", ctypes.c_int32),
                ("stage_rows", ctypes.c_int32)]


_lib = None


def load(path: str | None = None):
    """Load the HIP library (raises if it is absent: the product has no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise RepeatHipError(RP_EINVAL, f"{path} not built; run __graft_entry__.build() or pyrepeat_amd.build")
    L = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    i64p = ctypes.POINTER(ctypes.c_int64)
    L.rp_ctx_create.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.rp_ctx_destroy.argtypes = [vp]
    L.rp_last_error.argtypes = [vp]
    L.rp_last_error.restype = ctypes.c_char_p
    L.rp_ctx_hbm_in_use.argtypes = [vp, ctypes.POINTER(ctypes.c_uint64)]
    L.rp_fit_create.argtypes = [vp, ctypes.POINTER(RpSystem), ctypes.POINTER(RpOptions), ctypes.POINTER(vp)]
    L.rp_fit_destroy.argtypes = [vp]
    L.rp_fit_info_get.argtypes = [vp, ctypes.POINTER(RpFitInfo)]
    L.rp_fit_grid.argtypes = [vp, ctypes.POINTER(ctypes.c_uint8), i64p, dp, dp]
    L.rp_fit_ewald.argtypes = [vp, dp, dp, dp, dp]
    L.rp_fit_accumulate.argtypes = [vp]
    L.rp_fit_normal_equations.argtypes = [vp, dp, dp, dp, dp]
    L.rp_fit_solve.argtypes = [vp, ctypes.POINTER(RpConstraints), dp, dp, ctypes.POINTER(ctypes.c_int32)]
    L.rp_fit_phi_matvec.argtypes = [vp, dp, dp]
    L.rp_fit_phi.argtypes = [vp, ctypes.c_int64, dp, dp]
    L.rp_fit_phi_noalloc.argtypes = [vp, ctypes.c_int64, dp, dp]
    L.rp_fit_solve_report.argtypes = [vp, ctypes.POINTER(RpSolveReport)]
    L.rp_fit_timings.argtypes = [vp, dp]
    L.rp_fit_refine_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int32), dp]
    L.rp_ctx_profile.argtypes = [vp, ctypes.c_int]
    L.rp_ctx_profile_kinds.argtypes = [vp, ctypes.c_uint32]
    L.rp_ctx_kernel_stats.argtypes = [vp, ctypes.c_int, dp, i64p]
    L.rp_fit_refit.argtypes = [vp, ctypes.POINTER(RpConstraints), dp]
    L.rp_fit_reset.argtypes = [vp, ctypes.POINTER(RpSystem), ctypes.POINTER(RpOptions)]
    L.rp_server_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(vp)]
    L.rp_server_destroy.argtypes = [vp]
    L.rp_server_error.argtypes = [vp]
    L.rp_server_error.restype = ctypes.c_char_p
    L.rp_server_now.argtypes = [vp]
    L.rp_server_now.restype = ctypes.c_double
    L.rp_request_bytes.argtypes = [ctypes.POINTER(RpRequest), ctypes.POINTER(ctypes.c_uint64)]
    L.rp_submit.argtypes = [vp, ctypes.POINTER(RpRequest), ctypes.POINTER(ctypes.c_uint64)]
    L.rp_poll.argtypes = [vp, ctypes.c_uint64, ctypes.c_int32, ctypes.POINTER(RpResult), dp]
    L.rp_server_stats_get.argtypes = [vp, ctypes.POINTER(RpServerStats)]
    L.rp_cube_read_header.argtypes = [ctypes.c_char_p, ctypes.POINTER(RpCubeHeader)]
    L.rp_cube_read.argtypes = [ctypes.c_char_p, ctypes.POINTER(RpCubeHeader), ctypes.POINTER(ctypes.c_int32), dp, dp,
                               ctypes.c_int]
    L.rp_fit_qeq.argtypes = [vp, dp, dp, ctypes.c_double, dp, dp]
    L.rp_parse_values.argtypes = [ctypes.c_char_p, ctypes.c_int64, dp, ctypes.c_int64, ctypes.c_int, i64p]
    L.rp_chunk_plan.argtypes = [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                ctypes.POINTER(RpChunkPlanInfo)]
    for name in EXPORTED:
        if getattr(L, name).restype is None and name not in ("rp_last_error", "rp_server_error", "rp_server_now"):
            getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def ptr(a, ctype=ctypes.c_double):
    return None if a is None else a.ctypes.data_as(ctypes.POINTER(ctype))


def chunk_plan(n_keep: int, n_atoms: int, n_k: int, precision: int) -> dict:
    """The streaming geometry the accumulate of this size allocates and launches from (host only)."""
    info = RpChunkPlanInfo()
    check(load().rp_chunk_plan(int(n_keep), int(n_atoms), int(n_k), int(precision), ctypes.byref(info)))
    return {name: int(getattr(info, name)) for name, _ in RpChunkPlanInfo._fields_}


def check(rc, ctx=None):
    if rc != RP_OK:
        msg = load().rp_last_error(ctx).decode(errors="replace")
        if rc == RP_ESINGULAR:
            raise SingularKKTError(rc, msg)
        raise RepeatHipError(rc, msg)


class Context:
    """One device context (stream + HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        L = load()
        h = ctypes.c_void_p()
        check(L.rp_ctx_create(int(device), int(hbm_budget), ctypes.byref(h)))
        self.handle = h
        self.device = device
        self._fits = weakref.WeakSet()

    def hbm_in_use(self):
        v = ctypes.c_uint64()
        check(load().rp_ctx_hbm_in_use(self.handle, ctypes.byref(v)), self.handle)
        return int(v.value)

    def profile(self, enable=True, kinds=None):
        """Kernel timing with events on the stream; ``kinds`` (names of KERNEL_KINDS) limits it to
        those kinds, which keeps the event records off the other launches."""
        if kinds is None:
            check(load().rp_ctx_profile(self.handle, int(enable)), self.handle)
        else:
            mask = 0
            for k in kinds:
                mask |= 1 << KERNEL_KINDS[k]
            check(load().rp_ctx_profile_kinds(self.handle, mask if enable else 0), self.handle)

    def kernel_stats(self):
        out = {}
        for name, kind in KERNEL_KINDS.items():
            ms = ctypes.c_double()
            n = ctypes.c_int64()
            check(load().rp_ctx_kernel_stats(self.handle, kind, ctypes.byref(ms), ctypes.byref(n)), self.handle)
            out[name] = (ms.value, n.value)
        return out

    def close(self):
        """Destroys the context after its fits (the library defers it while any fit is alive)."""
        if getattr(self, "handle", None):
            for f in list(getattr(self, "_fits", ())):
                f.close()
            load().rp_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0, 0)
    return _default_ctx


def read_cube(path, n_threads=0):
    """``read_gauss_cube`` through the library's parser: header, atoms and the ESP in file order."""
    L = load()
    bpath = os.fsencode(path)
    h = RpCubeHeader()
    check(L.rp_cube_read_header(bpath, ctypes.byref(h)))
    n_grid = np.array(list(h.n_grid), dtype=np.int64)
    Z = np.zeros(h.n_atoms, dtype=np.int32)
    xyz = np.zeros((h.n_atoms, 3))
    esp = np.empty(int(np.prod(n_grid)))
    check(L.rp_cube_read(bpath, ctypes.byref(h), ptr(Z, ctypes.c_int32), ptr(xyz), ptr(esp), int(n_threads)))
    return {"n_atoms": int(h.n_atoms), "n_grid": n_grid, "atom_index": Z.astype(np.int64), "atom_pos": xyz,
            "axis_vector": np.array(list(h.axis)).reshape(3, 3), "origin": np.array(list(h.origin)), "esp_cube": esp}


def parse_values(text: bytes, cap=None, n_threads=0):
    """Every whitespace-separated value of ``text`` as float64 (bitwise ``float()``); the first ``cap``."""
    L = load()
    n = ctypes.c_int64()
    if cap is None:
        check(L.rp_parse_values(text, len(text), None, 0, int(n_threads), ctypes.byref(n)))
        cap = n.value
    out = np.empty(int(cap))
    check(L.rp_parse_values(text, len(text), ptr(out), int(cap), int(n_threads), ctypes.byref(n)))
    return out, int(n.value)


def make_system(n_grid, axis, atomic_numbers, atom_xyz, esp, esp_order=RP_ORDER_CUBE):
    """An ``rp_system`` over arrays; returns ``(struct, arrays)`` (keep ``arrays`` alive while it is used)."""
    Z = np.ascontiguousarray(atomic_numbers, dtype=np.int32)
    xyz = np.ascontiguousarray(atom_xyz, dtype=np.float64).reshape(-1, 3)
    esp = np.ascontiguousarray(esp)
    if esp.dtype not in (np.float32, np.float64):
        esp = esp.astype(np.float64)
    sysd = RpSystem()
    for i in range(3):
        sysd.n_grid[i] = int(n_grid[i])
    ax = np.asarray(axis, dtype=np.float64).ravel()
    for i in range(9):
        sysd.axis[i] = float(ax[i])
    sysd.n_atoms = len(Z)
    sysd.atomic_number = ptr(Z, ctypes.c_int32)
    sysd.atom_xyz = ptr(xyz)
    sysd.esp = esp.ctypes.data_as(ctypes.c_void_p)
    sysd.esp_dtype = RP_F32 if esp.dtype == np.float32 else RP_F64
    sysd.esp_order = int(esp_order)
    return sysd, (Z, xyz, esp)


RP_OPT_STREAM_REFINE = 1


def make_options(fit_type=1, precision=RP_PREC_FP64, vdw_factor=1.0, vdw_max=1000.0, r_cutoff=20.0, refine_steps=0,
                 flags=0):
    return RpOptions(int(fit_type), int(precision), float(vdw_factor), float(vdw_max), float(r_cutoff),
                     int(refine_steps), int(flags))


class Fit:
    """Device-resident fit (filter done at construction)."""

    def __init__(self, ctx: Context, n_grid, axis, atomic_numbers, atom_xyz, esp, esp_order=RP_ORDER_CUBE,
                 fit_type=1, precision=RP_PREC_FP64, vdw_factor=1.0, vdw_max=1000.0, r_cutoff=20.0, refine_steps=0,
                 flags=0):
        L = load()
        self.ctx = ctx
        sysd, self._arrays = make_system(n_grid, axis, atomic_numbers, atom_xyz, esp, esp_order)
        opt = make_options(fit_type, precision, vdw_factor, vdw_max, r_cutoff, refine_steps, flags)
        h = ctypes.c_void_p()
        check(L.rp_fit_create(ctx.handle, ctypes.byref(sysd), ctypes.byref(opt), ctypes.byref(h)), ctx.handle)
        self.handle = h
        ctx._fits.add(self)
        self.n_atoms = int(sysd.n_atoms)
        self.info = self._info()

    def reset(self, n_grid, axis, atomic_numbers, atom_xyz, esp, esp_order=RP_ORDER_CUBE, fit_type=1,
              precision=RP_PREC_FP64, vdw_factor=1.0, vdw_max=1000.0, r_cutoff=20.0, refine_steps=0):
        """Reload this fit with another system, reusing its device buffers."""
        sysd, arrays = make_system(n_grid, axis, atomic_numbers, atom_xyz, esp, esp_order)
        opt = make_options(fit_type, precision, vdw_factor, vdw_max, r_cutoff, refine_steps)
        check(load().rp_fit_reset(self.handle, ctypes.byref(sysd), ctypes.byref(opt)), self.ctx.handle)
        self._arrays = arrays
        self.n_atoms = int(sysd.n_atoms)
        self.info = self._info()

    def _info(self):
        inf = RpFitInfo()
        check(load().rp_fit_info_get(self.handle, ctypes.byref(inf)), self.ctx.handle)
        return inf

    def grid(self, mask=False):
        L = load()
        inf = self.info
        idx = np.empty(inf.n_keep, dtype=np.int64)
        v = np.empty(inf.n_keep, dtype=np.float64)
        xyz = np.empty((inf.n_keep, 3), dtype=np.float64)
        m = np.empty(inf.n_total, dtype=np.uint8) if mask else None
        check(L.rp_fit_grid(self.handle, ptr(m, ctypes.c_uint8), ptr(idx, ctypes.c_int64), ptr(v), ptr(xyz)),
              self.ctx.handle)
        return (xyz, v, idx, m) if mask else (xyz, v, idx)

    def ewald(self):
        inf = self.info
        kv = np.empty((inf.n_k, 3))
        kc = np.empty(inf.n_k)
        sh = np.empty((inf.n_shift, 3))
        sp = np.empty((inf.n_shift_pruned, 3))
        check(load().rp_fit_ewald(self.handle, ptr(kv), ptr(kc), ptr(sh), ptr(sp)), self.ctx.handle)
        return kv, kc, sh, sp

    def accumulate(self):
        check(load().rp_fit_accumulate(self.handle), self.ctx.handle)

    def normal_equations(self):
        N = self.n_atoms
        ata = np.empty((N, N))
        atb = np.empty(N)
        sp = np.empty(N)
        sv = np.empty(1)
        check(load().rp_fit_normal_equations(self.handle, ptr(ata), ptr(atb), ptr(sp), ptr(sv)), self.ctx.handle)
        return ata, atb, sp, float(sv[0])

    @staticmethod
    def _constraints(con):
        nr = int(con["n_reduced"])
        c = RpConstraints()
        c.q_tot = float(con["q_tot"])
        c.n_reduced = nr
        arrays = {}
        for name, dt, ct in (("red_ptr", np.int32, ctypes.c_int32), ("red_src", np.int32, ctypes.c_int32),
                             ("red_mult", np.float64, ctypes.c_double), ("weights", np.float64, ctypes.c_double),
                             ("diag_add", np.float64, ctypes.c_double), ("rhs_add", np.float64, ctypes.c_double),
                             ("expand_src", np.int32, ctypes.c_int32)):
            a = con.get(name)
            if a is None:
                setattr(c, name, None)
                continue
            arrays[name] = np.ascontiguousarray(a, dtype=dt)
            setattr(c, name, ptr(arrays[name], ct))
        return c, arrays

    def solve(self, con):
        """``con``: dict from :func:`pyrepeat_amd.params.constraint_arrays`."""
        N = self.n_atoms
        nr = int(con["n_reduced"])
        c, keep = self._constraints(con)
        q = np.empty(N)
        sol = np.empty(nr + 1)
        fb = ctypes.c_int32(0)
        check(load().rp_fit_solve(self.handle, ctypes.byref(c), ptr(q), ptr(sol), ctypes.byref(fb)), self.ctx.handle)
        self.last_solver = int(fb.value)
        self.last_solve_used_lu = self.last_solver != RP_SOLVER_CHOLESKY
        return q, sol

    def solve_report(self):
        """How the last solve went: solver (RP_SOLVER_*), refinement steps, last correction, error bound."""
        r = RpSolveReport()
        check(load().rp_fit_solve_report(self.handle, ctypes.byref(r)), self.ctx.handle)
        return {"solver": int(r.solver), "refine_steps": int(r.refine_steps),
                "last_correction": float(r.last_correction), "error_bound": float(r.error_bound),
                "within_target": bool(r.within_target), "sensitivity": float(r.sensitivity),
                "well_conditioned": bool(r.well_conditioned)}

    def refit(self, con, c_prepared=None):
        """Whole fit (filter, accumulate, solve) from the device-resident cube."""
        c, keep = c_prepared if c_prepared is not None else self._constraints(con)
        q = np.empty(self.n_atoms)
        check(load().rp_fit_refit(self.handle, ctypes.byref(c), ptr(q)), self.ctx.handle)
        self.info = self._info()
        return q

    def phi_matvec(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        out = np.empty(self.info.n_keep)
        check(load().rp_fit_phi_matvec(self.handle, ptr(q), ptr(out)), self.ctx.handle)
        return out

    def phi(self, grid_xyz):
        g = np.ascontiguousarray(grid_xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty((len(g), self.n_atoms))
        check(load().rp_fit_phi(self.handle, len(g), ptr(g), ptr(out)), self.ctx.handle)
        return out

    def phi_noalloc(self, grid_xyz):
        """``phi`` through the entry point that allocates no device memory (static per-GPU scratch)."""
        g = np.ascontiguousarray(grid_xyz, dtype=np.float64).reshape(-1, 3)
        out = np.empty((len(g), self.n_atoms))
        check(load().rp_fit_phi_noalloc(self.handle, len(g), ptr(g), ptr(out)), self.ctx.handle)
        return out

    def qeq(self, chi, j_self, q_tot=0.0, want_j=False):
        """QEq charges of the fit's atoms (and the N x N Coulomb matrix when ``want_j``)."""
        chi = np.ascontiguousarray(chi, dtype=np.float64)
        js = np.ascontiguousarray(j_self, dtype=np.float64)
        q = np.empty(self.n_atoms)
        J = np.empty((self.n_atoms, self.n_atoms)) if want_j else None
        check(load().rp_fit_qeq(self.handle, ptr(chi), ptr(js), float(q_tot), ptr(q), ptr(J)), self.ctx.handle)
        return (q, J) if want_j else q

    def refine_stats(self):
        """Refinement of the last solve (split form with refinement): steps run and max |dq| (e) of
        the last correction."""
        steps, last = ctypes.c_int32(0), ctypes.c_double(0.0)
        check(load().rp_fit_refine_stats(self.handle, ctypes.byref(steps), ctypes.byref(last)), self.ctx.handle)
        return int(steps.value), float(last.value)

    def timings(self):
        t = np.zeros(3)
        check(load().rp_fit_timings(self.handle, ptr(t)), self.ctx.handle)
        return {"filter_ms": t[0], "accumulate_ms": t[1], "solve_ms": t[2]}

    def close(self):
        if getattr(self, "handle", None):
            load().rp_fit_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Serving loop: concurrent fits in ``slots`` worker slots on one device within ``hbm_budget`` bytes.

    ``submit`` copies the request and returns a ticket; ``poll`` returns ``None`` while it is queued or
    running, else ``(charges, result)`` with ``result`` a dict of the server-clock times and status.
    """

    def __init__(self, device=0, slots=4, hbm_budget=200 * 10**9):
        L = load()
        h = ctypes.c_void_p()
        rc = L.rp_server_create(int(device), int(slots), int(hbm_budget), ctypes.byref(h))
        if rc != RP_OK:
            raise RepeatHipError(rc, L.rp_last_error(None).decode(errors="replace"))
        self.handle = h
        self._n_atoms = {}

    @staticmethod
    def request(system_args, con, options):
        """``(struct, keepalive)`` for ``system_args`` = (n_grid, axis, Z, xyz, esp[, esp_order])."""
        sysd, arrays = make_system(*system_args)
        c, carr = Fit._constraints(con)
        req = RpRequest()
        req.system = sysd
        req.options = options
        req.constraints = c
        return req, (arrays, carr)

    @staticmethod
    def request_bytes(system_args, con, options):
        req, keep = Server.request(system_args, con, options)
        v = ctypes.c_uint64()
        rc = load().rp_request_bytes(ctypes.byref(req), ctypes.byref(v))
        if rc != RP_OK:
            raise RepeatHipError(rc, load().rp_last_error(None).decode(errors="replace"))
        return int(v.value)

    def now(self):
        return float(load().rp_server_now(self.handle))

    def submit(self, system_args, con, options):
        req, keep = self.request(system_args, con, options)
        t = ctypes.c_uint64()
        rc = load().rp_submit(self.handle, ctypes.byref(req), ctypes.byref(t))
        if rc != RP_OK:
            raise RepeatHipError(rc, load().rp_server_error(self.handle).decode(errors="replace"))
        self._n_atoms[t.value] = int(req.system.n_atoms)
        return int(t.value)

    def poll(self, ticket, wait_ms=-1):
        res = RpResult()
        q = np.empty(self._n_atoms[ticket])
        rc = load().rp_poll(self.handle, ctypes.c_uint64(ticket), int(wait_ms), ctypes.byref(res), ptr(q))
        if rc == RP_EAGAIN:
            return None
        if rc != RP_OK:
            raise RepeatHipError(rc, "rp_poll: unknown ticket")
        del self._n_atoms[ticket]
        out = {"status": res.status, "slot": res.slot, "n_keep": res.n_keep, "t_submit": res.t_submit,
               "t_admit": res.t_admit, "t_done": res.t_done, "device_ms": list(res.device_ms),
               "bytes_bound": int(res.bytes_bound), "error": res.error.decode(errors="replace")}
        return q, out

    def stats(self):
        st = RpServerStats()
        check(load().rp_server_stats_get(self.handle, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in RpServerStats._fields_}

    def close(self):
        if getattr(self, "handle", None):
            load().rp_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
