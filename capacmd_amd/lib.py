"""ctypes bindings of the two C ABIs (include/capac_host.h, include/capac_gpu.h).

The shared libraries are built in-tree by capacmd_amd/build.py.  Loading is
strict: a missing libcapac_gpu.so raises instead of falling back to anything.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
HOST_SO = os.path.join(HERE, "libcapac_host.so")
GPU_SO = os.path.join(HERE, "libcapac_gpu.so")

NPOT = 15
NHC = 10

PREC_FP64 = 0
PREC_FP32_REFINED = 1

SOLVE_KINDS = {0: "converged", 1: "loosened", 2: "failed", 3: "zero_rhs", 4: "zero_rho"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_up = C.POINTER(C.c_uint)
_lp = C.POINTER(C.c_long)


class CapacParams(C.Structure):
    """Mirror of capac_params (include/capac_gpu.h)."""
    _fields_ = [
        ("n_atoms", C.c_int), ("n_metal", C.c_int), ("n_nps", C.c_int),
        ("n_types", C.c_int), ("n_mols", C.c_int), ("n_cstr", C.c_int),
        ("charge", _dp), ("mass", _dp), ("atomtype", _ip), ("mol_id", _ip),
        ("local_idx", _ip), ("excl_mask", _up),
        ("mol_first", _ip), ("mol_last", _ip), ("mol_cstr", _ip),
        ("cstr_i", _ip), ("cstr_j", _ip), ("cstr_r", _dp),
        ("np_first", _ip), ("np_last", _ip), ("np_charge", _dp),
        ("nb_funct", _ip), ("nb_c6", _dp), ("nb_c12", _dp), ("nb_a", _dp), ("nb_b", _dp),
        ("box", C.c_double * 3), ("rcut", C.c_double), ("dsf_alpha", C.c_double),
        ("use_coulomb", C.c_int), ("use_vdw", C.c_int),
        ("qsc_n", C.c_int), ("qsc_m", C.c_int),
        ("qsc_eps", C.c_double), ("qsc_c", C.c_double), ("qsc_a", C.c_double),
        ("use_cpff", C.c_int), ("cpff_polar", C.c_double), ("cpff_capac", C.c_double),
        ("efield", C.c_double * 3),
        ("dt", C.c_double), ("fix_metal", C.c_int), ("nvt", C.c_int),
        ("ref_temp", C.c_double), ("tau_t", C.c_double), ("n_heating", C.c_int),
        ("ndf", C.c_int),
        ("n_bonds", C.c_int), ("n_pairs14", C.c_int), ("n_angles", C.c_int),
        ("n_dihedrals", C.c_int), ("n_vsites", C.c_int),
        ("mol_terms", _ip), ("bond_atoms", _ip), ("bond_funct", _ip), ("bond_par", _dp),
        ("pair14_atoms", _ip), ("pair14_par", _dp),
        ("angle_atoms", _ip), ("angle_funct", _ip), ("angle_par", _dp),
        ("dihedral_atoms", _ip), ("dihedral_par", _dp),
        ("vsite_atoms", _ip), ("vsite_par", _dp),
    ]


class CapacTiming(C.Structure):
    _fields_ = [("t", C.c_double * NPOT), ("n_force_calls", C.c_longlong),
                ("solve_iters", C.c_longlong), ("nb_builds", C.c_longlong),
                ("solve_syncs", C.c_longlong)]


def dptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def iptr(a: Optional[np.ndarray]):
    if a is None:
        return None
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


_host = None
_gpu = None


def host_lib() -> C.CDLL:
    global _host
    if _host is None:
        if not os.path.exists(HOST_SO):
            raise RuntimeError(f"{HOST_SO} is not built; run capacmd_amd/build.py")
        lib = C.CDLL(HOST_SO)
        lib.capac_deck_load.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p,
                                        C.POINTER(C.c_void_p), C.c_char_p, C.c_int]
        lib.capac_deck_load.restype = C.c_int
        lib.capac_deck_free.argtypes = [C.c_void_p]
        lib.capac_deck_n_mat.argtypes = [C.c_void_p]
        lib.capac_deck_params.argtypes = [C.c_void_p]
        lib.capac_deck_params.restype = C.POINTER(CapacParams)
        lib.capac_deck_positions.argtypes = [C.c_void_p]
        lib.capac_deck_positions.restype = _dp
        lib.capac_deck_velocities.argtypes = [C.c_void_p]
        lib.capac_deck_velocities.restype = _dp
        lib.capac_deck_run.argtypes = [C.c_void_p, _ip, _ip, _dp, _dp]
        lib.capac_deck_minimiser.argtypes = [C.c_void_p, _ip, _dp, _dp]
        lib.capac_params_size.restype = C.c_int
        for fn in ("capac_write_gro_path",):
            getattr(lib, fn).argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, _dp, _dp,
                                         C.c_double, C.c_double, C.c_int]
        lib.capac_write_vec_pq_path.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, _dp, C.c_int]
        lib.capac_write_binary_path.argtypes = [C.c_char_p, C.c_char_p, C.c_void_p, _dp, C.c_int]
        if lib.capac_params_size() != C.sizeof(CapacParams):
            raise RuntimeError("capac_params layout mismatch between C and ctypes")
        _host = lib
    return _host


class Deck:
    """An input deck read by the C host library (init.gro, param.txt, mdset.txt)."""

    def __init__(self, gro: str, par: str, mds: str):
        lib = host_lib()
        h = C.c_void_p()
        err = C.create_string_buffer(1024)
        rc = lib.capac_deck_load(gro.encode(), par.encode(), mds.encode(), C.byref(h), err, 1024)
        if rc != 0:
            raise ValueError(err.value.decode(errors="replace").strip())
        self._h = h
        self._lib = lib
        self.params_ptr = lib.capac_deck_params(h)
        self.p = self.params_ptr.contents
        self.n_atoms = self.p.n_atoms
        self.n_metal = self.p.n_metal
        self.n_mat = lib.capac_deck_n_mat(h)
        n3 = 3 * self.n_atoms
        self.r0 = np.ctypeslib.as_array(lib.capac_deck_positions(h), shape=(n3,)).copy()
        self.v0 = np.ctypeslib.as_array(lib.capac_deck_velocities(h), shape=(n3,)).copy()
        ns, nsave = C.c_int(), C.c_int()
        vq0, vp = C.c_double(), C.c_double()
        lib.capac_deck_run(h, C.byref(ns), C.byref(nsave), C.byref(vq0), C.byref(vp))
        self.n_steps, self.n_save, self.vq0, self.vp = ns.value, nsave.value, vq0.value, vp.value
        es, el, et = C.c_int(), C.c_double(), C.c_double()
        kind = lib.capac_deck_minimiser(h, C.byref(es), C.byref(el), C.byref(et))
        self.run_type = {0: "md", 1: "em", 2: "cg"}.get(kind, "?")
        self.em_steps, self.em_length, self.em_tol = es.value, el.value, et.value

    @classmethod
    def from_dir(cls, d: str) -> "Deck":
        return cls(os.path.join(d, "init.gro"), os.path.join(d, "param.txt"),
                   os.path.join(d, "mdset.txt"))

    def array(self, name: str, n: int, dtype=np.float64) -> np.ndarray:
        ptr = getattr(self.p, name)
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

    @property
    def box(self) -> np.ndarray:
        return np.array(self.p.box[:])

    def close(self):
        if self._h:
            self._lib.capac_deck_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ----------------------------------------------------------------- GPU ABI

GPU_EXPORTS = [
    "capac_last_error", "capac_version", "capac_create", "capac_destroy",
    "capac_set_state", "capac_get_state", "capac_get_forces", "capac_set_box",
    "capac_force", "capac_neighbours", "capac_nonbonded", "capac_sutton_chen",
    "capac_cpim_rhs", "capac_cpim_assemble", "capac_cpim_solve", "capac_cpim_solve_noalloc",
    "capac_cpim_force",
    "capac_solve_stats", "capac_md_begin", "capac_md_run", "capac_md_observables",
    "capac_set_thermostat", "capac_get_timing", "capac_reset_timing", "capac_enable_timing",
    "capac_time_kernel", "capac_set_skin", "capac_skin", "capac_set_solve_tolerance",
    "capac_set_theta", "capac_get_theta", "capac_fit_frame",
    "capac_em_begin", "capac_em_step", "capac_em_direction",
    "capac_checkpoint_bytes", "capac_checkpoint_save", "capac_checkpoint_load",
    "capac_set_predictor", "capac_predictor", "capac_set_fp32_layout", "capac_fp32_layout",
    "capac_set_block_precond", "capac_block_precond", "capac_block_precond_inverse",
    "capac_block_precond_coarse",
    "capac_bonded", "capac_vsite_build", "capac_vsite_spread",
]
NTHETA = 6

LAYOUT_AUTO, LAYOUT_ROWS, LAYOUT_TILES = 0, 1, 2      # CAPAC_LAYOUT_*
BJ_EVERY_DEFAULT = 32       # CAPAC_BJ_EVERY_DEFAULT
KERNEL_SPMV, KERNEL_ASSEMBLE, KERNEL_PAIR, KERNEL_NEIGHBOUR, KERNEL_CPIM_MM, KERNEL_APPLY64, KERNEL_SOLVE = range(7)
DEFAULT_SKIN = 0.1          # nm, CAPAC_DEFAULT_SKIN


def gpu_lib() -> C.CDLL:
    """Loads libcapac_gpu.so; raises if it is missing (no fallback path)."""
    global _gpu
    if _gpu is None:
        if not os.path.exists(GPU_SO):
            raise RuntimeError(f"{GPU_SO} is not built; run capacmd_amd/build.py")
        lib = C.CDLL(GPU_SO)
        lib.capac_last_error.restype = C.c_char_p
        lib.capac_version.restype = C.c_char_p
        lib.capac_create.argtypes = [C.POINTER(CapacParams), C.c_int, C.c_int, C.POINTER(C.c_void_p)]
        lib.capac_destroy.argtypes = [C.c_void_p]
        lib.capac_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.capac_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.capac_get_forces.argtypes = [C.c_void_p, _dp]
        lib.capac_set_box.argtypes = [C.c_void_p, _dp]
        lib.capac_force.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.capac_neighbours.argtypes = [C.c_void_p, _ip, _ip, C.c_int]
        lib.capac_nonbonded.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.capac_bonded.argtypes = [C.c_void_p, _dp, _dp, _dp]
        lib.capac_vsite_build.argtypes = [C.c_void_p, _dp]
        lib.capac_vsite_spread.argtypes = [C.c_void_p, _dp, _dp]
        lib.capac_sutton_chen.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        lib.capac_cpim_rhs.argtypes = [C.c_void_p, _dp]
        lib.capac_cpim_assemble.argtypes = [C.c_void_p, _dp, _dp]
        lib.capac_cpim_solve.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, _ip]
        lib.capac_cpim_solve_noalloc.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp, _ip]
        lib.capac_cpim_force.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        lib.capac_solve_stats.argtypes = [C.c_void_p, _ip, _dp, _dp, _ip]
        lib.capac_md_begin.argtypes = [C.c_void_p]
        lib.capac_md_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int]
        lib.capac_md_observables.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        lib.capac_set_thermostat.argtypes = [C.c_void_p, _dp, C.c_double]
        lib.capac_get_timing.argtypes = [C.c_void_p, C.POINTER(CapacTiming)]
        lib.capac_reset_timing.argtypes = [C.c_void_p]
        lib.capac_enable_timing.argtypes = [C.c_void_p, C.c_int]
        lib.capac_time_kernel.argtypes = [C.c_void_p, C.c_int, C.c_int, _dp]
        lib.capac_set_skin.argtypes = [C.c_void_p, C.c_double]
        lib.capac_skin.argtypes = [C.c_void_p]
        lib.capac_skin.restype = C.c_double
        lib.capac_set_solve_tolerance.argtypes = [C.c_void_p, C.c_double]
        lib.capac_set_theta.argtypes = [C.c_void_p, _dp]
        lib.capac_get_theta.argtypes = [C.c_void_p, _dp]
        lib.capac_fit_frame.argtypes = [C.c_void_p, _dp, C.c_double, C.c_double, _dp, _dp]
        lib.capac_em_begin.argtypes = [C.c_void_p]
        lib.capac_em_step.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        lib.capac_em_direction.argtypes = [C.c_void_p, C.c_int]
        lib.capac_set_predictor.argtypes = [C.c_void_p, C.c_int]
        lib.capac_predictor.argtypes = [C.c_void_p]
        lib.capac_set_fp32_layout.argtypes = [C.c_void_p, C.c_int]
        lib.capac_fp32_layout.argtypes = [C.c_void_p]
        lib.capac_set_block_precond.argtypes = [C.c_void_p, C.c_int]
        lib.capac_block_precond.argtypes = [C.c_void_p]
        lib.capac_block_precond_inverse.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_float)]
        lib.capac_block_precond_coarse.argtypes = [C.c_void_p, _ip, _dp, C.POINTER(C.c_float)]
        lib.capac_checkpoint_bytes.argtypes = [C.c_void_p]
        lib.capac_checkpoint_bytes.restype = C.c_size_t
        lib.capac_checkpoint_save.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_size_t]
        lib.capac_checkpoint_load.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, _ip]
        _gpu = lib
    return _gpu


class CapacError(RuntimeError):
    pass


def _check(rc: int, what: str) -> int:
    if rc < 0:
        msg = gpu_lib().capac_last_error()
        raise CapacError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")
    return rc


class Context:
    """One device-resident box (capac_create / capac_destroy)."""

    def __init__(self, deck: Deck, precision: int = PREC_FP64, device: int = 0):
        self.lib = gpu_lib()
        self.deck = deck
        self.N = deck.n_atoms
        self.n_mat = deck.n_mat
        self.n_metal = deck.n_metal
        h = C.c_void_p()
        _check(self.lib.capac_create(deck.params_ptr, precision, device, C.byref(h)), "capac_create")
        self.h = h
        self.set_state(deck.r0, deck.v0, np.zeros(max(self.n_mat, 1)))

    def close(self):
        if getattr(self, "h", None):
            self.lib.capac_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # state
    def set_state(self, r=None, v=None, x=None):
        r = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        x = None if x is None or self.n_mat == 0 else np.ascontiguousarray(x, dtype=np.float64)
        _check(self.lib.capac_set_state(self.h, dptr(r), dptr(v), dptr(x)), "capac_set_state")

    def get_state(self):
        r = np.empty(3 * self.N); v = np.empty(3 * self.N); x = np.empty(max(self.n_mat, 1))
        _check(self.lib.capac_get_state(self.h, dptr(r), dptr(v), dptr(x)), "capac_get_state")
        return r, v, x[: self.n_mat]

    def forces(self):
        f = np.empty(3 * self.N)
        _check(self.lib.capac_get_forces(self.h, dptr(f)), "capac_get_forces")
        return f

    def set_skin(self, skin: float):
        """Verlet skin of the neighbour rows in nm (0: rebuild every call)."""
        _check(self.lib.capac_set_skin(self.h, float(skin)), "capac_set_skin")

    def set_solve_tolerance(self, crit: float):
        """BiCGSTAB stop test rTr < crit bTb (the reference's crit is 1e-8)."""
        _check(self.lib.capac_set_solve_tolerance(self.h, float(crit)), "capac_set_solve_tolerance")

    def set_predictor(self, order: int):
        """Warm start of the MD solves: 0 previous solution (the reference's),
        1 linear, 2 quadratic extrapolation (capac_set_predictor)."""
        _check(self.lib.capac_set_predictor(self.h, int(order)), "capac_set_predictor")

    @property
    def predictor(self) -> int:
        return int(self.lib.capac_predictor(self.h))

    def set_fp32_layout(self, layout: int):
        """Storage layout of the FP32 relay: LAYOUT_AUTO, LAYOUT_ROWS or
        LAYOUT_TILES (capac_set_fp32_layout)."""
        _check(self.lib.capac_set_fp32_layout(self.h, int(layout)), "capac_set_fp32_layout")

    @property
    def fp32_layout(self) -> int:
        return int(self.lib.capac_fp32_layout(self.h))

    def set_block_precond(self, every: int):
        """GMRES preconditioner of the FP32-refined path: block Jacobi over the
        relay's diagonal blocks of 64 metal atoms, inverted every `every`
        assemblies, or the Jacobi diagonal for 0 (capac_set_block_precond)."""
        _check(self.lib.capac_set_block_precond(self.h, int(every)), "capac_set_block_precond")

    @property
    def block_precond(self) -> int:
        return int(self.lib.capac_block_precond(self.h))

    def block_precond_inverse(self, block: int) -> np.ndarray:
        """Block `block`'s inverse as last formed, 256 x 256 FP32 (unknowns:
        dipoles of its 64 atoms, then their charges)."""
        out = np.zeros((256, 256), dtype=np.float32)
        _check(self.lib.capac_block_precond_inverse(self.h, int(block), out.ctypes.data_as(C.POINTER(C.c_float))),
               "capac_block_precond_inverse")
        return out

    def block_precond_coarse(self):
        """The coarse space as last formed: (Ac^-1 as n x n, basis as
        (ceil(Nm/64) * 64, 4) entries over the metal charges), or None when
        it is off or not formed."""
        n = C.c_int()
        _check(self.lib.capac_block_precond_coarse(self.h, C.byref(n), None, None), "capac_block_precond_coarse")
        if n.value == 0:
            return None
        nb = (self.n_metal + 63) // 64
        ainv = np.zeros((n.value, n.value))
        basis = np.zeros((nb * 64, 4), dtype=np.float32)
        _check(self.lib.capac_block_precond_coarse(self.h, C.byref(n), dptr(ainv),
                                                   basis.ctypes.data_as(C.POINTER(C.c_float))),
               "capac_block_precond_coarse")
        return ainv, basis

    def set_theta(self, theta):
        """Fitting parameters (alpha, c, R_q, R_p, eps_SC, c_SC); see capac_set_theta."""
        t = np.ascontiguousarray(theta, dtype=np.float64)
        assert t.shape == (NTHETA,)
        _check(self.lib.capac_set_theta(self.h, dptr(t)), "capac_set_theta")

    def get_theta(self) -> np.ndarray:
        t = np.empty(NTHETA)
        _check(self.lib.capac_get_theta(self.h, dptr(t)), "capac_get_theta")
        return t

    # energy minimisation
    def em_begin(self):
        _check(self.lib.capac_em_begin(self.h), "capac_em_begin")

    def em_step(self, length: float):
        pot = np.zeros(NPOT); fmax = C.c_double(); frms = C.c_double()
        _check(self.lib.capac_em_step(self.h, float(length), dptr(pot), C.byref(fmax), C.byref(frms)),
               "capac_em_step")
        return pot, fmax.value, frms.value

    def em_direction(self, cg: bool):
        _check(self.lib.capac_em_direction(self.h, int(bool(cg))), "capac_em_direction")

    # restart checkpoint
    def checkpoint(self, step: int) -> bytes:
        """Everything the MD loop carries between steps, as bytes
        (capac_checkpoint_save)."""
        n = self.lib.capac_checkpoint_bytes(self.h)
        buf = C.create_string_buffer(n)
        _check(self.lib.capac_checkpoint_save(self.h, int(step), buf, n), "capac_checkpoint_save")
        return buf.raw

    def restore(self, blob: bytes) -> int:
        """Loads a checkpoint (capac_checkpoint_load); returns its step."""
        step = C.c_int()
        _check(self.lib.capac_checkpoint_load(self.h, blob, len(blob), C.byref(step)), "capac_checkpoint_load")
        return step.value

    def minimise(self, cg: bool, steps: int, length: float, tol: float, on_step=None):
        """The reference's minimiser loop (mpi-main.c:706-865) after the first
        force call: capac_md_begin, then up to `steps` moves along s / F_max
        scaled by `length`, which grows by 1.2 after a step that lowers the
        energy and shrinks by 0.2 otherwise; stops when F_max < tol,
        F_rms < tol / 2 and |dE| < tol / 10.  cg False is steepest descent.
        on_step(step, fmax, frms, pot) runs after every step.  Returns the
        step lines [(step, F_max, E)] from step 0 and whether it converged."""
        self.md_begin()
        obs = self.observables()
        self.em_begin()
        e = obs["pot"][0]
        lines = [(0, obs["fmax"], e)]
        for step in range(1, steps + 1):
            pot, fmax, frms = self.em_step(length)
            delta = pot[0] - e
            e = pot[0]
            length *= 1.2 if delta <= 0.0 else 0.2
            lines.append((step, fmax, e))
            if on_step is not None:
                on_step(step, fmax, frms, pot)
            if fmax < tol and frms < tol * 0.5 and abs(delta) < tol * 0.1:
                return lines, True
            self.em_direction(cg)
        return lines, False

    def fit_frame(self, f_ref, e_ref: float, w_e: float, grad: bool = True):
        """Loss of one frame at the resident state and, with grad, dL/dtheta."""
        f = np.ascontiguousarray(f_ref, dtype=np.float64)
        assert f.shape == (3 * self.n_metal,)
        loss = C.c_double()
        g = np.zeros(NTHETA) if grad else None
        _check(self.lib.capac_fit_frame(self.h, dptr(f), float(e_ref), float(w_e), C.byref(loss),
                                        dptr(g)), "capac_fit_frame")
        return loss.value, g

    @property
    def skin(self) -> float:
        return float(self.lib.capac_skin(self.h))

    # operators
    def force(self):
        pot = np.zeros(NPOT); vir = np.zeros(9); f = np.empty(3 * self.N)
        _check(self.lib.capac_force(self.h, dptr(pot), dptr(vir), dptr(f)), "capac_force")
        return f, pot, vir

    def neighbours(self, max_per_atom: int = 2048):
        cnt = np.zeros(self.N, dtype=np.int32)
        lst = np.zeros(self.N * max_per_atom, dtype=np.int32)
        m = _check(self.lib.capac_neighbours(self.h, iptr(cnt), iptr(lst), max_per_atom),
                   "capac_neighbours")
        return cnt, lst.reshape(self.N, max_per_atom), m

    def neighbour_counts(self):
        """entries of each atom's full neighbour row (rc + skin), no rows"""
        cnt = np.zeros(self.N, dtype=np.int32)
        _check(self.lib.capac_neighbours(self.h, iptr(cnt), None, 0), "capac_neighbours")
        return cnt

    def nonbonded(self):
        f = np.empty(3 * self.N); pot = np.zeros(NPOT); vir = np.zeros(9)
        _check(self.lib.capac_nonbonded(self.h, dptr(f), dptr(pot), dptr(vir)), "capac_nonbonded")
        return f, pot, vir

    def bonded(self):
        """mpi_bonded at the resident positions: f, pot (2, 3, 4, 7, 8), vir"""
        f = np.empty(3 * self.N); pot = np.zeros(NPOT); vir = np.zeros(9)
        _check(self.lib.capac_bonded(self.h, dptr(f), dptr(pot), dptr(vir)), "capac_bonded")
        return f, pot, vir

    def vsite_build(self):
        """build_vsites on the resident positions; returns them (3N)"""
        r = np.empty(3 * self.N)
        _check(self.lib.capac_vsite_build(self.h, dptr(r)), "capac_vsite_build")
        return r

    def vsite_spread(self, f):
        """mpi_vsites: the forces on virtual sites moved to their atoms"""
        f_in = np.ascontiguousarray(f, dtype=np.float64)
        out = np.empty(3 * self.N)
        _check(self.lib.capac_vsite_spread(self.h, dptr(f_in), dptr(out)), "capac_vsite_spread")
        return out

    def sutton_chen(self):
        f = np.empty(3 * self.N); pot = np.zeros(NPOT); vir = np.zeros(9)
        isd = np.empty(max(self.n_metal, 1))
        _check(self.lib.capac_sutton_chen(self.h, dptr(f), dptr(pot), dptr(vir), dptr(isd)),
               "capac_sutton_chen")
        return f, pot, vir, isd[: self.n_metal]

    def cpim_rhs(self):
        b = np.empty(self.n_mat)
        _check(self.lib.capac_cpim_rhs(self.h, dptr(b)), "capac_cpim_rhs")
        return b

    def cpim_assemble(self, dense: bool = True):
        n4 = 4 * self.n_metal
        a = np.empty(n4 * n4) if dense else None
        diag = np.empty(self.n_mat)
        _check(self.lib.capac_cpim_assemble(self.h, dptr(a), dptr(diag)), "capac_cpim_assemble")
        return (a.reshape(n4, n4) if dense else None), diag

    def cpim_solve(self, b=None, x=None):
        it = C.c_int(); rtr = C.c_double(); btb = C.c_double(); kind = C.c_int()
        b = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        _check(self.lib.capac_cpim_solve(self.h, dptr(b), dptr(x), C.byref(it), C.byref(rtr),
                                         C.byref(btb), C.byref(kind)), "capac_cpim_solve")
        return x, it.value, rtr.value, btb.value, kind.value

    def cpim_solve_noalloc(self, b=None, x=None):
        """cpim_solve with the work space in the library's device array
        (capac_cpim_solve_noalloc: the call allocates nothing)."""
        it = C.c_int(); rtr = C.c_double(); btb = C.c_double(); kind = C.c_int()
        b = None if b is None else np.ascontiguousarray(b, dtype=np.float64)
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64).copy()
        _check(self.lib.capac_cpim_solve_noalloc(self.h, dptr(b), dptr(x), C.byref(it), C.byref(rtr),
                                                 C.byref(btb), C.byref(kind)), "capac_cpim_solve_noalloc")
        return x, it.value, rtr.value, btb.value, kind.value

    def cpim_force(self, x=None):
        f = np.empty(3 * self.N); pot = np.zeros(NPOT); vir = np.zeros(9)
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        _check(self.lib.capac_cpim_force(self.h, dptr(x), dptr(f), dptr(pot), dptr(vir)),
               "capac_cpim_force")
        return f, pot, vir

    def solve_stats(self):
        it = C.c_int(); rtr = C.c_double(); btb = C.c_double(); kind = C.c_int()
        _check(self.lib.capac_solve_stats(self.h, C.byref(it), C.byref(rtr), C.byref(btb),
                                          C.byref(kind)), "capac_solve_stats")
        return it.value, rtr.value, btb.value, kind.value

    # MD
    def md_begin(self):
        _check(self.lib.capac_md_begin(self.h), "capac_md_begin")

    def md_run(self, first_step: int, n_steps: int, n_save: int, stop_at_save: bool = False) -> int:
        return _check(self.lib.capac_md_run(self.h, first_step, n_steps, n_save, int(stop_at_save)),
                      "capac_md_run")

    def observables(self):
        ek = C.c_double(); t = C.c_double(); rt = C.c_double()
        vq = np.zeros(NHC); fmax = C.c_double(); frms = C.c_double()
        pot = np.zeros(NPOT); vir = np.zeros(9)
        _check(self.lib.capac_md_observables(self.h, C.byref(ek), C.byref(t), C.byref(rt), dptr(vq),
                                             C.byref(fmax), C.byref(frms), dptr(pot), dptr(vir)),
               "capac_md_observables")
        return dict(eksum=ek.value, temp=t.value, ref_temp=rt.value, vq=vq, fmax=fmax.value,
                    frms=frms.value, pot=pot, vir=vir)

    def set_thermostat(self, vq, vp: float = 0.0):
        vq = np.ascontiguousarray(vq, dtype=np.float64)
        _check(self.lib.capac_set_thermostat(self.h, dptr(vq), vp), "capac_set_thermostat")

    def timing(self):
        t = CapacTiming()
        _check(self.lib.capac_get_timing(self.h, C.byref(t)), "capac_get_timing")
        return np.array(t.t[:]), t.n_force_calls, t.solve_iters

    def solve_syncs(self) -> int:
        """Host synchronisations inside the solves since the last reset_timing."""
        t = CapacTiming()
        _check(self.lib.capac_get_timing(self.h, C.byref(t)), "capac_get_timing")
        return int(t.solve_syncs)

    def nb_builds(self) -> int:
        """Neighbour-row builds since the last reset_timing."""
        t = CapacTiming()
        _check(self.lib.capac_get_timing(self.h, C.byref(t)), "capac_get_timing")
        return int(t.nb_builds)

    def reset_timing(self):
        _check(self.lib.capac_reset_timing(self.h), "capac_reset_timing")

    def time_kernel(self, kernel: int, reps: int = 20) -> float:
        ms = C.c_double()
        _check(self.lib.capac_time_kernel(self.h, kernel, reps, C.byref(ms)), "capac_time_kernel")
        return ms.value

    def enable_timing(self, on: bool = True):
        _check(self.lib.capac_enable_timing(self.h, int(on)), "capac_enable_timing")
