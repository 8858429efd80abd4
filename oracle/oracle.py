"""ctypes wrapper of the CPU oracle (oracle/cpu_ref.c).

TEST INFRASTRUCTURE: imported only by tests/, bench.py's cpu_baseline leg and
__graft_entry__.smoke(), as the checker.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_dp = C.POINTER(C.c_double)
_lp = C.POINTER(C.c_long)
_ip = C.POINTER(C.c_int)

_libs = {}


def lib(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    if name not in _libs:
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        L.oref_create.argtypes = [C.c_void_p]
        L.oref_create.restype = C.c_void_p
        L.oref_destroy.argtypes = [C.c_void_p]
        L.oref_n_mat.argtypes = [C.c_void_p]
        L.oref_set_state.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.oref_get_state.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oref_set_int_pair_count.argtypes = [C.c_void_p, C.c_int]
        L.oref_set_theta.argtypes = [C.c_void_p, _dp]
        L.oref_reference_pair_end.argtypes = [C.c_int]
        L.oref_reference_pair_end.restype = C.c_long
        L.oref_nonb.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oref_qsc.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oref_bonded.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oref_vsite_build.argtypes = [C.c_void_p, _dp]
        L.oref_vsite_spread.argtypes = [C.c_void_p, _dp, _dp]
        L.oref_vec_ext.argtypes = [C.c_void_p, _dp, _dp]
        L.oref_relay.argtypes = [C.c_void_p, _dp]
        L.oref_relay.restype = C.c_long
        L.oref_relay_coo.argtypes = [C.c_void_p, _lp, _lp, _dp, _dp]
        L.oref_relay_dense.argtypes = [C.c_void_p, _dp]
        L.oref_relay_apply.argtypes = [C.c_void_p, _dp, _dp]
        L.oref_solve.argtypes = [C.c_void_p, _dp, _dp, _ip, _dp, _dp]
        L.oref_cpff_force.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.oref_force.argtypes = [C.c_void_p]
        L.oref_get_force_out.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.oref_solve_stats.argtypes = [C.c_void_p, _ip, _dp, _dp, _ip]
        L.oref_md_begin.argtypes = [C.c_void_p]
        L.oref_md_step.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.oref_md_observables.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.oref_set_thermostat.argtypes = [C.c_void_p, _dp]
        L.oref_timers.argtypes = [C.c_void_p, _dp]
        L.oref_set_solve_crit.argtypes = [C.c_void_p, C.c_double]
        L.oref_em_begin.argtypes = [C.c_void_p]
        L.oref_em_step.argtypes = [C.c_void_p, C.c_double, _dp, _dp, _dp]
        L.oref_em_direction.argtypes = [C.c_void_p, C.c_int]
        _libs[name] = L
    return _libs[name]


def _d(a):
    return None if a is None else a.ctypes.data_as(_dp)


class Oracle:
    """FP64 CPU restatement of the reference path for one deck."""

    def __init__(self, deck, omp: bool = False):
        self.L = lib(omp)
        self.deck = deck
        self.h = self.L.oref_create(C.cast(deck.params_ptr, C.c_void_p))
        self.N = deck.n_atoms
        self.Nm = deck.n_metal
        self.n_mat = deck.n_mat
        self.set_state(deck.r0, deck.v0, np.zeros(max(self.n_mat, 1)))

    def close(self):
        if self.h:
            self.L.oref_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, r=None, v=None, x=None):
        r = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        v = None if v is None else np.ascontiguousarray(v, dtype=np.float64)
        x = None if x is None else np.ascontiguousarray(x, dtype=np.float64)
        self.L.oref_set_state(self.h, _d(r), _d(v), _d(x))

    def get_state(self):
        r = np.empty(3 * self.N); v = np.empty(3 * self.N); f = np.empty(3 * self.N)
        x = np.empty(max(self.n_mat, 1))
        self.L.oref_get_state(self.h, _d(r), _d(v), _d(x), _d(f))
        return r, v, x[: self.n_mat], f

    def int_pair_count(self, on: bool):
        self.L.oref_set_int_pair_count(self.h, int(on))

    def set_theta(self, theta):
        """Fitting parameters (alpha, c, R_q, R_p, eps_SC, c_SC), widths free."""
        t = np.ascontiguousarray(theta, dtype=np.float64)
        self.L.oref_set_theta(self.h, t.ctypes.data_as(_dp))

    def nonb(self, r):
        f = np.zeros(3 * self.N); pot = np.zeros(15); vir = np.zeros(9)
        self.L.oref_nonb(self.h, _d(np.ascontiguousarray(r)), _d(f), _d(pot), _d(vir))
        return f, pot, vir

    def bonded(self, r):
        """mpi_bonded (force.c:1694-1904)"""
        f = np.zeros(3 * self.N); pot = np.zeros(15); vir = np.zeros(9)
        self.L.oref_bonded(self.h, _d(np.ascontiguousarray(r)), _d(f), _d(pot), _d(vir))
        return f, pot, vir

    def vsite_build(self, r):
        """build_vsites (force.c:1582-1629) on a copy of r"""
        out = np.array(r, dtype=np.float64)
        self.L.oref_vsite_build(self.h, _d(out))
        return out

    def vsite_spread(self, r, f):
        """mpi_vsites (force.c:1636-1692) on a copy of f"""
        out = np.array(f, dtype=np.float64)
        self.L.oref_vsite_spread(self.h, _d(np.ascontiguousarray(r)), _d(out))
        return out

    def qsc(self, r):
        f = np.zeros(3 * self.N); pot = np.zeros(15); vir = np.zeros(9)
        isd = np.zeros(max(self.Nm, 1))
        self.L.oref_qsc(self.h, _d(np.ascontiguousarray(r)), _d(f), _d(pot), _d(vir), _d(isd))
        return f, pot, vir, isd[: self.Nm]

    def vec_ext(self, r):
        b = np.zeros(self.n_mat)
        self.L.oref_vec_ext(self.h, _d(np.ascontiguousarray(r)), _d(b))
        return b

    def relay(self, r):
        nnz = self.L.oref_relay(self.h, _d(np.ascontiguousarray(r)))
        row = np.empty(nnz, dtype=np.int64); col = np.empty(nnz, dtype=np.int64)
        val = np.empty(nnz); diag = np.empty(self.n_mat)
        self.L.oref_relay_coo(self.h, row.ctypes.data_as(_lp), col.ctypes.data_as(_lp), _d(val), _d(diag))
        return row, col, val, diag

    def relay_dense(self):
        a = np.empty(self.n_mat * self.n_mat)
        self.L.oref_relay_dense(self.h, _d(a))
        return a.reshape(self.n_mat, self.n_mat)

    def relay_apply(self, x):
        """A x with the relay assembled by the last relay() call."""
        y = np.empty(self.n_mat)
        self.L.oref_relay_apply(self.h, _d(np.ascontiguousarray(x, dtype=np.float64)), _d(y))
        return y

    def solve(self, b, x0):
        x = np.ascontiguousarray(x0, dtype=np.float64).copy()
        it = C.c_int(); rtr = C.c_double(); btb = C.c_double()
        kind = self.L.oref_solve(self.h, _d(np.ascontiguousarray(b)), _d(x), C.byref(it),
                                 C.byref(rtr), C.byref(btb))
        return x, it.value, rtr.value, btb.value, kind

    def cpff_force(self, r, x):
        f = np.zeros(3 * self.N); pot = np.zeros(15); vir = np.zeros(9)
        self.L.oref_cpff_force(self.h, _d(np.ascontiguousarray(r)), _d(np.ascontiguousarray(x)),
                               _d(f), _d(pot), _d(vir))
        return f, pot, vir

    def force(self):
        rc = self.L.oref_force(self.h)
        f = np.empty(3 * self.N); pot = np.empty(15); vir = np.empty(9)
        b = np.empty(max(self.n_mat, 1))
        self.L.oref_get_force_out(self.h, _d(f), _d(pot), _d(vir), _d(b))
        return rc, f, pot, vir, b[: self.n_mat]

    def solve_stats(self):
        it = C.c_int(); rtr = C.c_double(); btb = C.c_double(); kind = C.c_int()
        self.L.oref_solve_stats(self.h, C.byref(it), C.byref(rtr), C.byref(btb), C.byref(kind))
        return it.value, rtr.value, btb.value, kind.value

    def md_begin(self):
        self.L.oref_md_begin(self.h)

    def md_step(self, step: int, n_save: int) -> int:
        return self.L.oref_md_step(self.h, step, n_save)

    def observables(self):
        ek = C.c_double(); t = C.c_double(); rt = C.c_double(); vq = np.zeros(10)
        fmax = C.c_double(); frms = C.c_double(); pot = np.zeros(15); vir = np.zeros(9)
        self.L.oref_md_observables(self.h, C.byref(ek), C.byref(t), C.byref(rt), _d(vq),
                                   C.byref(fmax), C.byref(frms), _d(pot), _d(vir))
        return dict(eksum=ek.value, temp=t.value, ref_temp=rt.value, vq=vq, fmax=fmax.value,
                    frms=frms.value, pot=pot, vir=vir)

    def set_thermostat(self, vq):
        self.L.oref_set_thermostat(self.h, _d(np.ascontiguousarray(vq, dtype=np.float64)))

    def set_solve_crit(self, crit: float):
        self.L.oref_set_solve_crit(self.h, float(crit))

    def em_begin(self):
        self.L.oref_em_begin(self.h)

    def em_step(self, length: float):
        pot = np.zeros(15); fmax = C.c_double(); frms = C.c_double()
        rc = self.L.oref_em_step(self.h, float(length), _d(pot), C.byref(fmax), C.byref(frms))
        return rc, pot, fmax.value, frms.value

    def em_direction(self, cg: bool):
        self.L.oref_em_direction(self.h, int(bool(cg)))

    def timers(self):
        t = np.zeros(15)
        self.L.oref_timers(self.h, _d(t))
        return t
