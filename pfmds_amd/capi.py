"""ctypes binding of the PFMDS step C-ABI declared in include/pfm.h.

`Context` wraps one handle of libpfmds.so (the gfx950 HIP library).  The same
class can wrap any library that exports the same entry points under another
prefix, which is how the tests drive the CPU oracle with identical calls.

The product path has no CPU fallback: `Context.gpu()` raises if the HIP
library is missing or no device is visible.
"""
import ctypes as C
import os

import numpy as np

# constants of include/pfm.h
OK, E_OVERFLOW, E_OUT_OF_CELL, E_GRAPHENE_NN, E_NHC_PARAMS, E_UNKNOWN_KIND, E_INVALID, E_HIP, E_NO_DEVICE = range(9)
KINDS = {"lj": 1, "lj1g": 2, "ljc": 3, "morsec": 4, "rjl": 5, "tb": 6, "sur": 7, "rebosc": 8}
SUR_NTHETA = 6929  # PFM_SUR_NTHETA
INTEGRATORS = {"nvt": 1, "nve": 2, "nvms": 3}
FP64, MIXED = 0, 1
EV_DONE, EV_CONVERGED = 0, 1

HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB = os.path.join(HERE, "lib", "libpfmds.so")
# captured steps on new GPU handles (pfm_set_graphs); a profiling run may
# switch them off for every handle it makes
GRAPHS_DEFAULT = True

_P = C.POINTER
_d = _P(C.c_double)
_i = _P(C.c_int)
_SIGS = {
    "create": [_P(C.c_void_p), C.c_int, _d, _d, _d, _d],
    "set_group": [C.c_void_p, C.c_int, C.c_int, _i],
    "set_roles": [C.c_void_p] + [C.c_int] * 6,
    "add_interaction": [C.c_void_p, C.c_int, _d, C.c_int, C.c_int, _i, _i, _d, _i],
    "add_nhc": [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double],
    "run": [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, _i, _i],
    "energies": [C.c_void_p, _d, _d, _d, _d, _d, _d],
    "get_state": [C.c_void_p, _d, _d, _d],
    "set_state": [C.c_void_p, _d, _d],
    "get_list": [C.c_void_p, C.c_int, C.c_int, _i, _i, _i, _i, C.c_int],
    "get_normals": [C.c_void_p, C.c_int, _d],
    "last_error": [C.c_void_p, C.c_char_p, C.c_int],
    "add_group_change": [C.c_void_p] + [C.c_int] * 5,
    "group_count": [C.c_void_p, C.c_int, _i],
    "pe_change": [C.c_void_p, _d],
    "get_chain": [C.c_void_p, C.c_int, _d, _d],
}
_libs = {}


class Record(C.Structure):
    """struct pfm_record of include/pfm.h."""
    _fields_ = [("step", C.c_int), ("integrator", C.c_int), ("has_energy", C.c_int), ("has_frame", C.c_int),
                ("sim_time", C.c_double), ("pe", C.c_double), ("ke", C.c_double), ("temp", C.c_double),
                ("force_sum", C.c_double * 3), ("mass_center", C.c_double * 3),
                ("mass_center_vel", C.c_double * 3), ("max_vel", C.c_double), ("e_inter", _d), ("e_nhc", _d),
                ("pos", _d), ("vel", _d), ("group_n", _i)]


RECORD_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _P(Record))


class PfmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def load_library(path, prefix):
    key = (os.path.abspath(path), prefix)
    if key in _libs:
        return _libs[key]
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is not built (run __graft_entry__.build())")
    lib = C.CDLL(path)
    for name, argtypes in _SIGS.items():
        fn = getattr(lib, f"{prefix}_{name}")
        fn.argtypes = argtypes
        fn.restype = C.c_int
    destroy = getattr(lib, f"{prefix}_destroy")
    destroy.argtypes = [C.c_void_p]
    if prefix == "pfm":
        lib.pfm_create.argtypes = _SIGS["create"] + [C.c_int]
        lib.pfm_create_batch.argtypes = [_P(C.c_void_p), C.c_int, _i, _d, _d, _d, _d, C.c_int]
        lib.pfm_create_batch.restype = C.c_int
        lib.pfm_replica_energies.argtypes = [C.c_void_p, _d, _d, _d, _d]
        lib.pfm_replica_energies.restype = C.c_int
        lib.pfm_set_timing.argtypes = [C.c_void_p, C.c_int]
        lib.pfm_set_timing.restype = C.c_int
        lib.pfm_kernel_time.argtypes = [C.c_void_p, C.c_char_p, _i, _d]
        lib.pfm_kernel_time.restype = C.c_int
        lib.pfm_set_tuning.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.pfm_set_tuning.restype = C.c_int
        lib.pfm_set_graphs.argtypes = [C.c_void_p, C.c_int]
        lib.pfm_set_graphs.restype = C.c_int
        lib.pfm_set_fusion.argtypes = [C.c_void_p, C.c_int]
        lib.pfm_set_fusion.restype = C.c_int
        lib.pfm_set_pass_lanes.argtypes = [C.c_void_p, C.c_int, C.c_int]
        lib.pfm_set_pass_lanes.restype = C.c_int
        lib.pfm_row_energies.argtypes = [C.c_void_p, C.c_int, _i, _d]
        lib.pfm_row_energies.restype = C.c_int
        lib.pfm_graph_count.argtypes = [C.c_void_p, _P(C.c_long), _P(C.c_long)]
        lib.pfm_graph_count.restype = C.c_int
        lib.pfm_sur_set_params.argtypes = [C.c_void_p, C.c_int, _d, C.c_int]
        lib.pfm_sur_set_params.restype = C.c_int
        lib.pfm_sur_descriptors.argtypes = [C.c_void_p, C.c_int, _d]
        lib.pfm_sur_descriptors.restype = C.c_int
        lib.pfm_sur_loss_grad.argtypes = [C.c_void_p, C.c_int, _d, C.c_double, C.c_double, C.c_double, _d, _d]
        lib.pfm_sur_loss_grad.restype = C.c_int
        lib.pfm_copy_state.argtypes = [C.c_void_p, C.c_void_p]
        lib.pfm_copy_state.restype = C.c_int
        lib.pfm_force_rmse.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _d]
        lib.pfm_force_rmse.restype = C.c_int
        lib.pfm_force_rmse_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_int, _d]
        lib.pfm_force_rmse_noalloc.restype = C.c_int
        lib.pfm_set_chain.argtypes = [C.c_void_p, C.c_int, _d, _d]
        lib.pfm_set_chain.restype = C.c_int
        lib.pfm_set_speculation.argtypes = [C.c_void_p, C.c_int]
        lib.pfm_set_speculation.restype = C.c_int
        lib.pfm_set_records.argtypes = [C.c_void_p, C.c_int, _i, C.c_int, _i, C.c_int, RECORD_FN, C.c_void_p]
        lib.pfm_set_records.restype = C.c_int
    _libs[key] = lib
    return lib


def _dp(a):
    return a.ctypes.data_as(_d)


def _ip(a):
    return a.ctypes.data_as(_i)


class Context:
    """One system behind the C-ABI (one pfm_ctx handle)."""

    def __init__(self, lib, prefix, box, pos, vel, mass, precision=FP64, n_atoms=None):
        """One system, or with `n_atoms` (atoms per replica) a batch whose
        `box` holds one row of three box lengths per replica."""
        self.lib, self.pfx = lib, prefix
        self.h = C.c_void_p()
        self.n = len(mass)
        self._pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        self._vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        box = np.ascontiguousarray(box, dtype=np.float64)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        self.n_inter = 0
        self.n_nhc = 0
        self.nhc_M = []
        self.n_rep = 1
        self._rec_fn = None
        self._rec_exc = None
        if n_atoms is not None:
            if prefix != "pfm":
                raise ValueError("batches need the HIP library")
            na = np.ascontiguousarray(n_atoms, dtype=np.int32)
            self.n_rep = len(na)
            rc = self.lib.pfm_create_batch(C.byref(self.h), self.n_rep, _ip(na), _dp(box.reshape(-1)),
                                           _dp(self._pos), _dp(self._vel), _dp(mass), int(precision))
            self._check(rc)
            return
        args = [C.byref(self.h), self.n, _dp(box), _dp(self._pos), _dp(self._vel), _dp(mass)]
        if prefix == "pfm":
            args.append(int(precision))
        rc = self._fn("create")(*args)
        self._check(rc)

    @classmethod
    def gpu(cls, box, pos, vel, mass, precision=FP64, path=GPU_LIB):
        c = cls(load_library(path, "pfm"), "pfm", box, pos, vel, mass, precision)
        if not GRAPHS_DEFAULT:
            c.set_graphs(False)
        return c

    @classmethod
    def gpu_batch(cls, n_atoms, boxes, pos, vel, mass, precision=FP64, path=GPU_LIB):
        c = cls(load_library(path, "pfm"), "pfm", boxes, pos, vel, mass, precision, n_atoms=n_atoms)
        if not GRAPHS_DEFAULT:
            c.set_graphs(False)
        return c

    def _fn(self, name):
        return getattr(self.lib, f"{self.pfx}_{name}")

    def last_error(self):
        buf = C.create_string_buffer(1024)
        self._fn("last_error")(self.h, buf, 1024)
        return buf.value.decode(errors="replace")

    def _check(self, rc):
        if rc != OK:
            raise PfmError(rc, self.last_error())

    def close(self):
        if self.h:
            self._fn("destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- setup -------------------------------------------------------------
    def set_group(self, gid, idx1):
        idx = np.ascontiguousarray(idx1, dtype=np.int32)
        self._check(self._fn("set_group")(self.h, gid, len(idx), _ip(idx)))

    def set_roles(self, all_moving, xyz_moving, z_moving, all_atoms, zero_momentum_period, invert_z=False):
        self._check(self._fn("set_roles")(self.h, all_moving, xyz_moving, z_moving, all_atoms,
                                          zero_momentum_period, int(bool(invert_z))))

    def add_interaction(self, name, params, lists):
        kind = KINDS.get(name, 0) if isinstance(name, str) else int(name)
        p = np.ascontiguousarray(params, dtype=np.float64)
        g = np.ascontiguousarray([x for row in lists for x in row[:2]], dtype=np.int32)
        nmax = np.ascontiguousarray([row[2] for row in lists], dtype=np.int32)
        rcut = np.ascontiguousarray([row[3] for row in lists], dtype=np.float64)
        per = np.ascontiguousarray([row[4] for row in lists], dtype=np.int32)
        self._check(self._fn("add_interaction")(self.h, kind, _dp(p), len(p), len(lists), _ip(g), _ip(nmax),
                                                _dp(rcut), _ip(per)))
        self.n_inter += 1

    def add_group_change(self, g_from, g_to, ts1, ts2, frec):
        """One runtime group change (change_particle_group_N, md_general.f90:82-94)."""
        self._check(self._fn("add_group_change")(self.h, int(g_from), int(g_to), int(ts1), int(ts2), int(frec)))

    def group_count(self, gid):
        """Active atoms of group gid after the last step."""
        n = C.c_int(0)
        self._check(self._fn("group_count")(self.h, int(gid), C.byref(n)))
        return n.value

    def pe_change(self):
        """PE - PE_prev of the last step entered (printed by md() after nvms)."""
        d = C.c_double(0.0)
        self._check(self._fn("pe_change")(self.h, C.byref(d)))
        return d.value

    def add_nhc(self, group, temperature, M, q1):
        self._check(self._fn("add_nhc")(self.h, group, temperature, M, q1))
        self.n_nhc += 1
        self.nhc_M.append(int(M))

    # -- stepping ----------------------------------------------------------
    def run(self, integrator, dt, step_begin, step_end, energy_at_end=True, ms_de=0.0):
        integ = INTEGRATORS[integrator] if isinstance(integrator, str) else int(integrator)
        last = C.c_int(-1)
        ev = C.c_int(0)
        rc = self._fn("run")(self.h, integ, float(dt), int(step_begin), int(step_end), int(bool(energy_at_end)),
                             float(ms_de), C.byref(last), C.byref(ev))
        if self._rec_exc is not None:
            exc, self._rec_exc = self._rec_exc, None
            raise exc
        self._check(rc)
        return last.value, ev.value

    def energies(self):
        e = np.zeros(max(self.n_inter, 1))
        enhc = np.zeros(max(self.n_nhc, 1))
        pe, ke, t, st = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        self._check(self._fn("energies")(self.h, _dp(e), C.byref(pe), C.byref(ke), C.byref(t), _dp(enhc),
                                         C.byref(st)))
        return dict(e_inter=e[: self.n_inter], pe=pe.value, ke=ke.value, temp=t.value,
                    e_nhc=enhc[: self.n_nhc], time=st.value)

    def state(self):
        pos = np.zeros((self.n, 3))
        vel = np.zeros((self.n, 3))
        frc = np.zeros((self.n, 3))
        self._check(self._fn("get_state")(self.h, _dp(pos), _dp(vel), _dp(frc)))
        return pos, vel, frc

    def set_state(self, pos, vel):
        p = np.ascontiguousarray(pos, dtype=np.float64).reshape(self.n, 3)
        v = np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        self._check(self._fn("set_state")(self.h, _dp(p), _dp(v)))

    def get_list(self, inter, k, stride=None):
        nrows = C.c_int(0)
        self._check(self._fn("get_list")(self.h, inter, k, C.byref(nrows), None, None, None, 0))
        n = nrows.value
        stride = stride or 1024
        nnum = np.zeros(max(n, 1), dtype=np.int32)
        less = np.zeros(max(n, 1), dtype=np.int32)
        nl = np.zeros((max(n, 1), stride), dtype=np.int32)
        self._check(self._fn("get_list")(self.h, inter, k, C.byref(nrows), _ip(nnum), _ip(nl), _ip(less), stride))
        return nnum[:n], nl[:n], less[:n]

    def normals(self, inter, n_rows):
        out = np.zeros((max(n_rows, 1), 3))
        self._check(self._fn("get_normals")(self.h, inter, _dp(out)))
        return out[:n_rows]

    # -- GPU-only extras ----------------------------------------------------
    def replica_energies(self):
        """Per-replica energies of the last evaluation (arrays over replicas)."""
        r, k = self.n_rep, max(self.n_inter, 1)
        e = np.zeros((r, k))
        pe, ke, t = np.zeros(r), np.zeros(r), np.zeros(r)
        self._check(self.lib.pfm_replica_energies(self.h, _dp(e), _dp(pe), _dp(ke), _dp(t)))
        return dict(e_inter=e[:, : self.n_inter], pe=pe, ke=ke, temp=t)

    def set_timing(self, on):
        self._check(self.lib.pfm_set_timing(self.h, int(bool(on))))

    def kernel_time(self, name):
        cnt = C.c_int(0)
        ms = C.c_double(0.0)
        self._check(self.lib.pfm_kernel_time(self.h, name.encode(), C.byref(cnt), C.byref(ms)))
        return cnt.value, ms.value

    def set_tuning(self, lanes_per_row=0, block=0):
        self._check(self.lib.pfm_set_tuning(self.h, int(lanes_per_row), int(block)))

    def row_energies(self, inter):
        """Per-row energies of interaction `inter` (1-based) at the last energy
        evaluation (pfm_row_energies)."""
        n = C.c_int()
        self._check(self.lib.pfm_row_energies(self.h, int(inter), C.byref(n), None))
        out = np.zeros(max(n.value, 1))
        self._check(self.lib.pfm_row_energies(self.h, int(inter), C.byref(n), _dp(out)))
        return out[:n.value]

    def set_fusion(self, on):
        """Co-scheduled passes over disjoint atoms (pfm_set_fusion)."""
        self._check(self.lib.pfm_set_fusion(self.h, int(bool(on))))

    def set_pass_lanes(self, inter, lanes):
        """Lanes per row of interaction `inter`'s separately launched passes
        (1-based; 0: the pfm_set_tuning count) (pfm_set_pass_lanes)."""
        self._check(self.lib.pfm_set_pass_lanes(self.h, int(inter), int(lanes)))

    def set_graphs(self, on):
        self._check(self.lib.pfm_set_graphs(self.h, int(bool(on))))

    def set_records(self, energy_periods=(), frame_periods=(), frame_slots=4, callback=None):
        """Output records (pfm_set_records): `callback(rec)` receives a dict
        per recorded step, in step order, during run(): step, integrator,
        sim_time; with an energy record e_inter, pe, ke, temp, e_nhc,
        force_sum, mass_center, mass_center_vel, max_vel; with a frame pos
        and vel ([n, 3], original order).  Arrays are copies.  An exception
        in the callback stops the run and is raised from run()."""
        ep = np.ascontiguousarray(list(energy_periods) or [0], dtype=np.int32)
        xp = np.ascontiguousarray(list(frame_periods) or [0], dtype=np.int32)
        if callback is None:
            self._check(self.lib.pfm_set_records(self.h, 0, _ip(ep), 0, _ip(xp), 0, RECORD_FN(), None))
            self._rec_fn = None
            return
        n, ni, nq = self.n, self.n_inter, self.n_nhc

        def tramp(_user, rp):
            try:
                r = rp.contents
                d = dict(step=r.step, integrator=r.integrator, sim_time=r.sim_time,
                         group_n=np.ctypeslib.as_array(r.group_n, (33,)).copy())
                if r.has_energy:
                    d.update(pe=r.pe, ke=r.ke, temp=r.temp, force_sum=np.array(r.force_sum[:]),
                             mass_center=np.array(r.mass_center[:]), mass_center_vel=np.array(r.mass_center_vel[:]),
                             max_vel=r.max_vel,
                             e_inter=np.ctypeslib.as_array(r.e_inter, (max(ni, 1),))[:ni].copy(),
                             e_nhc=np.ctypeslib.as_array(r.e_nhc, (nq,)).copy() if nq else np.zeros(0))
                if r.has_frame:
                    d["pos"] = np.ctypeslib.as_array(r.pos, (n, 3)).copy()
                    d["vel"] = np.ctypeslib.as_array(r.vel, (n, 3)).copy()
                callback(d)
                return 0
            except BaseException as e:  # noqa: BLE001 -- re-raised from run()
                self._rec_exc = e
                return 1

        fn = RECORD_FN(tramp)
        self._check(self.lib.pfm_set_records(self.h, len(energy_periods), _ip(ep), len(frame_periods), _ip(xp),
                                             int(frame_slots), fn, None))
        self._rec_fn = fn

    def copy_state_from(self, src):
        """Positions and velocities of `src` (same atoms, same GPU), device to device."""
        self._check(self.lib.pfm_copy_state(self.h, src.h))

    def force_rmse(self, other, group):
        """RMSE (eV/A) between `other`'s forces and this handle's over the 3n
        force components of this handle's group `group`."""
        r = C.c_double(0.0)
        self._check(self.lib.pfm_force_rmse(other.h, self.h, int(group), C.byref(r)))
        return r.value

    def force_rmse_noalloc(self, other, group):
        """force_rmse through pfm_force_rmse_noalloc (no device allocation;
        at most 1,048,576 atoms)."""
        r = C.c_double()
        self._check(self.lib.pfm_force_rmse_noalloc(other.h, self.h, int(group), C.byref(r)))
        return r.value

    def get_chain(self, c, M):
        """(x[M], v[M]) of Nose-Hoover chain c (1-based)."""
        x, v = np.zeros(max(M, 1)), np.zeros(max(M, 1))
        self._check(self._fn("get_chain")(self.h, int(c), _dp(x), _dp(v)))
        return x[:M], v[:M]

    def set_chain(self, c, x, v):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        self._check(self.lib.pfm_set_chain(self.h, int(c), _dp(x), _dp(v)))

    def set_speculation(self, on):
        """nvms blocks with one energy read-back per block (pfm_set_speculation)."""
        self._check(self.lib.pfm_set_speculation(self.h, int(bool(on))))

    def graph_count(self):
        """(blocks captured, blocks launched from graphs)"""
        a, b = C.c_long(0), C.c_long(0)
        self._check(self.lib.pfm_graph_count(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    # -- learned surrogate (interaction kind 'sur') ---------------------------
    def sur_set_params(self, inter, theta):
        t = np.ascontiguousarray(theta, dtype=np.float64)
        if t.size != SUR_NTHETA:
            raise ValueError(f"surrogate parameters: {t.size} values, expected {SUR_NTHETA}")
        self._check(self.lib.pfm_sur_set_params(self.h, int(inter), _dp(t), int(t.size)))

    def sur_descriptors(self, inter, n_rows):
        """Descriptors of the surrogate's group-1 atoms, [n_rows, 40] in group order."""
        g = np.zeros((max(n_rows, 1), 40))
        self._check(self.lib.pfm_sur_descriptors(self.h, int(inter), _dp(g)))
        return g[:n_rows]

    def sur_loss_grad(self, inter, f_ref, e_ref, w_e, w_f):
        """Loss {total, energy term, force term} of one reference frame at
        the current positions and its gradient over the parameters."""
        f = np.ascontiguousarray(f_ref, dtype=np.float64).reshape(self.n, 3)
        loss = np.zeros(3)
        grad = np.zeros(SUR_NTHETA)
        self._check(self.lib.pfm_sur_loss_grad(self.h, int(inter), _dp(f), float(e_ref), float(w_e), float(w_f),
                                               _dp(loss), _dp(grad)))
        return loss, grad
