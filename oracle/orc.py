"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the shock_amd package.

Arrays are numpy float64 of shape (ng0, ng1, ng2), C-contiguous, the reference's
ghost-padded k-fastest layout (src/Functions.c:56-103).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# SHX_ORACLE_LIB names another build of the same source (the mutation tests load deliberately
# altered copies of the restatement this way, to show that the pins catch the alteration)
_LIB = os.environ.get("SHX_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")

QUIRK_ZETA_VISC_J = 1
QUIRK_METRIC_DIAG3D = 2
QUIRKS_REFERENCE = QUIRK_ZETA_VISC_J | QUIRK_METRIC_DIAG3D
BC_PERIODIC, BC_MMS, BC_INFLOW_SUPERSONIC, BC_OUTFLOW_RIEMANN = 0, 1, 2, 3
BC_OUTFLOW_NORMAL, BC_OUTFLOW_RUDY = 4, 5
BC_WALL_SLIP, BC_WALL_NOSLIP = 6, 7
BC_INFLOW_SUB_NORMAL, BC_INFLOW_SUB_RIEMANN, BC_INFLOW_SUB_ISENTROPIC = 8, 9, 10
BC_FARFIELD, BC_WALL_ISOTHERMAL = 11, 12
WALLS = (BC_WALL_SLIP, BC_WALL_NOSLIP, BC_WALL_ISOTHERMAL)


class OrcCfg(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int), ("S", ctypes.c_int), ("n", ctypes.c_int * 3),
        ("time_order", ctypes.c_int), ("inviscid", ctypes.c_int), ("mms", ctypes.c_int),
        ("quirks", ctypes.c_uint),
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("T0", ctypes.c_double), ("eps_override", ctypes.c_double),
        ("bc", ctypes.c_int * 6), ("rho_in", ctypes.c_double), ("u_in", ctypes.c_double),
        ("v_in", ctypes.c_double), ("w_in", ctypes.c_double), ("p_in", ctypes.c_double),
        ("p_out", ctypes.c_double), ("alpha_nonref", ctypes.c_double), ("tau_step", ctypes.c_double),
        ("p_last", ctypes.c_void_p), ("rho_last", ctypes.c_void_p), ("u_last", ctypes.c_void_p),
        ("c_last", ctypes.c_void_p), ("corr", ctypes.c_void_p * 3), ("T_wall", ctypes.c_double),
        ("gam", ctypes.c_void_p),
        ("g", ctypes.c_int), ("ng", ctypes.c_int * 3), ("ntot", ctypes.c_long),
        ("upsilon", ctypes.c_double), ("psi", ctypes.c_double), ("gam_heat", ctypes.c_double),
        ("suth", ctypes.c_double), ("eps", ctypes.c_double),
        ("rk_a", ctypes.c_double * 4), ("rk_b", ctypes.c_double * 4), ("rk_q", ctypes.c_double * 4),
        ("rk_f", ctypes.c_double * 4), ("rk_m", ctypes.c_double * 4),
        ("zd_mid", ctypes.c_double * 10), ("zd_int", ctypes.c_double * 10), ("zd_der", ctypes.c_double * 11),
    ]


def build():
    """Compile liboracle.so from the C restatement (gcc, no GPU)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = ctypes.CDLL(_LIB)
        P = ctypes.c_void_p
        _lib.orc_define.argtypes = [ctypes.POINTER(OrcCfg)]
        _lib.orc_phi_w5.restype = ctypes.c_double
        _lib.orc_phi_w5.argtypes = [ctypes.c_double, P, P]
        _lib.orc_phi_w9.restype = ctypes.c_double
        _lib.orc_phi_w9.argtypes = [ctypes.c_double, P, P]
        _lib.orc_eigvec.argtypes = [P] + [ctypes.c_double] * 6 + [P, P]
        _lib.orc_cfl_dt.restype = ctypes.c_double
        _lib.orc_cfl_dt.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_double, P, P]
        _lib.orc_check_nan.restype = ctypes.c_int
        _lib.orc_check_metric.restype = ctypes.c_int
        _lib.orc_visc_fac.argtypes = [ctypes.c_int, P, P]
        _lib.orc_step.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_double, P, P, P, P, P]
        _lib.orc_step_b.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_double, P, P, P, P, P, P]
        _lib.orc_set_threads.argtypes = [ctypes.c_int]
        _lib.orc_rk_update.argtypes = [ctypes.POINTER(OrcCfg), ctypes.c_int, ctypes.c_double, P, P, P, P, P]
    return _lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def _bundle(arrs):
    """C array of double* for a list of numpy arrays (kept alive by the caller)."""
    t = (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])
    return t


class Oracle:
    """One block's configuration plus helpers that allocate padded fields."""

    def __init__(self, dims=2, S=5, n=(16, 16, 1), time_order=4, gamma=1.4, mach=0.5,
                 reynolds=1000.0, prandtl=0.72, T0=300.0, inviscid=False, mms=False,
                 quirks=QUIRKS_REFERENCE, eps_override=0.0, bc=(0, 0, 0, 0, 0, 0), inflow=(1.0, 1.0, 0.0, 0.0, 1.0),
                 outflow=(1.0, 0.0), T_wall=1.0):
        """outflow: (p_out, alpha_nonref) of the subsonic outflow writers; T_wall of isothermal walls."""
        c = OrcCfg()
        for f in range(6):
            c.bc[f] = bc[f]
        c.rho_in, c.u_in, c.v_in, c.w_in, c.p_in = inflow
        c.p_out, c.alpha_nonref = outflow
        c.T_wall = T_wall
        c.dims, c.S, c.time_order = dims, S, time_order
        for d in range(3):
            c.n[d] = n[d] if d < len(n) else 1
        c.inviscid, c.mms, c.quirks = int(inviscid), int(mms), quirks
        c.gamma, c.mach, c.reynolds, c.prandtl, c.T0 = gamma, mach, reynolds, prandtl, T0
        c.eps_override = eps_override
        if lib().orc_define(ctypes.byref(c)) != 0:
            raise ValueError("invalid oracle configuration")
        self.c = c
        self.shape = (c.ng[0], c.ng[1], c.ng[2])
        self.g = c.g
        self.corr = None
        if any(b in WALLS for b in bc[:2 * dims]):
            # BC_Corrector arrays: 1 everywhere, set on wall ghosts by the wall writers
            self.corr = [np.ones(self.shape) for _ in range(3)]
            for m in range(3):
                c.corr[m] = self.corr[m].ctypes.data
        self.gam = None
        if BC_WALL_ISOTHERMAL in bc[:2 * dims]:
            # per-node heat-conduction factor (pnt_config->Gamma), set on isothermal ghosts
            self.gam = np.full(self.shape, c.gam_heat)
            c.gam = self.gam.ctypes.data

    # ---- allocation
    def zeros(self, nf):
        return [np.zeros(self.shape) for _ in range(nf)]

    # ---- kernels
    def calc_values(self, met, prim):
        der = self.zeros(8)
        lib().orc_calc_values(ctypes.byref(self.c), _bundle(met), _bundle(prim), _bundle(der))
        return der

    def apply_bc(self, met, prim, der, bstate=None):
        """Face boundary writers (in place on prim and der)."""
        lib().orc_apply_bc(ctypes.byref(self.c), _bundle(met), _bundle(prim), _bundle(der),
                           _bundle(bstate) if bstate is not None else None)

    def ghost_periodic(self, fields):
        lib().orc_ghost_periodic(ctypes.byref(self.c), _bundle(fields), len(fields))

    def weno_dir(self, d, met, prim, der, q):
        lib().orc_weno_dir(ctypes.byref(self.c), d, _bundle(met), _bundle(prim), _bundle(der), _bundle(q))

    def weno_interface(self, d, met, prim, der, i, j, k):
        out = np.zeros(5)
        lib().orc_weno_interface(ctypes.byref(self.c), d, _bundle(met), _bundle(prim), _bundle(der),
                                 i, j, k, _ptr(out))
        return out

    def point_flux(self, d, met, prim, der, i, j, k):
        out = np.zeros(5)
        lib().orc_point_flux_at(ctypes.byref(self.c), d, _bundle(met), _bundle(prim), _bundle(der),
                                i, j, k, _ptr(out))
        return out

    def visc_derivs(self, prim, der):
        grad = self.zeros(12)
        lib().orc_visc_derivs(ctypes.byref(self.c), _bundle(prim), _bundle(der), _bundle(grad))
        return grad

    def visc_dir(self, d, met, prim, der, grad, q):
        lib().orc_visc_dir(ctypes.byref(self.c), d, _bundle(met), _bundle(prim), _bundle(der),
                           _bundle(grad), _bundle(q))

    def _step_start(self, tau, met, un, p_last, c_last):
        """Point the Rudy writer at the step-start state (kept alive on self)."""
        self._last = (un[0], un[1], p_last, c_last)
        self.c.tau_step = tau
        self.c.rho_last, self.c.u_last = un[0].ctypes.data, un[1].ctypes.data
        self.c.p_last, self.c.c_last = p_last.ctypes.data, c_last.ctypes.data

    def eos_pc(self, met, prim):
        """p and c of a state on every node (CalcValues, src/Functions.c:2445-2467)."""
        der = self.calc_values(met, prim)
        return der[0], der[2]

    def rhs(self, met, prim, src=None, bstate=None):
        """Q (5 fields) of one stage for state prim; refills prim's ghosts (periodic + face BCs).
        A Rudy face reads the input state as its step-start state, with tau = 0."""
        scratch = self.zeros(30)
        p0, c0 = self.eos_pc(met, prim)
        self._step_start(0.0, met, [a.copy() for a in prim[:2]], p0, c0)
        lib().orc_rhs_b(ctypes.byref(self.c), _bundle(met), _bundle(prim), _bundle(scratch),
                        _bundle(src) if src is not None else None,
                        _bundle(bstate) if bstate is not None else None)
        return scratch[20:25], scratch[:8], scratch[8:20]

    def step(self, tau, met, un, urk, src=None, scratch=None, bstate=None):
        """One CalcRungeKutta step in place on urk (un is the state at the start of the step)."""
        if scratch is None:
            scratch = self.zeros(30)
        lib().orc_step_b(ctypes.byref(self.c), ctypes.c_double(tau), _bundle(met), _bundle(un),
                         _bundle(urk), _bundle(scratch), _bundle(src) if src is not None else None,
                         _bundle(bstate) if bstate is not None else None)
        return scratch

    def run(self, tau, met, prim0, steps, src=None, bstate=None):
        """Advance prim0 by `steps` full steps (startSimulation semantics: U_lastStep <- U_RK, which
        copies rho, p, u, v, w, e on every node, src/Functions.c:2484-2508).  The initial state's
        p and c on every node (equation of state) start U_lastStep's p and c; c is never updated."""
        un = [a.copy() for a in prim0]
        urk = [a.copy() for a in prim0]
        scratch = self.zeros(30)
        p_last, c_last = self.eos_pc(met, prim0)
        for _ in range(steps):
            self._step_start(tau, met, un, p_last, c_last)
            self.step(tau, met, un, urk, src, scratch, bstate)
            for a, b in zip(un, urk):
                a[...] = b
            p_last[...] = scratch[0]
        self.last_p = p_last          # U_lastStep's stored p after the last step
        return urk

    def simulate(self, met, prim0, sim, src=None, bstate=None):
        """startSimulation's iteration loop (src/SHOCK.c:493-768) with its NaN rollback and time-step
        control: time += tau L0/u0; one step; checkNAN on the new state; on NaN, unless NumberResets
        rollbacks were taken (then stop), tau *= the decelerator, DistanceForward restarts and the
        older of two backups (iteration, time, state with p) is reloaded; otherwise the tau resetter
        and the tau accelerator as configured; then U_lastStep <- U_RK and a backup every
        DistanceBackward iterations (slot 1) and half-way between (slot 2), with slot 1 initially the
        starting state at iteration 1 and slot 2 at iteration 999999 (src/SHOCK.c:531-540).
        sim: dict with start, end, tau, time, time_last_action, L0, u0, accelerator, accel_factor,
        decel_factor, number_resets, distance_nan, distance_forward, distance_backward, and for
        sampling (CalcValuesForPost + WriteValuesFromUToFilm of U_lastStep) samples,
        iterations_between_samples, start_sampling.  Returns (state, report); report["film"] holds
        (time, [rho, u, v, w, p, gradRho, Mach, Lambda2] over the interior) per sample."""
        un = [a.copy() for a in prim0]
        urk = [a.copy() for a in prim0]
        scratch = self.zeros(30)
        p_last, c_last = self.eos_pc(met, prim0)
        tau = tau_start = sim["tau"]
        time, last = sim["time"], sim["time_last_action"]
        L0, u0 = sim["L0"], sim["u0"]
        dfwd, dback, dnan = sim["distance_forward"], sim["distance_backward"], sim["distance_nan"]
        end = sim["end"]
        backup = [([a.copy() for a in un], p_last.copy()), None]
        it_b, time_b = [1, 999999], [time, 0.0]
        rep = {"resets": 0, "nan_steps": 0, "accelerations": 0, "tau_resets": 0, "aborted": 0}
        # sampling (src/SHOCK.c:694-717): counter and sample index, rewound with the backups
        nsamp = sim.get("samples", 0)
        between, start_s = sim.get("iterations_between_samples", 1), sim.get("start_sampling", 0)
        counter, sample = 0, 0
        counter_b, sample_b = [0, 0], [0, 0]
        film = [None] * nsamp
        # probe histories (src/Functions.c:5381-5417) under index iteration - (start + 1)
        probes = list(sim.get("probes", ()))
        iters = sim["end"] - sim["start"]
        hist_t = np.zeros(iters)
        hist = np.zeros((4, len(probes), iters))
        sl = tuple(slice(self.g, self.g + self.c.n[d]) if (d < 2 or self.c.dims == 3) else slice(None)
                   for d in range(3))
        it = sim["start"] + 1
        last_run = sim["start"]
        while it <= end:
            last_run = it
            time += tau * L0 / u0
            self._step_start(tau, met, un, p_last, c_last)
            self.step(tau, met, un, urk, src, scratch, bstate)
            p_rk = scratch[0]
            if self.check_nan(urk) > 0:
                rep["nan_steps"] += 1
                if rep["resets"] >= sim["number_resets"]:
                    rep["aborted"] = 1
                    break
                rep["resets"] += 1
                tau *= sim["decel_factor"]
                dfwd = sim["distance_forward"]
                b = 0 if it_b[0] < it_b[1] else 1
                for a, bk in zip(urk, backup[b][0]):
                    a[...] = bk
                p_rk = backup[b][1]
                last = time
                time = time_b[b]
                it = it_b[b]
                sample, counter = sample_b[b], counter_b[b]
            elif (time - last > dnan * tau * L0 / u0 and sim["accelerator"] and it + dback < end
                  and tau < tau_start):
                dfwd = sim["distance_forward"]
                tau = tau_start
                last = time
                rep["tau_resets"] += 1
            elif (time - last > dfwd * tau * L0 / u0 and sim["accelerator"] and it + dback < end
                  and tau >= tau_start):
                dfwd = int(dfwd * 4.0)
                tau *= sim["accel_factor"]
                last = time
                rep["accelerations"] += 1
            for a, b_ in zip(un, urk):
                a[...] = b_
            p_last = p_rk.copy()
            if it % dback == 0:
                backup[0] = ([a.copy() for a in un], p_last.copy())
                time_b[0], it_b[0], sample_b[0], counter_b[0] = time, it, sample, counter
            elif (it + dback // 2) % dback == 0:
                backup[1] = ([a.copy() for a in un], p_last.copy())
                time_b[1], it_b[1], sample_b[1], counter_b[1] = time, it, sample, counter
            counter += 1
            if nsamp and counter % between == 0 and sample < nsamp and it >= start_s:
                post = self.calc_post(met, un, p_last)
                fields = list(un[:4]) + [p_last] + [post[1], post[0], post[2]]
                film[sample] = (time, [f[sl].copy() for f in fields])
                sample += 1
            row = it - (sim["start"] + 1)
            if probes and 0 <= row < iters:
                # a rollback below the start (the first backup is iteration 1) gives rows before
                # the first, which the reference writes out of bounds; they are not recorded
                hist_t[row] = time
                for q, n in enumerate(probes):
                    idx = np.unravel_index(n, self.shape)
                    hist[:, q, row] = (p_last[idx], un[1][idx], un[2][idx], un[3][idx])
            it += 1
        rep.update(iteration=last_run, time=time, tau=tau, samples=sample)
        if nsamp:
            rep["film"] = film[:sample]
        if probes:
            rep["history"] = (hist_t, hist)
        return urk, rep

    def rk_update(self, s, tau, met, un, urk, q, qsum):
        lib().orc_rk_update(ctypes.byref(self.c), s, ctypes.c_double(tau), _bundle(met), _bundle(un),
                            _bundle(urk), _bundle(q), _bundle(qsum))

    def check_metric(self, met):
        """Padded nodes with a negative Jacobian (check_Metric)."""
        return lib().orc_check_metric(ctypes.byref(self.c), _bundle(met))

    def cfl_dt(self, cfl, met, prim):
        return lib().orc_cfl_dt(ctypes.byref(self.c), cfl, _bundle(met), _bundle(prim))

    def check_nan(self, prim):
        return lib().orc_check_nan(ctypes.byref(self.c), _bundle(prim))

    def calc_post(self, met, prim, p):
        """Mach, gradRho, Lambda2 on interior nodes (CalcValuesForPost); p is U_lastStep's stored p."""
        out = self.zeros(3)
        lib().orc_calc_post(ctypes.byref(self.c), _bundle(met), _bundle(prim), _ptr(p), _bundle(out))
        return out

    def mms_error(self, prim, p, ex_rho, ex_p):
        """ErrorManufacturedSolution's (L2 rho, Linf rho, L2 p, Linf p)."""
        out = np.zeros(4)
        lib().orc_mms_error(ctypes.byref(self.c), _bundle(prim), _ptr(p), _ptr(ex_rho), _ptr(ex_p), _ptr(out))
        return tuple(out)

    def mirror_metric(self, met):
        """Wall and farfield faces: ghost metric copied from the mirror nodes (mirrorGhostCellsMetric), in place."""
        lib().orc_mirror_metric(ctypes.byref(self.c), _bundle(met))

    def metric_from_coords(self, xe, ye, ze):
        met = self.zeros(10)
        lib().orc_metric_from_coords(ctypes.byref(self.c), _ptr(xe), _ptr(ye), _ptr(ze), _bundle(met))
        return met


def phi_w5(eps, f, d):
    f = np.ascontiguousarray(f, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return lib().orc_phi_w5(eps, _ptr(f), _ptr(d))


def phi_w9(eps, f, d):
    f = np.ascontiguousarray(f, dtype=np.float64)
    d = np.ascontiguousarray(d, dtype=np.float64)
    return lib().orc_phi_w9(eps, _ptr(f), _ptr(d))


def is_w5(d):
    d = np.ascontiguousarray(d, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_is_w5(_ptr(d), _ptr(out))
    return out


def is_w9(f):
    f = np.ascontiguousarray(f, dtype=np.float64)
    out = np.zeros(5)
    lib().orc_is_w9(_ptr(f), _ptr(out))
    return out


def eigvec(nh, u, v, w, rho, c, gamma):
    nh = np.ascontiguousarray(nh, dtype=np.float64)
    L = np.zeros(25)
    R = np.zeros(25)
    lib().orc_eigvec(_ptr(nh), u, v, w, rho, c, gamma, _ptr(L), _ptr(R))
    return L.reshape(5, 5), R.reshape(5, 5)


def visc_fac(d, m):
    m = np.ascontiguousarray(m, dtype=np.float64)
    out = np.zeros(30)
    lib().orc_visc_fac(d, _ptr(m), _ptr(out))
    return out
