"""Python mirror of the C ABI in include/shockx.h (ctypes over libshockx.so).

Mirrors the reference's entry points for the RHS / RK path: CalcRungeKutta
(src/Functions.c:2034), checkNAN (src/Functions.c:5419), CreateMetric
(src/Functions.c:6388) and the state/metric arrays of strct_U / strct_mesh
(src/SHOCK.h:547-620).  Errors raise ShxError with the library's message; there is
no CPU fallback: without the built library every call fails.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libshockx.so")

QUIRK_ZETA_VISC_J = 1
QUIRK_METRIC_DIAG3D = 2
QUIRKS_REFERENCE = QUIRK_ZETA_VISC_J | QUIRK_METRIC_DIAG3D
BC_PERIODIC, BC_MANUFACTURED, BC_INFLOW_SUPERSONIC, BC_OUTFLOW_RIEMANN = 0, 1, 2, 3
BC_OUTFLOW_NORMAL, BC_OUTFLOW_RUDY = 4, 5
BC_WALL_SLIP, BC_WALL_NOSLIP = 6, 7
BC_INFLOW_SUB_NORMAL, BC_INFLOW_SUB_RIEMANN, BC_INFLOW_SUB_ISENTROPIC = 8, 9, 10
BC_FARFIELD, BC_WALL_ISOTHERMAL = 11, 12
BC_INTERFACE = 13
FP64, MIXED = 0, 1
NUM_KERNELS = 16


class ShxError(RuntimeError):
    """A negative status of libshockx; `code` is the SHX_ERR_* value."""

    def __init__(self, msg, code=None):
        super().__init__(msg)
        self.code = code


class Desc(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int), ("space_order", ctypes.c_int), ("time_order", ctypes.c_int),
        ("n", ctypes.c_int * 3),
        ("gamma", ctypes.c_double), ("mach", ctypes.c_double), ("reynolds", ctypes.c_double),
        ("prandtl", ctypes.c_double), ("T0", ctypes.c_double),
        ("inviscid", ctypes.c_int), ("manufactured", ctypes.c_int), ("quirks", ctypes.c_uint),
        ("bc", ctypes.c_int * 6), ("precision", ctypes.c_int), ("eps_override", ctypes.c_double),
        ("rho_inflow", ctypes.c_double), ("u_inflow", ctypes.c_double), ("v_inflow", ctypes.c_double),
        ("w_inflow", ctypes.c_double), ("p_inflow", ctypes.c_double),
        ("p_out", ctypes.c_double), ("alpha_nonref", ctypes.c_double), ("T_wall", ctypes.c_double),
    ]


class Sim(ctypes.Structure):
    _fields_ = [
        ("start_iteration", ctypes.c_int), ("end_iteration", ctypes.c_int), ("tau", ctypes.c_double),
        ("time", ctypes.c_double), ("time_last_action", ctypes.c_double), ("L0", ctypes.c_double),
        ("u0", ctypes.c_double), ("accelerator", ctypes.c_int), ("accel_factor", ctypes.c_double),
        ("decel_factor", ctypes.c_double), ("number_resets", ctypes.c_int), ("distance_nan", ctypes.c_int),
        ("distance_forward", ctypes.c_int), ("distance_backward", ctypes.c_int), ("samples", ctypes.c_int),
        ("iterations_between_samples", ctypes.c_int), ("start_sampling", ctypes.c_int),
        ("probes", ctypes.c_void_p), ("nprobes", ctypes.c_int),
    ]


class SimReport(ctypes.Structure):
    _fields_ = [
        ("iteration", ctypes.c_int), ("time", ctypes.c_double), ("tau", ctypes.c_double),
        ("resets", ctypes.c_int), ("nan_steps", ctypes.c_int), ("accelerations", ctypes.c_int),
        ("tau_resets", ctypes.c_int), ("aborted", ctypes.c_int), ("samples", ctypes.c_int),
    ]


EXPORTS = [
    "shx_last_error", "shx_device_count", "shx_create", "shx_destroy", "shx_padded_extent",
    "shx_build_metric", "shx_upload_metric", "shx_download_metric", "shx_upload_state",
    "shx_download_state", "shx_upload_source", "shx_calc_runge_kutta", "shx_run", "shx_rhs",
    "shx_check_nan", "shx_cfl_dt", "shx_synchronize", "shx_set_timing", "shx_kernel_times",
    "shx_copy_probe", "shx_upload_boundary_state",
    "shx_batch_create", "shx_batch_set_tau", "shx_batch_run", "shx_batch_tune", "shx_batch_synchronize",
    "shx_batch_destroy", "shx_batch_nan_counts", "shx_tune", "shx_check_metric", "shx_simulate", "shx_post", "shx_film_open",
    "shx_film_sample", "shx_film_count", "shx_film_download", "shx_history_download", "shx_mms_error",
    "shx_run_cfl", "shx_mms_error_static",
    "shx_blocks_create", "shx_blocks_transfer_metric", "shx_blocks_run", "shx_blocks_destroy",
]
ERR_NAN = -5
BATCH_CLASSES = ("ghost", "values", "grad", "weno", "visc", "rk")

_lib = None


def lib():
    """Load libshockx.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ShxError(f"{LIB_PATH} is missing: build it with `python -m shock_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        L.shx_last_error.restype = ctypes.c_char_p
        L.shx_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(P)]
        L.shx_destroy.argtypes = [P]
        L.shx_padded_extent.argtypes = [P, P]
        L.shx_build_metric.argtypes = [P, P, P, P]
        for f in ("shx_upload_metric", "shx_download_metric", "shx_upload_state", "shx_download_state",
                  "shx_upload_source", "shx_rhs", "shx_upload_boundary_state"):
            getattr(L, f).argtypes = [P, P]
        L.shx_calc_runge_kutta.argtypes = [P, ctypes.c_double]
        L.shx_run.argtypes = [P, ctypes.c_double, ctypes.c_int]
        L.shx_run_cfl.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.shx_simulate.argtypes = [P, ctypes.POINTER(Sim), ctypes.POINTER(SimReport)]
        L.shx_post.argtypes = [P, P]
        L.shx_film_open.argtypes = [P, ctypes.c_int]
        L.shx_film_sample.argtypes = [P, ctypes.c_double]
        L.shx_film_count.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.shx_film_download.argtypes = [P, ctypes.c_int, P, ctypes.POINTER(ctypes.c_double)]
        L.shx_history_download.argtypes = [P, P, P, P, P, P]
        L.shx_mms_error.argtypes = [P, P]
        L.shx_check_nan.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.shx_check_metric.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
        L.shx_cfl_dt.argtypes = [P, ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.shx_synchronize.argtypes = [P]
        L.shx_copy_probe.argtypes = [P, ctypes.c_int]
        L.shx_set_timing.argtypes = [P, ctypes.c_int]
        L.shx_kernel_times.argtypes = [P, P, P, P]
        L.shx_device_count.argtypes = [ctypes.POINTER(ctypes.c_int)]
        L.shx_batch_create.argtypes = [P, ctypes.c_int, P, ctypes.POINTER(P)]
        L.shx_batch_set_tau.argtypes = [P, P]
        L.shx_batch_run.argtypes = [P, ctypes.c_int]
        L.shx_batch_tune.argtypes = [P, ctypes.c_int, P]
        L.shx_batch_synchronize.argtypes = [P]
        L.shx_batch_nan_counts.argtypes = [P, P]
        L.shx_batch_destroy.argtypes = [P]
        L.shx_tune.argtypes = [P, ctypes.c_int, P]
        L.shx_blocks_create.argtypes = [P, ctypes.c_int, P, ctypes.c_int, ctypes.POINTER(P)]
        L.shx_blocks_transfer_metric.argtypes = [P]
        L.shx_blocks_run.argtypes = [P, ctypes.c_double, ctypes.c_int]
        L.shx_blocks_destroy.argtypes = [P]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        msg = lib().shx_last_error()
        raise ShxError(f"shockx error {rc}: {msg.decode() if msg else ''}", rc)


def _f64(a, shape):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.shape != shape:
        raise ValueError(f"field shape {a.shape} != padded shape {shape}")
    return a


class Case:
    """One block on the GPU (a shx_case handle)."""

    def __init__(self, dims=2, space_order=5, time_order=4, n=(64, 64, 1), gamma=1.4, mach=0.5,
                 reynolds=1000.0, prandtl=0.72, T0=300.0, inviscid=False, manufactured=False,
                 quirks=QUIRKS_REFERENCE, bc=(BC_PERIODIC,) * 6, precision=FP64, eps_override=0.0,
                 inflow=(1.0, 1.0, 0.0, 0.0, 1.0), outflow=(1.0, 0.0), T_wall=1.0):
        """inflow = (rho, u, v, w, p) of the inflow state (also the Riemann outflow's far field);
        outflow = (p_out, alpha_nonref) of the subsonic outflow writers; T_wall of isothermal walls."""
        d = Desc()
        d.dims, d.space_order, d.time_order = dims, space_order, time_order
        for a in range(3):
            d.n[a] = n[a] if a < len(n) else 1
        d.gamma, d.mach, d.reynolds, d.prandtl, d.T0 = gamma, mach, reynolds, prandtl, T0
        d.inviscid, d.manufactured, d.quirks = int(inviscid), int(manufactured), quirks
        for f in range(6):
            d.bc[f] = bc[f]
        d.precision, d.eps_override = precision, eps_override
        d.rho_inflow, d.u_inflow, d.v_inflow, d.w_inflow, d.p_inflow = inflow
        d.p_out, d.alpha_nonref = outflow
        d.T_wall = T_wall
        h = ctypes.c_void_p()
        _check(lib().shx_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.desc = d
        ng = (ctypes.c_int * 3)()
        _check(lib().shx_padded_extent(h, ng))
        self.shape = (ng[0], ng[1], ng[2])
        self.g = (space_order + 1) // 2

    def close(self):
        """Free the device case (refused while a batch holds it: close the batch first)."""
        if self.h:
            _check(lib().shx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- helpers
    def _ptrs(self, arrs):
        arrs = [_f64(a, self.shape) for a in arrs]
        return arrs, (ctypes.c_void_p * len(arrs))(*[a.ctypes.data for a in arrs])

    def _outs(self, nf):
        arrs = [np.zeros(self.shape) for _ in range(nf)]
        return arrs, (ctypes.c_void_p * nf)(*[a.ctypes.data for a in arrs])

    # ---- geometry
    def build_metric(self, xe, ye, ze):
        xe = np.ascontiguousarray(xe, dtype=np.float64)
        ye = np.ascontiguousarray(ye, dtype=np.float64)
        ze = np.ascontiguousarray(ze, dtype=np.float64)
        _check(lib().shx_build_metric(self.h, xe.ctypes.data, ye.ctypes.data, ze.ctypes.data))

    def upload_metric(self, met):
        keep, p = self._ptrs(met)
        _check(lib().shx_upload_metric(self.h, p))

    def download_metric(self):
        out, p = self._outs(10)
        _check(lib().shx_download_metric(self.h, p))
        return out

    # ---- state
    def upload_state(self, prim):
        keep, p = self._ptrs(prim)
        _check(lib().shx_upload_state(self.h, p))

    def download_state(self):
        out, p = self._outs(5)
        _check(lib().shx_download_state(self.h, p))
        return out

    def upload_boundary_state(self, bs):
        """rho, u, v, w, e, p on the padded grid for manufactured-solution faces."""
        keep, p = self._ptrs(bs)
        _check(lib().shx_upload_boundary_state(self.h, p))

    def upload_source(self, src):
        if src is None:
            _check(lib().shx_upload_source(self.h, None))
            return
        keep, p = self._ptrs(src)
        _check(lib().shx_upload_source(self.h, p))

    # ---- hot path
    def calc_runge_kutta(self, tau):
        _check(lib().shx_calc_runge_kutta(self.h, tau))

    def run(self, tau, nsteps):
        """nsteps fixed-tau steps, checkNAN after each (ShxError with code ERR_NAN on a NaN)."""
        _check(lib().shx_run(self.h, tau, nsteps))

    def run_cfl(self, cfl, nsteps):
        """nsteps steps with the CFL time step of each step's starting state; returns the time
        advanced (ShxError with code ERR_NAN on a NaN)."""
        t = ctypes.c_double(0.0)
        _check(lib().shx_run_cfl(self.h, cfl, nsteps, ctypes.byref(t)))
        return t.value

    def simulate(self, start, end, tau, time=0.0, time_last_action=0.0, L0=1.0, u0=1.0, accelerator=False,
                 accel_factor=1.1, decel_factor=0.7, number_resets=10, distance_nan=5000, distance_forward=1000,
                 distance_backward=200, samples=0, iterations_between_samples=1, start_sampling=0, probes=()):
        """Iterations start+1 .. end with the reference's NaN rollback and time-step control
        (defaults: the reference's tau:* defaults, src/Import.c:53-58), sampling into the film when
        samples > 0 (film_open first).  Returns the report as a dict."""
        pr = np.ascontiguousarray(probes, dtype=np.int64)
        sim = Sim(start, end, tau, time, time_last_action, L0, u0, int(accelerator), accel_factor, decel_factor,
                  number_resets, distance_nan, distance_forward, distance_backward, samples,
                  iterations_between_samples, start_sampling, pr.ctypes.data if pr.size else None, int(pr.size))
        rep = SimReport()
        _check(lib().shx_simulate(self.h, ctypes.byref(sim), ctypes.byref(rep)))
        return {f: getattr(rep, f) for f, _ in SimReport._fields_}

    def post(self):
        """Mach, gradRho, Lambda2 (padded arrays; interior nodes filled) of the current state."""
        out, p = self._outs(3)
        _check(lib().shx_post(self.h, p))
        return out

    def history(self, nprobes, iterations):
        """Probe histories of the last simulate(): time[iterations], and p, u, v, w [nprobes, iterations]."""
        t = np.zeros(iterations)
        f = [np.zeros((nprobes, iterations)) for _ in range(4)]
        _check(lib().shx_history_download(self.h, t.ctypes.data, *[a.ctypes.data for a in f]))
        return t, f

    def mms_error(self):
        """(L2 rho, Linf rho, L2 p, Linf p) as ErrorManufacturedSolution forms them."""
        out = np.zeros(4)
        _check(lib().shx_mms_error(self.h, out.ctypes.data))
        return tuple(out)

    def film_open(self, samples):
        _check(lib().shx_film_open(self.h, samples))

    def film_sample(self, time):
        _check(lib().shx_film_sample(self.h, time))

    def film_count(self):
        n = ctypes.c_int()
        _check(lib().shx_film_count(self.h, ctypes.byref(n)))
        return n.value

    def film_download(self, s):
        """Sample s: (time, [rho, u, v, w, p, gradRho, Mach, Lambda2]) over the interior nodes."""
        d = self.desc
        n = (d.n[0], d.n[1], d.n[2] if d.dims == 3 else 1)
        out = [np.zeros(n) for _ in range(8)]
        ptrs = (ctypes.c_void_p * 8)(*[a.ctypes.data for a in out])
        t = ctypes.c_double()
        _check(lib().shx_film_download(self.h, s, ptrs, ctypes.byref(t)))
        return t.value, out

    def rhs(self):
        out, p = self._outs(5)
        _check(lib().shx_rhs(self.h, p))
        return out

    def check_nan(self):
        c = ctypes.c_int()
        _check(lib().shx_check_nan(self.h, ctypes.byref(c)))
        return c.value

    def check_metric(self):
        """Number of padded nodes with a negative Jacobian (check_Metric)."""
        c = ctypes.c_int()
        _check(lib().shx_check_metric(self.h, ctypes.byref(c)))
        return c.value

    def cfl_dt(self, cfl):
        dt = ctypes.c_double()
        _check(lib().shx_cfl_dt(self.h, cfl, ctypes.byref(dt)))
        return dt.value

    def tune(self, reps=3):
        """Pick the fastest WENO tile variant per direction on the current state."""
        ch = (ctypes.c_int * 3)()
        _check(lib().shx_tune(self.h, reps, ch))
        return list(ch)

    def copy_probe(self, reps=1):
        _check(lib().shx_copy_probe(self.h, reps))

    def synchronize(self):
        _check(lib().shx_synchronize(self.h))

    def set_timing(self, on):
        _check(lib().shx_set_timing(self.h, int(on)))

    def kernel_times(self):
        names = (ctypes.c_char_p * NUM_KERNELS)()
        ms = (ctypes.c_double * NUM_KERNELS)()
        cnt = (ctypes.c_int * NUM_KERNELS)()
        _check(lib().shx_kernel_times(self.h, names, ms, cnt))
        return {names[i].decode(): (ms[i], cnt[i]) for i in range(NUM_KERNELS) if names[i] is not None}


class Interface(ctypes.Structure):
    """One face-to-face connection of a multi-block case (shx_interface)."""
    _fields_ = [("block", ctypes.c_int), ("face", ctypes.c_int), ("donor", ctypes.c_int),
                ("donor_face", ctypes.c_int), ("transform", ctypes.c_int * 3)]


class Blocks:
    """The blocks of one case on the device, joined at their SHX_BC_INTERFACE faces (shx_blocks).
    connections: (block, face, donor, donor_face, (t0, t1, t2)) per interface face, both sides
    listed.  Close the group before its blocks."""

    def __init__(self, blocks, connections):
        self.blocks = list(blocks)
        hs = (ctypes.c_void_p * len(self.blocks))(*[c.h.value for c in self.blocks])
        arr = (Interface * max(1, len(connections)))()
        for k, (b, f, d, fd, t) in enumerate(connections):
            arr[k].block, arr[k].face, arr[k].donor, arr[k].donor_face = b, f, d, fd
            for q in range(3):
                arr[k].transform[q] = t[q]
        h = ctypes.c_void_p()
        _check(lib().shx_blocks_create(hs, len(self.blocks), arr, len(connections), ctypes.byref(h)))
        self.h = h

    def transfer_metric(self):
        _check(lib().shx_blocks_transfer_metric(self.h))

    def run(self, tau, nsteps):
        _check(lib().shx_blocks_run(self.h, tau, nsteps))

    def close(self):
        if self.h:
            lib().shx_blocks_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Batch:
    """Many cases advanced together (shx_batch): one launch per kernel class and group of
    like cases per stage, replayed from a captured HIP graph.  The cases stay usable through their
    own handles; close the batch before the cases."""

    def __init__(self, cases, taus):
        self.cases = list(cases)
        hs = (ctypes.c_void_p * len(self.cases))(*[c.h.value for c in self.cases])
        t = np.ascontiguousarray(taus, dtype=np.float64)
        h = ctypes.c_void_p()
        _check(lib().shx_batch_create(hs, len(self.cases), t.ctypes.data, ctypes.byref(h)))
        self.h = h

    def set_tau(self, taus):
        t = np.ascontiguousarray(taus, dtype=np.float64)
        _check(lib().shx_batch_set_tau(self.h, t.ctypes.data))

    def run(self, nsteps):
        _check(lib().shx_batch_run(self.h, nsteps))

    def tune(self, steps=1):
        f = (ctypes.c_int * 6)()
        _check(lib().shx_batch_tune(self.h, steps, f))
        return dict(zip(BATCH_CLASSES, list(f)))

    def synchronize(self):
        """Waits for the batch; ShxError with code ERR_NAN when the fused checkNAN of a case found
        a non-finite density since the last call (nan_counts says which)."""
        _check(lib().shx_batch_synchronize(self.h))

    def nan_counts(self):
        """Per case, the running count of non-finite densities found by the fused checkNAN."""
        out = np.zeros(len(self.cases), dtype=np.int32)
        _check(lib().shx_batch_nan_counts(self.h, out.ctypes.data))
        return out

    def close(self):
        if self.h:
            lib().shx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
