"""ctypes binding of the C-ABI declared in include/topopt_gpu.h.

This is plumbing for the test harness and bench.py: it loads the in-tree
libtopopt_gpu.so and fails loudly if it is missing.  Arrays are exchanged
either as host numpy arrays (on_device = 0) or as torch device tensors
(on_device = 1, raw data pointers).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TT_LIB: another build of the same library (the sanitizer test run loads the host-sanitized build)
LIB_PATH = os.environ.get("TT_LIB") or os.path.join(_HERE, "libtopopt_gpu.so")

TT_PREC_FP64 = 0
TT_PREC_MIXED = 1
TT_OP_K, TT_OP_M, TT_OP_A = 0, 1, 2
TT_ERR_NOT_CONVERGED = -4
TT_ERR_NONFINITE = -5

EXPORTED = [
    "tt_default_params", "tt_create", "tt_destroy", "tt_last_error", "tt_set_density", "tt_op_apply",
    "tt_complete_initial_state", "tt_forward", "tt_get_state", "tt_objective_gradient", "tt_volume_constraint",
    "tt_pde_filter", "tt_filter_project", "tt_filter_gradients", "tt_mnd", "tt_mma_create", "tt_mma_destroy",
    "tt_mma_set_asymptotes", "tt_mma_set_outer_movelimit", "tt_mma_update", "tt_mma_design_change",
    "tt_mma_kkt_residual", "tt_mma_get_asymptotes", "tt_get_stats", "tt_reset_stats", "tt_synchronize",
    "tt_num_nodes", "tt_num_elements", "tt_time_op_apply", "tt_set_material", "tt_set_filter_radius",
    "tt_mma_get_state", "tt_mma_set_state", "tt_get_restart_vectors", "tt_set_restart_vectors",
    "tt_petsc_write_vecs", "tt_petsc_read_vecs", "tt_get_reduced_model", "tt_set_initial_displacement",
    "tt_volume_constraint_noalloc", "tt_time_fcg_apply", "tt_set_checkpoint", "tt_time_vcycle",
]


class Params(ctypes.Structure):
    _fields_ = [
        ("nelx", ctypes.c_int), ("nely", ctypes.c_int), ("nelz", ctypes.c_int),
        ("xc", ctypes.c_double * 6), ("nu", ctypes.c_double),
        ("sInt", ctypes.c_int), ("mInt", ctypes.c_int),
        ("Emin", ctypes.c_double), ("Emax", ctypes.c_double), ("Rmin", ctypes.c_double), ("Rmax", ctypes.c_double),
        ("penalK", ctypes.c_double), ("penalM", ctypes.c_double),
        ("dalpha", ctypes.c_double), ("dbeta", ctypes.c_double),
        ("ti", ctypes.c_int), ("nsteps", ctypes.c_int), ("t1", ctypes.c_double),
        ("objective", ctypes.c_int), ("boundaries", ctypes.c_int), ("loadloc", ctypes.c_int),
        ("loadsignal", ctypes.c_int), ("pulse_steps", ctypes.c_int), ("region", ctypes.c_int),
        ("nlvls", ctypes.c_int), ("precision", ctypes.c_int), ("rtol", ctypes.c_double), ("maxit", ctypes.c_int),
        ("cheb_degree", ctypes.c_int), ("checkpoint_slots", ctypes.c_int),
        ("rmin", ctypes.c_double), ("filter_rtol", ctypes.c_double), ("verbose", ctypes.c_int),
        ("graphs", ctypes.c_int),
        ("reduction", ctypes.c_int), ("basis", ctypes.c_int), ("omega0", ctypes.c_double),
        ("coarse_direct_max", ctypes.c_int), ("mg_coarse_pre", ctypes.c_int), ("mg_coarse_post", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("solves", ctypes.c_long), ("krylov_iters", ctypes.c_long), ("last_iters", ctypes.c_int),
        ("last_relres", ctypes.c_double), ("max_relres", ctypes.c_double), ("forward_steps", ctypes.c_long),
        ("adjoint_steps", ctypes.c_long), ("lambda_max", ctypes.c_double * 8), ("nlvls", ctypes.c_int),
    ]


class TTError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: run __graft_entry__.build() (the HIP path has no fallback)")
        L = ctypes.CDLL(LIB_PATH)
        vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
        sig = {
            "tt_default_params": [ctypes.POINTER(Params)],
            "tt_create": [ctypes.POINTER(Params), ctypes.POINTER(vp)],
            "tt_set_density": [vp, vp, ip],
            "tt_set_material": [vp] + [ctypes.c_double] * 8,
            "tt_set_filter_radius": [vp, ctypes.c_double],
            "tt_op_apply": [vp, ip, ip, vp, vp, ip],
            "tt_complete_initial_state": [vp], "tt_forward": [vp],
            "tt_get_state": [vp, ip, vp, vp, vp, ip],
            "tt_objective_gradient": [vp, dp, vp, ip],
            "tt_volume_constraint": [vp, vp, ctypes.c_double, dp, vp, ip],
            "tt_pde_filter": [vp, vp, vp, ip],
            "tt_filter_project": [vp, vp, vp, vp, ip, ctypes.c_double, ctypes.c_double, ip],
            "tt_filter_gradients": [vp, vp, vp, vp, ip, ctypes.c_double, ctypes.c_double, ip],
            "tt_mnd": [vp, vp, dp, ip],
            "tt_mma_create": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.POINTER(vp)],
            "tt_mma_set_asymptotes": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double],
            "tt_mma_set_outer_movelimit": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, vp, vp, vp, ip],
            "tt_mma_update": [vp, vp, vp, vp, vp, vp, vp, ip],
            "tt_mma_design_change": [vp, vp, vp, dp, ip],
            "tt_mma_kkt_residual": [vp, vp, vp, vp, vp, vp, vp, dp, dp, ip],
            "tt_mma_get_asymptotes": [vp, vp, vp, dp],
            "tt_time_op_apply": [vp, ip, ip, vp, vp, ip, dp],
            "tt_get_stats": [vp, ctypes.POINTER(Stats)], "tt_reset_stats": [vp], "tt_synchronize": [vp],
            "tt_mma_get_state": [vp, ctypes.POINTER(ctypes.c_int), vp, vp, vp, vp],
            "tt_mma_set_state": [vp, ctypes.c_int, vp, vp, vp, vp],
            "tt_get_restart_vectors": [vp, vp, vp, vp], "tt_set_restart_vectors": [vp, vp, vp, vp],
            "tt_petsc_write_vecs": [ctypes.c_char_p, ctypes.c_int, vp, vp],
            "tt_petsc_read_vecs": [ctypes.c_char_p, ctypes.c_int, vp, vp],
            "tt_get_reduced_model": [vp, vp, ip, vp, vp, vp, vp, vp, vp],
            "tt_set_initial_displacement": [vp, vp, ip],
            "tt_volume_constraint_noalloc": [vp, ctypes.c_long, ctypes.c_double, vp, vp, vp],
            "tt_time_fcg_apply": [vp, ip, dp],
            "tt_time_vcycle": [vp, ip, dp, dp],
            "tt_set_checkpoint": [vp, ip, vp, vp, vp, ip],
        }
        for name, args in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = ctypes.c_int
        L.tt_last_error.argtypes = [vp]
        L.tt_last_error.restype = ctypes.c_char_p
        L.tt_destroy.argtypes = [vp]
        L.tt_destroy.restype = None
        L.tt_mma_destroy.argtypes = [vp]
        L.tt_mma_destroy.restype = None
        L.tt_num_nodes.argtypes = [vp]
        L.tt_num_nodes.restype = ctypes.c_long
        L.tt_num_elements.argtypes = [vp]
        L.tt_num_elements.restype = ctypes.c_long
        _lib = L
    return _lib


def default_params(**kw):
    p = Params()
    lib().tt_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if k == "xc":
            for i, x in enumerate(v):
                p.xc[i] = x
        else:
            setattr(p, k, v)
    return p


def _ptr(a):
    """Raw pointer and on_device flag of a numpy array or torch tensor."""
    if a is None:
        return None, 0
    if isinstance(a, np.ndarray):
        return a.ctypes.data, 0
    return a.data_ptr(), int(a.is_cuda)


def volume_constraint_noalloc(xphys, volfrac, gx, dgdx=None, stream=0):
    """tt_volume_constraint_noalloc on device tensors (xphys, gx of one element, dgdx or None), queued
    on `stream` (a raw hipStream_t, 0 = the null stream); returns without synchronising."""
    rc = lib().tt_volume_constraint_noalloc(xphys.data_ptr(), xphys.numel(), volfrac, gx.data_ptr(),
                                            None if dgdx is None else dgdx.data_ptr(), stream or None)
    if rc != 0:
        raise TTError(rc, lib().tt_last_error(None).decode())


class Context:
    """One problem on the current GPU (mirrors the C-ABI one to one)."""

    def __init__(self, params):
        self.params = params
        h = ctypes.c_void_p()
        rc = lib().tt_create(ctypes.byref(params), ctypes.byref(h))
        if rc != 0:
            raise TTError(rc, lib().tt_last_error(None).decode())
        self.h = h
        self.nn = lib().tt_num_nodes(h)
        self.nel = lib().tt_num_elements(h)
        self.nd = 3 * self.nn

    def close(self):
        if getattr(self, "h", None):
            lib().tt_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise TTError(rc, lib().tt_last_error(self.h).decode())

    def set_material(self, Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta):
        self._chk(lib().tt_set_material(self.h, Emin, Emax, Rmin, Rmax, penalK, penalM, dalpha, dbeta))

    def set_filter_radius(self, rmin):
        self._chk(lib().tt_set_filter_radius(self.h, rmin))

    def set_density(self, x):
        p, dev = _ptr(x)
        self._chk(lib().tt_set_density(self.h, p, dev))

    def op_apply(self, which, u, y, prec=0):
        pu, dev = _ptr(u)
        py, _ = _ptr(y)
        self._chk(lib().tt_op_apply(self.h, which, prec, pu, py, dev))

    def time_op_apply(self, which, u, y, prec, nrep):
        ms = ctypes.c_double()
        pu, _ = _ptr(u)
        py, _ = _ptr(y)
        self._chk(lib().tt_time_op_apply(self.h, which, prec, pu, py, nrep, ctypes.byref(ms)))
        return ms.value

    def time_fcg_apply(self, nrep):
        ms = ctypes.c_double()
        self._chk(lib().tt_time_fcg_apply(self.h, nrep, ctypes.byref(ms)))
        return ms.value

    def time_vcycle(self, nrep):
        """(GPU ms, host enqueue ms) per multigrid preconditioner application."""
        g, h = ctypes.c_double(), ctypes.c_double()
        self._chk(lib().tt_time_vcycle(self.h, nrep, ctypes.byref(g), ctypes.byref(h)))
        return g.value, h.value

    def complete_initial_state(self):
        self._chk(lib().tt_complete_initial_state(self.h))

    def forward(self):
        self._chk(lib().tt_forward(self.h))

    def get_state(self, t, out=None):
        """(u, v, a) of step t as host arrays, or into `out`: three FP64 device tensors."""
        if out is not None:
            self._chk(lib().tt_get_state(self.h, t, *(o.data_ptr() for o in out), 1))
            return out
        u, v, a = (np.zeros(self.nd) for _ in range(3))
        self._chk(lib().tt_get_state(self.h, t, u.ctypes.data, v.ctypes.data, a.ctypes.data, 0))
        return u, v, a

    def set_checkpoint(self, t, u, v, a):
        """State of step t given in place of the forward sweep (host arrays of nd values)."""
        arrs = [np.ascontiguousarray(q, dtype=np.float64) for q in (u, v, a)]
        self._chk(lib().tt_set_checkpoint(self.h, t, *(q.ctypes.data for q in arrs), 0))

    def objective_gradient(self, dfdx=None):
        out = np.zeros(self.nel) if dfdx is None else dfdx
        fx = ctypes.c_double()
        p, dev = _ptr(out)
        self._chk(lib().tt_objective_gradient(self.h, ctypes.byref(fx), p, dev))
        return fx.value, out

    def volume_constraint(self, xphys, volfrac):
        dg = np.zeros(self.nel)
        gx = ctypes.c_double()
        p, dev = _ptr(np.ascontiguousarray(xphys, dtype=np.float64))
        self._chk(lib().tt_volume_constraint(self.h, p, volfrac, ctypes.byref(gx), dg.ctypes.data, 0))
        return gx.value, dg

    def pde_filter(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros_like(x)
        self._chk(lib().tt_pde_filter(self.h, x.ctypes.data, out.ctypes.data, 0))
        return out

    def filter_project(self, x, projection=False, beta=0.1, eta=0.15):
        x = np.ascontiguousarray(x, dtype=np.float64)
        xt = np.zeros_like(x)
        xp = np.zeros_like(x)
        self._chk(lib().tt_filter_project(self.h, x.ctypes.data, xt.ctypes.data, xp.ctypes.data, int(projection),
                                          beta, eta, 0))
        return xt, xp

    def filter_gradients(self, xtilde, dfdx, dgdx, projection=False, beta=0.1, eta=0.15):
        xt = np.ascontiguousarray(xtilde, dtype=np.float64)
        df = np.array(dfdx, dtype=np.float64)
        dg = np.array(dgdx, dtype=np.float64)
        self._chk(lib().tt_filter_gradients(self.h, xt.ctypes.data, df.ctypes.data, dg.ctypes.data, int(projection),
                                            beta, eta, 0))
        return df, dg

    def mnd(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        m = ctypes.c_double()
        self._chk(lib().tt_mnd(self.h, x.ctypes.data, ctypes.byref(m), 0))
        return m.value

    def stats(self):
        s = Stats()
        self._chk(lib().tt_get_stats(self.h, ctypes.byref(s)))
        return s

    def reset_stats(self):
        self._chk(lib().tt_reset_stats(self.h))

    def synchronize(self):
        self._chk(lib().tt_synchronize(self.h))

    def restart_vectors(self):
        """(a0, lambda_T, nu0): the physics restart vectors (LinearElasticity.cc:2474-2486)."""
        out = [np.zeros(self.nd) for _ in range(3)]
        self._chk(lib().tt_get_restart_vectors(self.h, *[o.ctypes.data for o in out]))
        return tuple(out)

    def set_initial_displacement(self, u0=None):
        """u0 (nd values, masked to the free dofs) for the next complete_initial_state; None resets to 0."""
        if u0 is None:
            self._chk(lib().tt_set_initial_displacement(self.h, None, 0))
            return
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        if u0.size != self.nd:
            raise ValueError("u0 has nd values")
        self._chk(lib().tt_set_initial_displacement(self.h, u0.ctypes.data, 0))

    def reduced_model(self, with_states=False):
        """The SOMMG reduced model (reduction = 1): dict of Q (basis x nd), Ut, Kr, Mr, Cr (basis x basis),
        Pr (basis) and, with_states, the reduced states (nsteps + 1, 3, basis) of the last forward sweep."""
        nb = self.params.basis
        Q = np.zeros((nb, self.nd))
        Ut, Kr, Mr, Cr = (np.zeros((nb, nb)) for _ in range(4))
        Pr = np.zeros(nb)
        rs = np.zeros((self.params.nsteps + 1, 3, nb)) if with_states else None
        self._chk(lib().tt_get_reduced_model(self.h, Q.ctypes.data, 0, Ut.ctypes.data, Kr.ctypes.data,
                                             Mr.ctypes.data, Cr.ctypes.data, Pr.ctypes.data,
                                             None if rs is None else rs.ctypes.data))
        out = dict(Q=Q, Ut=Ut, Kr=Kr, Mr=Mr, Cr=Cr, Pr=Pr)
        if with_states:
            out["rstates"] = rs
        return out

    def set_restart_vectors(self, a0=None, lamT=None, nu0=None):
        arrs = [None if v is None else np.ascontiguousarray(v, dtype=np.float64) for v in (a0, lamT, nu0)]
        for v in arrs:
            if v is not None and v.size != self.nd:
                raise ValueError("restart vectors have nd values")
        self._chk(lib().tt_set_restart_vectors(self.h, *[None if v is None else v.ctypes.data for v in arrs]))


def petsc_write_vecs(path, vecs):
    """Write vectors as a PETSc binary Vec file (VecView with a binary viewer, 32-bit PetscInt)."""
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in vecs]
    ptrs = (ctypes.c_void_p * max(1, len(arrs)))(*[a.ctypes.data for a in arrs])
    lens = (ctypes.c_long * max(1, len(arrs)))(*[a.size for a in arrs])
    rc = lib().tt_petsc_write_vecs(os.fsencode(path), len(arrs), ptrs, lens)
    if rc != 0:
        raise TTError(rc, lib().tt_last_error(None).decode())


def petsc_read_vecs(path, lens):
    """Read the first len(lens) vectors of a PETSc binary Vec file (VecLoad), checking their lengths."""
    out = [np.zeros(int(n)) for n in lens]
    ptrs = (ctypes.c_void_p * max(1, len(out)))(*[a.ctypes.data for a in out])
    ls = (ctypes.c_long * max(1, len(out)))(*[a.size for a in out])
    rc = lib().tt_petsc_read_vecs(os.fsencode(path), len(out), ptrs, ls)
    if rc != 0:
        raise TTError(rc, lib().tt_last_error(None).decode())
    return out


class MMA:
    """MMA with one constraint on the context's GPU (MMA.cc)."""

    def __init__(self, ctx, a=0.0, c=1000.0, d=0.0):
        self.ctx = ctx
        h = ctypes.c_void_p()
        ctx._chk(lib().tt_mma_create(ctx.h, a, c, d, ctypes.byref(h)))
        self.h = h
        self.n = ctx.nel

    def __del__(self):
        if getattr(self, "h", None):
            lib().tt_mma_destroy(self.h)
            self.h = None

    def set_asymptotes(self, init, dec, inc):
        self.ctx._chk(lib().tt_mma_set_asymptotes(self.h, init, dec, inc))

    def outer_movelimit(self, Xmin, Xmax, movlim, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        lo = np.zeros_like(x)
        hi = np.zeros_like(x)
        self.ctx._chk(lib().tt_mma_set_outer_movelimit(self.h, Xmin, Xmax, movlim, x.ctypes.data, lo.ctypes.data,
                                                       hi.ctypes.data, 0))
        return lo, hi

    def update(self, x, dfdx, gx, dgdx, xmin, xmax):
        x = np.array(x, dtype=np.float64)
        g = np.array([gx], dtype=np.float64).ravel()
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (dfdx, dgdx, xmin, xmax)]
        self.ctx._chk(lib().tt_mma_update(self.h, x.ctypes.data, arrs[0].ctypes.data, g.ctypes.data,
                                          arrs[1].ctypes.data, arrs[2].ctypes.data, arrs[3].ctypes.data, 0))
        return x

    def design_change(self, x, xold):
        x = np.ascontiguousarray(x, dtype=np.float64)
        xo = np.array(xold, dtype=np.float64)
        ch = ctypes.c_double()
        self.ctx._chk(lib().tt_mma_design_change(self.h, x.ctypes.data, xo.ctypes.data, ctypes.byref(ch), 0))
        return ch.value, xo

    def kkt_residual(self, x, dfdx, gx, dgdx, xmin, xmax):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (x, dfdx, dgdx, xmin, xmax)]
        g = np.array([gx], dtype=np.float64).ravel()
        n2 = ctypes.c_double()
        ni = ctypes.c_double()
        self.ctx._chk(lib().tt_mma_kkt_residual(self.h, arrs[0].ctypes.data, arrs[1].ctypes.data, g.ctypes.data,
                                                arrs[2].ctypes.data, arrs[3].ctypes.data, arrs[4].ctypes.data,
                                                ctypes.byref(n2), ctypes.byref(ni), 0))
        return n2.value, ni.value

    def state(self):
        """(k, xold1, xold2, U, L): the restart state of MMA::Restart (MMA.cc:307-350)."""
        arrs = [np.zeros(self.n) for _ in range(4)]
        k = ctypes.c_int()
        self.ctx._chk(lib().tt_mma_get_state(self.h, ctypes.byref(k), *[a.ctypes.data for a in arrs]))
        return (k.value, *arrs)

    def set_state(self, k, xo1, xo2, U, L):
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (xo1, xo2, U, L)]
        self.ctx._chk(lib().tt_mma_set_state(self.h, int(k), *[a.ctypes.data for a in arrs]))

    def asymptotes(self):
        L = np.zeros(self.n)
        U = np.zeros(self.n)
        lam = ctypes.c_double()
        self.ctx._chk(lib().tt_mma_get_asymptotes(self.h, L.ctypes.data, U.ctypes.data, ctypes.byref(lam)))
        return L, U, lam.value
