"""ctypes front end of the FP64 CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module; the product library never does.  The arithmetic lives in
topopt_oracle.c, which restates the reference's LinearElasticity.cc,
PDEFilter.cc, Filter.cc and MMA.cc (each C function cites its lines).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# TT_ORACLE_LIB: another build of the oracle (the sanitizer test run loads oracle/_asan/liboracle.so)
_LIB_PATH = os.environ.get("TT_ORACLE_LIB") or os.path.join(_HERE, "liboracle.so")


class Params(ctypes.Structure):
    _fields_ = [
        ("nelx", ctypes.c_int), ("nely", ctypes.c_int), ("nelz", ctypes.c_int),
        ("xc", ctypes.c_double * 6),
        ("nu", ctypes.c_double),
        ("sInt", ctypes.c_int), ("mInt", ctypes.c_int),
        ("Emin", ctypes.c_double), ("Emax", ctypes.c_double),
        ("Rmin", ctypes.c_double), ("Rmax", ctypes.c_double),
        ("penalK", ctypes.c_double), ("penalM", ctypes.c_double),
        ("dalpha", ctypes.c_double), ("dbeta", ctypes.c_double),
        ("ti", ctypes.c_int), ("nsteps", ctypes.c_int), ("t1", ctypes.c_double),
        ("objective", ctypes.c_int), ("boundaries", ctypes.c_int),
        ("loadloc", ctypes.c_int), ("loadsignal", ctypes.c_int), ("pulse_steps", ctypes.c_int),
        ("region", ctypes.c_int),
        ("rtol", ctypes.c_double), ("maxit", ctypes.c_int),
    ]


def cantilever_params(nelx=32, nely=16, nelz=16, nsteps=100, dt=0.01, lx=2.0, ly=1.0, lz=1.0,
                      objective=3, rtol=1e-12, pulse_steps=20, **kw):
    """The cantilever pulse problem: x=0 face clamped, y-line load at x=lx, z=0,
    half-sine over `pulse_steps`, SIMP p=3 stiffness, linear mass, Rayleigh C=1e-2 M + 1e-4 K."""
    p = Params()
    p.nelx, p.nely, p.nelz = nelx, nely, nelz
    for i, v in enumerate((0.0, lx, 0.0, ly, 0.0, lz)):
        p.xc[i] = v
    p.nu = 0.3
    p.sInt, p.mInt = 1, 0
    p.Emax, p.Rmax = 1.0, 1.0
    p.Emin, p.Rmin = 1e-9 * p.Emax, 1e-3 * p.Rmax
    p.penalK, p.penalM = 3.0, 1.0
    p.dalpha, p.dbeta = 1e-2, 1e-4
    p.ti = 1
    p.nsteps = nsteps
    p.t1 = dt * nsteps
    p.objective = objective
    p.boundaries = 1
    p.loadloc = 2
    p.loadsignal = 1
    p.pulse_steps = pulse_steps
    p.region = 2
    p.rtol = rtol
    p.maxit = 100000
    for k, v in kw.items():
        setattr(p, k, v)
    return p


def build():
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.POINTER(Params)
        dp = ctypes.POINTER(ctypes.c_double)
        sig = {
            "or_nn": (ctypes.c_long, [P]), "or_nel": (ctypes.c_long, [P]),
            "or_element_nodes": (None, [P, ctypes.c_long, ctypes.POINTER(ctypes.c_long)]),
            "or_hex8_ke": (None, [ctypes.c_double] * 4 + [dp]),
            "or_hex8_me": (None, [ctypes.c_double] * 3 + [dp]),
            "or_pdefilter_ke": (None, [ctypes.c_double] * 4 + [dp]),
            "or_newmark_consts": (None, [ctypes.c_int, ctypes.c_double, dp]),
            "or_interp": (None, [ctypes.c_int] + [ctypes.c_double] * 4 + [dp, dp]),
            "or_bc_mask": (None, [P, dp]), "or_load_location": (None, [P, dp]),
            "or_load_intensity": (ctypes.c_double, [P, ctypes.c_int]),
            "or_region": (None, [P, dp]),
            "or_apply": (None, [P, dp, ctypes.c_double, ctypes.c_double, dp, dp]),
            "or_apply_bc": (None, [P, dp, ctypes.c_double, ctypes.c_double, dp, dp, dp]),
            "or_pcg": (ctypes.c_int, [P, dp, ctypes.c_double, ctypes.c_double, dp, dp, dp,
                                      ctypes.c_double, ctypes.c_int, dp]),
            "or_forward": (ctypes.c_int, [P, dp, dp]),
            "or_forward_u0": (ctypes.c_int, [P, dp, dp, dp]),
            "or_objective_gradient": (ctypes.c_int, [P, dp, dp, dp, dp]),
            "or_objective": (ctypes.c_double, [P, dp, dp]),
            "or_pde_filter": (ctypes.c_int, [P, ctypes.c_double, dp, dp, ctypes.c_double]),
            "or_smooth_projection": (ctypes.c_double, [ctypes.c_double] * 3),
            "or_chainrule_projection": (ctypes.c_double, [ctypes.c_double] * 3),
            "or_filter_project": (ctypes.c_int, [P, ctypes.c_double, dp, dp, dp, ctypes.c_int,
                                                 ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "or_filter_gradients": (ctypes.c_int, [P, ctypes.c_double, dp, dp, dp, ctypes.c_int,
                                                   ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "or_mma_create": (ctypes.c_void_p, [ctypes.c_long, ctypes.c_int, dp, dp, dp, dp]),
            "or_mma_destroy": (None, [ctypes.c_void_p]),
            "or_mma_set_asymptotes": (None, [ctypes.c_void_p] + [ctypes.c_double] * 3),
            "or_mma_set_outer_movelimit": (None, [ctypes.c_long] + [ctypes.c_double] * 3 + [dp, dp, dp]),
            "or_mma_update": (None, [ctypes.c_void_p, dp, dp, dp, dp, dp, dp]),
            "or_mma_design_change": (ctypes.c_double, [ctypes.c_long, dp, dp]),
            "or_mma_kkt_residual": (None, [ctypes.c_void_p, dp, dp, dp, dp, dp, dp, dp, dp]),
            "or_mma_get_state": (None, [ctypes.c_void_p, dp, dp, dp]),
            "or_sommg_basis": (ctypes.c_int, [P, dp, ctypes.c_int, ctypes.c_double, dp, dp]),
            "or_sommg_reduce": (None, [P, dp, ctypes.c_int, dp, dp, dp, dp, dp]),
            "or_reduced_forward": (ctypes.c_int, [P, ctypes.c_int, dp, dp, dp, dp, dp]),
            "or_reduced_expand": (None, [ctypes.c_long, ctypes.c_int, ctypes.c_int, dp, dp, dp]),
            "or_reduced_objective_gradient": (ctypes.c_int, [P, dp, ctypes.c_int, dp, dp, dp, dp, dp, dp, dp]),
            "or_design_loop": (ctypes.c_int, [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, dp, dp, dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def nn(p):
    return lib().or_nn(ctypes.byref(p))


def nel(p):
    return lib().or_nel(ctypes.byref(p))


def element_nodes(p, e):
    out = (ctypes.c_long * 8)()
    lib().or_element_nodes(ctypes.byref(p), e, out)
    return list(out)


def hex8_ke(dx, dy, dz, nu):
    ke = np.zeros(576)
    lib().or_hex8_ke(dx, dy, dz, nu, _p(ke))
    return ke.reshape(24, 24)


def hex8_me(dx, dy, dz):
    me = np.zeros(576)
    lib().or_hex8_me(dx, dy, dz, _p(me))
    return me.reshape(24, 24)


def pdefilter_ke(dx, dy, dz, R):
    kf = np.zeros(64)
    lib().or_pdefilter_ke(dx, dy, dz, R, _p(kf))
    return kf.reshape(8, 8)


def newmark_consts(ti, dt):
    a = np.zeros(6)
    lib().or_newmark_consts(ti, dt, _p(a))
    return a


def interp_c(scheme, vmin, vmax, penal, x):
    """The C oracle's material interpolation (value, derivative) at a scalar x."""
    f = ctypes.c_double()
    df = ctypes.c_double()
    lib().or_interp(scheme, vmin, vmax, penal, x, ctypes.byref(f), ctypes.byref(df))
    return f.value, df.value


def bc_mask(p):
    N = np.zeros(3 * nn(p))
    lib().or_bc_mask(ctypes.byref(p), _p(N))
    return N


def load_location(p):
    F = np.zeros(3 * nn(p))
    lib().or_load_location(ctypes.byref(p), _p(F))
    return F


def load_intensity(p, step):
    return lib().or_load_intensity(ctypes.byref(p), step)


def region(p):
    P = np.zeros(nel(p))
    lib().or_region(ctypes.byref(p), _p(P))
    return P


def apply(p, xphys, cK, cM, u):
    xphys, u = _f64(xphys), _f64(u)
    y = np.zeros_like(u)
    lib().or_apply(ctypes.byref(p), _p(xphys), cK, cM, _p(u), _p(y))
    return y


def apply_bc(p, xphys, cK, cM, u, N=None):
    xphys, u = _f64(xphys), _f64(u)
    N = bc_mask(p) if N is None else _f64(N)
    y = np.zeros_like(u)
    lib().or_apply_bc(ctypes.byref(p), _p(xphys), cK, cM, _p(N), _p(u), _p(y))
    return y


def pcg(p, xphys, cK, cM, b, rtol=1e-12, maxit=100000, x0=None):
    xphys, b = _f64(xphys), _f64(b)
    N = bc_mask(p)
    x = np.zeros_like(b) if x0 is None else _f64(x0).copy()
    rel = ctypes.c_double()
    it = lib().or_pcg(ctypes.byref(p), _p(xphys), cK, cM, _p(N), _p(b), _p(x), rtol, maxit, ctypes.byref(rel))
    return x, it, rel.value


def forward(p, xphys, u0=None):
    """States (nsteps+1, 3, nd); u0: initial displacement (masked to the free dofs), default 0."""
    xphys = _f64(xphys)
    nd = 3 * nn(p)
    states = np.zeros((p.nsteps + 1) * 3 * nd)
    if u0 is None:
        st = lib().or_forward(ctypes.byref(p), _p(xphys), _p(states))
    else:
        st = lib().or_forward_u0(ctypes.byref(p), _p(xphys), _p(_f64(u0)), _p(states))
    if st != 0:
        raise RuntimeError("oracle forward: CG did not converge")
    return states.reshape(p.nsteps + 1, 3, nd)


def objective_gradient(p, xphys, states):
    xphys = _f64(xphys)
    states = _f64(states)
    fx = ctypes.c_double()
    df = np.zeros(nel(p))
    st = lib().or_objective_gradient(ctypes.byref(p), _p(xphys), _p(states), ctypes.byref(fx), _p(df))
    if st != 0:
        raise RuntimeError("oracle adjoint: CG did not converge")
    return fx.value, df


def objective(p, xphys, states):
    return lib().or_objective(ctypes.byref(p), _p(_f64(xphys)), _p(_f64(states)))


def pde_filter(p, rmin, x, rtol=1e-12):
    x = _f64(x)
    out = np.zeros_like(x)
    it = lib().or_pde_filter(ctypes.byref(p), rmin, _p(x), _p(out), rtol)
    if it < 0:
        raise RuntimeError("oracle PDE filter did not converge")
    return out


def smooth_projection(x, beta, eta):
    return lib().or_smooth_projection(x, beta, eta)


def chainrule_projection(x, beta, eta):
    return lib().or_chainrule_projection(x, beta, eta)


def filter_project(p, rmin, x, projection=False, beta=0.1, eta=0.15, rtol=1e-12):
    x = _f64(x)
    xt = np.zeros_like(x)
    xp = np.zeros_like(x)
    lib().or_filter_project(ctypes.byref(p), rmin, _p(x), _p(xt), _p(xp), int(projection), beta, eta, rtol)
    return xt, xp


def filter_gradients(p, rmin, xtilde, dfdx, dgdx, projection=False, beta=0.1, eta=0.15, rtol=1e-12):
    xt = _f64(xtilde)
    df = _f64(dfdx).copy()
    dg = _f64(dgdx).copy()
    lib().or_filter_gradients(ctypes.byref(p), rmin, _p(xt), _p(df), _p(dg), int(projection), beta, eta, rtol)
    return df, dg


class MMA:
    """Oracle MMA (MMA.cc) for m constraints; dgdx is (m, n)."""

    def __init__(self, n, m, x, a=None, c=None, d=None):
        self.n, self.m = n, m
        a = np.zeros(m) if a is None else _f64(a)
        c = np.full(m, 1000.0) if c is None else _f64(c)
        d = np.zeros(m) if d is None else _f64(d)
        self._keep = (a, c, d)
        self.h = lib().or_mma_create(n, m, _p(_f64(x)), _p(a), _p(c), _p(d))

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_mma_destroy(self.h)
            self.h = None

    def set_asymptotes(self, init, dec, inc):
        lib().or_mma_set_asymptotes(self.h, init, dec, inc)

    @staticmethod
    def outer_movelimit(Xmin, Xmax, movlim, x):
        x = _f64(x)
        xmin = np.zeros_like(x)
        xmax = np.zeros_like(x)
        lib().or_mma_set_outer_movelimit(len(x), Xmin, Xmax, movlim, _p(x), _p(xmin), _p(xmax))
        return xmin, xmax

    def update(self, x, dfdx, gx, dgdx, xmin, xmax):
        x = _f64(x).copy()
        lib().or_mma_update(self.h, _p(x), _p(_f64(dfdx)), _p(_f64(np.atleast_1d(gx))),
                            _p(_f64(dgdx)), _p(_f64(xmin)), _p(_f64(xmax)))
        return x

    @staticmethod
    def design_change(x, xold):
        x = _f64(x)
        xo = _f64(xold).copy()
        ch = lib().or_mma_design_change(len(x), _p(x), _p(xo))
        return ch, xo

    def kkt_residual(self, x, dfdx, gx, dgdx, xmin, xmax):
        n2 = ctypes.c_double()
        ni = ctypes.c_double()
        lib().or_mma_kkt_residual(self.h, _p(_f64(x)), _p(_f64(dfdx)), _p(_f64(np.atleast_1d(gx))),
                                  _p(_f64(dgdx)), _p(_f64(xmin)), _p(_f64(xmax)), ctypes.byref(n2),
                                  ctypes.byref(ni))
        return n2.value, ni.value

    def state(self):
        L = np.zeros(self.n)
        U = np.zeros(self.n)
        lam = np.zeros(self.m)
        lib().or_mma_get_state(self.h, _p(L), _p(U), _p(lam))
        return L, U, lam


def design_loop(p, rmin, volfrac, V0, maxit, projection=False, beta=0.1, eta=0.15, movlim=0.2):
    n = nel(p)
    x = np.zeros(n)
    xp = np.zeros(n)
    hist = np.zeros(5 * maxit)
    itr = lib().or_design_loop(ctypes.byref(p), rmin, volfrac, V0, maxit, int(projection), beta, eta, movlim,
                               _p(x), _p(xp), _p(hist))
    if itr < 0:
        raise RuntimeError("oracle design loop: a solve did not converge")
    return x, xp, hist.reshape(maxit, 5)[:itr]


# ---------------------------------------------------------------- SOMMG reduced-order model
def sommg_basis(p, xphys, nb, omega0):
    """Basis Q (nb x nd, orthonormal) and Gram-Schmidt factor Ut (SOMMG.cc:40-230)."""
    nd = 3 * nn(p)
    Q = np.zeros((nb, nd))
    Ut = np.zeros((nb, nb))
    if lib().or_sommg_basis(ctypes.byref(p), _p(_f64(xphys)), nb, omega0, _p(Q), _p(Ut)) != 0:
        raise RuntimeError("oracle SOMMG basis: CG did not converge")
    return Q, Ut


def sommg_reduce(p, xphys, Q):
    """Reduced Kr, Mr, Cr (nb x nb) and load Pr (SOMMG.cc:236-262)."""
    Q = _f64(Q)
    nb = Q.shape[0]
    Kr, Mr, Cr = (np.zeros((nb, nb)) for _ in range(3))
    Pr = np.zeros(nb)
    lib().or_sommg_reduce(ctypes.byref(p), _p(_f64(xphys)), nb, _p(Q), _p(Kr), _p(Mr), _p(Cr), _p(Pr))
    return Kr, Mr, Cr, Pr


def reduced_forward(p, Kr, Mr, Cr, Pr):
    """Reduced Newmark states (nsteps+1, 3, nb)."""
    nb = len(Pr)
    rs = np.zeros((p.nsteps + 1) * 3 * nb)
    if lib().or_reduced_forward(ctypes.byref(p), nb, _p(_f64(Kr)), _p(_f64(Mr)), _p(_f64(Cr)), _p(_f64(Pr)),
                                _p(rs)) != 0:
        raise RuntimeError("oracle reduced forward: singular system")
    return rs.reshape(p.nsteps + 1, 3, nb)


def reduced_expand(p, Q, rstates):
    """Full states Q r (nsteps+1, 3, nd)."""
    Q = _f64(Q)
    nb, nd = Q.shape
    full = np.zeros((p.nsteps + 1) * 3 * nd)
    lib().or_reduced_expand(nd, nb, p.nsteps, _p(Q), _p(_f64(rstates)), _p(full))
    return full.reshape(p.nsteps + 1, 3, nd)


def reduced_objective_gradient(p, xphys, Q, Kr, Mr, Cr, rstates):
    """Objective and sensitivities of the reduced model with the basis held fixed."""
    Q = _f64(Q)
    nb = Q.shape[0]
    fx = ctypes.c_double()
    df = np.zeros(nel(p))
    if lib().or_reduced_objective_gradient(ctypes.byref(p), _p(_f64(xphys)), nb, _p(Q), _p(_f64(Kr)),
                                           _p(_f64(Mr)), _p(_f64(Cr)), _p(_f64(rstates)), ctypes.byref(fx),
                                           _p(df)) != 0:
        raise RuntimeError("oracle reduced adjoint: singular system")
    return fx.value, df


def reduced_model(p, xphys, nb, omega0):
    """The whole reduced path for one design: basis, reduced matrices, states, objective, gradient."""
    Q, Ut = sommg_basis(p, xphys, nb, omega0)
    Kr, Mr, Cr, Pr = sommg_reduce(p, xphys, Q)
    rs = reduced_forward(p, Kr, Mr, Cr, Pr)
    fx, df = reduced_objective_gradient(p, xphys, Q, Kr, Mr, Cr, rs)
    return dict(Q=Q, Ut=Ut, Kr=Kr, Mr=Mr, Cr=Cr, Pr=Pr, rstates=rs, fx=fx, dfdx=df)


# ---------------------------------------------------------------- assembled forms (scipy checks)
def assemble(p, xphys, cK, cM):
    """Assembled CSR of sum_e (cK kfac KE + cM mfac ME) without BCs (for scipy cross-checks)."""
    import scipy.sparse as sp
    dx = (p.xc[1] - p.xc[0]) / p.nelx
    dy = (p.xc[3] - p.xc[2]) / p.nely
    dz = (p.xc[5] - p.xc[4]) / p.nelz
    ke = hex8_ke(dx, dy, dz, p.nu)
    me = hex8_me(dx, dy, dz)
    ne = nel(p)
    x = _f64(xphys)
    kf = interp(p.sInt, p.Emin, p.Emax, p.penalK, x)[0]
    mf = interp(p.mInt, p.Rmin, p.Rmax, p.penalM, x)[0]
    rows, cols, vals = [], [], []
    for e in range(ne):
        nodes = element_nodes(p, e)
        ed = np.array([3 * n + c for n in nodes for c in range(3)])
        blk = cK * kf[e] * ke + cM * mf[e] * me
        rows.append(np.repeat(ed, 24))
        cols.append(np.tile(ed, 24))
        vals.append(blk.ravel())
    nd = 3 * nn(p)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(nd, nd))


def interp(scheme, vmin, vmax, penal, x):
    x = np.asarray(x, dtype=np.float64)
    if scheme == 0:
        return vmin + x * (vmax - vmin), np.full_like(x, vmax - vmin)
    if scheme == 1:
        return vmin + x ** penal * (vmax - vmin), penal * x ** (penal - 1) * (vmax - vmin)
    return (vmin + x / (1 + penal * (1 - x)) * (vmax - vmin),
            (1 + penal) / (1 + penal * (1 - x)) ** 2 * (vmax - vmin))
