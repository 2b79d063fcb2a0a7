"""ctypes binding of the FP64 CPU oracle.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product path never imports
this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm_oracle.so")
REF_MRT_PATH = os.path.join(HERE, "_ref", "libref_cellfunctions.so")

SOLID, FLUID, VEL_W, PRES_E, VEL_N = 0, 1, 2, 3, 4
VEL_E, VEL_S, VEL_B, VEL_T = 5, 6, 7, 8
OUT_W, OUT_S, OUT_N, OUT_B, OUT_T = 9, 10, 11, 12, 13
SYM_W, SYM_E, SYM_S, SYM_N, SYM_B, SYM_T = 14, 15, 16, 17, 18, 19
BGK, TRT, MRT = 0, 1, 2
FORCE_NONE, FORCE_SHIFT, FORCE_GUO = 0, 1, 2
WALL_HALFWAY, WALL_ONNODE, WALL_HH, WALL_CURVED = 0, 1, 2, 3
INLET_ZOUHE, INLET_REF2D = 0, 1
OUTLET_PRESSURE, OUTLET_REF_VELOCITY, OUTLET_EXTRAP2, OUTLET_EXTRAP1, OUTLET_FEQ = 0, 1, 2, 3, 4
PARITY_MRT_RHO2, PARITY_MRT_NO_F18, PARITY_INLET_FROZEN, PARITY_OUTLET_V0 = 1, 2, 4, 8
PARITY_INLET3D_FLIP = 0x10


class OrcParams(C.Structure):
    _fields_ = [
        ("q", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("periodic", C.c_int * 3), ("collision", C.c_int), ("omega", C.c_double),
        ("trt_magic", C.c_double), ("mrt_s", C.c_double * 19), ("force_model", C.c_int),
        ("g", C.c_double * 3), ("wall_model", C.c_int), ("inlet_model", C.c_int),
        ("parity", C.c_uint), ("rho_out", C.c_double), ("nthreads", C.c_int),
        ("outlet_model", C.c_int),
    ]


class OrcCgParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("r_alpha", C.c_double), ("b_alpha", C.c_double),
        ("r_nu", C.c_double), ("b_nu", C.c_double), ("beta", C.c_double), ("A", C.c_double),
        ("g_limit", C.c_double), ("high_order", C.c_int), ("g", C.c_double),
        ("external_force", C.c_int), ("kappa", C.c_double), ("nthreads", C.c_int), ("nz", C.c_int),
        ("enhanced", C.c_int), ("walls", C.c_int), ("u_wall", C.c_double),
    ]


class OrcCgCpuParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("r_alpha", C.c_double), ("b_alpha", C.c_double),
        ("r_nu", C.c_double), ("b_nu", C.c_double), ("beta", C.c_double), ("A", C.c_double),
        ("g_limit", C.c_double), ("test_case", C.c_int), ("u_wall", C.c_double), ("nz", C.c_int),
    ]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        dp = P(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [P(OrcParams), P(C.c_uint8), P(C.c_uint32), dp, dp, dp]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_init.argtypes = [C.c_void_p, C.c_int, dp, dp, dp, dp]
        L.orc_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_get_f.argtypes = [C.c_void_p, dp]
        L.orc_set_f.argtypes = [C.c_void_p, dp]
        L.orc_get_macros.argtypes = [C.c_void_p, dp, dp, dp, dp]
        L.orc_mass.argtypes = [C.c_void_p]
        L.orc_mass.restype = C.c_double
        L.orc_residual_l2.argtypes = [C.c_void_p]
        L.orc_residual_l2.restype = C.c_double
        L.orc_collide_node.argtypes = [C.c_void_p, dp, C.c_double, dp, dp]
        L.orc_open_boundary.argtypes = [C.c_void_p, C.c_long, dp]
        L.orc_wall_node.argtypes = [C.c_void_p, C.c_long, dp]
        L.orc_set_wall_q.argtypes = [C.c_void_p, dp]
        L.orc_equilibrium.argtypes = [C.c_int, C.c_double, dp, dp]
        L.orc_lattice.argtypes = [C.c_int, P(C.c_int), P(C.c_int), P(C.c_int), P(C.c_int), dp]
        L.orc_mrt_matrices.argtypes = [C.c_int, C.c_double, dp, dp, dp]
        L.orc_cg_create.restype = C.c_void_p
        L.orc_cg_create.argtypes = [P(OrcCgParams)]
        L.orc_cg_destroy.argtypes = [C.c_void_p]
        L.orc_cg_init.argtypes = [C.c_void_p, dp, dp]
        L.orc_cg_init_f.argtypes = [C.c_void_p, dp, dp]
        L.orc_cg_get_w.argtypes = [C.c_void_p, dp]
        L.orc_cg_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_cg_get.argtypes = [C.c_void_p] + [dp] * 7
        L.orc_cg_laplace.argtypes = [C.c_void_p, dp]
        L.orc_cg_gradient.argtypes = [C.c_void_p, dp, dp, dp]
        _lib = L
    return _lib


def _dp(a):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel()


def lattice(q):
    cx = (C.c_int * q)(); cy = (C.c_int * q)(); cz = (C.c_int * q)(); opp = (C.c_int * q)()
    w = (C.c_double * q)()
    lib().orc_lattice(q, cx, cy, cz, opp, w)
    return (np.array(cx[:]), np.array(cy[:]), np.array(cz[:]), np.array(opp[:]), np.array(w[:]))


def mrt_matrices(q, omega):
    M = np.zeros(q * q); Mi = np.zeros(q * q); s = np.zeros(q)
    lib().orc_mrt_matrices(q, omega, _dp(M), _dp(Mi), _dp(s))
    return M.reshape(q, q), Mi.reshape(q, q), s


def equilibrium(q, rho, uvw):
    out = np.zeros(q)
    u = np.array(list(uvw) + [0.0] * (3 - len(uvw)), dtype=np.float64)
    lib().orc_equilibrium(q, rho, _dp(u), _dp(out))
    return out


class Sim:
    """One oracle simulation; arrays are (Q, nz, ny, nx) / (nz, ny, nx)."""

    def __init__(self, q, shape, node_class, bb=None, *, collision=BGK, omega=1.0, trt_magic=0.0,
                 mrt_s=None, force_model=FORCE_NONE, g=(0.0, 0.0, 0.0), wall_model=WALL_HALFWAY,
                 inlet_model=INLET_ZOUHE, parity=0, rho_out=1.0, periodic=(0, 0, 0), ubc=None,
                 vbc=None, wbc=None, nthreads=0, outlet_model=OUTLET_PRESSURE):
        nz, ny, nx = (1,) + tuple(shape) if len(shape) == 2 else tuple(shape)
        self.q, self.shape3 = q, (nz, ny, nx)
        self.n = nx * ny * nz
        p = OrcParams()
        p.q, p.nx, p.ny, p.nz = q, nx, ny, nz
        for i in range(3):
            p.periodic[i] = int(periodic[i])
            p.g[i] = float(g[i]) if i < len(g) else 0.0
        p.collision, p.omega, p.trt_magic = collision, omega, trt_magic
        if mrt_s is not None:
            for i, v in enumerate(mrt_s):
                p.mrt_s[i] = v
        p.force_model, p.wall_model, p.inlet_model = force_model, wall_model, inlet_model
        p.parity, p.rho_out, p.nthreads = parity, rho_out, nthreads
        p.outlet_model = outlet_model
        self.params = p
        cls = np.ascontiguousarray(node_class, dtype=np.uint8).ravel()
        assert cls.size == self.n
        bbv = np.zeros(self.n, np.uint32) if bb is None else np.ascontiguousarray(bb, dtype=np.uint32).ravel()
        self._keep = [_f64(ubc), _f64(vbc), _f64(wbc)]
        self.h = lib().orc_create(C.byref(p), cls.ctypes.data_as(C.POINTER(C.c_uint8)),
                                  bbv.ctypes.data_as(C.POINTER(C.c_uint32)),
                                  *[_dp(a) for a in self._keep])
        if not self.h:
            raise ValueError("oracle: bad parameters")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def init(self, mode=0, rho=None, u=None, v=None, w=None):
        arrs = [_f64(a) for a in (rho, u, v, w)]
        lib().orc_init(self.h, mode, *[_dp(a) for a in arrs])

    def step(self, n=1):
        lib().orc_step(self.h, int(n))

    def get_f(self):
        f = np.zeros(self.q * self.n)
        lib().orc_get_f(self.h, _dp(f))
        return f.reshape((self.q,) + self.shape3)

    def set_f(self, f):
        f = _f64(f)
        lib().orc_set_f(self.h, _dp(f))

    def macros(self):
        out = [np.zeros(self.n) for _ in range(4)]
        lib().orc_get_macros(self.h, *[_dp(a) for a in out])
        return [a.reshape(self.shape3) for a in out]

    def mass(self):
        return lib().orc_mass(self.h)

    def residual_l2(self):
        return lib().orc_residual_l2(self.h)

    def collide_node(self, f, rho, uvw):
        f = _f64(f); u = np.array(list(uvw) + [0.0] * (3 - len(uvw)))
        out = np.zeros(self.q)
        lib().orc_collide_node(self.h, _dp(f), rho, _dp(u), _dp(out))
        return out

    def open_boundary(self, ind, f):
        """Node-local open-boundary condition of dense node ind on streamed f[Q]."""
        g = np.array(f, dtype=np.float64)
        lib().orc_open_boundary(self.h, int(ind), _dp(g))
        return g

    def set_wall_q(self, q):
        """q (N x 19, or (..., 19) in node order): wall fraction of each wall link."""
        self._wq = np.ascontiguousarray(q, dtype=np.float64).reshape(-1)
        assert self._wq.size == 19 * self.n
        lib().orc_set_wall_q(self.h, _dp(self._wq))

    def wall_node(self, ind, f):
        """Wall condition (on-node or wet-node) of dense node ind on streamed f[Q]."""
        g = np.array(f, dtype=np.float64)
        lib().orc_wall_node(self.h, int(ind), _dp(g))
        return g


class CgSim:
    """Colour-gradient oracle on a periodic box: D2Q9 for CgSim(nx, ny) with
    arrays (ny, nx), D3Q19 for CgSim(nx, ny, nz=...) with arrays (nz, ny, nx)."""

    def __init__(self, nx, ny, *, nz=1, r_alpha=4 / 9, b_alpha=4 / 9, r_nu=1 / 6, b_nu=1 / 6, beta=0.7,
                 A=0.01125, g_limit=0.0, high_order=0, g=0.0, external_force=0, kappa=0.9,
                 nthreads=0, enhanced=0, walls=0, u_wall=0.0):
        p = OrcCgParams()
        p.nx, p.ny, p.r_alpha, p.b_alpha, p.r_nu, p.b_nu = nx, ny, r_alpha, b_alpha, r_nu, b_nu
        p.beta, p.A, p.g_limit, p.high_order, p.g = beta, A, g_limit, high_order, g
        p.external_force, p.kappa, p.nthreads, p.nz = external_force, kappa, nthreads, nz
        p.enhanced, p.walls, p.u_wall = enhanced, walls, u_wall
        self.params, self.nx, self.ny, self.nz = p, nx, ny, nz
        self.n = nx * ny * nz
        self.q = 19 if nz > 1 else 9
        self.shape = (nz, ny, nx) if nz > 1 else (ny, nx)
        self.h = lib().orc_cg_create(C.byref(p))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cg_destroy(self.h)
            self.h = None

    def init(self, r_rho, b_rho):
        r = _f64(r_rho); b = _f64(b_rho)
        lib().orc_cg_init(self.h, _dp(r), _dp(b))

    def init_f(self, r_f, b_f):
        r = _f64(r_f); b = _f64(b_f)
        lib().orc_cg_init_f(self.h, _dp(r), _dp(b))

    def step(self, n=1):
        lib().orc_cg_step(self.h, int(n))

    def get(self):
        n, q = self.n, self.q
        arrs = [np.zeros(q * n), np.zeros(q * n)] + [np.zeros(n) for _ in range(5)]
        lib().orc_cg_get(self.h, *[_dp(a) for a in arrs])
        shq = (q,) + self.shape
        keys = ["r_f", "b_f", "r_rho", "b_rho", "rho", "u", "v"]
        out = {k: a.reshape(shq if i < 2 else self.shape) for i, (k, a) in enumerate(zip(keys, arrs))}
        w = np.zeros(n)
        lib().orc_cg_get_w(self.h, _dp(w))
        out["w"] = w.reshape(self.shape)
        return out

    def laplace(self):
        out = np.zeros(5)
        lib().orc_cg_laplace(self.h, _dp(out))
        return out

    def gradient(self):
        g = [np.zeros(self.n) for _ in range(3)]
        lib().orc_cg_gradient(self.h, *[_dp(a) for a in g])
        if self.nz > 1:
            return tuple(a.reshape(self.shape) for a in g)
        return g[0].reshape(self.shape), g[1].reshape(self.shape)


class CgCpuSim:
    """The reference's CPU colour-gradient path on its full lattice: a second
    opinion beside the GPU formulation.  D2Q9 (mp2DColl, streamMP,
    peridicBoundaries, updateMacroMP of S/Multiphase.cu), arrays (ny, nx); with
    nz > 1 D3Q19 (mp3DColl, streamMP3D, peridicBoundaries3D, updateMacroMP3D),
    arrays (nz, ny, nx)."""

    def __init__(self, nx, ny, *, nz=1, r_alpha=4 / 9, b_alpha=4 / 9, r_nu=1 / 6, b_nu=1 / 6, beta=0.7, A=0.01125,
                 g_limit=0.0, test_case=1, u_wall=0.0):
        p = OrcCgCpuParams()
        p.nx, p.ny, p.r_alpha, p.b_alpha, p.r_nu, p.b_nu = nx, ny, r_alpha, b_alpha, r_nu, b_nu
        p.beta, p.A, p.g_limit, p.test_case, p.u_wall, p.nz = beta, A, g_limit, test_case, u_wall, nz
        self.nz, self.q = nz, 19 if nz > 1 else 9
        self.nx, self.ny, self.n = nx, ny, nx * ny * max(nz, 1)
        self.shape = (nz, ny, nx) if nz > 1 else (ny, nx)
        L = lib()
        L.orc_cgcpu_create.restype = C.c_void_p
        L.orc_cgcpu_create.argtypes = [C.POINTER(OrcCgCpuParams)]
        L.orc_cgcpu_destroy.argtypes = [C.c_void_p]
        dp = C.POINTER(C.c_double)
        L.orc_cgcpu_init.argtypes = [C.c_void_p, dp, dp]
        L.orc_cgcpu_step.argtypes = [C.c_void_p, C.c_int]
        L.orc_cgcpu_get.argtypes = [C.c_void_p] + [dp] * 7
        L.orc_cgcpu_get_coll.argtypes = [C.c_void_p] + [dp] * 4
        L.orc_cgcpu_init_f.argtypes = [C.c_void_p, dp, dp]
        L.orc_cgcpu_get_w.argtypes = [C.c_void_p, dp, dp]
        self.h = L.orc_cgcpu_create(C.byref(p))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_cgcpu_destroy(self.h)
            self.h = None

    def init(self, r_rho, b_rho):
        lib().orc_cgcpu_init(self.h, _dp(_f64(r_rho)), _dp(_f64(b_rho)))

    def step(self, n=1):
        lib().orc_cgcpu_step(self.h, int(n))

    def init_f(self, r_f, b_f):
        """colour populations (q, *shape) as the state before the first collision"""
        lib().orc_cgcpu_init_f(self.h, _dp(_f64(r_f)), _dp(_f64(b_f)))

    def get(self):
        n, q = self.n, self.q
        arrs = [np.zeros(q * n), np.zeros(q * n)] + [np.zeros(n) for _ in range(5)]
        lib().orc_cgcpu_get(self.h, *[_dp(a) for a in arrs])
        keys = ["r_f", "b_f", "r_rho", "b_rho", "rho", "u", "v"]
        out = {k: a.reshape((q,) + self.shape if i < 2 else self.shape) for i, (k, a) in enumerate(zip(keys, arrs))}
        if q == 19:
            w = np.zeros(n)
            lib().orc_cgcpu_get_w(self.h, _dp(w), None)
            out["w"] = w.reshape(self.shape)
        return out

    def collided(self):
        """the last step's collided colours and colour gradient (gx, gy[, gz])"""
        n, q = self.n, self.q
        arrs = [np.zeros(q * n), np.zeros(q * n), np.zeros(n), np.zeros(n)]
        lib().orc_cgcpu_get_coll(self.h, *[_dp(a) for a in arrs])
        out = [arrs[0].reshape((q,) + self.shape), arrs[1].reshape((q,) + self.shape),
               arrs[2].reshape(self.shape), arrs[3].reshape(self.shape)]
        if q == 19:
            gz = np.zeros(n)
            lib().orc_cgcpu_get_w(self.h, None, _dp(gz))
            out.append(gz.reshape(self.shape))
        return tuple(out)


def ref_mrt_tables(q, omega):
    """M and M^-1 S from the reference's own MRTInitializer2D/3D, when
    oracle/_ref was built (S/CellFunctions.cu:7-154)."""
    if not os.path.exists(REF_MRT_PATH):
        return None
    L = C.CDLL(REF_MRT_PATH)
    fn = L["_Z16MRTInitializer2DPdS_d" if q == 9 else "_Z16MRTInitializer3DPdS_d"]
    fn.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_double]
    M = np.zeros(q * q); MS = np.zeros(q * q)
    fn(_dp(M), _dp(MS), omega)
    return M.reshape(q, q), MS.reshape(q, q)
