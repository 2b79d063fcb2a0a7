"""Host-side binding of the C ABI (include/lbm.h) through ctypes.

Mirrors the solver interface of the reference (S/include/Arguments.h:86-135,
S/include/Iterate.h:57-65): a `Solver` is one configured lattice with its
populations on the GPU.  The HIP library is mandatory: importing works without
it, but creating a solver raises when thesis_amd/liblbm.so is missing or no
HIP device is visible -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbm.so")
ABI_VERSION = 3

SOLID, FLUID, VEL_W, PRES_E, VEL_N = 0, 1, 2, 3, 4
VEL_E, VEL_S, VEL_B, VEL_T = 5, 6, 7, 8
OUT_W, OUT_S, OUT_N, OUT_B, OUT_T = 9, 10, 11, 12, 13
SYM_W, SYM_E, SYM_S, SYM_N, SYM_B, SYM_T = 14, 15, 16, 17, 18, 19
BGK, TRT, MRT = 0, 1, 2
STORE_F64, STORE_F32, STORE_F16, STORE_BF16 = 0, 1, 2, 3
FORCE_NONE, FORCE_SHIFT, FORCE_GUO = 0, 1, 2
WALL_HALFWAY, WALL_ONNODE, WALL_HH, WALL_CURVED = 0, 1, 2, 3
INLET_ZOUHE, INLET_REF2D = 0, 1
OUTLET_PRESSURE, OUTLET_REF_VELOCITY, OUTLET_EXTRAP2, OUTLET_EXTRAP1, OUTLET_FEQ = 0, 1, 2, 3, 4
PARITY_MRT_RHO2, PARITY_MRT_NO_F18, PARITY_OUTLET_V0, PARITY_INLET3D_FLIP = 0x1, 0x2, 0x8, 0x10
PARITY_INLET_FROZEN = 0x4

STATUS = {0: "OK", 1: "DIVERGED", 2: "EIO", 3: "EINVAL", 4: "EDEVICE", 5: "ENOMEM", 6: "EUNSUPPORTED"}
STORAGE_NAMES = {"f64": STORE_F64, "f32": STORE_F32, "f16": STORE_F16, "bf16": STORE_BF16}

EXPORTED = [
    "lbm_abi_version", "lbm_device_count", "lbm_create", "lbm_destroy", "lbm_last_error", "lbm_init",
    "lbm_step", "lbm_run_to_residual", "lbm_reduce", "lbm_reduce_all", "lbm_residual_l2", "lbm_residual_l2_noalloc", "lbm_residual_fdrel", "lbm_get_macros", "lbm_get_f", "lbm_set_f",
    "lbm_time_steps", "lbm_time_launches", "lbm_device_bytes", "lbm_read_lattice", "lbm_read_wall_q", "lbm_latticegen",
    "lbm_latticegen_last_error", "lbm_write_lattice_binary",
    "lbm_cg_create", "lbm_cg_destroy", "lbm_cg_last_error", "lbm_cg_init", "lbm_cg_step", "lbm_cg_get_macros",
    "lbm_cg_get_f", "lbm_cg_reduce", "lbm_cg_time_steps", "lbm_cg_init_f", "lbm_cg_get_macros3d",
    "lbm_batch_create", "lbm_batch_destroy", "lbm_batch_last_error", "lbm_batch_init", "lbm_batch_run",
    "lbm_batch_get_macros", "lbm_batch_get_f", "lbm_batch_nodes",
]


class LbmError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("q", C.c_int32), ("nx", C.c_int32), ("ny", C.c_int32),
        ("nz", C.c_int32), ("periodic", C.c_int32 * 3), ("collision", C.c_int32),
        ("omega", C.c_double), ("trt_magic", C.c_double), ("mrt_s", C.c_double * 19),
        ("force_model", C.c_int32), ("force", C.c_double * 3), ("storage", C.c_int32),
        ("wall_model", C.c_int32), ("inlet_model", C.c_int32), ("parity_flags", C.c_uint32),
        ("rho_out", C.c_double), ("outlet_model", C.c_int32),
    ]


class CgParams(C.Structure):
    _fields_ = [
        ("abi_version", C.c_uint32), ("nx", C.c_int32), ("ny", C.c_int32), ("r_alpha", C.c_double),
        ("b_alpha", C.c_double), ("r_nu", C.c_double), ("b_nu", C.c_double), ("beta", C.c_double),
        ("A", C.c_double), ("g_limit", C.c_double), ("high_order", C.c_int32), ("kappa", C.c_double),
        ("storage", C.c_int32), ("nz", C.c_int32), ("enhanced", C.c_int32), ("walls", C.c_int32),
        ("u_wall", C.c_double), ("g", C.c_double), ("external_force", C.c_int32),
    ]


class Lattice(C.Structure):
    _fields_ = [
        ("node_class", C.POINTER(C.c_uint8)), ("bb_links", C.POINTER(C.c_uint32)),
        ("u_bc", C.POINTER(C.c_double)), ("v_bc", C.POINTER(C.c_double)), ("w_bc", C.POINTER(C.c_double)),
        ("wall_q", C.POINTER(C.c_double)),
    ]


class BatchInstance(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("node_class", C.POINTER(C.c_uint8)),
        ("bb_links", C.POINTER(C.c_uint32)), ("u_bc", C.POINTER(C.c_double)), ("v_bc", C.POINTER(C.c_double)),
        ("omega", C.c_double), ("trt_magic", C.c_double), ("rho_out", C.c_double),
    ]


_lib = None


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LbmError(4, f"{LIB_PATH} is missing: run `python -m thesis_amd.build` first")
    # PyTorch ships its own HIP and HSA runtimes.  A process must hold only
    # one: when this library's runtime (ROCm's) is loaded first, a later torch
    # finds no GPU, while the reverse order makes this library bind to the
    # runtime torch already loaded.  So torch, when installed, goes first.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    dp = P(C.c_double)
    vp = C.c_void_p
    L.lbm_abi_version.restype = C.c_int
    L.lbm_device_count.restype = C.c_int
    L.lbm_create.argtypes = [P(Params), P(Lattice), P(vp)]
    L.lbm_destroy.argtypes = [vp]
    L.lbm_destroy.restype = None
    L.lbm_last_error.argtypes = [vp]
    L.lbm_last_error.restype = C.c_char_p
    L.lbm_init.argtypes = [vp, C.c_int, dp, dp, dp, dp]
    L.lbm_step.argtypes = [vp, C.c_int64]
    L.lbm_run_to_residual.argtypes = [vp, C.c_double, C.c_int64, C.c_int64, P(C.c_int64), dp]
    L.lbm_reduce.argtypes = [vp, dp]
    L.lbm_reduce_all.argtypes = [vp, dp]
    L.lbm_residual_l2.argtypes = [vp, dp]
    L.lbm_residual_l2_noalloc.argtypes = [vp, dp]
    L.lbm_residual_fdrel.argtypes = [vp, dp]
    L.lbm_get_macros.argtypes = [vp, dp, dp, dp, dp]
    L.lbm_get_f.argtypes = [vp, dp]
    L.lbm_set_f.argtypes = [vp, dp]
    L.lbm_time_steps.argtypes = [vp, C.c_int64, dp]
    L.lbm_time_launches.argtypes = [vp, C.c_int64, dp]
    L.lbm_device_bytes.argtypes = [vp]
    L.lbm_device_bytes.restype = C.c_int64
    L.lbm_read_lattice.argtypes = [C.c_char_p, C.c_char_p, P(C.c_int32), P(C.c_int64), P(C.c_uint8),
                                   P(C.c_uint32), P(C.c_int64), P(C.c_int64), P(C.c_double), P(C.c_int32)]
    L.lbm_read_wall_q.argtypes = [C.c_char_p, C.c_char_p, dp]
    L.lbm_latticegen.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, dp, C.c_int32, C.c_char_p, C.c_char_p,
                                 C.c_char_p, P(C.c_int32), P(C.c_int64)]
    L.lbm_latticegen_last_error.restype = C.c_char_p
    L.lbm_write_lattice_binary.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p]
    L.lbm_cg_create.argtypes = [P(CgParams), P(vp)]
    L.lbm_cg_destroy.argtypes = [vp]
    L.lbm_cg_destroy.restype = None
    L.lbm_cg_last_error.argtypes = [vp]
    L.lbm_cg_last_error.restype = C.c_char_p
    L.lbm_cg_init.argtypes = [vp, dp, dp]
    L.lbm_cg_step.argtypes = [vp, C.c_int64]
    L.lbm_cg_get_macros.argtypes = [vp, dp, dp, dp, dp, dp]
    L.lbm_cg_get_macros3d.argtypes = [vp, dp, dp, dp, dp, dp, dp]
    L.lbm_cg_init_f.argtypes = [vp, dp, dp]
    L.lbm_cg_get_f.argtypes = [vp, dp, dp]
    L.lbm_cg_reduce.argtypes = [vp, dp]
    L.lbm_cg_time_steps.argtypes = [vp, C.c_int64, dp]
    L.lbm_batch_create.argtypes = [P(BatchInstance), C.c_int32, C.c_int32, C.c_int32, P(vp)]
    L.lbm_batch_destroy.argtypes = [vp]
    L.lbm_batch_destroy.restype = None
    L.lbm_batch_last_error.argtypes = [vp]
    L.lbm_batch_last_error.restype = C.c_char_p
    L.lbm_batch_init.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.lbm_batch_run.argtypes = [vp, C.c_double, C.c_int32, C.c_int64, P(C.c_int64), dp, dp]
    L.lbm_batch_get_macros.argtypes = [vp, C.c_int32, dp, dp, dp]
    L.lbm_batch_get_f.argtypes = [vp, C.c_int32, dp]
    L.lbm_batch_nodes.argtypes = [vp, P(C.c_int64)]
    L.lbm_batch_nodes.restype = C.c_int64
    for name in EXPORTED:
        fn = getattr(L, name)
        if fn.restype is C.c_int and name not in ("lbm_abi_version", "lbm_device_count", "lbm_batch_nodes"):
            fn.restype = C.c_int
    if L.lbm_abi_version() != ABI_VERSION:
        raise LbmError(3, "liblbm.so ABI version mismatch")
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _f64(a, n=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def device_count():
    return load().lbm_device_count()


def read_wall_q(node_path, bc_path):
    """Wall fractions (nz, ny, nx, 19) of the WALL links of a latticegen2 pair
    (S/GpuInit.cu:1088, :1205); 0 where there is no WALL link."""
    L = load()
    r = read_lattice(node_path, None)
    nx, ny, nz = r["dims"]
    q = np.zeros(nx * ny * nz * 19)
    st = L.lbm_read_wall_q(node_path.encode(), bc_path.encode() if bc_path else None, _dp(q))
    if st:
        raise LbmError(st, f"cannot read {node_path} / {bc_path}")
    return q.reshape(nz, ny, nx, 19)


def read_lattice(node_path, bc_path=None):
    """Read a latticegen2 pair (S/FilesReading.cu:270-347).

    Returns dict(dims=(nx, ny, nz), node_class, bb_links, nlinks, type_counts, dir_counts)
    with arrays shaped (nz, ny, nx)."""
    L = load()
    dims = (C.c_int32 * 3)()
    nl = C.c_int64()
    nb = bc_path.encode() if bc_path else None
    st = L.lbm_read_lattice(node_path.encode(), nb, dims, C.byref(nl), None, None, None, None, None, None)
    if st:
        raise LbmError(st, f"cannot read {node_path} / {bc_path}")
    nx, ny, nz = dims[:]
    n = nx * ny * nz
    cls = np.zeros(n, np.uint8)
    bb = np.zeros(n, np.uint32)
    tc = np.zeros(6, np.int64)
    dc = np.zeros(19, np.int64)
    delta = C.c_double()
    cyc = (C.c_int32 * 3)()
    st = L.lbm_read_lattice(node_path.encode(), nb, dims, C.byref(nl),
                            cls.ctypes.data_as(C.POINTER(C.c_uint8)), bb.ctypes.data_as(C.POINTER(C.c_uint32)),
                            tc.ctypes.data_as(C.POINTER(C.c_int64)), dc.ctypes.data_as(C.POINTER(C.c_int64)),
                            C.byref(delta), cyc)
    if st:
        raise LbmError(st, f"cannot read {node_path} / {bc_path}")
    return dict(dims=(nx, ny, nz), node_class=cls.reshape(nz, ny, nx), bb_links=bb.reshape(nz, ny, nx),
                nlinks=nl.value, type_counts=tc, dir_counts=dc, delta=delta.value, cyclic=tuple(cyc[:]))


def latticegen(vrt_path, bnd_path, nx, seed, out_prefix, complexity=1, mesh_path=None):
    """Generate <out_prefix>_Node.dat / _BC.dat from a STAR-CD geometry, as the
    reference's latticegen2 does (L/main.cpp:17-70).  Returns (dims, nlinks)."""
    L = load()
    dims = (C.c_int32 * 3)()
    nl = C.c_int64()
    sd = np.ascontiguousarray(seed, dtype=np.float64)
    st = L.lbm_latticegen(vrt_path.encode(), bnd_path.encode(), int(nx), _dp(sd), int(complexity),
                          (out_prefix + "_Node.dat").encode(), (out_prefix + "_BC.dat").encode(),
                          mesh_path.encode() if mesh_path else None, dims, C.byref(nl))
    if st:
        raise LbmError(st, L.lbm_latticegen_last_error().decode())
    return tuple(dims[:]), nl.value


def write_lattice_binary(node_path, bc_path, out_path):
    """Binary copy of a latticegen2 pair, readable by read_lattice(out_path)."""
    st = load().lbm_write_lattice_binary(node_path.encode(), bc_path.encode() if bc_path else None, out_path.encode())
    if st:
        raise LbmError(st, f"cannot convert {node_path} / {bc_path}")


class Solver:
    """One lattice on the GPU.  Arrays are (Q, nz, ny, nx) / (nz, ny, nx) in
    3D and (Q, ny, nx) / (ny, nx) in 2D, x fastest as in the reference
    (ind = i + j*n + k*n*m)."""

    def __init__(self, q, shape, node_class, bb=None, *, collision=BGK, omega=1.0, trt_magic=0.0,
                 mrt_s=None, force_model=FORCE_NONE, g=(0.0, 0.0, 0.0), wall_model=WALL_HALFWAY,
                 inlet_model=INLET_ZOUHE, parity=0, rho_out=1.0, periodic=(0, 0, 0), ubc=None,
                 vbc=None, wbc=None, storage=STORE_F64, outlet_model=OUTLET_PRESSURE, wall_q=None):
        L = load()
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        self.two_d = len(shape) == 2
        nz, ny, nx = (1,) + tuple(shape) if self.two_d else tuple(shape)
        self.q, self.shape3, self.n = q, (nz, ny, nx), nx * ny * nz
        self.shape = tuple(shape)
        p = Params()
        p.abi_version, p.q, p.nx, p.ny, p.nz = ABI_VERSION, q, nx, ny, nz
        for i in range(3):
            p.periodic[i] = int(periodic[i]) if i < len(periodic) else 0
            p.force[i] = float(g[i]) if i < len(g) else 0.0
        p.collision, p.omega, p.trt_magic = collision, float(omega), float(trt_magic)
        if mrt_s is not None:
            for i, s in enumerate(mrt_s):
                p.mrt_s[i] = s
        p.force_model, p.storage, p.wall_model, p.inlet_model = force_model, storage, wall_model, inlet_model
        p.parity_flags, p.rho_out, p.outlet_model = parity, float(rho_out), outlet_model
        self.params = p
        cls = np.ascontiguousarray(node_class, dtype=np.uint8).ravel()
        if cls.size != self.n:
            raise ValueError("node_class size does not match the lattice shape")
        bbv = None if bb is None else np.ascontiguousarray(bb, dtype=np.uint32).ravel()
        if bbv is not None and bbv.size != self.n:
            raise ValueError("bb size does not match the lattice shape")
        self._keep = [cls, bbv, _f64(ubc, self.n), _f64(vbc, self.n), _f64(wbc, self.n), _f64(wall_q, 19 * self.n)]
        lat = Lattice()
        lat.node_class = cls.ctypes.data_as(C.POINTER(C.c_uint8))
        lat.bb_links = bbv.ctypes.data_as(C.POINTER(C.c_uint32)) if bbv is not None else None
        lat.u_bc, lat.v_bc, lat.w_bc, lat.wall_q = (_dp(a) for a in self._keep[2:])
        h = C.c_void_p()
        st = L.lbm_create(C.byref(p), C.byref(lat), C.byref(h))
        if st:
            raise LbmError(st, "lbm_create failed (see stderr)")
        self.h = h
        self._L = L

    def _check(self, st):
        if st:
            raise LbmError(st, self._L.lbm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._L.lbm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, mode=0, rho=None, u=None, v=None, w=None):
        arrs = [_f64(a, self.n) for a in (rho, u, v, w)]
        self._check(self._L.lbm_init(self.h, mode, *[_dp(a) for a in arrs]))

    def step(self, n=1):
        self._check(self._L.lbm_step(self.h, int(n)))

    def run_to_residual(self, tol, every=100, max_steps=10 ** 6):
        done = C.c_int64()
        res = C.c_double()
        self._check(self._L.lbm_run_to_residual(self.h, tol, every, max_steps, C.byref(done), C.byref(res)))
        return done.value, res.value

    def reduce(self):
        out = np.zeros(4)
        self._check(self._L.lbm_reduce(self.h, _dp(out)))
        return dict(du2=out[0], u2=out[1], mass=out[2], nonfinite=out[3])

    def reduce_all(self):
        """lbm_reduce plus the largest |du|, |dv|, |dw|, |drho| since the last reduction."""
        out = np.zeros(8)
        self._check(self._L.lbm_reduce_all(self.h, _dp(out)))
        return dict(du2=out[0], u2=out[1], mass=out[2], nonfinite=out[3], max_du=out[4], max_dv=out[5],
                    max_dw=out[6], max_drho=out[7])

    def residual_l2(self):
        """The reference's L2 residual sqrt(sum (f - fColl)^2) of the last step."""
        out = C.c_double()
        self._check(self._L.lbm_residual_l2(self.h, C.byref(out)))
        return out.value

    def residual_l2_noalloc(self):
        """The same L2 residual with no device allocation (populations rebuilt
        in registers, partial sums in the library's static scratch)."""
        out = C.c_double()
        self._check(self._L.lbm_residual_l2_noalloc(self.h, C.byref(out)))
        return out.value

    def residual_fdrel(self):
        """The reference's FdRelDiff residual max |f - f_prev| / |f1| (f_prev:
        the previous call's populations, f1: the first call's; the first call
        returns 1)."""
        out = C.c_double()
        self._check(self._L.lbm_residual_fdrel(self.h, C.byref(out)))
        return out.value

    def macros(self):
        out = [np.zeros(self.n) for _ in range(4)]
        self._check(self._L.lbm_get_macros(self.h, *[_dp(a) for a in out]))
        sh = self.shape if self.two_d else self.shape3
        return [a.reshape(sh) for a in out]

    def get_f(self):
        f = np.zeros(self.q * self.n)
        self._check(self._L.lbm_get_f(self.h, _dp(f)))
        return f.reshape((self.q,) + (self.shape if self.two_d else self.shape3))

    def set_f(self, f):
        f = _f64(f, self.q * self.n)
        self._check(self._L.lbm_set_f(self.h, _dp(f)))

    def time_steps(self, n):
        ms = C.c_double()
        self._check(self._L.lbm_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value

    def time_launches(self, n):
        """Device time of each of n steps (ms), one event pair per launch."""
        out = np.zeros(int(n))
        self._check(self._L.lbm_time_launches(self.h, int(n), _dp(out)))
        return out

    def device_bytes(self):
        return self._L.lbm_device_bytes(self.h)


class CgSolver:
    """Colour-gradient two-phase step (GPU): D2Q9 for CgSolver(nx, ny) with
    arrays (ny, nx), D3Q19 for CgSolver(nx, ny, nz=...) with arrays
    (nz, ny, nx); periodic, or walls on the y rows / planes (walls=1 Zou-He
    velocity walls, the north one at u_wall; walls=2 bounce-back) and a body
    force g.  Same keyword arguments as the oracle's CgSim."""

    def __init__(self, nx, ny, *, nz=1, r_alpha=4 / 9, b_alpha=4 / 9, r_nu=1 / 6, b_nu=1 / 6, beta=0.7, A=0.01125,
                 g_limit=0.0, high_order=0, kappa=0.9, storage=STORE_F64, enhanced=0, walls=0, u_wall=0.0,
                 g=0.0, external_force=0):
        L = load()
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        p = CgParams()
        p.abi_version, p.nx, p.ny, p.nz, p.enhanced = ABI_VERSION, nx, ny, nz, enhanced
        p.r_alpha, p.b_alpha, p.r_nu, p.b_nu, p.beta, p.A = r_alpha, b_alpha, r_nu, b_nu, beta, A
        p.g_limit, p.high_order, p.kappa, p.storage = g_limit, high_order, kappa, storage
        p.walls, p.u_wall, p.g, p.external_force = walls, u_wall, g, external_force
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * max(nz, 1)
        self.q = 19 if nz > 1 else 9
        self.shape = (nz, ny, nx) if nz > 1 else (ny, nx)
        h = C.c_void_p()
        st = L.lbm_cg_create(C.byref(p), C.byref(h))
        if st:
            raise LbmError(st, "lbm_cg_create failed")
        self.h, self._L = h, L

    def _check(self, st):
        if st:
            raise LbmError(st, self._L.lbm_cg_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._L.lbm_cg_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, r_rho, b_rho):
        r, b = _f64(r_rho, self.n), _f64(b_rho, self.n)
        self._check(self._L.lbm_cg_init(self.h, _dp(r), _dp(b)))

    def init_f(self, r_f, b_f):
        """Start from colour populations (q, *shape)."""
        r, b = _f64(r_f, self.q * self.n), _f64(b_f, self.q * self.n)
        self._check(self._L.lbm_cg_init_f(self.h, _dp(r), _dp(b)))

    def step(self, n=1):
        self._check(self._L.lbm_cg_step(self.h, int(n)))

    def macros(self):
        out = [np.zeros(self.n) for _ in range(6)]
        self._check(self._L.lbm_cg_get_macros3d(self.h, *[_dp(a) for a in out]))
        keys = ["r_rho", "b_rho", "rho", "u", "v", "w"]
        d = dict(zip(keys, [a.reshape(self.shape) for a in out]))
        if self.q == 9:
            del d["w"]
        return d

    def get_f(self):
        r, b = np.zeros(self.q * self.n), np.zeros(self.q * self.n)
        self._check(self._L.lbm_cg_get_f(self.h, _dp(r), _dp(b)))
        return r.reshape((self.q,) + self.shape), b.reshape((self.q,) + self.shape)

    def reduce(self):
        out = np.zeros(12)
        self._check(self._L.lbm_cg_reduce(self.h, _dp(out)))
        keys = ["mass_r", "mass_b", "pin_sum", "pin_n", "pout_sum", "pout_n", "rho_in_sum", "rho_in_n",
                "rho_out_sum", "rho_out_n", "umax", "nonfinite"]
        return dict(zip(keys, out))

    def time_steps(self, n):
        ms = C.c_double()
        self._check(self._L.lbm_cg_time_steps(self.h, int(n), C.byref(ms)))
        return ms.value


class Batch:
    """Many D2Q9 channels advanced together in one persistent launch (GPU).

    `instances` is a list of dicts with node_class (ny, nx), bb (ny, nx)
    wall-link masks, ubc / vbc inlet velocities, omega, and optionally
    trt_magic and rho_out (the keys lattices.channel2d returns, plus omega)."""

    def __init__(self, instances, collision=TRT, storage=STORE_F64):
        L = load()
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        n = len(instances)
        arr = (BatchInstance * n)()
        self._keep = []
        self.shapes = []
        for k, d in enumerate(instances):
            cls = np.ascontiguousarray(d["node_class"], dtype=np.uint8)
            ny, nx = cls.shape
            bb = np.ascontiguousarray(d["bb"], dtype=np.uint32) if d.get("bb") is not None else None
            ubc = _f64(d.get("ubc"), nx * ny)
            vbc = _f64(d.get("vbc"), nx * ny)
            self._keep += [cls, bb, ubc, vbc]
            a = arr[k]
            a.nx, a.ny = nx, ny
            a.node_class = cls.ctypes.data_as(C.POINTER(C.c_uint8))
            a.bb_links = bb.ctypes.data_as(C.POINTER(C.c_uint32)) if bb is not None else None
            a.u_bc, a.v_bc = _dp(ubc), _dp(vbc)
            a.omega = float(d["omega"])
            a.trt_magic = float(d.get("trt_magic", 0.0))
            a.rho_out = float(d.get("rho_out", 1.0))
            self.shapes.append((ny, nx))
        h = C.c_void_p()
        st = L.lbm_batch_create(arr, n, collision, storage, C.byref(h))
        if st:
            raise LbmError(st, L.lbm_batch_last_error(None).decode())
        self.h, self._L, self.n = h, L, n
        self._keep = []

    def _check(self, st):
        if st:
            raise LbmError(st, self._L.lbm_batch_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self._L.lbm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def init(self, i=-1, rho=None, u=None, v=None):
        """Equilibrium start of instance i (-1: all, at rest)."""
        m = None if i < 0 else self.shapes[i][0] * self.shapes[i][1]
        r, uu, vv = _f64(rho, m), _f64(u, m), _f64(v, m)
        self._check(self._L.lbm_batch_init(self.h, int(i), _dp(r), _dp(uu), _dp(vv)))

    def run(self, tol, every=100, max_steps=100000):
        """One persistent launch; returns (steps, residuals, device ms)."""
        steps = np.zeros(self.n, np.int64)
        res = np.zeros(self.n)
        ms = C.c_double()
        st = self._L.lbm_batch_run(self.h, float(tol), int(every), int(max_steps),
                                   steps.ctypes.data_as(C.POINTER(C.c_int64)), _dp(res), C.byref(ms))
        self._check(st)
        return steps, res, ms.value

    def macros(self, i):
        ny, nx = self.shapes[i]
        out = [np.zeros(nx * ny) for _ in range(3)]
        self._check(self._L.lbm_batch_get_macros(self.h, int(i), *[_dp(a) for a in out]))
        return [a.reshape(ny, nx) for a in out]

    def get_f(self, i):
        ny, nx = self.shapes[i]
        f = np.zeros(9 * nx * ny)
        self._check(self._L.lbm_batch_get_f(self.h, int(i), _dp(f)))
        return f.reshape(9, ny, nx)

    def nodes(self):
        fl = C.c_int64()
        tot = self._L.lbm_batch_nodes(self.h, C.byref(fl))
        return tot, fl.value
