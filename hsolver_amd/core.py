"""Thin Python host over the C ABI: one class drives either libhsolver_amd (prefix hs_)
or any library exposing the same entry points under another prefix.

The method names follow the reference's solver structs (FluidSolver / MobSolver /
Solver: init, fluid_solve, solid_solve, exchange, set_for_next, step), so a host
loop reads like Solver::solve (solver.h:44-155).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class HSError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        super().__init__(f"{where} failed with code {code} {msg}")
        self.code = code


class Context:
    """Owns one solver context (device state for the GPU library)."""

    def __init__(self, lib: C.CDLL, prefix: str, grid: abi.GridArrays, params: abi.Params):
        self.lib = lib
        self.prefix = prefix
        self.grid = grid
        self.params = params
        self.nx, self.ny, self.nz = grid.nx, grid.ny, grid.nz
        self.h = C.c_void_p()
        self._call("create", C.byref(grid.struct), C.byref(params), C.byref(self.h), ctx=False)

    # -- plumbing -------------------------------------------------------------
    def _fn(self, name):
        return getattr(self.lib, f"{self.prefix}_{name}")

    def _call(self, name, *args, ctx=True):
        fn = self._fn(name)
        rc = fn(self.h, *args) if ctx else fn(*args)
        if rc != 0:
            msg = ""
            se = getattr(self.lib, f"{self.prefix}_strerror", None) or getattr(self.lib, "hs_strerror", None)
            if se is not None:
                se.restype = C.c_char_p
                msg = "(" + se(rc).decode() + ")"
            raise HSError(rc, f"{self.prefix}_{name}", msg)
        return rc

    def close(self):
        if self.h:
            self._fn("destroy")(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- reference solver methods ---------------------------------------------
    def init(self):
        self._call("init")

    def fluid_solve(self) -> dict:
        st = abi.Stats()
        self._call("fluid_solve", C.byref(st))
        return st.as_dict()

    def solid_solve(self) -> dict:
        st = abi.Stats()
        self._call("solid_solve", C.byref(st))
        return {"ts_sweeps": st.ts_sweeps, "ts_residual": st.ts_residual}

    def exchange(self):
        self._call("exchange")

    def set_for_next(self):
        self._call("set_for_next")

    def step(self) -> dict:
        st = abi.Stats()
        self._call("step", C.byref(st))
        return st.as_dict()

    def run(self, n: int) -> list:
        st = (abi.Stats * n)()
        self._call("run", n, st)
        return [s.as_dict() for s in st]

    def momentum(self, comp: int):
        s = C.c_int32()
        r = C.c_double()
        self._call("momentum", comp, C.byref(s), C.byref(r))
        return s.value, r.value

    def pressure_correction(self) -> dict:
        st = abi.Stats()
        self._call("pressure_correction", C.byref(st))
        return st.as_dict()

    def correct(self):
        self._call("correct")

    def fluid_energy(self):
        s = C.c_int32()
        r = C.c_double()
        self._call("fluid_energy", C.byref(s), C.byref(r))
        return s.value, r.value

    def cell_bc(self):
        self._call("cell_bc")

    def run_params(self) -> abi.Params:
        """The parameters the context runs with (by-grid choices filled in)."""
        p = abi.Params()
        self._call("get_params", C.byref(p))
        return p

    # -- data -----------------------------------------------------------------
    def size(self, which: int) -> int:
        fn = self._fn("array_size")
        fn.restype = C.c_int64
        return int(fn(self.h, which))

    def get(self, name: str) -> np.ndarray:
        which = abi.ARRAY_NAMES.index(name)
        out = np.empty(self.size(which), dtype=np.float64)
        self._call("get_array", which, abi.dptr(out), out.size)
        return out.reshape(abi.padded_shape(name, self.nx, self.ny, self.nz))

    def put(self, name: str, arr: np.ndarray):
        which = abi.ARRAY_NAMES.index(name)
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        if a.size != self.size(which):
            raise ValueError(f"{name}: expected {self.size(which)} values, got {a.size}")
        self._call("put_array", which, abi.dptr(a), a.size)

    def state(self, names=("u", "v", "w", "up", "vp", "wp", "p", "ptemp", "pprime", "tf", "ts")) -> dict:
        return {n: self.get(n) for n in names}

    def load_state(self, st: dict):
        for n, a in st.items():
            self.put(n, a)

    def fields(self) -> dict:
        nx, ny, nz = self.nx, self.ny, self.nz
        u = np.empty((nx + 1) * ny * nz)
        v = np.empty(nx * (ny + 1) * nz)
        w = np.empty(nx * ny * (nz + 1))
        p = np.empty(nx * ny * nz)
        t = np.empty(nx * ny * nz)
        self._call("get_fields", abi.dptr(u), abi.dptr(v), abi.dptr(w), abi.dptr(p), abi.dptr(t))
        return {"u": u.reshape(nz, ny, nx + 1), "v": v.reshape(nz, ny + 1, nx),
                "w": w.reshape(nz + 1, ny, nx), "p": p.reshape(nz, ny, nx), "t": t.reshape(nz, ny, nx)}

    def output(self) -> dict:
        """The reference's output records (Solver::solve solver.h:160-213): per cell, in
        global cell order, the minimum vertex, west/south/bottom face velocity, p, t and the
        p' row id of fluid cells (-1 on solid cells, whose velocity and p are 0)."""
        n = self.nx * self.ny * self.nz
        loc = np.empty(3 * n)
        vel = np.empty(3 * n)
        p = np.empty(n)
        t = np.empty(n)
        ids = np.empty(n, dtype=np.int32)
        self._call("get_output", abi.dptr(loc), abi.dptr(vel), abi.dptr(p), abi.dptr(t),
                   ids.ctypes.data_as(C.POINTER(C.c_int32)))
        return {"loc": loc.reshape(n, 3), "vel": vel.reshape(n, 3), "p": p, "t": t, "id": ids}

    # cell vectors (C layout) and padded cell vectors (P layout) handed to the C side,
    # which always reads or writes the full length: checked here before the call
    def _cells(self, arr: np.ndarray, name: str) -> np.ndarray:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        n = self.nx * self.ny * self.nz
        if a.size != n:
            raise ValueError(f"{name}: expected {n} cell values, got {a.size}")
        return a

    def _padded(self, arr: np.ndarray, name: str) -> np.ndarray:
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(-1)
        n = (self.nx + 2) * (self.ny + 2) * (self.nz + 2)
        if a.size != n:
            raise ValueError(f"{name}: expected {n} padded cell values, got {a.size}")
        return a

    def set_source(self, q: np.ndarray):
        a = self._cells(q, "q")
        self._call("set_source", abi.dptr(a))

    def set_conductivity(self, k: np.ndarray):
        a = self._cells(k, "k")
        self._call("set_conductivity", abi.dptr(a))

    # -- p' operator ------------------------------------------------------------
    def pressure_apply(self, x: np.ndarray) -> np.ndarray:
        xa = self._padded(x, "x")
        y = np.empty_like(xa)
        self._call("pressure_apply", abi.dptr(xa), abi.dptr(y))
        return y.reshape(abi.padded_shape("b", self.nx, self.ny, self.nz))

    def pressure_solve(self, b: np.ndarray):
        ba = self._padded(b, "b")
        x = np.empty_like(ba)
        st = abi.Stats()
        self._call("pressure_solve", abi.dptr(ba), abi.dptr(x), C.byref(st))
        return x.reshape(abi.padded_shape("b", self.nx, self.ny, self.nz)), st.as_dict()

    # -- energy equation on the frozen flow and its adjoint ----------------------
    # Cell vectors are (nz, ny, nx) float64 numpy arrays (host memory).
    def energy_solve(self, tol: float = 1e-12, maxit: int = 500) -> dict:
        info = abi.SolveInfo()
        self._call("energy_solve", C.c_double(tol), maxit, C.byref(info))
        return info.as_dict()

    def adjoint_energy(self, dJdT: np.ndarray, tol: float = 1e-12, maxit: int = 500):
        g = self._cells(dJdT, "dJdT")
        dq = np.empty_like(g)
        dk = np.empty_like(g)
        info = abi.SolveInfo()
        self._call("adjoint_energy", g.ctypes.data, dq.ctypes.data, dk.ctypes.data, C.c_double(tol), maxit,
                   abi.HS_MEM_HOST, C.byref(info))
        shp = (self.nz, self.ny, self.nx)
        return dq.reshape(shp), dk.reshape(shp), info.as_dict()

    def energy_solve_rhs(self, b: np.ndarray, transpose: bool = False, tol: float = 1e-12, maxit: int = 500):
        ba = self._cells(b, "b")
        x = np.empty_like(ba)
        info = abi.SolveInfo()
        self._call("energy_solve_rhs", 1 if transpose else 0, ba.ctypes.data, x.ctypes.data, C.c_double(tol), maxit,
                   abi.HS_MEM_HOST, C.byref(info))
        return x.reshape(self.nz, self.ny, self.nx), info.as_dict()

    def energy_apply(self, x: np.ndarray, transpose: bool = False) -> np.ndarray:
        xa = self._cells(x, "x")
        y = np.empty_like(xa)
        self._call("energy_apply", 1 if transpose else 0, xa.ctypes.data, y.ctypes.data, abi.HS_MEM_HOST)
        return y.reshape(self.nz, self.ny, self.nx)

    def energy_balance(self) -> tuple:
        """(heat input sum q V, |sum of energy residuals| / heat input) of the current field."""
        fn = self._fn("energy_balance")
        fn.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        h, r = C.c_double(), C.c_double()
        self._call("energy_balance", C.byref(h), C.byref(r))
        return h.value, r.value

    def amg_levels(self, which: int = 0) -> list:
        """(rows, off-diagonal entries) of each aggregated level: which = 0 the p' hierarchy,
        1 the energy hierarchy."""
        fn = self.lib.hs_amg_levels
        fn.argtypes = [C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
        buf = np.zeros(2 * 32, dtype=np.int64)
        n = fn(self.h, which, buf.ctypes.data, 32)
        if n < 0:
            raise HSError(n, "hs_amg_levels", "")
        return [(int(buf[2 * i]), int(buf[2 * i + 1])) for i in range(n)]

    def cells(self, name: str) -> np.ndarray:
        which = abi.ARRAY_NAMES.index(name)
        out = np.empty(self.nx * self.ny * self.nz)
        self._call("get_cells", which, out.ctypes.data, abi.HS_MEM_HOST)
        return out.reshape(self.nz, self.ny, self.nx)

    def put_cells(self, name: str, arr: np.ndarray):
        which = abi.ARRAY_NAMES.index(name)
        a = self._cells(arr, name)
        self._call("put_cells", which, a.ctypes.data, abi.HS_MEM_HOST)


_LIB = None


def library() -> C.CDLL:
    global _LIB
    if _LIB is None:
        _LIB = abi.load_library()
    return _LIB


def create(case_or_grid, params: abi.Params | None = None, q: np.ndarray | None = None) -> Context:
    """GPU context from a configs.Case (or a GridArrays + Params)."""
    if params is None:
        grid, params, q = case_or_grid.grid, case_or_grid.params, case_or_grid.q
    else:
        grid = case_or_grid
    ctx = Context(library(), "hs", grid, params)
    if q is not None and np.any(q):
        ctx.set_source(q)
    return ctx
