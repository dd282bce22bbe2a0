"""Drop-in solver object for the 3-D river reach, running on the GPU through
libwfs.so.

``LBM3D`` keeps the constructor and attribute surface of the reference's
solver class (floodsim/lbm3d.py:333-418):

    LBM3D(be, solid, Fx, Fy, Fz, uin, tau0, cs=0.17, cf_wall=None)

solid: bool (nz, ny, nx); Fx/Fy/Fz: body force per unit volume in lattice
units (nz, ny, nx); uin: inlet velocity in lattice units (nz, ny); cf_wall:
wall friction coefficient of wall-adjacent cells (tau = rho*cf*U^2) or None.
``rho ux uy uz`` are flat float32 arrays of the last step; ``Fx Fy Fz uin
uref`` may be replaced by float32 arrays of the right size, checked at the
next ``step`` with ``TypeError`` like the reference (lbm3d.py:392-394).

Differences by design: the state lives in GPU memory, so the field
attributes return host copies instead of live device views. One fetch
brings all four fields of the current step; later reads of any of them
before the next step reuse it, so a caller that reads ``ux``, ``uy``,
``uz`` and ``rho`` after a chunk (reach3d.py:412-418, 441-450) pays one
device-to-host copy of the fields, not six. The returned arrays are
read-only. ``xp`` is numpy: the arrays a caller gets and gives are host
arrays (the reference's ``xp`` is the array module of its backend,
lbm3d.py:348-351). ``be`` is accepted for signature compatibility and only a
torch backend is rejected, as in the reference (lbm3d.py:349-350).
Keyword-only extensions choose the population storage, the work-group width,
the device, the layout, in-place streaming and the gap between population
direction arrays (``pop_stagger``, bytes: a layout choice with no effect on
results, tuned by autotune.pop_stagger_for), and give the device-side cadence helpers
used by ``Run3D``.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


class LBM3D:
    engine = "hip-lbm"
    xp = np

    def __init__(self, be, solid, Fx, Fy, Fz, uin, tau0, cs=0.17, cf_wall=None, *, storage="f32", block_x=0,
                 device=0, layout="dense", streaming="ab", pop_stagger=0):
        if be is not None and getattr(be, "name", None) == "torch":
            raise RuntimeError("LBM3D: the 3-D model runs on the HIP library, not on a torch backend")
        if storage not in nat.STORAGE:
            raise ValueError(f"LBM3D: storage must be one of {sorted(nat.STORAGE)}")
        solid = np.asarray(solid)
        if solid.ndim != 3:
            raise ValueError("LBM3D: solid must be a (nz, ny, nx) array")
        self.be = be
        self.nz, self.ny, self.nx = solid.shape
        self.N = int(solid.size)
        self.tau0 = float(tau0)
        self.cs = float(cs)
        self.csmag2 = float(cs * cs)
        self.storage = storage
        self.nstep = 0
        self.closure = "smagorinsky"
        if streaming not in ("ab", "inplace"):
            raise ValueError("LBM3D: streaming must be 'ab' or 'inplace'")
        self.inplace = False
        if layout not in ("dense", "tiles"):
            raise ValueError("LBM3D: layout must be 'dense' or 'tiles'")
        self.layout = layout
        L = nat.lib()
        desc = nat.wfs_desc(self.nx, self.ny, self.nz, self.tau0, self.cs, nat.STORAGE[storage], int(block_x),
                            int(device))
        h = ctypes.c_void_p()
        sol = np.ascontiguousarray(solid.ravel(), dtype=np.uint8)
        if layout == "tiles":
            nat.check(L.wfs_create_tiled(ctypes.byref(desc), nat.ptr(sol), ctypes.byref(h)), "wfs_create_tiled")
        else:
            nat.check(L.wfs_create(ctypes.byref(desc), ctypes.byref(h)), "wfs_create")
        self._h = h
        # own copies, as the reference's device backend takes them (lbm3d.py:352-355): a caller
        # scaling L.Fx in place must not change its own arrays
        A = lambda a: np.array(a, dtype=np.float32, copy=True).ravel()  # noqa: E731
        cfw = A(cf_wall) if cf_wall is not None else None
        if cfw is not None and cfw.size != self.N:
            raise ValueError("LBM3D: cf_wall must have the shape of solid")
        self._uin = A(uin)
        if self._uin.size != self.ny * self.nz:
            raise ValueError("LBM3D: uin must be (nz, ny)")
        if pop_stagger:               # bytes between population direction arrays (layout only)
            nat.check(L.wfs_set_pop_stagger(h, int(pop_stagger)), "wfs_set_pop_stagger")
        self.pop_stagger = int(pop_stagger)
        if streaming == "inplace":    # before the upload: the second buffer is never allocated
            self.set_inplace(True)
        nat.check(L.wfs_upload(h, nat.ptr(sol), nat.ptr(cfw), nat.ptr(self._uin)), "wfs_upload")
        if Fx is None:        # forces come later as a cell list (set_force_cells)
            self._Fx = self._Fy = self._Fz = None
        else:
            self._Fx, self._Fy, self._Fz = A(Fx), A(Fy), A(Fz)
            if not (self._Fx.size == self._Fy.size == self._Fz.size == self.N):
                raise ValueError("LBM3D: forces must have the shape of solid")
            nat.check(L.wfs_set_force(h, nat.ptr(self._Fx), nat.ptr(self._Fy), nat.ptr(self._Fz)),
                      "wfs_set_force")
        self._uref = None          # host value assigned by a caller, uploaded at the next step
        self._dirty = set()
        self._gen = 0              # bumped by every call that may change the macroscopic fields
        self._macro_cache = None   # ((nstep, gen), (rho, ux, uy, uz)) of the last fetch

    # ------------------------------------------------------------ lifetime
    def close(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            nat.lib().wfs_destroy(h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("LBM3D: solver was closed")
        return self._h

    # ------------------------------------------------------------ replaceable inputs
    def _set_input(self, name, value):
        setattr(self, "_" + name, value)
        self._dirty.add(name)

    Fx = property(lambda s: s._Fx, lambda s, v: s._set_input("Fx", v))
    Fy = property(lambda s: s._Fy, lambda s, v: s._set_input("Fy", v))
    Fz = property(lambda s: s._Fz, lambda s, v: s._set_input("Fz", v))
    uin = property(lambda s: s._uin, lambda s, v: s._set_input("uin", v))

    @property
    def uref(self):
        out = np.empty(self.N, np.float32)
        nat.check(nat.lib().wfs_get_uref(self.handle, nat.ptr(out)))
        return out

    @uref.setter
    def uref(self, v):
        self._set_input("uref", v)

    def set_force(self, Fx, Fy, Fz):
        """Replace the body forces (lbm3d.py:386-388)."""
        A = lambda a: np.ascontiguousarray(np.asarray(a).ravel(), dtype=np.float32)  # noqa: E731
        self._Fx, self._Fy, self._Fz = A(Fx), A(Fy), A(Fz)
        self._dirty.update(("Fx", "Fy", "Fz"))
        self._flush()

    def set_force_cells(self, cells, fx, fy, fz):
        """Body forces (lattice units) given on a list of dense cell indices; every other cell gets none."""
        idx = np.ascontiguousarray(cells, dtype=np.int64)
        v = [np.ascontiguousarray(a, dtype=np.float32) for a in (fx, fy, fz)]
        if not all(a.size == idx.size for a in v):
            raise ValueError("set_force_cells: one value per cell and component")
        nat.check(nat.lib().wfs_set_force_cells(self.handle, idx.size, nat.ptr(idx), *(nat.ptr(a) for a in v)))

    @property
    def stored_cells(self):
        """Cells held on the device (padded rows, or the stored tiles x 64)."""
        return int(nat.lib().wfs_stored_cells(self.handle))

    def _flush(self):
        """Validate and upload replaced inputs (the checks of lbm3d.py:392-394)."""
        if not self._dirty:
            return
        for name in ("Fx", "Fy", "Fz", "uin", "uref"):
            if name not in self._dirty:
                continue
            a = getattr(self, "_" + name)
            want = self.ny * self.nz if name == "uin" else self.N
            if not isinstance(a, np.ndarray) or a.dtype != np.float32 or a.size != want:
                raise TypeError(f"LBM3D: force/inlet arrays must be float32 of the correct size "
                                f"(got {getattr(a, 'dtype', type(a))}, {getattr(a, 'size', '?')})")
        L, h = nat.lib(), self.handle
        c = lambda a: np.ascontiguousarray(a.ravel())  # noqa: E731
        if self._dirty & {"Fx", "Fy", "Fz"}:
            self._Fx, self._Fy, self._Fz = c(self._Fx), c(self._Fy), c(self._Fz)
            nat.check(L.wfs_set_force(h, nat.ptr(self._Fx), nat.ptr(self._Fy), nat.ptr(self._Fz)))
        if "uin" in self._dirty:
            self._uin = c(self._uin)
            nat.check(L.wfs_set_uin(h, nat.ptr(self._uin)))
        if "uref" in self._dirty:
            u = c(self._uref)
            nat.check(L.wfs_set_uref(h, nat.ptr(u)))
            self._uref = None
        self._dirty.clear()

    # ------------------------------------------------------------ stepping
    def step(self, n=1):
        """Advance n steps (lbm3d.py:390-412). Asynchronous on the GPU."""
        self._flush()
        self._gen += 1
        nat.check(nat.lib().wfs_step(self.handle, int(n)), "wfs_step")
        self.nstep += int(n)

    def synchronize(self):
        nat.check(nat.lib().wfs_synchronize(self.handle))

    # ------------------------------------------------------------ outputs
    def macro(self, dtype=np.float32):
        """(rho, ux, uy, uz) of the last step as flat host arrays."""
        out = [np.empty(self.N, dtype) for _ in range(4)]
        fn = nat.lib().wfs_get_macro_f64 if dtype == np.float64 else nat.lib().wfs_get_macro
        nat.check(fn(self.handle, *(nat.ptr(a) for a in out)))
        return tuple(out)

    def _fields(self):
        """(rho, ux, uy, uz) of the current step, fetched once per step and
        shared by the four field attributes (read-only arrays)."""
        key = (int(nat.lib().wfs_nstep(self.handle)), self._gen)
        if self._macro_cache is None or self._macro_cache[0] != key:
            arrs = self.macro()
            for a in arrs:
                a.flags.writeable = False
            self._macro_cache = (key, arrs)
        return self._macro_cache[1]

    rho = property(lambda s: s._fields()[0])
    ux = property(lambda s: s._fields()[1])
    uy = property(lambda s: s._fields()[2])
    uz = property(lambda s: s._fields()[3])

    def set_macro(self, rho, ux, uy, uz):
        """Overwrite the macroscopic fields the next step's wall model reads."""
        A = lambda a: np.ascontiguousarray(np.asarray(a, np.float32).ravel())  # noqa: E731
        arrs = [A(a) for a in (rho, ux, uy, uz)]
        if any(a.size != self.N for a in arrs):
            raise ValueError("LBM3D: macroscopic fields must hold N values")
        self._gen += 1
        nat.check(nat.lib().wfs_set_macro(self.handle, *(nat.ptr(a) for a in arrs)))

    def populations(self, dtype=np.float32):
        """Decoded populations f[i*N + cell] (19*N)."""
        out = np.empty(19 * self.N, dtype)
        fn = nat.lib().wfs_get_populations_f64 if dtype == np.float64 else nat.lib().wfs_get_populations
        nat.check(fn(self.handle, nat.ptr(out)))
        return out

    def set_populations(self, f):
        f = np.ascontiguousarray(np.asarray(f, np.float32).ravel())
        if f.size != 19 * self.N:
            raise ValueError("LBM3D: populations must hold 19*N values")
        nat.check(nat.lib().wfs_set_populations(self.handle, nat.ptr(f)))

    # ------------------------------------------------------------ device-side cadence helpers
    def set_force_scale(self, s):
        """Forces used by the step become F * s (FP32), e.g. a thrust ramp."""
        self._flush()
        nat.check(nat.lib().wfs_set_force_scale(self.handle, float(np.float32(s))))

    def set_bedwall(self, bedwall):
        bw = np.ascontiguousarray(np.asarray(bedwall).ravel(), dtype=np.uint8)
        if bw.size != self.N:
            raise ValueError("LBM3D: bedwall must have the shape of solid")
        nat.check(nat.lib().wfs_set_bedwall(self.handle, nat.ptr(bw)))

    def update_uref(self):
        """Depth-averaged column speed into uref on bed/bank cells (reach3d.py:412-418), on the GPU."""
        self._flush()
        nat.check(nat.lib().wfs_update_uref(self.handle))

    def accumulate(self):
        nat.check(nat.lib().wfs_accumulate(self.handle))

    def reset_sums(self):
        nat.check(nat.lib().wfs_reset_sums(self.handle))

    def sums(self, dtype=np.float32):
        """The six running sums (ux, uy, uz, rho, sp, uxx) and the sample count."""
        names = ("ux", "uy", "uz", "rho", "sp", "uxx")
        out = {k: np.empty(self.N, dtype) for k in names}
        n = ctypes.c_int32(0)
        fn = nat.lib().wfs_get_sums_f64 if dtype == np.float64 else nat.lib().wfs_get_sums
        nat.check(fn(self.handle, *(nat.ptr(out[k]) for k in names), ctypes.byref(n)))
        return out, int(n.value)

    def profile_x(self, field, z0=0, z1=None):
        """FP64 sum over fluid cells with z0 <= z < z1 of a field, per x."""
        out = np.empty(self.nx, np.float64)
        z1 = self.nz if z1 is None else z1
        nat.check(nat.lib().wfs_profile_x(self.handle, nat.FIELD[field], int(z0), int(z1), nat.ptr(out)))
        return out

    def total(self, field):
        out = ctypes.c_double(0.0)
        nat.check(nat.lib().wfs_total(self.handle, nat.FIELD[field], ctypes.byref(out)))
        return float(out.value)

    def check_finite(self):
        """Raises NonFiniteState if any fluid cell holds NaN/Inf."""
        nat.check(nat.lib().wfs_check_finite(self.handle))

    def timing(self, on=True):
        """Bracket each step-kernel launch with HIP events on the solver's stream."""
        nat.check(nat.lib().wfs_timing_enable(self.handle, 1 if on else 0))

    def timing_read(self):
        """(summed step-kernel milliseconds, launches) since the last read."""
        ms, n = ctypes.c_double(0.0), ctypes.c_int64(0)
        nat.check(nat.lib().wfs_timing_read(self.handle, ctypes.byref(ms), ctypes.byref(n)))
        return float(ms.value), int(n.value)

    def set_column_u(self, per_step=True):
        """per_step: the wall model's column speed is refreshed every step inside
        the step kernel from the previous step's velocity (update_uref and the
        uref attribute then do not reach the step); False: back to uref as the
        run cadence sets it. Starts from the current velocity; needs
        set_bedwall."""
        nat.check(nat.lib().wfs_set_column_u(self.handle, int(bool(per_step))))

    def set_inplace(self, on=True):
        """on: stream in place in one population buffer (the second is freed;
        two steps equal two two-buffer steps bit for bit); False: back to two
        buffers. populations, set_populations, forces and switching off need
        an even number of steps since switching on. nx - 1 must not be a
        multiple of 64.
        LBM3D(streaming="inplace") switches it on before the upload, so the
        second buffer is never allocated."""
        nat.check(nat.lib().wfs_set_inplace(self.handle, int(bool(on))))
        self.inplace = bool(on)

    # ------------------------------------------------------------ checkpoints
    def state(self):
        """The raw solver state (populations as stored, macroscopic fields,
        uref, time-mean sums, column buffers, counters) as a uint8 array."""
        n = ctypes.c_int64()
        nat.check(nat.lib().wfs_state_bytes(self.handle, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        nat.check(nat.lib().wfs_state_save(self.handle, nat.ptr(buf), n.value))
        return buf

    def load_state(self, buf):
        """Restore a state() of a solver built the same way; the run then
        continues bit for bit."""
        buf = np.ascontiguousarray(buf, dtype=np.uint8)
        self._gen += 1
        nat.check(nat.lib().wfs_state_load(self.handle, nat.ptr(buf), buf.size))
        self.nstep = int(nat.lib().wfs_nstep(self.handle))
        # uref comes with the state: a value assigned before the restore is dropped
        self._uref = None
        self._dirty.discard("uref")

    # ------------------------------------------------------------ particle tracer
    def particles_set(self, xyz, inlet_rows, seed=0):
        """Upload particle positions (n x 3, metres from the domain corner) and
        the rows (y + ny z) of the inlet cells they restart in."""
        p = np.ascontiguousarray(np.asarray(xyz, dtype=np.float64).reshape(-1, 3))
        rows = np.ascontiguousarray(np.asarray(inlet_rows, dtype=np.int32).ravel())
        self._np = len(p)
        nat.check(nat.lib().wfs_particles_set(self.handle, len(p), nat.ptr(p), rows.size, nat.ptr(rows),
                                             int(seed) & 0xFFFFFFFFFFFFFFFF))

    def particles_move(self, dt, vel, dx):
        """One tracer step on the device; returns how many particles restarted."""
        n = np.zeros(1, np.int64)
        nat.check(nat.lib().wfs_particles_move(self.handle, float(dt), float(vel), float(dx), nat.ptr(n)))
        return int(n[0])

    def particles_get(self):
        out = np.empty((getattr(self, "_np", 0), 3))
        if out.size:
            nat.check(nat.lib().wfs_particles_get(self.handle, nat.ptr(out)))
        return out

    # ------------------------------------------------------------ momentum budget
    def set_body_mask(self, body):
        """Mark the solid cells (e.g. hulls) whose force forces() reports apart."""
        b = np.ascontiguousarray(np.asarray(body).ravel(), dtype=np.uint8)
        if b.size != self.N:
            raise TypeError("set_body_mask: mask must have the shape of solid")
        nat.check(nat.lib().wfs_set_body_mask(self.handle, nat.ptr(b)))

    def forces(self):
        """FP64 momentum budget of the next step in lattice units: applied
        thrust, force on the marked body cells, force on the other walls
        (each a length-3 array)."""
        out = np.zeros(9)
        nat.check(nat.lib().wfs_forces(self.handle, nat.ptr(out)))
        return {"thrust": out[0:3].copy(), "body": out[3:6].copy(), "walls": out[6:9].copy()}

    # ------------------------------------------------------------ launch shape
    @property
    def block_x(self):
        """Work-group width of the step kernel (cells along x per work-group)."""
        return int(nat.lib().wfs_block_x(self.handle))

    @block_x.setter
    def block_x(self, v):
        nat.check(nat.lib().wfs_set_block_x(self.handle, int(v)))

    # ------------------------------------------------------------ sub-grid closure
    def set_closure(self, kind="smagorinsky", params=None, geo=None, hcol=None, u0=None):
        """Select the sub-grid closure: "smagorinsky" (the reference's constant
        Cs, floodsim/lbm3d.py:120-123) or "mlp", the learned per-cell C^2
        (``params``: 1377 float weights; ``geo``/``hcol``: the geometry codes of
        closure.geometry(solid); ``u0``: mean inlet speed, lattice units)."""
        if kind not in nat.CLOSURE:
            raise ValueError(f"unknown closure {kind!r}")
        if kind == "smagorinsky":
            nat.check(nat.lib().wfs_set_closure(self.handle, 0, None, None, None, 0.0))
            self.closure = kind
            return
        p = np.ascontiguousarray(params, dtype=np.float32)
        g = np.ascontiguousarray(geo, dtype=np.uint16)
        h = np.ascontiguousarray(hcol, dtype=np.uint8)
        if p.size != nat.MLP_PARAMS or g.size != self.N or h.size != self.nx * self.ny:
            raise TypeError("set_closure: params (1377), geo (nz, ny, nx) and hcol (ny, nx) sizes")
        nat.check(nat.lib().wfs_set_closure(self.handle, 1, nat.ptr(p), nat.ptr(g), nat.ptr(h), float(u0)))
        self.closure = kind

    def closure_field(self):
        """C^2 of every cell in the last step of the previous ``step`` call (learned closure)."""
        out = np.empty(self.N, np.float32)
        nat.check(nat.lib().wfs_get_closure_field(self.handle, nat.ptr(out)))
        return out

    def keep_closure_field(self, on=True):
        """Whether the learned step keeps the C^2 field closure_field reads
        (on by default; off frees 4 bytes per stored cell)."""
        nat.check(nat.lib().wfs_set_closure_field(self.handle, 1 if on else 0))

    # ------------------------------------------------------------ reference helpers
    def shape3(self, a):
        return np.asarray(a).reshape(self.nz, self.ny, self.nx)

    def memory_mb(self):
        """Device memory held by this solver (MB)."""
        return nat.lib().wfs_device_bytes(self.handle) / 1e6
