"""Host side of the device update: a Plan owns one ah_plan, its packed state
and its workspace as torch CUDA tensors, and moves states between the
reference layout (cells..., 8 conservative components) and the packed SoA
layout of the library (include/astrea_hip.h)."""
import ctypes

import numpy as np
import torch

from . import _lib as L
from . import options as O

BOUNDARY = {"edge": L.AH_EDGE, "outflow": L.AH_EDGE, "wrap": L.AH_WRAP, "periodic": L.AH_WRAP}


def infer_varmask(grids, dims, magnetic):
    """Smallest instantiated component set that holds every nonzero component
    of the initial states.  Components that are zero everywhere stay zero
    under the update for the classes named here (no transverse velocity or
    field source in 1D/2D hydro; v_z = B_z = 0 is preserved in 2D MHD)."""
    nz = np.zeros(8, bool)
    for g in grids:
        nz |= np.any(np.asarray(g).reshape(-1, 8) != 0, axis=0)
    if dims == 1:
        if not magnetic and not nz[[2, 3, 5, 6, 7]].any():
            return L.AH_VARS_H1
        return L.AH_VARS_ALL
    if not magnetic and not nz[[3, 5, 6, 7]].any():
        return L.AH_VARS_H2
    if not nz[[3, 7]].any():
        return L.AH_VARS_M6
    raise NotImplementedError("2D states with nonzero v_z or B_z have no device path")


def solver_varmask(mask, dims, solver):
    """Widen an inferred component set to one the solver is instantiated for
    (``solver`` an AH_* enum): HLLD and entropy-stable in 1D run on all eight
    components, HLLD in 2D on the six of 2D MHD (on hydrodynamic data its
    field components stay zero, as they do in the reference's arrays)."""
    if dims == 1 and solver in (L.AH_HLLD, L.AH_ES):
        return L.AH_VARS_ALL
    if dims == 2 and solver == L.AH_HLLD and mask == L.AH_VARS_H2:
        return L.AH_VARS_M6
    return mask


class Plan:
    def __init__(self, dims, subgrid, solver, timestep, problems, *, limiter="minmod", flattener="off",
                 precision="fp64", varmask=L.AH_VARS_H1, compat="astrea", magnetic_2d=False, device="cuda:0",
                 solver_category=None, av=False, ppm_author="mc", tile_s=0, fuse2d=False):
        self.lib = L.load()
        self.dims = dims
        self.device = torch.device(device)
        self._cap = None
        sch = L.AhScheme()
        sch.dims = dims
        sch.recon = O.recon_of(subgrid)
        sch.limiter = O.LIMITERS[limiter]
        sch.flattener = O.FLATTENERS[flattener]
        sch.solver = O.solver_of(solver, solver_category)
        sch.rk = O.rk_of(timestep)
        sch.precision = L.AH_SPLIT if precision == "split" else L.AH_FP64
        sch.varmask = varmask
        sch.magnetic_2d = 1 if magnetic_2d else 0
        sch.compat = L.AH_COMPAT_PHYSICAL if compat == "physical" else L.AH_COMPAT_ASTREA
        sch.av = 1 if av else 0
        sch.tile_s = int(tile_s)
        sch.fuse2d = 1 if fuse2d else 0
        sch.ppm_author = {"mc": L.AH_PPM_MC, "c": L.AH_PPM_C, "ph": L.AH_PPM_PH}[ppm_author]
        self.scheme = sch
        arr = (L.AhProblem * len(problems))()
        for i, p in enumerate(problems):
            arr[i].n0 = int(p["n0"])
            arr[i].n1 = int(p.get("n1", 1)) if dims == 2 else 1
            arr[i].boundary = BOUNDARY[p.get("boundary", "edge")]
            arr[i].checkpoints = int(p.get("checkpoints", 1))
            arr[i].dx = float(p["dx"])
            arr[i].gamma = float(p["gamma"])
            arr[i].cfl = float(p["cfl"])
            arr[i].t_end = float(p["t_end"])
            arr[i].max_steps = int(p.get("max_steps", -1))
        self.problems = arr
        self.nprob = len(problems)
        h = ctypes.c_void_p()
        L.check(self.lib.ah_plan_create(ctypes.byref(sch), self.nprob, arr, ctypes.byref(h)), "ah_plan_create")
        self.h = h
        total, nv = ctypes.c_longlong(), ctypes.c_int()
        vidx = (ctypes.c_int * 8)()
        offs = (ctypes.c_longlong * self.nprob)()
        L.check(self.lib.ah_plan_layout(h, ctypes.byref(total), ctypes.byref(nv), vidx, offs), "ah_plan_layout")
        self.total, self.nv = total.value, nv.value
        self.varidx = [vidx[i] for i in range(self.nv)]
        self.offsets = [offs[i] for i in range(self.nprob)]
        ws = ctypes.c_size_t()
        L.check(self.lib.ah_plan_workspace_size(h, ctypes.byref(ws)), "ah_plan_workspace_size")
        self.state = torch.zeros((self.nv, self.total), dtype=torch.float64, device=self.device)
        self.workspace = torch.empty(max(ws.value, 1), dtype=torch.uint8, device=self.device)
        L.check(self.lib.ah_plan_bind(h, ctypes.c_void_p(self.state.data_ptr()),
                                      ctypes.c_void_p(self.workspace.data_ptr()), ws.value), "ah_plan_bind")

    # ---- helpers -----------------------------------------------------------
    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def shape(self, i):
        p = self.problems[i]
        return (p.n0,) if self.dims == 1 else (p.n0, p.n1)

    def close(self):
        if getattr(self, "h", None):
            self.lib.ah_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state in/out --------------------------------------------------------
    def load(self, grids):
        """Pack conservative states given in the reference layout (numpy
        arrays, or CUDA tensors packed on the device)."""
        if all(isinstance(g, torch.Tensor) for g in grids):
            for i, g in enumerate(grids):
                assert tuple(g.shape) == self.shape(i) + (8,), (g.shape, self.shape(i))
                flat = g.reshape(-1, 8).to(self.device, torch.float64)
                o, n = self.offsets[i], flat.shape[0]
                dead = [c for c in range(8) if c not in self.varidx]
                if dead and bool((flat[:, dead] != 0).any()):
                    raise ValueError(f"problem {i} has nonzero components {dead} outside the plan's variable set")
                for v, c in enumerate(self.varidx):
                    self.state[v, o:o + n] = flat[:, c]
            return
        host = np.zeros((self.nv, self.total))
        for i, g in enumerate(grids):
            g = np.asarray(g, dtype=np.float64)
            assert g.shape == self.shape(i) + (8,), (g.shape, self.shape(i))
            flat = g.reshape(-1, 8)
            o = self.offsets[i]
            for v, c in enumerate(self.varidx):
                host[v, o:o + flat.shape[0]] = flat[:, c]
            dead = [c for c in range(8) if c not in self.varidx]
            if dead and np.any(flat[:, dead] != 0):
                raise ValueError(f"problem {i} has nonzero components {dead} outside the plan's variable set")
        self.state.copy_(torch.from_numpy(host).to(self.device))

    def unpack(self, i=None, state=None):
        """States of one or all problems in the reference layout."""
        host = (self.state if state is None else state).detach().cpu().numpy()
        out = []
        for k in range(self.nprob) if i is None else [i]:
            n = int(np.prod(self.shape(k)))
            g = np.zeros((n, 8))
            o = self.offsets[k]
            for v, c in enumerate(self.varidx):
                g[:, c] = host[v, o:o + n]
            out.append(g.reshape(self.shape(k) + (8,)))
        return out if i is None else out[0]

    # ---- time stepping -------------------------------------------------------
    def reset(self):
        L.check(self.lib.ah_plan_reset(self.h, self._stream()), "ah_plan_reset")

    def step(self, nsteps=1):
        L.check(self.lib.ah_plan_step(self.h, int(nsteps), self._stream()), "ah_plan_step")

    def step_graph(self, nsteps, nlaunch=1):
        """Replay a captured hipGraph of `nsteps` steps `nlaunch` times (the
        current torch stream must not be the default stream)."""
        L.check(self.lib.ah_plan_step_graph(self.h, int(nsteps), int(nlaunch), self._stream()), "ah_plan_step_graph")

    def time_export(self):
        arr = (L.AhTimeState * self.nprob)()
        L.check(self.lib.ah_plan_time_export(self.h, arr, self._stream()), "ah_plan_time_export")
        return [(a.t, a.dt, a.steps, a.idx, a.active, a.failed, a.parity) for a in arr]

    def time_import(self, states):
        arr = (L.AhTimeState * self.nprob)()
        for i, s in enumerate(states):
            arr[i] = L.AhTimeState(*s)
        L.check(self.lib.ah_plan_time_import(self.h, arr, self._stream()), "ah_plan_time_import")

    # ---- initial states built on the device (ah_plan_init_ic) ---------------
    def _high_order(self):
        return self.scheme.recon in (L.AH_PPM, L.AH_WENO3, L.AH_WENO5, L.AH_WENO7)

    def _init_ic(self, i, kind, regions, x0, y0, high_order):
        ic = L.AhIc()
        ic.kind = kind
        ic.high_order = int(self._high_order() if high_order is None else high_order)
        ic.x0, ic.y0 = float(x0), float(y0)
        for k, w in enumerate(regions):
            for c in range(8):
                ic.w[k][c] = float(w[c])
        L.check(self.lib.ah_plan_init_ic(self.h, int(i), ctypes.byref(ic), self._stream()), "ah_plan_init_ic")

    def init_riemann(self, i, left, right, x0=.5, high_order=None):
        """Problem i (1D) starts from the Riemann data left = (rho, u, p) for
        x <= x0 and right otherwise, built on the device (constructor.py:11-109);
        high_order defaults to the plan's reconstruction (PPM/WENO)."""
        reg = [[l[0], l[1], 0, 0, l[2], 0, 0, 0] for l in (left, right)]
        self._init_ic(i, L.AH_IC_RIEMANN_1D, reg, x0, 0., high_order)

    def init_quadrants(self, i, states, x0=.5, y0=.5, high_order=None):
        """Problem i (2D) starts from four-quadrant data, states[k] = (rho, u,
        v, p) for k = 0 (x <= x0, y < y0), 1 (x > x0, y < y0), 2 (x <= x0,
        y >= y0), 3 (x > x0, y >= y0), built on the device."""
        reg = [[s[0], s[1], s[2], 0, s[3], 0, 0, 0] for s in states]
        self._init_ic(i, L.AH_IC_QUADRANTS_2D, reg, x0, y0, high_order)

    def snapshot(self, i, out=None):
        """Problem i's primitive snapshot as the reference saves it each step
        (astrea.py:47): convert_conservative of the state, transposed by
        ortho_axis; a device tensor of shape (n0, 8) in 1D, (n1, n0, 8) in 2D."""
        shape = self.shape(i)[::-1] + (8,)
        if out is None:
            out = torch.empty(shape, dtype=torch.float64, device=self.device)
        if tuple(out.shape) != shape or out.dtype != torch.float64 or not out.is_contiguous():
            raise ValueError(f"snapshot buffer must be a contiguous float64 tensor of shape {shape}")
        L.check(self.lib.ah_plan_snapshot(self.h, int(i), ctypes.c_void_p(out.data_ptr()), self._stream()),
                "ah_plan_snapshot")
        return out

    def capture(self, i, every="step", slots=64):
        """Let the step program take problem i's snapshots (ah_plan_set_capture):
        after every step (every="step") or after the steps whose dt was
        clamped to a checkpoint (every="checkpoint"), into a device ring of
        `slots` snapshots; every=None stops it.  ``drain`` reads them."""
        modes = {None: L.AH_CAPTURE_OFF, "step": L.AH_CAPTURE_EVERY_STEP, "checkpoint": L.AH_CAPTURE_CHECKPOINTS}
        if every not in modes:
            raise ValueError(f"every must be one of {list(modes)}")
        mode = modes[every]
        if mode == L.AH_CAPTURE_OFF:
            L.check(self.lib.ah_plan_set_capture(self.h, int(i), mode, None, 0, None, None, self._stream()),
                    "ah_plan_set_capture")
            self._cap = None
            return
        if slots <= 0:
            raise ValueError("slots must be positive")
        shape = self.shape(i)[::-1] + (8,)
        ring = torch.empty((slots,) + shape, dtype=torch.float64, device=self.device)
        times = torch.empty(slots, dtype=torch.float64, device=self.device)
        ctl = torch.empty(2, dtype=torch.int64, device=self.device)
        L.check(self.lib.ah_plan_set_capture(self.h, int(i), mode, ctypes.c_void_p(ring.data_ptr()), int(slots),
                                             ctypes.c_void_p(times.data_ptr()), ctypes.c_void_p(ctl.data_ptr()),
                                             self._stream()), "ah_plan_set_capture")
        self._cap = dict(ring=ring, times=times, ctl=ctl, slots=slots, read=0)

    def drain(self):
        """The snapshots captured since the last drain, oldest first, as
        (t, ring slot view) pairs; a slot is overwritten once the step
        program has taken `slots` more snapshots.  Synchronises the stream."""
        c = self._cap
        if c is None:
            raise L.AstreaHipError("drain: no capture set (Plan.capture)")
        n = int(c["ctl"][0].item())
        if n - c["read"] > c["slots"]:
            raise L.AstreaHipError(f"capture ring overran: {n - c['read']} snapshots for {c['slots']} slots")
        times = c["times"].cpu()
        out = [(float(times[k % c["slots"]]), c["ring"][k % c["slots"]]) for k in range(c["read"], n)]
        c["read"] = n
        return out

    def bytes_per_cell(self):
        """Device bytes per cell of this plan (state + workspace)."""
        return (self.state.numel() * 8 + self.workspace.numel()) / max(self.total, 1)

    def compact(self):
        n = ctypes.c_int()
        L.check(self.lib.ah_plan_compact(self.h, ctypes.byref(n), self._stream()), "ah_plan_compact")
        return n.value

    def set_blocks(self, nblocks):
        L.check(self.lib.ah_plan_set_blocks(self.h, int(nblocks)), "ah_plan_set_blocks")

    def status(self):
        st = (L.AhProblemStatus * self.nprob)()
        L.check(self.lib.ah_plan_status(self.h, st, self._stream()), "ah_plan_status")
        return [dict(t=s.t, dt=s.dt, steps=s.steps, active=bool(s.active), failed=bool(s.failed)) for s in st]

    def status_async(self):
        """Queue the statuses of this point of the stream (ah_plan_status_async
        into device memory, then a copy to pinned host memory) and return at
        once; ``PendingStatus.result()`` waits for that point only."""
        n = self.nprob * ctypes.sizeof(L.AhProblemStatus)
        dev = torch.empty(n, dtype=torch.uint8, device=self.device)
        L.check(self.lib.ah_plan_status_async(self.h, ctypes.c_void_p(dev.data_ptr()), self._stream()),
                "ah_plan_status_async")
        host = torch.empty(n, dtype=torch.uint8, pin_memory=True)
        host.copy_(dev, non_blocking=True)
        ev = torch.cuda.Event()
        ev.record(torch.cuda.current_stream(self.device))
        return PendingStatus(host, dev, ev, self.nprob)

    def run(self, chunk=32, max_chunks=10 ** 9):
        """Step until no problem is active; the host looks at the device time
        state once per chunk of steps."""
        for _ in range(max_chunks):
            self.step(chunk)
            st = self.status()
            if not any(s["active"] for s in st):
                return st
        return self.status()

    # ---- kernel timing ---------------------------------------------------------
    def profile(self, capacity):
        L.check(self.lib.ah_plan_profile(self.h, int(capacity)), "ah_plan_profile")

    def profile_read(self):
        ms = (ctypes.c_double * 8)()
        cnt = (ctypes.c_int * 8)()
        L.check(self.lib.ah_plan_profile_read(self.h, ms, cnt), "ah_plan_profile_read")
        return {k: (ms[i], cnt[i]) for i, k in enumerate(L.KERNEL_KINDS) if cnt[i]}

    # ---- parity entry points (single-problem plans) ----------------------------
    def _ptr(self, t):
        return ctypes.c_void_p(t.data_ptr()) if t is not None else None

    def evolve_space(self, parity=0, state=None):
        """evolvers.evolve_space: interface fluxes per physical axis in the
        reference's sweep-frame layout ((n_a + 1, n_other, 8) in 2D) and
        eigmax per axis; with constrained transport also 'mag_corner' (the
        corner EMF in the state's (x, y) order, as the reference attaches it
        to the axis-1 sweep; the axis-0 entry holds its transpose) and
        'face_avg' (the sweep's face-valued B component, (x, y) order)."""
        assert self.nprob == 1
        p = self.problems[0]
        X = self.state if state is None else state
        n0, n1 = p.n0, p.n1
        f0 = torch.zeros((self.nv, (n0 + 1) * n1), dtype=torch.float64, device=self.device)
        f1 = torch.zeros((self.nv, (n1 + 1) * n0), dtype=torch.float64, device=self.device)
        eig = (ctypes.c_double * 2)()
        ct = bool(self.scheme.magnetic_2d)
        if ct:
            corner = torch.zeros((n0, n1), dtype=torch.float64, device=self.device)
            faces = torch.zeros((2, n0, n1), dtype=torch.float64, device=self.device)
            L.check(self.lib.ah_evolve_space_ct(self.h, 0, int(parity), self._ptr(X), self._ptr(f0), self._ptr(f1),
                                                eig, self._ptr(corner), self._ptr(faces[0]), self._ptr(faces[1]),
                                                self._stream()), "ah_evolve_space_ct")
        else:
            L.check(self.lib.ah_evolve_space(self.h, 0, int(parity), self._ptr(X), self._ptr(f0), self._ptr(f1),
                                             eig, self._stream()), "ah_evolve_space")
        out = {}
        for ax, (f, ns, nt) in enumerate([(f0, n0, n1), (f1, n1, n0)][:self.dims]):
            h = f.cpu().numpy()
            g = np.zeros(((ns + 1) * nt, 8))
            for v, c in enumerate(self.varidx):
                g[:, c] = h[v]
            shape = (ns + 1,) if self.dims == 1 else (ns + 1, nt)
            out[ax] = {"flux": g.reshape(shape + (8,)), "eigmax": eig[ax], "raw": f}
            if ct:
                e3 = corner.cpu().numpy()
                out[ax]["mag_corner"] = e3 if ax == 1 else np.ascontiguousarray(e3.T)
                out[ax]["face_avg"] = faces[ax].cpu().numpy()
                out[ax]["raw_corner"], out[ax]["raw_face"] = corner, faces[ax]
        return out

    def compute_L(self, fluxes, parity=0):
        """evolvers.compute_L (with the EMF rows of constrained transport)
        in the packed state layout."""
        L_ = torch.zeros_like(self.state)
        f0 = fluxes[0]["raw"]
        f1 = fluxes[1]["raw"] if self.dims == 2 else f0
        if self.scheme.magnetic_2d:
            L.check(self.lib.ah_compute_L_ct(self.h, 0, int(parity), self._ptr(f0), self._ptr(f1),
                                             self._ptr(fluxes[1]["raw_corner"]), self._ptr(L_), self._stream()),
                    "ah_compute_L_ct")
        else:
            L.check(self.lib.ah_compute_L(self.h, 0, int(parity), self._ptr(f0), self._ptr(f1), self._ptr(L_),
                                          self._stream()), "ah_compute_L")
        return L_

    def evolve_time(self, fluxes, dt, parity=0, state=None):
        """evolvers.evolve_time on ``state`` (default: the plan's state) in
        place: L0 from ``fluxes`` (as evolve_space returns them), the field
        reassignment and refinement of constrained transport, the plan's
        Runge-Kutta stages.  Returns the updated state tensor."""
        X = self.state if state is None else state
        f0 = fluxes[0]["raw"]
        f1 = fluxes[1]["raw"] if self.dims == 2 else f0
        ct = bool(self.scheme.magnetic_2d)
        corner = fluxes[1]["raw_corner"] if ct else None
        face0 = fluxes[0]["raw_face"] if ct else None
        face1 = fluxes[1]["raw_face"] if ct else None
        L.check(self.lib.ah_evolve_time(self.h, 0, int(parity), self._ptr(X), self._ptr(f0), self._ptr(f1),
                                        self._ptr(corner), self._ptr(face0), self._ptr(face1), float(dt),
                                        self._stream()), "ah_evolve_time")
        return X


class PendingStatus:
    """Statuses queued by ``Plan.status_async``."""

    def __init__(self, host, dev, event, nprob):
        self._host, self._dev, self._event, self._n = host, dev, event, nprob

    def ready(self):
        return self._event.query()

    def result(self):
        self._event.synchronize()
        arr = (L.AhProblemStatus * self._n).from_buffer_copy(self._host.numpy().tobytes())
        return [dict(t=s.t, dt=s.dt, steps=s.steps, active=bool(s.active), failed=bool(s.failed)) for s in arr]
