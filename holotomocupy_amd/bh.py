"""Fused BH gradient on gfx950 (the served path of the gradient evaluation).

``BHGradient(rec)`` binds the propagator and shift plans of a ``Rec`` and
evaluates, for a chunk of angles, exactly the quantities of
Rec.gradients_cascade (src/holotomocupy/rec_mpi.py:327-353) through the C ABI
``htamd_bh_gradient``: three FFT passes per (angle, distance), the F3 recompute
done once per pixel, deterministic reductions.
"""
import ctypes

import torch

from . import _lib


class BHGradient:
    def __init__(self, rec):
        self.rec = rec
        self.device = rec.device
        plan = ctypes.c_void_p()
        _lib.call("htamd_bh_plan_create", rec.cl_prop._plan, rec.cl_shift._plan, ctypes.byref(plan))
        self._plan = plan
        self._ws = None
        self._scales = {}

    def __del__(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value and _lib._lib is not None:
            _lib._lib.htamd_bh_plan_destroy(plan)
            self._plan = None

    def workspace_bytes(self, nangles):
        nb = ctypes.c_int64()
        _lib.call("htamd_bh_workspace_size", self._plan, int(nangles), ctypes.byref(nb))
        return nb.value

    def workspace(self, nangles):
        nb = self.workspace_bytes(nangles)
        if self._ws is None or self._ws.numel() < nb:
            self._ws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        return self._ws

    def scales(self, nangles):
        """Per-angle (2/data_size, rho_obj^2, rho_pos^2, 0) and per-probe rho_prb^2."""
        key = nangles
        if key not in self._scales:
            r = self.rec
            s = torch.tensor([2.0 / r.data_size, r.rho_sq["obj"], r.rho_sq["pos"], 0.0], dtype=torch.float32)
            self._scales[key] = (s.repeat(nangles, 1).to(self.device),
                                 torch.tensor([r.rho_sq["prb"]], dtype=torch.float32, device=self.device))
        return self._scales[key]

    def __call__(self, prb, proj, pos, data, mag, gproj, gpos, gprb, accumulate_prb=True,
                 objective=None, probe_index=None, scales=None, rho_prb=None, workspace=None):
        """Write gproj / gpos and accumulate (or overwrite) gprb for the angles
        of proj; all arguments are device tensors with the layouts of
        include/htamd.h.  objective: optional float64 device tensor [1]."""
        B = proj.shape[0]
        for t, name in ((prb, "prb"), (proj, "proj"), (pos, "pos"), (data, "data"), (mag, "mag"),
                        (gproj, "gproj"), (gpos, "gpos"), (gprb, "gprb")):
            if not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"{name} must be a contiguous tensor on {self.device}")
        if scales is None or rho_prb is None:
            s, rp = self.scales(B)
            scales = s if scales is None else scales
            rho_prb = rp if rho_prb is None else rho_prb
        ws = self.workspace(B) if workspace is None else workspace
        _lib.call("htamd_bh_gradient", self._plan, B, _lib.ptr(proj), _lib.ptr(pos), _lib.ptr(mag),
                  _lib.ptr(prb), _lib.ptr(probe_index) if probe_index is not None else None,
                  _lib.ptr(data), _lib.ptr(scales), _lib.ptr(rho_prb), 1 if accumulate_prb else 0,
                  _lib.ptr(gproj), _lib.ptr(gpos), _lib.ptr(gprb),
                  _lib.ptr(objective) if objective is not None else None,
                  _lib.ptr(ws), ws.numel(), _lib.stream_ptr())

    def call_tab(self, tables, nangles, nprobes, pos, mag, probe_index, scales, rho_prb, gpos, workspace,
                 accumulate_prb=False, objective=None):
        """htamd_bh_gradient_tab: the projections, data, object gradients,
        probes and probe gradients of the batch addressed through `tables`
        (int64 device tensor of 3 nangles + 2 nprobes addresses); pos, mag,
        probe_index, scales, rho_prb and gpos packed as for __call__."""
        if tables.dtype != torch.int64 or tables.numel() != 3 * nangles + 2 * nprobes:
            raise ValueError(f"tables must hold {3 * nangles + 2 * nprobes} int64 addresses")
        _lib.call("htamd_bh_gradient_tab", self._plan, int(nangles), int(nprobes), _lib.ptr(tables), _lib.ptr(pos),
                  _lib.ptr(mag), _lib.ptr(probe_index), _lib.ptr(scales), _lib.ptr(rho_prb),
                  1 if accumulate_prb else 0, _lib.ptr(gpos), _lib.ptr(objective) if objective is not None else None,
                  _lib.ptr(workspace), workspace.numel(), _lib.stream_ptr())

    def hessian_workspace_bytes(self, nangles, same):
        """same: False (general form), True (z = y) or 2 (the pair form)."""
        nb = ctypes.c_int64()
        mode = 2 if same == 2 and same is not True else (1 if same else 0)
        _lib.call("htamd_bh_hessian_workspace_size", self._plan, int(nangles), mode, ctypes.byref(nb))
        return nb.value

    def hessian_workspace(self, nangles, same):
        """The object's own bilinear-form workspace (grown on demand; calls on
        different streams must pass their own)."""
        nb = self.hessian_workspace_bytes(nangles, same)
        if getattr(self, "_hws", None) is None or self._hws.numel() < nb:
            self._hws = None  # release the old one first
            self._hws = torch.empty(nb, dtype=torch.uint8, device=self.device)
        return self._hws

    def hessian(self, x, y, z, data, mag, out, accumulate=False, scale=None, workspace=None):
        """Fused bilinear form of the data term (Rec.hessian_cascade,
        rec_mpi.py:250-303) for the angles of x['proj']: out (float64 device
        tensor [1]) (+)= scale sum [...] (htamd_bh_hessian), scale = 2/data_size
        by default.  x, y, z are {'prb', 'proj', 'pos'} dicts of device tensors;
        z is y means the same direction (three propagations instead of four)."""
        same = z is y
        B = x["proj"].shape[0]
        parts = [x["proj"], x["pos"], mag, x["prb"], y["proj"], y["pos"], y["prb"]]
        if not same:
            parts += [z["proj"], z["pos"], z["prb"]]
        for t in parts + [data, out]:
            if not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"bh hessian: every operand must be a contiguous tensor on {self.device}")
        ws = self.hessian_workspace(B, same) if workspace is None else workspace
        if ws.numel() < self.hessian_workspace_bytes(B, same):
            raise ValueError("bh hessian: workspace too small")
        zz = [None, None, None] if same else [_lib.ptr(z["proj"]), _lib.ptr(z["pos"]), _lib.ptr(z["prb"])]
        _lib.call("htamd_bh_hessian", self._plan, B, _lib.ptr(x["proj"]), _lib.ptr(x["pos"]), _lib.ptr(mag),
                  _lib.ptr(x["prb"]), _lib.ptr(y["proj"]), _lib.ptr(y["pos"]), _lib.ptr(y["prb"]), *zz,
                  _lib.ptr(data), float(2.0 / self.rec.data_size if scale is None else scale), _lib.ptr(out),
                  1 if accumulate else 0, _lib.ptr(ws), ws.numel(), _lib.stream_ptr())
        return out

    def hessian_pair(self, x, y, z, data, mag, out, accumulate=False):
        """out[0] (+)= H(y, z), out[1] (+)= H(z, z) in one fused pass
        (htamd_bh_hessian_pair): the two bilinear forms of the BH loop's beta
        step (rec_mpi.py:186-188).  out: float64 device tensor [2]."""
        B = x["proj"].shape[0]
        parts = [x["proj"], x["pos"], mag, x["prb"], y["proj"], y["pos"], y["prb"], z["proj"], z["pos"], z["prb"]]
        for t in parts + [data, out]:
            if not t.is_contiguous() or t.device != self.device:
                raise ValueError(f"bh hessian: every operand must be a contiguous tensor on {self.device}")
        if out.numel() < 2:
            raise ValueError("bh hessian pair: out needs two elements")
        ws = self.hessian_workspace(B, 2)
        _lib.call("htamd_bh_hessian_pair", self._plan, B, _lib.ptr(x["proj"]), _lib.ptr(x["pos"]), _lib.ptr(mag),
                  _lib.ptr(x["prb"]), _lib.ptr(y["proj"]), _lib.ptr(y["pos"]), _lib.ptr(y["prb"]),
                  _lib.ptr(z["proj"]), _lib.ptr(z["pos"]), _lib.ptr(z["prb"]), _lib.ptr(data),
                  float(2.0 / self.rec.data_size), _lib.ptr(out), 1 if accumulate else 0, _lib.ptr(ws), ws.numel(),
                  _lib.stream_ptr())
        return out
