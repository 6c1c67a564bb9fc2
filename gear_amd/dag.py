"""Expression programs on the device: plan, batches and the persistent
CROWN bound (include/gearhip.h, gh_dag_*).

`DagPlan(programs)` compiles a list of `gear_amd.program.Program` into one
device-resident plan; `bound(xLs, xUs)` bounds every row of every program
over its own batch of boxes in one persistent launch.  It replaces, for
programs in this node vocabulary, BoundedModule.compute_bounds(...,
method='CROWN', bound_upper=False) (auto_LiRPA/bound_general.py:1275-1593)
as called per instance by input_split/bounding.py:316-326.
"""
import ctypes

import numpy as np
import torch

from . import _lib


class GhProg(ctypes.Structure):
    """gh_prog of include/gearhip_program.h (host pointers)."""
    _fields_ = [("n", ctypes.c_int32), ("V", ctypes.c_int32), ("R", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("nodes", ctypes.c_void_p), ("out", ctypes.c_void_p)]


def default_waves(device=None):
    """Resident wavefronts of a full persistent launch: 12 per CU (3 per
    SIMD, the kernel's register-limited occupancy)."""
    props = torch.cuda.get_device_properties(device or torch.cuda.current_device())
    return 12 * props.multi_processor_count


class DagPlan:
    def __init__(self, programs, device="cuda"):
        self.programs = list(programs)
        self.device = torch.device(device)
        if self.device.type == "cuda" and self.device.index is None:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.n = [P.n for P in self.programs]
        self.R = [P.R for P in self.programs]
        arrs = [P.arrays() for P in self.programs]
        progs = (GhProg * max(1, len(arrs)))()
        for k, (P, (nodes, out)) in enumerate(zip(self.programs, arrs)):
            progs[k] = GhProg(P.n, P.V, P.R, 0, nodes.ctypes.data, out.ctypes.data)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().gh_dag_plan_create(len(arrs), ctypes.addressof(progs),
                                                     ctypes.byref(handle)), "dag plan")
        self._h = handle
        jobs, slots, tape = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(_lib.lib().gh_dag_plan_info(handle, ctypes.byref(jobs), ctypes.byref(slots),
                                               ctypes.byref(tape)), "dag plan info")
        self.jobs, self.max_slots, self.tape_bytes = jobs.value, slots.value, tape.value
        asl = ctypes.c_int64()
        _lib.check(_lib.lib().gh_dag_plan_aslots(handle, ctypes.byref(asl)), "dag plan aslots")
        self.max_aslots = asl.value
        k = max(1, len(arrs))
        st, io, steps = (ctypes.c_int64 * k)(), (ctypes.c_int64 * k)(), (ctypes.c_int64 * k)()
        _lib.check(_lib.lib().gh_dag_plan_stats(handle, st, io, steps), "dag plan stats")
        # per program, per box: state bytes, I/O bytes, backward node visits
        self.state_bytes = list(st)[:len(arrs)]
        self.io_bytes = list(io)[:len(arrs)]
        self.steps = list(steps)[:len(arrs)]
        fl, tr = (ctypes.c_int64 * k)(), (ctypes.c_int64 * k)()
        _lib.check(_lib.lib().gh_dag_plan_work(handle, fl, tr), "dag plan work")
        # per program, per box: algorithmic FP64 operations, transcendentals
        self.flops = list(fl)[:len(arrs)]
        self.transcendentals = list(tr)[:len(arrs)]
        self._ws = None
        self._ws_waves = None

    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().gh_dag_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- device buffers
    def workspace(self, waves):
        if isinstance(waves, tuple):
            waves = waves[0]
        if self._ws is None or self._ws_waves != waves:
            nbytes = _lib.lib().gh_dag_workspace_bytes(self._h, waves)
            self._ws = torch.empty(max(nbytes, 256), dtype=torch.uint8, device=self.device)
            self._ws_waves = waves
        return self._ws

    def make_batches(self, xLs, xUs, lbs, counts, planes=None):
        """Device array of gh_dag_batch: per program (xL, xU, lb, count*,
        lA, lbias, undef); planes: per program (lA (cap, R, n), lbias
        (cap, R), undef (cap, R) uint8) or None."""
        rows = []
        for k, (xL, xU, lb, cnt) in enumerate(zip(xLs, xUs, lbs, counts)):
            for t in (xL, xU, lb):
                if t.dtype != torch.float64 or not t.is_contiguous() or t.device != self.device:
                    raise _lib.GearHipError("batch tensors must be contiguous float64 on the plan's device")
            if xL.shape[-1] != self.n[k] or lb.shape[-1] != self.R[k]:
                raise _lib.GearHipError(f"program {k}: box width {xL.shape[-1]} / rows {lb.shape[-1]} "
                                        f"do not match n={self.n[k]}, R={self.R[k]}")
            pl = [0, 0, 0]
            if planes is not None and planes[k] is not None:
                lA, lbias, und = planes[k]
                cap = xL.shape[0]
                if lA.shape != (cap, self.R[k], self.n[k]) or lbias.shape != (cap, self.R[k]) \
                        or und.shape != (cap, self.R[k]) or und.dtype != torch.uint8 \
                        or not (lA.is_contiguous() and lbias.is_contiguous() and und.is_contiguous()):
                    raise _lib.GearHipError(f"program {k}: plane buffers must be contiguous "
                                            f"(B, R, n), (B, R) float64 and (B, R) uint8")
                pl = [lA.data_ptr(), lbias.data_ptr(), und.data_ptr()]
            rows.append([xL.data_ptr(), xU.data_ptr(), lb.data_ptr(), cnt.data_ptr(), *pl])
        return torch.tensor(np.asarray(rows, dtype=np.uint64).view(np.int64).reshape(-1, 7),
                            dtype=torch.int64).to(self.device)

    # LDS per compute unit the A areas may take (of 160 KB)
    LDS_BUDGET = 144 << 10

    def lds_slots_for(self, waves):
        """A-area slots kept in LDS per wavefront at `waves` resident
        wavefronts: the plan's largest A area, within the LDS budget."""
        props = torch.cuda.get_device_properties(self.device)
        per_cu = max(1, -(-int(waves) // props.multi_processor_count))
        fit = self.LDS_BUDGET // (per_cu * 64 * 8)
        return int(max(0, min(self.max_aslots, fit, 64)))

    def launch(self, batches, waves, stream=None, lds_slots=None):
        ws = self.workspace(waves)
        if isinstance(waves, tuple):
            waves, lds_slots = waves
        if lds_slots is None:
            lds_slots = self.lds_slots_for(waves)
        _lib.check(_lib.lib().gh_dag_crown_f64(self._h, batches.data_ptr(), ws.data_ptr(),
                                               ws.numel(), waves, int(lds_slots),
                                               _lib.stream_ptr(stream)),
                   "dag crown")

    def tune_waves(self, batches, total_boxes, tuner=None):
        """(resident wavefronts, A-area LDS slots) for this plan and batch
        size, picked by the autotuner among 2, 4, 8 and 12 waves per compute
        unit, each with its A areas in LDS (as many slots as fit) or in the
        global workspace."""
        from .autotune import default_tuner
        tuner = tuner or default_tuner()
        props = torch.cuda.get_device_properties(self.device)
        cus = props.multi_processor_count
        cands = []
        for k in (2, 4, 8, 12):
            w = k * cus
            cands.append((w, 0))
            ls = self.lds_slots_for(w)
            if ls > 0:
                cands.append((w, ls))
        bucket = 1 << max(0, int(total_boxes).bit_length() - 1)
        key = ("dag_crown_f64_lds", getattr(props, "gcnArchName", ""), cus, self.jobs,
               self.max_slots, self.tape_bytes, bucket)
        choice, _ = tuner.pick(key, cands, lambda c: self.launch(batches, c[0], lds_slots=c[1]))
        return tuple(choice)

    def graph(self, batches, waves):
        """A captured replay of the bound over `batches` (a
        torch.cuda.CUDAGraph).  Counts are read on the device at replay, so
        the same graph serves changing batch sizes up to each capacity.
        waves: a wave count or a (waves, lds_slots) pair from tune_waves."""
        w = waves[0] if isinstance(waves, tuple) else waves
        self.workspace(w)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            self.launch(batches, waves)     # settles lazy library state outside capture
        torch.cuda.current_stream().wait_stream(s)
        with torch.cuda.graph(g):
            self.launch(batches, waves)
        return g

    def bound(self, xLs, xUs, waves=None, return_A=False, stream=None):
        """lb per program (B_k, R_k), FP64; with return_A also each row's
        lower plane and undefined flag: (lbs, lAs (B_k, R_k, n_k), lbiases
        (B_k, R_k), undefs (B_k, R_k) bool)."""
        waves = waves or default_waves(self.device)
        xLs = [x.contiguous() for x in xLs]
        xUs = [x.contiguous() for x in xUs]
        d = dict(dtype=torch.float64, device=self.device)
        lbs = [torch.empty(x.shape[0], r, **d) for x, r in zip(xLs, self.R)]
        planes = None
        if return_A:
            planes = [(torch.empty(x.shape[0], r, n, **d), torch.empty(x.shape[0], r, **d),
                       torch.empty(x.shape[0], r, dtype=torch.uint8, device=self.device))
                      for x, r, n in zip(xLs, self.R, self.n)]
        counts = torch.tensor([x.shape[0] for x in xLs], dtype=torch.int64, device=self.device)
        cnt_views = [counts[k:k + 1] for k in range(len(xLs))]
        batches = self.make_batches(xLs, xUs, lbs, cnt_views, planes)
        self.launch(batches, waves, stream)
        self._keep = (batches, counts, xLs, xUs)
        if not return_A:
            return lbs
        return (lbs, [p[0] for p in planes], [p[1] for p in planes],
                [p[2].bool() for p in planes])
