"""Branch and bound over expression programs (nonlinear programs whose rows
are expression DAGs, include/gearhip_program.h), on the device.

The loop is the reordered B&B step of the QCQP loop (gear_amd/bab.py,
reordered_step: batch_branch_and_bound.py:395-560 with reorder_bab) with
the expression-program bound in place of the QCQP bound: every picked
domain is bounded by the persistent CROWN kernel with its rows' planes
(BoundedModule.compute_bounds with return_A and needed_A_dict, as
input_split/bounding.py:316-326 calls it), never below its stored key,
branched on the widest input, bisected, its children keyed by the
objective plane (batch_branch_and_bound.py:189-225), pruned against the
incumbent and the constraint thresholds, and appended to a device
frontier.  Rows whose relaxation is undefined carry the vacuous plane
(A = 0, bias -inf), so their children keep the parent's bound.

The incumbent comes from ALM-PGD (the reference's ALMPGDSolver, complete_
verifier/nlp_solver/alm.py) run on the program's values as PyTorch tensor
operations with autograd, as the reference runs it on its model: the fused
ALM kernel covers QCQPs only.
"""
import torch

from .alm import ALMConfig, ALMPGDSolver
from .bab import BaBConfig, QCQPBranchAndBound
from .dag import DagPlan
from .program import (ADD, ADDC, CONST, COS, EXP, INPUT, LOG, MUL, MULC, NEG, POW, RECIP, SIN,
                      SQR, SUB)


class DeviceProgram:
    """A program's rows as differentiable tensor operations on the device:
    values(x) (N, n) -> (N, R).  n, S = R and rhs as DeviceQCQP."""

    def __init__(self, prog, device):
        self.prog = prog
        self.n, self.S = prog.n, prog.R
        self.device = torch.device(device)
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=self.device)
        self.rhs = t(prog.rhs)
        self.xl, self.xu = t(prog.xl), t(prog.xu)

    def values(self, x):
        vals = []
        for op, a, b, c in self.prog.nodes:
            if op == INPUT:
                v = x[:, a]
            elif op == CONST:
                v = torch.full((x.shape[0],), c, dtype=x.dtype, device=x.device)
            elif op == ADD:
                v = vals[a] + vals[b]
            elif op == SUB:
                v = vals[a] - vals[b]
            elif op == NEG:
                v = -vals[a]
            elif op == MULC:
                v = vals[a] * c
            elif op == ADDC:
                v = vals[a] + c
            elif op == MUL:
                v = vals[a] * vals[b]
            elif op == SQR:
                v = vals[a] * vals[a]
            elif op == POW:
                v = vals[a] ** c
            elif op == EXP:
                v = torch.exp(vals[a])
            elif op == LOG:
                v = torch.log(vals[a])
            elif op == RECIP:
                v = 1.0 / vals[a]
            elif op == SIN:
                v = torch.sin(vals[a])
            elif op == COS:
                v = torch.cos(vals[a])
            else:
                raise ValueError(op)
            vals.append(v)
        return torch.stack([vals[o] for o in self.prog.out], dim=1)


class BoundedProgram:
    """The bound of one expression program, with the call shape of
    BoundedQCQP.bound: lb (B, R) [, lA (B, P, n), lbias (B, P)] for the
    leading P = plane_rows rows."""

    large = False

    def __init__(self, prog, device="cuda"):
        self.prog = prog
        self.plan = DagPlan([prog], device=device)
        self.device = self.plan.device
        self.n, self.S = prog.n, prog.R
        self.dev = DeviceProgram(prog, self.device)
        self.last_undefined = None

    def bound(self, x_L, x_U, return_A=False, out=None, stream=None, plane_rows=None,
              rows=None):
        if rows is not None:
            x_L, x_U = x_L.index_select(0, rows), x_U.index_select(0, rows)
        if not return_A:
            lb = self.plan.bound([x_L], [x_U], stream=stream)[0]
        else:
            lbs, lAs, lbs_b, und = self.plan.bound([x_L], [x_U], return_A=True, stream=stream)
            lb = lbs[0]
            self.last_undefined = und[0]
            P = self.S if plane_rows is None else int(plane_rows)
            lA, lbias = lAs[0][:, :P].contiguous(), lbs_b[0][:, :P].contiguous()
        if out is not None:
            out.copy_(lb)
            lb = out
        return (lb, lA, lbias) if return_A else lb


class _ProgramProblem:
    """The program with the problem interface the loop reads (n, S, rhs,
    box, values for the FP64 replay of incumbents)."""

    def __init__(self, prog):
        import numpy as np
        self.prog = prog
        self.n, self.S, self.m = prog.n, prog.R, prog.R - 1
        self.rhs = np.asarray(prog.rhs, dtype=np.float64)
        self.xl = np.asarray(prog.xl, dtype=np.float64)
        self.xu = np.asarray(prog.xu, dtype=np.float64)

    def values(self, X):
        return self.prog.eval(X)


def program_alm_config():
    """ALM settings for expression programs: the presets' schedule with a
    restart budget sized for op-by-op autograd."""
    return ALMConfig(br_budget=256, outer_max_iter=10, inner_steps=100)


class ProgramBranchAndBound(QCQPBranchAndBound):
    """Device-resident B&B over one expression program (the reordered loop,
    FP64 bounds with planes, widest branching, child keys, prune)."""

    def __init__(self, prog, cfg: BaBConfig = None, device="cuda"):
        cfg = cfg or BaBConfig(alm=program_alm_config(), alm_interval=0)
        if not cfg.reorder or cfg.clip != "none" or cfg.dual_lp_iters:
            raise ValueError("expression programs run the reordered loop with box "
                             "concretization and without the LP merge")
        self.prog = prog
        net = BoundedProgram(prog, device)
        self._init_common(_ProgramProblem(prog), cfg, net)
        self.alm = ALMPGDSolver(cfg.alm, backend="torch")
