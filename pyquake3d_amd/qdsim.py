"""Device-resident quasi-dynamic earthquake-cycle stepper (one GPU, no MPI).

Mirrors the reference's ``QDsim`` time step (src/QDsim.py:892-996,
``simu_forward``) with the same semantics:
  * one operator application per step on the slip-rate deficit
    (V1 - Vpl cos rake0, V2 - Vpl sin rake0), frozen for every RK stage and
    every rejected attempt (:902-922);
  * Dormand-Prince 5(4) attempt with the max-relative error of psi and of
    (tau1, tau2, sigma) divided by the 4th-order values, RelTol (1e-4 in the
    reference, :1122-1123), hnew = 0.9 h (tol/err)^0.2 (:1001-1134);
  * accept if max(cond1, cond2) < 1 and neither is NaN; dtnext = min(hnew1,
    hnew2, 1.5 h); otherwise h = max(0.5 h, min(hnew1, hnew2)), abort when
    h < 1e-15 or after 20 rejections (:931-958);
  * commit: regularized law per component, rake, V floor 1e-30, slip (:960-993).
The RK attempt is one fused HIP kernel (all 7 stages per cell in registers);
the operator is an ``HMatrix`` (FP32 or FP64 storage) or a ``DenseOperator``.
All device arrays are in the operator's cluster order.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _lib
from .problem import Problem


class StepAborted(RuntimeError):
    """dt fell below 1e-15 s or more than 20 attempts were rejected."""


class DenseOperator:
    """The reference's plain method: the six traction operators stored densely
    in FP64 (the dense-block products of blocks_process_MVM,
    src/Hmatrix.py:1445-1538, applied as in src/QDsim.py:906-917), on this
    library's block matvec. It is an ``HMatrix`` with admissibility eta = 0,
    so every leaf-cluster pair is a near-field block assembled densely on the
    device and no block is compressed; like any HMatrix it works in cluster
    order (``perm``)."""

    def __init__(self, mesh, mu, lambda_, device="cuda", dtype=torch.float64, halfspace=False,
                 jud_ele_order=False, leaf=64):
        from .hmatrix import HMatrix
        self.mesh, self.mu, self.lambda_ = mesh, float(mu), float(lambda_)
        self.halfspace, self.jud_ele_order = bool(halfspace), bool(jud_ele_order)
        self.H = HMatrix.from_mesh(mesh, mu, lambda_, leaf=leaf, eta=0.0, dtype=dtype,
                                   device=device, halfspace=halfspace,
                                   jud_ele_order=jud_ele_order)
        if self.H.stats.n_lowrank:
            raise RuntimeError("an eta = 0 partition has no admissible block")
        self.n = self.H.n
        self.perm = self.H.perm
        self.device = self.H.device
        self.dtype = dtype
        self.matvec_bytes = self.H.matvec_bytes

    @property
    def accumulate(self):
        return self.H.accumulate

    def _bufs(self, nrhs):
        return self.H._bufs(nrhs)

    def matvec(self, v1, v2, out=None, stream=None, nrhs=1, instance_major=False, planes=3):
        """As ``HMatrix.matvec`` (cluster order)."""
        return self.H.matvec(v1, v2, out=out, stream=stream, nrhs=nrhs,
                             instance_major=instance_major, planes=planes)

    def algorithmic_bytes(self, nrhs=1, planes=3):
        return self.H.algorithmic_bytes(nrhs, planes)

    def matrix(self) -> torch.Tensor:
        """[[A1s, A1d], [A2s, A2d], [Bs, Bd]] as one (3N, 2N) FP64 device
        matrix in the original element order, assembled by the traction-block
        kernel (for checks against host restatements of the step)."""
        from .tde import DeviceMesh
        dm = DeviceMesh.from_mesh(self.mesh, self.mu, self.lambda_, device=self.device,
                                  halfspace=self.halfspace, jud_ele_order=self.jud_ele_order)
        K = dm.traction_block()
        n = self.n
        M = torch.empty((3 * n, 2 * n), dtype=torch.float64, device=self.device)
        for c in range(3):
            M[c * n:(c + 1) * n, :n] = K[c]
            M[c * n:(c + 1) * n, n:] = K[3 + c]
        return M


class QDsim:
    """Time stepper on device. ``problem`` arrays are in original element
    order; ``op`` provides ``perm`` (cluster order) and ``matvec``."""

    def __init__(self, problem: Problem, op, reltol: float = 1e-4, device="cuda",
                 tableau: str = "ref"):
        """tableau "ref": the reference's coefficients (a63 = -46732/5247,
        src/QDsim.py:1019); "dp5": the published Dormand-Prince tableau."""
        if tableau not in ("ref", "dp5"):
            raise ValueError("tableau must be 'ref' or 'dp5'")
        self.a63 = -46732. / 5247. if tableau == "ref" else 46732. / 5247.
        self.p = problem
        self.op = op
        self.reltol = float(reltol)
        dev = torch.device(device)
        self.device = dev
        self.n = n = problem.n
        perm = np.asarray(op.perm)
        self.perm = perm

        def up(x):
            return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64)[perm])).to(dev)

        Tno, Tt1o, Tt2o, state, slipv1, slipv2, slipv = problem.initial_state()
        self.Tno, self.Tt1o, self.Tt2o, self.state = up(Tno), up(Tt1o), up(Tt2o), up(state)
        self.slipv1, self.slipv2, self.slipv = up(slipv1), up(slipv2), up(slipv)
        self.rake = up(problem.rake)
        self.slip1 = torch.zeros(n, dtype=torch.float64, device=dev)
        self.slip2 = torch.zeros(n, dtype=torch.float64, device=dev)
        self.a, self.b, self.dc, self.f0 = up(problem.a), up(problem.b), up(problem.dc), up(problem.f0)
        rake0 = np.asarray(problem.rake, dtype=np.float64)
        self.pl1 = up(problem.Vpl * np.cos(rake0))
        self.pl2 = up(problem.Vpl * np.sin(rake0))
        self.shear_loading = up(problem.shear_loading)
        self.normal_loading = up(problem.normal_loading)
        self.fix_Tn = bool(problem.fix_Tn)
        self.V0 = float(problem.V0)
        self.mu = float(problem.mu)
        self.Cs = float(problem.Cs)
        self.cellpar = torch.tensor([self.V0, self.mu, self.Cs], dtype=torch.float64, device=dev)
        # candidate state, deficits, operator output (AdotV1, AdotV2, dsigmadt)
        self.y = torch.empty((4, n), dtype=torch.float64, device=dev)
        self.x = torch.empty((2, n), dtype=torch.float64, device=dev)
        self.K = torch.empty((3, n), dtype=torch.float64, device=dev)
        self.hdev = torch.zeros(1, dtype=torch.float64, device=dev)
        self.err = torch.zeros(4, dtype=torch.int64, device=dev)
        self.vmax = torch.zeros(1, dtype=torch.int64, device=dev)
        self.err_host = torch.zeros(4, dtype=torch.int64)
        if dev.type == "cuda":
            self.err_host = self.err_host.pin_memory()
        self.time = 0.0
        self.htry = 1e-3
        self.maxslipv0 = float(np.max(slipv))
        self.Relerrormax1 = self.Relerrormax2 = 0.0
        self.Relerrormax1_last = self.Relerrormax2_last = 0.0
        self.nsteps = 0
        self.nattempts = 0
        self.nrejects = 0
        self.timing = False          # record HIP events around every operator product
        self.matvec_events = []

    # -------------------------------------------------------------- kernels
    def _stream(self):
        return _lib.stream_handle()

    def compute_K(self):
        """Deficit -> operator -> (AdotV1, AdotV2, dsigmadt) on device."""
        s = self._stream()
        _lib.call("pq3d_deficit", self.n, _lib.ptr(self.slipv1), _lib.ptr(self.slipv2),
                  _lib.ptr(self.pl1), _lib.ptr(self.pl2), _lib.ptr(self.x[0]),
                  _lib.ptr(self.x[1]), s)
        # with a fixed normal traction the Bs, Bd products are not needed
        # (src/QDsim.py:906-909)
        planes = 2 if self.fix_Tn else 3
        self._kcalls = getattr(self, "_kcalls", 0) + 1
        if self.timing and self._kcalls % max(int(getattr(self, "timing_every", 1)), 1) == 0:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            self.op.matvec(self.x[0], self.x[1], out=self.K, planes=planes)
            e1.record()
            self.matvec_events.append((e0, e1))
        else:
            self.op.matvec(self.x[0], self.x[1], out=self.K, planes=planes)
        _lib.call("pq3d_add_loading", self.n, _lib.ptr(self.K[2]), _lib.ptr(self.normal_loading),
                  int(self.fix_Tn), s)

    def _attempt(self, h):
        """One RK45 attempt; returns (Relerrormax1, Relerrormax2 parts)."""
        self.hdev.fill_(h)
        s = self._stream()
        _lib.call("pq3d_rk45_attempt", self.n, 0, _lib.ptr(self.hdev), _lib.ptr(self.Tno),
                  _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o), _lib.ptr(self.state),
                  _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.dc), _lib.ptr(self.f0),
                  _lib.ptr(self.K[0]), _lib.ptr(self.K[1]), _lib.ptr(self.K[2]),
                  _lib.ptr(self.shear_loading), _lib.ptr(self.slipv), _lib.ptr(self.cellpar),
                  _lib.ptr(self.y[0]), _lib.ptr(self.y[1]), _lib.ptr(self.y[2]),
                  _lib.ptr(self.y[3]), _lib.ptr(self.err), self.a63, 1, s)
        self.err_host.copy_(self.err, non_blocking=True)
        if self.device.type == "cuda":
            torch.cuda.current_stream().synchronize()
        e = self.err_host.numpy().view(np.float64)
        self.nattempts += 1
        return float(e[0]), float(e[1]), float(e[2]), float(e[3])

    def _commit(self, h):
        self.hdev.fill_(h)
        _lib.call("pq3d_rk45_commit", self.n, 0, _lib.ptr(self.hdev), 0, _lib.ptr(self.a),
                  _lib.ptr(self.cellpar), _lib.ptr(self.y[0]), _lib.ptr(self.y[1]),
                  _lib.ptr(self.y[2]), _lib.ptr(self.y[3]), _lib.ptr(self.Tno),
                  _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o), _lib.ptr(self.state),
                  _lib.ptr(self.slipv1), _lib.ptr(self.slipv2), _lib.ptr(self.slipv),
                  _lib.ptr(self.rake), _lib.ptr(self.slip1), _lib.ptr(self.slip2),
                  _lib.ptr(self.vmax), 1, self._stream())

    # ---------------------------------------------------- reference surface
    def RungeKutte_solve_Dormand_Prince(self, h):
        """One attempt (src/QDsim.py:1001-1134): returns
        (condition1, condition2, hnew1, hnew2); the candidate stays on device."""
        e_state, e_t1, e_t2, e_s = self._attempt(h)
        self.Relerrormax1 = np.float64(e_state) + 1e-10
        self.Relerrormax2 = max(np.float64(e_t1), np.float64(e_t2), np.float64(e_s)) + 1e-10
        RelTol1 = RelTol2 = self.reltol
        condition1 = self.Relerrormax1 / RelTol1
        condition2 = self.Relerrormax2 / RelTol2
        hnew1 = h * 0.9 * (RelTol1 / self.Relerrormax1) ** 0.2
        hnew2 = h * 0.9 * (RelTol2 / self.Relerrormax2) ** 0.2
        return condition1, condition2, hnew1, hnew2

    def simu_forward(self, dttry):
        """One accepted time step (src/QDsim.py:892-996): returns (h, dtnext)."""
        self.compute_K()
        nrjct = 0
        h = dttry
        while True:
            condition1, condition2, hnew1, hnew2 = self.RungeKutte_solve_Dormand_Prince(h)
            if max(condition1, condition2) < 1.0 and not (np.isnan(condition1) or np.isnan(condition2)):
                dtnext = min(hnew1, hnew2)
                dtnext = min(1.5 * h, dtnext)
                self.Relerrormax1_last = self.Relerrormax1
                self.Relerrormax2_last = self.Relerrormax2
                break
            nrjct = nrjct + 1
            self.nrejects += 1
            dtnext = min(hnew1, hnew2)
            h = max(0.5 * h, dtnext)
            if h < 1.e-15 or nrjct > 20:
                raise StepAborted(f"dt is too small (h={h:g}, rejections={nrjct})")
        self.time = self.time + h
        self._commit(h)
        self.nsteps += 1
        return h, dtnext

    def derivative(self, Tno, Tt1o, Tt2o, state):
        """RHS at the given cell arrays (device tensors, cluster order) with the
        current frozen K (src/QDsim.py:760-813): (dstatedt, dsigmadt/1e6,
        dtau1dt/1e6, dtau2dt/1e6)."""
        out = torch.empty((4, self.n), dtype=torch.float64, device=self.device)
        _lib.call("pq3d_qd_rhs", self.n, _lib.ptr(Tno), _lib.ptr(Tt1o), _lib.ptr(Tt2o),
                  _lib.ptr(state), _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.dc),
                  _lib.ptr(self.f0), _lib.ptr(self.K[0]), _lib.ptr(self.K[1]),
                  _lib.ptr(self.K[2]), _lib.ptr(self.shear_loading), _lib.ptr(self.slipv),
                  self.V0, self.mu, self.Cs, _lib.ptr(out), self._stream())
        return out[0], out[1], out[2], out[3]

    # ------------------------------------------------------------- helpers
    def matvec_ms(self) -> list:
        """Durations (ms) of the recorded operator products (synchronises)."""
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in self.matvec_events]

    def max_slipv(self) -> float:
        return float(self.slipv.max())

    def to_original(self, x: torch.Tensor) -> np.ndarray:
        out = np.empty(self.n)
        out[self.perm] = x.detach().cpu().numpy()
        return out

    def snapshot(self) -> dict:
        g = self.to_original
        return dict(time=self.time, Tno=g(self.Tno), Tt1o=g(self.Tt1o), Tt2o=g(self.Tt2o),
                    state=g(self.state), slipv1=g(self.slipv1), slipv2=g(self.slipv2),
                    slipv=g(self.slipv), rake=g(self.rake), slip1=g(self.slip1),
                    slip2=g(self.slip2))

    SNAPSHOT_FIELDS = ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "rake",
                       "slip1", "slip2")

    def restore(self, snap: dict) -> None:
        """Continue from a snapshot in original element order (``snapshot()``,
        or ``writers.restart_fields`` of a VTK file, as the reference's
        restart from a VTK file, src/QDsim.py:656-682). The simulated time is
        kept unless the snapshot holds one."""
        for k in self.SNAPSHOT_FIELDS:
            x = np.asarray(snap[k], dtype=np.float64)
            if x.shape != (self.n,):
                raise ValueError(f"snapshot field {k} must have one value per element")
            getattr(self, k).copy_(torch.from_numpy(np.ascontiguousarray(x[self.perm])))
        if "time" in snap:
            self.time = float(snap["time"])
        self.maxslipv0 = float(np.max(snap["slipv"]))

    def ouputVTK(self, fname) -> None:
        """The current state as the reference's VTK file (QDsim.ouputVTK,
        src/QDsim.py:1142-1289; the method keeps the reference's name)."""
        from .writers import vtk_cells, write_vtk
        write_vtk(fname, self.p.mesh.nodelst, self.p.mesh.elelst,
                  vtk_cells(self.snapshot(), self.p))

    def run(self, nsteps: int, dttry=None, log=None, vmax_every=1):
        """Run ``nsteps`` accepted steps; returns rows (step, h, t, maxV).
        ``log``: a ``writers.StateLog`` that gets the reference's state.txt
        line of every step (src/main.py:177)."""
        dt = self.htry if dttry is None else dttry
        rows = []
        for i in range(nsteps):
            h, dt = self.simu_forward(dt)
            if log is not None:
                log.step(self.nsteps - 1, h, float(self.slipv1.abs().max()),
                         float(self.slipv2.abs().max()), self.time)
            if vmax_every and (i % vmax_every == 0):
                rows.append((self.nsteps, h, self.time, self.max_slipv()))
        self.htry = dt
        return rows


class EnsembleQDsim:
    """NB independent simulations on one fault mesh (different friction
    fields / initial states), stepped together: one multi-RHS operator product
    per step serves all of them (the operator is read once), and every
    instance keeps its own step size and the reference's accept/reject
    controller. Cell fields are instance-major (cell i of instance b at
    b*n + i, cluster order), so a wavefront of the cell kernels belongs to
    one instance and a retry round costs only the instances still
    attempting; the deficit is written interleaved (i*NB + b) for the
    product, which returns instance-major.

    A lockstep iteration ends when every instance has accepted its step; the
    retry rounds recompute only the instances that were rejected (the others
    pass h < 0 and keep their accepted candidate)."""

    def __init__(self, problems, op, reltol: float = 1e-4, device="cuda", tableau: str = "ref"):
        nb = len(problems)
        # a MultiHMatrix (meshes of different size): one block-diagonal
        # product per step; otherwise one shared mesh and an NB-RHS product
        self.ragged = bool(getattr(op, "ragged", False))
        if self.ragged:
            sizes = np.asarray(op.sizes)
            if len(sizes) != nb or any(p.n != s for p, s in zip(problems, sizes)):
                raise ValueError("problem b must have the mesh of the operator's part b")
            n, ne, self.nrhs = None, int(op.n), 1
        else:
            if nb not in (1, 2, 4, 8, 16):
                raise ValueError("ensemble size must be 1, 2, 4, 8 or 16")
            n = problems[0].n
            for p in problems:
                if p.n != n:
                    raise ValueError("all instances must share the mesh of the operator")
            ne, self.nrhs = n * nb, nb
        if tableau not in ("ref", "dp5"):
            raise ValueError("tableau must be 'ref' or 'dp5'")
        self.a63 = -46732. / 5247. if tableau == "ref" else 46732. / 5247.
        self.nb, self.n, self.ne = nb, n, ne
        self.op = op
        self.reltol = float(reltol)
        dev = torch.device(device)
        self.device = dev
        perm = np.asarray(op.perm)
        self.perm = perm

        perms = op.perms if self.ragged else [perm] * nb
        self.off = np.concatenate([[0], np.cumsum([len(pb) for pb in perms])]).astype(np.int64)

        def inter(arrs):
            # instance b's cells, in its mesh's cluster order, one after another
            A = np.concatenate([np.asarray(x, dtype=np.float64)[pb]
                                for x, pb in zip(arrs, perms)])
            return torch.from_numpy(np.ascontiguousarray(A)).to(dev)

        inits = [p.initial_state() for p in problems]
        (self.Tno, self.Tt1o, self.Tt2o, self.state, self.slipv1, self.slipv2,
         self.slipv) = (inter([s[k] for s in inits]) for k in range(7))
        self.rake = inter([p.rake for p in problems])
        self.slip1 = torch.zeros(self.ne, dtype=torch.float64, device=dev)
        self.slip2 = torch.zeros(self.ne, dtype=torch.float64, device=dev)
        self.a = inter([p.a for p in problems])
        self.b = inter([p.b for p in problems])
        self.dc = inter([p.dc for p in problems])
        self.f0 = inter([p.f0 for p in problems])
        self.pl1 = inter([p.Vpl * np.cos(p.rake) for p in problems])
        self.pl2 = inter([p.Vpl * np.sin(p.rake) for p in problems])
        self.shear_loading = inter([p.shear_loading for p in problems])
        self.normal_loading = inter([p.normal_loading for p in problems])
        if len({bool(p.fix_Tn) for p in problems}) != 1:
            raise ValueError("fix_Tn must be the same for all instances")
        self.fix_Tn = bool(problems[0].fix_Tn)
        inst = np.repeat(np.arange(nb, dtype=np.int32), np.diff(self.off))
        self.inst = torch.from_numpy(inst).to(dev)
        self.cellpar = torch.tensor([[p.V0, p.mu, p.Cs] for p in problems],
                                    dtype=torch.float64, device=dev).reshape(-1)
        self.y = torch.empty((4, self.ne), dtype=torch.float64, device=dev)
        self.x = torch.empty((2, self.ne), dtype=torch.float64, device=dev)
        self.K = torch.empty((3, self.ne), dtype=torch.float64, device=dev)
        self.hdev = torch.zeros(nb, dtype=torch.float64, device=dev)
        self.err = torch.zeros(4 * nb, dtype=torch.int64, device=dev)
        self.vmax = torch.zeros(nb, dtype=torch.int64, device=dev)
        self.err_host = torch.zeros(4 * nb, dtype=torch.int64)
        self.vmax_host = torch.zeros(nb, dtype=torch.int64)
        if dev.type == "cuda":
            self.err_host = self.err_host.pin_memory()
            self.vmax_host = self.vmax_host.pin_memory()
        self.time = np.zeros(nb)
        self.dt = np.full(nb, 1e-3)
        self.nsteps = 0
        self.nattempts = 0
        self.nrejects = np.zeros(nb, dtype=np.int64)
        self.timing = False          # record HIP events around every operator product
        self.matvec_events = []

    def compute_K(self):
        s = _lib.stream_handle()
        op = self.op
        # the 16-RHS matrix-core product takes its inputs rounded to FP32: the
        # deficit goes straight into its weight rows
        w32 = (self.nrhs == 16 and getattr(op, "accumulate", None) == "mfma" and
               getattr(op, "esize", 8) == 4 and hasattr(op, "weights"))
        if w32:
            _lib.call("pq3d_deficit_w32", self.n, self.nrhs, _lib.ptr(self.slipv1),
                      _lib.ptr(self.slipv2), _lib.ptr(self.pl1), _lib.ptr(self.pl2),
                      _lib.ptr(op.weights()), s)
        elif self.nrhs > 1:
            _lib.call("pq3d_deficit_interleave", self.n, self.nrhs, _lib.ptr(self.slipv1),
                      _lib.ptr(self.slipv2), _lib.ptr(self.pl1), _lib.ptr(self.pl2),
                      _lib.ptr(self.x[0]), _lib.ptr(self.x[1]), s)
        else:
            _lib.call("pq3d_deficit", self.ne, _lib.ptr(self.slipv1), _lib.ptr(self.slipv2),
                      _lib.ptr(self.pl1), _lib.ptr(self.pl2), _lib.ptr(self.x[0]),
                      _lib.ptr(self.x[1]), s)
        x1, x2 = (None, None) if w32 else (self.x[0], self.x[1])
        planes = 2 if self.fix_Tn else 3
        self._kcalls = getattr(self, "_kcalls", 0) + 1
        if self.timing and self._kcalls % max(int(getattr(self, "timing_every", 1)), 1) == 0:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            op.matvec(x1, x2, out=self.K, nrhs=self.nrhs, instance_major=True, planes=planes)
            e1.record()
            self.matvec_events.append((e0, e1))
        else:
            op.matvec(x1, x2, out=self.K, nrhs=self.nrhs, instance_major=True, planes=planes)
        _lib.call("pq3d_add_loading", self.ne, _lib.ptr(self.K[2]),
                  _lib.ptr(self.normal_loading), int(self.fix_Tn), s)

    def _attempt(self, h):
        self.hdev.copy_(torch.from_numpy(np.asarray(h, dtype=np.float64)))
        _lib.call("pq3d_rk45_attempt", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hdev),
                  _lib.ptr(self.Tno), _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o),
                  _lib.ptr(self.state), _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.dc),
                  _lib.ptr(self.f0), _lib.ptr(self.K[0]), _lib.ptr(self.K[1]),
                  _lib.ptr(self.K[2]), _lib.ptr(self.shear_loading), _lib.ptr(self.slipv),
                  _lib.ptr(self.cellpar), _lib.ptr(self.y[0]), _lib.ptr(self.y[1]),
                  _lib.ptr(self.y[2]), _lib.ptr(self.y[3]), _lib.ptr(self.err), self.a63,
                  self.nb, _lib.stream_handle())
        self.err_host.copy_(self.err, non_blocking=True)
        if self.device.type == "cuda":
            torch.cuda.current_stream().synchronize()
        self.nattempts += 1
        return self.err_host.numpy().view(np.float64).reshape(self.nb, 4)

    def step(self):
        """One accepted step of every instance (the reference controller per
        instance); returns the accepted h per instance."""
        self.compute_K()
        h = self.dt.copy()
        done = np.zeros(self.nb, dtype=bool)
        nrj = np.zeros(self.nb, dtype=np.int64)
        dtnext = np.zeros(self.nb)
        tol = self.reltol
        while not done.all():
            # instances that accepted in an earlier round are left out (h < 0)
            e = self._attempt(np.where(done, -1.0, h))
            for b in np.where(~done)[0]:
                R1 = np.float64(e[b, 0]) + 1e-10
                R2 = max(np.float64(e[b, 1]), np.float64(e[b, 2]), np.float64(e[b, 3])) + 1e-10
                c1, c2 = R1 / tol, R2 / tol
                hn1 = h[b] * 0.9 * (tol / R1) ** 0.2
                hn2 = h[b] * 0.9 * (tol / R2) ** 0.2
                if max(c1, c2) < 1.0 and not (np.isnan(c1) or np.isnan(c2)):
                    dtnext[b] = min(min(hn1, hn2), 1.5 * h[b])
                    done[b] = True
                else:
                    nrj[b] += 1
                    h[b] = max(0.5 * h[b], min(hn1, hn2))
                    if h[b] < 1e-15 or nrj[b] > 20:
                        raise StepAborted(f"instance {b}: dt is too small (h={h[b]:g})")
        self.nrejects += nrj
        self.hdev.copy_(torch.from_numpy(h))
        _lib.call("pq3d_rk45_commit", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hdev), 0,
                  _lib.ptr(self.a), _lib.ptr(self.cellpar), _lib.ptr(self.y[0]),
                  _lib.ptr(self.y[1]), _lib.ptr(self.y[2]), _lib.ptr(self.y[3]),
                  _lib.ptr(self.Tno), _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o),
                  _lib.ptr(self.state), _lib.ptr(self.slipv1), _lib.ptr(self.slipv2),
                  _lib.ptr(self.slipv), _lib.ptr(self.rake), _lib.ptr(self.slip1),
                  _lib.ptr(self.slip2), _lib.ptr(self.vmax), self.nb, _lib.stream_handle())
        self.time += h
        self.dt = dtnext
        self.nsteps += 1
        return h

    def matvec_ms(self) -> list:
        """Durations (ms) of the recorded operator products (synchronises)."""
        torch.cuda.synchronize()
        return [a.elapsed_time(b) for a, b in self.matvec_events]

    STATE_FIELDS = ("Tno", "Tt1o", "Tt2o", "state", "slipv1", "slipv2", "slipv", "rake", "slip1",
                    "slip2")

    def save_state(self) -> dict:
        """Copies of the evolving fields, the simulated times and the next
        step sizes: ``load_state`` on an ensemble of the same problems
        continues from exactly this point."""
        d = {k: getattr(self, k).clone() for k in self.STATE_FIELDS}
        d["time"], d["dt"] = self.time.copy(), self.dt.copy()
        return d

    def load_state(self, d: dict) -> None:
        for k in self.STATE_FIELDS:
            getattr(self, k).copy_(d[k])
        self.time, self.dt = d["time"].copy(), d["dt"].copy()

    def max_slipv(self) -> np.ndarray:
        """Max slip rate of every instance after the last commit."""
        self.vmax_host.copy_(self.vmax)
        if self.device.type == "cuda":
            torch.cuda.current_stream().synchronize()
        return self.vmax_host.numpy().view(np.float64).copy()

    def instance(self, name: str, b: int) -> np.ndarray:
        """Field ``name`` of instance b in original element order."""
        o, e = int(self.off[b]), int(self.off[b + 1])
        x = getattr(self, name)[o:e].cpu().numpy()
        out = np.empty(e - o)
        out[self.op.perms[b] if self.ragged else self.perm] = x
        return out


class DeviceEnsemble(EnsembleQDsim):
    """EnsembleQDsim whose accept/reject controller runs on the device.

    One ``step()`` enqueues the operator product, ``rounds`` rounds of
    (attempt of every instance still attempting, the reference's decision,
    src/QDsim.py:924-958) and the commit of the accepted instances, with no
    host synchronisation. An instance that needs more than ``rounds``
    attempts keeps its state and reduced h and continues in the next
    ``step()``: the frozen operator product is recomputed from the same state
    (identical values), so the accepted step is the one the reference loop
    would take. ``sync()`` copies time, next dt and step counts to the host
    and raises StepAborted for an instance whose dt fell below 1e-15 s or
    that was rejected more than 20 times.

    Run control on the device (no host round trip per step):
      ``t_end``         simulated time (s) at which every instance stops
                        (it finishes once its time reaches t_end; the
                        others carry on);
      ``vup``/``vdown`` event detection: an event starts after the first
                        accepted step whose max slip rate exceeds vup while
                        armed, re-armed below vdown; ``events()`` returns the
                        start times, ``peak_slipv()`` the largest slip rate;
      ``mlp``           a trained ``learned_dt.StepMLP``: each fresh step's
                        first attempt is the network's proposal (the
                        estimator still accepts or rejects every attempt, a
                        rejection falls back to a PI controller);
      ``record``        log one training row per accepted step (features of
                        the step's start, log10 of the accepted h over the
                        reference controller's dtnext) into a device buffer
                        of ``record`` rows (``samples()``).
    """

    def __init__(self, problems, op, reltol: float = 1e-4, device="cuda", tableau: str = "ref",
                 rounds: int = 8, dt0: float = 1e-3, t_end: float = math.inf,
                 vup: float = 1e-2, vdown: float = 1e-4, mlp=None, record: int = 0,
                 margin: float = 0.0, hi: float = 0.0):
        super().__init__(problems, op, reltol=reltol, device=device, tableau=tableau)
        dev = self.device
        nb = self.nb
        self.rounds = int(rounds)
        self.hcur = torch.full((nb,), float(dt0), dtype=torch.float64, device=dev)
        self.dtnext = torch.zeros(nb, dtype=torch.float64, device=dev)
        self.flag = torch.zeros(nb, dtype=torch.uint8, device=dev)
        self.nrj = torch.zeros(nb, dtype=torch.int32, device=dev)
        self.status = torch.zeros(nb, dtype=torch.int32, device=dev)
        self.tdev = torch.zeros(nb, dtype=torch.float64, device=dev)
        self.nacc = torch.zeros(nb, dtype=torch.int64, device=dev)
        self.iterations = 0
        self.naccepted = np.zeros(nb, dtype=np.int64)
        W = _lib.controller_widths()
        self._W = W
        run = np.zeros((nb, W["run"]))
        run[:, 0] = float(t_end)
        run[:, 1] = 0.0                  # armed once max V falls below vdown
        run[:, 4], run[:, 5] = float(vup), float(vdown)
        self.run = torch.from_numpy(run).to(dev)
        self.mlp = mlp
        self.record_cap = int(record)
        self.learned = mlp is not None or self.record_cap > 0
        if self.learned:
            ls = np.zeros((nb, W["lstate"]))
            ls[:, 1] = 0.5               # last accepted max(c1, c2) before the first step
            ls[:, 2] = 1.0               # last dtnext / h before the first step
            self.lstate = torch.from_numpy(ls).to(dev)
            self.facc = torch.zeros(4 * nb, dtype=torch.int64, device=dev)
            # the slots hold the features of every instance's current state:
            # computed in full once (and after a state load), then kept
            # current by the commit of the accepted instances
            self._feat_current = False
            self.ncell = torch.from_numpy(np.diff(self.off).astype(np.int64)).to(dev)
            self.V0 = float(problems[0].V0)
            if len({float(p.V0) for p in problems}) != 1:
                raise ValueError("the learned proposal needs one V0 for all instances")
        else:
            self.lstate = None
        self.samples_buf = (torch.zeros((self.record_cap, W["sample"]), dtype=torch.float32,
                                        device=dev) if self.record_cap > 0 else None)
        if mlp is not None:
            norm = np.concatenate([mlp.x_mu, mlp.x_sd, [mlp.y_mu, mlp.y_sd, margin, hi]])
            self.norm = torch.from_numpy(norm.astype(np.float32)).to(dev)
        else:
            self.norm = None

    _graph = None
    _graph_key = None
    _graph_steps = 1

    def _graph_state(self):
        # the matrix-core product's configuration selects the step lists the
        # graph points at (the autotuner frees the ones it does not keep)
        # and a graph recorded with current feature slots leaves out their
        # full reduction, so a state load (slots no longer current) drops it
        cfg = getattr(self.op, "tuning", None)
        return (getattr(self.op, "accumulate", None), self.mlp is not None, self.rounds,
                None if cfg is None else tuple(sorted(cfg.items())),
                getattr(self, "_feat_current", None))

    def capture(self, steps=1):
        """Record ``steps`` consecutive iterations (operator product, attempt
        rounds, commit/advance: device launches on fixed buffers, no host
        synchronisation) as one HIP graph; every later ``step()`` replays it
        and advances ``steps`` iterations, so the host launches once per
        ``steps`` iterations. The controller runs on the device, so an
        instance that finishes inside the graph stays finished for the rest
        of it (its later iterations skip it); the caller reads the finish
        flags between replays. Recording runs nothing, so the state is
        unchanged. The operator's scratch buffers are created first; a later
        change of the operator's accumulation mode, of its matrix-core
        configuration (``tuning``) or of ``rounds`` drops the graph. With
        the learned proposal the feature slots are brought up to date first
        (outside the graph); a later state load drops the graph."""
        if self.timing:
            raise ValueError("capture needs timing off (events are not recorded into the graph)")
        steps = int(steps)
        if steps < 1:
            raise ValueError("steps must be >= 1")
        if hasattr(self.op, "_bufs"):
            self.op._bufs(self.nrhs)
        if self.learned and not self._feat_current:
            self._features(_lib.stream_handle())
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for _ in range(steps):
                self._enqueue_step()
        torch.cuda.synchronize(self.device)
        self._graph = g
        self._graph_steps = steps
        self._graph_key = self._graph_state()

    def release_graph(self):
        self._graph = None
        self._graph_key = None

    def step(self):
        """One iteration, or ``steps`` of them when a graph recorded by
        ``capture(steps)`` is replayed."""
        if self._graph is not None and self._graph_key != self._graph_state():
            self.release_graph()
        if self._graph is not None:
            self._graph.replay()
            self.iterations += self._graph_steps
        else:
            self._enqueue_step()
            self.iterations += 1

    def _enqueue_step(self):
        s = _lib.stream_handle()
        if self.learned:
            if not self._feat_current:
                self._features(s)
            _lib.call("pq3d_learned_propose", self.nb, _lib.ptr(self.facc), _lib.ptr(self.ncell),
                      _lib.ptr(self.flag), _lib.ptr(self.nrj), _lib.ptr(self.hcur),
                      _lib.ptr(self.lstate), _lib.ptr(self.mlp.prm) if self.mlp else None,
                      _lib.ptr(self.norm), s)
        self.compute_K()
        _lib.call("pq3d_rk45_rounds", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hcur),
                  _lib.ptr(self.flag), _lib.ptr(self.nrj), _lib.ptr(self.dtnext),
                  _lib.ptr(self.status), self.rounds, self.reltol, _lib.ptr(self.Tno),
                  _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o), _lib.ptr(self.state),
                  _lib.ptr(self.a), _lib.ptr(self.b), _lib.ptr(self.dc), _lib.ptr(self.f0),
                  _lib.ptr(self.K[0]), _lib.ptr(self.K[1]), _lib.ptr(self.K[2]),
                  _lib.ptr(self.shear_loading), _lib.ptr(self.slipv), _lib.ptr(self.cellpar),
                  _lib.ptr(self.y[0]), _lib.ptr(self.y[1]), _lib.ptr(self.y[2]),
                  _lib.ptr(self.y[3]), _lib.ptr(self.err), self.a63, self.nb,
                  _lib.ptr(self.lstate), s)
        _lib.call("pq3d_rk45_advance", self.ne, _lib.ptr(self.inst), _lib.ptr(self.hcur),
                  _lib.ptr(self.flag), _lib.ptr(self.nrj), _lib.ptr(self.dtnext),
                  _lib.ptr(self.tdev), _lib.ptr(self.nacc), _lib.ptr(self.a),
                  _lib.ptr(self.cellpar), _lib.ptr(self.y[0]), _lib.ptr(self.y[1]),
                  _lib.ptr(self.y[2]), _lib.ptr(self.y[3]), _lib.ptr(self.Tno),
                  _lib.ptr(self.Tt1o), _lib.ptr(self.Tt2o), _lib.ptr(self.state),
                  _lib.ptr(self.slipv1), _lib.ptr(self.slipv2), _lib.ptr(self.slipv),
                  _lib.ptr(self.rake), _lib.ptr(self.slip1), _lib.ptr(self.slip2),
                  _lib.ptr(self.vmax), _lib.ptr(self.run), _lib.ptr(self.lstate),
                  _lib.ptr(self.samples_buf), self.record_cap, self.nb,
                  _lib.ptr(self.b) if self.learned else None,
                  _lib.ptr(self.f0) if self.learned else None,
                  self.V0 if self.learned else 0.0,
                  _lib.ptr(self.facc) if self.learned else None, s)

    def _features(self, s):
        """Feature slots of every instance from the current state (one
        reduction over all cells)."""
        _lib.call("pq3d_features", self.ne, self.nb, _lib.ptr(self.inst),
                  _lib.ptr(self.slipv), _lib.ptr(self.state), _lib.ptr(self.b),
                  _lib.ptr(self.f0), self.V0, _lib.ptr(self.facc), s)
        self._feat_current = True

    DEVICE_FIELDS = ("hcur", "dtnext", "flag", "nrj", "status", "tdev", "nacc", "run")

    def save_state(self) -> dict:
        self.sync()
        d = super().save_state()
        d.update({k: getattr(self, k).clone() for k in self.DEVICE_FIELDS})
        if self.lstate is not None:
            d["lstate"] = self.lstate.clone()
        return d

    def load_state(self, d: dict) -> None:
        super().load_state(d)
        for k in self.DEVICE_FIELDS:
            if k in d:
                getattr(self, k).copy_(d[k])
        if self.lstate is not None and "lstate" in d:
            self.lstate.copy_(d["lstate"])
        if self.learned:
            self._feat_current = False
        self.sync()

    def sync(self):
        """Host copies of time, next dt and accepted-step counts; raises
        StepAborted if an instance aborted."""
        st = self.status.cpu().numpy()
        self.time = self.tdev.cpu().numpy().copy()
        self.dt = self.hcur.cpu().numpy().copy()
        self.naccepted = self.nacc.cpu().numpy().copy()
        self.nsteps = int(self.naccepted.min())
        if st.any():
            b = int(np.argmax(st != 0))
            raise StepAborted(f"instance {b}: dt is too small or more than 20 rejections")
        return self.time

    def finished(self) -> np.ndarray:
        """Which instances have reached t_end (flag 3)."""
        return self.flag.cpu().numpy() == 3

    def aborted(self) -> np.ndarray:
        return self.status.cpu().numpy() != 0

    RUN_EV0 = 8

    def events(self) -> list:
        """Event start times (s) of every instance (at most the record's
        capacity; ``event_counts()`` has the full counts)."""
        r = self.run.cpu().numpy()
        cnt = r[:, 2].astype(np.int64)
        e0 = self.RUN_EV0
        cap = r.shape[1] - e0
        return [r[b, e0:e0 + min(int(cnt[b]), cap)].copy() for b in range(self.nb)]

    def rejected(self) -> np.ndarray:
        """Rejected attempts of every instance's accepted steps so far."""
        return self.run[:, 6].cpu().numpy().astype(np.int64)

    def event_counts(self) -> np.ndarray:
        return self.run[:, 2].cpu().numpy().astype(np.int64)

    def peak_slipv(self) -> np.ndarray:
        return self.run[:, 3].cpu().numpy().copy()

    def samples(self):
        """(features, targets) of the recorded training rows so far."""
        if self.samples_buf is None:
            raise ValueError("construct with record > 0 to log training rows")
        acc = self.nacc.cpu().numpy()
        rows = np.arange(self.record_cap)
        k, b = rows // self.nb, rows % self.nb
        valid = k < acc[b]
        S = self.samples_buf.cpu().numpy()[valid]
        return S[:, :6].astype(np.float64), S[:, 6].astype(np.float64)
