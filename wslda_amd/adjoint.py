"""Gradients of a final-time observable with respect to the external-potential
parameters, by the adjoint (reverse-mode) PECE step of libwslda_td.so.

The forward run is the modular step (pca.c:1401-1570, non-gradient form) from
a start state with zero ABM history.  Every K steps it keeps a checkpoint (y,
history, potentials).  The backward sweep takes the checkpoints in reverse:
each K-step segment is recomputed with the taped step (which also records y,
the norms, <H> values, densities and potentials of every step), then the
adjoint steps run backwards through the segment.  The last taped_segments
segments are taped by the forward run itself and not recomputed.

The tape keeps one state array per step (y); the history entries the
backward steps read are rebuilt from it (tdw_adj_history_entry: H y - <H> y
with the step's potentials and <H>, the H application the step used), one per step,
in a window of h + 1 entries.  That costs one H application per backward
step and halves the tape, so twice as many segments fit in memory untouched
by the recompute.  tape_history=True keeps the entries on the tape instead
(two state arrays per step).

The observable is a linear functional of the final densities,
O = sum_r g(r) . (rho_a, rho_b, Re nu, Im nu)(r), given by the 4*N array g,
or a function of the final state with its own seed (state_observable), such
as the current through a plane of the lattice (PlaneCurrent).

On 3D lattices the segments that are not taped may run the fused step
instead (forward_precision = td.FP64 or td.SPLIT).  In split precision the
checkpoints keep the FP32 history (3.5 state arrays instead of 6); every
taped segment, the recompute and the backward sweep stay FP64 on the modular
step, so the gradient is the FP64 map's, linearised along the split run.
"""
import ctypes as C
import math

import numpy as np

from . import td

NPAR = 7   # per Gaussian: amp, cx, cy, cz, wx, wy, wz


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class AdjointProblem:
    """Forward run with checkpoints and the adjoint sweep on one context.

    params: td.Params (SLDA, FP64); wf0: [2n, N] complex128 start state
    (normalised here); pots0: 12*N start potentials; fbeta/kt/cnt: state
    weights; nsteps: steps; K: checkpoint interval; it0: step index of the
    start; dens_grad: 4*N observable weights (see module doc);
    taped_segments: how many of the last K-step segments the forward run
    tapes directly (each holds K+1 state arrays, 2K+1 with tape_history),
    so the backward sweep recomputes only the others."""

    def __init__(self, params, wf0, pots0, fbeta, kt=None, cnt=None, nsteps=100, K=10, it0=0, dens_grad=None,
                 device=0, taped_segments=1, state_observable=None, forward_precision=None, tape_history=False):
        import torch
        self.torch = torch
        self.params = params
        self.n = wf0.shape[0] // 2
        self.ctx = td.TDContext(params, self.n, device)
        self.ctx.set_states(fbeta, kt, cnt)
        # forward_precision (td.FP64 / td.SPLIT): the segments that are not
        # taped run the fused step on a context of their own
        self.ctx_fwd = None
        if forward_precision is not None:
            pf = td.Params.from_buffer_copy(params)
            pf.precision = forward_precision
            self.ctx_fwd = td.TDContext(pf, self.n, device)
            self.ctx_fwd.set_states(fbeta, kt, cnt)
            if self.ctx_fwd.L.tdw_fused_supported(self.ctx_fwd.h) != 1:
                self.ctx_fwd.close()
                raise td.TDWError("the fused step does not run this lattice / functional: "
                                  "no fused forward pass for the adjoint")
        self.dev = "cuda:%d" % device
        self.wf0 = torch.as_tensor(wf0).to(self.dev).to(torch.complex128).clone()
        self.pots0 = torch.as_tensor(np.ascontiguousarray(pots0, dtype=np.float64)).to(self.dev)
        self.nsteps, self.K, self.it0 = nsteps, K, it0
        self.N = self.ctx.N
        self.h = {td.AB3AM4: 3, td.AB4AM5: 4, td.AB5AM5: 5}[params.scheme]
        self.dens_grad = None if dens_grad is None else torch.as_tensor(
            np.ascontiguousarray(dens_grad, dtype=np.float64)).to(self.dev)
        self.ckpts = []
        self.run = None
        self.taped = {}
        self._tps = []
        self.taped_segments = taped_segments
        self.tape_history = tape_history
        # the history buffer a taped step writes while the run's own h
        # entries are all live (AB5AM5)
        self._fspare = None
        # state_observable(run) -> (O, dO/dy* as a [2n, N] tensor, paired by
        # Re sum conj(seed) dy): an observable of the final state itself
        self.state_observable = state_observable
        self.fbeta = np.asarray(fbeta, dtype=np.float64)

    # -- parameters -------------------------------------------------------
    def set_vext(self, vext):
        self.params.vext = vext
        self.ctx.set_params(self.params)
        if self.ctx_fwd is not None:
            pf = self.ctx_fwd.params
            pf.vext = vext
            self.ctx_fwd.set_params(pf)

    # -- forward ----------------------------------------------------------
    def _new_run(self, wf, pots, hist=None):
        """An FP64 modular run from (y, potentials, history); FP32 history
        entries (a split checkpoint) are widened."""
        R = td.TDRun(self.ctx, wf.clone(), pots.clone().cpu().numpy())
        if hist is not None:
            for j in range(self.h):
                R.f[j].copy_(hist[j])
        return R

    def _fwd_run(self):
        """The run of the segments that are not taped: the fused step on
        ctx_fwd when a forward precision was given, else the modular step."""
        if self.ctx_fwd is None:
            return self._new_run(self.wf0, self.pots0)
        return td.TDRun(self.ctx_fwd, self.wf0.clone(), self.pots0.clone().cpu().numpy(), fused=True)

    def _tape(self, i=0):
        """Tape buffer i: one K-step segment (allocated on first use)."""
        while len(self._tps) <= i:
            torch = self.torch
            n, N, K = self.n, self.N, self.K
            z = lambda: torch.zeros((2 * n, N), dtype=torch.complex128, device=self.dev)
            ty = [z() for _ in range(K + 1)]
            tf = [z() for _ in range(K)] if self.tape_history else []
            tscal = torch.zeros((K, 4 * n), dtype=torch.float64, device=self.dev)
            tgrid = torch.zeros((K, 4, 12 * N), dtype=torch.float64, device=self.dev)
            tapes = []
            for r in range(K):
                t = td.TapeStep()
                t.n = n
                t.scal = tscal[r].data_ptr()
                t.dens_p, t.pots_p = tgrid[r, 0].data_ptr(), tgrid[r, 1].data_ptr()
                t.dens_c, t.pots_c = tgrid[r, 2].data_ptr(), tgrid[r, 3].data_ptr()
                tapes.append(t)
            self._tps.append((ty, tf, tscal, tgrid, tapes))
        return self._tps[i]

    def _z(self):
        return self.torch.zeros((2 * self.n, self.N), dtype=self.torch.complex128, device=self.dev)

    def _free_entry(self, R):
        """A history buffer of the run that the next step does not read (the
        entry its rotation dropped last), for the new entry."""
        if self._fspare is None:
            self._fspare = self._z()
        live = {R.st.f[j] for j in range(self.h)}
        for t in list(R.f) + [self._fspare]:
            if t.data_ptr() not in live:
                return t
        raise td.TDWError("no free history buffer")

    def _taped_step(self, R, k, r, i=0):
        """Step k (row r of tape buffer i) with y written straight into the
        tape (tdw_adj_forward_step_to), and the new history entry into the
        tape (tape_history) or into a free history buffer of the run.  At
        r == 0 the run's y is first copied into the tape's row 0."""
        ty, tf, _, _, tapes = self._tape(i)
        if r == 0 and R.st.wf != ty[0].data_ptr():
            ty[0].copy_(self._tensor_at(R, R.st.wf))
            R.st.wf = ty[0].data_ptr()
        fnext = tf[r] if self.tape_history else self._free_entry(R)
        td.check(td.lib().tdw_adj_forward_step_to(self.ctx.h, C.byref(R.st), self.it0 + k, _p(ty[r + 1]),
                                                  _p(fnext), C.byref(tapes[r])), "tdw_adj_forward_step_to")

    def _tensor_at(self, R, ptr):
        """The tensor (the run's own, the spare history buffer or a tape row)
        whose data is at ptr."""
        spare = [self._fspare] if self._fspare is not None else []
        for t in [R.wf] + list(R.f) + spare:
            if t.data_ptr() == ptr:
                return t
        for (ty, tf, _, _, _) in self._tps:
            for t in list(ty) + list(tf):
                if t.data_ptr() == ptr:
                    return t
        raise td.TDWError("state pointer outside the run and its tape")

    def _restore(self, R):
        """Points the run's y and history back at its own tensors (copying the
        values out of the tape), so that it can step and be checkpointed."""
        own = {t.data_ptr() for t in R.f}
        if R.st.wf != R.wf.data_ptr():
            R.wf.copy_(self._tensor_at(R, R.st.wf))
            R.st.wf = R.wf.data_ptr()
        ptrs = [R.st.f[j] for j in range(5)]
        if all(p in own for p in ptrs):
            return
        # values of every slot first (a slot may still hold one of the run's
        # own tensors in a shifted position), then write them back in order
        vals = [self._tensor_at(R, p).clone() for p in ptrs]
        for j in range(5):
            R.f[j].copy_(vals[j])
            R.st.f[j] = R.f[j].data_ptr()
        del vals

    def forward(self, keep_checkpoints=True):
        """Runs nsteps steps; returns the run (final y, densities, potentials).
        With checkpoints, the last taped_segments segments run with the taped
        step, so the backward sweep does not recompute them."""
        # release the previous run and checkpoints before allocating anew
        self.run = None
        self.ckpts = []
        self.taped = {}
        R = self._fwd_run()
        R.initial_measure()
        K = self.K
        nseg = (self.nsteps + K - 1) // K
        first_taped = max(0, nseg - self.taped_segments)
        for k in range(self.nsteps):
            seg, r = divmod(k, K)
            # a taped segment that follows another one needs no checkpoint
            # on the light tape: the backward sweep rebuilds its start
            # history from the segment before
            if keep_checkpoints and r == 0 and (self.tape_history or seg <= first_taped):
                if R.fused:
                    R.finalize()
                self.ckpts.append((k, R.wf.clone(), [t.clone() for t in R.history()[:self.h]], R.pots.clone()))
            if keep_checkpoints and seg >= first_taped:
                if R.ctx is not self.ctx:
                    # the taped segments run the FP64 modular step from the
                    # checkpoint just taken
                    _, y0, hist0, pots0 = self.ckpts[-1]
                    del R
                    R = self._new_run(y0, pots0, hist0)
                self._taped_step(R, k, r, seg - first_taped)
                if r == K - 1 or k == self.nsteps - 1:
                    self.taped[seg] = seg - first_taped
                    self._restore(R)
            else:
                R.step(self.it0 + k, grad_last=False)
        if R.fused:
            R.finalize()
        R.ctx.sync()
        self.run = R
        return R

    def observable(self, R=None):
        R = R or self.run
        if self.state_observable is not None:
            return float(self.state_observable(R)[0])
        d = R.dens
        N = self.N
        g = self.dens_grad
        return float((g[:N] * d[2 * N:3 * N]).sum() + (g[N:2 * N] * d[7 * N:8 * N]).sum()
                     + (g[2 * N:3 * N] * d[0:2 * N:2]).sum() + (g[3 * N:4 * N] * d[1:2 * N:2]).sum())

    # -- backward ---------------------------------------------------------
    def backward(self):
        """dO/d(amp, cx, cy, cz, wx, wy, wz) of every Gaussian slot, shape
        (TDW_MAX_GAUSS, 7); needs forward() first."""
        torch = self.torch
        n, N, h, K = self.n, self.N, self.h, self.K
        dev = self.dev
        z = lambda: torch.zeros((2 * n, N), dtype=torch.complex128, device=dev)
        ly = z()
        lf = [z() for _ in range(5)]
        w = [z() for _ in range(5)]
        lpots = torch.zeros(4 * N, dtype=torch.float64, device=dev)
        lgrid = torch.zeros(12 * N, dtype=torch.float64, device=dev)
        theta = torch.zeros(7 * td.MAX_GAUSS, dtype=torch.float64, device=dev)
        A = td.AdjState()
        A.n = n
        A.ly = ly.data_ptr()
        for j in range(5):
            A.lf[j] = lf[j].data_ptr()
            A.w[j] = w[j].data_ptr()
        A.lpots, A.lgrid, A.theta_grad = lpots.data_ptr(), lgrid.data_ptr(), theta.data_ptr()
        if self.state_observable is not None:
            # dL/dy' of the last step's output state
            ly.copy_(self.state_observable(self.run)[1])
            torch.cuda.synchronize()
        L = self.ctx.L
        nseg = (self.nsteps + K - 1) // K
        ck = {k0 // K: (k0, y0, hist0, pots0) for (k0, y0, hist0, pots0) in self.ckpts}
        # light tape: the history entries rebuilt so far, by the global step
        # that wrote them (at most h + 1 live), and their free buffers
        win = {}
        pool = [] if self.tape_history else [z() for _ in range(h + 1)]
        torch.cuda.synchronize()

        def rebuilt(kg, yrow, tape):
            if kg not in win:
                buf = pool.pop() if pool else z()
                # w[2] is free between backward steps: the rebuild's scratch
                td.check(L.tdw_adj_history_entry(self.ctx.h, self.it0 + kg, C.byref(tape), _p(yrow), _p(buf),
                                                 _p(w[2])), "tdw_adj_history_entry")
                win[kg] = buf
            return win[kg]

        last = True
        for seg in reversed(range(nseg)):
            k0 = seg * K
            ks = min(K, self.nsteps - k0)
            hist0 = pots0 = None
            if seg in ck:
                _, y0, hist0, pots0 = ck[seg]
                # a split checkpoint's FP32 history, widened for this segment
                hist0 = [t if t.dtype == torch.complex128 else t.to(torch.complex128) for t in hist0]
            if seg in self.taped:
                ti = self.taped.pop(seg)
            else:
                # taped recompute of the segment (every taped buffer is free by now)
                ti = 0
                R = self._new_run(y0, pots0, hist0)
                for r in range(ks):
                    self._taped_step(R, k0 + r, r, ti)
                del R
            ty, tf, tscal, tgrid, tapes = self._tape(ti)
            # without a checkpoint (a taped segment after another on the light
            # tape) the start comes from the segment before, still on the tape
            prev = None if seg in ck else self._tape(self.taped[seg - 1])

            def entry(kg):
                q = kg - k0
                if q >= 0:
                    return tf[q] if self.tape_history else rebuilt(kg, ty[q + 1], tapes[q])
                if hist0 is not None:
                    return hist0[-q - 1]
                return rebuilt(kg, prev[0][K + q + 1], prev[4][K + q])

            # adjoint sweep through the segment
            for r in reversed(range(ks)):
                kg = k0 + r
                for key in [key for key in win if key > kg]:
                    pool.append(win.pop(key))
                fnew = entry(kg)
                hist = [entry(kg - j) for j in range(1, h + 1)]
                fp = (C.c_void_p * h)(*[t.data_ptr() for t in hist])
                pots_prev = tgrid[r - 1, 3] if r > 0 else (pots0 if pots0 is not None else prev[3][K - 1, 3])
                dg = self.dens_grad if (last and r == ks - 1 and self.state_observable is None) else None
                td.check(L.tdw_adj_backward_step(self.ctx.h, C.byref(A), self.it0 + kg, _p(ty[r]), fp,
                                                 _p(ty[r + 1]), _p(fnew), C.byref(tapes[r]), _p(pots_prev),
                                                 _p(dg)), "tdw_adj_backward_step")
            last = False
        self.ctx.sync()
        self.lambda_y0 = ly
        return theta.cpu().numpy().reshape(td.MAX_GAUSS, NPAR)

    def close(self):
        self.ckpts = []
        self.run = None
        self._tps = []
        self.ctx.close()
        if self.ctx_fwd is not None:
            self.ctx_fwd.close()


def beams_vext(theta, switch=None):
    """Gaussian beams exp(-((x-x_k)^2 + (y-y_k)^2) / (2 s_k^2)) with
    theta = (A_1, x_1, y_1, s_1, A_2, x_2, y_2, s_2, ...)."""
    th = np.asarray(theta, dtype=np.float64).reshape(-1, 4)
    gs = []
    for A, x, y, s in th:
        w = -0.5 / (s * s)
        gs.append((A, (x, y, 0.0), (w, w, 0.0)))
    return td.gauss_vext(gs, switch=switch)


def beams_grad(theta, g):
    """Chain rule from the Gaussian-slot gradient g (MAX_GAUSS x 7) to theta."""
    th = np.asarray(theta, dtype=np.float64).reshape(-1, 4)
    out = np.zeros_like(th)
    for k, (A, x, y, s) in enumerate(th):
        dw_ds = 1.0 / (s ** 3)
        out[k] = (g[k, 0], g[k, 1], g[k, 2], (g[k, 4] + g[k, 5]) * dw_ds)
    return out.ravel()


def _make_function():
    import torch

    class VextObservable(torch.autograd.Function):
        """O(theta): the observable of an AdjointProblem after its forward run
        with the Gaussian beams theta; backward by the adjoint sweep."""

        @staticmethod
        def forward(fctx, theta, problem):
            th = theta.detach().cpu().numpy()
            problem.set_vext(beams_vext(th))
            problem.forward()
            fctx.problem = problem
            fctx.theta = th
            return torch.tensor(problem.observable(), dtype=torch.float64)

        @staticmethod
        def backward(fctx, grad_out):
            g = fctx.problem.backward()
            gt = torch.as_tensor(beams_grad(fctx.theta, g), dtype=torch.float64)
            return grad_out * gt, None

    return VextObservable


_fn = None


def vext_observable(theta, problem):
    """torch.autograd entry point: O = vext_observable(theta, problem)."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(theta, problem)


def disk_weights(params, center, radius):
    """dens_grad of the particle number inside a disk (2D) / ball (3D) /
    interval (1D): VE * 1_{|r - c| < radius} on rho_a and rho_b."""
    dim = params.dim
    nx, ny, nz = params.nx, params.ny, params.nz
    x = params.dx * (np.arange(nx) - nx // 2)
    if dim == 1:
        r2 = (x - center[0]) ** 2
        ve = params.dx * params.dy * params.dz * ny * nz
    elif dim == 2:
        y = params.dy * (np.arange(ny) - ny // 2)
        X, Y = np.meshgrid(x, y, indexing="ij")
        r2 = ((X - center[0]) ** 2 + (Y - center[1]) ** 2).ravel()
        ve = params.dx * params.dy * params.dz * nz
    else:
        y = params.dy * (np.arange(ny) - ny // 2)
        zz = params.dz * (np.arange(nz) - nz // 2)
        X, Y, Z = np.meshgrid(x, y, zz, indexing="ij")
        r2 = ((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2).ravel()
        ve = params.dx * params.dy * params.dz
    m = (r2 < radius * radius).astype(np.float64) * ve
    N = m.size
    g = np.zeros(4 * N)
    g[:N] = m
    g[N:2 * N] = m
    return g


class PlaneCurrent:
    """O = sum over the plane z = z_index of (j_a + j_b) . e_axis dA: the
    particle current through a lattice plane normal to `axis` (0, 1, 2),
    with the densities' currents j_a = sum_n f_n Im(u_n* d u_n),
    j_b = -sum_n (1 - f_n) Im(v_n* d v_n) (the spectral first derivative with
    the Nyquist mode zero, as calculate_densities takes it).  Its seed for
    the adjoint, with O paired to a state change by Re sum conj(seed) dpsi:
    u_n: -i f_n (g d u_n + d(g u_n)), v_n: +i (1 - f_n)(g d v_n + d(g v_n)),
    g the plane's area weight (d is antisymmetric on the periodic lattice)."""

    def __init__(self, ctx, fbeta, axis=2, index=None):
        import torch
        self.ctx = ctx
        self.torch = torch
        p = ctx.params
        self.axis = axis
        shape = (p.nx, p.ny, p.nz)
        sp = (p.dx, p.dy, p.dz)
        idx = shape[axis] // 2 if index is None else index
        g = np.zeros(shape)
        area = sp[(axis + 1) % 3] * sp[(axis + 2) % 3]
        sl = [slice(None)] * 3
        sl[axis] = idx
        g[tuple(sl)] = area
        dev = "cuda:%d" % ctx.device
        self.g = torch.as_tensor(g.ravel()).to(dev)
        self.f = torch.as_tensor(np.asarray(fbeta, dtype=np.float64)).to(dev)

    def _d(self, a):
        """Spectral first derivative along the axis of the rows of a."""
        t = self.torch
        outs = [t.zeros_like(a) for _ in range(4)]
        self.ctx.compute_derivatives(a, outs[0], outs[1], outs[2], outs[3])
        self.ctx.sync()
        return outs[self.axis]

    def __call__(self, run):
        t = self.torch
        y = run.wf
        n = y.shape[0] // 2
        dy = self._d(y)
        gy = y * self.g[None, :]
        dgy = self._d(gy)
        f = self.f[:n, None]
        fm = 1.0 - f
        ja = (f * (y[:n].conj() * dy[:n]).imag).sum(dim=0)
        jb = -(fm * (y[n:].conj() * dy[n:]).imag).sum(dim=0)
        O = float(((ja + jb) * self.g).sum())
        seed = t.empty_like(y)
        seed[:n] = -1j * f * (self.g[None, :] * dy[:n] + dgy[:n])
        seed[n:] = 1j * fm * (self.g[None, :] * dy[n:] + dgy[n:])
        return O, seed


def barrier_vext(theta, axis=2):
    """A Gaussian barrier theta = (height, width) across the lattice centre,
    normal to `axis`: height exp(-s^2 / (2 width^2))."""
    h, w = float(theta[0]), float(theta[1])
    ww = [0.0, 0.0, 0.0]
    ww[axis] = -0.5 / (w * w)
    return td.gauss_vext([(h, (0.0, 0.0, 0.0), tuple(ww))])


def barrier_grad(theta, g, axis=2):
    """Chain rule from the Gaussian-slot gradient g (MAX_GAUSS x 7) to
    (height, width): the exponent coefficient is -1/(2 w^2)."""
    w = float(theta[1])
    return np.array([g[0, 0], g[0, 4 + axis] / (w ** 3)])
