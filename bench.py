"""Benchmark of the td-W-SLDA step on one MI355X.

Metric: quasiparticle-wavefunction steps per second (states x ABM steps /
wall time of the step loop, init and measurements excluded), FP64 and
split precision.  Default workload: the plain 3D uniform SLDA problem
(24^3 lattice, the uniform solution's 16308 states below its cutoff,
AB5AM5, FP64), then the same in split precision, then the trapped vortex
(32^3, 24576 states) in split precision and FP64.

Prints ONE JSON line (rank 0).  Extra objects:
  roofline     -- dominant kernel: algorithmic HBM bytes per launch over its
                  average launch time (HIP events on the library stream over
                  the timed region), against 8 TB/s;
  cpu_baseline -- the CPU FP64 oracle (oracle/tdw_oracle.c, OpenMP) timed on
                  the host on a bounded sample of the same workload.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def nhist(scheme):
    return {34: 3, 45: 4, 55: 5}[scheme]


def algorithmic_units(name, scheme, split=False):
    """A kernel's share of the compulsory traffic of one PECE step, in units
    of S (one state's u and v in FP64) per state and launch.

    The compulsory traffic of an FP64 step (AB5AM5) is B_qp = 13 S: the
    predictor reads y, f_1..f_5 and writes y_p and the partial corrector c
    (8 S); the corrector reads y_p and c and writes y (3 S); the history
    refresh reads y and writes the newest f (2 S).  In the split form the
    history entries and the partial corrector are FP32 (S/2 each), so the
    same passes need 5 S, 2.5 S and 1.5 S (B_qp = 9 S at AB5AM5).  FFTs,
    densities, norms and <H> are taken as fused into those passes, so
    whatever a kernel moves beyond its share counts against its achieved
    bandwidth."""
    h = nhist(scheme)
    if split:
        return {
            "f_pred": 1 + 0.5 * h + 1 + 0.5,
            "f_corr": 1 + 0.5 + 1,
            "f_hyz": 0.75,       # two launches per step: H of y_p (share 0), history refresh (share 1.5)
        }.get(name, 0)
    return {
        "f_pred": (h + 1) + 2,
        "f_corr": 3,
        "f_hyz": 1,          # two launches per step: H of y_p (share 0), history refresh (share 2)
        "abm_step1": (h + 1) + 2,
        "abm_step4": 3,
    }.get(name, 0)


def design_units(name, scheme, split=False):
    """Bytes (units of S per state and launch) this implementation's kernels
    must move: reads and writes of state-sized arrays (FP64 S; the split
    form's history, T / H scratch and the FP32 copies of the prediction and
    of the corrected state are S/2)."""
    h = nhist(scheme)
    if split:
        return {
            # y (twice: its pencil for the x kinetic part of f_1, then the
            # points), f_1..f_h; c, y_p (FP32), f_1 completed
            "f_pred": 2 + 0.5 * h + 1 + 0.5 + 0.5,
            "f_corr": 1 + 0.5 + 0.5 + 1 + 0.5,             # c, H y_p, y_p (FP32); y, y (FP32)
            "f_hyz": 1.0,                                   # y (FP32); H
            "densities": 0.5,                               # y (FP32)
            "f_fixup": 2,
        }.get(name, 0)
    return {
        "f_pred": (h + 2) + 3,   # read y (twice, as above), f_1..f_h; write y_p, c, f_1 (completed)
        "f_corr": 4,             # read y_p, H y_p, c; write y
        "f_hyz": 2,              # read y; write H y
        "densities": 1,          # read y
        "f_fixup": 2,
    }.get(name, 0)


def modular_units(name, split=False):
    """The ASLDA step's kernels (3D, FP64): each kernel's own compulsory
    reads and writes of state-sized arrays, in S per state and launch (its
    algorithmic bytes at its boundaries).  The folded step of tdw_step_fused
    (AB5AM5) is a_pred 9 (y, f1..f5 read; f1 completed, y_p and the partial
    corrector written), a_corr 4, and twice the density x and yz passes
    1 + 1 (gradients kept in LDS) and the H x and yz passes 2 + 3 -- 27 S
    against the 13 S compulsory for a whole step.  The modular step
    (tdw_step) on the same lattices: abm_step1 8, the corrector with the
    predictor's <H> subtraction folded in 4, twice norms 1 + scale 2, the
    same density and H passes, and the history entry's <H> subtraction 3 --
    35 S.  The modular kernels of other lattices: derivatives 3 + 5,
    densities 5, alpha_wf 2, lap(alpha psi) 2 + 3, apply_h 7 + 3,
    abm_step4 3."""
    if split:
        # the split form: FP32 history, prediction copy, H and T (0.5 S
        # each), FP64 y and partial corrector: a_pred 1 + 2.5 read, 0.5 +
        # 1 + 0.5 written; a_corr 1 + 0.5 + 0.5 read, 1 + 0.5 written; the
        # density and H passes on FP32 arrays -- 16 S per state-step
        return {"a_pred": 5.5, "a_corr": 3.5, "dens_gx": 0.5, "dens_gyz": 0.5, "ah_x": 1,
                "ah_yz": 1.5}.get(name, 0)
    return {"a_pred": 9, "a_corr": 4, "abm_step1": 8, "abm_step4": 3, "abm_step4q": 4, "norms": 1, "scale": 2,
            "dens_gx": 1, "dens_gyz": 1, "ah_x": 2,
            "ah_yz": 3, "sub_qpe": 3, "fft_outer_deriv": 3, "fft_inner_deriv": 5, "densities": 5, "alpha_wf": 2,
            "fft_outer": 2, "fft_inner": 3, "apply_h": 10}.get(name, 0)


def step_units(scheme, split=False):
    if split:
        return 0.5 * nhist(scheme) + 6.5   # 8, 8.5, 9 S
    return nhist(scheme) + 8     # B_qp: 11, 12, 13 S for AB3AM4, AB4AM5, AB5AM5


def copy_calibration(args):
    """One device copy of args.copy_calib_mib MiB at the end of a run: the
    reference dispatch for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    if not args.copy_calib_mib:
        return 0
    import torch
    nel = args.copy_calib_mib * 1024 * 1024 // 8
    a = torch.ones(nel, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    return nel * 8


def reserve_release(gb):
    """Allocates, touches and frees gb of device memory before the run's own
    allocations.  On a fresh box the first large allocations of a process
    have landed where the fused predictor runs ~15% slower (14.2 vs 12.2 ms
    per launch at 24^3 x 16384, constant over the run); allocations made after
    one reserve-and-release cycle have not.  No timed work is changed."""
    import torch
    x = torch.empty(int(gb * 1e9), dtype=torch.uint8, device="cuda:0")
    x.fill_(0)
    torch.cuda.synchronize()
    del x
    torch.cuda.empty_cache()


def run_gpu(args, prec):
    import numpy as np
    import torch
    from wslda_amd import problems, td

    dev = "cuda:0"
    t0 = time.time()
    w = problems.config2(torch, precision=prec, n_states=args.states, device=dev,
                         functional=td.SLDAE if args.sldae else td.ASLDA if args.aslda else td.SLDA,
                         sclgth=-1.1 if args.sldae else 0.0)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"], w["kt"], w["cnt"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=args.fused)
    del w["wf"]
    R.initial_measure()
    ctx.sync()
    setup_s = time.time() - t0
    it = 0
    meas = args.meas_every

    def run_steps(k):
        # the corrector of every meas-th step and of the loop's last step
        # computes gradients (tau, j), as the last step of a measurement does
        nonlocal it
        for q in range(k):
            last = ((it + 1) % meas == 0) or q == k - 1
            R.step(it, grad_last=last)
            it += 1

    run_steps(args.warmup)
    ctx.sync()
    # energy at the start of the timed loop: the warm-up ended on a gradient
    # step, so densities (tau included) and potentials belong to the state
    it_e0 = it
    e0 = R.energy(it)
    torch.cuda.synchronize()
    stream = torch.cuda.ExternalStream(ctx.stream_ptr())
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ctx.timing(True)
    ctx.timing_report()   # reset
    ev0.record(stream)
    wall0 = time.time()
    run_steps(args.steps)
    R.finalize()
    ev1.record(stream)
    ev1.synchronize()
    ctx.sync()
    wall = time.time() - wall0
    ms = ev0.elapsed_time(ev1)
    rep = ctx.timing_report()
    ctx.timing(False)
    e1 = R.energy(it)
    N = ctx.N
    qps = n * args.steps / (ms / 1000.0)
    # dominant kernel: largest total time in the timed loop
    dom = max(rep.items(), key=lambda kv: kv[1][0])
    dname, (dtot, dcnt) = dom
    davg = dtot / dcnt / 1000.0
    S = 2 * N * 16
    if args.aslda:
        dbytes = modular_units(dname, prec == td.SPLIT) * S * n
        moved = dbytes
    else:
        dbytes = algorithmic_units(dname, p.scheme, prec == td.SPLIT) * S * n
        moved = design_units(dname, p.scheme, prec == td.SPLIT) * S * n
    achieved = dbytes / davg / 1e9 if dbytes else None
    step_bytes = step_units(p.scheme, prec == td.SPLIT) * S * n
    out = dict(qps=qps, ms=ms, wall=wall, n=n, N=N, rep=rep, dname=dname, davg=davg, dbytes=dbytes,
               achieved=achieved, moved=moved, step_bytes=step_bytes, setup_s=setup_s, scheme=p.scheme,
               e0=e0.tolist(), e1=e1.tolist(), it_e0=it_e0, it_e1=it, name=w.get("name"), params=p,
               shape=(p.nx, p.ny, p.nz),
               uniform=w["uniform"], pots=w["pots"], fbeta=w["fbeta"], R=R, ctx=ctx)
    return out


def host_cores():
    """Every core the process may run on (the affinity mask, else
    os.cpu_count())."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(g, args):
    """Oracle on the host: a sample of the same workload (the first states of
    the GPU problem), a few PECE steps."""
    import ctypes as C
    import numpy as np
    from oracle import pyoracle as po
    threads = host_cores()
    po.lib().tdwo_set_threads(threads)
    p = g["params"]
    cfg = po.Cfg()
    C.memmove(C.addressof(cfg), C.addressof(p), C.sizeof(cfg))
    n, N = g["n"], g["N"]
    ns = args.cpu_states
    # the initial (noisy, normalised) state of the GPU run is not kept; use the
    # current GPU state: same lattice, same functional, same work per state
    wf = g["R"].wf
    sample = np.concatenate([wf[:ns].cpu().numpy(), wf[n:n + ns].cpu().numpy()])
    P = po.Problem(cfg, sample, g["fbeta"][:ns], pots=g["R"].pots.cpu().numpy())
    P.initial_measure()
    P.step(0, grad_last=0)   # warm
    t = time.time()
    k = 0
    while True:
        P.step(1 + k, grad_last=0)
        k += 1
        if time.time() - t > args.cpu_seconds or k >= args.cpu_steps:
            break
    dt = time.time() - t
    return dict(value=ns * k / dt, unit="qp-steps/s", cores=threads, kind="port",
                host_nproc=os.cpu_count(), cpu_model=cpu_model(),
                sample="%s lattice, %d of %d states, %d PECE steps after 1 warm-up, %d OpenMP threads "
                       "(all host cores), %.1f s"
                       % ("x".join(str(x) for x in (cfg.nx, cfg.ny, cfg.nz)), ns, n, k, threads, dt))


# HBM bytes per launch of the dominant kernel, keyed by (kernel, lattice,
# states, scheme, precision): FETCH_SIZE x 2 + WRITE_SIZE (KiB x 1024) of one
# launch, from the counter runs of `bench.py --steps 4 --warmup 1 --no-cpu
# --meas-every 1000 --copy-calib-mib 4096 --trap-states 0` (the 4096 MiB copy
# reads FETCH_SIZE = half its bytes and WRITE_SIZE = all of them)
MEASURED_TRAFFIC = {
    ("f_pred", (24, 24, 24), 16308, 55, "fp64"): (24629179 * 2 + 21140835) * 1024,
    ("f_pred", (24, 24, 24), 16308, 55, "split"): (15591706 * 2 + 14094742) * 1024,
}


def summarize(g, args, mode):
    S = 2 * g["N"] * 16
    out = {
        "value": g["qps"],
        "ms_per_step": g["ms"] / args.steps,
        "roofline": {"bound": "hbm", "kernel": g["dname"],
                     "achieved": g["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": (g["achieved"] / HBM_PEAK_GBS) if g["achieved"] else None,
                     "traffic": MEASURED_TRAFFIC.get((g["dname"], tuple(g["shape"]), g["n"], g["scheme"], mode)),
                     "avg_launch_ms": g["davg"] * 1e3, "algorithmic_bytes_per_launch": g["dbytes"],
                     "design_bytes_per_launch": g["moved"],
                     "step_algorithmic_bytes": g["step_bytes"],
                     "step_level_frac": g["step_bytes"] / (g["ms"] / args.steps / 1e3) / 1e9 / HBM_PEAK_GBS},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in sorted(g["rep"].items(), key=lambda kv: -kv[1][0])},
    }
    if args.traffic_bytes:
        out["roofline"]["traffic"] = args.traffic_bytes
    e0, e1 = sum(g["e0"][:7]), sum(g["e1"][:7])
    drift = abs(e1 - e0) / abs(e0)
    out["energy_drift"] = {"E_start": e0, "E_end": e1, "steps": g["it_e1"] - g["it_e0"], "rel": drift,
                           "rel_per_1000_steps": drift * 1000.0 / max(1, g["it_e1"] - g["it_e0"])}
    return out


def density_errors(d, ref, N, currents=True):
    """Relative error per density field of the gate (rho_a, rho_b, |nu|,
    tau_a, tau_b, j): max-abs difference over the field's own max-abs value
    (nu as a complex field; j = j_a + j_b, the total current, as one vector
    field).  currents=False leaves j out (a uniform state, whose currents are
    zero up to noise)."""
    import numpy as np
    out = {}
    nu, nur = d[0:2 * N:2] + 1j * d[1:2 * N:2], ref[0:2 * N:2] + 1j * ref[1:2 * N:2]
    out["nu"] = float(np.abs(nu - nur).max() / np.abs(nur).max())
    for k, i in (("rho_a", 2), ("tau_a", 3), ("rho_b", 7), ("tau_b", 8)):
        r = ref[i * N:(i + 1) * N]
        out[k] = float(np.abs(d[i * N:(i + 1) * N] - r).max() / np.abs(r).max())
    if currents:
        j, jr = d[4 * N:7 * N] + d[9 * N:12 * N], ref[4 * N:7 * N] + ref[9 * N:12 * N]
        out["j"] = float(np.abs(j - jr).max() / np.abs(jr).max())
    return out


def spin_current_errors(d, ref, N):
    """j_a and j_b each against its own maximum (reported beside the gate)."""
    import numpy as np
    out = {}
    for k, i0 in (("j_a", 4), ("j_b", 9)):
        r = ref[i0 * N:(i0 + 3) * N]
        out[k] = float(np.abs(d[i0 * N:(i0 + 3) * N] - r).max() / np.abs(r).max())
        out["|%s|_inf" % k] = float(np.abs(r).max())
    return out


def run_adjoint(args, calibrate=True):
    """2D SLDA 64x64 with 4096 states, two Gaussian beams, particle number in
    a disk of radius 10 dx after args.adj_steps steps, checkpoints every
    args.adj_k steps: time of the plain forward run, of the checkpointed
    forward + adjoint sweep, and a central-difference check (h = 1e-6) of
    two beam parameters."""
    import numpy as np
    import torch
    from wslda_amd import adjoint, problems
    s = problems.config4(n_states=args.adj_states)
    p = s["params"]
    theta = s["theta"]
    g = adjoint.disk_weights(p, s["center"], s["radius"])
    prob = adjoint.AdjointProblem(p, s["wf"], s["pots"], s["fbeta"], s["kt"], s["cnt"], nsteps=args.adj_steps,
                                  K=args.adj_k, dens_grad=g, taped_segments=args.adj_taped,
                                  tape_history=args.adj_tape_history == 1)
    prob.set_vext(adjoint.beams_vext(theta))
    prob.forward(keep_checkpoints=False)     # warm-up
    torch.cuda.synchronize()
    t0 = time.time()
    prob.forward(keep_checkpoints=False)
    torch.cuda.synchronize()
    t_fwd = time.time() - t0
    O = prob.observable()
    torch.cuda.reset_peak_memory_stats()
    t0 = time.time()
    prob.forward(keep_checkpoints=True)
    torch.cuda.synchronize()
    t_fwd_ck = time.time() - t0
    t0 = time.time()
    gslot = prob.backward()
    torch.cuda.synchronize()
    t_bwd = time.time() - t0
    peak = torch.cuda.max_memory_allocated()
    grad = adjoint.beams_grad(theta, gslot)
    fd = {}
    h = 1e-6
    for idx, name in ((0, "A_1"), (3, "width_1"), (5, "x_2")) if args.adj_fd else ():
        vals = []
        for sg in (1, -1):
            th = theta.copy()
            th[idx] += sg * h
            prob.set_vext(adjoint.beams_vext(th))
            prob.forward(keep_checkpoints=False)
            vals.append(prob.observable())
        fd[name] = {"adjoint": float(grad[idx]), "fd": (vals[0] - vals[1]) / (2 * h)}
    gmax = float(np.abs(grad).max())
    err = max(abs(v["adjoint"] - v["fd"]) for v in fd.values()) / gmax if fd else None
    n = prob.n
    prob.close()
    if calibrate:
        copy_calibration(args)
    return {
        "metric": "adjoint gradient cost / forward cost, and gradient error vs central finite differences",
        "value": (t_fwd_ck + t_bwd) / t_fwd, "unit": "x forward", "higher_is_better": False,
        "n_gpus": 1, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "td-SLDA 2D 64x64 lattice (8-point transverse box), %d qp states, AB5AM5, two "
                               "Gaussian beams (seeded), N in a disk r < 10 dx after %d steps, checkpoints every %d, "
                               "last %d segments taped in the forward run (y per step)" % (n, args.adj_steps, args.adj_k,
                                                                                           args.adj_taped)},
        "forward_s": t_fwd, "forward_with_checkpoints_s": t_fwd_ck, "backward_s": t_bwd,
        "backward_over_forward": t_bwd / t_fwd, "observable": O, "peak_device_GB": peak / 1e9,
        "grad_theta": [float(x) for x in grad], "fd_check": fd, "max_rel_err_vs_fd": err,
        "qp_steps_per_s_forward": n * args.adj_steps / t_fwd,
    }


def trap_leg(args, prec, nst, steps, timed, build=None, name="trap-vortex"):
    """A 3D workload (default: the 32^3 trapped vortex) in one precision:
    initial measurement, three steps (the last a gradient step), then
    `steps` timed steps with gradient steps every meas_every and at the end.
    Every array is released before returning."""
    import torch
    from wslda_amd import problems, td
    sys.stderr.write("%s: %s, %s states, %d steps\n" % (name, "split" if prec == td.SPLIT else "fp64",
                                                        "all" if nst < 0 else nst, steps))
    sys.stderr.flush()
    build = build or problems.config3
    w = build(torch, precision=prec, first=None if nst < 0 else nst)
    p = w["params"]
    n = w["wf"].shape[0] // 2
    ctx = td.TDContext(p, n, 0)
    ctx.set_states(w["fbeta"])
    R = td.TDRun(ctx, w["wf"], w["pots"], fused=True)
    del w["wf"]
    R.initial_measure()
    for it in range(3):
        R.step(it, grad_last=(it == 2))
    ctx.sync()
    e0 = R.energy(3)
    stream = torch.cuda.ExternalStream(ctx.stream_ptr())
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    if timed:
        ctx.timing(True)
        ctx.timing_report()
    ev0.record(stream)
    for it in range(3, 3 + steps):
        R.step(it, grad_last=((it + 1) % args.meas_every == 0) or it == 2 + steps)
    R.finalize()
    ev1.record(stream)
    ev1.synchronize()
    ms = ev0.elapsed_time(ev1)
    rep = ctx.timing_report() if timed else {}
    ctx.timing(False)
    e1 = R.energy(3 + steps)
    d = R.dens.cpu().numpy()
    N = ctx.N
    ctx.close()
    del R, ctx
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    return dict(qps=n * steps / (ms / 1e3), ms=ms / steps, n=n, N=N, dens=d, e0=e0, e1=e1, rep=rep)


def trap_pair(args):
    """Split precision and FP64 on the whole trapped-vortex workload (32^3,
    24586 states), one after the other (each alone fills most of the HBM):
    both throughputs, the speedup, and the split run's density errors
    against the FP64 run after the same steps from the same seeded start --
    the current j included, measured against its own maximum (the vortex
    carries real currents)."""
    from wslda_amd import td
    steps = args.trap_steps or args.steps
    sp = trap_leg(args, td.SPLIT, args.trap_states, steps, True)
    fp = trap_leg(args, td.FP64, args.trap_states, steps, True)
    err = density_errors(sp["dens"], fp["dens"], fp["N"], currents=True)
    E64, Esp = fp["e1"][:7].sum(), sp["e1"][:7].sum()
    drift = lambda r: abs(r["e1"][:7].sum() - r["e0"][:7].sum()) / abs(r["e0"][:7].sum())
    N = fp["N"]
    return {
        "workload": "td-SLDA 3D 32^3, harmonic trap omega = 0.1 eF on from t = 0, vortex line imprinted on u, the "
                    "uniform solution's %d plane-wave BdG states below its cutoff, seeded noise in a Gaussian "
                    "envelope, AB5AM5, %d timed steps" % (sp["n"], steps),
        "split_value": sp["qps"], "fp64_value": fp["qps"], "unit": "qp-steps/s",
        "split_ms_per_step": sp["ms"], "fp64_ms_per_step": fp["ms"], "speedup_vs_fp64": sp["qps"] / fp["qps"],
        "density_rel_err_vs_fp64": err, "max_density_rel_err": max(err.values()),
        "current_per_spin_rel_err": spin_current_errors(sp["dens"], fp["dens"], N),
        "j_inf_fp64": float(abs(fp["dens"][4 * N:7 * N] + fp["dens"][9 * N:12 * N]).max()),
        "energy_rel_diff_vs_fp64": abs(Esp - E64) / abs(E64),
        "energy_drift_fp64": drift(fp), "energy_drift_split": drift(sp), "drift_steps": steps,
        "split_kernels_ms_per_step": {k: v[0] / steps for k, v in sorted(sp["rep"].items(), key=lambda kv: -kv[1][0])},
        "fp64_kernels_ms_per_step": {k: v[0] / steps for k, v in sorted(fp["rep"].items(), key=lambda kv: -kv[1][0])},
    }


def tube_pair(args):
    """The tube Josephson junction (32 x 32 x 96, barrier 0.3 eF, pi phase
    step) in split precision and FP64 on one GPU's share of its 12062
    states (--tube-states, default 6031: half of them, a two-GPU split),
    with the split run's density errors against the FP64 run."""
    from wslda_amd import problems, td
    steps = args.trap_steps or args.steps
    sp = trap_leg(args, td.SPLIT, args.tube_states, steps, True, build=problems.config5, name="tube")
    fp = trap_leg(args, td.FP64, args.tube_states, steps, True, build=problems.config5, name="tube")
    err = density_errors(sp["dens"], fp["dens"], fp["N"], currents=True)
    E64, Esp = fp["e1"][:7].sum(), sp["e1"][:7].sum()
    N = fp["N"]
    return {
        "workload": "td-SLDA 3D tube 32x32x96, Gaussian barrier 0.3 eF of width 3 dx across z, pi phase step "
                    "imprinted on u, %d of the uniform solution's 12062 states below its cutoff (one GPU's "
                    "share), AB5AM5, %d timed steps" % (sp["n"], steps),
        "split_value": sp["qps"], "fp64_value": fp["qps"], "unit": "qp-steps/s",
        "split_ms_per_step": sp["ms"], "fp64_ms_per_step": fp["ms"], "speedup_vs_fp64": sp["qps"] / fp["qps"],
        "density_rel_err_vs_fp64": err, "max_density_rel_err": max(err.values()),
        "current_per_spin_rel_err": spin_current_errors(sp["dens"], fp["dens"], N),
        "energy_rel_diff_vs_fp64": abs(Esp - E64) / abs(E64),
        "split_kernels_ms_per_step": {k: v[0] / steps for k, v in sorted(sp["rep"].items(), key=lambda kv: -kv[1][0])},
        "fp64_kernels_ms_per_step": {k: v[0] / steps for k, v in sorted(fp["rep"].items(), key=lambda kv: -kv[1][0])},
    }


def run_tube(args):
    """--tube: the tube pair as its own JSON line."""
    t = tube_pair(args)
    return {
        "metric": "qp-wavefunction steps/s, split precision, tube Josephson junction",
        "value": t["split_value"], "unit": "qp-steps/s", "n_gpus": 1, "steps": args.trap_steps or args.steps,
        "ms_per_step": t["split_ms_per_step"], "higher_is_better": True, "dtype": "f32-operator/f64-state",
        "data": "synthetic", "config": {"workload": t["workload"] + ", split precision"}, "tube": t,
    }


def run_tube_adjoint(args):
    """--tube-adjoint: the tube's barrier-current gradient on one GPU's
    memory share (--tadj-states, default 360 states), --adj-steps steps
    (default 1000) with checkpoints every --adj-k (default 40): the forward
    run in split precision (checkpoints with FP32 history), the last segment
    taped and the backward sweep in FP64.  Times: the plain split forward,
    the fused FP64 forward, the checkpointed forward and the backward; the
    peak device memory."""
    import numpy as np
    import torch
    from wslda_amd import adjoint, problems, td
    nst = args.tadj_states
    # one taped segment fits either way at this size, so the tape keeps the
    # history entries rather than rebuilding them in the sweep
    full_tape = args.adj_tape_history != 0
    w = problems.config5(torch, precision=td.FP64, first=nst)
    theta = np.array([0.3 * w["eF"], 3.0])

    def problem(fp, taped):
        p = td.Params.from_buffer_copy(w["params"])
        prob = adjoint.AdjointProblem(p, w["wf"], w["pots"], w["fbeta"], nsteps=args.adj_steps, K=args.adj_k,
                                      taped_segments=taped, forward_precision=fp, tape_history=full_tape)
        prob.state_observable = adjoint.PlaneCurrent(prob.ctx, w["fbeta"], axis=2)
        prob.set_vext(adjoint.barrier_vext(theta))
        return prob

    def timed(fn):
        torch.cuda.synchronize()
        t0 = time.time()
        r = fn()
        torch.cuda.synchronize()
        return time.time() - t0, r

    # the fused FP64 forward, for reference
    prob = problem(td.FP64, 1)
    prob.forward(keep_checkpoints=False)
    t_f64, _ = timed(lambda: prob.forward(keep_checkpoints=False))
    O64 = prob.observable()
    prob.close()
    del prob
    torch.cuda.empty_cache()
    torch.cuda.reset_peak_memory_stats()
    prob = problem(td.SPLIT, args.adj_taped)
    prob.forward(keep_checkpoints=False)     # warm-up
    t_fwd, _ = timed(lambda: prob.forward(keep_checkpoints=False))
    O_split = prob.observable()
    t_ck, _ = timed(lambda: prob.forward(keep_checkpoints=True))
    O = prob.observable()
    ckpt_bytes = sum(t.numel() * t.element_size() for c in prob.ckpts for t in [c[1]] + list(c[2]) + [c[3]])
    t_bwd, gslot = timed(prob.backward)
    grad = adjoint.barrier_grad(theta, gslot)
    peak = torch.cuda.max_memory_allocated()
    n = prob.n
    prob.close()
    S = 2 * 16 * 32 * 32 * 96
    return {
        "metric": "adjoint gradient cost / split forward cost, tube Josephson junction",
        "value": (t_ck + t_bwd) / t_fwd, "unit": "x split forward", "higher_is_better": False,
        "n_gpus": 1, "dtype": "f32-operator/f64-state forward, f64 backward", "data": "synthetic",
        "config": {"workload": "td-SLDA 3D tube 32x32x96, barrier 0.3 eF width 3 dx, pi phase step, %d of the "
                               "uniform solution's 12062 states (one GPU's memory share), AB5AM5, %d steps, "
                               "checkpoints every %d, last %d segment(s) taped (%s); observable: the current "
                               "through the barrier plane at the final time" % (
                                   n, args.adj_steps, args.adj_k, args.adj_taped,
                                   "y and history entries" if full_tape else "y per step")},
        "split_forward_s": t_fwd, "fp64_fused_forward_s": t_f64, "forward_with_checkpoints_s": t_ck,
        "backward_s": t_bwd, "over_fp64_forward": (t_ck + t_bwd) / t_f64,
        "observable": O, "observable_split_forward": O_split, "observable_fp64_forward": O64,
        "grad_height_width": [float(x) for x in grad],
        "checkpoint_GB": ckpt_bytes / 1e9, "checkpoint_state_arrays": ckpt_bytes / (S * n),
        "peak_device_GB": peak / 1e9,
    }


def run_trap(args):
    """--trap: the trapped-vortex split / FP64 pair as its own JSON line."""
    t = trap_pair(args)
    return {
        "metric": "qp-wavefunction steps/s, split precision, trapped vortex",
        "value": t["split_value"], "unit": "qp-steps/s", "n_gpus": 1, "steps": t["drift_steps"],
        "ms_per_step": t["split_ms_per_step"], "higher_is_better": True, "dtype": "f32-operator/f64-state",
        "data": "synthetic", "config": {"workload": t["workload"] + ", split precision"}, "trap_vortex": t,
    }


def run_aslda(args):
    """ASLDA (density-dependent effective mass, current corrections) on the
    3D uniform workload: the folded step (tdw_step_fused) with gradients on
    both half-steps, FP64, its dominant kernel against the HBM roofline, and
    the oracle's ASLDA step on a sample of the same states."""
    from wslda_amd import td
    if args.reserve_gb > 0:
        reserve_release(args.reserve_gb)
    g = run_gpu(args, td.FP64)
    head = summarize(g, args, "fp64")
    roof = head["roofline"]
    roof["unit_note"] = ("achieved = the kernel's own compulsory state-array reads and writes per launch "
                         "/ average launch time")
    res = {
        "metric": "qp-wavefunction steps/s, %s (gradients every half-step), FP64" % ("SLDAE" if args.sldae
                                                                                    else "ASLDA"),
        "value": head["value"], "unit": "qp-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"], "higher_is_better": True, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "td-%s 3D uniform, 24x24x24 lattice, %d qp states, AB5AM5, plane-wave BdG "
                               "+ seeded 1e-6 noise, folded step" % ("SLDAE (a kF = -1.1)" if args.sldae
                                                                      else "ASLDA", g["n"])},
        "roofline": roof, "kernels_ms_per_step": head["kernels_ms_per_step"],
        "energy_drift": head["energy_drift"],
        "step_state_traffic_S": sum(modular_units(k) * v[1] for k, v in g["rep"].items()) / args.steps,
        "step_compulsory_S": step_units(g["scheme"]),
    }
    if args.mode == "both":
        # the same workload in split precision, from the same start
        gs = run_gpu(args, td.SPLIT)
        sp = summarize(gs, args, "split")
        err = density_errors(gs["R"].dens.cpu().numpy(), g["R"].dens.cpu().numpy(), g["N"], currents=False)
        e64, es = sum(g["e1"][:7]), sum(gs["e1"][:7])
        res["split"] = {"value": sp["value"], "ms_per_step": sp["ms_per_step"],
                        "speedup_vs_fp64": sp["value"] / res["value"], "dtype": "f32-operator/f64-state",
                        "density_rel_err_vs_fp64": err, "max_density_rel_err": max(err.values()),
                        "energy_rel_diff_vs_fp64": abs(es - e64) / abs(e64), "energy_drift": sp["energy_drift"],
                        "roofline": sp["roofline"], "kernels_ms_per_step": sp["kernels_ms_per_step"],
                        "step_state_traffic_S": sum(modular_units(k, True) * v[1]
                                                    for k, v in gs["rep"].items()) / args.steps}
        gs["ctx"].close()
        del gs
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(g, args)
    g["ctx"].close()
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--states", type=int, default=None,
                    help="states of the 24^3 workload: the uniform solution's whole state set below its "
                         "cutoff (16308) by default; another value must equal that set's size")
    ap.add_argument("--mode", choices=["both", "fp64", "split"], default="both",
                    help="both: the FP64 run (value) and the split-precision run on the same initial state, "
                         "with the split run's density errors against FP64")
    ap.add_argument("--fused", type=int, default=1)
    ap.add_argument("--reserve-gb", type=float, default=120.0,
                    help="reserve, touch and release this much device memory before the run's arrays are "
                         "allocated (0: off)")
    ap.add_argument("--meas-every", type=int, default=100,
                    help="steps between measurements (the corrector of the last step of each computes gradients)")
    ap.add_argument("--cpu-states", type=int, default=256)
    ap.add_argument("--cpu-steps", type=int, default=200)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--traffic-bytes", type=float, default=0.0,
                    help="HBM bytes per launch of the dominant kernel from FETCH_SIZE + WRITE_SIZE counter runs "
                         "(overrides the measured table)")
    ap.add_argument("--copy-calib-mib", type=int, default=0,
                    help="after the run, one device-to-device copy of this many MiB (a plain copy kernel "
                         "of known traffic, to scale FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--adjoint", action="store_true",
                    help="instead of the step benchmark: the adjoint workload (2D 64x64, 4096 states, two beams)")
    ap.add_argument("--adj-steps", type=int, default=500)
    ap.add_argument("--adj-k", type=int, default=50)
    ap.add_argument("--adj-states", type=int, default=4096)
    ap.add_argument("--adj-fd", type=int, default=1, help="central-difference check of three parameters (0: off)")
    ap.add_argument("--adj-taped", type=int, default=9,
                    help="segments the forward run tapes directly (each K+1 state arrays of HBM: y per step; "
                         "the backward sweep rebuilds the history entries)")
    ap.add_argument("--trap", action="store_true",
                    help="instead of the step benchmark: split precision on the trapped vortex (32^3)")
    ap.add_argument("--trap-states", type=int, default=-1,
                    help="states of the trapped-vortex pair in the default line (-1: the workload's whole "
                         "state set, 24586 states; 0: leave the pair out; n: its first n states)")
    ap.add_argument("--trap-steps", type=int, default=0, help="timed steps of the trapped-vortex pair "
                                                              "(0: --steps)")
    ap.add_argument("--tube", action="store_true",
                    help="the tube Josephson junction (32x32x96) in split precision and FP64 as its own line")
    ap.add_argument("--tube-states", type=int, default=6031,
                    help="states of the tube runs (one GPU's share of its 12062; -1: all)")
    ap.add_argument("--line-tube-steps", type=int, default=50,
                    help="timed steps of the tube pair in the default line (0: leave it out)")
    ap.add_argument("--line-adjoint", type=int, default=1,
                    help="the 2D adjoint workload (cost and FD check) in the default line (0: leave it out)")
    ap.add_argument("--tube-adjoint", action="store_true",
                    help="instead of the step benchmark: the tube's barrier-current gradient (split forward, "
                         "FP64 backward) on one GPU's memory share; --adj-steps, --adj-k and --adj-taped apply "
                         "(defaults here 1000, 40, 1)")
    ap.add_argument("--tadj-states", type=int, default=360, help="states of --tube-adjoint")
    ap.add_argument("--adj-tape-history", type=int, default=-1,
                    help="1: the adjoint tape keeps the history entries (2 state arrays per step); 0: y alone, "
                         "the entries rebuilt in the backward sweep (-1: the workload's default)")
    ap.add_argument("--aslda", action="store_true",
                    help="instead of the step benchmark: the ASLDA step on the 3D uniform workload (24^3)")
    ap.add_argument("--sldae", action="store_true",
                    help="as --aslda with the SLDAE functional (scattering length a kF = -1.1)")
    args = ap.parse_args()
    if args.sldae:
        args.aslda = True
    if args.gpus != 1:
        sys.stderr.write("bench.py: this build runs on one GPU only (the density all-reduce across "
                         "GPUs is a design, not code); run with --gpus 1\n")
        sys.exit(2)
    if args.adjoint:
        print(json.dumps(run_adjoint(args)))
        return
    if args.trap:
        print(json.dumps(run_trap(args)))
        return
    if args.tube:
        print(json.dumps(run_tube(args)))
        return
    if args.tube_adjoint:
        if "--adj-steps" not in sys.argv:
            args.adj_steps = 1000
        if "--adj-k" not in sys.argv:
            args.adj_k = 40
        if "--adj-taped" not in sys.argv:
            args.adj_taped = 1
        print(json.dumps(run_tube_adjoint(args)))
        return
    if args.aslda:
        print(json.dumps(run_aslda(args)))
        return
    from wslda_amd import td
    primary = "split" if args.mode == "split" else "fp64"
    if args.reserve_gb > 0:
        reserve_release(args.reserve_gb)
    g = run_gpu(args, td.SPLIT if primary == "split" else td.FP64)
    n = g["n"]
    head = summarize(g, args, primary)
    res = {
        "metric": "qp-wavefunction steps/s, FP64 vs split-prec at rel density err <=1e-6",
        "value": head["value"],
        "unit": "qp-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if primary == "fp64" else "f32-operator/f64-state",
        "data": "synthetic",
        "config": {"workload": "td-SLDA 3D uniform, %dx%dx%d lattice, %d qp states (the uniform solution's whole "
                               "state set below its cutoff; 16384 is not the size of any such set on this "
                               "lattice), AB5AM5, plane-wave BdG + seeded 1e-6 noise, gradients every %d steps"
                               % (24, 24, 24, n, args.meas_every),
                   "lattice": [24, 24, 24], "states": n, "scheme": "AB5AM5", "precision": primary,
                   "fused": bool(args.fused)},
        "roofline": head["roofline"],
        "kernels_ms_per_step": head["kernels_ms_per_step"],
        "energy_drift": head["energy_drift"],
    }
    if args.mode == "both":
        gs = run_gpu(args, td.SPLIT)
        sp = summarize(gs, args, "split")
        dens64 = g["R"].dens.cpu().numpy()
        denss = gs["R"].dens.cpu().numpy()
        err = density_errors(denss, dens64, g["N"], currents=False)
        e64, es = sum(g["e1"][:7]), sum(gs["e1"][:7])
        res["split"] = {"value": sp["value"], "ms_per_step": sp["ms_per_step"],
                        "speedup_vs_fp64": sp["value"] / res["value"],
                        "dtype": "f32-operator/f64-state",
                        "density_rel_err_vs_fp64": err, "max_density_rel_err": max(err.values()),
                        "energy_rel_diff_vs_fp64": abs(es - e64) / abs(e64),
                        "after_steps": gs["it_e1"],
                        "energy_drift": sp["energy_drift"], "roofline": sp["roofline"],
                        "kernels_ms_per_step": sp["kernels_ms_per_step"]}
        gs["ctx"].close()
        del gs
    if not args.no_cpu and primary == "fp64":
        res["cpu_baseline"] = cpu_baseline(g, args)
    g["ctx"].close()
    del g
    if args.mode == "both" and args.trap_states != 0:
        import torch
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        res["trap_vortex"] = trap_pair(args)
    if args.mode == "both" and args.line_tube_steps > 0:
        import copy
        import torch
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        a2 = copy.copy(args)
        a2.trap_steps = args.line_tube_steps
        res["tube"] = tube_pair(a2)
    if args.mode == "both" and args.line_adjoint:
        import torch
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        adj = run_adjoint(args, calibrate=False)
        res["adjoint"] = {k: adj[k] for k in ("config", "value", "unit", "forward_s", "forward_with_checkpoints_s",
                                              "backward_s", "fd_check", "max_rel_err_vs_fd",
                                              "peak_device_GB")}
    nb = copy_calibration(args)
    if nb:
        res["copy_calib_bytes"] = nb
    print(json.dumps(res))


if __name__ == "__main__":
    main()
