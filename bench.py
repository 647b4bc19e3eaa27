"""Benchmark of the fused D2Q9 time step on one MI355X.

Workload (one "step" = one lattice iteration of the whole domain): channel
8192 x 2048 with 64 axis-aligned square obstacles (IBB links, drag/lift terms
every iteration), uniform Zou-He inlet u = 0.04 (ramp sigma = 1000
iterations), Zou-He zero-pressure outlet, no-slip walls, Re = 2000 on the
height, TRT with Lambda = 1/4.  Inputs are synthetic and resident in HBM
before the timed region.

The headline value is the MLUPS of the FP16-offset storage form (f - w stored
as IEEE half, FP32 arithmetic).  The FP64, FP32, BF16-offset, f16e (half-
precision f - w - E(v) with the node velocity v in a half pair), f16e8 (the
same with v as an 8-bit pair) and Q16 (16-bit fixed point of f - w) forms are
timed on the same workload and reported under "forms", each with its
fraction of the 8 TB/s HBM roofline (algorithmic bytes per lattice update:
18 x storage bytes, 44 for f16e, 40 for f16e8), its drag error against the
FP64 run and its velocity-field L2 error.  Further sections: the 200x200
cavity, the Turek cylinder through the drop-in lattice/run, the Turek
cylinder per form (shedding statistics), a 256-cavity batch and the adjoint;
optionally the reference's Turek app against its published table.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms f16,f64,...] [--sections ...]
"""
import argparse
import json
import math
import os
import sys
import time
import types

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
STORE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2, "f16e": 2, "q16": 2, "f16e8": 2}
# algorithmic bytes per lattice update: nine populations read and written
# (18 x storage bytes); f16e and f16e8 also read and write their velocity word
BYTES_PER_UPDATE = {f: 18 * b for f, b in STORE_BYTES.items()}
BYTES_PER_UPDATE["f16e"] = 18 * 2 + 8
BYTES_PER_UPDATE["f16e8"] = 18 * 2 + 4


def build_case(ny=2048, aspect=4, n_obs=64, seed=0):
    """Geometry and parameters of the obstacle-array workload (host only)."""
    from lbm_amd.app.configs import obstacle_array
    from lbm_amd import geometry
    from lbm_amd.shapes import generate_shape
    app = obstacle_array(ny=ny, aspect=aspect, n_obs=n_obs, seed=seed)
    lat = types.SimpleNamespace(nx=app.nx, ny=app.ny, x_min=app.x_min, x_max=app.x_max,
                                y_min=app.y_min, y_max=app.y_max, dx=app.dx, IBB=True,
                                lattice=np.zeros((app.nx, app.ny)))
    out_dir = os.path.join("/tmp", "lbm_amd_bench_shapes") + "/"
    for k, o in enumerate(app.obstacles):
        s = generate_shape(o.n_pts, o.pos, o.type, o.size, o.name, o.n_spts, out_dir)
        o.set_polygon(s.curve_pts)
        o.set_tag(k + 1)
        area, bnd, ibb = geometry.add_obstacle(lat, o, verbose=False)
        o.fill(area, bnd, ibb)
    tau_p = app.tau_lbm
    tau_m = 0.25 / (tau_p - 0.5) + 0.5
    return app, lat, 1.0 / tau_p, 1.0 / tau_m


def inlet_rows(app, n_rows):
    """Wall arrays of iterations 0..n_rows-1 (uniform ramped inlet)."""
    nx, ny = app.nx, app.ny
    zeros_x = np.zeros((2, nx))
    zeros_y = np.zeros((2, ny))
    rho_r = np.full(ny, app.rho_lbm)
    rows = []
    prev = None
    for it in range(n_rows):
        ret = (1.0 - math.exp(-it ** 2 / (2.0 * app.sigma ** 2)))
        ul = np.zeros((2, ny))
        ul[0, :] = ret * app.u_lbm
        if prev is not None and np.array_equal(prev, ul):
            ul = prev
        prev = ul
        rows.append((zeros_x, ul, zeros_x, zeros_y, rho_r))
    return rows


# Q16 storage: velocity bound of the range of f - w (lbm_amd.domain.q16_range),
# in units of the case's inlet speed u_lbm
Q16_U_FACTOR = {"array": 2.5, "turek": 1.5}


def q16_for(case, u_lbm):
    from lbm_amd.domain import q16_range
    return q16_range(Q16_U_FACTOR[case] * u_lbm)


# the obstacle-array configuration runs 10 000 iterations: the accuracy of
# every form is measured there, whatever the number of timed iterations
ACC_ITERATIONS = 10000


def chunks(it, end, graph_len):
    """(first iteration, length) of the graph replays covering it..end-1."""
    out = []
    while it < end:
        n = min(graph_len, end - it)
        out.append((it, n))
        it += n
    return out


def run_form(app, lat, om_p, om_m, storage, steps, warmup, graph_len, cpt=0, vector_width=0,
             acc_iterations=ACC_ITERATIONS):
    """One storage form on the obstacle array: iterations 1..warmup-1 untimed,
    the next `steps` timed (their graphs captured and uploaded beforehand,
    so the timed region holds launches only), then on to the configuration's
    length untimed; the drag of its last 1000 iterations, the final velocity
    field and the density sum are returned for the accuracy comparison."""
    import torch
    from lbm_amd import _lib
    from lbm_amd.domain import FusedDomain, pack_inlets
    nx, ny = app.nx, app.ny
    total = max(warmup + steps, acc_iterations)    # iterations 0 .. total-1
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, om_p, om_m, bc, [(o.boundary, o.ibb) for o in app.obstacles],
                      ibb=True, storage=storage, ring=total + 3, cells_per_thread=cpt,
                      vector_width=vector_width,
                      q16_range=q16_for("array", app.u_lbm) if storage == "q16" else None)
    pool, offs = pack_inlets(inlet_rows(app, total + 2))
    dom.set_inlets(0, offs, pool)
    w = torch.tensor([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4, dtype=torch.float64)
    g0 = w[:, None, None].expand(9, nx, ny).contiguous().cuda()
    dom.init(g0)
    it = 1
    for i0, n in chunks(it, max(warmup, 1), graph_len):
        dom.steps(i0, n)
    it = max(warmup, 1)
    timed = chunks(it, it + steps, graph_len)
    for i0, n in timed:
        dom.prepare(i0, n)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    for i0, n in timed:
        dom.steps(i0, n)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    it += steps
    # untimed: on to the configuration's last iteration, with its moments
    last = total - 1
    for i0, n in chunks(it, last, graph_len):
        dom.steps(i0, n)
    rho = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
    u = torch.empty((2, nx, ny), dtype=torch.float64, device="cuda")
    dom.step(last, rho, u)
    dom.drag_only(last)
    n_avg = min(1000, last)
    drag = dom.drag_read(last - n_avg + 1, n_avg)
    g = torch.empty((9, nx, ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    mass = dom.mass(last)
    dom.epilogue(last, g, gup, rho, u)
    ux = u.cpu().numpy()
    speed = float(np.sqrt((ux ** 2).sum(axis=0)).max())
    del g, gup, rho, u, g0
    dbytes = dom.device_bytes()
    dom.close()
    torch.cuda.synchronize()
    mlups = nx * ny * steps / (ms * 1e-3) / 1e6
    return dict(ms=ms, wall=wall, mlups=mlups, drag=drag, u=ux, mass=mass, device_bytes=dbytes,
                max_speed=speed,
                n_links=int(sum(len(o.boundary) for o in app.obstacles)), iterations=total)


def copy_ceiling(gib=2, reps=10):
    """HBM copy ceiling: the copy probes (16-byte accesses, plain and
    non-temporal) moving `gib` GiB with 4096 and 32 768 work-groups (read +
    write bytes / time); the best configuration is the ceiling."""
    import torch
    from lbm_amd import _lib
    L = _lib.load()
    nbytes = gib << 30
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    st = _lib.stream_ptr()
    runs = {}
    for name, fn in (("k_copy16", L.lbm_copy_probe), ("k_copy16_nt", L.lbm_copy_probe_nt)):
        for blocks in (4096, 32768):
            for _ in range(3):
                _lib.check(fn(b.data_ptr(), a.data_ptr(), nbytes, blocks, st))
            torch.cuda.synchronize()
            s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            s.record()
            for _ in range(reps):
                _lib.check(fn(b.data_ptr(), a.data_ptr(), nbytes, blocks, st))
            e.record()
            torch.cuda.synchronize()
            runs[f"{name}/{blocks}"] = round(2 * nbytes * reps / (s.elapsed_time(e) * 1e-3) / 1e9,
                                             1)
    del a, b
    best = max(runs, key=runs.get)
    return dict(GBps=runs[best], best=best, runs=runs)


def cpu_baseline(app, lat, om_p, om_m, iterations):
    """The oracle (strict-order C restatement of the reference kernels, node
    loops on OMP_NUM_THREADS threads) on the same workload for a few
    iterations, driven in the reference's call order (run.py:24-54)."""
    from oracle import oracle
    nx, ny = app.nx, app.ny
    w = np.array([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4)
    g = np.ascontiguousarray(np.broadcast_to(w[:, None, None], (9, nx, ny)))
    geq = np.zeros_like(g)
    gup = np.zeros_like(g)
    rho = np.ones((nx, ny))
    u = np.zeros((2, nx, ny))
    rows = inlet_rows(app, iterations)
    t0 = time.perf_counter()
    for it in range(iterations):
        ub, ul, ut, ur, rr = rows[it]
        oracle.macro(g, rho, u)
        oracle.equilibrium(u, rho, geq)
        oracle.col_str(g, geq, gup, om_p, om_m)
        for o in app.obstacles:
            oracle.bounce_back_obstacle(True, o.boundary, o.ibb, gup, g)
        oracle.zou_he("bottom", u, rho, g, side=np.ascontiguousarray(ub))
        oracle.zou_he("left", u, rho, g, side=np.ascontiguousarray(ul))
        oracle.zou_he("top", u, rho, g, side=np.ascontiguousarray(ut))
        oracle.zou_he("right_pressure", u, rho, g, side=np.ascontiguousarray(ur),
                      rho_right=np.ascontiguousarray(rr))
        for c in ("bottom_left", "top_left", "top_right", "bottom_right"):
            oracle.zou_he(c, u, rho, g)
        for o in app.obstacles:
            oracle.drag_sums(o.boundary, gup, g)
    dt = time.perf_counter() - t0
    return nx * ny * iterations / dt / 1e6, dt, oracle.threads()


def turek_case(ny=200):
    """Schaefer-Turek channel (turek200 app): geometry, rates, inlet profile."""
    from lbm_amd import geometry
    from lbm_amd.app.configs import turek200
    from lbm_amd.shapes import generate_shape
    app = turek200(iterations=200000)
    app.L_lbm = ny
    app.compute_lbm_parameters()
    app.obstacles[0].size = app.r_cyl
    app.it_max = 199999
    lat = types.SimpleNamespace(nx=app.nx, ny=app.ny, x_min=app.x_min, x_max=app.x_max,
                                y_min=app.y_min, y_max=app.y_max, dx=app.dx, IBB=True,
                                lattice=np.zeros((app.nx, app.ny)))
    o = app.obstacles[0]
    sh = generate_shape(o.n_pts, o.pos, o.type, o.size, o.name, o.n_spts,
                        "/tmp/lbm_amd_bench_shapes/")
    o.set_polygon(sh.curve_pts)
    o.set_tag(1)
    o.fill(*geometry.add_obstacle(lat, o, verbose=False))
    prof = np.zeros((2, app.ny))
    for j in range(app.ny):
        prof[:, j] = app.poiseuille(lat_coords(app, 0, j))
    return app, o, prof


def lat_coords(app, i, j):
    dx = (app.x_max - app.x_min) / (app.nx - 1)
    dy = (app.y_max - app.y_min) / (app.ny - 1)
    return [app.x_min + i * dx, app.y_min + j * dy]


def run_turek(app, o, prof, storage, iterations):
    """Turek channel through a one-case batch (separable inlet); returns the
    per-iteration Cd/Cl (reference convention), the final velocity field and
    the timed MLUPS."""
    import torch
    from lbm_amd import _lib
    from lbm_amd.batch import Batch
    tau_m = 0.25 / (app.tau_lbm - 0.5) + 0.5
    case = dict(nx=app.nx, ny=app.ny, om_p=1.0 / app.tau_lbm, om_m=1.0 / tau_m,
                bc=[_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE],
                profiles=[None, prof, None, None, np.ones(app.ny)], links=o.boundary,
                ibb=o.ibb, ibb_on=np.ones(len(o.boundary), np.uint8),
                obs_offsets=[0, len(o.boundary)])
    b = Batch([case], storage=storage, ring=iterations + 2,
              q16_range=q16_for("turek", app.u_lbm) if storage == "q16" else None)
    amp = np.zeros((iterations + 1, 1, 5))
    for it in range(iterations + 1):
        ret = (1.0 - math.exp(-it ** 2 / (2.0 * app.sigma ** 2)))
        amp[it, 0, 1] = ret * app.u_lbm
        amp[it, 0, 4] = app.rho_lbm
    b.set_amplitudes(0, amp)
    b.init()
    last = iterations - 1
    plan = chunks(1, last, 1000)
    for i0, n in plan:
        b.prepare(i0, n)
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    for i0, n in plan:
        b.steps(i0, n)
    s1.record()
    torch.cuda.synchronize()
    ms = s0.elapsed_time(s1)
    dom = b.case(0)
    rho = torch.empty((app.nx, app.ny), dtype=torch.float64, device="cuda")
    u = torch.empty((2, app.nx, app.ny), dtype=torch.float64, device="cuda")
    g = torch.empty((9, app.nx, app.ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    dom.step(last, rho, u)
    dom.drag_only(last)
    fxy = dom.drag_read(0, iterations)[:, 0, :]
    dom.epilogue(last, g, gup, rho, u)
    ux = u.cpu().numpy()
    b.close()
    scale = -2.0 / (app.rho_lbm * app.D_lbm * app.u_avg ** 2)
    return dict(cd=fxy[:, 0] * scale, cl=fxy[:, 1] * scale, u=ux,
                mlups=app.nx * app.ny * (last - 1) / (ms * 1e-3) / 1e6)


def shedding_stats(app, cd, cl, tail=50000):
    """Statistics of the last `tail` iterations: max |Cd|, max |Cl|, mean Cd,
    the standard deviation of Cl and the Strouhal number f D / U, f from the
    peak of the Hann-windowed lift spectrum refined by a parabola through the
    log-magnitudes of the peak bin and its neighbours (robust to the jitter a
    16-bit storage adds, which mean-crossing counts are not)."""
    cd, cl = np.asarray(cd[-tail:]), np.asarray(cl[-tail:])
    c = cl - cl.mean()
    sp = np.abs(np.fft.rfft(c * np.hanning(len(c))))
    k = int(np.argmax(sp[1:-1])) + 1
    a, b, g = np.log(sp[k - 1] + 1e-300), np.log(sp[k] + 1e-300), np.log(sp[k + 1] + 1e-300)
    den = a - 2 * b + g
    kk = k + (0.5 * (a - g) / den if den != 0 else 0.0)
    st = kk / len(c) * app.D_lbm / app.u_avg
    return dict(max_cd=float(np.abs(cd).max()), max_cl=float(np.abs(cl).max()),
                mean_cd=float(cd.mean()), std_cl=float(c.std()), strouhal=float(st))


def section_turek(forms, iterations, log):
    app, o, prof = turek_case()
    out = {}
    ref = None
    for f in forms:
        log(f"# turek200 {f}: {iterations} iterations")
        r = run_turek(app, o, prof, f, iterations)
        st = shedding_stats(app, r["cd"], r["cl"], tail=min(50000, iterations // 2))
        entry = dict(mlups=round(r["mlups"], 1), **st)
        if f == "f64":
            ref = (r, st)
        elif ref is not None:
            r64, st64 = ref
            for key in ("max_cd", "max_cl", "mean_cd", "std_cl", "strouhal"):
                entry[key + "_rel_err"] = abs(st[key] - st64[key]) / abs(st64[key])
        out[f] = entry
        log(f"# turek200 {f}: {entry}")
    out["config"] = "turek200: 1073x200, IBB cylinder D=48, u_max=0.05, Re=100, %d iterations" \
        % iterations
    return out


def section_dropin(iterations, log):
    """The Turek channel through the drop-in lattice/run -- the turek200 app's
    own hooks (drag/lift, moving averages, drag_lift file every iteration),
    wall inputs from device amplitude tables, chunks pipelined with the host
    replay -- timed over the run loop (lattice.loop_time), FP64."""
    import tempfile
    from lbm_amd.app.configs import turek200
    from lbm_amd.lattice import lattice
    from lbm_amd.run import run
    app = turek200("f64", iterations=iterations)
    app.verbose = False
    app.write_png = False
    app.results_dir = tempfile.mkdtemp(prefix="lbm_amd_dropin_") + "/"
    lat = lattice(app)
    last = run(lat, app)
    loop = lat.loop_time
    out = dict(config="turek200 app through lbm_amd.lattice + lbm_amd.run, FP64, %d iterations"
                      % (last + 1),
               iterations=last + 1, loop_s=round(loop, 3),
               mlups=round(app.nx * app.ny * (last + 1) / loop / 1e6, 1),
               host_replay="chunked (turek.replay_chunk), overlapped with the next chunk")
    log(f"# dropin: {out}")
    return out


# the reference's published Turek results (README.md, "Turek benchmark"): ny -> Re -> (Cd, Cl)
PUBLISHED_TUREK = {100: {20: (5.7111, 0.0285), 100: (3.5409, 1.0696)},
                   200: {20: (5.6250, 0.0212), 100: (3.2959, 1.0253)}}
SCHAEFER_TUREK = {20: (5.5800, 0.0107), 100: (3.2300, 1.0000)}


def section_published(log, cap=3000000):
    """The reference's Turek app itself (lbm/src/app/turek.py: stop = 'obs', obs_cv_ct = 1e-3,
    obs_cv_nb = 1000) through the drop-in lattice/run, FP64, at ny = 100 and 200, Re = 20
    (2D-1) and 100 (2D-2), against the values the reference's README publishes: the final
    Cd/Cl for 2D-1, the maxima of |Cd| and |Cl| for 2D-2 (the code's Cd is negative: drag
    along +x with links pointing into the solid; magnitudes are compared)."""
    import tempfile
    from lbm_amd.app.turek import turek
    from lbm_amd.lattice import lattice
    from lbm_amd.run import run
    out = {}
    for ny, by_re in PUBLISHED_TUREK.items():
        for re_, (cd_pub, cl_pub) in by_re.items():
            app = turek()
            app.L_lbm = ny
            app.Re_lbm = float(re_)
            app.compute_lbm_parameters()
            app.obstacles[0].size = app.r_cyl
            app.it_max = cap
            app.verbose = False
            app.write_png = False
            app.results_dir = tempfile.mkdtemp(prefix="lbm_amd_turek_") + "/"
            lat = lattice(app)
            last = run(lat, app)
            cd = np.abs(app.drag_buff.buff[2:])
            cl = np.abs(app.lift_buff.buff[2:])
            tail = slice(int(len(cd) * 4 // 5), None)
            if re_ == 20:
                got = (float(cd[-1]), float(cl[-1]))
            else:
                got = (float(cd[tail].max()), float(cl[tail].max()))
            key = f"ny{ny}_re{re_}"
            out[key] = dict(iterations=last + 1, converged=bool(app.drag_buff.obs_cv and
                                                                 app.lift_buff.obs_cv),
                            loop_s=round(lat.loop_time, 2), cd=got[0], cl=got[1],
                            cd_published=cd_pub, cl_published=cl_pub,
                            cd_rel_diff=(got[0] - cd_pub) / cd_pub,
                            cl_rel_diff=(got[1] - cl_pub) / cl_pub,
                            cd_schaefer_turek=SCHAEFER_TUREK[re_][0],
                            cl_schaefer_turek=SCHAEFER_TUREK[re_][1],
                            max_over_whole_run=(float(cd.max()), float(cl.max())))
            log(f"# published {key}: {out[key]}")
    out["config"] = ("turek app (reference defaults, stop='obs'), FP64, through lattice/run; "
                     "2D-1: final |Cd|, |Cl|; 2D-2: max |Cd|, |Cl| over the last fifth")
    return out


def section_cavity(iterations, log):
    """Lid-driven cavity 200x200, FP64 (launch-bound: 40 000 nodes)."""
    import torch
    from lbm_amd.batch import Batch, cavity_amplitudes, cavity_ensemble
    cases = cavity_ensemble(n_cases=1, seed=0, n_lo=200, n_hi=200, re_lo=100.0, re_hi=100.0)
    b = Batch(cases, storage="f64", ring=iterations + 2)
    b.set_amplitudes(0, cavity_amplitudes(cases, 0, iterations + 1))
    b.init()
    plan = chunks(1, iterations, 1000)
    for i0, n in plan:
        b.prepare(i0, n)
    torch.cuda.synchronize()
    s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s0.record()
    for i0, n in plan:
        b.steps(i0, n)
    s1.record()
    torch.cuda.synchronize()
    ms = s0.elapsed_time(s1)
    b.close()
    out = dict(config="cavity 200x200, u_lid=0.1, Re=100, FP64, %d iterations" % iterations,
               mlups=round(200 * 200 * (iterations - 1) / (ms * 1e-3) / 1e6, 1),
               us_per_iteration=round(ms * 1e3 / (iterations - 1), 2))
    log(f"# cavity200: {out}")
    return out


def section_batch(iterations, log):
    """256 seeded cavities (n ~ U{64..512}, Re ~ logU[100,1000]) in one launch
    per iteration, FP32, autotuned; per-case error vs the FP64 solves."""
    import torch
    from lbm_amd.batch import Batch, autotune, cavity_amplitudes, cavity_ensemble
    cases = cavity_ensemble(n_cases=256, seed=0)
    best, tried = autotune(cases, "f32", steps=20, cache="/tmp/lbm_amd_autotune.json", log=log)
    amp = cavity_amplitudes(cases, 0, iterations + 1)
    res = {}
    for storage, kw in (("f32", best), ("f64", {})):
        b = Batch(cases, storage=storage, ring=iterations + 2, **kw)
        b.set_amplitudes(0, amp)
        b.init()
        last = iterations - 1
        plan = chunks(1, last, 1000)
        for i0, n in plan:
            b.prepare(i0, n)
        torch.cuda.synchronize()
        s0, s1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        s0.record()
        for i0, n in plan:
            b.steps(i0, n)
        s1.record()
        torch.cuda.synchronize()
        ms = s0.elapsed_time(s1)
        us = []
        for c, cs in enumerate(cases):
            n = cs["nx"]
            rho = torch.empty((n, n), dtype=torch.float64, device="cuda")
            u = torch.empty((2, n, n), dtype=torch.float64, device="cuda")
            b.case(c).step(last, rho, u)
            us.append(u.cpu().numpy())
        res[storage] = dict(ms=ms, u=us, mlups=b.cells * (last - 1) / (ms * 1e-3) / 1e6,
                            grid=b.grid, cpt=b.cpt)
        b.close()
        log(f"# batch {storage}: {res[storage]['mlups']:.0f} MLUPS")
    errs = [float(np.linalg.norm(a - r) / np.linalg.norm(r))
            for a, r in zip(res["f32"]["u"], res["f64"]["u"])]
    cells = int(sum(c["nx"] * c["ny"] for c in cases))
    return dict(config="256 cavities, n~U{64..512}, Re~logU[100,1000], u_lid=0.1, %d iterations, "
                       "one launch per iteration" % iterations,
                cells=cells, f32_mlups=round(res["f32"]["mlups"], 1),
                f32_hbm_frac=round(res["f32"]["mlups"] * 1e6 * 72 / 8e12, 4),
                f64_mlups=round(res["f64"]["mlups"], 1),
                autotuned=dict(best, candidates=len(tried)),
                u_l2_rel_err_vs_f64=dict(max=max(errs), median=float(np.median(errs))))


def section_adjoint(log):
    """Adjoint: 1024x256 channel, solid-fraction disk + noise, 2000 iterations,
    loss over the last 1000, checkpoints every 100; FP64 and FP32.  Forward and
    backward wall times (host clock, device synchronised) are the medians of
    three pairs after a warm-up pair."""
    import torch
    from lbm_amd.adjoint import PbbEngine, cd_scale, cylinder_case
    case = cylinder_case()
    out = {"config": "1024x256 channel, disk D=40 at (200,128) as solid fraction 1, U[0,0.2] "
                     "noise in a 60x60 box, parabolic inlet u_max=0.05, Re=100, 2000 "
                     "iterations, loss = mean drag of the last 1000, checkpoint every 100"}
    grads = {}
    for storage in ("f64", "f32"):
        eng = PbbEngine(case, storage=storage)
        dt = torch.float64 if storage == "f64" else torch.float32
        s = torch.tensor(case["s"], dtype=dt, device="cuda")
        eng.forward(s, case["u_max"])           # warm-up
        eng.backward()
        torch.cuda.synchronize()
        fw, bw = [], []
        for _ in range(3):                      # medians of three forward/backward pairs
            t0 = time.perf_counter()
            loss = eng.forward(s, case["u_max"])
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            gs, gu = eng.backward()
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            fw.append(t1 - t0)
            bw.append(t2 - t1)
        f_med, b_med = float(np.median(fw)), float(np.median(bw))
        grads[storage] = (loss, gs.cpu().numpy(), gu)
        out[storage] = dict(loss=loss, cd=loss * cd_scale(case), forward_ms=round(f_med * 1e3, 2),
                            backward_ms=round(b_med * 1e3, 2),
                            adjoint_to_forward=round(b_med / f_med, 3),
                            forward_ms_runs=[round(t * 1e3, 2) for t in fw],
                            backward_ms_runs=[round(t * 1e3, 2) for t in bw],
                            dL_du_max=gu)
        eng.close()
        log(f"# adjoint {storage}: {out[storage]}")
    l64, g64, u64 = grads["f64"]
    l32, g32, u32 = grads["f32"]
    out["f32_vs_f64"] = dict(loss_rel_err=abs(l32 - l64) / abs(l64),
                             grad_s_rel_l2=float(np.linalg.norm(g32 - g64) / np.linalg.norm(g64)),
                             grad_umax_rel_err=abs(u32 - u64) / abs(u64))
    return out


FORM_OF_STORE = {"0": "f64", "1": "f32", "2": "f16", "3": "bf16", "4": "f16e", "5": "q16",
                 "6": "f16e8"}


def summarize_profile(pdir, nodes, copy_bytes=2 << 30):
    """Turn the rocprofv3 output of the three profiling runs under `pdir`
    (trace/: --kernel-trace --stats, fetch/: --pmc FETCH_SIZE, write/: --pmc
    WRITE_SIZE, commands.txt: the commands as run) into the text summary and
    the traffic figures under profiles/.  FETCH_SIZE and WRITE_SIZE (KiB per
    dispatch) are scaled by what they read for the copy kernel, which moves
    `copy_bytes` each way per dispatch."""
    import csv
    import glob
    import re

    def rows(sub, suffix):
        out = []
        for f in glob.glob(os.path.join(pdir, sub, "**", "*" + suffix), recursive=True):
            with open(f) as fh:
                out += list(csv.DictReader(fh))
        return out

    def form_of(name):
        m = re.search(r"k_step(_v4|_w7|_w8)?<(\d), (\d)>", name)
        return (FORM_OF_STORE[m.group(2)], m.group(1) == "_v4", int(m.group(3))) if m else None

    stats = {r["Name"]: r for r in rows("trace", "kernel_stats.csv")}
    counters = {}
    for sub, cname in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
        acc = {}
        for r in rows(sub, "counter_collection.csv"):
            if r["Counter_Name"] == cname:
                acc.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        counters[cname] = {k: sum(v) / len(v) for k, v in acc.items()}
    copy = [k for k in counters["FETCH_SIZE"] if "k_copy16" in k]
    assert copy, "no copy-kernel dispatch in the FETCH_SIZE run"
    fetch_factor = copy_bytes / 1024.0 / counters["FETCH_SIZE"][copy[0]]
    write_factor = copy_bytes / 1024.0 / counters["WRITE_SIZE"][copy[0]]
    with open(os.path.join(pdir, "commands.txt")) as fh:
        commands = fh.read().strip()
    lines = ["# Fused D2Q9 step kernels on the 8192x2048 obstacle array (64 IBB squares)",
             "", "Commands (repository root; rocprofv3 run from /tmp with TMPDIR=/tmp):", "",
             "```", commands, "```", "",
             f"Counter calibration on the copy kernel (k_copy16, {copy_bytes >> 20} MiB read and "
             f"{copy_bytes >> 20} MiB written per dispatch): FETCH_SIZE x {fetch_factor:.3f}, "
             f"WRITE_SIZE x {write_factor:.3f}.", "",
             "| kernel | form | calls | avg us | min us | max us | algorithmic MB/launch "
             "| HBM read MB | HBM write MB | traffic / algorithmic | algorithmic GB/s |",
             "|---|---|---|---|---|---|---|---|---|---|---|"]
    traffic = {}
    for name, st in sorted(stats.items()):
        fm = form_of(name)
        if fm is None:
            continue
        form, v4, cpt = fm
        alg = float(BYTES_PER_UPDATE[form]) * nodes
        avg_us = float(st["AverageNs"]) / 1e3
        fb = counters["FETCH_SIZE"].get(name)
        wb = counters["WRITE_SIZE"].get(name)
        rd = fb * 1024.0 * fetch_factor if fb is not None else None
        wr = wb * 1024.0 * write_factor if wb is not None else None
        tot = rd + wr if rd is not None and wr is not None else None
        lines.append(
            f"| `{name}` | {form}{' 4-node' if v4 else ''} cpt {cpt} | {st['Calls']} | "
            f"{avg_us:.2f} | {float(st['MinNs']) / 1e3:.2f} | {float(st['MaxNs']) / 1e3:.2f} | "
            f"{alg / 1e6:.1f} | {rd / 1e6 if rd else float('nan'):.1f} | "
            f"{wr / 1e6 if wr else float('nan'):.1f} | "
            f"{tot / alg if tot else float('nan'):.3f} | {alg / (avg_us * 1e-6) / 1e9:.1f} |")
        traffic[form] = {"kernel": name, "avg_duration_us": round(avg_us, 3),
                         "algorithmic_bytes_per_launch": alg,
                         "hbm_read_bytes_per_launch": rd, "hbm_write_bytes_per_launch": wr,
                         "hbm_bytes_per_launch": tot}
    lines += ["", "Other kernels (kernel trace):", "",
              "| kernel | calls | avg us |", "|---|---|---|"]
    for name, st in sorted(stats.items()):
        if form_of(name) is None:
            lines.append(f"| `{name[:90]}` | {st['Calls']} | {float(st['AverageNs']) / 1e3:.2f} |")
    os.makedirs(os.path.join(ROOT, "profiles"), exist_ok=True)
    with open(os.path.join(ROOT, "profiles", "k_step_obstacle_array_8192x2048.md"), "w") as fh:
        fh.write("\n".join(lines) + "\n")
    out = dict(traffic.get("f16", {}), fetch_factor=fetch_factor, write_factor=write_factor,
               forms=traffic)
    with open(os.path.join(ROOT, "profiles", "fused_f16_obstacle_array_traffic.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    print("\n".join(lines))


def qualified_form(forms_out, turek, bound=1e-3):
    """The fastest 16-bit storage form (on the obstacle array) whose Turek
    max-Cd and max-Cl relative errors against FP64 are both within `bound`."""
    if not turek:
        return None
    best = None
    for f, e in forms_out.items():
        t = turek.get(f)
        if STORE_BYTES[f] != 2 or not t or "max_cd_rel_err" not in t:
            continue
        if t["max_cd_rel_err"] <= bound and t["max_cl_rel_err"] <= bound:
            if best is None or e["mlups"] > forms_out[best]["mlups"]:
                best = f
    if best is None:
        return {"form": None, "bound": bound}
    e, t = forms_out[best], turek[best]
    return {"form": best, "bound": bound, "mlups": e["mlups"], "hbm_frac": e["hbm_frac"],
            "bytes_per_update": e["bytes_per_update"],
            "speedup_vs_f64": e.get("speedup_vs_f64"),
            "turek_max_cd_rel_err": t["max_cd_rel_err"],
            "turek_max_cl_rel_err": t["max_cl_rel_err"]}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--forms", default="f16,f64,f32,bf16,f16e,f16e8,q16")
    ap.add_argument("--graph-len", type=int, default=500)
    ap.add_argument("--cpt", type=int, default=0)
    ap.add_argument("--vector-width", type=int, default=0,
                    help="16-bit forms: 4 nodes per work-item (4), one (1) or automatic (0)")
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--cpu-iterations", type=int, default=160)
    ap.add_argument("--accuracy-iterations", type=int, default=ACC_ITERATIONS,
                    help="iterations of the obstacle-array run whose final state and drag "
                         "are compared with FP64 (the configuration's 10 000; 0 = only the "
                         "timed ones, for profiling runs)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--sections", default="array,cavity,dropin,turek,batch,adjoint",
                    help="array (8192x2048 obstacle array, the headline), cavity (200x200), "
                         "dropin (Turek through lattice/run), turek (cylinder per form), batch "
                         "(256 cavities), adjoint; optional: published (the reference app's "
                         "Turek runs against its README's values)")
    ap.add_argument("--turek-iterations", type=int, default=200000)
    ap.add_argument("--batch-iterations", type=int, default=5000)
    ap.add_argument("--summarize-profile", metavar="DIR",
                    help="write profiles/ summaries from rocprofv3 output under DIR and exit")
    args = ap.parse_args()
    if args.summarize_profile:
        summarize_profile(args.summarize_profile, nodes=8192 * args.ny)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (the lattice would be "
              "split into x-slabs with one halo column per cut; not built)", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    app, lat, om_p, om_m = build_case(ny=args.ny)
    forms = [f.strip() for f in args.forms.split(",") if f.strip()]
    res = {}
    for f in forms:
        print(f"# form {f}: {args.warmup} warmup + {args.steps} timed iterations",
              file=sys.stderr, flush=True)
        res[f] = run_form(app, lat, om_p, om_m, f, args.steps, args.warmup, args.graph_len,
                          args.cpt, args.vector_width, acc_iterations=args.accuracy_iterations)
        print(f"# form {f}: {res[f]['mlups']:.0f} MLUPS, {res[f]['ms'] / args.steps * 1e3:.1f} "
              f"us/iteration", file=sys.stderr, flush=True)
    ref = res.get("f64")
    # accuracy against the FP64 run of the same configuration (10 000
    # iterations): time-averaged drag of each obstacle over the last 1000
    # iterations, velocity field of the last iteration
    forms_out = {}
    for f, r in res.items():
        bpu = BYTES_PER_UPDATE[f]
        gbs = r["mlups"] * 1e6 * bpu / 1e9
        entry = {"mlups": round(r["mlups"], 1), "us_per_iteration": round(r["ms"] / args.steps * 1e3, 2),
                 "algorithmic_GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                 "bytes_per_update": bpu, "mass": r["mass"], "max_speed": r["max_speed"],
                 "accuracy_iterations": r["iterations"]}
        if ref is not None and f != "f64":
            # time-averaged momentum-exchange force of each obstacle over the last
            # iterations; errors relative to the largest FP64 drag of the array
            fx = r["drag"][:, :, 0].mean(axis=0)
            fx64 = ref["drag"][:, :, 0].mean(axis=0)
            fy = r["drag"][:, :, 1].mean(axis=0)
            fy64 = ref["drag"][:, :, 1].mean(axis=0)
            scale = float(np.max(np.abs(fx64)))
            entry["drag_err_rel_max_drag"] = float(np.max(np.abs(fx - fx64)) / scale)
            entry["lift_err_rel_max_drag"] = float(np.max(np.abs(fy - fy64)) / scale)
            entry["total_drag_rel_err"] = float(abs(fx.sum() - fx64.sum()) / abs(fx64.sum()))
            entry["u_l2_rel_err"] = float(np.linalg.norm(r["u"] - ref["u"]) /
                                          np.linalg.norm(ref["u"]))
            entry["speedup_vs_f64"] = round(r["mlups"] / ref["mlups"], 3)
        forms_out[f] = entry
    head = "f16" if "f16" in res else forms[0]
    h = res[head]
    achieved = h["mlups"] * 1e6 * BYTES_PER_UPDATE[head] / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fused_f16_obstacle_array_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        print(f"# cpu baseline: oracle, {args.cpu_iterations} iterations", file=sys.stderr,
              flush=True)
        mlups_cpu, secs, threads = cpu_baseline(app, lat, om_p, om_m, args.cpu_iterations)
        cpu = {"value": round(mlups_cpu, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
               "sample": f"{args.cpu_iterations} iterations of the same 8192x2048 obstacle-array "
                         f"workload, FP64, oracle/d2q9_oracle.c (OpenMP node loops, {threads} "
                         f"threads) driven in the reference's call order ({secs:.1f} s)"}
    copy = copy_ceiling()
    sections = set(x.strip() for x in args.sections.split(","))

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    extra = {}
    if "cavity" in sections:
        extra["cavity"] = section_cavity(20001, log)
    if "dropin" in sections:
        extra["dropin"] = section_dropin(args.turek_iterations, log)
    if "published" in sections:
        extra["published_turek"] = section_published(log)
    if "turek" in sections:
        extra["turek"] = section_turek(["f64", "f32", "f16", "f16e", "f16e8", "bf16", "q16"],
                                       args.turek_iterations, log)
    if "dropin" in extra and "turek" in extra and "f64" in extra["turek"]:
        extra["dropin"]["vs_batch_path_f64"] = round(
            extra["dropin"]["mlups"] / extra["turek"]["f64"]["mlups"], 3)
    if "batch" in sections:
        extra["batch"] = section_batch(args.batch_iterations, log)
    if "adjoint" in sections:
        extra["adjoint"] = section_adjoint(log)
    line = {
        "metric": "D2Q9 MLUPS (FP16-offset storage, 8192x2048 obstacle array) and % of HBM roofline",
        "value": round(h["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(h["ms"] / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "storage": head + ("-offset" if head in ("f16", "bf16") else ""),
        "data": "synthetic",
        "config": {"workload": "obstacle_array: channel 8192x2048, 64 IBB squares (side 32-96, "
                               "seed 0), Zou-He velocity inlet / pressure outlet / walls / corners, "
                               "TRT, drag+lift terms every iteration",
                   "nx": app.nx, "ny": app.ny, "obstacles": len(app.obstacles),
                   "links": h["n_links"], "iterations_timed": args.steps},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
        "copy_ceiling_GBps": copy["GBps"],
        "copy_ceiling": copy,
        "cpu_baseline": cpu,
        "forms": forms_out,
    }
    line.update(extra)
    line["accuracy_qualified"] = qualified_form(forms_out, extra.get("turek"))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
