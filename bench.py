"""Benchmark of the fused D2Q9 time step on one MI355X.

Workload (one "step" = one lattice iteration of the whole domain): channel
8192 x 2048 with 64 axis-aligned square obstacles (IBB links, drag/lift terms
every iteration), uniform Zou-He inlet u = 0.04 (ramp sigma = 1000
iterations), Zou-He zero-pressure outlet, no-slip walls, Re = 2000 on the
height, TRT with Lambda = 1/4.  Inputs are synthetic and resident in HBM
before the timed region.

The headline value is the MLUPS of the FP16-offset storage form (f - w stored
as IEEE half, FP32 arithmetic).  The FP64, FP32 and BF16-offset forms are
timed on the same workload and reported under "forms", each with its
fraction of the 8 TB/s HBM roofline (18 x storage bytes per lattice update),
its drag error against the FP64 run and its velocity-field L2 error.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--forms f16,f64,f32,bf16]
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
STORE_BYTES = {"f64": 8, "f32": 4, "f16": 2, "bf16": 2}


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


def run_form(app, lat, om_p, om_m, storage, steps, warmup, graph_len, cpt=0):
    import torch
    from lbm_amd import _lib
    from lbm_amd.domain import FusedDomain, pack_inlets
    nx, ny = app.nx, app.ny
    total = warmup + steps
    bc = [_lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_VELOCITY, _lib.BC_PRESSURE]
    dom = FusedDomain(nx, ny, om_p, om_m, bc, [(o.boundary, o.ibb) for o in app.obstacles],
                      ibb=True, storage=storage, ring=total + 3, cells_per_thread=cpt)
    pool, offs = pack_inlets(inlet_rows(app, total + 2))
    dom.set_inlets(0, offs, pool)
    w = torch.tensor([4. / 9.] + [1. / 9.] * 4 + [1. / 36.] * 4, dtype=torch.float64)
    g0 = w[:, None, None].expand(9, nx, ny).contiguous().cuda()
    dom.init(g0)
    it = 1
    while it < warmup:
        n = min(graph_len, warmup - it)
        dom.steps(it, n)
        it += n
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record()
    end_it = it + steps
    while it < end_it:
        n = min(graph_len, end_it - it)
        dom.steps(it, n)
        it += n
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    last = it - 1
    # outputs of the run, after the timed region: one more iteration with its
    # moments stored, the drag of the last 1000 iterations, the final fields
    rho = torch.empty((nx, ny), dtype=torch.float64, device="cuda")
    u = torch.empty((2, nx, ny), dtype=torch.float64, device="cuda")
    dom.step(it, rho, u)
    last = it
    dom.drag_only(last)
    n_avg = min(1000, steps)
    drag = dom.drag_read(last - n_avg + 1, n_avg)
    g = torch.empty((9, nx, ny), dtype=torch.float64, device="cuda")
    gup = torch.empty_like(g)
    mass = dom.mass(last)
    dom.epilogue(last, g, gup, rho, u)
    ux = u.cpu().numpy()
    del g, gup, rho, u, g0
    dbytes = dom.device_bytes()
    dom.close()
    torch.cuda.synchronize()
    mlups = nx * ny * steps / (ms * 1e-3) / 1e6
    return dict(ms=ms, wall=wall, mlups=mlups, drag=drag, u=ux, mass=mass, device_bytes=dbytes,
                n_links=int(sum(len(o.boundary) for o in app.obstacles)))


def copy_ceiling_gbs(gib=2, blocks=4096, reps=10):
    """HBM copy ceiling: lbm_copy_probe moving `gib` GiB (read + write bytes / time)."""
    import torch
    from lbm_amd import _lib
    L = _lib.load()
    nbytes = gib << 30
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    st = _lib.stream_ptr()
    for _ in range(3):
        _lib.check(L.lbm_copy_probe(b.data_ptr(), a.data_ptr(), nbytes, blocks, st))
    torch.cuda.synchronize()
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    s.record()
    for _ in range(reps):
        _lib.check(L.lbm_copy_probe(b.data_ptr(), a.data_ptr(), nbytes, blocks, st))
    e.record()
    torch.cuda.synchronize()
    gbs = 2 * nbytes * reps / (s.elapsed_time(e) * 1e-3) / 1e9
    del a, b
    return gbs


def cpu_baseline(app, lat, om_p, om_m, iterations):
    """The oracle (strict-order C restatement of the reference kernels, one
    thread) on the same workload for a few iterations, driven in the
    reference's call order (run.py:24-54)."""
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
    return nx * ny * iterations / dt / 1e6, dt


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--forms", default="f16,f64,f32,bf16")
    ap.add_argument("--graph-len", type=int, default=500)
    ap.add_argument("--cpt", type=int, default=0)
    ap.add_argument("--ny", type=int, default=2048)
    ap.add_argument("--cpu-iterations", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(the domain decomposition is described in DESIGN.md)", file=sys.stderr)
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
                          args.cpt)
        print(f"# form {f}: {res[f]['mlups']:.0f} MLUPS, {res[f]['ms'] / args.steps * 1e3:.1f} "
              f"us/iteration", file=sys.stderr, flush=True)
    ref = res.get("f64")
    # drag error: time-averaged drag of each obstacle over the last 1000 iterations
    forms_out = {}
    for f, r in res.items():
        sb = STORE_BYTES[f]
        gbs = r["mlups"] * 1e6 * 18 * sb / 1e9
        entry = {"mlups": round(r["mlups"], 1), "us_per_iteration": round(r["ms"] / args.steps * 1e3, 2),
                 "algorithmic_GBps": round(gbs, 1), "hbm_frac": round(gbs / HBM_PEAK_GBS, 4),
                 "bytes_per_update": 18 * sb, "mass": r["mass"]}
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
    sb = STORE_BYTES[head]
    achieved = h["mlups"] * 1e6 * 18 * sb / 1e9
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "fused_f16_obstacle_array_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            traffic = json.load(fh).get("hbm_bytes_per_launch")
    cpu = None
    if not args.no_cpu:
        print(f"# cpu baseline: oracle, {args.cpu_iterations} iterations", file=sys.stderr,
              flush=True)
        mlups_cpu, secs = cpu_baseline(app, lat, om_p, om_m, args.cpu_iterations)
        cpu = {"value": round(mlups_cpu, 3), "unit": "MLUPS", "cores": 1, "kind": "port",
               "sample": f"{args.cpu_iterations} iterations of the same 8192x2048 obstacle-array "
                         f"workload, FP64, oracle/d2q9_oracle.c driven in the reference's call "
                         f"order ({secs:.1f} s)"}
    copy_gbs = copy_ceiling_gbs()
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
        "copy_ceiling_GBps": round(copy_gbs, 1),
        "cpu_baseline": cpu,
        "forms": forms_out,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
