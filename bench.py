#!/usr/bin/env python3
"""Benchmark of the staggered hot path on one MI355X.

Headline metric: HISQ Dslash GB/s and fraction of HBM peak on a 32^4
lattice, plus FP64 vs mixed-precision CG time to a 1e-10 true residual.

A "step" is one single-parity even-odd Dslash over the 32^4 lattice's 524288
even sites (FP64, fat 1-hop + Naik 3-hop, Naik links stored as 12 reals),
inputs resident in HBM.  `value` is algorithmic bytes / average Dslash time,
with 2017 bytes per site: 8 fat links x 18 x 8 B + 8 Naik links x 12 x 8 B +
1 B of Naik row-3 signs + 48 B input + 48 B output.

The CPU baseline is the oracle's OpenMP Dslash (same site loop as the
reference's ks_long_dslash.cpp:564-668) on a bounded sample on this host.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "milc-qphix_amd"))
# OpenMP placement of the CPU baseline's threads (read when the OpenMP runtime
# starts, so set before anything loads it); an explicit setting wins.
os.environ.setdefault("OMP_PROC_BIND", "close")
os.environ.setdefault("OMP_PLACES", "cores")

import numpy as np  # noqa: E402

METRIC = "HISQ Dslash GB/s (% HBM peak); mixed vs FP64 CG time to 1e-10 residual, 32^4"
HBM_PEAK_GBS = 8000.0
DIMS = (32, 32, 32, 32)


def site_bytes(real_bytes, n_long, spinor_bytes):
    """Algorithmic bytes per output site of one single-parity Dslash."""
    sign = 1 if n_long in (12, 8) else 0
    return real_bytes * (8 * 18 + 8 * n_long) + sign + 2 * spinor_bytes


def measured_traffic():
    """HBM bytes per launch of the FP64 plain Dslash (FETCH_SIZE + WRITE_SIZE, each
    scaled by the copy-kernel factor) from the committed rocprofv3 summary of this
    benchmark's Dslash run, or None when that summary is absent."""
    path = os.path.join(ROOT, "profiles", "dslash_32x4_counters.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return int(round(d["fetch_bytes"] + d["write_bytes"]))
    except (OSError, KeyError, ValueError):
        return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_fields(lat, F):
    fat, lng = F.make_links(lat, "weak", 1002, eps=0.3)
    raw = F.raw_link_set(lat, fat, lng)
    del fat, lng
    src = F.to_raw(F.gaussian_source(lat, 2002))
    return raw, src


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_cores():
    """(physical cores of the host, CPUs this process may run on): the first
    from the (package, core) pairs in sysfs, the second from the affinity mask."""
    phys = set()
    base = "/sys/devices/system/cpu"
    try:
        for d in os.listdir(base):
            if not (d.startswith("cpu") and d[3:].isdigit()):
                continue
            try:
                with open(f"{base}/{d}/topology/physical_package_id") as f:
                    pkg = f.read().strip()
                with open(f"{base}/{d}/topology/core_id") as f:
                    core = f.read().strip()
                phys.add((pkg, core))
            except OSError:
                continue
    except OSError:
        pass
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 0
    return (len(phys) or None), aff


def cpu_baseline(dims, raw, src, budget_s):
    """Oracle Dslash (FP64) timed on the host: two warm-ups, then a bounded
    number (5 to 50) of individually timed applications; returns the median."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    orc = O.Oracle(dims, *raw)
    vh = int(np.prod(dims)) // 2
    half = np.ascontiguousarray(src[6 * vh:])
    t = time.perf_counter()
    orc.dslash(half, 0)  # warm-ups
    orc.dslash(half, 0)
    first = (time.perf_counter() - t) / 2
    reps = max(5, min(50, int(budget_s / max(first, 1e-3))))
    ts = []
    for _ in range(reps):
        t = time.perf_counter()
        orc.dslash(half, 0)
        ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    threads = O.lib().oracle_num_threads()
    orc.close()
    return dt, reps, threads


def cg_compare(A, s, L64, lat, src, raw_links):
    """FP64 CG vs mixed-precision CG (FP32 and FP16-stored sloppy spinors, FP32
    sloppy links with 8- or 12-real long links or 16-bit fixed-point sloppy
    links, FP64 reliable updates), each to a 1e-10
    true relative residual (resid is the squared residual, 1e-20), m = 0.02 in
    M = m^2 - D_eo D_oe (API mass m/2)."""
    mass, resid = 0.01, 1e-20
    b = src.copy()
    b[6 * lat.vh:] = 0
    vb = A.Vector(s, b, A.EVEN, "D")
    out = {}
    bnorm2 = float(b[:6 * lat.vh] @ b[:6 * lat.vh])
    x64 = None
    # mixed_*: FP32 sloppy links with 8-real long links (the default); *_r12:
    # 12-real long links; *_l16: 16-bit fixed-point links
    for name, bits, lbits, recon in (("f64", None, None, 8), ("mixed_f32", 32, 32, 8),
                                     ("mixed_f16", 16, 32, 8), ("mixed_f32_r12", 32, 32, 12),
                                     ("mixed_f16_r12", 16, 32, 12), ("mixed_f32_l16", 32, 16, 12),
                                     ("mixed_f16_l16", 16, 16, 12)):
        vx = A.Vector(s, None, A.EVEN, "D")
        if lbits is not None:
            s.lib.QPHIX_MI355X_set_sloppy_link_bits(lbits)
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(recon)
        t = time.perf_counter()
        if bits is None:
            ret, info, res = A.invert(L64, vx, vb, mass, A.EVEN, resid, niter=5000, nrestart=10)
        else:
            ret, info, res = A.invert_mixed(L64, vx, vb, mass, A.EVEN, resid, sloppy_bits=bits,
                                            niter=5000, nrestart=10)
        sec = time.perf_counter() - t
        s.lib.QPHIX_MI355X_set_sloppy_link_bits(32)
        s.lib.QPHIX_MI355X_set_sloppy_long_recon(8)
        x = vx.extract()[:6 * lat.vh]
        rec = {"seconds": round(sec, 4), "iters": res.final_iter, "final_rsq": res.final_rsq,
               "status": info.status}
        if bits is None:
            x64 = x
        else:
            rec["reliable_updates"] = res.final_restart
            rec["sloppy_links"] = "FP32, long links as 8 reals" if (lbits == 32 and recon == 8) else \
                ("FP32, long links as 12 reals" if lbits == 32 else "16-bit fixed point")
            rec["rel_err_vs_f64"] = float(np.linalg.norm(x - x64) / np.linalg.norm(x64))
            rec["speedup_vs_f64"] = round(out["f64"]["seconds"] / sec, 3)
        out[name] = rec
        log(f"[bench] CG {name}: {sec:.3f}s {res.final_iter} iters rsq {res.final_rsq:.2e}")
        vx.free()
    # Pure FP32 CG (the reference's F build: FP32 links, vectors and solver
    # scalars; no FP64 anywhere) with the same target: it stagnates, so it runs
    # to its iteration cap.  Reported: where it stagnates, measured with the
    # FP64 operator, and its error against the FP64 solution.
    raw32 = [a.astype(np.float32) for a in raw_links]
    L32 = A.Links(s, *raw32, prec="F")
    del raw32
    vb32 = A.Vector(s, b.astype(np.float32), A.EVEN, "F")
    vx32 = A.Vector(s, None, A.EVEN, "F")
    t = time.perf_counter()
    ret, info, res = A.invert(L32, vx32, vb32, mass, A.EVEN, resid, niter=1000, nrestart=4)
    sec = time.perf_counter() - t
    x32 = vx32.extract()[:6 * lat.vh].astype(np.float64)
    vx = A.Vector(s, np.concatenate([x32, np.zeros(6 * lat.vh)]), A.EVEN, "D")
    tmp = A.Vector(s, None, A.EVENODD, "D")
    A.dslash(L64, tmp, vx, A.ODD)
    A.dslash(L64, vx_out := A.Vector(s, None, A.EVEN, "D"), tmp, A.EVEN)
    ddx = vx_out.extract()[:6 * lat.vh]
    r = b[:6 * lat.vh] - (4 * mass * mass * x32 - ddx)
    true_rsq = float(r @ r) / bnorm2
    out["f32"] = {"seconds": round(sec, 4), "iters": res.final_iter, "final_rsq": res.final_rsq,
                  "true_rsq_fp64_operator": true_rsq,
                  "rel_err_vs_f64": float(np.linalg.norm(x32 - x64) / np.linalg.norm(x64)),
                  "status": info.status, "stagnates": bool(true_rsq > resid)}
    log(f"[bench] CG f32 (pure): {sec:.3f}s {res.final_iter} iters, solver rsq {res.final_rsq:.2e}, "
        f"FP64-operator rsq {true_rsq:.2e}, rel err vs f64 {out['f32']['rel_err_vs_f64']:.1e}")
    for o in (L32, vb32, vx32, vx, tmp, vx_out, vb):
        o.free()
    return out


def batch_multishift(A, F, nrhs=8, nshift=12, dims=(32, 32, 32, 64)):
    """32^3x64, 12 shifts log-spaced in [1e-4, 1e-1], 8 Gaussian sources:
    batched mixed-precision multi-shift (FP32 iteration, FP64 reliable updates
    and per-shift refinement) to a 1e-10 true residual per shift, against the
    FP64 single-source multi-shift (the reference's algorithm) on one source."""
    lat = F.Lattice(dims)
    log(f"[bench] batched multi-shift: generating {dims} links")
    fat, lng = F.make_links(lat, "weak", 1003, eps=0.3)
    raw = F.raw_link_set(lat, fat, lng)
    del fat, lng
    sig = np.logspace(-4, -1, nshift)
    masses = list(np.sqrt(sig) / 2)
    resid = 1e-20
    out = {"lattice": list(dims), "nrhs": nrhs, "nshift": nshift, "sigma": [float(x) for x in sig]}
    with A.Session(dims) as s:
        L = A.Links(s, *raw)
        del raw
        bs = []
        for r in range(nrhs):
            b = F.to_raw(F.gaussian_source(lat, 3000 + r))
            b[6 * lat.vh:] = 0
            bs.append(A.Vector(s, b, A.EVEN))
        xs = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
        s.lib.QPHIX_MI355X_sync()
        t = time.perf_counter()
        ret, info, res = A.invert_multi_rhs(L, xs, bs, masses, A.EVEN, resid, niter=20000, nrestart=2)
        sec = time.perf_counter() - t
        worst = max(res[r][j].final_rsq for r in range(nrhs) for j in range(nshift))
        # per solution: batched steps until its shift retired, plus its refinement iterations
        sol_iters = [res[r][j].final_iter for r in range(nrhs) for j in range(nshift)]
        out["mixed_batched"] = {"seconds": round(sec, 3), "seconds_per_source": round(sec / nrhs, 3),
                                "multishift_steps": ret, "max_final_rsq": worst,
                                "solution_iters_total": int(sum(sol_iters)),
                                "solution_iters_max": int(max(sol_iters)), "status": info.status}
        log(f"[bench] batched mixed multi-shift: {sec:.2f}s for {nrhs} sources x {nshift} shifts, "
            f"{ret} steps, worst rsq {worst:.2e}")
        x64 = [A.Vector(s, None, A.EVEN) for _ in masses]
        t = time.perf_counter()
        ret64, info64, res64 = A.invert_multi(L, x64, bs[0], masses, A.EVEN, resid, niter=20000,
                                              nrestart=2)
        sec64 = time.perf_counter() - t
        errs = []
        for j in range(nshift):
            a = xs[0][j].extract()[:6 * lat.vh]
            b = x64[j].extract()[:6 * lat.vh]
            errs.append(float(np.linalg.norm(a - b) / np.linalg.norm(b)))
        out["fp64_single_source"] = {"seconds": round(sec64, 3), "iters": ret64,
                                     "final_rsq": res64[0].final_rsq}
        out["max_rel_err_vs_fp64"] = max(errs)
        out["speedup_per_source_vs_fp64"] = round(sec64 / (sec / nrhs), 2)
        log(f"[bench] FP64 multi-shift, one source: {sec64:.2f}s {ret64} iters; "
            f"speed-up per source {out['speedup_per_source_vs_fp64']}x, max rel err {max(errs):.1e}")
        for o in [L] + bs + x64 + [v for row in xs for v in row]:
            o.free()
    return out


def force_24(A, F, dims=(24, 24, 24, 24)):
    """24^4, FP64: the Symanzik gauge force (beta = 6.0, u0 = 0.86) on weak-field
    SU(3) links, and the fermion-force outer products of
    phi^dag (M^dag M + sigma_k)^{-1} phi for 4 shifts from the FP64 multi-shift
    solutions.  Fields are resident in HBM; the gauge force is timed as the
    synchronous API call (host clock, 20 calls after one warm-up), the fermion
    force as device time."""
    lat = F.Lattice(dims)
    out = {"lattice": list(dims)}
    log(f"[bench] forces at {dims}: generating fields")
    with A.Session(dims) as s:
        U = F.gauge_field(lat, "weak", 31, eps=0.3)
        G = A.Gauge(s, F.to_raw(U))
        del U
        Fm = A.Force(s, np.zeros(72 * lat.V))
        coef = F.symanzik_coeffs(6.0, 0.86)

        def timed():
            A.gauge_force(G, Fm, coef, 0.02)
            reps = 20
            t = time.perf_counter()
            for _ in range(reps):
                A.gauge_force(G, Fm, coef, 0.02)
            return (time.perf_counter() - t) / reps

        # staple-product kernel (the default for SU(3) links): 10.5 products per
        # link to form the 24 plaquette staples of every site, 51 per link to
        # combine them (plaquettes, 3 rectangle shapes, parallelograms) and W =
        # U S; loop-walk kernel: 139 products per link (walks ending on the same
        # link share it) and U S; naive: 181 products, every loop on its own
        sec = timed()
        s.lib.QPHIX_MI355X_set_gauge_force_staples(0)
        sec_walk = timed()
        s.lib.QPHIX_MI355X_set_gauge_force_staples(1)
        flop = 61.5 * 216.0 * 4 * lat.V
        naive = 181 * 216.0 * 4 * lat.V
        out["gauge_force"] = {"ms": round(sec * 1e3, 3), "kernel": "staple products",
                              "GFLOPs": round(flop / sec / 1e9, 1),
                              "frac_fp64_peak": round(flop / sec / 78.6e12, 4),
                              "flop_per_link": 61.5 * 216, "naive_flop_per_link": 181 * 216,
                              "naive_equivalent_GFLOPs": round(naive / sec / 1e9, 1),
                              "loop_walk_kernel_ms": round(sec_walk * 1e3, 3),
                              "loop_walk_flop_per_link": 140 * 216}
        log(f"[bench] gauge force {sec * 1e3:.2f} ms (loop-walk kernel {sec_walk * 1e3:.2f} ms), "
            f"{flop / sec / 1e9:.0f} GFLOP/s executed ({naive / sec / 1e9:.0f} naive-equivalent)")
        G.free()
        Fm.free()

        fat, lng = F.make_links(lat, "weak", 41, eps=0.3)
        L = A.Links(s, *F.raw_link_set(lat, fat, lng))
        del fat, lng
        phi = F.to_raw(F.gaussian_source(lat, 43, parity=0))
        vb = A.Vector(s, phi, A.EVEN)
        masses = [0.05, 0.1, 0.2, 0.4]
        fcoef = [0.5, 0.25, 0.125, 0.0625]
        xs = [A.Vector(s, None, A.EVEN) for _ in masses]
        ret, info, res = A.invert_multi(L, xs, vb, masses, A.EVEN, 1e-20, niter=4000)
        t_dev, t_k = A.fermion_force_time(L, fcoef, xs, reps=10)
        nbytes = (len(masses) * 96 + 2304) * lat.vh
        out["fermion_force"] = {"terms": len(masses), "device_ms": round(t_dev, 4),
                                "oprod_kernel_ms": round(t_k, 4),
                                "oprod_bytes": nbytes,
                                "oprod_GBs": round(nbytes / (t_k * 1e-3) / 1e9, 1),
                                "solve_iters": ret}
        log(f"[bench] fermion force: {t_dev:.3f} ms device ({t_k:.3f} ms outer-product kernel, "
            f"{nbytes / (t_k * 1e-3) / 1e9:.0f} GB/s)")
        for o in [L, vb] + xs:
            o.free()
    return out


def combined_48(A, F, dims=(48, 48, 48, 96), nrhs=4, nshift=16, mass=0.0102):
    """One GPU's share of a production lattice: 48^3x96 weak-field links, 16-shift
    mixed multi-shift on 4 Gaussian sources (FP64 reliable updates, 1e-10 per
    shift), then the fermion-force outer products of all 64 solutions; against
    the FP64 path (FP64 multi-shift per source + the same force)."""
    lat = F.Lattice(dims)

    sig = np.logspace(np.log10(4 * mass * mass), 0.0, nshift)
    masses = list(np.sqrt(sig) / 2)
    coef = list(1.0 / (1.0 + np.arange(nshift)))
    out = {"lattice": list(dims), "nrhs": nrhs, "nshift": nshift, "mass": mass,
           "sigma": [float(x) for x in sig]}
    with A.Session(dims) as s:
        log(f"[bench] combined {dims}: generating links")
        fat, lng = A.weak_links_on_device(s, lat, 51, 0.3)
        log("[bench]   raw link arrays")
        raw = F.raw_link_set(lat, fat, lng)
        del fat, lng
        log("[bench]   links to the device")
        L = A.Links(s, *raw)
        del raw
        bs = [A.Vector(s, F.to_raw(F.gaussian_source(lat, 5000 + r, parity=0)), A.EVEN)
              for r in range(nrhs)]
        log("[bench] combined: mixed batched multi-shift")
        xm = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
        s.lib.QPHIX_MI355X_sync()
        t = time.perf_counter()
        ret, info, res = A.invert_multi_rhs(L, xm, bs, masses, A.EVEN, 1e-20, niter=20000, nrestart=2)
        t_solve = time.perf_counter() - t
        flat = [x for row in xm for x in row]
        cf = [c for _ in bs for c in coef]
        t_force, _ = A.fermion_force_time(L, cf, flat, reps=1)
        worst = max(r.final_rsq for row in res for r in row)
        out["mixed"] = {"solve_s": round(t_solve, 3), "force_ms": round(t_force, 3),
                        "end_to_end_s": round(t_solve + t_force * 1e-3, 3), "steps": ret,
                        "worst_final_rsq": worst, "status": info.status}
        log(f"[bench] combined mixed: solve {t_solve:.2f}s, force {t_force:.1f} ms, worst rsq {worst:.2e}")
        o1m, o3m, _ = A.fermion_force_oprod(L, cf, flat)
        for x in flat:
            x.free()
        log("[bench] combined: FP64 multi-shift per source")
        x64 = [[A.Vector(s, None, A.EVEN) for _ in masses] for _ in bs]
        t = time.perf_counter()
        for k, (b, row) in enumerate(zip(bs, x64)):
            A.invert_multi(L, row, b, masses, A.EVEN, 1e-20, niter=20000, nrestart=2)
            log(f"[bench]   FP64 source {k} done ({time.perf_counter() - t:.1f}s)")
        t64 = time.perf_counter() - t
        flat = [x for row in x64 for x in row]
        t_force64, _ = A.fermion_force_time(L, cf, flat, reps=1)
        o1d, o3d, _ = A.fermion_force_oprod(L, cf, flat)
        out["fp64"] = {"solve_s": round(t64, 3), "force_ms": round(t_force64, 3),
                       "end_to_end_s": round(t64 + t_force64 * 1e-3, 3)}
        err1 = float(np.linalg.norm(o1m - o1d) / np.linalg.norm(o1d))
        err3 = float(np.linalg.norm(o3m - o3d) / np.linalg.norm(o3d))
        out["force_rel_err_mixed_vs_fp64"] = {"fat": err1, "long": err3}
        out["speedup_end_to_end"] = round(out["fp64"]["end_to_end_s"] / out["mixed"]["end_to_end_s"], 2)
        log(f"[bench] combined FP64: solve {t64:.2f}s; force rel err fat {err1:.1e} long {err3:.1e}; "
            f"speed-up {out['speedup_end_to_end']}x")
        for x in flat + [L] + bs:
            x.free()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--only-batch", action="store_true",
                    help="run only the batched multi-shift measurement (prints its JSON)")
    ap.add_argument("--no-force", action="store_true")
    ap.add_argument("--only-force", action="store_true",
                    help="run only the 24^4 gauge- and fermion-force measurement (prints its JSON)")
    ap.add_argument("--combined", action="store_true",
                    help="run only the 48^3x96 multi-shift + fermion-force case (prints its JSON)")
    ap.add_argument("--no-combined", action="store_true",
                    help="leave out the 48^3x96 multi-shift + fermion-force case")
    ap.add_argument("--only-cg", action="store_true",
                    help="run only the 32^4 FP64 / mixed / FP32 CG comparison (prints its JSON)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs one lattice on one GPU; --gpus {args.gpus} is not "
              "supported (multi-GPU sharding is design-only)", file=sys.stderr)
        return 2

    from pyqphix import api as A
    from pyqphix import fields as F

    if args.only_batch:
        print(json.dumps({"multishift_batched": batch_multishift(A, F)}), flush=True)
        return 0
    if args.only_force:
        print(json.dumps({"forces_24": force_24(A, F)}), flush=True)
        return 0
    if args.combined:
        print(json.dumps({"combined_48x96": combined_48(A, F)}), flush=True)
        return 0

    lat = F.Lattice(DIMS)
    log("[bench] generating 32^4 weak-field links")
    t = time.perf_counter()
    raw, src = make_fields(lat, F)
    log(f"[bench] fields ready in {time.perf_counter() - t:.1f}s")
    if args.only_cg:
        with A.Session(DIMS) as s:
            L64 = A.Links(s, *raw, prec="D")
            res = cg_compare(A, s, L64, lat, src, raw)
            L64.free()
        print(json.dumps({"cg": res}), flush=True)
        return 0

    s = A.Session(DIMS)
    lib = s.lib
    ncu = __import__("ctypes").c_int()
    lib.QPHIX_MI355X_device_info(__import__("ctypes").byref(ncu), None)

    # achieved copy bandwidth (read + write), for reference
    copy_bytes = 2 << 30
    copy_ms = lib.QPHIX_MI355X_copy_time(copy_bytes, 20)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9

    L64 = A.Links(s, *raw, prec="D")
    vin = A.Vector(s, src, A.EVENODD, "D")
    vout = A.Vector(s, None, A.EVENODD, "D")
    rb = site_bytes(8, L64.recon, 48)
    log(f"[bench] FP64 links recon={L64.recon}, {rb} B/site")

    A.dslash_time(L64, vout, vin, A.EVEN, args.warmup)
    lib.QPHIX_MI355X_sync()
    t0 = time.perf_counter()
    ms_ev = A.dslash_time(L64, vout, vin, A.EVEN, args.steps)
    lib.QPHIX_MI355X_sync()
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    gbs = rb * lat.vh / (ms_ev * 1e-3) / 1e9
    gflops = 1146.0 * lat.vh / (ms_ev * 1e-3) / 1e9
    log(f"[bench] FP64 Dslash {ms_ev * 1e3:.1f} us  {gbs:.0f} GB/s  {gflops:.0f} GFLOP/s")

    extra = {"plain_copy_kernel_GBs": round(copy_gbs, 1),
             "plain_copy_note": "one grid-stride 16-byte copy of 2 GiB (read + write); a floor for "
                                "the card, not its bandwidth ceiling (8 TB/s peak is the roofline)",
             "n_cu": ncu.value}
    extra["dslash_f64"] = {"us": round(ms_ev * 1e3, 2), "GBs": round(gbs, 1),
                           "GFLOPs": round(gflops, 1), "bytes_per_site": rb, "recon": L64.recon}

    # FP32 Dslash (the mixed-precision solver's operator)
    raw32 = [r.astype(np.float32) for r in raw]
    L32 = A.Links(s, *raw32, prec="F")
    vin32 = A.Vector(s, src.astype(np.float32), A.EVENODD, "F")
    vout32 = A.Vector(s, None, A.EVENODD, "F")
    A.dslash_time(L32, vout32, vin32, A.EVEN, args.warmup)
    ms32 = A.dslash_time(L32, vout32, vin32, A.EVEN, args.steps)
    rb32 = site_bytes(4, L32.recon, 24)
    extra["dslash_f32"] = {"us": round(ms32 * 1e3, 2),
                           "GBs": round(rb32 * lat.vh / (ms32 * 1e-3) / 1e9, 1),
                           "bytes_per_site": rb32, "recon": L32.recon}
    log(f"[bench] FP32 Dslash {ms32 * 1e3:.1f} us  {extra['dslash_f32']['GBs']} GB/s")
    del raw32

    # Sloppy Dslash with 16-bit fixed-point links (per link 16-bit elements and
    # a float scale: 80 B per fat pair, 64 B per 12-real Naik pair), FP32 vectors
    A.dslash_sloppy(L64, vout32, vin32, A.EVEN, 16, args.warmup)
    ms16 = A.dslash_sloppy(L64, vout32, vin32, A.EVEN, 16, args.steps)
    rb16 = 4 * 80 + 4 * (64 if L64.recon == 12 else 80) + (1 if L64.recon == 12 else 0) + 2 * 24
    extra["dslash_l16"] = {"us": round(ms16 * 1e3, 2),
                           "GBs": round(rb16 * lat.vh / (ms16 * 1e-3) / 1e9, 1),
                           "bytes_per_site": rb16, "recon": L64.recon}
    log(f"[bench] 16-bit-link Dslash {ms16 * 1e3:.1f} us  {extra['dslash_l16']['GBs']} GB/s")

    # The mixed solver's FP32 sloppy Dslash: FP32 fat links, long links in their
    # 8-real form (recon8), FP32 vectors
    A.dslash_sloppy(L64, vout32, vin32, A.EVEN, 32, args.warmup)
    ms8 = A.dslash_sloppy(L64, vout32, vin32, A.EVEN, 32, args.steps)
    rb8 = site_bytes(4, 8, 24)
    extra["dslash_f32_long8"] = {"us": round(ms8 * 1e3, 2),
                                 "GBs": round(rb8 * lat.vh / (ms8 * 1e-3) / 1e9, 1),
                                 "bytes_per_site": rb8, "recon": 8}
    log(f"[bench] FP32 Dslash, 8-real long links {ms8 * 1e3:.1f} us  {extra['dslash_f32_long8']['GBs']} GB/s")

    if not args.no_cg:
        extra["cg"] = cg_compare(A, s, L64, lat, src, raw)

    cpu = None
    if not args.no_cpu:
        dt, reps, threads = cpu_baseline(DIMS, raw, src, args.cpu_seconds)
        phys, aff = host_cores()
        cpu = {"value": round(rb * lat.vh / dt / 1e9, 2), "unit": "GB/s", "cores": threads,
               "cores_note": "OpenMP threads used; the affinity mask allows "
                             f"{aff} CPUs of a host with {phys} physical cores",
               "affinity_cpus": aff, "host_physical_cores": phys,
               "kind": "port",
               "sample": f"median of {reps} FP64 single-parity Dslash applications (after 2 "
                         f"warm-ups) on the 32^4 weak-field lattice, oracle OpenMP site loop, "
                         f"{dt * 1e3:.1f} ms each; reported baseline only",
               "cpu_model": cpu_model()}
        log(f"[bench] CPU oracle Dslash {dt * 1e3:.1f} ms ({threads} threads)")

    out = {
        "metric": METRIC, "value": round(gbs, 1), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "32^4 even-odd HISQ/asqtad Dslash (fat 1-hop + Naik 3-hop), "
                               "weak-field links eps=0.3, FP64, Naik links as 12 reals",
                   "lattice": list(DIMS), "sites_per_step": lat.vh},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": measured_traffic()},
        "cpu_baseline": cpu,
        "extra": extra,
    }
    for o in (L64, vin, vout, L32, vin32, vout32):
        o.free()
    s.close()
    if not args.no_batch:
        extra["multishift_batched"] = batch_multishift(A, F)
    if not args.no_force:
        extra["forces_24"] = force_24(A, F)
    if not args.no_combined:
        extra["combined_48x96"] = combined_48(A, F)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
