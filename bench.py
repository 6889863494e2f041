"""Benchmark: FP64 brick stencil sweeps at 512^3 on one MI355X.

Workload (the headline): 7-point star, FP64, 512^3 interior cells plus an
8-cell ghost zone (66^3 bricks of 8x8x8), inputs U[-1,1], per-axis
symmetric coefficients U(0,1) normalised to sum 1, periodic ghost-brick
refresh every sweep, ping-pong sweeps.  One "step" is one sweep over all
262144 interior bricks whose stores also refresh the periodic ghost bricks
of its output (the fused form: the input's ghosts are refreshed once before
the first step; the separate-refresh form is reported in by_stencil); the
reported value is
interior cells per second of the timed steps (GStencil/s), timed with HIP
events recorded on the stream the kernels run on.

The same JSON line also carries:
  roofline     -- the sweep kernel alone (no refresh): algorithmic bytes
                  (16 B/point: one FP64 read + one write) / its mean launch
                  time, against the 8 TB/s HBM3E peak; `measured_copy_gbs`
                  is a plain device copy kernel on the same box;
  cpu_baseline -- the reference's CPU brick loops with OpenMP on the host
                  cores, per stencil, with the CPU model and core counts:
                  its own 7-point and cond loops (stencils/3axis.cpp:51-66,
                  :123-140) compiled from its headers, the oracle's
                  restatement of the 13/25/125-point tap sets;
  by_stencil   -- 13/25-point, 125-point cube and cond at the same size,
                  the FP32 / BF16 forms of the 7-point star, FP32 25-point
                  and cube, and the 7/25-point block kernels with the
                  i-direction taps from DPP lane shifts (*_lane_shift);
  precision    -- FP32 and BF16-storage forms against the FP64 sweep after
                  1, 10 and 100 sweeps at 512^3 and 1024^3 (errors computed
                  in FP64 on the GPU) and their speed-up over FP64 per step
                  (separate refresh kernel each sweep; for the 7-point star
                  also with the refresh fused into the sweep, the headline's
                  form);
  serving      -- continuous batching of independent solves (edges 32..256,
                  7/13/25/125-point, FP64/FP32, 1..200 sweeps): Poisson
                  arrivals at 200 req/s (p50/p99 latency), a saturating run
                  (all requests present at t=0) against the single-grid
                  512^3 rates, and a sample's errors against the same solve
                  alone in FP64.

Usage: python bench.py [--gpus 1] [--steps 100] [--warmup 10]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# default seconds of untimed sweeps before each secondary case (see time_sweeps)
WARM_S = 0.5


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=100)
    p.add_argument("--warmup", type=int, default=10)
    p.add_argument("--n", type=int, default=512, help="interior cells per axis (multiple of 8)")
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-seconds", type=float, default=20.0,
                   help="host seconds for the CPU baseline, split over its five stencils")
    p.add_argument("--only", default="", help="comma list of extra cases to run (default all)")
    p.add_argument("--no-extra", action="store_true")
    p.add_argument("--no-precision", action="store_true")
    p.add_argument("--no-serve", action="store_true")
    p.add_argument("--no-tune", action="store_true",
                   help="skip the autotuner table (brick shapes, folds, schedules at n^3)")
    p.add_argument("--warm-s", type=float, default=WARM_S,
                   help="seconds of untimed sweeps before each by-stencil and precision timing")
    p.add_argument("--serve-requests", type=int, default=600)
    p.add_argument("--serve-rate", type=float, default=200.0)
    p.add_argument("--serve-tune", type=int, default=0,
                   help="1: each request's schedule from the autotuner per size class")
    p.add_argument("--precision-1024", action="store_true", default=True,
                   help="also run the precision study at twice the size (1024^3)")
    return p.parse_args()


def time_sweeps(bk, st, a, b, g, stream, steps, warmup, refresh=True, warm_s=0.0, **kw):
    """ms per sweep over `steps` ping-pong sweeps (events on `stream`).
    `warmup` untimed sweeps first; with warm_s > 0, untimed sweeps go on
    for at least warm_s seconds: the GPU's clock rises over the first
    fraction of a second of sustained compute-bound work (the FP64 cube runs
    about a quarter slower in its first ~20 ms), and a run of hundreds of
    sweeps spends its time at the sustained rate."""
    from bricklib_amd import Event
    if warmup:
        bk.sweeps(st, a, b, g, warmup, refresh=refresh, stream=stream, **kw)
    stream.sync()
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < warm_s:
        bk.sweeps(st, a, b, g, 10, refresh=refresh, stream=stream, **kw)
        stream.sync()
    e0, e1 = Event(), Event()
    e0.record(stream)
    bk.sweeps(st, a, b, g, steps, refresh=refresh, stream=stream, **kw)
    e1.record(stream)
    e1.sync()
    return e0.elapsed_ms(e1) / steps


def copy_peak(bk, stream, nbytes=2 << 30, reps=20):
    """Measured HBM bandwidth (read + write bytes / time) of three device
    copies of `nbytes`: the library's brick copy (ghost-refresh) kernel,
    hipMemcpy device-to-device and the library's streaming copy kernel.
    The best of them is the measured peak."""
    from bricklib_amd import DeviceBuffer, Event
    from bricklib_amd import _lib
    import ctypes as C
    nbr = nbytes // 4096
    buf = DeviceBuffer(2 * nbr * 4096)
    pairs = np.empty(2 * nbr, np.uint32)
    pairs[0::2] = np.arange(nbr, 2 * nbr, dtype=np.uint32)
    pairs[1::2] = np.arange(nbr, dtype=np.uint32)
    dp = DeviceBuffer.from_array(pairs)
    lib = bk.load()

    def run():
        _lib.check(lib.bk_ghost_refresh(0, C.c_void_p(buf.ptr), C.c_size_t(512),
                                        C.c_size_t(512), C.c_void_p(dp.ptr), C.c_uint32(nbr),
                                        C.c_void_p(int(stream))))
    def d2d():
        _lib.check(lib.bk_memcpy_d2d(C.c_void_p(buf.ptr + nbr * 4096), C.c_void_p(buf.ptr),
                                     C.c_size_t(nbr * 4096), C.c_void_p(int(stream))))

    def scopy():
        _lib.check(lib.bk_stream_copy(C.c_void_p(buf.ptr + nbr * 4096), C.c_void_p(buf.ptr),
                                      C.c_size_t(nbr * 4096), C.c_void_p(int(stream))))

    out = {}
    for name, fn in (("brick_copy_kernel", run), ("hip_memcpy_d2d", d2d),
                     ("stream_copy_kernel", scopy)):
        fn()
        stream.sync()
        e0, e1 = Event(), Event()
        e0.record(stream)
        for _ in range(reps):
            fn()
        e1.record(stream)
        e1.sync()
        ms = e0.elapsed_ms(e1) / reps
        out[name] = round(2 * nbr * 4096 / (ms * 1e-3) / 1e9, 1)
    return max(out.values()), out


def valu_peak(bk, stream, warm_s=0.3):
    """Sustained VALU rate of the part (TFLOP/s, FP64 and FP32 without
    packed instructions): the best of the library's register-only
    multiply-add probe over 1 and 8 independent chains per lane at 4 and 8
    waves per SIMD, after warm_s seconds of untimed probe launches."""
    from bricklib_amd import DeviceBuffer, Event
    from bricklib_amd import _lib
    import ctypes as C
    lib = bk.load()
    n = 8 * 1024 * 4 * 64  # up to 8 waves on each SIMD of up to 1024 CUs
    out = DeviceBuffer(n * 8)
    fl = C.c_double(0)
    best = {}
    for dt, name in ((bk.BK_F64, "f64"), (bk.BK_F32, "f32")):
        def launch(chains, waves, iters):
            _lib.check(lib.bk_valu_probe(dt, chains, waves, iters, C.c_void_p(out.ptr), n,
                                         C.byref(fl), C.c_void_p(int(stream))))
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < warm_s:
            launch(8, 8, 2000)
            stream.sync()
        rates = {}
        for chains in (1, 8):
            for waves in (4, 8):
                iters = 8000 // waves
                launch(chains, waves, iters)
                e0, e1 = Event(), Event()
                e0.record(stream)
                for _ in range(3):
                    launch(chains, waves, iters)
                e1.record(stream)
                e1.sync()
                rates[f"chains{chains}_waves{waves}"] = round(
                    3 * fl.value / (e0.elapsed_ms(e1) * 1e-3) / 1e12, 2)
        best[name] = {"tflops": max(rates.values()), "runs": rates}
    return best


def host_cpu():
    """CPU model, physical cores and logical CPUs of this host (/proc/cpuinfo)."""
    model, phys, logical = "unknown", set(), 0
    try:
        with open("/proc/cpuinfo") as f:
            pid = cid = None
            for line in f:
                k, _, v = line.partition(":")
                k, v = k.strip(), v.strip()
                if k == "processor":
                    logical += 1
                elif k == "model name" and model == "unknown":
                    model = v
                elif k == "physical id":
                    pid = v
                elif k == "core id":
                    cid = v
                    phys.add((pid, cid))
    except OSError:
        pass
    return {"cpu_model": model, "physical_cores": len(phys) or None, "logical_cpus": logical,
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline(nb, seconds):
    """The reference's CPU brick loops on the host, OpenMP over the interior
    bricks, each timed like time_func (stencils/stencils.h:40-53: one
    warm-up, then whole sweeps with the periodic refresh until its share of
    `seconds` has passed), over the full grid.  The 7-point star and cond
    run the reference's own loops (stencils/3axis.cpp:51-66, :123-140)
    through its own Brick accessor, compiled from its headers (oracle/_ref;
    kind "reference"); the 13/25/125-point stencils exist in the reference
    only as generated code, so they run the oracle's restatement of their
    tap sets in the reference's order (kind "port"), as do all five when
    oracle/_ref is absent.  `value` is the 7-point star's rate."""
    import ctypes as C

    import oracle as O
    from bricklib_amd import stencils as S
    g = O.Grid(nb)
    a = g.storage()
    b = g.storage()
    g.fill_uniform(a, 1234)
    g.refresh(a)
    R = O.ref() if O.ref_available() else None
    acc = O.lib().or_brick_sweep_7pt_accessor
    pts = g.interior.size * 512
    by = {}
    names = ("7pt", "13pt", "25pt", "125pt", "cond")
    for name in names:
        # cond in the reference's form (p = sum c max(x, 0), |p|) for both kinds
        st = S.Stencil.make(name, seed=1, cond_mode=0)
        c = np.ascontiguousarray(st.coeff)
        ref_loop = R is not None and name in ("7pt", "cond")

        def sweep(src, dst):
            if ref_loop:
                fn = R.ref_sweep_7pt if name == "7pt" else R.ref_sweep_cond
                fn(O.P(g.adj), g.nbricks, O.P(g.interior), g.interior.size, O.P(src),
                   C.c_size_t(512), O.P(dst), C.c_size_t(512), O.P(c))
            elif name == "7pt":
                acc(O.P(g.adj, O.u32p), O.P(g.interior, O.u32p), g.interior.size,
                    O.P(src, O.dblp), 512, O.P(dst, O.dblp), 512, O.P(c, O.dblp))
            else:
                g.sweep(st.kind, c, src, dst, mode=O.MODE_REF, theta=st.theta,
                        cond_mode=st.cond_mode)
        sweep(a, b)  # warm-up
        n, src, dst = 0, a, b
        t0 = time.perf_counter()
        while True:
            g.refresh(src)
            sweep(src, dst)
            src, dst = dst, src
            n += 1
            el = time.perf_counter() - t0
            if el >= seconds / len(names) or n >= 1000:
                break
        by[name] = {"gstencil_s": round(pts * n / el / 1e9, 4), "sweeps": n,
                    "seconds": round(el, 2), "kind": "reference" if ref_loop else "port"}
    info = host_cpu()
    threads = int(O.lib().or_num_threads())
    kind = by["7pt"]["kind"]
    return {"value": by["7pt"]["gstencil_s"], "unit": "GStencil/s", "cores": threads,
            "kind": kind,
            "sample": f"FP64 {8 * nb[0]}^3 grid, whole sweeps with the periodic refresh, "
                      f"{seconds / len(names):.1f} s per stencil (value: the 7-point star, "
                      + ("the reference's stencils/3axis.cpp brick loop compiled from its "
                         "headers" if kind == "reference" else "the oracle's restatement")
                      + ")",
            "threads": threads, **info, "by_stencil": by}


def serve_bench(bk, S, args):
    """Batched serving: Poisson arrivals at --serve-rate (p50/p99 latency),
    then the same kind of stream all present at t=0 (saturating) for the
    throughput, compared with the single-grid rates of each (stencil,
    dtype) at 512^3; per-request errors of a sample against FP64 solves of
    the same requests run alone (the FP64 GPU oracle)."""
    from bricklib_amd import BK_F32, BK_F64, precision as PR, serve as SV
    from bricklib_amd.batch import Request
    out = {}
    # single-grid reference rates at 512^3 in the serving loop's own field
    # layout (the wrapped adjacency: no ghost bricks, no refresh), default
    # schedule, so the efficiency compares like with like
    serve_ghost = 0
    g = bk.PeriodicGrid((64, 64, 64), ghost=serve_ghost)
    stream = bk.Stream()
    rates = {}
    for name in ("7pt", "13pt", "25pt", "125pt"):
        for dt in (BK_F64, BK_F32):
            st = S.Stencil.make(name, seed=3)
            fa = bk.Brick(g.info, g.field_storage(dt), 0)
            fb = bk.Brick(g.info, g.field_storage(dt), 0)
            g.fill_uniform(fa, 3, stream=stream)
            # warmed like the by-stencil cases: a cold GPU would understate
            # the single-grid rates (the FP64 cube by about a third) and so
            # overstate the batched efficiency
            ms = time_sweeps(bk, st, fa, fb, g, stream, 10, 2, refresh=bool(serve_ghost),
                             warm_s=args.warm_s)
            rates[(S.NAMES[name], dt)] = g.cells / (ms * 1e-3)
            del fa, fb
    del g
    out["single_grid_gstencil_s"] = {f"{SV.stencil_name(k)}_{PR.DTYPE_NAMES[d]}": round(v / 1e9, 2)
                                     for (k, d), v in rates.items()}
    out["layout"] = ("wrapped adjacency, no ghost bricks or refresh, for the batched requests "
                     "and the single-grid rates alike" if serve_ghost == 0 else
                     "one-brick ghost zone refreshed every sweep, batched and single grid alike")

    # real-time Poisson stream
    reqs = SV.workload(args.serve_requests, rate=args.serve_rate, seed=2024)
    # chunks of 16 ticks: at this load a chunk's GPU time is short and the
    # host's round trip per chunk adds to every request's latency (p50 / p99
    # 6.2 / 24.4 ms with chunks of 8, 4.8 / 21.7 with 16, 4.4 / 22.5 with 32;
    # the saturating run alike for all)
    srv = SV.Server(hbm_budget=200e9, chunk=16, ghost=serve_ghost, tune=args.serve_tune)
    t_pw = time.perf_counter()
    srv.prewarm(reqs)
    out["prewarm_s"] = round(time.perf_counter() - t_pw, 3)
    sample = set(r.rid for r in reqs[::max(1, len(reqs) // 12)])
    st = srv.serve(reqs, realtime=True, keep_results=sample)
    rt = st.summary()
    rt["rate_req_s"] = args.serve_rate
    rt["offered_gstencil_s"] = round(sum(r.cells * r.steps for r in reqs) /
                                     max(reqs[-1].arrival, 1e-9) / 1e9, 2)
    rt["schedules"] = sorted({f"{SV.stencil_name(r.stencil.kind)}:{r.schedule[0]}/{r.schedule[1]}"
                              for r in reqs})
    out["poisson"] = rt
    print(f"[bench] serving poisson: {rt}", file=sys.stderr, flush=True)

    # per-request error of the sample against the same request alone in FP64
    errs = []
    for r in reqs:
        if r.rid not in sample:
            continue
        ref = Request(rid=-1, nb=r.nb, stencil=r.stencil, dtype=BK_F64, steps=r.steps,
                      seed=r.seed)
        srv.serve([ref], realtime=False, keep_results=True)
        gr = bk.PeriodicGrid(r.nb, ghost=serve_ghost)
        a = bk.Brick(gr.info, bk.BrickStorage(_View(r.result_ptr, r.field_bytes), gr.nbricks,
                                              512, r.dtype), 0)
        b = bk.Brick(gr.info, bk.BrickStorage(_View(ref.result_ptr, ref.field_bytes),
                                              gr.nbricks, 512, BK_F64), 0)
        c = PR.compare(gr, a, b)
        errs.append({"rid": r.rid, "stencil": SV.stencil_name(r.stencil.kind),
                     "dtype": PR.DTYPE_NAMES[r.dtype], "steps": r.steps, "nb": list(r.nb),
                     "max_rel": c["max_rel"], "rms_rel": c["rms_rel"],
                     "underflow": c["underflow"]})
        r.release()
        ref.release()
    out["sample_errors_vs_fp64"] = errs

    # saturating run: the same mix with every request present at t=0
    reqs2 = SV.workload(args.serve_requests // 2, rate=args.serve_rate, seed=77)
    srv.prewarm(reqs2)
    st2 = srv.serve(reqs2, realtime=False)
    sat = st2.summary()
    single_time = sum(r.cells * r.steps / rates[(r.stencil.kind, r.dtype)] for r in reqs2)
    sat["single_grid_time_s"] = round(single_time, 4)
    sat["efficiency_vs_single_grid"] = round(single_time / st2.elapsed, 4)
    out["saturating"] = sat
    print(f"[bench] serving saturating: {sat}", file=sys.stderr, flush=True)

    # the same saturating mix with the other schedule policy (autotuned per
    # size class, or the fixed defaults)
    alt = not args.serve_tune
    srv2 = SV.Server(hbm_budget=200e9, chunk=16, ghost=serve_ghost, tune=alt)
    t_pw = time.perf_counter()
    reqs3 = SV.workload(args.serve_requests // 2, rate=args.serve_rate, seed=77)
    srv2.prewarm(reqs3)
    pw = time.perf_counter() - t_pw
    st3 = srv2.serve(reqs3, realtime=False)
    sat2 = st3.summary()
    sat2["prewarm_s"] = round(pw, 3)
    sat2["efficiency_vs_single_grid"] = round(single_time / st3.elapsed, 4)
    sat2["schedules"] = sorted({f"{SV.stencil_name(r.stencil.kind)}:{r.schedule[0]}/"
                                f"{r.schedule[1]}" for r in reqs3})
    out["saturating_tuned" if alt else "saturating_default_schedules"] = sat2
    print(f"[bench] serving saturating ({'tuned' if alt else 'default'} schedules): {sat2}",
          file=sys.stderr, flush=True)
    # and with the stars one brick per wave (the batched schedule before the
    # 2x2 brick blocks)
    srv3 = SV.Server(hbm_budget=200e9, chunk=16, ghost=serve_ghost, star_plan=None)
    reqs4 = SV.workload(args.serve_requests // 2, rate=args.serve_rate, seed=77)
    srv3.prewarm(reqs4)
    st4 = srv3.serve(reqs4, realtime=False)
    sat3 = st4.summary()
    sat3["efficiency_vs_single_grid"] = round(single_time / st4.elapsed, 4)
    out["saturating_stars_per_brick"] = sat3
    print(f"[bench] serving saturating (stars one brick per wave): {sat3}", file=sys.stderr,
          flush=True)
    return out


class _View:
    """A non-owning view of device memory that a Request owns."""

    def __init__(self, ptr, nbytes):
        self.ptr = ptr
        self.nbytes = nbytes


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print("bench.py: runs on one GPU only; sharding the sweep across GPUs is designed "
              "but neither implemented nor measured", file=sys.stderr)
        sys.exit(2)
    import bricklib_amd as bk
    from bricklib_amd import stencils as S

    if bk.device_count() < 1:
        print("bench.py: no HIP device", file=sys.stderr)
        sys.exit(1)
    n = args.n
    nb = (n // 8,) * 3
    stream = bk.Stream()
    t_setup = time.perf_counter()
    g = bk.PeriodicGrid(nb, stream=stream)
    pts = g.cells
    stream.sync()

    def fields(dtype):
        sa, sb = g.field_storage(dtype), g.field_storage(dtype)
        a, b = bk.Brick(g.info, sa, 0), bk.Brick(g.info, sb, 0)
        g.fill_uniform(a, 1234, stream=stream)
        g.fill_uniform(b, 4321, stream=stream)
        return a, b

    a, b = fields(bk.BK_F64)
    st7 = S.Stencil.make("7pt", seed=1)
    print(f"[bench] setup {time.perf_counter() - t_setup:.1f}s, {g.nbricks} bricks, "
          f"{g.npairs} ghost bricks", file=sys.stderr, flush=True)

    # headline: K timed steps, each a sweep whose stores also refresh the
    # periodic ghost bricks of its output (the input's ghosts are refreshed
    # once first): bk_sweeps_fused, the same bits as a separate refresh
    # before every sweep (that form is by_stencil's 7pt_f64_separate_refresh)
    pl7 = g.auto_plan(st7.kind, bk.BK_F64)
    sched7 = ("one brick per wave" if pl7 is None else
              f"2x2 brick blocks ({pl7.shape} plan, {pl7.kcol} level(s) per unit), one "
              f"workgroup of four waves per block")
    ms_step = time_sweeps(bk, st7, a, b, g, stream, args.steps, args.warmup, fused=True,
                          warm_s=args.warm_s)
    # the dominant kernel alone, for the roofline
    ms_kernel = time_sweeps(bk, st7, a, b, g, stream, args.steps, 2, refresh=False)
    peak_copy, copies = copy_peak(bk, stream)
    alg_bytes = 16 * pts
    achieved = alg_bytes / (ms_kernel * 1e-3) / 1e9
    print(f"[bench] 7pt f64: {ms_step:.4f} ms/step, kernel {ms_kernel:.4f} ms, "
          f"{achieved:.0f} GB/s, copy peak {peak_copy:.0f} GB/s", file=sys.stderr, flush=True)

    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"star7_f64_{n}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("hbm_bytes_per_launch")

    by = {}
    cur = {}
    vpeak = {}
    if not args.no_extra:
        vpeak = valu_peak(bk, stream)
        print(f"[bench] measured VALU peak: {vpeak}", file=sys.stderr, flush=True)
        only = set(x for x in args.only.split(",") if x)
        # the *_lane_shift cases: the same default schedule (2x2 brick
        # blocks) with the i-direction taps from DPP lane shifts, each right
        # after its LDS-tile twin (the clocks after the VALU-heavy cube and
        # probe differ: compared across them, the two forms swap places)
        cases = [("7pt_f64_tile_taps", "7pt", bk.BK_F64, {"lane_shift": 0}),
                 ("7pt_f64_lane_shift", "7pt", bk.BK_F64, {"lane_shift": 1}),
                 ("13pt_f64", "13pt", bk.BK_F64, {}), ("25pt_f64", "25pt", bk.BK_F64, {}),
                 ("25pt_f64_lane_shift", "25pt", bk.BK_F64, {"lane_shift": 1}),
                 ("cond_f64", "cond", bk.BK_F64, {}), ("125pt_f64", "125pt", bk.BK_F64, {}),
                 ("7pt_f32", "7pt", bk.BK_F32, {}), ("25pt_f32", "25pt", bk.BK_F32, {}),
                 ("25pt_f32_lane_shift", "25pt", bk.BK_F32, {"lane_shift": 1}),
                 ("7pt_bf16", "7pt", bk.BK_BF16, {}), ("125pt_f32", "125pt", bk.BK_F32, {})]
        cur = {bk.BK_F64: (a, b)}
        for name, sname, dt, kw in cases:
            if only and name not in only:
                continue
            if dt not in cur:
                cur = {bk.BK_F64: (a, b), dt: fields(dt)}
            fa, fb = cur[dt]
            st = S.Stencil.make(sname, seed=1)
            steps = max(10, args.steps // (5 if sname == "125pt" else 1))
            ms_s = time_sweeps(bk, st, fa, fb, g, stream, steps, 2, warm_s=args.warm_s, **kw)
            ms_k = time_sweeps(bk, st, fa, fb, g, stream, steps, 1, refresh=False,
                               warm_s=args.warm_s / 2, **kw)
            es = {bk.BK_F64: 8, bk.BK_F32: 4, bk.BK_BF16: 2}[dt]
            fl = S.flops_per_point(st.kind)
            gb = 2 * es * pts / (ms_k * 1e-3) / 1e9
            tf = fl * pts / (ms_k * 1e-3) / 1e12
            peak_tf = FP64_PEAK_TFLOPS if dt == bk.BK_F64 else FP32_PEAK_TFLOPS
            by[name] = {"gstencil_s": round(pts / (ms_s * 1e-3) / 1e9, 2),
                        "kernel_gstencil_s": round(pts / (ms_k * 1e-3) / 1e9, 2),
                        "ms_per_step": round(ms_s, 4), "kernel_ms": round(ms_k, 4),
                        "hbm_gbs": round(gb, 1), "hbm_frac": round(gb / HBM_PEAK_GBS, 4),
                        "tflops": round(tf, 2), "valu_frac": round(tf / peak_tf, 4)}
            vm = vpeak.get("f64" if dt == bk.BK_F64 else "f32", {}).get("tflops")
            if sname == "125pt" and vm:
                # the measured peak is the non-packed rate for both dtypes
                by[name]["valu_frac_of_measured"] = round(tf / vm, 4)
            print(f"[bench] {name}: {by[name]}", file=sys.stderr, flush=True)
        # the same 7-point FP64 loop in other forms: a separate refresh
        # kernel before every sweep (the reference's order,
        # weak/main.hip.cpp:249-279), the same in the BrickDecomp layout
        # (refresh = contiguous ghost runs from skin runs), the refresh once
        # per 8 sweeps (sweeps in between over every brick,
        # weak/main.cpp:26-36), and the wrapped-adjacency layout (no ghost
        # bricks, no refresh)
        if not only or "7pt_f64_forms" in only:
            cur.clear()
            ms_deep = time_sweeps(bk, st7, a, b, g, stream, args.steps, 2, ghost_iters=8)
            ms_sep = time_sweeps(bk, st7, a, b, g, stream, args.steps, 2)
            from bricklib_amd.decomp import DecompGrid
            gd = DecompGrid(nb, stream=stream)
            da = bk.Brick(gd.info, gd.field_storage(bk.BK_F64), 0)
            db = bk.Brick(gd.info, gd.field_storage(bk.BK_F64), 0)
            gd.fill_uniform(da, 1234, stream=stream)
            ms_decomp = time_sweeps(bk, st7, da, db, gd, stream, args.steps, 2)
            del da, db, gd
            gw = bk.PeriodicGrid(nb, stream=stream, ghost=0)
            wa = bk.Brick(gw.info, gw.field_storage(bk.BK_F64), 0)
            wb = bk.Brick(gw.info, gw.field_storage(bk.BK_F64), 0)
            gw.fill_uniform(wa, 1234, stream=stream)
            ms_wrap = time_sweeps(bk, st7, wa, wb, gw, stream, args.steps, 2)
            del wa, wb, gw
            for name, ms_s in (("7pt_f64_separate_refresh", ms_sep),
                               ("7pt_f64_brickdecomp_layout", ms_decomp),
                               ("7pt_f64_refresh_every_8", ms_deep),
                               ("7pt_f64_wrapped_no_ghost", ms_wrap)):
                by[name] = {"gstencil_s": round(pts / (ms_s * 1e-3) / 1e9, 2),
                            "ms_per_step": round(ms_s, 4)}
                print(f"[bench] {name}: {by[name]}", file=sys.stderr, flush=True)

    tuned = {}
    if not args.no_tune:
        # the autotuner's table at n^3 (FP64, ghost layer + refresh every
        # sweep): every brick layout (8x8x8, 4x8x16 in the 4x16 and the 8x8
        # fold) and every schedule of each, fastest first
        cur.clear()
        for sname, dt, layouts in (("7pt", bk.BK_F64, (0, 1, 2)), ("13pt", bk.BK_F64, (0, 1, 2)),
                                   ("25pt", bk.BK_F64, (0, 1, 2)), ("125pt", bk.BK_F64, (0, 1, 2)),
                                   ("7pt", bk.BK_F32, (0,)), ("25pt", bk.BK_F32, (0,)),
                                   ("7pt", bk.BK_BF16, (0,))):
            st = S.Stencil.make(sname, seed=1)
            t0 = time.perf_counter()
            res = bk.tune(st, (n, n, n), dt, ghost=1, steps=10, rounds=2,
                          warm_s=args.warm_s / 2, layouts=layouts, stream=stream)
            key = f"{sname}_{'f64' if dt == bk.BK_F64 else 'f32' if dt == bk.BK_F32 else 'bf16'}"
            tuned[key] = {"best": bk.cand_name(res[0]),
                          "ms_per_step": {bk.cand_name(c): round(c.ms, 4) for c in res},
                          "tune_s": round(time.perf_counter() - t0, 2)}
            print(f"[bench] tune {key}: {tuned[key]}", file=sys.stderr, flush=True)

    prec = {}
    if not args.no_precision:
        from bricklib_amd import precision as PR
        cur.clear()
        del a, b
        for pn in sorted(set([n, 2 * n])) if args.precision_1024 else [n]:
            for sname, dts in (("7pt", (bk.BK_F32, bk.BK_BF16)), ("125pt", (bk.BK_F32,))):
                key = f"{sname}_{pn}"
                prec[key] = PR.study(pn, sname, dtypes=dts, sweeps=(1, 10, 100), stream=stream)
                gp = bk.PeriodicGrid((pn // 8,) * 3, stream=stream)
                stp = S.Stencil.make(sname, seed=7)
                times, fused = {}, {}
                for dt in (bk.BK_F64, bk.BK_F32) + ((bk.BK_BF16,) if sname == "7pt" else ()):
                    fa = bk.Brick(gp.info, gp.field_storage(dt), 0)
                    fb = bk.Brick(gp.info, gp.field_storage(dt), 0)
                    gp.fill_uniform(fa, 7, stream=stream)
                    times[dt] = time_sweeps(bk, stp, fa, fb, gp, stream, 20, 2,
                                            warm_s=args.warm_s)
                    if sname == "7pt":
                        # the headline's form: the refresh folded into the
                        # sweep's stores (stars only)
                        fused[dt] = time_sweeps(bk, stp, fa, fb, gp, stream, 20, 2,
                                                warm_s=args.warm_s / 2, fused=True)
                    del fa, fb
                prec[key]["ms_per_step"] = {PR.DTYPE_NAMES[d]: round(t, 4) for d, t in times.items()}
                prec[key]["speedup_vs_f64"] = {PR.DTYPE_NAMES[d]: round(times[bk.BK_F64] / t, 3)
                                               for d, t in times.items()}
                if fused:
                    prec[key]["ms_per_step_fused_refresh"] = {
                        PR.DTYPE_NAMES[d]: round(t, 4) for d, t in fused.items()}
                    prec[key]["speedup_vs_f64_fused_refresh"] = {
                        PR.DTYPE_NAMES[d]: round(fused[bk.BK_F64] / t, 3) for d, t in fused.items()}
                del gp
                print(f"[bench] precision {key}: {prec[key]}", file=sys.stderr, flush=True)

    serving = {}
    if not args.no_serve:
        serving = serve_bench(bk, S, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(nb, args.cpu_seconds)
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    value = pts / (ms_step * 1e-3) / 1e9
    line = {
        "metric": "GStencil/s + % HBM roofline at 512^3 (7pt/125pt); p99 latency batched solves",
        "value": round(value, 3),
        "unit": "GStencil/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded U[-1,1] inputs, seeded coefficients)",
        "config": {"workload": f"7-point star FP64, {n}^3 interior + 8-cell ghost, 8x8x8 bricks "
                               f"(fold 8x8), periodic ghost-brick refresh every sweep (fused into "
                               f"the sweep's stores), ping-pong sweeps", "interior_cells": pts,
                   "schedule": sched7,
                   "bricks": g.nbricks, "ghost_bricks": g.npairs},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "kernel_ms": round(ms_kernel, 5),
                     "algorithmic_bytes_per_launch": alg_bytes,
                     "measured_copy_gbs": round(peak_copy, 1), "copy_kernels_gbs": copies,
                     "frac_of_measured_copy": round(achieved / peak_copy, 4)},
        "cpu_baseline": cpu,
        "valu_peak_measured": vpeak,
        "by_stencil": by,
        "precision": prec,
        "tuner": tuned,
        "serving": serving,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
