"""Benchmark: Wilson-clover dslash on one MI355X at 32^3 x 64.

One step = the even-odd clover operator on one checkerboard, i.e. the two
hot kernels back to back, each over 1,048,576 sites:
    t   = A_ee^-1 D_eo psi_o          clov_dslash (cb 0)
    out = A_oo chi_o - beta D_oe t    clov_dslash_achimbdpsi (cb 1)
FP32 storage and arithmetic, 12-real links, packed Hermitian clover.
Inputs are synthetic (seeded random SU(3) links, Gaussian spinors, random
clover blocks and their FP64-computed inverses) and resident in HBM before
timing starts; two input/output sets alternate so the Infinity Cache cannot
hold a step's working set.

Prints one JSON line: metric, value = GFLOP/s of the step, the roofline of
the dominant kernel from HIP-event timing, every reduced-precision kernel's
error against the FP64 kernel, "cg" = FP64 vs mixed-precision CG time to a
true residual of 1e-10 (the 32^3x64 Wilson-clover solve), "cfg4" = the
48^3x96 twisted-mass clover case (dslash throughput and error, FP64 vs
FP16-mixed CG, the pseudofermion force and its error), "shard" = one GPU's
share of the 32^3x64 operator split in t over 2, 4 and 8 GPUs (interior and
face kernels, critical path), and the CPU baseline (the reference's own
AVX-512/AVX2 kernels on the host cores).
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "dslash GFLOP/s + % HBM roofline @32^3x64; mixed vs FP64 CG time to 1e-10"
DIMS = (32, 32, 32, 64)
PEAK_HBM_GBS = 8000.0
FLOP_CLOVER = 1824          # flop/site, clover dslash
FLOP_CLOVER_ACHI = 1872     # flop/site, clover a*chi - b*D*psi (A chi + hop)
BYTES_F32 = 4


def algorithmic_bytes(recon, clover_reals, chi, sbytes):
    """Compulsory bytes per output site: input spinor once, 8 links, the
    clover blocks, chi, the output spinor."""
    return (24 + 8 * recon + clover_reals + (24 if chi else 0) + 24) * sbytes


def time_kernel(fn, n):
    import torch
    s, e = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    fn(0)
    torch.cuda.synchronize()
    s.record()
    for i in range(n):
        fn(i)
    e.record()
    e.synchronize()
    return s.elapsed_time(e) / n * 1e-3


def _host_flags():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("flags"):
                return set(line.split(":", 1)[1].split())
    except OSError:
        pass
    return set()


def host_cpu():
    """The host CPUs this process may run on: model, logical CPUs and
    physical cores (distinct (package, core) pairs among them)."""
    try:
        allowed = os.sched_getaffinity(0)
    except AttributeError:
        allowed = set(range(os.cpu_count() or 1))
    model, cores = None, set()
    try:
        text = open("/proc/cpuinfo").read()
    except OSError:
        text = ""
    for block in text.split("\n\n"):
        d = {}
        for line in block.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                d[k.strip()] = v.strip()
        if "processor" not in d or int(d["processor"]) not in allowed:
            continue
        model = model or d.get("model name")
        cores.add((d.get("physical id", "0"), d.get("core id", d["processor"])))
    return {"model": model, "logical_cpus": len(allowed), "physical_cores": len(cores) or len(allowed)}


# The reference's ISA configurations timed as the CPU baseline (VECLEN =
# SOALEN, codegen/jinja/isa.js): name, host flags needed, storage bytes.
CPU_CONFIGS = (("avx512_f32", {"avx512f", "fma"}, 4), ("avx512_f64", {"avx512f", "fma"}, 8),
               ("avx2_f32", {"avx2", "fma"}, 4))


def _cpu_baseline_child(seconds_per_config, dims):
    """Runs in a child process that loads no torch (and so no second OpenMP
    runtime) with OMP_PROC_BIND=close, OMP_PLACES=cores: the reference's
    generated clover kernel under an OpenMP block loop, for each of the
    AVX-512 FP32, AVX-512 FP64 and AVX2 FP32 builds the host supports, about
    `seconds_per_config` of calls each (median call time), and a STREAM
    triad on the same threads.  The inputs are Gaussian arrays of the
    workload's shapes (FP64, converted by the reference driver): the
    kernels' instruction and memory streams do not depend on the values."""
    import numpy as np
    from oracle import ref
    flags = _host_flags()
    cfgs = [(c, b) for c, need, b in CPU_CONFIGS if need <= flags and ref.available(c)]
    nthr = int(os.environ["OMP_NUM_THREADS"])
    if not cfgs:
        return None
    rng = np.random.default_rng(5)
    V = int(np.prod(dims))
    U = rng.standard_normal((V, 4, 3, 3, 2))
    psi = rng.standard_normal((V, 4, 3, 2))
    inv = rng.standard_normal((V, 2, 36))
    vcb = V // 2
    triad = ref.stream_triad(cfgs[0][0], nthr)
    by = {}
    for cfg, sb in cfgs:
        _, t = ref.apply(cfg, dims, psi, U, 1, 0, kind=1, clov=inv, recon=12, nthreads=nthr, reps=3)
        reps = max(5, int(seconds_per_config / max(float(np.median(t)), 1e-6)))
        _, t = ref.apply(cfg, dims, psi, U, 1, 0, kind=1, clov=inv, recon=12, nthreads=nthr, reps=reps)
        med = float(np.median(t))
        gbs = vcb * (24 + 96 + 72 + 24) * sb / med / 1e9  # algorithmic bytes, 12-real links
        by[cfg] = {"gflops": round(vcb * FLOP_CLOVER / med / 1e9, 2), "ms_median": round(med * 1e3, 3),
                   "ms_min": round(float(np.min(t)) * 1e3, 3), "calls": reps,
                   "veclen": ref.soalen(cfg), "gbs": round(gbs, 1),
                   "host_roofline_frac": round(gbs / triad, 3)}
    return {"triad": triad, "by": by}


def cpu_baseline(seconds_per_config=6.0, dims=DIMS):
    """The reference's own generated clover kernel (oracle/_ref, built from
    the qphix-codegen sources) on the host cores, on the whole 32^3 x 64
    lattice (0.9-1.8 GB of fields per call, so the kernel streams from host
    memory rather than from caches), in the AVX-512 FP32, AVX-512 FP64 and
    AVX2 FP32 builds the host supports; the headline value is the fastest
    FP32 build.
    It runs in a child process with OMP_PROC_BIND=close, OMP_PLACES=cores
    and one thread per CPU this process may use (OMP_NUM_THREADS when set):
    in this process torch's own OpenMP runtime would, under those settings,
    pin the main thread to one core as it loads, and the reference's
    separate runtime would then see a one-core affinity mask.
    Falls back to the FP64 C oracle (one thread, a bounded sample of sites)
    if no reference build is present."""
    import numpy as np
    cpu = host_cpu()
    nthr = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or cpu["logical_cpus"]
    env = dict(os.environ, OMP_PROC_BIND="close", OMP_PLACES="cores", OMP_NUM_THREADS=str(nthr))
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-baseline-child",
           "--cpu-seconds", str(seconds_per_config), "--lattice", ",".join(map(str, dims))]
    r = subprocess.run(cmd, env=env, capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"CPU baseline child failed:\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
    child = json.loads(r.stdout.strip().splitlines()[-1])
    if child is not None:
        by, triad = child["by"], child["triad"]
        vcb = int(np.prod(dims)) // 2
        # the headline is the strongest FP32 build on this host (AVX2 has
        # beaten AVX-512 on hosts whose AVX-512 units run at half width)
        f32 = [c for c in by if c.endswith("_f32")] or list(by)
        head = max(f32, key=lambda c: by[c]["gflops"])
        h = by[head]
        return {"value": h["gflops"], "unit": "GFLOP/s", "cores": nthr, "kind": "reference",
                "physical_cores": cpu["physical_cores"], "logical_cpus": cpu["logical_cpus"],
                "cpu_model": cpu["model"], "omp_proc_bind": "close", "omp_places": "cores",
                "host_triad_gbs": round(triad, 1), "host_roofline_frac": h["host_roofline_frac"],
                "by_config": by,
                "sample": f"reference {head} clov_dslash (VECLEN=SOALEN={h['veclen']}, 12-real links) "
                          f"on {'x'.join(map(str, dims))} ({vcb} sites/cb), median of {h['calls']} calls "
                          f"({h['ms_median']:.2f} ms), OpenMP {nthr} threads pinned one per core place; "
                          f"also " + ", ".join(c for c in by if c != head)}
    import __graft_entry__ as ge
    from oracle import oracle as orc
    L = ge.load_package().lattice
    dims = (16, 16, 16, 32)
    U = L.random_gauge(dims, 0.3, 5).numpy()
    psi = L.random_spinor(dims, 6).numpy()
    clov = L.invert_clover(L.random_clover(dims, 7)).numpy()
    sites = np.arange(0, U.shape[0], 2, dtype=np.int64)
    done, t0, n = 0, time.perf_counter(), 2048
    while time.perf_counter() - t0 < seconds_per_config:
        chunk = sites[(done % len(sites)):(done % len(sites)) + n]
        orc.dslash_sites(dims, chunk, psi, U, 1, kind=orc.CLOVER, clov=clov)
        done += len(chunk)
    dt = time.perf_counter() - t0
    return {"value": round(done * FLOP_CLOVER / dt / 1e9, 4), "unit": "GFLOP/s", "cores": 1,
            "kind": "port", "physical_cores": cpu["physical_cores"], "cpu_model": cpu["model"],
            "sample": f"FP64 C oracle, clover dslash at {done} sites of a 16^3x32 lattice, "
                      f"{dt:.1f} s, 1 thread"}


def profile_traffic():
    """HBM bytes per clov_dslash launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE x its copy-kernel calibration + WRITE_SIZE), or None."""
    path = os.path.join(ROOT, "profiles", "clov_dslash_f32_32x64_pmc.txt")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1])
    except OSError:
        pass
    return None


def tm_kernel(q, geom, s, dev, n):
    """The twisted-mass variant of the 32^3x64 operator: tm_dslash, FP32, 12-real links,
    mu = 0.1, kappa = 0.1350 (alpha = 1/(2 kappa); the kernels take mu/alpha
    and alpha/(alpha^2 + mu^2))."""
    alpha, mu = 1.0 / (2 * 0.1350), 0.1
    op = q.TMDslash(geom, alpha - 4.0, mu)
    vcb = geom.vcb

    def k(i):
        op.dslash(s["t0"], s["psi1"], s["g0"], 1, 0)

    for i in range(5):
        k(i)
    t = time_kernel(k, n)
    b = algorithmic_bytes(12, 0, False, BYTES_F32) * vcb
    return {"us": round(t * 1e6, 2), "gflops": round(vcb * 1392 / t / 1e9, 1),
            "hbm_frac": round(b / t / 1e9 / PEAK_HBM_GBS, 4), "flop_per_site": 1392}


def inplace_kernels(q, geom, sets, n):
    """clov_dslash FP32, 12-real links, reading and writing QPhiX
    FourSpinorBlock arrays in place (qphix_amd_qphix_clov_dslash; the links
    and clover blocks are the native launch's) for the reference's AVX-512
    FP32 layouts, veclen 16 with soalen 16, 8 and 4: HIP-event average and
    fraction of 8 TB/s by the same algorithmic bytes as the native kernel.
    The two input sets alternate, as for the native kernel."""
    import torch
    out = {}
    op = q.ClovDslash(geom)
    b = algorithmic_bytes(12, 72, False, BYTES_F32) * geom.vcb
    for S in (16, 8, 4):
        lay = q.qphix_layout(q.F32, 16, S, True)
        qs = []
        for s in sets:
            psi_q = q.QPhiXSpinor(geom.spinor_to_qphix(s["psi1"], lay), lay)
            qs.append((psi_q, q.QPhiXSpinor(torch.empty_like(psi_q.data), lay), s))

        def k(i):
            psi_q, res_q, s = qs[i & 1]
            op.dslash(res_q, psi_q, s["g0"], s["inv0"], 1, 0)

        for i in range(3):
            k(i)
        t = time_kernel(k, n)
        out[f"soalen{S}"] = {"us": round(t * 1e6, 2), "hbm_frac": round(b / t / 1e9 / PEAK_HBM_GBS, 4)}
        del qs
    torch.cuda.empty_cache()
    return out


def inplace_other_precisions(q, geom, dev, n):
    """clov_dslash on QPhiX arrays in place in FP64 (the reference's AVX-512
    FP64 layouts, veclen 8 with soalen 8 and 4) and FP16 (its low-precision
    build's, veclen 16 with soalen 16 and 8), each beside the native layout
    in the same precision, two input sets alternating: HIP-event averages
    and fractions of 8 TB/s by the native kernel's algorithmic bytes."""
    import torch
    L = q.lattice
    dims = geom.dims
    lex = []
    for k in range(2):
        U = L.random_gauge(dims, 0.3, 1000 + 10 * k, device=dev)
        inv = L.invert_clover(L.random_clover(dims, 1003 + 10 * k, device=dev))
        lex.append((U, inv, L.random_spinor(dims, 1001 + 10 * k, device=dev)))
    op = q.ClovDslash(geom)
    out = {}
    for name, prec, sb, lays in (("f64", q.F64, 8, ((8, 8), (8, 4))), ("f16", q.F16, 2, ((16, 16), (16, 8)))):
        fs = [dict(g0=geom.pack_gauge(U, prec, 12, 0), i0=geom.pack_clover(inv, prec, 0),
                   p1=geom.pack_spinor(psi, prec, 1), t0=geom.new_spinor(prec, cb=0)) for U, inv, psi in lex]
        b = algorithmic_bytes(12, 72, False, sb) * geom.vcb

        def kn(i):
            f = fs[i & 1]
            op.dslash(f["t0"], f["p1"], f["g0"], f["i0"], 1, 0)

        for i in range(3):
            kn(i)
        t = time_kernel(kn, n)
        res = {"native": {"us": round(t * 1e6, 2), "hbm_frac": round(b / t / 1e9 / PEAK_HBM_GBS, 4)}}
        for V, S in lays:
            lay = q.qphix_layout(prec, V, S, True)
            for f in fs:
                f["pq"] = q.QPhiXSpinor(geom.spinor_to_qphix(f["p1"], lay), lay)
                f["rq"] = q.QPhiXSpinor(torch.empty_like(f["pq"].data), lay)

            def kq(i):
                f = fs[i & 1]
                op.dslash(f["rq"], f["pq"], f["g0"], f["i0"], 1, 0)

            for i in range(3):
                kq(i)
            t = time_kernel(kq, n)
            res[f"veclen{V}_soalen{S}"] = {"us": round(t * 1e6, 2), "hbm_frac": round(b / t / 1e9 / PEAK_HBM_GBS, 4)}
            for f in fs:
                del f["pq"], f["rq"]
        out[name] = res
        del fs
    del lex
    torch.cuda.empty_cache()
    return out


def _err(out, ref):
    d = (out - ref).abs().max().item()
    return {"max_abs": d, "max_abs_rel": d / ref.abs().max().item()}


def errors_vs_f64(q, geom, dev):
    """Per-site max-abs difference of each reduced-precision kernel from the
    FP64 kernel on the same inputs (the FP64 kernels equal the FP64 oracle
    to 1e-13 in the tests), absolute and relative to the largest output
    element: clov_dslash in FP32 and in the CG's sloppy FP16 / BF16 storage
    (12-real links), and the twisted-mass tm_dslash (mu 0.1, kappa 0.1350)
    in FP32."""
    import torch
    L = q.lattice
    U = L.random_gauge(DIMS, 0.3, 1000, device=dev)
    inv = L.invert_clover(L.random_clover(DIMS, 1003, device=dev))
    psi = L.random_spinor(DIMS, 1001, device=dev)
    alpha, mu = 1.0 / (2 * 0.1350), 0.1
    ops = {"clov_dslash": (q.ClovDslash(geom), True), "tm_dslash": (q.TMDslash(geom, alpha - 4.0, mu), False)}
    precs = {"clov_dslash": (("f32", q.F32), ("f16", q.F16), ("bf16", q.BF16)), "tm_dslash": (("f32", q.F32),)}
    out = {}
    for name, (op, clover) in ops.items():
        res = {}
        for pname, prec in (("f64", q.F64),) + precs[name]:
            t = geom.new_spinor(prec, cb=0)
            args = (t, geom.pack_spinor(psi, prec, 1), geom.pack_gauge(U, prec, 12, 0))
            if clover:
                args = args + (geom.pack_clover(inv, prec, 0),)
            op.dslash(*args, 1, 0)
            res[pname] = geom.unpack_spinor(t)
        out[name] = {p: _err(res[p], res["f64"]) for p, _ in precs[name]}
        del res
    del U, inv, psi
    torch.cuda.empty_cache()
    return out


def copy_bandwidth(dev, nbytes=1 << 30, reps=10):
    """Device-to-device copy of nbytes (read + write), GB/s: the box's
    practical ceiling for a stream that reads and writes."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a)
    t = time_kernel(lambda i: b.copy_(a), reps)
    del a, b
    torch.cuda.empty_cache()
    return round(2 * nbytes / t / 1e9, 1)


def triad_bandwidth(dev, nbytes=1 << 30, reps=10):
    """STREAM triad c = a + s b on the device (torch's vectorized elementwise
    kernel, three 1 GiB arrays, 2 reads : 1 write), GB/s."""
    import torch
    a = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    b = torch.empty_like(a).fill_(2.0)
    c = torch.empty_like(a)
    t = time_kernel(lambda i: torch.add(a, b, alpha=1.5, out=c), reps)
    del a, b, c
    torch.cuda.empty_cache()
    return round(3 * nbytes / t / 1e9, 1)


def sloppy_kernels(q, dev):
    """The CG's sloppy operator kernels (FP16 and BF16 storage, FP32
    arithmetic, 12-real links) at 32^3x64, HIP-event averages."""
    import torch
    L = q.lattice
    geom = q.Geometry(DIMS, t_boundary=-1.0)
    vcb = geom.vcb
    U = L.random_gauge(DIMS, 0.3, 1000, device=dev)
    clov = L.random_clover(DIMS, 1003, device=dev)
    inv = L.invert_clover(clov)
    psi = L.random_spinor(DIMS, 1001, device=dev)
    op = q.ClovDslash(geom)
    out = {}
    for name, prec in (("f16", q.F16), ("bf16", q.BF16)):
        sets = [dict(g0=geom.pack_gauge(U, prec, 12, 0), g1=geom.pack_gauge(U, prec, 12, 1),
                     inv0=geom.pack_clover(inv, prec, 0), clov1=geom.pack_clover(clov, prec, 1),
                     psi1=geom.pack_spinor(psi, prec, 1), t0=geom.new_spinor(prec, cb=0),
                     out1=geom.new_spinor(prec, cb=1)) for _ in range(2)]

        def kd(i):
            s = sets[i & 1]
            op.dslash(s["t0"], s["psi1"], s["g0"], s["inv0"], 1, 0)

        def ka(i):
            s = sets[i & 1]
            op.dslashAChiMinusBDPsi(s["out1"], s["t0"], s["psi1"], s["g1"], s["clov1"], 0.25, 1, 1)

        for k in (kd, ka):
            for i in range(5):
                k(i)
        td, ta = time_kernel(kd, 100), time_kernel(ka, 100)
        bd, ba = algorithmic_bytes(12, 72, False, 2) * vcb, algorithmic_bytes(12, 72, True, 2) * vcb
        out[name] = {"clov_dslash_us": round(td * 1e6, 2),
                     "clov_dslash_hbm_frac": round(bd / td / 1e9 / PEAK_HBM_GBS, 4),
                     "clov_dslash_achimbdpsi_us": round(ta * 1e6, 2),
                     "clov_dslash_achimbdpsi_hbm_frac": round(ba / ta / 1e9 / PEAK_HBM_GBS, 4)}
        del sets
    del U, clov, inv, psi
    torch.cuda.empty_cache()
    return out


CG_REPEATS = 3


def _solve(q, geom, op, b, mixed):
    """CG_REPEATS solves of the same system (the reductions are deterministic,
    so every solve gives the same iterate); returns the last solution, its
    result record and the median solve time."""
    times = []
    for _ in range(CG_REPEATS):
        x = geom.new_spinor(q.F64, cb=1)
        res = q.InvCG(op, max_iters=20000)(x, b, 1e-10, mixed=mixed)
        times.append(res["seconds"])
    return x, res, sorted(times)[len(times) // 2]


def cg_compare(q, dev):
    """The 32^3x64 solve: even-odd Wilson-clover M^dag M, kappa = 0.1350,
    csw = 1.0, links sigma = 0.2; FP64 CG vs mixed-precision reliable-update
    CG (FP16 and BF16 storage, FP32 arithmetic, FP64 true residuals) to a
    true relative residual of 1e-10 on the same Gaussian right-hand side;
    time to solution is the median of CG_REPEATS solves."""
    import torch
    L = q.lattice
    geom = q.Geometry(DIMS, t_boundary=-1.0)
    U = L.random_gauge(DIMS, 0.2, 3000, device=dev)
    b = L.random_spinor(DIMS, 3001, device=dev)
    A = q.clover_from_links(geom, U, 1.0 / (2 * 0.1350), 1.0)
    bf = geom.pack_spinor(b, q.F64, 1)
    out, sols = {}, {}
    for name, prec in (("f64", q.F64), ("f16", q.F16), ("bf16", q.BF16)):
        op = q.EvenOddCloverOperator(geom, U, A, prec_sloppy=prec, recon_sloppy=12)
        x, res, sec = _solve(q, geom, op, bf, prec != q.F64)
        sols[name] = x.data
        out[name] = {"iters": res["iters"], "seconds": round(sec, 4),
                     "reliable_updates": res["reliable_updates"], "rsd_final": res["rsd_final"],
                     "rsd_true_recomputed": true_residual(q, geom, op, x, bf)}
        op.close()
    for name in ("f16", "bf16"):
        out[name]["speedup_vs_f64"] = round(out["f64"]["seconds"] / out[name]["seconds"], 3)
        out[name]["rel_err_vs_f64_solution"] = float(
            ((sols[name] - sols["f64"]).norm() / sols["f64"].norm()).item())
    del sols
    torch.cuda.empty_cache()
    return out


def true_residual(q, geom, op, x, b):
    """|b - M^dag M x| / |b| recomputed in FP64 through the operator's own
    FP64 application (outside the CG's loop)."""
    mx, mdmx = geom.new_spinor(q.F64, cb=1), geom.new_spinor(q.F64, cb=1)
    op.apply(mx, x)
    op.apply(mdmx, mx, dagger=True)
    return float(((b.data - mdmx.data).norm() / b.data.norm()).item())


FORCE_REPEATS = 3


def _force(q, op, U, x, csw):
    """The whole pseudofermion force at x (hopping + clover terms, FP64 force
    kernels; the hopping term alone with csw None): one call that sizes the
    operator's workspaces, then the median wall time of FORCE_REPEATS calls
    (each synchronises on return)."""
    import torch
    F = op.force(torch.zeros_like(U), x, U, project=True, csw=csw)
    times = []
    for _ in range(FORCE_REPEATS):
        G = torch.zeros_like(U)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        op.force(G, x, U, project=True, csw=csw)
        times.append(time.perf_counter() - t0)
        del G
    return F, sorted(times)[len(times) // 2]


def shard(q, dev, n=100):
    """The 32^3x64 clover operator of the headline line split in t over N
    GPUs (N = 2, 4, 8), on one GPU's share: a local lattice 32^3 x (64/N)
    whose t wrap-around hops are dropped (qphix_amd_set_comm), FP32, 12-real
    links.  Times (HIP events around replays of a HIP graph of the calls,
    so that no host launch time sits between the small face kernels) of the
    interior clov_dslash (cb 0) and clov_dslash_achimbdpsi (cb 1), and for
    each of the two the face pack and the face finish of both t faces (one
    launch each, fb = 2); the bytes of one face message.  The exchange
    itself needs a second GPU and is not timed.  critical_path_us: both
    kernels' interiors and finishes in sequence, the packs and the exchange
    overlapped with the interiors; two_stream_step_us: the step measured
    with each kernel's packs on a second stream, joined before its finishes;
    one_stream_step_us: the six launches in order on one stream (no
    exchange in either);
    efficiency: the same two kernels on the whole 32^3x64 lattice, divided
    by N, over the critical path and over the one-stream step."""
    import torch
    L = q.lattice
    prec = q.F32
    beta = 0.25

    def fields(dims, seed, comm):
        geom = q.Geometry(dims, t_boundary=-1.0)
        if comm:
            geom.set_comm([0, 0, 0, 1])
        U = L.random_gauge(dims, 0.3, seed, device=dev)
        clov = L.random_clover(dims, seed + 1, device=dev)
        f = dict(g0=geom.pack_gauge(U, prec, 12, 0), g1=geom.pack_gauge(U, prec, 12, 1),
                 inv0=geom.pack_clover(L.invert_clover(clov), prec, 0), clov1=geom.pack_clover(clov, prec, 1),
                 psi1=geom.pack_spinor(L.random_spinor(dims, seed + 2, device=dev), prec, 1),
                 t0=geom.new_spinor(prec, cb=0), out1=geom.new_spinor(prec, cb=1))
        del U, clov
        return geom, f

    def timed(fn, reps=50):
        """us per fn(), from replays of a HIP graph of reps calls (no host
        launch time between the small face kernels)."""
        st = torch.cuda.Stream()
        st.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(st):
            fn(0)
        torch.cuda.current_stream().wait_stream(st)
        torch.cuda.synchronize()
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            for i in range(reps):
                fn(i)
        g.replay()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n // 10):
            g.replay()
        e1.record()
        e1.synchronize()
        return e0.elapsed_time(e1) * 1e3 / (reps * (n // 10))

    geom, f = fields(DIMS, 1500, False)
    op = q.ClovDslash(geom)
    whole = timed(lambda i: (op.dslash(f["t0"], f["psi1"], f["g0"], f["inv0"], 1, 0),
                             op.dslashAChiMinusBDPsi(f["out1"], f["t0"], f["psi1"], f["g1"], f["clov1"],
                                                     beta, 1, 1)))
    del geom, f, op
    torch.cuda.empty_cache()
    out = {"workload": "32^3x64 clover operator (clov_dslash + clov_dslash_achimbdpsi, FP32, 12-real links) "
                       "split in t, one GPU's share",
           "whole_lattice_step_us": round(whole, 2), "parts": {}}
    for N in (2, 4, 8):
        dims = (DIMS[0], DIMS[1], DIMS[2], DIMS[3] // N)
        geom, f = fields(dims, 1510 + N, True)
        op = q.ClovDslash(geom)
        # both t faces, packed and finished in one launch each (one buffer per kernel)
        buf, buf2 = geom.new_face(prec, 3, both=True), geom.new_face(prec, 3, both=True)
        t_d = timed(lambda i: op.dslash(f["t0"], f["psi1"], f["g0"], f["inv0"], 1, 0))
        t_a = timed(lambda i: op.dslashAChiMinusBDPsi(f["out1"], f["t0"], f["psi1"], f["g1"], f["clov1"],
                                                      beta, 1, 1))
        # the first kernel's faces: half-spinors of psi (cb 1), finished into t (cb 0) through
        # the clover epilogue; the second's: of t, finished into out (cb 1) with +coeff
        p_d = timed(lambda i: q.face_pack(geom, buf, f["psi1"], 1, 0, 3, 2))
        f_d = timed(lambda i: q.face_finish(geom, 1, f["t0"], buf, f["g0"], 1, 0, 3, 2, 1.0, clov=f["inv0"]))
        p_a = timed(lambda i: q.face_pack(geom, buf, f["t0"], 1, 1, 3, 2))
        f_a = timed(lambda i: q.face_finish(geom, 0, f["out1"], buf, f["g1"], 1, 1, 3, 2, beta))
        # the whole step with the design's overlap, on two streams: each kernel's
        # faces packed on a second stream while its interior runs, joined before
        # its finishes (the second kernel's packs read t after the first's
        # finishes); the exchange itself is absent
        side = torch.cuda.Stream()

        def overlapped(i):
            cur = torch.cuda.current_stream()
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                q.face_pack(geom, buf, f["psi1"], 1, 0, 3, 2)
            op.dslash(f["t0"], f["psi1"], f["g0"], f["inv0"], 1, 0)
            cur.wait_stream(side)
            q.face_finish(geom, 1, f["t0"], buf, f["g0"], 1, 0, 3, 2, 1.0, clov=f["inv0"])
            side.wait_stream(cur)
            with torch.cuda.stream(side):
                q.face_pack(geom, buf2, f["t0"], 1, 1, 3, 2)
            op.dslashAChiMinusBDPsi(f["out1"], f["t0"], f["psi1"], f["g1"], f["clov1"], beta, 1, 1)
            cur.wait_stream(side)
            q.face_finish(geom, 0, f["out1"], buf2, f["g1"], 1, 1, 3, 2, beta)

        def sequential(i):
            q.face_pack(geom, buf, f["psi1"], 1, 0, 3, 2)
            op.dslash(f["t0"], f["psi1"], f["g0"], f["inv0"], 1, 0)
            q.face_finish(geom, 1, f["t0"], buf, f["g0"], 1, 0, 3, 2, 1.0, clov=f["inv0"])
            q.face_pack(geom, buf2, f["t0"], 1, 1, 3, 2)
            op.dslashAChiMinusBDPsi(f["out1"], f["t0"], f["psi1"], f["g1"], f["clov1"], beta, 1, 1)
            q.face_finish(geom, 0, f["out1"], buf2, f["g1"], 1, 1, 3, 2, beta)

        t_ov = timed(overlapped)
        t_seq = timed(sequential)
        face_bytes = geom.face_sites(3) * 12 * BYTES_F32
        # the packs and the exchange overlapped with the interiors: each kernel's
        # interior, then its finishes
        crit = t_d + f_d + t_a + f_a
        out["parts"][str(N)] = {
            "local_lattice": list(dims), "sites_per_cb": geom.vcb,
            "interior_us": {"clov_dslash": round(t_d, 2), "clov_dslash_achimbdpsi": round(t_a, 2)},
            "face_pack_us_two_faces": {"clov_dslash": round(p_d, 2), "clov_dslash_achimbdpsi": round(p_a, 2)},
            "face_finish_us_two_faces": {"clov_dslash": round(f_d, 2), "clov_dslash_achimbdpsi": round(f_a, 2)},
            "face_message_bytes": face_bytes,
            "critical_path_us": round(crit, 2),
            "two_stream_step_us": round(t_ov, 2),
            "one_stream_step_us": round(t_seq, 2),
            "whole_lattice_step_over_n_us": round(whole / N, 2),
            "efficiency_if_exchange_hidden": round(whole / N / crit, 3),
            "efficiency_one_stream_measured": round(whole / N / t_seq, 3),
        }
        del geom, f, op, buf, buf2, side
        torch.cuda.empty_cache()
    return out


def cfg4(q, dev, cg=True):
    """The 48^3x96 (5,308,416 sites/cb) twisted-mass clover workload, mu = 0.05,
    kappa = 0.1355, csw = 1.5 (clover from the links), links sigma = 0.2.
      * tm_clov_dslash and its achimbdpsi form in half storage (FP16 spinors,
        12-real FP16 links, FP16 full clover blocks, FP32 arithmetic) and in
        FP32: HIP-event time, GFLOP/s, fraction of 8 TB/s by algorithmic
        bytes, and the per-site error against the FP64 kernel;
      * FP64 CG vs mixed-precision CG (FP16 storage) to a true residual of
        1e-10 on M^dag M of the even-odd TM-clover operator (median of
        CG_REPEATS solves), each solution's true residual recomputed in FP64
        outside the solver, and the two solutions' difference;
      * the pseudofermion force at each solution (hopping and clover terms,
        and the hopping term alone; median of FORCE_REPEATS calls), and the
        force from the mixed solution against the force from the FP64
        solution."""
    import torch
    L = q.lattice
    dims = (48, 48, 48, 96)
    kappa, mu, csw = 0.1355, 0.05, 1.5
    alpha = 1.0 / (2 * kappa)
    geom = q.Geometry(dims, t_boundary=-1.0)
    vcb = geom.vcb
    U = L.random_gauge(dims, 0.2, 4000, device=dev)
    A = q.clover_from_links(geom, U, alpha, csw)
    phi = L.random_spinor(dims, 4001, device=dev)
    out = {"workload": "48^3x96 twisted-mass clover, mu=0.05, kappa=0.1355, csw=1.5 (clover from links), "
                       "sigma=0.2", "sites_per_cb": vcb}
    # --- dslash throughput and error against FP64
    flop_d, flop_a = 1872, 1920
    op_k = q.TMClovDslash(geom)
    B = q.tmclover_blocks(geom, A, mu, 1, True)
    F = q.tmclover_blocks(geom, A, mu, 1, False)
    res = {}
    for name, prec in (("f64", q.F64), ("f16", q.F16), ("f32", q.F32)):
        f = dict(g0=geom.pack_gauge(U, prec, 12, 0), g1=geom.pack_gauge(U, prec, 12, 1),
                 B0=geom.pack_full_clover(B, prec, 0), F1=geom.pack_full_clover(F, prec, 1),
                 psi=geom.pack_spinor(phi, prec, 1), t=geom.new_spinor(prec, cb=0),
                 o=geom.new_spinor(prec, cb=1))
        op_k.dslash(f["t"], f["psi"], f["g0"], f["B0"], 1, 0)
        res[name] = geom.unpack_spinor(f["t"])
        if name != "f64":
            def kd(i):
                op_k.dslash(f["t"], f["psi"], f["g0"], f["B0"], 1, 0)

            def ka(i):
                op_k.dslashAChiMinusBDPsi(f["o"], f["t"], f["psi"], f["g1"], f["F1"], 0.25, 1, 1)

            for k in (kd, ka):
                for i in range(3):
                    k(i)
            td, ta = time_kernel(kd, 50), time_kernel(ka, 50)
            sb = 2 if prec == q.F16 else 4
            bd = (24 + 96 + 144 + 24) * sb * vcb
            ba = (24 + 96 + 144 + 24 + 24) * sb * vcb
            out[f"dslash_{name}"] = {
                "tm_clov_dslash_us": round(td * 1e6, 1),
                "tm_clov_dslash_gflops": round(vcb * flop_d / td / 1e9, 1),
                "tm_clov_dslash_hbm_frac": round(bd / td / 1e9 / PEAK_HBM_GBS, 4),
                "tm_clov_achimbdpsi_us": round(ta * 1e6, 1),
                "tm_clov_achimbdpsi_gflops": round(vcb * flop_a / ta / 1e9, 1),
                "tm_clov_achimbdpsi_hbm_frac": round(ba / ta / 1e9 / PEAK_HBM_GBS, 4),
                "tm_clov_dslash_error_vs_f64": _err(res[name], res["f64"])}
        del f
        torch.cuda.empty_cache()
    del B, F, res
    torch.cuda.empty_cache()
    if not cg:
        return out
    # --- CG: FP64 vs mixed (FP16 storage), then the force at each solution
    b = geom.pack_spinor(phi, q.F64, 1)
    forces, sols = {}, {}
    for name, prec in (("f64", q.F64), ("f16", q.F16)):
        op = q.EvenOddCloverOperator(geom, U, A, mu=mu, prec_sloppy=prec, recon_sloppy=12)
        x, r, sec = _solve(q, geom, op, b, prec != q.F64)
        out[f"cg_{name}"] = {"iters": r["iters"], "seconds": round(sec, 4),
                             "reliable_updates": r["reliable_updates"], "rsd_final": r["rsd_final"],
                             "rsd_true_recomputed": true_residual(q, geom, op, x, b)}
        forces[name], ft = _force(q, op, U, x, csw)
        out[f"cg_{name}"]["force_seconds"] = round(ft, 4)
        out[f"cg_{name}"]["hopping_force_seconds"] = round(_force(q, op, U, x, None)[1], 4)
        sols[name] = x.data
        op.close()
        del x
        torch.cuda.empty_cache()
    out["cg_f16"]["speedup_vs_f64"] = round(out["cg_f64"]["seconds"] / out["cg_f16"]["seconds"], 3)
    out["cg_f16"]["rel_err_vs_f64_solution"] = float(
        ((sols["f16"] - sols["f64"]).norm() / sols["f64"].norm()).item())
    out["force_rel_err_mixed_vs_f64"] = float(
        ((forces["f16"] - forces["f64"]).abs().max() / forces["f64"].abs().max()).item())
    del forces, sols, U, A, phi, b
    torch.cuda.empty_cache()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-cg", action="store_true", help="skip the CG solves (32^3x64 and 48^3x96)")
    ap.add_argument("--no-cfg4", action="store_true", help="skip the 48^3x96 twisted-mass clover block")
    ap.add_argument("--no-shard", action="store_true", help="skip the split-lattice (one GPU's share) block")
    ap.add_argument("--workload", default="clover32", choices=["clover32", "tmclover48", "shard"],
                    help="clover32: the 32^3x64 clover operator (the headline line, with the "
                         "48^3x96 block); tmclover48: the 48^3x96 twisted-mass clover block alone; "
                         "shard: one GPU's share of the 32^3x64 operator split in t over 2, 4, 8 GPUs")
    ap.add_argument("--kernels-only", action="store_true",
                    help="launch clov_dslash, clov_dslash_achimbdpsi, Wilson dslash and clov_dslash "
                         "on QPhiX arrays in place (soalen 16, 8, 4) 10 times each plus a 1 GiB device copy (for "
                         "rocprofv3 --pmc runs and their copy-kernel calibration), print nothing")
    ap.add_argument("--lattice", default=None,
                    help="with --kernels-only: Lx,Ly,Lz,Lt (default 32,32,32,64)")
    ap.add_argument("--traversal", type=int, default=-1,
                    help="with --kernels-only: work-group traversal order (-1 automatic)")
    ap.add_argument("--cpu-baseline-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-seconds", type=float, default=6.0, help=argparse.SUPPRESS)
    args = ap.parse_args()
    if args.cpu_baseline_child:  # the CPU-baseline leg's own process (no torch)
        dims = tuple(int(v) for v in args.lattice.split(",")) if args.lattice else DIMS
        print(json.dumps(_cpu_baseline_child(args.cpu_seconds, dims)), flush=True)
        return
    if (args.lattice or args.traversal != -1) and not args.kernels_only:
        print("--lattice and --traversal apply to --kernels-only runs", file=sys.stderr)
        sys.exit(2)
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__ as ge
    q = ge.load_package()
    L = q.lattice
    dev = "cuda:0"
    torch.cuda.set_device(0)
    if args.workload == "tmclover48":
        print(json.dumps(cfg4(q, dev, cg=not args.no_cg)), flush=True)
        return
    if args.workload == "shard":
        print(json.dumps(shard(q, dev)), flush=True)
        return
    prec, recon = q.F32, 12
    dims = tuple(int(v) for v in args.lattice.split(",")) if args.lattice else DIMS
    geom = q.Geometry(dims, t_boundary=-1.0)
    geom.set_traversal(args.traversal)
    vcb = geom.vcb

    sets = []
    for k in range(2):
        U = L.random_gauge(dims, 0.3, 1000 + 10 * k, device=dev)
        clov = L.random_clover(dims, 1003 + 10 * k, device=dev)
        inv = L.invert_clover(clov)
        psi = L.random_spinor(dims, 1001 + 10 * k, device=dev)
        s = dict(g0=geom.pack_gauge(U, prec, recon, 0), g1=geom.pack_gauge(U, prec, recon, 1),
                 inv0=geom.pack_clover(inv, prec, 0), clov1=geom.pack_clover(clov, prec, 1),
                 psi1=geom.pack_spinor(psi, prec, 1), t0=geom.new_spinor(prec, cb=0),
                 out1=geom.new_spinor(prec, cb=1))
        del U, clov, inv, psi
        sets.append(s)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    op = q.ClovDslash(geom)
    beta = 0.25

    def step(i):
        s = sets[i & 1]
        op.dslash(s["t0"], s["psi1"], s["g0"], s["inv0"], 1, 0)
        op.dslashAChiMinusBDPsi(s["out1"], s["t0"], s["psi1"], s["g1"], s["clov1"], beta, 1, 1)

    if args.kernels_only:
        for i in range(10):
            op.dslash(sets[i & 1]["t0"], sets[i & 1]["psi1"], sets[i & 1]["g0"], sets[i & 1]["inv0"], 1, 0)
        for i in range(10):
            s = sets[i & 1]
            op.dslashAChiMinusBDPsi(s["out1"], s["t0"], s["psi1"], s["g1"], s["clov1"], 0.25, 1, 1)
        wil = q.Dslash(geom)
        for i in range(10):
            s = sets[i & 1]
            wil.dslash(s["t0"], s["psi1"], s["g0"], 1, 0)
        # clov_dslash on QPhiX arrays in place (veclen 16, soalen 16, 8, 4:
        # 10 launches each, in that order)
        for S in (16, 8, 4):
            lay = q.qphix_layout(q.F32, 16, S, True)
            psi_q = q.QPhiXSpinor(geom.spinor_to_qphix(sets[0]["psi1"], lay), lay)
            res_q = q.QPhiXSpinor(torch.empty_like(psi_q.data), lay)
            for i in range(10):
                op.dslash(res_q, psi_q, sets[0]["g0"], sets[0]["inv0"], 1, 0)
            del psi_q, res_q
        del sets
        a = torch.empty(256 * 1024 * 1024, dtype=torch.float32, device=dev).fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        return

    for i in range(args.warmup):
        step(i)
    torch.cuda.synchronize()
    t_step = time_kernel(step, args.steps)

    # the dominant kernel alone (clover dslash), same stream, HIP events
    def k_dslash(i):
        s = sets[i & 1]
        op.dslash(s["t0"], s["psi1"], s["g0"], s["inv0"], 1, 0)

    def k_achi(i):
        s = sets[i & 1]
        op.dslashAChiMinusBDPsi(s["out1"], s["t0"], s["psi1"], s["g1"], s["clov1"], beta, 1, 1)

    t_d = time_kernel(k_dslash, args.steps)
    t_a = time_kernel(k_achi, args.steps)
    tm = tm_kernel(q, geom, sets[0], dev, args.steps)
    inplace = inplace_kernels(q, geom, sets, args.steps)
    inplace_lp = inplace_other_precisions(q, geom, dev, args.steps)
    err = errors_vs_f64(q, geom, dev)
    # naive count (every neighbour spinor read from memory), as often quoted
    b_naive = (8 * 24 + 8 * recon + 72 + 24) * BYTES_F32 * vcb
    b_d = algorithmic_bytes(recon, 72, False, BYTES_F32) * vcb
    b_a = algorithmic_bytes(recon, 72, True, BYTES_F32) * vcb
    flops = vcb * (FLOP_CLOVER + FLOP_CLOVER_ACHI)
    gflops = flops / t_step / 1e9
    achieved = b_d / t_d / 1e9

    line = {
        "metric": METRIC,
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(t_step * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "even-odd Wilson-clover operator: clov_dslash + clov_dslash_achimbdpsi, "
                        "FP32, 12-real links, 32^3x64 (1,048,576 sites/cb)",
            "lattice": list(DIMS), "sites_per_cb": vcb, "recon": recon,
            "flop_per_site": {"clov_dslash": FLOP_CLOVER, "clov_dslash_achimbdpsi": FLOP_CLOVER_ACHI},
            "kernel_us": {"clov_dslash": round(t_d * 1e6, 2),
                          "clov_dslash_achimbdpsi": round(t_a * 1e6, 2)},
            "kernel_gflops": {"clov_dslash": round(vcb * FLOP_CLOVER / t_d / 1e9, 1),
                              "clov_dslash_achimbdpsi": round(vcb * FLOP_CLOVER_ACHI / t_a / 1e9, 1)},
            "kernel_hbm_frac": {"clov_dslash": round(b_d / t_d / 1e9 / PEAK_HBM_GBS, 4),
                                "clov_dslash_achimbdpsi": round(b_a / t_a / 1e9 / PEAK_HBM_GBS, 4)},
            "parallelism": "single GPU",
            "tm_dslash": tm,
            "qphix_inplace_clov_dslash": inplace,
            "qphix_inplace_clov_dslash_f64_f16": inplace_lp,
            "errors_vs_f64": err,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": profile_traffic(),
                     "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, "
                                       "profiles/clov_dslash_f32_32x64_pmc.txt (not measured in this run)",
                     "kernel": "clov_dslash f32 recon12",
                     "algorithmic_bytes_per_launch": b_d,
                     "naive_bytes_per_launch": b_naive,
                     "naive_gbs": round(b_naive / t_d / 1e9, 1)},
        "cpu_baseline": None,
        "cg": None,
    }
    del sets
    torch.cuda.empty_cache()
    line["roofline"]["measured_copy_gbs"] = copy_bandwidth(dev)
    line["roofline"]["measured_triad_gbs"] = triad_bandwidth(dev)
    if not args.no_cg:
        line["cg"] = cg_compare(q, dev)
        line["cg"]["sloppy_kernels"] = sloppy_kernels(q, dev)
    if not args.no_cfg4:
        line["cfg4"] = cfg4(q, dev, cg=not args.no_cg)
    if not args.no_shard:
        line["shard"] = shard(q, dev)
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
