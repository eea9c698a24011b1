"""Benchmark of the chimeraCL PIC step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config C2]
                    [--no-cpu] [--no-mixed] [--no-diff] [--align-every K]
                    [--compare-align K]

Workload (default): the medium laser-wakefield configuration, Nz=2048,
Nr=256, modes m=0,1, 2x2x8 particles per cell (16.6M electrons plus as many
immobile ions), dt=dz, seeded synthetic plasma and an a0=2 Gaussian laser.
A step is one full PIC step (push, sort, J and rho deposit, transforms,
grad, PSATD + damping, B restore, backward transforms, gather + push); the
timed steps replay a HIP graph of the step with every input resident in HBM.

The same workload runs twice, in the plain FP64 form (value) and in the
mixed form (FP32 particles and FFT, split-BF16 3-term DHT products); the
two runs advance in lockstep so that the mixed form's E, B, J and momentum
errors and both charge-conservation residuals are measured against the FP64
run at steps 100 and 500.  The FP64 run's error against the CPU FP64 oracle
is measured after the first 32 steps, with the oracle's throughput on 30 of
them (a bounded sample, ~15 s) timed on this host's cores.  The differentiable form is
measured on the two differentiable workloads (Nz=1024, Nr=128, 100 steps,
FP64, gradients against central finite differences; Nz=4096, Nr=256, 400
steps, mixed, gradients against the FP64 adjoint), with the backward /
forward time ratio.  The small (C1, FP64 vs the CPU oracle after 5 steps;
mixed vs FP64 at steps 100 and 200) and large (C3, M=2 with a density ramp;
mixed vs FP64 at steps 100 and 1000) configurations report the same errors
and charge residuals ("configs").

roofline: the longer of the two particle kernels of a step (the gather +
push and the electrons' J deposit, both under "kernels"), its algorithmic
bytes over its mean duration in the timed steps.  HIP graphs here cannot hold
event nodes, so the state before the timed steps is saved, the timed steps
run (graph replay, wall clock), the state is restored and the same steps are
rerun eagerly with HIP events around both kernels in each step; the rerun
must end bit-identical to the timed run.  traffic: HBM bytes per launch from
the rocprofv3 counter summary in profiles/.

Particle storage: every 20 steps the electrons are permuted into their
sorted order (PIC_loop(align_every=20), the reference's align gather over all
particles), so that their memory order keeps following their cells, as the
reference's own moving-window runs realign at every frame step.  Each
electron keeps its unreordered index for the gather's Np_stay skip, so the
run is the reference's step with the same particles skipped; only the
summation order within a cell follows the storage (the reference's in-cell
order is nondeterministic).  The "align_compare" section times both forms
again without any reordering (align_every=0, the reference loop's storage
order) over the same steps.

Printed: one JSON line.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFS = {"f64": 78.6, "f32": 157.3, "bf16": 2500.0}


DEFAULTS = {  # per workload: diagnostics steps, CPU-oracle sample (untimed, timed steps)
    "C1": {"diag": "100,200", "cpu": (2, 30)},
    "C2": {"diag": "100,500", "cpu": (2, 20)},
    "C3": {"diag": "100,1000", "cpu": (1, 3)},
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="C3",
                    help="headline workload (default C3, the largest one-GPU configuration)")
    ap.add_argument("--secondary", default="C2",
                    help="second workload timed the same way ('' to skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=None,
                    help="timed CPU-oracle steps (default per workload: C3 3, C2 20, C1 30)")
    ap.add_argument("--cpu-warmup", type=int, default=None)
    ap.add_argument("--kernel-reps", type=int, default=20)
    ap.add_argument("--diag-steps", default=None,
                    help="steps with the mixed-vs-FP64 report (default per workload: C3 "
                         "100,1000, C2 100,500)")
    ap.add_argument("--align-every", type=int, default=20,
                    help="permute the electrons into their sorted order every K steps (0: never)")
    ap.add_argument("--compare-align", type=int, default=0,
                    help="also time both forms with align_every=K (-1: skip)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy three times (HBM counter calibration)")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the error report of the small configuration (C1)")
    ap.add_argument("--no-window", action="store_true", help="skip the moving-window section")
    ap.add_argument("--gather-only", type=int, default=0, metavar="STEPS",
                    help="only run STEPS eager steps of both forms (no graphs), then three "
                         "gather + push launches of each (for counter collection)")
    return ap.parse_args()


def calibrate_copy():
    """A plain device-to-device copy of 1 GiB, three times: reads 1 GiB and
    writes 1 GiB per launch, the reference for FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"calibrate_copy_bytes": a.numel() * 4, "launches": 3}), flush=True)


def gather_only(args):
    """Both forms of the workload, `args.gather_only` eager PIC steps each
    (no HIP graph), then three gather + push launches per form (with the
    next first half push, as in the timed steps)."""
    import torch
    out = {}
    for prec in ("f64", "mixed"):
        sm = Sim(args.config, prec, False, args.align_every)
        sm.loop.graph = False
        sm.step(args.gather_only)
        for _ in range(3):  # as in the timed steps, with the next first half push
            sm.solver._gather_and_push(sm.eons, ["E", "B"], prepush=True)
        torch.cuda.synchronize()
        out[prec] = {"steps": args.gather_only, "gather_launches": 3, "electrons": sm.eons.Args["Np"]}
        del sm
        gc.collect()
        torch.cuda.empty_cache()
    print(json.dumps(out), flush=True)


def host_info():
    """CPU model, usable cores and the thread settings of the oracle."""
    info = {"cores_available": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                info["cpu_model"] = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        cfg = np.show_config(mode="dicts")
        blas = cfg.get("Build Dependencies", {}).get("blas", {})
        info["blas"] = f"{blas.get('name')} {blas.get('version')}"
    except Exception:  # numpy without the dict mode
        info["blas"] = None
    return info


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def load_arrays(cfg_name):
    """The seeded plasma of a workload (host float64 arrays, cell order)."""
    from chimeracl_amd import config as C
    c = C.CONFIGS[cfg_name]
    t = time.time()
    arr = C.load_plasma(C.grid_config(cfg_name), c["ppc"], c["dens"], seed=0, order="cell")
    log(f"[bench] {cfg_name}: loaded {arr['x'].size} particles in {time.time() - t:.1f}s")
    return arr


class Sim:
    def __init__(self, cfg_name, precision, with_oracle, align_every=0, arr=None, graph=True):
        import chimeracl_amd as CA
        from chimeracl_amd import config as C
        from chimeracl_amd.pic_loop import counter_collection_active
        from chimeracl_amd.methods.generic_methods_hip import Communicator
        self.comm = Communicator(precision=precision)
        g = self.g = C.grid_config(cfg_name)
        laser = C.laser_config(cfg_name)
        self.solver = CA.Solver(dict(g), self.comm)
        CA.add_gausian_pulse(self.solver, dict(laser))
        eo, io = C.species_config(cfg_name, self.solver.Args)
        if arr is None:
            arr = load_arrays(cfg_name)
        self.eons, self.ions = CA.Particles(dict(eo), self.comm), CA.Particles(dict(io), self.comm)
        self.eons.add_particles(source=C.ParticleSource(arr))
        self.ions.add_particles(source=C.ParticleSource(arr))
        prof = C.density_profile(cfg_name, g)
        if prof is not None:
            self.eons.dens_profile(*prof)
            self.ions.dens_profile(*prof)
        self.loop = CA.PIC_loop(solvers=[self.solver], species=[self.eons, self.ions],
                                graph=graph and not counter_collection_active(),
                                align_every=align_every)
        self.steps = 0
        self.orc = None
        if with_oracle:
            from oracle import oracle as O
            s = O.OracleSolver(dict(g))
            O.add_gausian_pulse(s, dict(laser))
            e, i = O.OracleParticles(dict(eo)), O.OracleParticles(dict(io))
            e.add_particles(arr)
            i.add_particles(arr)
            if prof is not None:
                e.dens_profile(*prof)
                i.dens_profile(*prof)
            self.orc = (s, e, i, O.OracleLoop(s, [e, i], align_every=align_every))

    def step(self, n=1):
        if n > 0:
            self.loop.run(n)
        self.steps += n


def oracle_errors(sim, M):
    s, e = sim.orc[0], sim.orc[1]
    out = {}
    for f in ("E", "B", "J"):
        num = den = 0.0
        for c in "xyz":
            for m in range(M + 1):
                a = sim.solver.DataDev[f"{f}{c}_m{m}"].detach().cpu().numpy()[1:]
                b = s.f[f"{f}{c}_m{m}"][1:]
                num += float(np.linalg.norm((a - b).ravel()) ** 2)
                den += float(np.linalg.norm(b.ravel()) ** 2)
        out[f] = (num / den) ** 0.5 if den > 0 else num ** 0.5
    num = den = 0.0
    for k in ("px", "py", "pz"):
        a = by_pid(sim.eons, k).double().cpu().numpy()
        b = e.data[k] if "pid" not in e.data else np.empty_like(e.data[k])
        if "pid" in e.data:
            b[e.data["pid"].astype(np.int64)] = e.data[k]
        num += float(np.linalg.norm(a - b) ** 2)
        den += float(np.linalg.norm(b) ** 2)
    out["p"] = (num / den) ** 0.5
    return out


# ---------------------------------------------------------------- diagnostics (FP64)
def _rel(pairs):
    num = den = 0.0
    for a, b in pairs:
        d = (a.to(b.dtype) - b)
        num += float((d.abs() ** 2).sum())
        den += float((b.abs() ** 2).sum())
    return (num / den) ** 0.5 if den > 0 else num ** 0.5


def mixed_errors(mx, fp, M):
    """Relative L2 of the mixed run against the FP64 run: real-space E, B, J
    (rows 1..Nr-1, every component and mode) and the electron momenta."""
    import torch
    out = {}
    for f in ("E", "B", "J"):
        pairs = []
        for c in "xyz":
            for m in range(M + 1):
                a = mx.solver.DataDev[f"{f}{c}_m{m}"][1:]
                b = fp.solver.DataDev[f"{f}{c}_m{m}"][1:]
                pairs.append((a.to(torch.complex128 if b.is_complex() else torch.float64), b))
        out[f] = _rel(pairs)
    out["p"] = _rel([(by_pid(mx.eons, k).double(), by_pid(fp.eons, k)) for k in ("px", "py", "pz")])
    return out


def by_pid(parts, k):
    """Particle component k in the numbering without reordering (the two
    forms reorder their storage independently)."""
    import torch
    a = parts.DataDev[k]
    pid = parts.DataDev.get("pid")
    if pid is None:
        return a
    out = torch.empty_like(a)
    out[pid.long()] = a
    return out


def add_div_scratch(solver):
    import torch
    a, D = solver.Args, solver.DataDev
    for n in ["divJ"] + [f"Jd{c}" for c in "xyz"]:
        for m in range(a["M"] + 1):
            D[f"{n}_fb_m{m}"] = torch.zeros((a["Nr"] - 1, a["Nx"]), dtype=torch.complex128,
                                            device=solver.comm.device)


def charge_residual(calc, sim, rho_prev):
    """|| (rho^(n+1) - rho^n)/dt + div J^(n+1/2) || / || div J || over the
    spectral planes (smoothed rho and J of the last two steps), in FP64 with
    the FP64 solver's operators (field_div, transformer_methods_cl.py:135-183)."""
    import torch
    M, D, Dc = sim.solver.Args["M"], sim.solver.DataDev, calc.DataDev
    for c in "xyz":
        for m in range(M + 1):
            Dc[f"Jd{c}_fb_m{m}"].copy_(D[f"J{c}_fb_m{m}"].to(torch.complex128))
    calc.field_div("Jd", "divJ")
    dt = sim.solver.Args["dt"]
    num = den = 0.0
    for m in range(M + 1):
        r = (D[f"rho_fb_m{m}"].to(torch.complex128) - rho_prev[m]) / dt + Dc[f"divJ_fb_m{m}"]
        num += float((r.abs() ** 2).sum())
        den += float((Dc[f"divJ_fb_m{m}"].abs() ** 2).sum())
    return (num / den) ** 0.5


def energies(sim):
    """Field energy sum 2 pi r dr dx (F_0^2 + 2 sum_m |F_m|^2) over E and B
    (rows 1..Nr-1) and electron kinetic energy sum w (gamma - 1), FP64."""
    import torch
    s, a = sim.solver, sim.solver.Args
    M = a["M"]
    r = torch.as_tensor(a["Rgrid"][1:], dtype=torch.float64, device=s.comm.device)[:, None]
    dv = 2 * np.pi * a["dr"] * a["dx"]
    wf = 0.0
    for f in ("E", "B"):
        for c in "xyz":
            wf += float((r * s.DataDev[f"{f}{c}_m0"][1:].double() ** 2).sum()) * dv
            for m in range(1, M + 1):
                v = s.DataDev[f"{f}{c}_m{m}"][1:]
                wf += 2 * float((r * (v.real.double() ** 2 + v.imag.double() ** 2)).sum()) * dv
    d = sim.eons.DataDev
    u2 = d["px"].double() ** 2 + d["py"].double() ** 2 + d["pz"].double() ** 2
    wk = float((d["w"].double() * (torch.sqrt(1 + u2) - 1)).sum())
    return {"field": wf, "kinetic": wk}


def rho_copy(sim):
    import torch
    M = sim.solver.Args["M"]
    return [sim.solver.DataDev[f"rho_fb_m{m}"].to(torch.complex128).clone() for m in range(M + 1)]


# ---------------------------------------------------------------- kernels
def kernel_table(sim, reps, w):
    """HIP-event timings (burst_time, on the stream the kernels run on) of the
    step's phases, with their algorithmic HBM bytes (or DHT flops) per call."""
    import torch
    s, e = sim.solver, sim.eons
    a = s.Args
    Np, M, Nx, Nr = e.Args["Np"], a["M"], a["Nx"], a["Nr"]
    Nk = Nr - 1
    r1 = Nr * Nx * w * (1 + 2 * M)          # one real-space component, all modes
    s1 = 2 * w * Nk * Nx * (M + 1)          # one spectral component, all modes

    def timeit(fn):
        return burst_time(fn, reps)

    rows = {}

    def add(name, t, nbytes=None, flops=None, peak=None):
        row = {"avg_us": t * 1e6}
        if nbytes is not None:
            row["bytes"] = nbytes
            row["GB/s"] = nbytes / t / 1e9
            row["hbm_frac"] = row["GB/s"] / HBM_PEAK_GBS
        if flops is not None:
            row["flops"] = flops
            row["TFLOP/s"] = flops / t / 1e12
            if peak:
                row["mfma_frac"] = row["TFLOP/s"] / peak
        rows[name] = row

    add("gather+push (a)", timeit(lambda: s._gather_and_push(e, ["E", "B"])), Np * 10 * w + 6 * r1)
    add("push_x+cell keys (a,b)", timeit(lambda: e.push_coords(mode="half", grid=s)), Np * (10 * w + 4))
    add("cell keys + radix sort (b)", timeit(lambda: e.index_sort(s)), Np * (3 * w + 12))
    add("J deposit (b)", timeit(lambda: s.depose_currents(species=[e])), Np * (8 * w + 4) + 3 * r1)
    add("rho deposit, electrons (b)", timeit(lambda: s.depose_scalar(e, "w", "rho", -1)), Np * (4 * w + 4) + r1)
    dflops = lambda nf: nf * (1 + 2 * M) * 2.0 * Nk * Nk * Nx
    mixed = s.comm.mixed
    peak = MFMA_PEAK_TFS["f32"] if mixed else MFMA_PEAK_TFS["f64"]
    add("forward transform rho,J + smoothing (c)",
        timeit(lambda: s.fb_transform_smooth(scals=["rho"], vects=["J"])), 4 * r1 + 4 * s1, dflops(4), peak)
    add("backward transform E,B (c)", timeit(lambda: s.fb_transform(vects=["E", "B"], dir=1)),
        6 * s1 + 6 * r1, dflops(6), peak)
    add("rot/w^2 (c)", timeit(lambda: s.restore_B_fb()), 6 * s1, 2 * (4 * M + 2) * 2.0 * Nk * Nk * Nx, peak)
    add("PSATD + damping (d)", timeit(lambda: s.damp_edges(("E", "G"), advance_with=("J", "dN0", "dN1"))),
        21 * s1 + 1.5 * s1)
    return rows


def pmc_traffic(config, form, kernel="gather"):
    """(HBM bytes per launch, source file) of one of the roofline kernels
    (gather, deposit_J, dht) on this workload: read from the rocprofv3
    counter summaries committed in profiles/ (files *_<config>_pmc.md, lines
    `traffic_bytes_<kernel>: {form: bytes}`; DRAM request counters for the
    4096 x 512 workload, copy-calibrated FETCH_SIZE / WRITE_SIZE for the
    2048 x 256 one), not measured in this run; (None, None) if absent."""
    import glob
    keys = [f"traffic_bytes_{kernel}:"] + (["traffic_bytes:"] if kernel == "gather" else [])
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", f"*_{config.lower()}_pmc.md"))):
        try:
            for line in open(path):
                for key in keys:
                    if line.startswith(key):
                        v = json.loads(line.split(":", 1)[1]).get(form)
                        if v is not None:
                            return v, os.path.relpath(path, ROOT)
        except OSError:
            continue
    return None, None


def burst_time(fn, reps):
    """Mean GPU time of fn: reps calls captured back to back in a HIP graph,
    replayed between two HIP events on the stream it runs on (no host gaps
    between the launches)."""
    import torch
    fn()  # buffers exist before the capture
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    cap = torch.cuda.Stream()
    cap.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(cap):
        with torch.cuda.graph(g, stream=cap):
            for _ in range(reps):
                fn()
    torch.cuda.current_stream().wait_stream(cap)
    st = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    g.replay()  # first replay uploads the graph
    a.record(st)
    g.replay()
    b.record(st)
    torch.cuda.synchronize()
    del g
    return a.elapsed_time(b) * 1e-3 / reps


def snapshot(sim):
    """(tensor, copy) pairs of everything the step reads and writes: the
    solver's and the species' device data."""
    import torch
    out, seen = [], set()
    for owner in (sim.solver, sim.eons, sim.ions):
        for v in owner.DataDev.values():
            if isinstance(v, torch.Tensor) and v.is_cuda and id(v) not in seen:
                seen.add(id(v))
                out.append((v, v.clone()))
    return out


def restore(snap):
    for live, saved in snap:
        live.copy_(saved)


def fingerprint(sim):
    """Copies of the fields and electron data that a step leaves behind."""
    d, e = sim.solver.DataDev, sim.eons.DataDev
    names = [k for k in d if k[0] in "EBJ" and "_m" in k and "_fb" not in k]
    return [d[k].clone() for k in sorted(names)] + \
        [e[k].clone() for k in ("x", "y", "z", "px", "py", "pz", "g_inv", "sort_indx")]


def replay_timed(sim, steps):
    """Runs `steps` PIC steps eagerly (PIC_loop.run without graphs: the
    kernels the graphs replay, in the same order) with a pair of HIP events
    around the electrons' gather + push and around their J deposit in every
    step, on the stream they are launched on, and around every DHT product
    launch (recorded by the library itself on the launch's stream,
    chm_dht_timing); returns the mean, min and max of each duration (ms).
    The host enqueues ahead of the GPU, so each pair brackets the kernel
    alone."""
    import torch
    from chimeracl_amd import _lib
    s, e = sim.solver, sim.eons
    kinds = {"_gather_and_push": "gather", "depose_vector_push": "deposit_J",
             "depose_vector": "deposit_J"}
    ev = {"gather": [], "deposit_J": []}
    flag = {"gather": [], "deposit_J": []}

    def wrap(name):
        orig = getattr(s, name)
        kind = kinds[name]

        def timed(parts, *a, **kw):
            if parts is not e:
                return orig(parts, *a, **kw)
            a0, b0 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a0.record()
            r = orig(parts, *a, **kw)
            b0.record()
            ev[kind].append((a0, b0))
            flag[kind].append(bool(kw.get("prepush")) if kind == "gather"
                              else name == "depose_vector_push")
            return r
        return timed

    for name in kinds:
        setattr(s, name, wrap(name))
    graph, sim.loop.graph = sim.loop.graph, False
    _lib.lib.chm_dht_timing(1)
    try:
        sim.loop.run(steps)  # eager, the same step sequence as the timed run
        torch.cuda.synchronize()
        nmax = 64 * steps
        ms, fl = np.zeros(nmax), np.zeros(nmax)
        n = int(_lib.lib.chm_dht_timing_collect(ms.ctypes.data, fl.ctypes.data, None, nmax))
    finally:
        _lib.lib.chm_dht_timing(0)
        for name in kinds:
            delattr(s, name)
        sim.loop.graph = graph
    out = {}
    for kind, pairs in ev.items():
        t = [a.elapsed_time(b) for a, b in pairs]
        out[kind] = {"mean_ms": float(np.mean(t)), "min_ms": float(np.min(t)),
                     "max_ms": float(np.max(t)), "steps": len(t),
                     ("prepush_steps" if kind == "gather" else "push_steps"): int(sum(flag[kind]))}
    if n > 0:
        out["dht"] = {"mean_ms": float(ms[:n].mean()), "min_ms": float(ms[:n].min()),
                      "max_ms": float(ms[:n].max()), "launches": n, "steps": steps,
                      "launches_per_step": n / steps, "flops_total": float(fl[:n].sum()),
                      "ms_total": float(ms[:n].sum())}
    return out


def _roof(kernel, nbytes, gm, ident, what):
    t = gm["mean_ms"] * 1e-3
    achieved = nbytes / t / 1e9
    return {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_us": t * 1e6, "launches_per_step": 1.0,
            "ms_per_step": gm["mean_ms"], "algorithmic_bytes": nbytes, "algorithmic": what,
            "measured": f"HIP events around the kernel in each of the {gm['steps']} timed steps "
                        f"(an eager rerun of the same steps from the saved state), min "
                        f"{gm['min_ms'] * 1e3:.0f} us, max {gm['max_ms'] * 1e3:.0f} us, rerun "
                        f"bit-identical to the timed run: {ident}"}


def window_roofline(sim, gm, w, ident):
    """Roofline of the gather + push over the timed steps: algorithmic bytes
    over its mean duration in those steps.  Per electron: x, y, z, px, py, pz
    read and px, py, pz, g_inv written (10 words); in a step whose gather also
    applies the next first half push, the pushed x, y, z and the key written
    as well (3 words + 4 bytes)."""
    Np = sim.eons.Args["Np"]
    nbytes = Np * 10 * w + Np * (3 * w + 4) * gm["prepush_steps"] / gm["steps"]
    return _roof("k_gather_sorted (gather + Boris push)", nbytes, gm, ident,
                 "10 words per electron, + 3 words + 4 B with the next half push")


def deposit_roofline(sim, dm, w, ident):
    """Roofline of the electrons' J deposit pass (per-cell sums with the
    fused half push, their node sums and the push of the particles beyond
    the last full chunk: the three launches of one deposit call) over the
    timed steps.  Per electron: x, y, z, px, py, pz, g_inv, w read (8
    words); with the fused half push, the pushed x, y, z and the key written
    (3 words + 4 bytes); per launch the J planes written once (3 components,
    modes 0..M, real words Nr Nx (1 + 2M))."""
    Np = sim.eons.Args["Np"]
    g = sim.g
    planes = 3 * g["Nr"] * g["Nx"] * (1 + 2 * g["M"]) * w
    nbytes = Np * 8 * w + Np * (3 * w + 4) * dm["push_steps"] / dm["steps"] + planes
    return _roof("J deposit pass: k_cell_sums + k_node_add + k_push_keys_tail",
                 nbytes, dm, ident,
                 "8 words per electron, + 3 words + 4 B with the half push, + J planes")


def dht_roofline(sim, dm, ident):
    """Roofline of the DHT products (k_gemm / k_gemm_bf16x3) over the timed
    steps: the algorithmic flops of every launch (2 M K N per product;
    16 + 32M real products of (Nr-1)^2 Nx per step) over the launches' summed
    durations, i.e. mean flops per launch over the mean launch duration."""
    mixed = sim.comm.mixed
    tf = dm["flops_total"] / (dm["ms_total"] * 1e-3) / 1e12
    # the split-BF16 product does three BF16 products per FP32-equivalent
    # one: its matrix-core work, against the BF16 peak, is what bounds it
    work = 3 * tf if mixed else tf
    peak = MFMA_PEAK_TFS["bf16"] if mixed else MFMA_PEAK_TFS["f64"]
    out = {"kernel": "k_gemm_bf16x3 (split-BF16 DHT products)" if mixed
           else "k_gemm<double> (FP64 DHT products)",
           "bound": "mfma", "achieved": work, "peak": peak, "unit": "TFLOP/s", "frac": work / peak,
           "traffic": None, "avg_launch_us": dm["mean_ms"] * 1e3,
           "launches_per_step": dm["launches_per_step"],
           "ms_per_step": dm["ms_total"] / dm["steps"],
           "algorithmic_flops": (3 if mixed else 1) * dm["flops_total"] / dm["launches"],
           "algorithmic": "2 M K N flops per product (per step (16 + 32 M) 2 (Nr-1)^2 Nx)"
                          + (", times 3 BF16 products per product" if mixed else ""),
           "measured": f"HIP events recorded by the library around each of the {dm['launches']} "
                       f"product launches of the {dm['steps']} timed steps (an eager rerun from "
                       f"the saved state), min {dm['min_ms'] * 1e3:.0f} us, max "
                       f"{dm['max_ms'] * 1e3:.0f} us, rerun bit-identical to the timed run: {ident}"}
    if mixed:
        out["peak_note"] = ("BF16 matrix-core work (3 BF16 products per FP32-equivalent product) "
                            "against the dense BF16 MFMA peak")
        out["fp32_equivalent_tflops"] = tf
        out["fp32_equivalent_over_fp32_mfma_peak"] = tf / MFMA_PEAK_TFS["f32"]
    return out


def dominant_roofline(sim, tm, w, config, form):
    """The roofline object of the kernel family that takes the most time per
    step among the DHT products, the gather + push and the J deposit pass,
    with all three beside it."""
    ident = tm["replay_identical"]
    r = {"gather": window_roofline(sim, tm["gather"], w, ident),
         "deposit_J": deposit_roofline(sim, tm["deposit_J"], w, ident)}
    if "dht" in tm:
        r["dht"] = dht_roofline(sim, tm["dht"], ident)
    for k in r:
        r[k]["traffic"], src = pmc_traffic(config, form, k)
        r[k]["traffic_source"] = (f"{src}: rocprofv3 counter run of the same kernels, "
                                  "not measured in this run") if src else None
    top = max(r, key=lambda k: r[k]["ms_per_step"])
    out = dict(r[top])
    out["kernels"] = r
    return out


def moving_window(sim, events=4, steps_per_event=20):
    """Frame (frame.py:4-64) on the FP64 run: window shifts of 20 cells at
    c with plasma injection on the GPU, sort and align of both species,
    every 20 steps; times the frame events and the eager steps between."""
    import torch
    import chimeracl_amd as CA
    s = sim.solver
    frame = CA.Frame({"Velocity": 1.0, "dt": s.Args["dt"], "Steps": steps_per_event})
    sim.ions.Args["InjectorSource"] = sim.eons
    loop = CA.PIC_loop(solvers=[s], species=[sim.eons, sim.ions])  # events driven here
    t_ev = t_all = 0.0
    n0 = sim.eons.Args["Np"]
    for e in range(events):
        torch.cuda.synchronize()
        t0 = time.time()
        frame.shift_grids(grids=[s])
        frame.inject_plasma(species=[sim.eons, sim.ions], grid=s)
        torch.cuda.synchronize()
        t1 = time.time()
        for _ in range(steps_per_event - 1):
            loop.step()
        torch.cuda.synchronize()
        t_ev += t1 - t0
        t_all += time.time() - t0
    nsteps = events * steps_per_event
    return {"events": events, "steps": nsteps, "ms_per_event": t_ev / events * 1e3,
            "ms_per_step_with_window": t_all / nsteps * 1e3,
            "particle_steps_per_s": sim.eons.Args["Np"] * nsteps / t_all,
            "electrons_before": n0, "electrons_after": sim.eons.Args["Np"],
            "note": "eager steps (a window step is not graph-captured)"}


# ---------------------------------------------------------------- differentiable form
def diff_section():
    import torch
    import chimeracl_amd.adjoint as A
    out = {}

    def grads(sim, a0, R, kv, n, objective):
        rg = sim.solver.Args["Rgrid"]
        pa = torch.tensor(a0, dtype=torch.float64, requires_grad=True)
        pR = torch.tensor(R, dtype=torch.float64, requires_grad=True)
        pk = torch.tensor(kv, dtype=torch.float64, requires_grad=True)
        L = objective(A.run(sim, pa, pR, pk, n), rg)
        L.backward()
        g = np.concatenate([[pa.grad.item(), pR.grad.item()], pk.grad.numpy()])
        res = {"objective": float(L.detach()), "steps": n, "checkpoint_every": sim.ckpt,
               "electrons": sim.Np, "forward_s": sim.timings["forward"],
               "backward_s": sim.timings["backward"],
               "backward_over_forward": sim.timings["backward"] / sim.timings["forward"]}
        return g, res

    def build(name, precision):
        sim, a0, R, kv, n = A.from_config(name, precision)
        sim.forward(a0, R, kv, 2)  # eager step + graph capture
        return sim, a0, R, kv, n

    def free(sim):
        del sim
        gc.collect()
        torch.cuda.empty_cache()

    def field_state(sim, a0, R, kv, n):
        """The forward run's final E, B, J planes and electron momenta at the
        nominal parameters (FP64 copies), and the charge residual of its last
        step (rho of step n-1 from a run of n-1 steps; FP64 operators of the
        run's FP64 twin solver)."""
        M = sim.solver.Args["M"]
        with torch.no_grad():
            sim.forward(a0, R, kv, n - 1)
        torch.cuda.synchronize()
        rho_prev = [sim.solver.DataDev[f"rho_fb_m{m}"].to(torch.complex128).clone()
                    for m in range(M + 1)]
        with torch.no_grad():
            sim.forward(a0, R, kv, n)
        torch.cuda.synchronize()
        d, e = sim.solver.DataDev, sim.eons.DataDev
        st = {f: [d[f"{f}{c}_m{m}"][1:].to(torch.complex128 if m else torch.float64).clone()
                  for c in "xyz" for m in range(M + 1)] for f in ("E", "B", "J")}
        st["p"] = [e[k].double().clone() for k in ("px", "py", "pz")]
        calc = sim._twin_solver()
        if "divJ_fb_m0" not in calc.DataDev:
            add_div_scratch(calc)
        st["charge_residual"] = charge_residual(calc, sim, rho_prev)
        return st

    def state_errs(a, ref):
        return {k: _rel(list(zip(a[k], ref[k]))) for k in ("E", "B", "J", "p")}

    def state_report(a, ref, n):
        return {"step": n, "rel_l2": state_errs(a, ref),
                "charge_residual": {"fp64": ref["charge_residual"],
                                    "mixed": a["charge_residual"]}}

    gam = lambda outs, rg: A.objective_gamma2(*outs[:3], threshold=5.0, tau=0.25)
    wake = lambda outs, rg: A.objective_wake_energy(list(outs[3:]), rg, ncols=256)

    # C4, FP64: gradients against central finite differences (step 1e-4 relative)
    sim, a0, R, kv, n = build("C4", "f64")
    g64, res = grads(sim, a0, R, kv, n, gam)
    p0 = np.concatenate([[a0, R], kv])
    fd = np.zeros_like(p0)
    t0 = time.time()
    for j in range(len(p0)):
        h = 1e-4 * abs(p0[j])
        dp = np.zeros_like(p0)
        dp[j] = h
        with torch.no_grad():
            fp = float(gam(sim.forward(*(p0 + dp)[:2], (p0 + dp)[2:], n), None))
            fm = float(gam(sim.forward(*(p0 - dp)[:2], (p0 - dp)[2:], n), None))
        fd[j] = (fp - fm) / (2 * h)
    log(f"[bench] C4 finite differences ({2 * len(p0)} runs) in {time.time() - t0:.1f}s")
    kn_fd, kn_g = fd[2:], g64[2:]
    big = np.abs(kn_fd) > 1e-3 * np.abs(kn_fd).max()
    res["fd_check"] = {
        "step": "1e-4 relative, central",
        "a0": {"adjoint": g64[0], "fd": fd[0], "rel_err": abs(g64[0] - fd[0]) / abs(fd[0])},
        "w0": {"adjoint": g64[1], "fd": fd[1], "rel_err": abs(g64[1] - fd[1]) / abs(fd[1])},
        "knots_max_err_over_max_grad": float(np.abs(kn_g - kn_fd).max() / np.abs(kn_fd).max()),
        "knots_median_rel_err_significant": float(np.median(np.abs(kn_g - kn_fd)[big] / np.abs(kn_fd)[big])),
        "knots_significant": int(big.sum()),
        "knots": len(kn_fd)}
    out["C4_f64"] = res
    ref_state = field_state(sim, a0, R, kv, n)
    free(sim)
    sim, a0, R, kv, n = build("C4", "mixed")
    gmx, res = grads(sim, a0, R, kv, n, gam)
    res["grad_err_vs_fp64_adjoint"] = _grad_errs(gmx, g64)
    res["fields_vs_fp64_forward"] = state_report(field_state(sim, a0, R, kv, n), ref_state, n)
    out["C4_mixed"] = res
    del ref_state
    free(sim)

    # C5: the mixed adjoint against the FP64 adjoint, after 50 steps and
    # after the full 400 (by then ~1e5 electrons have left the valid cells
    # at the right edge, which makes the objective rough: see the docs)
    ref = {}
    sim, a0, R, kv, n = build("C5", "f64")
    for ns in (50, n):
        ref[ns] = grads(sim, a0, R, kv, ns, wake)
    ref_state = field_state(sim, a0, R, kv, n)
    free(sim)
    sim, a0, R, kv, n = build("C5", "mixed")
    for ns in (50, n):
        g, res = grads(sim, a0, R, kv, ns, wake)
        res["grad_err_vs_fp64_adjoint"] = _grad_errs(g, ref[ns][0])
        res["fp64"] = {k: ref[ns][1][k] for k in ("objective", "forward_s", "backward_s",
                                                   "backward_over_forward")}
        out[f"C5_mixed_{ns}steps"] = res
    out[f"C5_mixed_{n}steps"]["fields_vs_fp64_forward"] = state_report(
        field_state(sim, a0, R, kv, n), ref_state, n)
    del ref_state
    free(sim)
    return out


def aligned_section(args, k, diag_at):
    """Both forms again with PIC_loop(align_every=k) (k = 0: never
    reordered): same workload, warmup and timed window as the main
    measurement (timed from the first diagnostics step diag_at[0] on)."""
    import torch
    out = {"align_every": k}
    first = min(diag_at or [0])
    for name, prec in (("f64", "f64"), ("mixed", "mixed")):
        if prec == "mixed" and args.no_mixed:
            continue
        sm = Sim(args.config, prec, False, k)
        sm.step(max(args.warmup, first))
        torch.cuda.synchronize()
        t0 = time.time()
        sm.loop.run(args.steps)
        torch.cuda.synchronize()
        dt = time.time() - t0
        Ne = sm.eons.Args["Np"]
        out[name] = {"ms_per_step": dt / args.steps * 1e3, "value": Ne * args.steps / dt,
                     "unit": "particle-steps/s", "steps": args.steps}
        log(f"[bench] {name}, align_every={k}: {args.steps} steps in {dt:.3f}s: "
            f"{dt / args.steps * 1e3:.3f} ms/step, {Ne * args.steps / dt:.4g} particle-steps/s")
        del sm
        gc.collect()
        torch.cuda.empty_cache()
    return out


def config_section(name, diag_at, oracle_steps=0, align_every=0):
    """Configuration `name` in both forms, in lockstep: at the steps in
    diag_at, the mixed run's relative L2 errors against the FP64 run (E, B,
    J, momenta), both charge residuals and energies; with oracle_steps, the
    FP64 run against the CPU FP64 oracle after that many steps."""
    import torch
    from chimeracl_amd import config as C
    t_all = time.time()
    fp = Sim(name, "f64", oracle_steps > 0, align_every)
    mx = Sim(name, "mixed", False, align_every)
    M = fp.solver.Args["M"]
    add_div_scratch(fp.solver)
    cfg = C.CONFIGS[name]
    out = {"workload": f"{name}: " + cfg["desc"], "Nz": cfg["Nx"], "Nr": cfg["Nr"], "modes": M + 1,
           "ppc": list(cfg["ppc"]), "electrons": fp.eons.Args["Np"], "steps": max(diag_at),
           "profile": cfg.get("profile"), "align_every": align_every}
    targets = sorted(set(diag_at) | ({oracle_steps} if oracle_steps else set()))
    t_gpu = 0.0
    for t in targets:
        torch.cuda.synchronize()
        t0 = time.time()
        for sm in (fp, mx):
            sm.step(t - 1 - sm.steps)
        torch.cuda.synchronize()
        t_gpu += time.time() - t0
        prev = [rho_copy(sm) for sm in (fp, mx)]
        for sm in (fp, mx):
            sm.step(1)
        torch.cuda.synchronize()
        if t == oracle_steps:
            t0 = time.time()
            oloop = fp.orc[3]
            for _ in range(oracle_steps):
                oloop.step()
            out["fp64_vs_oracle"] = {"step": t, "rel_l2": oracle_errors(fp, M),
                                     "oracle_s": time.time() - t0}
            fp.orc = None
        if t in diag_at:
            out[f"step_{t}"] = {
                "fp64": {"charge_residual": charge_residual(fp.solver, fp, prev[0]),
                         "energy": energies(fp)},
                "mixed": {"charge_residual": charge_residual(fp.solver, mx, prev[1]),
                          "energy": energies(mx), "rel_l2_vs_fp64": mixed_errors(mx, fp, M)}}
        log(f"[bench] {name} step {t}: {out.get(f'step_{t}', out.get('fp64_vs_oracle'))}")
    out["gpu_seconds_both_forms"] = t_gpu
    out["section_seconds"] = time.time() - t_all
    del fp, mx
    gc.collect()
    torch.cuda.empty_cache()
    return out


def _grad_errs(g, ref):
    return {"a0": abs(g[0] - ref[0]) / abs(ref[0]), "w0": abs(g[1] - ref[1]) / abs(ref[1]),
            "knots_rel_l2": float(np.linalg.norm(g[2:] - ref[2:]) / np.linalg.norm(ref[2:]))}


# ---------------------------------------------------------------- main
def measure(args, name, diag_at, want_cpu, cpu_sample, with_kernels=True, with_window=True):
    """One workload in both forms: the CPU-oracle baseline on a bounded sample
    (and the FP64 run's error against it), the lockstep advance with the
    mixed-vs-FP64 diagnostics at diag_at, args.steps timed graph-replayed
    steps per form after args.warmup steps (and after the first diagnostics
    step), the same steps rerun eagerly with HIP events (rooflines), the
    per-phase kernel table and the moving-window run."""
    import torch
    from chimeracl_amd import config as C
    cfg = C.CONFIGS[name]
    M = cfg["M"]
    arr = load_arrays(name)
    fp = Sim(name, "f64", want_cpu, args.align_every, arr=arr)
    mx = None if args.no_mixed else Sim(name, "mixed", False, args.align_every, arr=arr)
    del arr
    gc.collect()
    sims = [s for s in (fp, mx) if s is not None]
    Ne = fp.eons.Args["Np"]
    add_div_scratch(fp.solver)
    torch.cuda.synchronize()

    # ---- CPU baseline (oracle) on a bounded sample: cpu_warmup untimed
    # steps, then cpu_steps timed steps; the FP64 GPU run is compared with it
    # after the same steps
    cpu = errs = None
    if want_cpu:
        from oracle import oracle as O
        cw, cs = cpu_sample
        nthreads = O.lib().orc_num_threads()
        s, e, i, oloop = fp.orc
        for _ in range(cw):
            oloop.step()
        oloop.Timer = dict.fromkeys(oloop.Timer, 0.0)
        t0 = time.time()
        for _ in range(cs):
            oloop.step()
        tc = time.time() - t0
        nst = cw + cs
        for sm in sims:
            sm.step(nst)  # eager first step, captured graph from the second
        torch.cuda.synchronize()
        errs = oracle_errors(fp, M)
        errs["after_steps"] = nst
        cpu = {"value": Ne * cs / tc, "unit": "particle-steps/s", "cores": nthreads,
               "kind": "port", "pic_steps_per_s": cs / tc,
               "phases_s_per_step": {k: v / cs for k, v in oloop.Timer.items()},
               "host": host_info(),
               "sample": f"{cs} full PIC steps (after {cw} untimed) of the same workload ({Ne} "
                         f"electrons, {fp.g['Nx']}x{fp.g['Nr']}, M={M}) by the CPU FP64 oracle "
                         f"(C/OpenMP particle loops, numpy FFT/BLAS), {tc:.1f}s"}
        log(f"[bench] {name} oracle {cs} steps in {tc:.1f}s on {nthreads} threads; errors {errs}")
        fp.orc = None
        del s, e, i, oloop
        gc.collect()

    diags = {}

    def advance_to(n):
        """Lockstep advance of both runs to step n, with the diagnostics of
        the mixed run against the FP64 run at the requested steps."""
        while fp.steps < n:
            nxt = min([d for d in diag_at if d > fp.steps] + [n])
            k = nxt - fp.steps
            if nxt in diag_at and k >= 1:
                for sm in sims:
                    sm.step(k - 1)
                prev = [rho_copy(sm) for sm in sims]
                for sm in sims:
                    sm.step(1)
                torch.cuda.synchronize()
                d = {"fp64": {"charge_residual": charge_residual(fp.solver, fp, prev[0]),
                              "energy": energies(fp)}}
                if mx is not None:
                    d["mixed"] = {"charge_residual": charge_residual(fp.solver, mx, prev[1]),
                                  "energy": energies(mx),
                                  "rel_l2_vs_fp64": mixed_errors(mx, fp, M)}
                diags[str(nxt)] = d
                log(f"[bench] {name} step {nxt}: {d}")
            else:
                for sm in sims:
                    sm.step(k)

    # ---- warmup, then the timed steps of each form (graph replay)
    for sm in sims:
        sm.step(args.warmup)
    first_diag = min(diag_at) if diag_at else 0
    advance_to(max(first_diag, fp.steps))
    timing, tm = {}, {}
    for form, sm in (("f64", fp), ("mixed", mx)):
        if sm is None:
            continue
        snap = snapshot(sm)
        it0 = sm.loop.it
        torch.cuda.synchronize()
        t0 = time.time()
        sm.loop.run(args.steps)
        torch.cuda.synchronize()
        dt = time.time() - t0
        timing[form] = dt
        # the same steps again from the saved state, eagerly, with HIP events
        # around the gather, the J deposit and every DHT product (the step
        # is deterministic, so the rerun ends in the same state: checked)
        after = fingerprint(sm)
        restore(snap)
        del snap
        sm.loop.it = it0
        tm[form] = replay_timed(sm, args.steps)
        same = fingerprint(sm)
        tm[form]["replay_identical"] = all(torch.equal(x, y) for x, y in zip(after, same))
        log(f"[bench] {name} {form}: eager rerun identical: {tm[form]['replay_identical']}; "
            f"gather {tm[form]['gather']}; J deposit {tm[form]['deposit_J']}; "
            f"DHT {tm[form].get('dht')}")
        del after, same
        sm.steps += args.steps
        log(f"[bench] {name} {form}: {args.steps} steps in {dt:.3f}s: "
            f"{dt / args.steps * 1e3:.3f} ms/step, {Ne * args.steps / dt:.4g} particle-steps/s")
    if mx is not None and mx.steps != fp.steps:
        raise RuntimeError("lockstep broken")
    if diag_at:
        advance_to(max(max(diag_at), fp.steps))

    dt = timing["f64"]
    res = {"workload": f"{name}: " + cfg["desc"], "Nz": cfg["Nx"], "Nr": cfg["Nr"], "modes": M + 1,
           "ppc": list(cfg["ppc"]), "electrons": Ne, "dt": C.DZ, "align_every": args.align_every,
           "profile": cfg.get("profile"),
           "f64": {"value": Ne * args.steps / dt, "unit": "particle-steps/s",
                   "ms_per_step": dt / args.steps * 1e3, "pic_steps_per_s": args.steps / dt,
                   "roofline": dominant_roofline(fp, tm["f64"], 8, name, "f64")},
           "fp64_oracle_error": errs, "cpu_baseline": cpu, "diagnostics": diags,
           "t_roof_frac": {}}
    b = roof_bound(fp.g, Ne)
    res["t_roof_frac"]["f64"] = b["f64_ms"] / res["f64"]["ms_per_step"]
    res["t_roof_ms"] = b
    if mx is not None:
        dtm = timing["mixed"]
        res["mixed"] = {"value": Ne * args.steps / dtm, "unit": "particle-steps/s",
                        "ms_per_step": dtm / args.steps * 1e3, "pic_steps_per_s": args.steps / dtm,
                        "speedup_vs_f64": dt / dtm,
                        "dtype": "f32 particles/FFT, split-bf16x3 DHT",
                        "roofline": dominant_roofline(mx, tm["mixed"], 4, name, "mixed")}
        res["t_roof_frac"]["mixed"] = b["mixed_ms"] / res["mixed"]["ms_per_step"]
    from chimeracl_amd.pic_loop import counter_collection_active
    if with_kernels and not counter_collection_active():  # the bursts are graph replays
        res["kernels"] = {"f64": kernel_table(fp, args.kernel_reps, 8)}
        if mx is not None:
            res["kernels"]["mixed"] = kernel_table(mx, args.kernel_reps, 4)
    if with_window:
        res["moving_window"] = moving_window(fp)
    del fp, mx, sims
    gc.collect()
    torch.cuda.empty_cache()
    return res


def roof_bound(g, Ne):
    """Lower bound of a step's time: (B_part + B_grid) / 8 TB/s + F_DHT /
    P_MFMA, with B_part = N_e (21 w + 24), B_grid = 20 r1 + 53 s1 + 3 w S (M+1),
    F_DHT = (16 + 32 M) 2 (Nr-1)^2 Nx (10 2 (Nr-1)^2 Nx for M = 0)."""
    Nx, Nr, M = g["Nx"], g["Nr"], g["M"]
    S, K = (Nr - 1) * Nx, M + 1
    F = (16 + 32 * M if M > 0 else 10) * 2.0 * (Nr - 1) ** 2 * Nx
    out = {"F_DHT": F}
    for form, w, peak in (("f64", 8, MFMA_PEAK_TFS["f64"]), ("mixed", 4, MFMA_PEAK_TFS["f32"])):
        r1 = Nr * Nx * w * (1 + 2 * M)
        s1 = 2 * w * S * K
        bp, bg = Ne * (21 * w + 24), 20 * r1 + 53 * s1 + 3 * w * S * K
        out[f"{form}_bytes"] = bp + bg
        out[f"{form}_ms"] = ((bp + bg) / (HBM_PEAK_GBS * 1e9) + F / (peak * 1e12)) * 1e3
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.gather_only:
        gather_only(args)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: particle sharding over several "
                          "GPUs (an all-reduce of the deposited J and rho) is not implemented",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch  # noqa: F401

    def diag_for(name):
        spec = args.diag_steps if (args.diag_steps is not None and name == args.config) \
            else DEFAULTS.get(name, {}).get("diag", "")
        return [int(v) for v in spec.split(",") if v]

    def cpu_for(name):
        cw, cs = DEFAULTS.get(name, {}).get("cpu", (2, 10))
        if name == args.config:
            cw = cw if args.cpu_warmup is None else args.cpu_warmup
            cs = cs if args.cpu_steps is None else args.cpu_steps
        return cw, cs

    head = measure(args, args.config, diag_for(args.config), not args.no_cpu, cpu_for(args.config),
                   with_window=not args.no_window)
    second = None
    if args.secondary and args.secondary != args.config:
        second = measure(args, args.secondary, diag_for(args.secondary), False, None,
                         with_kernels=False, with_window=False)
    aligned = (aligned_section(args, args.compare_align, diag_for(args.config))
               if args.compare_align >= 0 and args.compare_align != args.align_every else None)
    configs = None
    if not args.no_configs and "C1" not in (args.config, args.secondary):
        configs = {"C1": config_section("C1", [100, 200], 0 if args.no_cpu else 5, args.align_every)}
    diff = None
    if not args.no_diff:
        t0 = time.time()
        diff = diff_section()
        log(f"[bench] differentiable section in {time.time() - t0:.1f}s")

    f64 = head["f64"]
    config = {k: head[k] for k in ("workload", "Nz", "Nr", "modes", "ppc", "electrons", "dt",
                                    "align_every", "profile")}
    out = {
        "metric": "particle-steps/s",
        "value": f64["value"],
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "pic_steps_per_s": f64["pic_steps_per_s"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded plasma with the workload's density profile, Gaussian laser)",
        "config": config,
        "fp64_oracle_error": head["fp64_oracle_error"],
        "roofline": f64["roofline"],
        "cpu_baseline": head["cpu_baseline"],
        "mixed": head.get("mixed"),
        "t_roof_frac": head["t_roof_frac"],
        "t_roof_ms": head["t_roof_ms"],
        "diagnostics": head["diagnostics"],
        "kernels": head.get("kernels"),
        "moving_window": head.get("moving_window"),
        "secondary": second,
        "align_compare": aligned,
        "configs": configs,
        "differentiable": diff,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
