"""Benchmark of the MI355X density stage of NewGalFinder.

One step = the whole per-halo density stage on device-resident float32
inputs: TSC deposit -> FFT Gaussian smoothing -> TSC read-back -> exact
32-NN lists (int64) + W4 kNN density.  Workload (default): one halo of
4,194,304 particles (NFW c=10, r200=0.4, plus 20 Plummer subhaloes with
log-uniform mass fractions 0.1-3%, a=0.01), 256^3 grid over the unit box,
RG = 1.5 cells, k = 32; synthetic seeded data.

Prints ONE JSON line (rank 0, N=1):
  value          particles/s through the stage (inputs already in HBM)
  roofline       dominant kernel (knn_packed_k, the kNN query) algorithmic
                 bytes per launch / its mean launch time (HIP events on the
                 launch stream); its VALU issue fraction from the committed
                 counters
  cpu_baseline   the reference's own C/OpenMP code compiled from its sources
                 (oracle/_ref, "reference") on the same halo, timed on this
                 host's cores (kNN list builder on a cell sample, scaled)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, public peak
# launch schedule measured on MI355X by python -m galaxyfinder_amd.tune
TUNED_SCHEDULE = os.path.join(ROOT, "galaxyfinder_amd", "schedule_mi355x.json")
CALIB_COPY_BYTES = 1 << 29  # plain device copy (negation) that scales the HBM byte counters
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "knn_packed_k_traffic.txt")
COUNTER_FILE = os.path.join(ROOT, "profiles", "knn_packed_k_counters.md")
# wave64 VALU instructions the chip can issue per second: 256 CUs x 4 SIMDs x
# 2.4 GHz, one wave64 instruction per SIMD every 4 cycles
VALU_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4


def measured_valu(n, k):
    """VALU wave-instructions per knn_packed_k launch from the committed
    counter summary, when it was taken on this workload (N and k in its
    title line)."""
    try:
        text = open(COUNTER_FILE).read()
    except OSError:
        return None
    if ("N=%s" % format(n, ",")) not in text.splitlines()[0] or ("k=%d" % k) not in text:
        return None
    for line in text.splitlines():
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 2 and cells[1] == "SQ_INSTS_VALU":
            return float(cells[2])
    return None


def measured_traffic(n, k):
    """HBM bytes per knn_packed_k launch from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE, each scaled by the same counter's reading for
    a plain copy of CALIB_COPY_BYTES), when it was taken on this workload."""
    try:
        with open(TRAFFIC_FILE) as f:
            kv = dict(line.split(":", 1) for line in f
                      if ":" in line and not line.startswith("#"))
        kv = {a.strip(): b.strip() for a, b in kv.items()}
        if int(kv["particles"]) != n or int(kv["k"]) != k:
            return None
        return int(float(kv["traffic_bytes_per_launch"]))
    except (OSError, KeyError, ValueError):
        return None


def stage_bytes(n, g_cells, k):
    """Algorithmic HBM bytes per stage (FP32 positions/masses, int64 ids)."""
    return {
        "tsc": 16 * n + 4 * g_cells,
        "smooth": 8 * g_cells,
        "findden": 16 * n + 4 * g_cells,
        "knn": 16 * n + 4 * n + 8 * k * n,
    }


def cpu_baseline(pos, mass, spec, k, knn_stride):
    """Time the reference's own C/OpenMP code (oracle/_ref, compiled from the
    reference sources) on the same halo, on every host core:
    assign_density_TSC (tsc_omp2.c:166) and findDen (tsc_omp2.c:40) on the
    whole halo; the serial octree build (new_Make_Tree_Near, nnost.c:199) and
    lagFindStellarCore's warm-started Find_Near_PreCopied list builder
    (nnost.c:1361-1529) on the cells whose flat index is a multiple of
    knn_stride, its time scaled to every particle.  gaussian_Smoothing
    (gsmooth.c) needs FFTW, which is absent, so the smoothing leg is the
    oracle's FP64-FFT restatement.  Without the reference build, the whole
    leg is the oracle's restatement ("port")."""
    import ctypes
    from oracle import oracle as O
    n = len(pos)
    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    g = (nx, ny, nz, spec["xmin"], spec["ymin"], spec["zmin"], spec["cellsize"])
    nth = O.num_threads()
    R = O.ref_lib()
    x, y, z = (np.ascontiguousarray(pos[:, a], dtype=np.float64) for a in range(3))
    m = np.ascontiguousarray(mass, dtype=np.float64)
    st = {}
    t0 = time.perf_counter()
    if R is not None:
        R.ref_init(8 << 30)
        den = np.zeros((nz, ny, 2 * (nx // 2 + 1)), dtype=np.float32)
        R.ref_tsc(O._p(x, O._dp), O._p(y, O._dp), O._p(z, O._dp), O._p(m, O._dp), None, n,
                  O._p(den, O._fp), *g, 5, nth)
    else:
        den, _ = O.tsc_deposit(pos, mass, g, nthreads=nth)
    t1 = time.perf_counter()
    sm = O.gsmooth(den, nx, ny, nz, spec["cellsize"], spec["RG"])
    t2 = time.perf_counter()
    if R is not None:
        rb = np.zeros(n, dtype=np.float32)
        R.ref_findden(O._p(x, O._dp), O._p(y, O._dp), O._p(z, O._dp), n, O._p(rb, O._fp),
                      *g[3:], O._p(sm, O._fp), *g[:3], nth)
    else:
        O.findden(pos, g, sm)
    t3 = time.perf_counter()
    st["tsc"], st["smooth"], st["findden"] = t1 - t0, t2 - t1, t3 - t2
    if R is not None:
        idx = np.empty((n, k), dtype=np.int64)
        nq, tt = ctypes.c_longlong(), ctypes.c_double()
        R.ref_knn_lag(O._p(x, O._dp), O._p(y, O._dp), O._p(z, O._dp), n, k, *g[3:], *g[:3],
                      O._p(idx, O._lp), nth, knn_stride, ctypes.byref(nq), ctypes.byref(tt))
        t4 = time.perf_counter()
        st["knn_tree"] = tt.value
        st["knn_lists_scaled"] = (t4 - t3 - tt.value) * n / max(nq.value, 1)
        kind = "reference"
        sample = ("the reference's own C/OpenMP code (oracle/_ref, gcc -O2 -fopenmp from "
                  "NewGalFinder/tsc_omp2.c and nnost.c) on the same %d-particle halo: "
                  "assign_density_TSC + findDen on all particles, the serial octree build, "
                  "and lagFindStellarCore's warm-started Find_Near_PreCopied list builder on "
                  "every %d-th grid cell (%d queries, time scaled to all particles); "
                  "smoothing by the oracle's FP64-FFT restatement (gsmooth.c needs FFTW, "
                  "absent); %.1f s of CPU wall time" % (n, knn_stride, nq.value, t4 - t0))
    else:
        sel = np.random.default_rng(0).choice(n, min(1 << 18, n), replace=False)
        tb0 = time.perf_counter()
        O.knn(pos[sel[:64]], k, cpos=pos)
        tb1 = time.perf_counter()
        O.knn(pos[sel], k, cpos=pos)
        t4 = time.perf_counter()
        st["knn_tree"] = tb1 - tb0
        st["knn_lists_scaled"] = max(t4 - tb1 - (tb1 - tb0), 0.0) * n / len(sel)
        kind = "port"
        sample = ("the oracle's C/OpenMP restatement (reference build absent) on the same "
                  "%d-particle halo, k-d tree kNN queried for %d particles (time scaled); "
                  "%.1f s of CPU wall time" % (n, len(sel), t4 - t0))
    total = sum(st.values())
    return {"value": n / total, "unit": "particles/s", "cores": nth, "kind": kind,
            "sample": sample, "stages_s": {a: round(b, 4) for a, b in st.items()}}


def peaks_bench(x, y, z, m, g, den, rho_t, nbr, rho_k, spec, reps=10, cpu=True):
    """The integer peak stage on the headline halo's own stage outputs: grid
    cores (nnost.c:1115-1271; densph = the read-back densities), finddenpeak
    (subhaloden.mod6.c:416-492; the W4 kNN densities and 32-NN lists),
    MergingPeak of the grid cores (subhaloden.mod6.c:341-415, link 4e-3),
    and the watershed on the finddenpeak cores (FindCoreDensity, :1167-1475;
    the shells, :572-678 and :2995-3024; FP64 masses).
    The synthetic halo's densities are not in the reference's units, so the
    thresholds are the medians of the positive grid values and of the kNN
    densities (the reference's PEAKTHRESHOLD = 2000 would select nothing;
    for the reference's finddenpeak, whose threshold is compiled in, the
    densities are scaled so that the same share passes it).
    GPU times: HIP events on the launch stream, median of `reps`.  CPU: the
    reference's own finddenpeak and MergingPeak (oracle/_ref) and the
    oracle's grid cores, on the same inputs."""
    import torch
    from galaxyfinder_amd import density as D
    n, k = x.numel(), nbr.shape[1]
    ng = spec["nx"]
    # every 4th particle a star
    types = torch.from_numpy(np.where(np.arange(n) % 4 == 0, 1, 2).astype(np.int32)).cuda()
    den_h = den.cpu().numpy()
    cells = den_h[..., :ng]
    thr_g = float(np.quantile(cells[cells > 0], 0.5))
    thr_k = float(np.quantile(rho_k.cpu().numpy(), 0.5))
    st = torch.cuda.current_stream()

    def timed(fn):
        fn()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(reps)]
        for a, b in ev:
            a.record(st)
            fn()
            b.record(st)
        torch.cuda.synchronize()
        D.dev_status()
        return float(np.median([a.elapsed_time(b) for a, b in ev]))

    out = {"thresholds": {"grid": thr_g, "knn": thr_k}}
    res = {}

    def gc():
        res["gc"] = D.dev_grid_cores(x, y, z, rho_t, g, den, threshold=thr_g, nbuff=10)
    t = timed(gc)
    core = res["gc"][0][:int(res["gc"][1].item())].clone()
    gb = (4 * den.numel() + 20 * n) / (t * 1e-3) / 1e9
    out["grid_cores"] = {"ms": round(t, 4), "cores": int(core.numel()),
                         "alg_gbs": round(gb, 1)}

    def fdp():
        res["fd"] = D.dev_finddenpeak(rho_k, nbr, types, 0, thr_k)
    t = timed(fdp)
    gb = (8 + 8 * k) * n / (t * 1e-3) / 1e9
    out["finddenpeak"] = {"ms": round(t, 4), "cores": int(res["fd"][1].item()),
                          "alg_gbs": round(gb, 1)}
    key = rho_t[core.long()].contiguous()

    def mp():
        res["mp"] = D.dev_merging_peak(x, y, z, core, key)
    t = timed(mp)
    out["merging_peak"] = {"ms": round(t, 4), "cores_in": int(core.numel()),
                           "cores_kept": int(res["mp"][1].item())}
    # the watershed of subhalo_den on the finddenpeak cores: FindCoreDensity,
    # the shell thresholds, GetShellParticlesPeaks over all of them and
    # SaveRemainingParticles2LastShell (host-orchestrated: wall time around
    # synchronised calls, median of 3)
    peak = res["fd"][0][:int(res["fd"][1].item())].to(torch.int32).contiguous()
    m64 = m.double()

    def wall(fn, n=3):
        ts = []
        for _ in range(n):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    def fcd():
        res["fcd"] = D.dev_find_core_density(rho_k, nbr, peak, types, m64)
    t_fcd = wall(fcd)
    o = res["fcd"]
    thr = D.dev_shell_thresholds(rho_k, o["coredensity"])

    def shells(nthr):
        def run():
            fl, hid = o["flag"].clone(), o["haloid"].clone()
            res["sh"] = D.dev_shell_particles_peaks(rho_k, nbr, peak, o["coredensity"],
                                                    thr[:nthr], fl, hid)
        return run
    t_sh = wall(shells(len(thr)))
    nshell = int(res["sh"][0])
    t_sh6 = wall(shells(6))
    out["watershed"] = {
        "cores": int(peak.numel()), "find_core_density_ms": round(t_fcd, 3),
        "members": int(o["nummem"].sum().item()), "thresholds": len(thr),
        "shells_ms": round(t_sh, 2), "shells": nshell,
        "shells_upper6_ms": round(t_sh6, 2)}
    if cpu:
        from oracle import oracle as O
        R = O.ref_lib()
        pos = np.stack([a.cpu().numpy() for a in (x, y, z)], axis=1).astype(np.float64)
        rt_h, rk_h, nb_h = rho_t.cpu().numpy(), rho_k.cpu().numpy(), nbr.cpu().numpy()
        ty_h = types.cpu().numpy()
        grid = (ng, spec["ny"], spec["nz"], spec["xmin"], spec["ymin"], spec["zmin"],
                spec["cellsize"])
        c = {}
        t0 = time.perf_counter()
        gcore = O.grid_cores(pos, ty_h, 5, rt_h, den_h, grid, thr_g, 10)
        c["grid_cores_oracle_s"] = round(time.perf_counter() - t0, 4)
        if R is not None:
            x_, y_, z_ = (np.ascontiguousarray(pos[:, a]) for a in range(3))
            cap = n
            pk = np.zeros(cap, dtype=np.int32)
            cx, cy, cz, cd = (np.zeros(cap, dtype=np.float32) for _ in range(4))
            nbc = np.ascontiguousarray(nb_h)
            R.ref_init(8 << 30)
            # the reference's threshold is compiled in (PEAKTHRESHOLD): the
            # densities are scaled so that the same share passes it
            rk_s = (rk_h.astype(np.float64) * (O.PEAKTHRESHOLD / thr_k)).astype(np.float32)
            t0 = time.perf_counter()
            R.ref_finddenpeak(O._p(rk_s, O._fp), k, O._p(nbc, O._lp), n, O._p(x_, O._dp),
                              O._p(y_, O._dp), O._p(z_, O._dp), O._p(ty_h, O._ip), 0,
                              O._p(pk, O._ip), O._p(cx, O._fp), O._p(cy, O._fp),
                              O._p(cz, O._fp), O._p(cd, O._fp), cap)
            c["finddenpeak_reference_s"] = round(time.perf_counter() - t0, 4)
            t0 = time.perf_counter()
            O.ref_merging_peak(pos, gcore, rt_h[gcore])
            c["merging_peak_reference_s"] = round(time.perf_counter() - t0, 4)
            # the reference's FindCoreDensity on all threads, and with its
            # shells over the six upper thresholds (the lower ones take the
            # reference hours: every sweep scans the core array per peak)
            pk_h = peak.cpu().numpy()
            m_h = m64.cpu().numpy()
            t0 = time.perf_counter()
            O.ref_core_shells(rk_h, nb_h, pk_h, ty_h, m_h, nthreads=O.num_threads())
            c["find_core_density_reference_s"] = round(time.perf_counter() - t0, 4)
            t0 = time.perf_counter()
            O.ref_core_shells(rk_h, nb_h, pk_h, ty_h, m_h, thr=thr[:6],
                              nthreads=O.num_threads())
            c["find_core_density_shells_upper6_reference_s"] = round(
                time.perf_counter() - t0, 4)
        c["threads"] = O.num_threads()
        out["cpu"] = c
    return out


def host_api_bench(pos, mass, spec, reps=3):
    """The drop-in entry points with host buffers (the reference's own call
    sequence, level A of the integration notes) on the headline halo:
    assign_density_TSC -> gaussian_Smoothing -> findDen -> findsphdensity,
    each copying its inputs to the device and its outputs back (PCIe
    included), 88-byte particle records in FP64; and pgalf_lag_density_stage
    (level B: the same stage in one device residency, with the reference's
    own grid sizing).  Wall time per call, median of `reps`."""
    from galaxyfinder_amd import density as D
    n, k = len(pos), spec["k"]
    nx, ny, nz = spec["nx"], spec["ny"], spec["nz"]
    g0 = (spec["xmin"], spec["ymin"], spec["zmin"], spec["cellsize"])
    bp = D.make_particles(pos.astype(np.float64), mass.astype(np.float64))
    den = np.zeros(D.padded_shape(nx, ny, nz), dtype=np.float32)
    densph = np.zeros(n, dtype=np.float32)
    nbr = np.zeros((n, k), dtype=np.int64)
    rho = np.zeros(n, dtype=np.float32)
    calls = {
        "assign_density_TSC": lambda: D.assign_density_TSC(bp, n, den, nx, ny, nz, *g0, 5),
        "gaussian_Smoothing": lambda: D.gaussian_Smoothing(den, nx, ny, nz, spec["cellsize"],
                                                           spec["RG"]),
        "findDen": lambda: D.findDen(bp, n, densph, *g0, den, nx, ny, nz),
        "findsphdensity": lambda: D.findsphdensity(bp, n, nbr, k, rho),
        "pgalf_lag_density_stage": lambda: D.lag_density_stage(bp, k, 5),
    }
    ms = {}
    for name, fn in calls.items():
        fn()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append((time.perf_counter() - t0) * 1e3)
        ms[name] = round(float(np.median(ts)), 2)
    seq = sum(ms[a] for a in ("assign_density_TSC", "gaussian_Smoothing", "findDen",
                              "findsphdensity"))
    return {"workload": "one halo, N=%d, %d^3 grid, RG=%.1f cells, k=%d; host buffers "
                        "(88-byte records)" % (n, nx, spec["rg_cells"], k),
            "ms": ms, "sequence_ms": round(seq, 2),
            "sequence_particles_per_s": round(n / (seq * 1e-3), 1),
            "lag_density_stage_particles_per_s": round(
                n / (ms["pgalf_lag_density_stage"] * 1e-3), 1)}


def error_report():
    """Observed errors of the GPU stage against FP64 on the one-halo
    workloads: the deposit and the smoothed grid (normwise over the grid,
    pointwise over cells >= 1e-3 of the maximum), the read-back and the W4
    kNN densities (normwise and pointwise over all particles).  Two FP64
    references: the reference's own arithmetic carried in FP64 (FP32 TSC
    weights as tsc_omp2.c forms them, sums and transforms in FP64) and a
    stage computed in FP64 throughout (FP64 weights)."""
    import torch
    from galaxyfinder_amd import density as D, halos
    from oracle import oracle as O

    def err(got, ref, mask_rel=None):
        got = np.asarray(got, np.float64).ravel()
        ref = np.asarray(ref, np.float64).ravel()
        nrm = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
        m = np.abs(ref) >= (mask_rel or 0.0) * np.abs(ref).max()
        m &= ref != 0
        return {"normwise": nrm, "pointwise": float((np.abs(got - ref)[m] / np.abs(ref)[m]).max())}

    out = {}
    for cfg in (1, 4, 2):
        pos, mass, spec = halos.single_halo_config(cfg)
        n, nx, k = len(pos), spec["nx"], spec["k"]
        grid = (nx, spec["ny"], spec["nz"], spec["xmin"], spec["ymin"], spec["zmin"],
                spec["cellsize"])
        g = D.grid_struct(*grid)
        dev = torch.device("cuda")
        x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3))
        m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
        shape = D.padded_shape(nx, spec["ny"], spec["nz"])
        dep = torch.empty(shape, dtype=torch.float32, device=dev)
        D.dev_tsc(x, y, z, m, g, dep)
        dep_h = dep.cpu().numpy()[..., :nx].astype(np.float64)
        den = torch.empty(shape, dtype=torch.float32, device=dev)
        rt = torch.empty(n, dtype=torch.float32, device=dev)
        nbr = torch.empty((n, k), dtype=torch.int64, device=dev)
        rk = torch.empty(n, dtype=torch.float32, device=dev)
        D.dev_density_stage(x, y, z, m, g, spec["RG"], k, den, rt, nbr, rk)
        D.dev_status()
        den_h = den.cpu().numpy()[..., :nx].astype(np.float64)
        rt_h, rk_h, idx = rt.cpu().numpy(), rk.cpu().numpy(), nbr.cpu().numpy()
        pos64 = pos.astype(np.float64)
        # the reference's arithmetic in FP64: its FP32 weights, FP64 sums
        dref, _ = O.tsc_deposit(pos64, mass, grid, fp64=True)
        dref = np.ascontiguousarray(dref[..., :nx])
        sref = O.gsmooth_f64(dref, spec["cellsize"], spec["RG"])
        # FP64 throughout
        rt64, rk64, F64 = O.stage_forward_d(pos64, mass, grid, spec["RG"], idx, weighted=False)
        out["halo_%d" % n] = {
            "deposit_vs_fp64_sums": err(dep_h, dref, 1e-3),
            "deposit_vs_fp64_weights": err(dep_h, O.tsc_deposit_d(pos64, mass, grid), 1e-3),
            "smoothed_vs_fp64_sums": err(den_h, sref, 1e-3),
            "smoothed_vs_fp64_stage": err(den_h, F64, 1e-3),
            "readback_vs_fp64_stage": err(rt_h, rt64),
            "knn_density_vs_fp64": err(rk_h, rk64)}
        print("[bench] errors: %d-particle halo done" % n, file=sys.stderr, flush=True)
    return out


def catalog_bench(nhalo=2000, seed=3, cpu=True):
    """The catalogue path on a synthetic FoF catalogue written to a scratch
    directory: nhalo halos of halos.catalogue_vec (dN/dn ~ n^-2 on
    [100, 65536]) of DM (1e8 Msun/h) and one star in five (1e6 Msun/h), each
    scaled to 0.2 cMpc/h across in a 100 cMpc/h box.  Times, on the host,
    the loader (read, Dump2BP, mklocalize, grid sizing) and the writer
    (every particle in subhalo 0); on the GPU, the batched density stage over
    the loaded catalogue with its lagFindStellarCore grids.  CPU: the
    reference's own FREAD + Dump2BP + mklocalize and write_data (oracle/_ref)
    on the same files."""
    import ctypes
    import shutil
    import tempfile
    import torch
    from galaxyfinder_amd import catalog as CAT, density as D, halos
    pos, mass, off, _, _ = halos.catalogue_vec(seed, nhalo=nhalo, alpha=2.0, nmin=100,
                                                nmax=65536)
    rng = np.random.default_rng(seed)
    d = tempfile.mkdtemp(prefix="pgalf_cat_")
    try:
        centres = rng.uniform(5.0, 95.0, (nhalo, 3))
        recs = []
        for h in range(nhalo):
            a, e = int(off[h]), int(off[h + 1])
            r = np.zeros(e - a, dtype=CAT.PM_DTYPE)
            p = (pos[a:e].astype(np.float64) - 0.5) * 0.2 + centres[h]
            r["x"], r["y"], r["z"] = p[:, 0], p[:, 1], p[:, 2]
            r["mass"] = 1.0e8
            star = np.arange(e - a) % 5 == 0
            r["mass"][star] = 1.0e6
            recs.append((r[~star], r[:0], np.zeros(0, dtype=np.uint8), r[star]))
        hp, mp = os.path.join(d, "FoF_halo_cat.00001"), os.path.join(d, "FoF_member_particle.00001")
        CAT.write_fof_catalog(hp, mp, [100.0, 0.7, 0.3, 0.048, 0.7, 1.0, 1.0], recs)
        mbytes = os.path.getsize(mp)
        t0 = time.perf_counter()
        cat = CAT.Catalog(hp, mp)
        b = cat.next_batch(copy_records=False)
        t_load = time.perf_counter() - t0
        n = int(b["offsets"][-1])
        grids = CAT.Catalog.grid_tuples(b)
        dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()  # noqa: E731
        x, y, z = (dev(b[k].astype(np.float32)) for k in "xyz")
        m = dev(b["mass"].astype(np.float32))
        bt = D.Batch(b["offsets"], grids, CAT.GAUSSIAN_SMOOTHING_LENGTH, 32)
        den = torch.empty(bt.grid_floats, dtype=torch.float32, device="cuda")
        rt = torch.empty(n, dtype=torch.float32, device="cuda")
        nbr = torch.empty((n, 32), dtype=torch.int64, device="cuda")
        rk = torch.empty(n, dtype=torch.float32, device="cuda")
        bt.run(x, y, z, m, den, rt, nbr, rk)
        bt.status()
        reps = 3
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            bt.run(x, y, z, m, den, rt, nbr, rk)
        torch.cuda.synchronize()
        t_gpu = (time.perf_counter() - t0) / reps
        bt.status()
        bt.close()
        outs = [os.path.join(d, f) for f in ("GALFIND.DATA", "GALCATALOG.LIST", "background")]
        t0 = time.perf_counter()
        w = CAT.Writer(*outs)
        offs, roffs = b["offsets"], b["record_offsets"]
        for h in range(len(offs) - 1):
            a, e = offs[h], offs[h + 1]
            w.halo(b["type"][a:e], b["records"][roffs[h]:roffs[h + 1]],
                   np.zeros(e - a, dtype=np.int32), 1)
        w.close()
        t_write = time.perf_counter() - t0
        cat.close()
        out = {"workload": "%d halos, %d particles (dN/dn ~ n^-2 on [100, 65536]), member "
                           "file %.1f MB, lagFindStellarCore grids, RG = 0.008, k = 32"
                           % (nhalo, n, mbytes / 1e6),
               "halos": len(offs) - 1, "particles": n,
               "grid_cells": int(sum(g[0] * g[1] * g[2] for g in grids)),
               "load_s": round(t_load, 4), "load_mb_per_s": round(mbytes / 1e6 / t_load, 1),
               "gpu_stage_ms": round(t_gpu * 1e3, 3),
               "gpu_particles_per_s": round(n / t_gpu, 1),
               "write_s": round(t_write, 4)}
        if cpu:
            from oracle import oracle as O
            R = O.ref_lib()
            if R is not None:
                R.ref_init(8 << 30)
                xs, ys, zs, ms = (np.zeros(n) for _ in range(4))
                ty = np.zeros(n, dtype=np.int32)
                l2 = np.zeros(n, dtype=np.float32)
                ro = np.zeros(nhalo + 2, dtype=np.int64)
                org = np.zeros(3 * (nhalo + 2))
                pp = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
                R.ref_read_catalog.restype = ctypes.c_longlong
                t0 = time.perf_counter()
                R.ref_read_catalog(hp.encode(), mp.encode(), nhalo + 1, ctypes.c_longlong(n),
                                   pp(xs), pp(ys), pp(zs), pp(ms), pp(ty), pp(l2), pp(ro),
                                   pp(org))
                out["load_reference_s"] = round(time.perf_counter() - t0, 4)
                p2 = np.zeros(n, dtype=np.int32)
                mpk = np.ones(nhalo, dtype=np.int32)
                refo = [f + ".ref" for f in outs]
                R.ref_write_catalog.restype = ctypes.c_longlong
                t0 = time.perf_counter()
                R.ref_write_catalog(hp.encode(), mp.encode(), pp(p2), pp(mpk),
                                    *(f.encode() for f in refo))
                out["write_reference_s"] = round(time.perf_counter() - t0, 4)
                out["files_identical_to_reference"] = all(
                    open(a_, "rb").read() == open(b_, "rb").read() for a_, b_ in zip(outs, refo))
        return out
    finally:
        shutil.rmtree(d, ignore_errors=True)


def batch_bench(steps, warmup, plain_pass=True, watershed=True, cpu=True):
    """Batched forward over the 10000-halo catalogue (dN/dn ~ n^-2 on
    [100, 1048576], grids 32^3-256^3 from n, RG = 2 cells, k = 32): eager
    and HIP-graph runs of the batch, and one pass of the single-halo stage
    called once per halo on the same device-resident particles."""
    import torch
    from galaxyfinder_amd import density as D
    from galaxyfinder_amd import halos

    t0 = time.perf_counter()
    pos, mass, off, grids, rg = halos.catalogue(config_id=3)
    gen_s = time.perf_counter() - t0
    n, nh, k = len(pos), len(off) - 1, 32
    print("[bench] batch workload: %d halos, %d particles (generated in %.1f s)"
          % (nh, n, gen_s), file=sys.stderr, flush=True)
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3))
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    b = D.Batch(off, grids, rg, k)
    den = torch.empty(b.grid_floats, dtype=torch.float32, device=dev)
    rho_t = torch.empty(n, dtype=torch.float32, device=dev)
    nbr = torch.empty((n, k), dtype=torch.int64, device=dev)
    rho_k = torch.empty(n, dtype=torch.float32, device=dev)

    def timed(graph):
        for _ in range(warmup):
            b.run(x, y, z, m, den, rho_t, nbr, rho_k, graph=graph)
        b.status()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run(x, y, z, m, den, rho_t, nbr, rho_k, graph=graph)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b.status()
        return (t1 - t0) * 1e3 / steps

    eager_ms = timed(False)
    graph_ms = timed(True)
    D.timing(enable=True, reset=True)
    for _ in range(steps):
        b.run(x, y, z, m, den, rho_t, nbr, rho_k, graph=False)
    torch.cuda.synchronize()
    phases = {}
    for ph in ("batch_stage", "batch_tsc", "batch_smooth", "batch_findden", "batch_knn",
               "knn_prep", "knn_query"):
        tot, cnt = D.timing_get(ph)
        if cnt:
            phases[ph] = round(tot / cnt, 4)
    D.timing(enable=False, reset=True)
    print("[bench] batch: eager %.3f ms, graph %.3f ms" % (eager_ms, graph_ms),
          file=sys.stderr, flush=True)
    out = {
        "workload": "%d halos, %d particles, dN/dn ~ n^-2 on [100, 1048576], grids "
                    "32^3-256^3 (pow2 >= 4 n^1/3), RG = 2 cells, k = %d" % (nh, n, k),
        "halos": nh, "particles": n,
        "value": round(n / (min(eager_ms, graph_ms) * 1e-3), 1),
        "unit": "particles/s",
        "ms_per_step_eager": round(eager_ms, 4),
        "ms_per_step_graph": round(graph_ms, 4),
        "phase_ms": phases,
    }
    if plain_pass:
        # the single-halo stage called once per halo (views into the same
        # device arrays; one pass, after a few warm-up halos)
        views = []
        for h in range(nh):
            a, e = int(off[h]), int(off[h + 1])
            g = D.grid_struct(*grids[h])
            views.append((x[a:e], y[a:e], z[a:e], m[a:e], g, float(rg[h]),
                          b.grid_view(den, h), rho_t[a:e], nbr[a:e], rho_k[a:e]))
        for v in views[:20]:
            D.dev_density_stage(v[0], v[1], v[2], v[3], v[4], v[5], k, v[6], v[7], v[8], v[9])
        D.dev_status()
        t0 = time.perf_counter()
        for v in views:
            D.dev_density_stage(v[0], v[1], v[2], v[3], v[4], v[5], k, v[6], v[7], v[8], v[9])
        torch.cuda.synchronize()
        plain_ms = (time.perf_counter() - t0) * 1e3
        D.dev_status()
        out["plain_per_halo_ms"] = round(plain_ms, 2)
        out["plain_per_halo_value"] = round(n / (plain_ms * 1e-3), 1)
        out["speedup_vs_per_halo"] = round(plain_ms / min(eager_ms, graph_ms), 2)
        print("[bench] per-halo plain pass: %.1f ms (%.1fx)" % (plain_ms, out["speedup_vs_per_halo"]),
              file=sys.stderr, flush=True)
    if watershed:
        out["watershed"] = catalog_watershed_bench(off, rho_k, nbr, m, cpu=cpu)
    b.close()
    return out


def catalog_watershed_bench(off, rho_k, nbr, m, cpu=True, stride=10, ref_max=20000):
    """The watershed of subhalo_den (FindCoreDensity, the shell thresholds,
    GetShellParticlesPeaks + SaveRemainingParticles2LastShell) of every halo
    of the batched catalogue in one call (pgalf_dev_catalog_watershed), on
    the batch's own W4 kNN densities and 32-NN lists.  Cores: each halo's
    finddenpeak cores above its median density (device finddenpeak on the
    densities divided by the halo's median, threshold 1); every 4th particle
    a star; FP64 masses.
      * one_call_ms: the catalogue call (wall time around the synchronised
        call, median of 3);
      * per_halo: the one-halo device entry points (dev_find_core_density,
        dev_shell_thresholds, dev_shell_particles_peaks) called halo by halo
        on every `stride`-th halo in size order, and that time times
        `stride` as the estimate for the catalogue;
      * reference: the halos of that sample with at most `ref_max`
        particles, gathered into one sub-catalogue: the catalogue call on
        it, the reference's own functions (oracle/_ref; FindCoreDensity on
        all threads, the shells serial) halo by halo, and whether every
        output is equal."""
    import torch
    from galaxyfinder_amd import density as D, halos
    off = np.asarray(off, np.int64)
    nh, n = len(off) - 1, int(off[-1])
    dev = rho_k.device
    cnt = np.diff(off)
    coff, peaks, types = halos.catalogue_cores(off, rho_k, nbr)
    peak_t = torch.from_numpy(np.concatenate(peaks)).to(dev)
    m64 = m.double()

    def wall(fn, reps=3):
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            ts.append((time.perf_counter() - t0) * 1e3)
        return float(np.median(ts))

    res = {}

    def cat_call():
        res["o"] = D.dev_catalog_watershed(off, coff, rho_k, nbr, peak_t, types, m64)
    t_cat = wall(cat_call)
    o = res["o"]
    assert not o["status"].any(), "catalogue watershed: halo status %s" % (
        np.unique(o["status"]),)
    out = {"workload": "the batched catalogue's %d halos, %d particles; finddenpeak cores "
                       "above each halo's median W4 density (%d cores); every 4th particle "
                       "a star; FP64 masses" % (nh, n, int(coff[-1])),
           "cores": int(coff[-1]), "shells": int(o["nshell"].sum()),
           "dropped_cores": int((o["nummem"][:int(coff[-1])] == 0).sum().item()),
           "one_call_ms": round(t_cat, 2)}
    print("[bench] catalogue watershed: %d cores, %d shells, %.1f ms" % (
        out["cores"], out["shells"], t_cat), file=sys.stderr, flush=True)
    # the one-halo path on every stride-th halo in size order
    sample = np.argsort(cnt, kind="stable")[::stride]

    def one(h):
        a, e = int(off[h]), int(off[h + 1])
        ca, ce = int(coff[h]), int(coff[h + 1])
        den, nb, p = rho_k[a:e], nbr[a:e], peak_t[ca:ce]
        oh = D.dev_find_core_density(den, nb, p, types[a:e], m64[a:e])
        thr = D.dev_shell_thresholds(den, oh["coredensity"])
        D.dev_shell_particles_peaks(den, nb, p, oh["coredensity"], thr, oh["flag"],
                                    oh["haloid"])
    for h in sample[:20]:
        one(h)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for h in sample:
        one(h)
    torch.cuda.synchronize()
    t_one = (time.perf_counter() - t0) * 1e3
    out["per_halo"] = {"sample_halos": len(sample),
                       "sample_particles": int(cnt[sample].sum()),
                       "sample_ms": round(t_one, 1),
                       "estimate_ms": round(t_one * stride, 1),
                       "one_call_speedup": round(t_one * stride / t_cat, 2)}
    print("[bench] one-halo path on %d halos: %.1f ms (x%d = %.0f ms)" % (
        len(sample), t_one, stride, t_one * stride), file=sys.stderr, flush=True)
    if cpu:
        from oracle import oracle as O
        if O.ref_lib() is not None:
            sub = [int(h) for h in sorted(sample) if cnt[h] <= ref_max]
            idx = np.concatenate([np.arange(off[h], off[h + 1]) for h in sub])
            soff = np.zeros(len(sub) + 1, np.int64)
            soff[1:] = np.cumsum(cnt[sub])
            scoff = np.zeros(len(sub) + 1, np.int64)
            scoff[1:] = np.cumsum([len(peaks[h]) for h in sub])
            it = torch.from_numpy(idx).to(dev)
            s_den, s_nbr = rho_k[it].contiguous(), nbr[it].contiguous()
            s_ty, s_m = types[it].contiguous(), m64[it].contiguous()
            s_pk = torch.from_numpy(np.concatenate([peaks[h] for h in sub])).to(dev)

            def sub_call():
                res["s"] = D.dev_catalog_watershed(soff, scoff, s_den, s_nbr, s_pk, s_ty, s_m)
            t_sub = wall(sub_call)
            so = {k_: (v.cpu().numpy() if hasattr(v, "cpu") else v) for k_, v in res["s"].items()}
            den_h, nb_h = s_den.cpu().numpy(), s_nbr.cpu().numpy()
            ty_h, m_h = s_ty.cpu().numpy(), s_m.cpu().numpy()
            same = True
            t_ref = 0.0
            for j in range(len(sub)):
                a, e = int(soff[j]), int(soff[j + 1])
                ca, ce = int(scoff[j]), int(scoff[j + 1])
                p = peaks[sub[j]]
                # thresholds from an untimed first FindCoreDensity (the
                # reference forms them inline in subhalo_den)
                r, _ = O.ref_core_shells(den_h[a:e], nb_h[a:e], p, ty_h[a:e], m_h[a:e],
                                         nthreads=O.num_threads())
                thr = O.shell_thresholds(den_h[a:e], r["coredensity"])
                t0 = time.perf_counter()
                r, rsh = O.ref_core_shells(den_h[a:e], nb_h[a:e], p, ty_h[a:e], m_h[a:e],
                                           thr=thr, nthreads=O.num_threads())
                t_ref += time.perf_counter() - t0
                for key in ("coredensity", "nummem", "numstar", "starmass"):
                    same &= np.array_equal(np.ascontiguousarray(so[key][ca:ce]).view(np.int32),
                                           np.ascontiguousarray(r[key]).view(np.int32))
                same &= np.array_equal(so["flag"][a:e], r["flag"])
                same &= np.array_equal(so["haloid"][a:e], r["haloid"])
                gs = D.catalog_shells_of(so, soff, j)
                same &= len(gs) == len(rsh) and all(
                    np.array_equal(g0, r0) and np.array_equal(g1, r1)
                    for (g0, g1), (r0, r1) in zip(gs, rsh))
            out["reference"] = {
                "sample": "%d halos of the size-ordered sample with <= %d particles, %d "
                          "particles, %d cores" % (len(sub), ref_max, len(idx), int(scoff[-1])),
                "gpu_one_call_ms": round(t_sub, 2),
                "reference_s": round(t_ref, 3),
                "reference_threads": O.num_threads(),
                "equal": bool(same)}
            print("[bench] reference watershed sample: %.2f s (GPU one call %.1f ms), equal=%s"
                  % (t_ref, t_sub, same), file=sys.stderr, flush=True)
    return out


def grad_bench(steps, warmup):
    """Forward and forward+backward of the differentiable stage on one halo
    of 1048576 particles (NFW c=8 + 10 subhalos), 128^3 grid, RG = 2 cells,
    k = 32, mass-weighted W4 density; loss = sum w_t log rho_tsc + sum w_k
    log rho_knn with seeded Gaussian weights."""
    import torch
    from galaxyfinder_amd import density as D
    from galaxyfinder_amd import halos
    from galaxyfinder_amd.autograd import density_stage

    pos, mass, spec = halos.single_halo_config(4)
    n, k = len(pos), spec["k"]
    dev = torch.device("cuda")
    xyz = [torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3)]
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    g = D.grid_struct(spec["nx"], spec["ny"], spec["nz"], spec["xmin"], spec["ymin"],
                      spec["zmin"], spec["cellsize"])
    rng = np.random.default_rng(20261203 + 4)
    wt = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)
    wk = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)

    def fwd():
        with torch.no_grad():
            return density_stage(xyz[0], xyz[1], xyz[2], m, g, spec["RG"], k, True)

    def fwd_bwd():
        leaves = [t.detach().requires_grad_() for t in xyz + [m]]
        rt, rk, _, _ = density_stage(leaves[0], leaves[1], leaves[2], leaves[3], g,
                                     spec["RG"], k, True)
        loss = (wt * torch.log(rt)).sum() + (wk * torch.log(rk)).sum()
        loss.backward()
        return leaves

    def timed(fn):
        for _ in range(warmup):
            fn()
        D.dev_status()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        D.dev_status()
        return (t1 - t0) * 1e3 / steps

    f_ms = timed(fwd)
    fb_ms = timed(fwd_bwd)
    D.timing(enable=True, reset=True)
    for _ in range(steps):
        fwd_bwd()
    torch.cuda.synchronize()
    phases = {}
    for ph in ("bwd", "bwd_tsc", "bwd_knn"):
        tot, cnt = D.timing_get(ph)
        if cnt:
            phases[ph] = round(tot / cnt, 4)
    D.timing(enable=False, reset=True)
    print("[bench] grad: fwd %.3f ms, fwd+bwd %.3f ms" % (f_ms, fb_ms), file=sys.stderr,
          flush=True)
    return {
        "workload": "one halo, N=%d, %d^3 grid, RG=%.1f cells, k=%d, mass-weighted W4; "
                    "loss sum w log rho over both densities" % (n, spec["nx"], spec["rg_cells"], k),
        "particles": n,
        "fwd_ms": round(f_ms, 4),
        "fwd_bwd_ms": round(fb_ms, 4),
        "fwd_bwd_over_fwd": round(fb_ms / f_ms, 3),
        "value": round(n / (fb_ms * 1e-3), 1),
        "unit": "particles/s (fwd+bwd)",
        "phase_ms": phases,
    }


def c5_bench(steps, warmup, check_halos=24):
    """Batched forward+backward with peak climbing over 50000 halos (dN/dn ~
    n^-1.9 on [64, 2097152], grids 32^3-256^3, RG = 2 cells, k = 32,
    mass-weighted W4; loss sum w_t log rho_tsc + sum w_k log rho_knn; 5
    normalised gradient-ascent steps of 0.5 cell), eager and from one HIP
    graph; membership (peak cells) compared with the single-halo path on the
    largest halos and a seeded sample."""
    import torch
    from galaxyfinder_amd import density as D
    from galaxyfinder_amd import halos

    t0 = time.perf_counter()
    pos, mass, off, grids, rg = halos.catalogue_vec(5)
    gen_s = time.perf_counter() - t0
    n, nh, k = len(pos), len(off) - 1, 32
    print("[bench] c5 workload: %d halos, %d particles (generated in %.1f s)" % (nh, n, gen_s),
          file=sys.stderr, flush=True)
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3))
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    del pos
    rng = np.random.default_rng(20261203 + 5)
    wt = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)
    wk = torch.from_numpy(rng.normal(size=n).astype(np.float32)).to(dev)
    b = D.Batch(off, grids, rg, k)
    out = b.run_grad(x, y, z, m, wt, wk, rho_mode=2, climb_iters=5, climb_step=0.5)
    b.status()

    def timed(fn):
        for _ in range(warmup):
            fn()
        b.status()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        b.status()
        return (t1 - t0) * 1e3 / steps

    fwd_ms = timed(lambda: b.run(x, y, z, m, out["den"], out["rho_tsc"], out["nbr"],
                                 out["rho_knn"]))
    full_ms = timed(lambda: b.run_grad(x, y, z, m, wt, wk, rho_mode=2, climb_iters=5,
                                       climb_step=0.5, out=out))
    graph_ms = timed(lambda: b.run_grad(x, y, z, m, wt, wk, rho_mode=2, climb_iters=5,
                                        climb_step=0.5, graph=True, out=out))
    print("[bench] c5: fwd %.1f ms, fwd+bwd+climb %.1f ms (graph %.1f ms)"
          % (fwd_ms, full_ms, graph_ms), file=sys.stderr, flush=True)
    # membership against the single-halo path
    counts = np.diff(off)
    pick = list(np.argsort(counts)[-4:]) + list(
        np.random.default_rng(5).choice(nh, check_halos - 4, replace=False))
    peak = out["peak"].cpu().numpy()
    same = True
    for h in pick:
        a, e = int(off[h]), int(off[h + 1])
        g = D.grid_struct(*grids[h])
        den = torch.empty(D.padded_shape(grids[h][0], grids[h][1], grids[h][2]),
                          dtype=torch.float32, device=dev)
        D.dev_tsc(x[a:e], y[a:e], z[a:e], m[a:e], g, den)
        D.dev_smooth(den, g, float(rg[h]))
        _, _, _, pk = D.dev_peak_climb(x[a:e], y[a:e], z[a:e], g, den, 5, 0.5)
        same &= bool(np.array_equal(pk.cpu().numpy(), peak[a:e]))
    D.dev_status()
    b.close()
    return {
        "workload": "%d halos, %d particles, dN/dn ~ n^-1.9 on [64, 2097152], grids 32^3-256^3,"
                    " RG = 2 cells, k = %d, mass-weighted W4, 5 climbing steps of 0.5 cell"
                    % (nh, n, k),
        "halos": nh, "particles": n,
        "fwd_ms": round(fwd_ms, 3),
        "fwd_bwd_climb_ms": round(full_ms, 3),
        "fwd_bwd_climb_graph_ms": round(graph_ms, 3),
        "fwd_bwd_climb_over_fwd": round(min(full_ms, graph_ms) / fwd_ms, 3),
        "value": round(n / (min(full_ms, graph_ms) * 1e-3), 1),
        "unit": "particles/s (fwd+bwd+climb)",
        "membership_identical_to_single_halo_path": same,
        "membership_halos_checked": len(pick),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", type=int, default=2, choices=[1, 2, 4])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--knn-stride", type=int, default=16,
                    help="CPU baseline: run the reference list builder on every n-th cell")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched 10000-halo catalogue measurement")
    ap.add_argument("--batch-only", action="store_true",
                    help="only the batched catalogue measurement (no per-halo pass)")
    ap.add_argument("--catalog-only", action="store_true",
                    help="only the FoF catalogue measurement")
    ap.add_argument("--no-catalog", action="store_true",
                    help="skip the FoF catalogue load / batch / write measurement")
    ap.add_argument("--errors", action="store_true",
                    help="observed errors of the stage's outputs against the FP64 oracle on "
                         "the 131072-, 4194304- and 1048576-particle halos (one JSON line)")
    ap.add_argument("--host-api", action="store_true",
                    help="only the drop-in entry points with host buffers (PCIe-inclusive "
                         "times) on the headline halo")
    ap.add_argument("--no-host-api", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) measurement")
    ap.add_argument("--peaks-only", action="store_true",
                    help="one density stage on the headline halo, then only the peak stage "
                         "and the watershed (the 'peaks' object; for profiling)")
    ap.add_argument("--no-peaks", action="store_true",
                    help="skip the integer peak stage measurement")
    ap.add_argument("--no-grad", action="store_true",
                    help="skip the forward+backward measurement")
    ap.add_argument("--c5", action="store_true",
                    help="only the 50000-halo batched fwd+bwd with peak climbing (about 20 s "
                         "of catalogue generation on the host)")
    ap.add_argument("--no-c5", action="store_true",
                    help="skip the 50000-halo batched fwd+bwd with peak climbing")
    ap.add_argument("--calib-copy", action="store_true",
                    help="also run a plain %d-byte device copy (counter calibration)"
                    % CALIB_COPY_BYTES)
    ap.add_argument("--schedule", default=None,
                    help="launch schedule (JSON of python -m galaxyfinder_amd.tune); default: "
                         "%s when present, else the library defaults" % TUNED_SCHEDULE)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build runs the density stage on a "
              "single MI355X (halos are independent; multi-GPU sharding is a host-side "
              "split of the halo catalogue, not run here).", file=sys.stderr)
        sys.exit(2)

    import torch
    from galaxyfinder_amd import density as D
    from galaxyfinder_amd import halos

    sched_path = args.schedule or (TUNED_SCHEDULE if os.path.exists(TUNED_SCHEDULE) else None)
    if sched_path:
        from galaxyfinder_amd import tune
        tune.load(sched_path)
        print("[bench] launch schedule %s" % sched_path, file=sys.stderr, flush=True)
    if args.catalog_only:
        print(json.dumps({"catalog": catalog_bench(cpu=not args.no_cpu_baseline)}), flush=True)
        return
    if args.c5:
        print(json.dumps({"c5": c5_bench(args.steps, args.warmup)}), flush=True)
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_bench(args.steps, args.warmup, plain_pass=False,
                                               cpu=not args.no_cpu_baseline)}),
              flush=True)
        return
    if args.errors:
        print(json.dumps({"errors": error_report()}), flush=True)
        return
    pos, mass, spec = halos.single_halo_config(args.config)
    n, k = len(pos), spec["k"]
    if args.host_api:
        print(json.dumps({"host_api": host_api_bench(pos, mass, spec)}), flush=True)
        return
    print("[bench] workload config %d: N=%d grid=%d^3 RG=%.2f cells k=%d"
          % (args.config, n, spec["nx"], spec["rg_cells"], k), file=sys.stderr, flush=True)
    dev = torch.device("cuda")
    x, y, z = (torch.from_numpy(np.ascontiguousarray(pos[:, a])).to(dev) for a in range(3))
    m = torch.from_numpy(np.ascontiguousarray(mass)).to(dev)
    g = D.grid_struct(spec["nx"], spec["ny"], spec["nz"], spec["xmin"], spec["ymin"],
                      spec["zmin"], spec["cellsize"])
    den = torch.empty(D.padded_shape(spec["nx"], spec["ny"], spec["nz"]), dtype=torch.float32,
                      device=dev)
    rho_t = torch.empty(n, dtype=torch.float32, device=dev)
    nbr = torch.empty((n, k), dtype=torch.int64, device=dev)
    rho_k = torch.empty(n, dtype=torch.float32, device=dev)

    def step():
        D.dev_density_stage(x, y, z, m, g, spec["RG"], k, den, rho_t, nbr, rho_k)

    if args.peaks_only:
        step()
        D.dev_status()
        print(json.dumps({"peaks": peaks_bench(x, y, z, m, g, den, rho_t, nbr, rho_k, spec,
                                               reps=3, cpu=not args.no_cpu_baseline)}),
              flush=True)
        return
    for _ in range(args.warmup):
        step()
    D.dev_status()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    D.dev_status()
    ms_step = (t1 - t0) * 1e3 / args.steps
    value = n * args.steps / (t1 - t0)
    print("[bench] timed %d steps: %.3f ms/step" % (args.steps, ms_step), file=sys.stderr,
          flush=True)

    # per-phase HIP-event timing on the launch stream (separate pass)
    D.timing(enable=True, reset=True)
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    phases = {}
    for ph in ("stage", "tsc", "smooth", "findden", "knn", "knn_prep", "knn_query",
               "knn_packed"):
        tot, cnt = D.timing_get(ph)
        if cnt:
            phases[ph] = tot / cnt
    D.timing(enable=False, reset=True)
    if args.calib_copy:
        src = torch.ones(CALIB_COPY_BYTES // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(2):
            torch.neg(src, out=dst)  # elementwise kernel: reads and writes CALIB_COPY_BYTES
        torch.cuda.synchronize()
        del src, dst
    g_cells = spec["nx"] * spec["ny"] * spec["nz"]
    bts = stage_bytes(n, g_cells, k)
    q_ms = phases.get("knn_packed", float("nan"))
    achieved = bts["knn"] / (q_ms * 1e-3) / 1e9
    valu = measured_valu(n, k)
    roofline = {
        "bound": "hbm",
        "kernel": "knn_packed_k",
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(n, k),
        "bytes_per_launch": bts["knn"],
        "ms_per_launch": round(q_ms, 4),
        "valu": None if valu is None else {
            "insts_per_launch": valu,
            "peak_insts_per_s": VALU_ISSUE_PER_S,
            "frac": round(valu / VALU_ISSUE_PER_S / (q_ms * 1e-3), 4)},
    }
    stage_alg = sum(bts.values())
    out = {
        "metric": "particles/s through density stage (deposit+smooth+read-back+kNN lists+"
                  "W4 density), one halo",
        "value": round(value, 1),
        "unit": "particles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (grid, positions) / f64 (kNN distances)",
        "data": "synthetic (seeded NFW+Plummer halo)",
        "config": {"workload": "one halo, N=%d, %d^3 grid, RG=%.1f cells, k=%d"
                               % (n, spec["nx"], spec["rg_cells"], k),
                   "particles": n, "grid": [spec["nx"], spec["ny"], spec["nz"]], "k": k,
                   "schedule": os.path.relpath(sched_path, ROOT) if sched_path else "defaults"},
        "phase_ms": {k_: round(v, 4) for k_, v in phases.items()},
        "stage_roofline_frac": round(stage_alg / (ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "roofline": roofline,
    }
    if not args.no_host_api:
        out["host_api"] = host_api_bench(pos, mass, spec)
    if not args.no_peaks:
        out["peaks"] = peaks_bench(x, y, z, m, g, den, rho_t, nbr, rho_k, spec,
                                   cpu=not args.no_cpu_baseline)
    if not args.no_catalog:
        out["catalog"] = catalog_bench(cpu=not args.no_cpu_baseline)
    if not args.no_batch:
        out["batch"] = batch_bench(max(3, args.steps // 2), 2, cpu=not args.no_cpu_baseline)
    if not args.no_grad:
        out["grad"] = grad_bench(max(3, args.steps // 2), 2)
    if not args.no_c5:
        out["c5"] = c5_bench(max(3, args.steps // 2), 2)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(pos.astype(np.float64), mass.astype(np.float64),
                                           spec, k, args.knn_stride)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
