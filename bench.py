"""Benchmark of the dictionary-generation + matching path on one MI355X.

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Workload (the full-size plain path): EPG-SSFP dictionary of the grid
T1 = 100:20:2000+2300:300:5000 ms, T2 = 20:5:100+110:10:300+400:200:2000 ms,
B0 = -40:2:40 Hz, B1 = 0.8:0.1:1.2 (871,455 atoms in 5 B1 slices), T = 1000
seeded synthetic timepoints, w = 30 states, complex64, normalised; then FP32
matching of 240 x 240 voxels (PD ~ U(0.3, 1), SNR 25, smooth B1 map, each
voxel matched against the slice its B1 map selects).

A step is one dictionary generation (value: atoms/s); the matcher is timed
over the same number of steps (match_voxels_per_s).  Inputs are resident in
HBM before the timed regions; timing uses HIP events on the stream the
kernels run on.  Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "atoms/s gen; voxels/s matched, BF16-MFMA vs FP32 at ≥99.9% argmax agree"
PEAK_FP32 = 157.3          # TFLOP/s, MI355X FP32 (vector packed / matrix), datasheet
FLOP_PER_STATE_SSFP = 102  # algorithmic flops per (atom, state, TR)
GRID = ("100:20:2000+2300:300:5000", "20:5:100+110:10:300+400:200:2000", "-40:2:40", "0.8:0.1:1.2")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(seq_raw, params, natoms, T, w, rows_idx, gpu_rows, slice_atoms, slice_sig,
                 gpu_idx, budget_s=10.0):
    """FP64 oracle (C, OpenMP over all host cores) on a bounded seeded sample
    of the same workload.  Generation: the atoms rows_idx of B1 slice 0
    (whose GPU rows are gpu_rows, compared with the oracle's), repeated until
    about budget_s of CPU work.  Matching: slice_sig voxels against every atom
    of slice 0 (slice_atoms, the GPU's complex64 dictionary rows), compared
    with the GPU's picks gpu_idx."""
    import numpy as np
    import oracle
    seq = oracle.seq_prepare("epg_ssfp", seq_raw)
    sub = params[rows_idx]
    n = len(rows_idx)
    t0 = oracle.wall()
    atoms = oracle.epg("epg_ssfp", seq, sub, n, w)
    done, dt_gen = n, oracle.wall() - t0
    rel = float(np.max(np.abs(gpu_rows - atoms).max(axis=1) / np.abs(atoms).max(axis=1)))
    del atoms
    reps = 0
    while dt_gen < budget_s and reps < 64:
        t0 = oracle.wall()
        oracle.epg("epg_ssfp", seq, sub, n, w)
        dt_gen += oracle.wall() - t0
        done += n
        reps += 1
    match = {}
    if slice_sig is not None:
        t0 = oracle.wall()
        ridx, _, _ = oracle.match_slice(slice_atoms, slice_sig)
        dt_m = oracle.wall() - t0
        match = dict(match_value=len(slice_sig) / dt_m, match_unit="voxels/s",
                     match_argmax_equal_gpu=float(np.mean(ridx == gpu_idx)),
                     match_sample=f"{len(slice_sig)} voxels of B1 slice 0 against all "
                                  f"{len(slice_atoms)} atoms of the slice ({dt_m:.1f} s)")
    cpu_model = ""
    try:
        with open("/proc/cpuinfo") as f:
            cpu_model = next((l.split(":", 1)[1].strip() for l in f if l.startswith("model name")), "")
    except OSError:
        pass
    return dict(value=done / dt_gen, unit="atoms/s", cores=oracle.num_threads(), kind="port",
                cpu_model=cpu_model, nproc=os.cpu_count(), affinity=len(os.sched_getaffinity(0)),
                sample=f"epg_ssfp FP64 oracle: {n} seeded atoms of B1 slice 0 x T={T} x w={w}, "
                       f"{done} atoms in {dt_gen:.1f} s",
                gpu_atoms_max_rel_err=rel, gpu_atoms_checked=n, **match)


def make_signals(torch, np, synth, d_atoms, params, natoms, T, nx=240, ny=240, snr=25.0, seed=0):
    """240x240 voxels: atom of the slice nearest the smooth B1 map value,
    PD ~ U(0.3, 1), complex Gaussian noise at the given SNR (on the GPU)."""
    V = nx * ny
    b1 = synth.b1_map(nx, ny, seed=seed).ravel()
    slice_b1 = params[::natoms, 3]
    d = np.abs(b1[:, None] - slice_b1[None, :])
    ks = np.argmin(np.where(d < 10.0, d, 10.0), axis=1)
    rng = np.random.default_rng(1 + 1000 * seed)
    truth = ks * natoms + rng.integers(0, natoms, V)
    pd = rng.uniform(0.3, 1.0, V).astype(np.float32)
    d_truth = torch.from_numpy(truth).cuda()
    clean = d_atoms[d_truth] * torch.from_numpy(pd).cuda()[:, None]
    rms = clean.abs().pow(2).mean(dim=1).sqrt()
    g = torch.Generator(device="cuda").manual_seed(2 + 1000 * seed)
    noise = torch.complex(torch.randn(V, T, device="cuda", generator=g),
                          torch.randn(V, T, device="cuda", generator=g))
    d_sig = (clean + noise * (rms / (snr * 2 ** 0.5))[:, None]).to(torch.complex64).contiguous()
    return d_sig, torch.from_numpy(b1).cuda(), truth, V, ks


def profile_run(sm, torch, np, synth, raw, T, w, d_params, A, natoms, params, d_atoms, d_norm2,
                stream):
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_atoms, d_norm2=d_norm2,
               normalize=0, stream=stream)
    torch.cuda.synchronize()
    log("profile: EPG launch done")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_atoms, d_norm2=d_norm2,
               stream=stream)
    d_sig, d_b1, truth, V, _ = make_signals(torch, np, synth, d_atoms, params, natoms, T)
    d_maps = torch.empty((V, 5), device="cuda")
    sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps,
                d_b1map=d_b1, stream=stream)
    torch.cuda.synchronize()
    log("profile: matching done")
    del d_sig, d_b1, d_maps
    torch.cuda.empty_cache()

    def timed1(fn, k):
        fn()
        torch.cuda.synchronize()
        return float("nan")   # rates are not reported in profile mode
    differentiation(sm, torch, np, raw, T, w, d_atoms, d_params, params, A, natoms, stream,
                    timed1, 0)
    log("profile: differentiation done")
    torch.cuda.empty_cache()
    precision_form(sm, torch, np, synth, raw, T, w, stream, 1, timed1, 0)
    log("profile: precision form done")
    src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")   # 1 GiB
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log("profile: 1 GiB calibration copy done")
    return 0


PEAK_BF16 = 2500.0         # TFLOP/s, MI355X dense BF16 MFMA
LOWP_PER_SLICE = 100000


def lowp_params(sm, np):
    """The 500,000-atom precision workload: the first 100,000 atoms (the
    reference order: T1 fastest, then T2, B0) of each of the five B1 slices
    of the 871,455-atom grid, i.e. B0 -40 .. 6 Hz at 2 Hz steps."""
    params, A, natoms = sm.build_params(*GRID)
    rows = (np.arange(A // natoms)[:, None] * natoms
            + np.arange(LOWP_PER_SLICE)[None, :]).ravel()
    return np.ascontiguousarray(params[rows]), len(rows), LOWP_PER_SLICE


def precision_form(sm, torch, np, synth, raw, T, w, stream, steps, timed, oracle_voxels=256):
    """BF16 storage + BF16 MFMA + rescore of the top 8/16/32/64 candidates,
    and the split-BF16 forms, against the FP32 matcher on 240x240x8 voxels;
    map errors against the FP64 oracle on a seeded voxel subset."""
    params, A, natoms = lowp_params(sm, np)
    log(f"precision form: {A} atoms ({A // natoms} slices x {natoms}), 460800 voxels")
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_atoms, stream=stream)
    ld = sm.lowp_ld(T)
    d_hi = torch.empty((A, ld), dtype=torch.int32, device="cuda")
    d_lo = torch.empty((A, ld), dtype=torch.int32, device="cuda")
    prep_ms = timed(lambda: sm.lowp_prepare(d_atoms, T, A, d_hi, d_lo, stream=stream), 1)
    sigs, b1s, truths = [], [], []
    for sl in range(8):
        d_sig, d_b1, truth, V1, _ = make_signals(torch, np, synth, d_atoms, params, natoms, T,
                                                 seed=sl)
        sigs.append(d_sig)
        b1s.append(d_b1)
        truths.append(truth)
    d_sig = torch.cat(sigs)
    d_b1 = torch.cat(b1s)
    truth = np.concatenate(truths)
    del sigs, b1s
    V = d_sig.shape[0]
    res = {"workload": f"epg_ssfp {A} atoms (first {natoms} of each B1 slice of the 871455-atom "
                       f"grid: B0 -40..6 Hz at 2 Hz) x T={T}, 240x240x8 = {V} voxels, "
                       f"B1-map slice selection", "atoms": A, "voxels": V,
           "dictionary_bytes": {"c64": A * T * 8, "bf16": A * ld * 4, "split_bf16": A * ld * 8},
           "lowp_prepare_ms": prep_ms}
    out = {}
    r32 = sm.MRF_MATCH_RESCORE_FP32
    for name, mode, k in (("fp32", sm.MRF_MATCH_FP32, 1),
                          ("bf16_top8", sm.MRF_MATCH_BF16_TOP8, steps),
                          ("bf16_top16", sm.MRF_MATCH_BF16_TOP16, steps),
                          ("bf16_top32", sm.MRF_MATCH_BF16_TOP32, steps),
                          ("bf16_top64", sm.MRF_MATCH_BF16_TOP64, steps),
                          ("bf16_top32_fp32_rescore", sm.MRF_MATCH_BF16_TOP32 | r32, steps),
                          ("bf16_top64_fp32_rescore", sm.MRF_MATCH_BF16_TOP64 | r32, steps),
                          ("split2", sm.MRF_MATCH_SPLIT2, steps),
                          ("split3", sm.MRF_MATCH_SPLIT3, steps)):
        d_maps = torch.empty((V, 5), device="cuda")
        d_idx = torch.empty(V, dtype=torch.int32, device="cuda")

        def run():
            sm.match_ex(mode, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps, d_b1map=d_b1,
                        d_idx=d_idx, d_lowp=d_hi, d_lowp_lo=d_lo, stream=stream)
        run()
        torch.cuda.synchronize()
        ms = timed(run, k)
        out[name] = (d_maps.cpu().numpy(), d_idx.cpu().numpy())
        flops = 8.0 * T * natoms * V
        res[name] = {"voxels_per_s": V / (ms * 1e-3), "ms_per_call": ms, "calls": k,
                     "tflops_algorithmic": flops / (ms * 1e-3) / 1e12,
                     "truth_agree": float(np.mean(out[name][1] == truth))}
        log(f"precision form {name}: {ms:.1f} ms/call")
    i32 = out["fp32"][1]
    lowp_names = ("bf16_top8", "bf16_top16", "bf16_top32", "bf16_top64", "bf16_top32_fp32_rescore",
                  "bf16_top64_fp32_rescore", "split2", "split3")
    for name in lowp_names:
        res[name]["argmax_agree_fp32"] = float(np.mean(out[name][1] == i32))
        res[name]["speedup_vs_fp32"] = res[name]["voxels_per_s"] / res["fp32"]["voxels_per_s"]
    # MFMA work of the split forms is 2x / 3x the algorithmic flops
    for name, nprod in (("split2", 2), ("split3", 3)):
        res[name]["tflops_mfma"] = res[name]["tflops_algorithmic"] * nprod
    # BF16_TOP32 with the dictionary streamed in 1 GiB blocks (generation,
    # BF16 conversion and matching of each block; no dictionary in memory)
    d_maps = torch.empty((V, 5), device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")

    def streamed():
        sm.dict_match_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_sig, V, d_maps,
                         d_b1map=d_b1, d_idx=d_idx, mode=sm.MRF_MATCH_BF16_TOP32, stream=stream)
    streamed()
    torch.cuda.synchronize()
    st_ms = timed(streamed, 1)
    st_idx = d_idx.cpu().numpy()
    res["bf16_top32_streamed"] = {
        "ms_gen_and_match": st_ms, "voxels_per_s": V / (st_ms * 1e-3),
        "argmax_agree_whole_dictionary": float(np.mean(st_idx == out["bf16_top32"][1])),
        "argmax_agree_fp32": float(np.mean(st_idx == out["fp32"][1]))}
    log(f"precision form bf16_top32 streamed (generation included): {st_ms:.1f} ms")
    # the precision form's headline: the fastest one-product BF16 mode that
    # meets the >= 99.9 % agreement bar
    kernels = {"bf16_top8": "match_bf16_dma_kernel<3> + rescore_kernel<8,24>",
               "bf16_top16": "match_bf16_dma_kernel<3> + rescore_kernel<16,24>",
               "bf16_top32": "match_bf16_dma_kernel<3> + rescore_kernel<32,24>",
               "bf16_top64": "match_bf16_dma_kernel<3> + rescore_kernel<64,24>"}
    ok = [n for n in kernels if res[n]["argmax_agree_fp32"] >= 0.999]
    best = max(ok, key=lambda n: res[n]["voxels_per_s"]) if ok else "bf16_top64"
    res["headline_mode"] = best
    res["roofline_bf16"] = {"bound": "mfma", "achieved": res[best]["tflops_algorithmic"],
                            "peak": PEAK_BF16, "unit": "TFLOP/s",
                            "frac": res[best]["tflops_algorithmic"] / PEAK_BF16,
                            "traffic": measured_traffic("match_" + best),
                            "mfma_busy": measured_counters("match_" + best, "mfma"),
                            "kernel": kernels[best], "mode": best}
    # FP64 oracle on a seeded subset: T1/T2/B0/PD errors of every path, and
    # on a sample of the voxels where a low-precision pick differs from the
    # FP32 pick: which of the two the exact (FP64) argmax chooses.
    if oracle_voxels:
        import oracle
        atoms = d_atoms.cpu().numpy()
        b1_host = d_b1.cpu().numpy()
        slice_b1 = params[::natoms, 3]

        def fp64_match(vox):
            sig = d_sig[torch.from_numpy(vox).cuda()].cpu().numpy()
            ks = np.array([oracle.b1_slice(v, slice_b1) for v in b1_host[vox]])
            ref = np.zeros((len(vox), 5))
            ridx = np.zeros(len(vox), np.int64)
            for k in np.unique(ks):
                sel = ks == k
                i, pd, _ = oracle.match_slice(atoms[k * natoms:(k + 1) * natoms], sig[sel])
                ridx[sel] = i + k * natoms
                ref[sel, :4] = params[ridx[sel]]
                ref[sel, 4] = pd
            return ridx, ref
        rng = np.random.default_rng(7)
        sub = np.sort(rng.choice(V, oracle_voxels, replace=False))
        t0 = time.time()
        _, ref = fp64_match(sub)
        dt_cpu = time.time() - t0
        log(f"FP64 oracle on {oracle_voxels} voxels: {dt_cpu:.1f} s")
        res["cpu_fp64_match"] = {"voxels_per_s": oracle_voxels / dt_cpu, "voxels": oracle_voxels,
                                 "cores": oracle.num_threads(), "kind": "port",
                                 "sample": f"{oracle_voxels} seeded voxels, each against its "
                                           f"B1 slice ({natoms} atoms), FP64 oracle"}
        errs = {}
        for name in ("fp32",) + lowp_names:
            m = out[name][0][sub]
            errs[name] = {"T1_mae_ms": float(np.mean(np.abs(m[:, 0] - ref[:, 0]))),
                          "T2_mae_ms": float(np.mean(np.abs(m[:, 1] - ref[:, 1]))),
                          "B0_mae_hz": float(np.mean(np.abs(m[:, 2] - ref[:, 2]))),
                          "PD_max_rel": float(np.max(np.abs(m[:, 4] - ref[:, 4]) / ref[:, 4])),
                          "map_exact": float(np.mean(np.all(m[:, :4] == ref[:, :4], axis=1)))}
        res["fp64_oracle_subset"] = {"voxels": oracle_voxels, "errors": errs}
        arb = {}
        for name in lowp_names:
            diff = np.nonzero(out[name][1] != i32)[0]
            if len(diff) == 0:
                arb[name] = {"disagree": 0}
                continue
            # every disagreement, or a seeded sample of 4,096 where there are
            # more (about 20 s of FP64 matching on 16 host threads)
            samp = np.sort(rng.choice(diff, min(len(diff), 4096), replace=False))
            t0 = time.time()
            ridx, _ = fp64_match(samp)
            lowp_right = float(np.mean(out[name][1][samp] == ridx))
            fp32_right = float(np.mean(i32[samp] == ridx))
            # FP64 score margin by which the FP32 pick loses where it is wrong:
            # (|<a_best, s>|^2 - |<a_fp32, s>|^2) / |<a_best, s>|^2
            sig = d_sig[torch.from_numpy(samp).cuda()].cpu().numpy().astype(np.complex128)
            wrong = i32[samp] != ridx
            def score(ix):
                return np.abs(np.sum(np.conj(atoms[ix].astype(np.complex128)) * sig[wrong], axis=1)) ** 2
            gap = ((score(ridx[wrong]) - score(i32[samp][wrong])) / score(ridx[wrong])
                   if np.any(wrong) else np.zeros(1))
            arb[name] = {"disagree": len(diff), "sampled": len(samp),
                         "fp64_picks_lowp": lowp_right, "fp64_picks_fp32": fp32_right,
                         "fp32_wrong_fp64_margin_max": float(np.max(gap)),
                         "fp32_wrong_fp64_margin_median": float(np.median(gap)),
                         "agree_fp32_or_fp64": 1.0 - len(diff) / V * (1.0 - lowp_right)}
            res[name]["agree_fp32_or_fp64"] = arb[name]["agree_fp32_or_fp64"]
            log(f"FP64 arbitration {name}: {len(diff)} disagreements, lowp right "
                f"{lowp_right:.3f}, fp32 right {fp32_right:.3f} ({time.time() - t0:.1f} s)")
        res["fp64_arbitration"] = arb
        del atoms
    del d_atoms, d_hi, d_lo, d_sig
    torch.cuda.empty_cache()
    return res


# Algorithmic FP32 flops per (atom, EPG state, TR) of the differentiation
# kernels: the Jacobian propagates the 102-flop SSFP step for the signal and 4
# tangents plus their source terms (B1: 78, T1/T2: 24, B0: 40); the adjoint
# counts one forward, one adjoint and one parameter contraction step.
FLOP_JAC = 5 * 102 + 78 + 24 + 40
FLOP_ADJ = 3 * 102


def differentiation(sm, torch, np, raw, T, w, d_atoms, d_params, params, A, natoms, stream,
                    timed, fd_atoms=1024):
    """Jacobian, adjoint and Gauss-Newton over 240x240 off-grid voxels, and
    the CRB design-loss gradient over 64 tissues (w = 60)."""
    V = 240 * 240
    rng = np.random.default_rng(21)
    truth = np.stack([rng.uniform(300, 2500, V), rng.uniform(30, 300, V),
                      rng.uniform(-30, 30, V), rng.uniform(0.85, 1.15, V)], 1).astype(np.float32)
    d_truth = torch.from_numpy(truth).cuda()
    d_s = torch.empty((V, T), dtype=torch.complex64, device="cuda")
    d_J = torch.empty((V, 4, T), dtype=torch.complex64, device="cuda")
    run_jac = lambda: sm.epg_jacobian(raw, d_truth, V, w, d_sig=d_s, d_jac=d_J, stream=stream)
    run_jac()
    jac_ms = timed(run_jac, 1)
    # Jacobian against FP64 central differences (h = 1e-4 relative) of the
    # oracle forward on 1,024 atoms
    fd_err = None
    if fd_atoms:
        import concurrent.futures as cf
        import oracle
        sel = np.sort(np.random.default_rng(23).choice(V, fd_atoms, replace=False))
        Jg = d_J[torch.from_numpy(sel).cuda()].cpu().numpy()

        def one(i):
            th = truth[sel[i]].astype(np.float64)
            worst = 0.0
            for d in range(4):
                h = 1e-4 * abs(th[d]) if th[d] != 0 else 1e-4
                tp, tm = th.copy(), th.copy()
                tp[d] += h
                tm[d] -= h
                fd = (oracle.dm_jacobian(raw, tp, w, False) - oracle.dm_jacobian(raw, tm, w, False)) / (2 * h)
                worst = max(worst, float(np.abs(Jg[i, d] - fd).max() / np.abs(fd).max()))
            return worst
        t0 = time.time()
        with cf.ThreadPoolExecutor(max_workers=16) as ex:
            errs_fd = list(ex.map(one, range(fd_atoms)))
        fd_err = {"atoms": fd_atoms, "max_rel_err": float(max(errs_fd)),
                  "median_rel_err": float(np.median(errs_fd)), "h_rel": 1e-4}
        log(f"Jacobian FD check on {fd_atoms} atoms: max rel err {max(errs_fd):.2e} "
            f"({time.time() - t0:.1f} s)")
    del d_J
    # noisy off-grid signals (SNR 25 on unit-norm atoms, PD ~ U(0.3, 1))
    s_hat = d_s / torch.linalg.vector_norm(d_s, dim=1, keepdim=True)
    pd = torch.from_numpy(rng.uniform(0.3, 1.0, V).astype(np.float32)).cuda()
    ph = torch.from_numpy(rng.uniform(0, 2 * np.pi, V).astype(np.float32)).cuda()
    sigma = 1.0 / (25.0 * np.sqrt(T))
    gen = torch.Generator(device="cuda")
    gen.manual_seed(22)
    noise = torch.complex(torch.randn((V, T), generator=gen, device="cuda"),
                          torch.randn((V, T), generator=gen, device="cuda")) * (sigma / np.sqrt(2))
    d_x = (s_hat * (pd * torch.exp(1j * ph))[:, None] + noise * pd[:, None]).to(torch.complex64).contiguous()
    del s_hat, noise
    d_maps = torch.empty((V, 5), device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")
    sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, d_x, V, d_maps, d_idx=d_idx,
                stream=stream)
    torch.cuda.synchronize()
    start = params[d_idx.cpu().numpy()]
    iters = 10
    d_th = torch.from_numpy(start.copy()).cuda()
    d_cost = torch.empty((V, iters + 1), device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record(stream)
    sm.gn_refine(raw, d_x, V, w, iters, d_th, d_cost=d_cost, stream=stream)
    e1.record(stream)
    torch.cuda.synchronize()
    gn_ms = e0.elapsed_time(e1)
    ref = d_th.cpu().numpy()
    cost = d_cost.cpu().numpy()

    def errs(est):
        return {"T1_median_rel": float(np.median(np.abs(est[:, 0] - truth[:, 0]) / truth[:, 0])),
                "T2_median_rel": float(np.median(np.abs(est[:, 1] - truth[:, 1]) / truth[:, 1])),
                "B0_median_abs_hz": float(np.median(np.abs(est[:, 2] - truth[:, 2]))),
                "B1_median_abs": float(np.median(np.abs(est[:, 3] - truth[:, 3])))}
    # adjoint of a random cotangent over the same voxels
    d_cot = torch.complex(torch.randn((V, T), generator=gen, device="cuda"),
                          torch.randn((V, T), generator=gen, device="cuda")).contiguous()
    g_th = torch.empty((V, 4), device="cuda")
    g_seq = torch.zeros((2, T), device="cuda")
    run_adj = lambda: sm.epg_adjoint(raw, d_truth, V, w, d_cot, d_grad_theta=g_th,
                                     d_grad_seq=g_seq, stream=stream)
    run_adj()
    adj_ms = timed(run_adj, 1)
    del d_cot, d_x, d_s
    # CRB design loss over 64 tissues, w = 60
    nt, wc = 64, 60
    tis = np.stack([rng.uniform(300, 2500, nt), rng.uniform(30, 300, nt), rng.uniform(-30, 30, nt),
                    rng.uniform(0.85, 1.15, nt)], 1).astype(np.float32)
    d_tis = torch.from_numpy(tis).cuda()
    d_loss = torch.empty(nt, dtype=torch.float64, device="cuda")
    g_crb = torch.zeros((2, T), device="cuda")
    run_crb = lambda: sm.crb_loss_grad(raw, d_tis, nt, wc, [1.0, 1.0, 0.0, 0.0], d_loss,
                                       d_grad_seq=g_crb, stream=stream)
    run_crb()
    crb_ms = timed(run_crb, 1)
    crb_fd = None
    if fd_atoms:
        # gradient of the summed loss w.r.t. 32 random flip-angle entries
        # against FP64 central differences of the oracle loss
        import concurrent.futures as cf
        import oracle
        g_crb.zero_()
        run_crb()
        torch.cuda.synchronize()
        gfa = g_crb[0].cpu().numpy()
        ent = np.sort(np.random.default_rng(24).choice(T, 32, replace=False))
        t64 = tis.astype(np.float64)
        w64 = np.array([1.0, 1.0, 0.0, 0.0])

        def fd1(t):
            r1, r2 = raw.copy(), raw.copy()
            hh = 1e-3 * max(1.0, abs(float(raw[0, t])))
            r1[0, t] = np.float32(raw[0, t] + hh)
            r2[0, t] = np.float32(raw[0, t] - hh)
            return (oracle.dm_crb(r1, t64, wc, w64) - oracle.dm_crb(r2, t64, wc, w64)) / (
                float(r1[0, t]) - float(r2[0, t]))
        t0 = time.time()
        with cf.ThreadPoolExecutor(max_workers=16) as ex:
            fds = np.array(list(ex.map(fd1, ent)))
        crb_fd = {"entries": 32, "max_rel_err": float(np.max(np.abs(gfa[ent] - fds)) /
                                                       float(np.max(np.abs(fds))))}
        log(f"CRB flip-angle FD check: max rel err {crb_fd['max_rel_err']:.2e} "
            f"({time.time() - t0:.1f} s)")
    jf = float(V) * T * w * FLOP_JAC
    af = float(V) * T * w * FLOP_ADJ
    res = {
        "workload": f"240x240 = {V} off-grid voxels (T1 U(300,2500), T2 U(30,300), B0 U(-30,30), "
                    f"B1 U(0.85,1.15)), T={T}, w={w}; CRB over 64 tissues, w=60",
        "jacobian": {"ms": jac_ms, "atoms_per_s": V / (jac_ms * 1e-3),
                     "tflops_algorithmic": jf / (jac_ms * 1e-3) / 1e12,
                     "frac_fp32_peak": jf / (jac_ms * 1e-3) / 1e12 / PEAK_FP32,
                     "flop_per_state_tr": FLOP_JAC},
        "adjoint": {"ms": adj_ms, "atoms_per_s": V / (adj_ms * 1e-3),
                    "tflops_algorithmic": af / (adj_ms * 1e-3) / 1e12,
                    "frac_fp32_peak": af / (adj_ms * 1e-3) / 1e12 / PEAK_FP32,
                    "flop_per_state_tr": FLOP_ADJ, "checkpoint_every": 4,
                    "lanes_per_atom": 8 if w <= 32 else 16,
                    "segment_states": "first and last in registers, middle two in LDS",
                    "traffic_per_call": measured_traffic("adjoint"),
                    "valu_issue": measured_counters("adjoint", "valu")},
        "gauss_newton": {"ms": gn_ms, "iters": iters, "voxels_per_s": V / (gn_ms * 1e-3),
                         "start_from": "FP32 grid match (all B1 slices)",
                         "errors_grid_match": errs(start), "errors_refined": errs(ref),
                         "cost_monotone": bool(np.all(np.diff(cost, axis=1) <= 0)),
                         "median_cost_ratio": float(np.median(cost[:, -1] / cost[:, 0]))},
        "crb": {"ms": crb_ms, "tissues": nt, "w": wc,
                "loss": float(d_loss.sum().item()), "fd_check": crb_fd},
        "jacobian_fd_check": fd_err,
    }
    log(f"differentiation: jacobian {jac_ms:.1f} ms, adjoint {adj_ms:.1f} ms, "
        f"GN x{iters} {gn_ms:.1f} ms, CRB {crb_ms:.1f} ms")
    del d_truth, g_th, g_seq
    torch.cuda.empty_cache()
    return res


GRID_C1 = ("100:20:2000", "20:10:300", "0", "1")
FLOP_PER_STATE_FISP = 78


def precision_check(sm, torch, np, d_atoms, d_params, params, T, A, natoms, d_sig, V, d_b1,
                    fp32_ms, fp32_idx, fp32_maps, stream, timed, oracle_voxels):
    """BF16_TOP32, BF16_TOP64 and SPLIT2 on a configuration's own dictionary and signals:
    time, ratio to the FP32 matcher, argmax agreement with it, and (on a
    seeded voxel subset) map errors against the FP64 oracle argmax over the
    same complex64 dictionary, with the FP32 matcher's beside them."""
    ld = sm.lowp_ld(T)
    d_hi = torch.empty((A, ld), dtype=torch.int32, device="cuda")
    d_lo = torch.empty((A, ld), dtype=torch.int32, device="cuda")
    sm.lowp_prepare(d_atoms, T, A, d_hi, d_lo, stream=stream)
    res, maps = {}, {"fp32": fp32_maps}
    for name, mode in (("bf16_top32", sm.MRF_MATCH_BF16_TOP32), ("bf16_top64", sm.MRF_MATCH_BF16_TOP64),
                       ("split2", sm.MRF_MATCH_SPLIT2)):
        d_maps = torch.empty((V, 5), device="cuda")
        d_idx = torch.empty(V, dtype=torch.int32, device="cuda")

        def run():
            sm.match_ex(mode, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps, d_b1map=d_b1,
                        d_idx=d_idx, d_lowp=d_hi, d_lowp_lo=d_lo, stream=stream)
        run()
        ms = timed(run, 2)
        idx = d_idx.cpu().numpy()
        maps[name] = d_maps.cpu().numpy()
        res[name] = {"ms_per_call": ms, "speedup_vs_fp32": fp32_ms / ms,
                     "argmax_agree_fp32": float(np.mean(idx == fp32_idx))}
    del d_hi, d_lo
    if oracle_voxels:
        import oracle
        sub = np.sort(np.random.default_rng(41).choice(V, min(V, oracle_voxels), replace=False))
        sig = d_sig[torch.from_numpy(sub).cuda()].cpu().numpy()
        slices = A // natoms
        if d_b1 is not None:
            slice_b1 = params[::natoms, 3]
            ks = np.array([oracle.b1_slice(float(b), slice_b1) for b in d_b1.cpu().numpy()[sub]])
        else:
            ks = np.zeros(len(sub), np.int64)
        ridx = np.zeros(len(sub), np.int64)
        t0 = time.time()
        for k in np.unique(ks):
            sel = ks == k
            if d_b1 is None and slices > 1:
                rows = d_atoms.cpu().numpy()
            else:
                rows = d_atoms[k * natoms:(k + 1) * natoms].cpu().numpy()
            i, _, _ = oracle.match_slice(rows, sig[sel])
            ridx[sel] = i + (0 if d_b1 is None else k * natoms)
        ref = params[ridx]
        for name, m in maps.items():
            mm = m[sub]
            e = {"map_exact": float(np.mean(np.all(mm[:, :4] == ref, axis=1))),
                 "T1_mae_ms": float(np.mean(np.abs(mm[:, 0] - ref[:, 0]))),
                 "T2_mae_ms": float(np.mean(np.abs(mm[:, 1] - ref[:, 1]))),
                 "B0_mae_hz": float(np.mean(np.abs(mm[:, 2] - ref[:, 2])))}
            res.setdefault(name, {})["fp64_oracle"] = e
        res["fp64_oracle_voxels"] = int(len(sub))
        log(f"precision check: FP64 oracle on {len(sub)} voxels, {time.time() - t0:.1f} s")
    return res


def config_fisp(sm, torch, np, synth, T, stream, timed, cpu):
    """The CPU-runnable configuration: EPG-FISP, T1 100:20:2000, T2 20:10:300,
    w = 20, T = 1000; 64 x 64 voxels (PD ~ U(0.5, 1), SNR 30 on the atom RMS)
    matched in FP32.  Atoms are normalised (explicit flag; the reference
    normalises only SSFP/ROA) so that the argmax recovers the truth."""
    w = 20
    raw = synth.csv_columns(synth.sequence(T, seed=0), "epg_fisp")
    params, A, natoms = sm.build_params(*GRID_C1)
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    run = lambda: sm.dict_ex("epg_fisp", raw, T, d_params, A, natoms, w, d_atoms, normalize=1,
                             stream=stream)
    run()
    gen_ms = timed(run, 5)
    rng = np.random.default_rng(30)
    V = 64 * 64
    truth = rng.integers(0, A, V)
    pd = rng.uniform(0.5, 1.0, V).astype(np.float32)
    clean = d_atoms[torch.from_numpy(truth).cuda()] * torch.from_numpy(pd).cuda()[:, None]
    rms = clean.abs().pow(2).mean(dim=1).sqrt()
    g = torch.Generator(device="cuda").manual_seed(31)
    noise = torch.complex(torch.randn(V, T, device="cuda", generator=g),
                          torch.randn(V, T, device="cuda", generator=g))
    d_sig = (clean + noise * (rms / (30.0 * 2 ** 0.5))[:, None]).to(torch.complex64).contiguous()
    d_maps = torch.empty((V, 5), device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")
    mrun = lambda: sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, d_sig, V,
                               d_maps, d_idx=d_idx, stream=stream)
    mrun()
    m_ms = timed(mrun, 5)
    idx = d_idx.cpu().numpy()
    prec = precision_check(sm, torch, np, d_atoms, d_params, params, T, A, natoms, d_sig, V, None,
                           m_ms, idx, d_maps.cpu().numpy(), stream, timed, V if cpu else 0)
    res = {"workload": f"epg_fisp {A} atoms (T1 {GRID_C1[0]}, T2 {GRID_C1[1]}) x T={T} x w={w}, "
                       f"FP32 matching of 64x64 voxels", "atoms": A, "voxels": V,
           "gen_ms": gen_ms, "gen_atoms_per_s": A / (gen_ms * 1e-3),
           "gen_tflops_algorithmic": A * T * w * FLOP_PER_STATE_FISP / (gen_ms * 1e-3) / 1e12,
           "match_ms": m_ms, "match_voxels_per_s": V / (m_ms * 1e-3),
           "match_truth_agree": float(np.mean(idx == truth)), "precision": prec}
    if cpu:
        import oracle
        seq = oracle.seq_prepare("epg_fisp", raw)
        t0 = oracle.wall()
        atoms = oracle.epg("epg_fisp", seq, params, natoms, w)
        dt_gen = oracle.wall() - t0
        a32 = (atoms / np.linalg.norm(atoms, axis=1, keepdims=True)).astype(np.complex64)
        sig = d_sig.cpu().numpy()
        t0 = oracle.wall()
        ridx, _, _ = oracle.match_slice(a32, sig)
        dt_m = oracle.wall() - t0
        res["cpu"] = {"gen_atoms_per_s": A / dt_gen, "match_voxels_per_s": V / dt_m,
                      "cores": oracle.num_threads(), "kind": "port",
                      "sample": "full configuration (all atoms, all voxels), FP64 oracle",
                      "argmax_equal_gpu": float(np.mean(ridx == idx))}
    log(f"FISP config: gen {gen_ms:.3f} ms, match {m_ms:.3f} ms")
    return res


def roa_iso(sm, torch, np, synth, T, stream, timed, check):
    """ROA Bloch variant with 100 isochromats on a 50,000-atom seeded subset
    of the SSFP grid, T = 1000; with `check`, 64 of the atoms against the
    FP64 oracle."""
    raw = synth.csv_columns(synth.sequence(T, seed=0), "roa_ssfp")
    params, A, natoms = sm.build_params(*GRID)
    rng = np.random.default_rng(40)
    sub = params[np.sort(rng.choice(natoms, 50000, replace=False))]
    n = len(sub)
    d_params = torch.from_numpy(np.ascontiguousarray(sub)).cuda()
    d_atoms = torch.empty((n, T), dtype=torch.complex64, device="cuda")
    run = lambda: sm.dict_ex("roa_ssfp", raw, T, d_params, n, n, 30, d_atoms, n_iso=100,
                             iso_spread_hz=100.0, stream=stream)
    run()
    ms = timed(run, 2)
    flops = float(n) * 100 * T * 31
    log(f"ROA 100 isochromats: {ms:.2f} ms")
    res = {"workload": f"roa_ssfp {n} atoms x 100 isochromats (100 Hz band) x T={T}",
           "ms": ms, "atoms_per_s": n / (ms * 1e-3),
           "tflops_algorithmic": flops / (ms * 1e-3) / 1e12,
           "frac_fp32_peak": flops / (ms * 1e-3) / 1e12 / PEAK_FP32, "flop_per_iso_tr": 31}
    if check:
        import oracle
        pick = np.sort(rng.choice(n, 64, replace=False))
        ref = oracle.roa(oracle.seq_prepare("roa_ssfp", raw), sub[pick], 64, n_iso=100,
                         spread_hz=100.0)
        got = d_atoms[torch.from_numpy(pick).cuda()].cpu().numpy()
        res["oracle_max_rel_err_64_atoms"] = float(
            np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
        log(f"ROA 100 isochromats vs FP64 oracle (64 atoms): {res['oracle_max_rel_err_64_atoms']:.2e}")
    return res


def svd_compressed(sm, torch, np, T, d_atoms, d_params, A, natoms, d_sig, d_b1, V, idx32, truth,
                   match_ms, timed, stream, energy=0.99999, k_max=256):
    """Rank-K compressed matching (per-B1-slice SVD bases, the MATLAB
    energy rule at `energy`) of the same voxels: basis, dictionary
    compression and compressed matching times, and agreement with the
    uncompressed FP32 matcher."""
    ns = A // natoms
    d_basis = torch.empty((ns * k_max, T), dtype=torch.complex64, device="cuda")
    t0 = time.time()
    K, sv = sm.svd_basis(d_atoms, T, A, natoms, d_basis, k_max, True, energy, stream=stream)
    basis_s = time.time() - t0
    d_cat = torch.empty((A, K), dtype=torch.complex64, device="cuda")
    comp_ms = timed(lambda: sm.svd_compress(d_basis, K, T, d_atoms, A, natoms, d_cat,
                                            stream=stream), 1)
    d_maps = torch.empty((V, 5), device="cuda")
    d_idx = torch.empty(V, dtype=torch.int32, device="cuda")

    def run():
        sm.svd_match_ex(d_basis, K, T, d_cat, d_params, A, natoms, d_sig, V, d_maps,
                        d_b1map=d_b1, d_idx=d_idx, stream=stream)
    run()
    torch.cuda.synchronize()
    ms = timed(run, 2)
    i = d_idx.cpu().numpy()
    res = {"energy_frac": energy, "K": K, "basis_s": basis_s, "compress_dictionary_ms": comp_ms,
           "match_ms": ms, "voxels_per_s": V / (ms * 1e-3),
           "speedup_vs_fp32_match": match_ms / ms,
           "argmax_agree_fp32": float(np.mean(i == idx32)),
           "truth_agree": float(np.mean(i == truth))}
    log(f"SVD-compressed matching: K={K}, basis {basis_s:.2f} s, compress {comp_ms:.1f} ms, "
        f"match {ms:.1f} ms, agree fp32 {res['argmax_agree_fp32']:.4f}")
    del d_basis, d_cat, d_maps, d_idx
    return res


def measured_traffic(key: str):
    """HBM bytes per launch of a kernel from the committed rocprofv3 --pmc
    summary (profiles/traffic.json, FETCH_SIZE + WRITE_SIZE scaled by the
    1 GiB copy calibration), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            t = json.load(f)[key]
        return t["fetch_bytes"] + t["write_bytes"]
    except (OSError, KeyError, ValueError):
        return None


def measured_counters(key: str, group: str):
    """A kernel's VALU or MFMA counter summary from profiles/traffic.json
    (rocprofv3 --pmc over bench.py --profile), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            c = json.load(f)[key][group]
    except (OSError, KeyError, ValueError):
        return None
    if group == "valu":
        return {"valu_insts_per_launch": c["valu_insts"],
                "insts_per_cycle_per_simd": c["ipc_per_simd"],
                "dual_issued_frac": c["dual_issued_frac"],
                "issue_ceiling_per_cycle_per_simd": 0.5}
    return {"mfma_busy_frac": c["mfma_busy_frac"]}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--small", action="store_true", help="reduced grid (debugging only)")
    ap.add_argument("--skip-match", action="store_true")
    ap.add_argument("--skip-lowp", action="store_true", help="skip the BF16 precision form")
    ap.add_argument("--skip-diff", action="store_true", help="skip the differentiation form")
    ap.add_argument("--oracle-voxels", type=int, default=4096,
                    help="voxels of the precision form checked against the FP64 oracle")
    ap.add_argument("--profile", action="store_true",
                    help="one EPG launch, one FP32 matching call, the precision-form matchers "
                         "and a 1 GiB device copy (the HBM-counter calibration), for rocprofv3 "
                         "runs; prints no JSON")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the whole path runs on a single MI355X "
              "(there is no multi-GPU path); run with --gpus 1", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    import snapmrf_amd as sm
    from snapmrf_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    sm.lib()
    torch.cuda.init()
    stream = torch.cuda.current_stream()
    T, w = 1000, 30
    grid = GRID if not args.small else ("100:100:2000", "20:20:300", "-40:20:40", "0.8:0.2:1.2")
    seq = synth.sequence(T, seed=0)
    raw = synth.csv_columns(seq, "epg_ssfp")
    params, A, natoms = sm.build_params(*grid)
    l_b1 = A // natoms
    log(f"grid: {A} atoms ({l_b1} B1 slices x {natoms}), T={T}, w={w}")
    d_params = torch.from_numpy(params).cuda()
    d_atoms = torch.empty((A, T), dtype=torch.complex64, device="cuda")
    d_norm2 = torch.empty(A, dtype=torch.float64, device="cuda")

    def gen(normalize=-1):
        sm.dict_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_atoms, d_norm2=d_norm2,
                   normalize=normalize, stream=stream)

    def timed(fn, k):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record(stream)
        for _ in range(k):
            fn()
        e1.record(stream)
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / k

    if args.profile:
        return profile_run(sm, torch, np, synth, raw, T, w, d_params, A, natoms, params,
                           d_atoms, d_norm2, stream)

    for _ in range(args.warmup):
        gen()
    torch.cuda.synchronize()
    log("generation warm-up done")
    t0 = time.time()
    gen_ms = timed(gen, args.steps)
    log(f"generation: {gen_ms:.2f} ms/step ({time.time() - t0:.1f} s wall)")
    # EPG kernel alone (no normalisation pass) for the roofline
    kern_ms = timed(lambda: gen(normalize=0), args.steps)
    gen()   # leave the dictionary normalised
    torch.cuda.synchronize()
    log(f"EPG kernel: {kern_ms:.2f} ms/launch")
    # rows of B1 slice 0 that the CPU baseline recomputes with the FP64 oracle
    cb_rows = np.sort(np.random.default_rng(0).choice(natoms, min(natoms, 20000), replace=False))
    cb_gpu = None if args.no_cpu else d_atoms[torch.from_numpy(cb_rows).cuda()].cpu().numpy()
    cb_slice = cb_sig = cb_idx = None
    flops = float(A) * T * w * FLOP_PER_STATE_SSFP
    achieved = flops / (kern_ms * 1e-3) / 1e12

    out = {
        "metric": METRIC, "value": A / (gen_ms * 1e-3), "unit": "atoms/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": gen_ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"epg_ssfp dictionary {A} atoms x T={T} x w={w} (c64, normalised) "
                               f"+ fp32 matching of 240x240 voxels against the B1-map slice",
                   "atoms": A, "b1_slices": l_b1, "T": T, "w": w, "voxels": 240 * 240,
                   "grid": {"T1": grid[0], "T2": grid[1], "B0": grid[2], "B1": grid[3]}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": measured_traffic("epg"),
                     "kernel": "epg_fused_kernel<30,false,true,false>",
                     # the S/D form executes 24 of the 102 algorithmic flop per
                     # state and TR as VALU instructions; how busy the SIMDs are
                     # is in the VALU issue counters
                     "valu_issue": measured_counters("epg", "valu"),
                     "algorithmic_flops_per_launch": flops, "ms_per_launch": kern_ms},
        "ref_published_atoms_per_s": 10765,
    }

    if not args.skip_match:
        d_sig, d_b1, truth, V, vox_ks = make_signals(torch, np, synth, d_atoms, params, natoms, T)
        d_maps = torch.empty((V, 5), device="cuda")
        d_idx = torch.empty(V, dtype=torch.int32, device="cuda")

        def match():
            sm.match_ex(sm.MRF_MATCH_FP32, d_atoms, d_params, T, A, natoms, d_sig, V, d_maps,
                        d_b1map=d_b1, d_idx=d_idx, stream=stream)

        match()
        torch.cuda.synchronize()
        log("matching warm-up done")
        msteps = max(1, args.steps // 2)
        match_ms = timed(match, msteps)
        log(f"matching fp32: {match_ms:.1f} ms/step")
        idx = d_idx.cpu().numpy()
        if not args.no_cpu:
            # 256 voxels whose B1 map selects slice 0, for the CPU matcher
            sel = np.sort(np.random.default_rng(1).choice(np.nonzero(vox_ks == 0)[0], 256,
                                                           replace=False))
            cb_sig = d_sig[torch.from_numpy(sel).cuda()].cpu().numpy()
            cb_idx = idx[sel]
            cb_slice = d_atoms[:natoms].cpu().numpy()
        mflops = 8.0 * T * natoms * V
        ach_m = mflops / (match_ms * 1e-3) / 1e12
        out["match_voxels_per_s"] = V / (match_ms * 1e-3)
        out["match_ms_per_step"] = match_ms
        out["match_steps"] = msteps
        out["match_truth_agree"] = float(np.mean(idx == truth))
        out["roofline_match"] = {"bound": "mfma", "achieved": ach_m, "peak": PEAK_FP32,
                                 "unit": "TFLOP/s", "frac": ach_m / PEAK_FP32,
                                 "traffic": measured_traffic("match"),
                                 "kernel": "match_fp32_kernel<true>",
                                 "algorithmic_flops_per_launch": mflops,
                                 # three real products per complex product:
                                 # the MFMA pipes issue 6 of the 8 flops
                                 "mfma_flops_per_launch": 0.75 * mflops,
                                 "frac_mfma_issued": 0.75 * ach_m / PEAK_FP32,
                                 "mfma_busy": measured_counters("match", "mfma")}
        if not args.skip_lowp:
            out["precision_c2"] = precision_check(
                sm, torch, np, d_atoms, d_params, params, T, A, natoms, d_sig, V, d_b1, match_ms,
                idx, d_maps.cpu().numpy(), stream, timed,
                0 if (args.no_cpu or args.oracle_voxels == 0) else 256)
            log(f"precision on this workload: {json.dumps(out['precision_c2'])[:200]}")
        # the reference interface: MRF_match with a host image and a B1-map
        # file (PCIe upload and file read included, wall clock)
        import tempfile
        h_img = d_sig.cpu().numpy()
        with tempfile.TemporaryDirectory() as tdir:
            b1_path = os.path.join(tdir, "b1.ra")
            sm.ra_write(b1_path, d_b1.cpu().numpy())
            d_maps_h = torch.empty((V, 5), device="cuda")
            torch.cuda.synchronize()
            t0 = time.time()
            sm.MRF_match(d_maps_h, h_img, d_atoms, d_params, T, T, A, natoms, V, 4, 5, b1_path)
            host_ms = (time.time() - t0) * 1e3
        out["match_host_image"] = {
            "ms": host_ms, "voxels_per_s": V / (host_ms * 1e-3),
            "image_bytes": int(h_img.nbytes),
            "maps_equal_device_path": bool(torch.equal(d_maps_h, d_maps))}
        log(f"MRF_match with a host image: {host_ms:.1f} ms")
        del h_img, d_maps_h
        # the same generation + matching streamed: the dictionary exists one
        # block (at most 1 GiB) at a time
        d_maps_s = torch.empty((V, 5), device="cuda")
        d_idx_s = torch.empty(V, dtype=torch.int32, device="cuda")

        def streamed():
            sm.dict_match_ex("epg_ssfp", raw, T, d_params, A, natoms, w, d_sig, V, d_maps_s,
                             d_b1map=d_b1, d_idx=d_idx_s, stream=stream)
        streamed()
        torch.cuda.synchronize()
        st_ms = timed(streamed, 1)
        block = (1 << 30) // (l_b1 * T * 8) // 128 * 128
        out["streamed_gen_match"] = {
            "ms": st_ms, "separate_gen_plus_match_ms": gen_ms + match_ms,
            "block_atoms_per_slice": block, "block_bytes": block * l_b1 * T * 8,
            "dictionary_bytes": A * T * 8,
            "idx_equal_whole_dictionary": bool(np.array_equal(d_idx_s.cpu().numpy(), idx)),
            "maps_equal_whole_dictionary": bool(torch.equal(d_maps_s, d_maps))}
        log(f"streamed generation + matching: {st_ms:.1f} ms "
            f"(equal: {out['streamed_gen_match']['idx_equal_whole_dictionary']})")
        del d_maps_s, d_idx_s
        out["svd_compressed"] = svd_compressed(sm, torch, np, T, d_atoms, d_params, A, natoms,
                                               d_sig, d_b1, V, idx, truth, match_ms, timed,
                                               stream)
    if not args.skip_diff:
        out["differentiation"] = differentiation(sm, torch, np, raw, T, w, d_atoms, d_params,
                                                 params, A, natoms, stream, timed,
                                                 0 if args.no_cpu else 1024)
    out["fisp_config"] = config_fisp(sm, torch, np, synth, T, stream, timed, not args.no_cpu)
    out["roa_isochromats"] = roa_iso(sm, torch, np, synth, T, stream, timed, not args.no_cpu)
    if not args.skip_lowp:
        if not args.skip_match:
            del d_sig, d_b1, d_maps, d_idx
        del d_atoms, d_norm2
        torch.cuda.empty_cache()
        out["precision_form"] = precision_form(sm, torch, np, synth, raw, T, w, stream,
                                               args.steps, timed, args.oracle_voxels)
    if not args.no_cpu:
        log("CPU baseline (FP64 oracle on host cores)...")
        cb = cpu_baseline(raw, params, natoms, T, w, cb_rows, cb_gpu, cb_slice, cb_sig, cb_idx)
        out["cpu_baseline"] = cb
        log(f"CPU baseline: {cb['value']:.1f} atoms/s, {cb.get('match_value', 0):.2f} voxels/s "
            f"on {cb['cores']} threads; GPU rows vs FP64 oracle: max rel err "
            f"{cb['gpu_atoms_max_rel_err']:.2e} on {cb['gpu_atoms_checked']} atoms")
        out["gen_parity"] = {"atoms_checked": cb["gpu_atoms_checked"],
                             "max_rel_err_vs_fp64_oracle": cb["gpu_atoms_max_rel_err"],
                             "tolerance": 1e-5}
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
