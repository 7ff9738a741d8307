"""Benchmark of the hot path on one MI355X.

Workload (the configuration the metric is quoted on): 2-centre Cartesian-Gaussian
overlaps of B = 65536 atom pairs over 3 synthetic double-zeta-shaped species
(9s4p1d -> 3s2p1d, N = 15 Cartesian functions, K = 9 primitive slots, exponents
log-uniform in [0.1, 7000], mixed-sign general contractions), FP32.  One step =
rrule(batch_overlap) on (ps): the forward overlap S, then the (zeta, D) pullback of a
fixed cotangent dS, sharing one plan, all through libaok.so; inputs resident in HBM.

Also reported (extra keys): forward-only and pullback-only pairs/s, the plain-FP64
GPU form (the reference decomposition) and the speedup over it, the FP32 error
against FP64, and atomic-orbital evaluation at nX = 2^24 (30 orbitals, 3 species)
in GB/s against the 8 TB/s HBM peak.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-extra]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "atomicorbitalkernels.jl_amd"))

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3          # MI355X FP32 vector, packed FMA
FP32_UNPACKED_TFLOPS = 78.6       # the same without packed FP32 (the kernels are built without it)
# HBM bytes per launch of the C1 kernels: rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE
# runs of this bench's C1 step (--no-extra; KiB per dispatch, median over the launches), each
# scaled by what the same counter reads for the bench's 2 GiB device copy (FETCH_SIZE reads
# 1048589 KiB for 2097152 KiB moved: x2.0; WRITE_SIZE reads 2097152 KiB: x1.0).
PMC_KIB = {"ovl2c_bwd_fast": (124037.0, 5926.6), "ovl2c_fwd_fast": (8356.1, 116779.4)}
PMC_SCALE = (2.0, 1.0)
# FLOPs per pair the C1 kernels execute: rocprofv3 --pmc SQ_INSTS_VALU_FLOPS_FP32 and
# SQ_INSTS_VALU_FLOPS_FP64 (per wave instruction, x 64 lanes), per launch / 65536 pairs;
# VALU busy: SQ_ACTIVE_INST_VALU * 4 / (1024 SIMDs * GRBM_GUI_ACTIVE / 8 XCDs).  Reported
# beside the algorithmic rate, never priced.
EXEC_FLOP_PER_PAIR = {"ovl2c_bwd_fast": (3.582e+09 + 1.004e+08) / 65536, "ovl2c_fwd_fast": (4.032e+08 + 1.287e+08) / 65536}
VALU_BUSY = {"ovl2c_bwd_fast": 0.74, "ovl2c_fwd_fast": 0.41}
METRIC = "2-ctr overlap pairs/s + (ζ,D)-grad pairs/s, FP32@1e-6 vs plain FP64; AO GB/s"


# ---- algorithmic FLOP model of the Obara-Saika 2-centre overlap -----------------------------
# Counted per primitive pair (a, b) of a shell pair (la, lb), FMA = 2 FLOPs, exp = 1:
#  * geometry and prefactor: mu R^2, exp, prefactor product (3); PA, PB per axis when
#    la + lb > 0 (6);
#  * 1-D tables s_ij (i <= la, j <= lb) per axis: vertical recursion
#    s_{i+1,0} = PA s_i0 + i/(2p) s_{i-1,0} (3 FLOPs per entry beyond s_10 = PA) to i = la + lb,
#    horizontal transfer s_{i,j+1} = s_{i+1,j} + AB s_ij (one FMA per entry);
#  * forward: the prefactor folded into the x table ((la+1)(lb+1) products), then per output
#    element s_x s_y (1) and an FMA into the accumulator (2): 3 FLOPs;
#  * pullback (real x real): tables to (la + 2, lb) and (la, lb + 2) from one vertical
#    recursion to la + lb + 2 and the horizontal transfer to lb + 2; per element the value
#    sum g s_x s_y s_z (4) and the bra- and ket-raised sums (8 each); 8 FLOPs of epilogue
#    (the four cotangent updates); s-s pairs use the closed form of the raised sums
#    (|PA|^2 + 3/(2p), 4 FLOPs each) instead of tables;
#  * pullback, padded representative x real: the value sum only (4 per element) on the
#    forward's tables, 2 FLOPs of epilogue.
def _nbf(l):
    return (l + 1) * (l + 2) // 2


def _os_tables(la, lb):
    return 3 * max(la + lb - 1, 0) + 2 * sum(la + lb - j + 1 for j in range(1, lb + 1))


def os_flops_fwd(la, lb):
    geo = 3 + (6 if la + lb > 0 else 0)
    pref = (la + 1) * (lb + 1) if la + lb > 0 else 0
    return geo + 3 * _os_tables(la, lb) + pref + 3 * _nbf(la) * _nbf(lb)


def os_flops_bwd(la, lb):
    if la == 0 and lb == 0:
        return 3 + 2 * 4 + 4 + 8 + 8
    vrr = 3 * (la + lb + 1)
    hrr = 2 * sum(la + lb + 3 - j for j in range(1, lb + 3))
    return 9 + 3 * (vrr + hrr) + 20 * _nbf(la) * _nbf(lb) + 8


def os_flops_pad(la, lb):
    return os_flops_fwd(la, lb) + _nbf(la) * _nbf(lb) + 2


def os_model_flops(basis, sidxA, sidxB):
    """Algorithmic FLOPs of one forward and one (zeta, D) pullback launch over the batch:
    forward per real x real primitive pair; pullback per real x real pair plus per (distinct
    padded exponent x real) pair, the padded slots' coefficient cotangents."""
    import numpy as np

    NZ = basis.NZ
    nreal = (basis.D != 0).sum(axis=1)  # (ns, NZ)
    npad = np.zeros_like(nreal)
    for s in range(basis.ns):
        for g in range(NZ):
            pad = basis.zeta[s, basis.D[s, :, g] == 0, g]
            npad[s, g] = len(set(pad.tolist()))
    fwd = np.zeros((NZ, NZ))
    bwd = np.zeros((NZ, NZ))
    for a in range(NZ):
        for b in range(NZ):
            for sa in range(basis.ns):
                for sb in range(basis.ns):
                    la, lb = basis.ls[sa], basis.ls[sb]
                    rr = nreal[sa, a] * nreal[sb, b]
                    fwd[a, b] += rr * os_flops_fwd(la, lb)
                    bwd[a, b] += rr * os_flops_bwd(la, lb) + \
                        (npad[sa, a] * nreal[sb, b] + nreal[sa, a] * npad[sb, b]) * os_flops_pad(la, lb)
    cnt = np.zeros((NZ, NZ))
    np.add.at(cnt, (np.asarray(sidxA) - 1, np.asarray(sidxB) - 1), 1)
    return float((fwd * cnt).sum()), float((bwd * cnt).sum())


def _cpu_model():
    """The host CPU as /proc/cpuinfo names it, with its family/model numbers (a masked
    "AMD EPYC" name still identifies the generation by them) and the logical CPUs this
    process may use."""
    info = {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if ":" not in line:
                    if info:
                        break
                    continue
                k, v = (x.strip() for x in line.split(":", 1))
                info.setdefault(k, v)
    except OSError:
        return "unknown"
    name = info.get("model name", "unknown")
    fam, mod = info.get("cpu family"), info.get("model")
    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return f"{name} (family {fam} model {mod}, {ncpu} logical CPUs available to the process)"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=300)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the kernels run on a single MI355X); --gpus {args.gpus} is not "
              f"supported; the library shards by batch element and each GPU runs this same kernel on its share.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import aok
    from aok import configs

    def log(msg):
        print(f"[bench] {msg}", file=sys.stderr, flush=True)

    dev = torch.device("cuda:0")
    B = 65536
    d = configs.c1(1, B=B)
    cb = d["basis"]
    spec, _, _ = cb.ao_maps()
    zlist = (1, 2, 3)

    def orbital(dtype):
        R = aok.GaussianTypeRadials(aok.fortran_from_numpy(cb.zeta, dtype, dev),
                                    aok.fortran_from_numpy(cb.D, dtype, dev), spec, None, zlist)
        return aok.AtomicOrbitals(R, 2, length_unit="bohr")

    def inputs(dtype):
        XA = (torch.as_tensor(d["posA"], dtype=dtype, device=dev).contiguous(),
              torch.as_tensor(d["sidxA"], dtype=torch.int32, device=dev))
        XB = (torch.as_tensor(d["posB"], dtype=dtype, device=dev).contiguous(),
              torch.as_tensor(d["sidxB"], dtype=torch.int32, device=dev))
        return XA, XB

    ctx = aok.context(0)
    ctx.set_timing(True)
    orb32 = orbital(torch.float32)
    ps32 = orb32.static_params()
    XA32, XB32 = inputs(torch.float32)
    g = torch.Generator(device=dev).manual_seed(7)
    dS32 = torch.randn((B, cb.N, cb.N), generator=g, device=dev, dtype=torch.float32).permute(2, 1, 0)

    def step():
        # rrule(batch_overlap): forward S, then the pullback of dS to (zeta, D)
        S, pb = aok.rrule_batch_overlap(orb32, XA32, XB32, ps32, ctx=ctx)
        dps = pb(dS32)[4]
        return S, dps

    # stream-ordered execution: each call returns once its kernels are enqueued on the
    # context stream (torch's default stream); the timed region ends with aok_sync, which
    # also reports any device-detected error.  The dominant kernels' HIP-event times are
    # recorded during the timed region and read back after it.
    ctx.set_blocking(False)
    log(f"C1 warmup {args.warmup}")
    for _ in range(args.warmup):
        step()
    ctx.sync()
    ctx.kernel_times()
    torch.cuda.synchronize()
    log(f"C1 timed {args.steps} steps")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    ctx.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.set_blocking(True)
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern = {"ovl2c_fwd_fast": [], "ovl2c_bwd_fast": []}
    for name, ms in ctx.kernel_times():
        kern.setdefault(name, []).append(ms)
    assert len(kern["ovl2c_fwd_fast"]) == args.steps and len(kern["ovl2c_bwd_fast"]) == args.steps
    value = B / (ms_step * 1e-3)
    fwd_ms = float(np.mean(kern["ovl2c_fwd_fast"]))
    bwd_ms = float(np.mean(kern["ovl2c_bwd_fast"]))
    flops_fwd, flops_bwd = os_model_flops(cb, d["sidxA"], d["sidxB"])
    dom = "ovl2c_bwd_fast" if bwd_ms >= fwd_ms else "ovl2c_fwd_fast"
    dom_ms = max(fwd_ms, bwd_ms)
    dom_flops = flops_bwd if dom == "ovl2c_bwd_fast" else flops_fwd
    exec_flops = EXEC_FLOP_PER_PAIR[dom] * B
    achieved = dom_flops / (dom_ms * 1e-3) / 1e12
    fk, wk = PMC_KIB[dom]
    traffic = (fk * PMC_SCALE[0] + wk * PMC_SCALE[1]) * 1024.0
    out = {
        "metric": METRIC, "value": value, "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "2c overlap forward + (zeta,D) pullback, B=65536 pairs, 3 species 9s4p1d->3s2p1d "
                               "(N=15, K=9), zeta log-uniform [0.1,7000], FP32", "global_batch": B,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic, "kernel": dom,
                     "kernel_ms": dom_ms, "algorithmic_flop_per_launch": dom_flops,
                     "algorithmic_flop_per_pair": dom_flops / B,
                     "flop_model": "Obara-Saika count per primitive pair (bench.py os_flops_*): geometry + "
                                   "1-D tables + per-element contraction, FMA = 2",
                     "frac_of_unpacked_fp32": achieved / FP32_UNPACKED_TFLOPS,
                     "executed_flop_per_launch": exec_flops,
                     "executed_TFLOPs": exec_flops / (dom_ms * 1e-3) / 1e12,
                     "valu_busy": VALU_BUSY[dom],
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "note": "FP32 VALU kernel (no MFMA: the primitive contraction has no operand shared "
                             "across pairs); peak = published packed FP32 rate"},
        "fwd_kernel_ms": fwd_ms, "bwd_kernel_ms": bwd_ms,
        "fwd_kernel_pairs_per_s": B / (fwd_ms * 1e-3), "grad_kernel_pairs_per_s": B / (bwd_ms * 1e-3),
    }

    def graph_step_ms(orb, XA, XB, ps, dS, n, knobs=None):
        """The same rrule step recorded once into a HIP graph (OverlapStepGraph: plan update,
        forward, pullback, normalisation VJP) and replayed n times, stream-ordered."""
        gctx = aok.Context.side_stream(0)
        for k, v in (knobs or {}).items():
            gctx.tune(k, v)
        gs = aok.OverlapStepGraph(orb, XA, XB, ps, ctx=gctx)
        gs.dS.copy_(dS)
        torch.cuda.synchronize()
        gs.ctx.set_blocking(False)
        for _ in range(2):
            gs.replay()
        gs.ctx.sync()
        ts = time.perf_counter()
        for _ in range(n):
            gs.replay()
        gs.ctx.sync()
        ms = (time.perf_counter() - ts) * 1e3 / n
        gs.close()
        gctx.close()
        return ms

    subsets = {}  # small GPU FP32 results for the CPU leg's oracle-anchored checks
    if not args.no_extra:
        # position cotangents of the same step (rrule with wrt_positions): their kernel alone
        log("C1 position cotangents")
        _, pbx = aok.rrule_batch_overlap(orb32, XA32, XB32, ps32, ctx=ctx, wrt_positions=True)
        pbx(dS32)
        ctx.kernel_times()
        for _ in range(5):
            pbx(dS32)
        dxk = [ms for name, ms in ctx.kernel_times() if name == "ovl2c_dx_fast"]
        out["c1_position_cotangents"] = {"kernel": "ovl2c_dx_fast", "kernel_ms": float(np.mean(dxk)),
                                         "pairs_per_s": B / (float(np.mean(dxk)) * 1e-3)}
        log("C1 graph-replayed step")
        gms = graph_step_ms(orb32, XA32, XB32, ps32, dS32, args.steps)
        out["c1_graph_step"] = {"ms_per_step": gms, "pairs_per_s": B / (gms * 1e-3),
                                "kernels": "plan update (3) + forward + pullback (3), one hipGraphLaunch per step"}
        # FP32 vs FP64 (fast FP64 kernels, themselves pinned to the oracle by the tests)
        log("C1 FP64 reference run for the FP32 error")
        orb64 = orbital(torch.float64)
        XA64, XB64 = inputs(torch.float64)
        # the FP64 kernels (pinned to the oracle by the tests) fed the FP32-rounded inputs,
        # so the difference is the FP32 arithmetic alone
        ps64 = {"Rnl": {"zeta": aok.as_fortran(orb32.Rnl.zeta.double()), "D": aok.as_fortran(orb32.Rnl.D.double())},
                "Ylm": {}}
        XA64r = (XA32[0].double().contiguous(), XA32[1])
        XB64r = (XB32[0].double().contiguous(), XB32[1])
        S32, dps32 = step()
        S64, pb64 = aok.rrule_batch_overlap(orb64, XA64r, XB64r, ps64, ctx=ctx)
        dps64 = pb64(aok.as_fortran(dS32.double()))[4]
        fp32 = {"definition": "max|x32 - x64| / max|x64| per output (overlap blocks: max over pairs of the "
                              "block's value), FP64 kernels on the same FP32-rounded inputs"}
        num = (S32.double() - S64).abs().amax(dim=(0, 1))
        den = S64.abs().amax(dim=(0, 1)).clamp_min(1e-300)
        fp32["c1_S_blockwise"] = float((num / den).max())
        big = S64.abs() >= 1e-3 * S64.abs().amax(dim=(0, 1), keepdim=True)
        rel = ((S32.double() - S64).abs() / S64.abs().clamp_min(1e-300))[big]
        fp32["c1_S_elementwise_tau1e-3"] = float(rel.max())
        for k in ("zeta", "D"):
            g32, g64 = dps32["Rnl"][k].double(), dps64["Rnl"][k]
            fp32[f"c1_d{k}"] = float((g32 - g64).abs().max() / g64.abs().max())
        # position cotangents (dA; dB = -dA)
        dA32 = pbx(dS32)[2].double()
        _, pbx64 = aok.rrule_batch_overlap(orb64, XA64r, XB64r, ps64, ctx=ctx, wrt_positions=True)
        dA64 = pbx64(aok.as_fortran(dS32.double()))[2]
        fp32["c1_dA"] = float((dA32 - dA64).abs().max() / dA64.abs().max())
        out["fp32_blockwise_max_rel_err"] = fp32["c1_S_blockwise"]
        out["fp32_error"] = fp32
        del S64, pb64, pbx64
        # plain FP64 GPU form (reference decomposition) -- the >= 4x denominator
        log("C1 plain FP64 forward/pullback")
        dS64 = dS32.double()
        dS64 = aok.as_fortran(dS64)
        tp = {}
        for name, fn in (("fwd", lambda: aok.batch_overlap(orb64, XA64, XB64, orb64.static_params(), mode="plain",
                                                           ctx=ctx)),
                         ("bwd", lambda: aok.overlap_pullback_ps(dS64, orb64, XA64, XB64, orb64.static_params(),
                                                                 mode="plain", ctx=ctx))):
            fn()
            torch.cuda.synchronize()
            ts = time.perf_counter()
            ks = []
            for _ in range(3):
                fn()
                ks.append(ctx.last_kernel_ms())
            torch.cuda.synchronize()
            tp[name] = ((time.perf_counter() - ts) * 1e3 / 3, float(np.mean(ks)))
        out["plain_f64_fwd_ms"], out["plain_f64_fwd_kernel_ms"] = tp["fwd"]
        out["plain_f64_bwd_ms"], out["plain_f64_bwd_kernel_ms"] = tp["bwd"]
        out["plain_f64_step_pairs_per_s"] = B / ((tp["fwd"][0] + tp["bwd"][0]) * 1e-3)
        out["speedup_vs_plain_f64"] = {
            "fwd": tp["fwd"][1] / fwd_ms, "bwd": tp["bwd"][1] / bwd_ms,
            "fwd+bwd": (tp["fwd"][1] + tp["bwd"][1]) / (fwd_ms + bwd_ms),
            "step_wall": (tp["fwd"][0] + tp["bwd"][0]) / ms_step}
        # HBM ceilings measured with torch's own copy and fill kernels (2 GiB buffers): the
        # copy kernel is also the scale reference for the FETCH_SIZE / WRITE_SIZE counters
        log("HBM copy / fill ceilings")
        nb_cal = 2 << 30
        src = torch.ones(nb_cal // 4, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        for _ in range(3):
            dst.copy_(src)
            dst.fill_(0.5)
        ev[0].record()
        for _ in range(10):
            dst.copy_(src)
        ev[1].record()
        ev[2].record()
        for _ in range(10):
            dst.fill_(0.5)
        ev[3].record()
        torch.cuda.synchronize()
        out["hbm_ceilings"] = {"copy_GB_per_s": 2 * nb_cal * 10 / (ev[0].elapsed_time(ev[1]) * 1e-3) / 1e9,
                               "fill_GB_per_s": nb_cal * 10 / (ev[2].elapsed_time(ev[3]) * 1e-3) / 1e9,
                               "bytes": nb_cal, "kernels": "torch copy_ / fill_ (float32)"}
        del src, dst
        # AO evaluation at nX = 2^24, 30 orbitals, 3 species, FP32 (HBM-bound)
        log("AO nX=2^24 evaluation")
        c4 = configs.c4b_basis()
        spec4, _, _ = c4.ao_maps()
        R4 = aok.GaussianTypeRadials(aok.fortran_from_numpy(c4.zeta, torch.float32, dev),
                                     aok.fortran_from_numpy(c4.D, torch.float32, dev), spec4, None, (1, 2, 3))
        orb4 = aok.AtomicOrbitals(R4, 3, length_unit="bohr")
        nX = 2 ** 24
        gx = torch.Generator(device=dev).manual_seed(4)
        X = 3.0 * torch.randn((nX, 3), generator=gx, device=dev, dtype=torch.float32)
        sidx = torch.randint(1, 4, (nX,), generator=gx, device=dev, dtype=torch.int32)
        for _ in range(3):
            P = aok.evaluate(orb4, X, sidx=sidx, ctx=ctx)
        ks = []
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(10):
            P = aok.evaluate(orb4, X, sidx=sidx, ctx=ctx)
            ks.append(ctx.last_kernel_ms())
        torch.cuda.synchronize()
        wall = (time.perf_counter() - ts) * 1e3 / 10
        bytes_alg = nX * (12 + 4 + 4 * len(orb4))
        kms = float(np.mean(ks))
        out["ao_c4b"] = {"nX": nX, "nOrb": len(orb4), "kernel_ms": kms, "wall_ms": wall, "kernel": "ao_eval",
                         "GB_per_s": bytes_alg / (kms * 1e-3) / 1e9,
                         "frac_of_hbm_peak": bytes_alg / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "frac_of_fill_ceiling": bytes_alg / (kms * 1e-3) / 1e9 / out["hbm_ceilings"]["fill_GB_per_s"],
                         "algorithmic_bytes": bytes_alg}
        # evaluate_ed and the rrule pullback (dX, dzeta, dD) at the same size
        gx2 = torch.Generator(device=dev).manual_seed(5)
        dPin = torch.randn((len(orb4), nX), generator=gx2, device=dev, dtype=torch.float32).t()
        for name, fn, nbytes in (
                ("ao_c4b_eval_ed", lambda: aok.evaluate_ed(orb4, X, sidx=sidx, ctx=ctx),
                 nX * (12 + 4 + 16 * len(orb4))),
                ("ao_c4b_pullback", lambda: aok.rrule_evaluate(orb4, X, sidx=sidx, ctx=ctx)[1](dPin),
                 nX * (12 + 4 + 4 * len(orb4) + 12))):
            log(name)
            fn()
            torch.cuda.synchronize()
            ks = []
            ts = time.perf_counter()
            for _ in range(5):
                fn()
                if ctx.last_kernel_name() in ("ao_eval_ed", "ao_pullback"):
                    ks.append(ctx.last_kernel_ms())
            torch.cuda.synchronize()
            wall = (time.perf_counter() - ts) * 1e3 / 5
            kms = float(np.mean(ks))
            out[name] = {"kernel": ctx.last_kernel_name(), "kernel_ms": kms, "wall_ms": wall,
                         "GB_per_s": nbytes / (kms * 1e-3) / 1e9,
                         "frac_of_hbm_peak": nbytes / (kms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                         "algorithmic_bytes": nbytes}
        # FP32 errors of every AO output against the FP64 kernels on the same rounded inputs
        log("AO FP64 reference run for the FP32 error")
        R4d = aok.GaussianTypeRadials(aok.as_fortran(orb4.Rnl.zeta.double()), aok.as_fortran(orb4.Rnl.D.double()),
                                      spec4, None, (1, 2, 3))
        orb4d = aok.AtomicOrbitals(R4d, 3, length_unit="bohr")
        Xd = X.double()

        def rel_err(a, b):
            return float((a.double() - b).abs().max() / b.abs().max())

        P32, dP32 = aok.evaluate_ed(orb4, X, sidx=sidx, ctx=ctx)
        P64, dP64 = aok.evaluate_ed(orb4d, Xd, sidx=sidx, ctx=ctx)
        fp32["ao_c4b_P"] = rel_err(P32, P64)
        fp32["ao_c4b_dP"] = rel_err(dP32, dP64)
        del P32, dP32, P64, dP64
        _, _, dX32, dq32, _ = aok.rrule_evaluate(orb4, X, sidx=sidx, ctx=ctx)[1](dPin)
        _, _, dX64, dq64, _ = aok.rrule_evaluate(orb4d, Xd, sidx=sidx, ctx=ctx)[1](aok.as_fortran(dPin.double()))
        fp32["ao_c4b_dX"] = rel_err(dX32, dX64)
        fp32["ao_c4b_dzeta"] = rel_err(dq32["Rnl"]["zeta"], dq64["Rnl"]["zeta"])
        fp32["ao_c4b_dD"] = rel_err(dq32["Rnl"]["D"], dq64["Rnl"]["D"])
        # the first 20000 points through the FP32 kernels, kept on the host for the CPU leg's
        # oracle-anchored check
        ns_ = 20000
        Xs, ss = X[:ns_].contiguous(), sidx[:ns_].contiguous()
        dPs = dPin[:ns_].t().contiguous().t()
        Ps, dPgs = aok.evaluate_ed(orb4, Xs, sidx=ss, ctx=ctx)
        _, _, dXs, dqs, _ = aok.rrule_evaluate(orb4, Xs, sidx=ss, ctx=ctx)[1](dPs)
        host = lambda t: t.double().cpu().numpy()
        subsets["ao"] = {"basis": c4, "X": host(Xs), "sidx": ss.cpu().numpy(), "dP": host(dPs), "P32": host(Ps),
                         "dP32": host(dPgs), "dX32": host(dXs), "dz32": host(dqs["Rnl"]["zeta"]),
                         "dD32": host(dqs["Rnl"]["D"])}
        del P, X, Xd, dPin, dX32, dX64
        # C2: uneven batch, 6 species (union spec N = 50, K = 10, l <= 3), Zipf species pairs
        log("C2 B=2^18 rrule step")
        d2 = configs.c2(2, B=2 ** 18)
        cb2 = d2["basis"]
        spec2, _, _ = cb2.ao_maps()

        def orbital2(dtype):
            R2 = aok.GaussianTypeRadials(aok.fortran_from_numpy(cb2.zeta, dtype, dev),
                                         aok.fortran_from_numpy(cb2.D, dtype, dev), spec2, None, tuple(range(1, 7)))
            return aok.AtomicOrbitals(R2, max(cb2.ls), length_unit="bohr")

        B2 = d2["posA"].shape[0]
        o2 = orbital2(torch.float32)
        p2 = o2.static_params()
        X2A = (torch.as_tensor(d2["posA"], dtype=torch.float32, device=dev).contiguous(),
               torch.as_tensor(d2["sidxA"], dtype=torch.int32, device=dev))
        X2B = (torch.as_tensor(d2["posB"], dtype=torch.float32, device=dev).contiguous(),
               torch.as_tensor(d2["sidxB"], dtype=torch.int32, device=dev))
        g2 = torch.Generator(device=dev).manual_seed(9)
        dS2 = torch.randn((B2, cb2.N, cb2.N), generator=g2, device=dev, dtype=torch.float32).permute(2, 1, 0)
        # launch-shape knobs chosen by the autotuner on this workload (stripe capacity,
        # light/heavy chunk split), then the timed steps with them
        log("C2 autotune")
        tknobs, tms, tres = aok.autotune.tune_overlap(o2, X2A, X2B, p2, dS2, ctx=ctx, save=False, steps=2,
                                                      trials=2)
        for _ in range(2):
            S2, pb2 = aok.rrule_batch_overlap(o2, X2A, X2B, p2, ctx=ctx)
            pb2(dS2)
        torch.cuda.synchronize()
        fk, bk = [], []
        ts = time.perf_counter()
        for _ in range(5):
            S2, pb2 = aok.rrule_batch_overlap(o2, X2A, X2B, p2, ctx=ctx)
            fk.append(ctx.last_kernel_ms())
            pb2(dS2)
            bk.append(ctx.last_kernel_ms())
        torch.cuda.synchronize()
        wall2 = (time.perf_counter() - ts) * 1e3 / 5
        log("C2 graph-replayed step")
        gwall2 = graph_step_ms(o2, X2A, X2B, p2, dS2, 5, tknobs)
        f2f, f2b = os_model_flops(cb2, d2["sidxA"], d2["sidxB"])
        o2d = orbital2(torch.float64)
        S2d = aok.batch_overlap(o2d, (X2A[0].double().contiguous(), X2A[1]), (X2B[0].double().contiguous(), X2B[1]),
                                {"Rnl": {"zeta": aok.as_fortran(o2.Rnl.zeta.double()),
                                         "D": aok.as_fortran(o2.Rnl.D.double())}, "Ylm": {}}, ctx=ctx)
        num = (S2.double() - S2d).abs().amax(dim=(0, 1))
        den = S2d.abs().amax(dim=(0, 1)).clamp_min(1e-300)
        err2 = float((num / den).max())
        fp32["c2_S_blockwise"] = err2
        del S2d
        # C2 (zeta, D) cotangents against the FP64 pullback of the same rounded inputs
        dq32 = pb2(dS2)[4]
        _, pb2d = aok.rrule_batch_overlap(o2d, (X2A[0].double().contiguous(), X2A[1]),
                                          (X2B[0].double().contiguous(), X2B[1]),
                                          {"Rnl": {"zeta": aok.as_fortran(o2.Rnl.zeta.double()),
                                                   "D": aok.as_fortran(o2.Rnl.D.double())}, "Ylm": {}}, ctx=ctx)
        dq64 = pb2d(aok.as_fortran(dS2.double()))[4]
        for k in ("zeta", "D"):
            fp32[f"c2_d{k}"] = float((dq32["Rnl"][k].double() - dq64["Rnl"][k]).abs().max() /
                                     dq64["Rnl"][k].abs().max())
        del pb2d, dq64
        aok.batch_overlap(o2d, (X2A[0].double().contiguous(), X2A[1]), (X2B[0].double().contiguous(), X2B[1]),
                          o2d.static_params(), mode="plain", ctx=ctx)
        plain2 = ctx.last_kernel_ms()
        out["ovl2c_c2"] = {
            "B": B2, "N": cb2.N, "NZ": cb2.NZ, "step_ms": wall2, "pairs_per_s": B2 / (wall2 * 1e-3),
            "graph_step_ms": gwall2, "graph_pairs_per_s": B2 / (gwall2 * 1e-3),
            "autotune": {"knobs": tknobs, "best_ms": tms,
                         "candidates": [[k["ovl_cap"], k["ovl_split"], k["ovl_order"], round(m, 4)] for k, m in tres],
                         "candidate_fields": ["ovl_cap", "ovl_split", "ovl_order", "step_ms"]},
            "fwd_kernel_ms": float(np.mean(fk)), "bwd_kernel_ms": float(np.mean(bk)),
            "fwd_model_TFLOPs": f2f / (np.mean(fk) * 1e-3) / 1e12, "bwd_model_TFLOPs": f2b / (np.mean(bk) * 1e-3) / 1e12,
            "fp32_blockwise_max_rel_err": err2, "plain_f64_fwd_kernel_ms": plain2,
            "fwd_speedup_vs_plain_f64": plain2 / float(np.mean(fk))}
        del S2, dS2, pb2
        ctx.tune("ovl_cap", 0)
        ctx.tune("ovl_split", 0)
        ctx.tune("ovl_order", 0)
        # 3-centre overlaps, B = 2^14 triples over the same 3 species, FP32
        log("3c B=2^14")
        c4 = configs.c4a()
        coef3 = aok.compile_coef(orb32.Rnl.zeta, orb32.Rnl.D, cb.ls, ctx=ctx)
        nbf3 = [(l + 1) * (l + 2) // 2 for l in cb.ls]
        off3 = [0]
        for n in nbf3:
            off3.append(off3[-1] + n)
        basis3 = aok.CartesianGTOBasis(2, cb.ns, cb.K, cb.ls, nbf3, off3, off3[-1], orb32.Rnl.zeta, coef3, zlist, 1.0)
        X3 = [(torch.as_tensor(c4[f"pos{c}"], dtype=torch.float32, device=dev).contiguous(),
               torch.as_tensor(c4[f"sidx{c}"], dtype=torch.int32, device=dev)) for c in "ABC"]
        V = aok.batch_overlap_3c(basis3, *X3, ctx=ctx)
        ks = []
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(5):
            V = aok.batch_overlap_3c(basis3, *X3, ctx=ctx)
            ks.append(ctx.last_kernel_ms())
        torch.cuda.synchronize()
        wall3 = (time.perf_counter() - ts) * 1e3 / 5
        B3 = c4["posA"].shape[0]
        out["ovl3c_c4a"] = {"B": B3, "kernel_ms": float(np.mean(ks)), "wall_ms": wall3,
                            "triples_per_s": B3 / (wall3 * 1e-3), "kernel_triples_per_s": B3 / (np.mean(ks) * 1e-3)}
        basis3d = aok.CartesianGTOBasis(2, cb.ns, cb.K, cb.ls, nbf3, off3, off3[-1],
                                        aok.as_fortran(orb32.Rnl.zeta.double()), aok.as_fortran(coef3.double()),
                                        zlist, 1.0)
        X3d = [(x[0].double().contiguous(), x[1]) for x in X3]
        V64 = aok.batch_overlap_3c(basis3d, *X3d, ctx=ctx)
        num3 = (V.double() - V64).abs().amax(dim=(0, 1, 2))
        fp32["c4a_V_blockwise"] = float((num3 / V64.abs().amax(dim=(0, 1, 2)).clamp_min(1e-300)).max())
        del V64
        # the 3-centre (coef, alpha) pullback of a random cotangent at the same size
        log("3c pullback B=2^14")
        g3 = torch.Generator(device=dev).manual_seed(11)
        dT = torch.randn(tuple(reversed(V.shape)), generator=g3, device=dev, dtype=torch.float32).permute(3, 2, 1, 0)
        aok.overlap_3c_pb_coef_alpha(dT, basis3, *X3, ctx=ctx)
        ks = []
        torch.cuda.synchronize()
        ts = time.perf_counter()
        for _ in range(3):
            aok.overlap_3c_pb_coef_alpha(dT, basis3, *X3, ctx=ctx)
            ks.append(ctx.last_kernel_ms())
        torch.cuda.synchronize()
        wall3b = (time.perf_counter() - ts) * 1e3 / 3
        # FP32 against the FP64 pullback of the same FP32-rounded inputs, relative to the
        # largest |cotangent| of each output
        dc32, da32 = aok.overlap_3c_pb_coef_alpha(dT, basis3, *X3, ctx=ctx)
        dc64, da64 = aok.overlap_3c_pb_coef_alpha(aok.as_fortran(dT.double()), basis3d, *X3d, ctx=ctx)
        err3 = max(float((dc32.double() - dc64).abs().max() / dc64.abs().max()),
                   float((da32.double() - da64).abs().max() / da64.abs().max()))
        out["ovl3c_c4a_pullback"] = {"B": B3, "kernel": "ovl3c_bwd_rows", "kernel_ms": float(np.mean(ks)),
                                     "wall_ms": wall3b, "triples_per_s": B3 / (wall3b * 1e-3),
                                     "fp32_max_rel_err": err3}
        fp32["c4a_dcoef_dalpha"] = err3
        # (zeta, D) cotangents through the normalisation VJP, against FP64 on the same inputs
        _, pb3 = aok.rrule_batch_overlap_3c(orb32, *X3, orb32.static_params(), ctx=ctx)
        dq3 = pb3(dT)[5]
        _, pb3d = aok.rrule_batch_overlap_3c(orb64, *X3d, ps64, ctx=ctx)
        dq3d = pb3d(aok.as_fortran(dT.double()))[5]
        for k in ("zeta", "D"):
            fp32[f"c4a_d{k}"] = float((dq3["Rnl"][k].double() - dq3d["Rnl"][k]).abs().max() /
                                      dq3d["Rnl"][k].abs().max())
        del pb3, pb3d
        # the first 256 triples through the FP32 kernels, kept for the CPU leg's oracle check
        n3 = 256
        X3s = [(x[0][:n3].contiguous(), x[1][:n3].contiguous()) for x in X3]
        dTs = dT[..., :n3]
        V3s = aok.batch_overlap_3c(basis3, *X3s, ctx=ctx)
        dc3s, da3s = aok.overlap_3c_pb_coef_alpha(dTs, basis3, *X3s, ctx=ctx)
        h = lambda t: t.double().cpu().numpy()
        subsets["c4a"] = {"pos": [h(x[0]) for x in X3s], "sidx": [x[1].cpu().numpy() for x in X3s], "ls": cb.ls,
                          "coef": h(coef3), "alpha": h(orb32.Rnl.zeta), "dT": h(dTs), "V32": h(V3s),
                          "dc32": h(dc3s), "da32": h(da3s)}
        del V, dT, dc64, da64

    if not args.no_cpu:
        # The reference's KernelAbstractions CPU backend cannot run here (Julia is absent); the
        # baseline is the FP64 C restatement of the same kernels, timed on the host cores with
        # OpenMP over the (b, s_a, s_b) work items (KA's CPU split) and on one thread.
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O

        nthr = O.num_threads()
        log(f"CPU baseline (C oracle, FP64, {nthr} threads and 1 thread)")

        def r32(a):
            return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)

        zr, Dr = r32(cb.zeta), r32(cb.D)
        pAr, pBr = r32(d["posA"]), r32(d["posB"])
        coef = O.compile_coef(zr, Dr, cb.ls)

        def cpu_step(nb, omp):
            dSn = r32(np.random.default_rng(0).standard_normal((cb.N, cb.N, nb)))
            ts = time.perf_counter()
            Sc = O.ovl2c(pAr[:nb], pBr[:nb], d["sidxA"][:nb], d["sidxB"][:nb], cb.ls, coef, zr, omp=omp)
            gc = O.pullback_ps_2c(dSn, pAr[:nb], pBr[:nb], d["sidxA"][:nb], d["sidxB"][:nb], cb.ls, zr, Dr, omp=omp)
            return time.perf_counter() - ts, Sc, gc, dSn

        nb_mt, nb_st = 32768, 4096
        t_mt, S_c, g_c, dS_c = cpu_step(nb_mt, True)
        t_st, _, _, _ = cpu_step(nb_st, False)
        out["cpu_baseline"] = {
            "value": nb_mt / t_mt, "unit": "pairs/s", "cores": nthr, "kind": "port",
            "cpu_model": _cpu_model(),
            "sample": f"first {nb_mt} of the {B} C1 pairs: forward + (zeta,D) pullback, FP64 C restatement of "
                      f"the reference kernels, OpenMP static over (b, s_a, s_b) on {nthr} threads ({t_mt:.1f} s)",
            "single_thread": {"value": nb_st / t_st, "unit": "pairs/s", "cores": 1,
                              "sample": f"first {nb_st} pairs, same work, serial ({t_st:.1f} s)"}}
        # the same pairs on the GPU (FP32 fast path) against these oracle outputs: the bench's
        # oracle-anchored error check (the FP64-kernel comparisons above cover the full sizes)
        sub = slice(0, nb_mt)
        XAs = (XA32[0][sub].contiguous(), XA32[1][sub].contiguous())
        XBs = (XB32[0][sub].contiguous(), XB32[1][sub].contiguous())
        Ss, pbs = aok.rrule_batch_overlap(orb32, XAs, XBs, ps32, ctx=ctx)
        gs = pbs(aok.fortran_from_numpy(dS_c, torch.float32, dev))[4]["Rnl"]
        Sh = Ss.double().cpu().numpy()
        errS = float(np.max(np.max(np.abs(Sh - S_c), axis=(0, 1)) / np.maximum(np.max(np.abs(S_c), axis=(0, 1)),
                                                                                    1e-300)))
        errz = float(np.max(np.abs(gs["zeta"].double().cpu().numpy() - g_c[0])) / np.max(np.abs(g_c[0])))
        errD = float(np.max(np.abs(gs["D"].double().cpu().numpy() - g_c[1])) / np.max(np.abs(g_c[1])))
        out.setdefault("fp32_error", {})["c1_vs_oracle_subset"] = {
            "pairs": nb_mt, "S_blockwise": errS, "dzeta": errz, "dD": errD}

        def rel(a, b):
            return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

        if "ao" in subsets:  # AO C4b: first 20000 points, FP32 GPU against the FP64 oracle
            a = subsets["ao"]
            c4o = a["basis"]
            _, iR4, iY4 = c4o.ao_maps()
            z4, D4 = r32(c4o.zeta), r32(c4o.D)
            poly4 = np.zeros(c4o.ns, np.int64)
            Pr, dPr = O.ao_eval(a["X"], a["sidx"], z4, D4, poly4, np.array(iR4), np.array(iY4), 3, grad=True)
            dXr, dzr, dDr = O.ao_pullback(a["X"], a["sidx"], z4, D4, poly4, np.array(iR4), np.array(iY4), 3, a["dP"])
            out["fp32_error"]["ao_c4b_vs_oracle_subset"] = {
                "points": len(a["sidx"]), "P": rel(a["P32"], Pr),
                "dP": max(rel(a["dP32"][:, :, c], dPr[:, :, c]) for c in range(3)),
                "dX": rel(a["dX32"], dXr), "dzeta": rel(a["dz32"], dzr), "dD": rel(a["dD32"], dDr)}
        if "c4a" in subsets:  # 3c C4a: first 256 triples, FP32 GPU against the FP64 oracle
            t = subsets["c4a"]
            pa, pb, pc = t["pos"]
            sa_, sb_, sc_ = t["sidx"]
            Vr = O.ovl3c(pa, pb, pc, sa_, sb_, sc_, t["ls"], t["coef"], t["alpha"])
            dcr, dar = O.ovl3c_pb(t["dT"], pa, pb, pc, sa_, sb_, sc_, t["ls"], t["coef"], t["alpha"])
            num = np.max(np.abs(t["V32"] - Vr), axis=(0, 1, 2))
            den = np.maximum(np.max(np.abs(Vr), axis=(0, 1, 2)), 1e-300)
            out["fp32_error"]["c4a_vs_oracle_subset"] = {
                "triples": len(sa_), "V_blockwise": float(np.max(num / den)),
                "dcoef": rel(t["dc32"], dcr), "dalpha": rel(t["da32"], dar)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
