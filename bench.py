#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X.

Workload (one step): a 3D type-1 and a 3D type-2 F32 transform at tol 1e-6
over 128^3 modes (fine grid 256^3) and M = 10^7 points, half uniform in
[-pi, pi)^3 and half a Gaussian cluster (sigma 0.05) around (0.3, -1.1, 2.0)
wrapped into the box, strengths / coefficients complex N(0,1).  Each
transform includes its own point binning (encode = set_points + execute),
the rocFFT fine-grid FFT and the (pre)deconvolution.  Inputs are resident
in HBM before timing starts.

value = M / (seconds per step): points per second through both transforms,
with the type-1 spread on the matrix cores (FP16 hi/lo split products,
FP32 accumulation) and bins summed in coordinate order, so the type-1 output
is bitwise identical for any ordering of the points (checked on one shuffle,
"bitwise"); the same step with the plain FP32 spreader ("plain") and with
the reference's index order ("index_order") is timed beside it.  The roofline object reports the dominant kernel of the
path, from HIP events
around that kernel over the timed region; cpu_baseline times the CPU port of
the reference's CpuNufftPlan on a bounded sample of the same workload.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "NU points/s for 3D type-1/2 at tol 1e-6, MFMA vs plain GPU; % HBM/MFMA roofline"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
FP16_MFMA_PEAK_TFLOPS = 2500.0
# HBM bytes per launch of the dominant kernels on the default workload,
# from the FETCH_SIZE / WRITE_SIZE counter passes over `bench.py
# --profile-step` (scaled by the same counters over a 1 GiB copy);
# profiles/traffic_3d_t1t2_128modes_1e7pts.md holds the table.  Reported
# only when the run uses the default workload.
MEASURED_TRAFFIC = {
    "mfma_spread": 2.5836e9,
    "k_interp_tiled3": 2.4193e9,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--modes", type=int, default=128)
    ap.add_argument("--eps", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error check")
    ap.add_argument("--cpu-points", type=int, default=0,
                    help="points of the CPU baseline sample (0: the whole workload)")
    ap.add_argument("--cpu-reps", type=int, default=5,
                    help="executes of each transform the CPU baseline takes the median of")
    ap.add_argument("--spread", default="mfma_f16x3",
                    choices=["mfma_f16x3", "mfma_bf16x3", "mfma_f16"],
                    help="matrix-core spread of the timed type-1 (the plain FP32 spread is "
                         "timed beside it)")
    ap.add_argument("--profile-step", action="store_true",
                    help="for counter passes: a 1 GiB device copy (the byte-counter "
                         "calibration) then two untimed steps of the headline workload")
    ap.add_argument("--config", type=int, default=2,
                    help="2: the headline workload (default); 1, 3, 4, 5: the other "
                         "configurations, reported as extra JSON lines")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: there is no multi-GPU path (type 1 would "
              "shard points with one reduce-scatter of the fine grid, type 2 points with a "
              "replicated grid)", file=sys.stderr)
        return 2
    import numpy as np
    import torch

    import wgpunufft_amd as W
    from wgpunufft_amd import synthetic as S

    if args.config != 2:
        from wgpunufft_amd import bench_configs
        return bench_configs.run(args)
    if args.profile_step:
        return profile_step(args, W, S, np, torch)

    dev = torch.device("cuda:0")
    M = args.points
    N = args.modes
    n_modes = [N, N, N]
    t0 = time.time()
    x = S.config2_points(M)
    c = S.complex_normal(S.config_seed(2, 3), M)
    fm = S.complex_normal(S.config_seed(2, 4), N ** 3)
    xt = torch.from_numpy(x).to(dev)
    ct = torch.from_numpy(c).to(dev)
    ft = torch.from_numpy(fm).to(dev)
    f_out = torch.empty(N ** 3, dtype=torch.complex64, device=dev)
    f_plain = torch.empty(N ** 3, dtype=torch.complex64, device=dev)
    c_out = torch.empty(M, dtype=torch.complex64, device=dev)
    cfg = W.NufftConfig(n_modes, args.eps)
    # the workload asks for bitwise-identical type-1 output across point
    # orderings: bins are summed in coordinate order
    p1 = W.NufftPlan.type1_gpu(cfg).set_point_order("coordinates")
    p1m = W.NufftPlan.type1_gpu(cfg).set_spread_method(args.spread).set_point_order("coordinates")
    p1i = W.NufftPlan.type1_gpu(cfg).set_spread_method(args.spread)
    p2 = W.NufftPlan.type2_gpu(cfg)
    print(f"# data ready in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)

    def leg(t1_plan, f_dst):
        """Times K steps (type 1 through t1_plan, then type 2); returns ms per
        step and the per-stage milliseconds averaged over the K steps."""
        def step():
            t1_plan.encode_gpu(xt, ct, out=f_dst)
            p2.encode_gpu(xt, ft, out=c_out)

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t1_plan.timing_history(True)
        p2.timing_history(True)
        torch.cuda.synchronize()
        t_start = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t_end = time.perf_counter()
        st1, st2 = t1_plan.stage_times(), p2.stage_times()
        t1_plan.timing_history(False)
        p2.timing_history(False)
        K = args.steps
        stages = {"type1": {k: (None if v is None else round(v / K, 4)) for k, v in st1.items()},
                  "type2": {k: (None if v is None else round(v / K, 4)) for k, v in st2.items()}}
        return (t_end - t_start) * 1e3 / K, stages

    ms_plain, stages_plain = leg(p1, f_plain)
    print(f"# plain leg {ms_plain:.3f} ms/step", file=sys.stderr, flush=True)
    f_index = torch.empty_like(f_out)
    ms_index, stages_index = leg(p1i, f_index)
    print(f"# {args.spread} index-order leg {ms_index:.3f} ms/step", file=sys.stderr, flush=True)
    ms_mfma, stages_mfma = leg(p1m, f_out)
    print(f"# {args.spread} leg {ms_mfma:.3f} ms/step", file=sys.stderr, flush=True)
    # bitwise repeat and point-order independence of the timed type 1
    f_rep = p1m.encode_gpu(xt, ct)
    perm = torch.from_numpy(S.shuffle_perm(S.config_seed(2, 9), M)).to(dev)
    f_shuf = p1m.encode_gpu(xt[perm].contiguous(), ct[perm].contiguous())
    torch.cuda.synchronize()
    bits = lambda t: t.view(torch.float32).view(torch.int32)
    bitwise = {"repeat": bool(torch.equal(bits(f_rep), bits(f_out))),
               "shuffled_points": bool(torch.equal(bits(f_shuf), bits(f_out)))}
    del f_rep, f_shuf, perm

    info1 = p1.info()
    w = info1["width"]
    nf = int(np.prod(info1["fine"]))
    sweep = info1["run"] == 0
    tiled = p2.info()["bin_side"] == [8, 8, 8]
    # algorithmic work per launch (one vector)
    spread_flops = 4.0 * w ** 3 * M
    interp_bytes = M * (4 * 3 + 8) + 8 * nf

    def spread_roof(t_ms, name, peak, unit_note, bound="mfma"):
        tf = spread_flops / (t_ms * 1e-3) / 1e12
        return {"kernel": name, "bound": bound, "achieved": round(tf, 3), "peak": peak,
                "unit": "TFLOP/s", "frac": round(tf / peak, 5), "traffic": None,
                "avg_ms": round(t_ms, 4),
                "algorithmic": f"4*w^3*M = {spread_flops:.4g} flop (w={w}, M={M}); {unit_note}"}

    t_mfma = stages_mfma["type1"]["spread_kernel"]
    t_plain = stages_plain["type1"]["spread_kernel"]
    t_interp = stages_mfma["type2"]["interp"]
    roof_mfma = spread_roof(t_mfma, f"mfma_spread (type-1 matrix-core spread, {args.spread}: "
                                    "k_spread_mfma3 16x16x8 tiles of the light bins, "
                                    "k_reduce_tiles, k_dense_windows + k_dense_reduce + "
                                    "k_dense_add for the dense bins; one event span)",
                            FP16_MFMA_PEAK_TFLOPS, "peak: dense FP16 matrix cores")
    tile = "x".join(str(v) for v in info1["block"])
    roof_plain = spread_roof(t_plain,
                             f"k_spread_rows (type-1 plain FP32 row tiles, {tile} cells)" if sweep
                             else "k_spread (type-1 output-stationary spread)",
                             FP32_PEAK_TFLOPS, "peak: FP32 vector", bound="fp32-vector")
    interp_gbs = interp_bytes / (t_interp * 1e-3) / 1e9
    roof_interp = {"kernel": "k_interp_tiled3 (type-2 LDS-tiled interpolation)" if tiled
                   else "k_interp (type-2 interpolation)", "bound": "hbm",
                   "achieved": round(interp_gbs, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                   "frac": round(interp_gbs / HBM_PEAK_GBS, 4), "traffic": None,
                   "avg_ms": round(t_interp, 4),
                   "algorithmic": f"M*(4d+8)+8*N_f = {interp_bytes:.4g} B"}
    def sort_roof(t_ms, name, nbins):
        # a bin sort reads every coordinate once and writes the permutation
        # and the bin offsets once; the radix passes move more than that
        b = M * (4 * 3 + 4) + 4 * (nbins + 1)
        gbs = b / (t_ms * 1e-3) / 1e9
        return {"kernel": name, "bound": "hbm", "achieved": round(gbs, 2), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": None,
                "avg_ms": round(t_ms, 4),
                "algorithmic": f"M*(4d+4)+4*(nbins+1) = {b:.4g} B (nbins={nbins})"}

    def key_bits(nbins):
        return max(1, int(nbins - 1).bit_length())

    info2 = p2.info()
    # type 1 in coordinate order: bin key, then 5 bits of in-bin position per
    # axis; type 2: bin key, then the start cell inside the bin
    bits1 = key_bits(info1["nbins"]) + 3 * min(5, (64 - key_bits(info1["nbins"])) // 3)
    bits2 = key_bits(info2["nbins"]) + sum(int(s).bit_length() - 1 for s in info2["bin_side"])
    roof_sort1 = sort_roof(stages_mfma["type1"]["sort"],
                           f"type-1 bin sort (stage: radix keys, {-(-bits1 // 8)} passes of "
                           "64-bit keys, bin offsets, large-group merges)",
                           info1["nbins"])
    roof_sort2 = sort_roof(stages_mfma["type2"]["sort"],
                           f"type-2 bin sort (stage: radix keys, {-(-bits2 // 8)} passes of "
                           "32-bit keys, bin offsets)", info2["nbins"])
    default_workload = (M, N, args.eps) == (10_000_000, 128, 1e-6)
    for roof in (roof_mfma, roof_interp):
        key = roof["kernel"].split(" ")[0]
        if (default_workload and MEASURED_TRAFFIC.get(key) is not None
                and args.spread == "mfma_f16x3"):
            roof["traffic"] = MEASURED_TRAFFIC[key]
            roof["traffic_unit"] = "B per launch (FETCH_SIZE+WRITE_SIZE, copy-calibrated)"
    dominant, other = (roof_mfma, roof_interp) if t_mfma >= t_interp else (roof_interp, roof_mfma)

    def pts(stages, key):
        ms = sum(v for k, v in stages[key].items() if v is not None and k != "spread_kernel")
        return M / (ms * 1e-3) if ms else None

    result = {
        "metric": METRIC,
        "value": M / (ms_mfma * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_mfma, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded SplitMix64)",
        "config": {"workload": f"3D type-1 + type-2 F32, {N}^3 modes, sigma 2, tol {args.eps:g}, "
                               f"M={M} (half uniform, half Gaussian sigma 0.05 cluster)",
                   "n_modes": n_modes, "fine": info1["fine"], "width": w,
                   "points": M, "bin_geometry_t1": info1["bin_side"],
                   "spreader": f"{args.spread} matrix-core tiles (type 1)",
                   "point_order": "coordinates (type-1 output independent of point order)",
                   "interp": "LDS-tiled" if tiled else "per-point gather"},
        "bitwise": bitwise,
        "roofline": dominant,
        "roofline_other": [other, roof_plain, roof_sort1, roof_sort2],
        "stages_ms": stages_mfma,
        "type1_points_per_s": pts(stages_mfma, "type1"),
        "type2_points_per_s": pts(stages_mfma, "type2"),
        "plain": {"value": M / (ms_plain * 1e-3), "ms_per_step": round(ms_plain, 4),
                  "spreader": f"plain FP32 row tiles ({tile})" if sweep
                              else "plain FP32 output-stationary",
                  "stages_ms": stages_plain,
                  "type1_points_per_s": pts(stages_plain, "type1")},
        "index_order": {"value": M / (ms_index * 1e-3), "ms_per_step": round(ms_index, 4),
                        "note": "same step with bins summed in ascending point index (the "
                                "reference's order; bitwise repeatable, not point-order "
                                "independent)",
                        "stages_ms": stages_index},
        "mfma_vs_plain": {"spread_kernel": round(t_plain / t_mfma, 3),
                          "type1": round(pts(stages_mfma, "type1") / pts(stages_plain, "type1"), 3),
                          "step": round(ms_plain / ms_mfma, 3)},
    }

    if not args.no_check:
        result["error"] = accuracy(W, S, np, torch, dev, x, c, fm, n_modes, f_out, c_out)
        result["plain"]["error"] = accuracy(W, S, np, torch, dev, x, c, fm, n_modes, f_plain,
                                            c_out)
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args, np, S, n_modes)
    print(json.dumps(result), flush=True)
    return 0


def profile_step(args, W, S, np, torch):
    """A 1 GiB device-to-device copy, whose FETCH_SIZE / WRITE_SIZE give the
    counters' byte scale, then two steps (matrix-core type 1 and type 2) of
    the headline workload; nothing is timed."""
    dev = torch.device("cuda:0")
    src = torch.empty(1 << 28, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst
    M, N = args.points, args.modes
    x = torch.from_numpy(S.config2_points(M)).to(dev)
    c = torch.from_numpy(S.complex_normal(S.config_seed(2, 3), M)).to(dev)
    fm = torch.from_numpy(S.complex_normal(S.config_seed(2, 4), N ** 3)).to(dev)
    cfg = W.NufftConfig([N] * 3, args.eps)
    p1 = W.NufftPlan.type1_gpu(cfg).set_spread_method(args.spread).set_point_order("coordinates")
    p2 = W.NufftPlan.type2_gpu(cfg)
    for _ in range(2):
        p1.encode_gpu(x, c)
        p2.encode_gpu(x, fm)
    torch.cuda.synchronize()
    print("profile step done", flush=True)
    return 0


def accuracy(W, S, np, torch, dev, x, c, fm, n_modes, f_out, c_out):
    """Relative l2 error against FP64 direct sums on 2,000 seeded modes
    (type 1) and 2,000 seeded points (type 2), on the GPU direct kernel."""
    nm = int(np.prod(n_modes))
    sel_m = np.sort(S.splitmix64(S.config_seed(2, 7), 2000) % np.uint64(nm)).astype(np.int64)
    sel_p = np.sort(S.splitmix64(S.config_seed(2, 8), 2000) % np.uint64(len(x))).astype(np.int64)
    xd = torch.from_numpy(x.astype(np.float64)).to(dev)
    cd = torch.from_numpy(c.astype(np.complex128)).to(dev)
    fd = torch.from_numpy(fm.astype(np.complex128)).to(dev)
    r1 = W.reference_type1_f64(n_modes, xd, cd, modes=torch.from_numpy(sel_m).to(dev))
    r2 = W.reference_type2_f64(n_modes, xd, fd, points=torch.from_numpy(sel_p).to(dev))
    torch.cuda.synchronize()
    g1 = f_out.cpu().numpy()[sel_m].astype(np.complex128)
    g2 = c_out.cpu().numpy()[sel_p].astype(np.complex128)
    e1 = float(np.linalg.norm(g1 - r1.cpu().numpy()) / np.linalg.norm(r1.cpu().numpy()))
    e2 = float(np.linalg.norm(g2 - r2.cpu().numpy()) / np.linalg.norm(r2.cpu().numpy()))
    return {"type1_rel_l2": e1, "type2_rel_l2": e2, "sample": "2000 modes / 2000 points"}


def cpu_baseline(args, np, S, n_modes):
    """CPU port of the reference's CpuNufftPlan (oracle/cpu_nufft.c), timed
    on this host's cores on the same workload (its first cpu_points points
    when a sample is asked for): medians of cpu_reps executes of type 1
    (split into its point sort, prepare_points, and the rest,
    execute_prepared) and of type 2 (src/cpu.rs; benches/nufft_cpu_bench.rs
    of the reference reports the same three figures)."""
    try:
        from oracle import cpu_nufft
    except Exception as e:  # the port is optional for the GPU numbers
        return {"value": None, "unit": "points/s", "error": str(e)}
    m = args.cpu_points or args.points
    full = S.config2_points(args.points)
    half = args.points // 2
    # keep the workload's uniform / clustered mix
    x = full if m == args.points else np.concatenate([full[: m // 2],
                                                      full[half: half + m - m // 2]])
    c = S.complex_normal(S.config_seed(2, 3), m)
    fm = S.complex_normal(S.config_seed(2, 4), int(np.prod(n_modes)))
    threads = cpu_nufft.threads()
    st1 = cpu_nufft.time_type1_stages(x, c, n_modes, args.eps, threads, reps=args.cpu_reps)
    t2 = cpu_nufft.time_type2(x, fm, n_modes, args.eps, threads, reps=args.cpu_reps)
    t1 = st1["execute_s"]
    sample = ("the whole workload" if m == args.points else
              f"first {m // 2} uniform + {m - m // 2} clustered points of the workload")
    out = {"value": m / (t1 + t2), "unit": "points/s", "cores": threads, "kind": "port",
           "sample": f"{sample}, same 128^3 grid; median of {args.cpu_reps} executes each of "
                     f"type 1 and type 2",
           "threads_note": "threads = the job's CPU share on this host (OMP_NUM_THREADS), "
                           "at most every online CPU",
           "type1_s": t1, "type1_prepare_s": st1["prepare_s"],
           "type1_execute_prepared_s": st1["execute_prepared_s"], "type2_s": t2,
           "type1_points_per_s": m / t1, "type2_points_per_s": m / t2}
    out.update(cpu_nufft.host_facts())
    return out


if __name__ == "__main__":
    sys.exit(main())
