"""Throughput of the correlator-node data path on one MI355X.

Main workload (N=1, `value`): 8 stations x 2 circular polarisations, 8
subbands of 32 MHz sampled at 64 Ms/s, 2-bit, 4096 channels, 1 s integrations,
delay offsets U[-5, 5] ms and rates U[-3e-6, 3e-6], plain FP32 X-step. One
step = one 1 s integration of all 8 subbands (8 slices x 16 streams x 64
Msamples = 8.19e9 station-samples); --steps 10 covers the full 10 s job.

Also reported: the same workload with the mixed-precision (split-FP16 MFMA)
X-step, and the 32-station workload (32 x 2 pols, 4 subbands of 64 MHz at 128
Ms/s, 16384 channels, 0.5 s integrations) with the plain FP32 and the
mixed-precision X-steps and their ratio.

The HEADER control records come from the input-node delay rule on the host;
the 2-bit payload is synthesised straight into device memory, so the timed
region starts with the input resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "sfxc-1_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP16_DENSE_PEAK_TFLOPS = 2500.0
KERNELS = ["stats", "fengine", "channelise", "xstep", "normalise"]
PREC = {"fp32": 0, "fp16": 1, "bf16": 2, "split_fp16": 3}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subbands", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-usec", type=int, default=125000)
    ap.add_argument("--contexts", type=int, default=2, help="contexts (HIP streams) slices alternate over")
    ap.add_argument("--main-only", action="store_true", help="skip the mixed-precision and 32-station runs")
    ap.add_argument("--c2-steps", type=int, default=2, help="0.5 s integrations of the 32-station workload")
    return ap.parse_args()


def kernel_bytes(job, M, nseg, n_samples, nfold, B, fused):
    """Algorithmic HBM bytes per launch of each kernel. fengine: unpack +
    delay correction (+ channeliser when fused: payload in, spectra out)."""
    nc = job.fft_size_correlation
    bits = job.station_streams[0].bits_per_sample
    spectra = M * nseg * (nc + 1) * 8
    return {
        "stats": M * n_samples * bits / 8,
        "fengine": M * n_samples * bits / 8 + (spectra if fused else M * n_samples * 4),
        "channelise": M * n_samples * 4 + spectra,
        "xstep": spectra + nfold * B * (nc + 1) * 16,
        "normalise": B * (nc + 1) * 8 * 2,
    }


def cpu_baseline(sample_usec):
    """The FP32 oracle (C restatement, one core) on one slice of the main
    workload shortened to sample_usec."""
    from oracle import oracle as O
    from sfxg import configs, synth

    jobs, raw_start = configs.config1_jobs(0, n_integrations=1, t_int_usec=sample_usec)
    job = jobs[0]
    rng = np.random.default_rng(7)
    hdrs = []
    nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
    for i, st in enumerate(job.station_streams):
        ctrl, off, n = synth.header_stream(1 << 40, raw_start, job.delay_models[i], job.start_time, nsamp,
                                           st.sample_rate, st.bits_per_sample, elided=True)
        payload = rng.integers(0, 256, n, dtype=np.uint8)
        hdrs.append(synth.join_elided(ctrl, payload))
    O.lib("f32")
    t0 = time.perf_counter()
    O.run_slice(job, hdrs, lut=0)
    dt = time.perf_counter() - t0
    ss = len(job.station_streams) * nsamp
    return {"value": ss / dt / 1e9, "unit": "station-Gsamples/s", "cores": 1, "kind": "port",
            "sample": f"one {sample_usec / 1e6:g} s slice of subband 0 ({len(job.station_streams)} streams, "
                      f"{ss:.3g} station-samples), FP32 oracle, {dt:.1f} s"}


class Workload:
    """Device-resident input of n_subbands x K slices."""

    def __init__(self, torch, capi, synth, jobs_fn, K, NSB, dev, seed_base):
        t0 = time.perf_counter()
        self.per_sb = [jobs_fn(sb, K) for sb in range(NSB)]
        self.K, self.NSB = K, NSB
        self.inputs = {}
        for sb in range(NSB):
            jobs, raw_start = self.per_sb[sb]
            for k, job in enumerate(jobs):
                nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
                lst = []
                for i, st in enumerate(job.station_streams):
                    ctrl, off, n = synth.header_stream(1 << 40, raw_start, job.delay_models[i], job.start_time,
                                                       nsamp, st.sample_rate, st.bits_per_sample, elided=True)
                    buf = torch.empty(n, dtype=torch.uint8, device=f"cuda:{dev}")
                    n_int = int(synth.integer_delay(job.delay_models[i], job.start_time, st.sample_rate))
                    capi.synthesize_raw(dev, buf.data_ptr(), n, st.bits_per_sample, off * 4, n_int,
                                        common_seed=seed_base + 2 * sb + (st.polarisation == "L"),
                                        stream_seed=(seed_base + sb * 64 + i) * 4096 + k, rho=0.02)
                    lst.append((ctrl, buf, n))
                self.inputs[(k, sb)] = lst
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0
        job = self.per_sb[0][0][0]
        self.job = job
        self.M = len(job.station_streams)
        self.nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
        self.station_samples_per_step = NSB * self.M * self.nsamp

    def run(self, torch, capi, precision, W, n_ctx, dev):
        ctxs = [capi.Context(dev) for _ in range(max(1, n_ctx))]
        for c in ctxs:
            c.set_options(lut=1, precision=precision)
        counter = [0]

        def run_slice(k, sb, ctx=None):
            ctx = ctx or ctxs[counter[0] % len(ctxs)]
            counter[0] += 1
            ctx.set_parameters(self.per_sb[sb][0][k])
            for i, (ctrl, buf, n) in enumerate(self.inputs[(k, sb)]):
                ctx.push_stream_elided(i, ctrl, buf.data_ptr(), n)
            ctx.correlate()

        for w in range(W):
            for sb in range(self.NSB):
                run_slice(w % self.K, sb)
        for c in ctxs:
            c.synchronize()
            c.kernel_times(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(self.K):
            for sb in range(self.NSB):
                run_slice(k, sb)
        for c in ctxs:
            c.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for c in ctxs:
            c.kernel_times(reset=True)
        # kernel durations without overlap: one step again on one context,
        # each slice waited for
        for sb in range(self.NSB):
            run_slice(0, sb, ctxs[0])
            ctxs[0].synchronize()
        kt = ctxs[0].kernel_times(reset=True)
        for c in ctxs:
            c.close()
        return wall, kt


def summarise(wl, wall, kt):
    job = wl.job
    nseg = job.number_ffts_in_integration()
    B = len(job.baselines())
    nfold = max(1, round(job.integration_time / job.sub_integration_time))
    fused = kt[2] == 0.0
    kb = kernel_bytes(job, wl.M, nseg, wl.nsamp, nfold, B, fused)
    value = wl.K * wl.station_samples_per_step / wall / 1e9
    xflops = 8.0 * B * (job.fft_size_correlation + 1) * nseg
    return {
        "value": round(value, 4),
        "ms_per_step": round(wall / wl.K * 1e3, 3),
        "kernel_ms_isolated": {n: round(float(a), 3) for n, a in zip(KERNELS, kt)},
        "kernel_GBs": {n: round(kb[n] / (a * 1e-3) / 1e9, 1) for n, a in zip(KERNELS, kt) if a > 0},
        "xstep_TFLOPs": round(xflops / (kt[3] * 1e-3) / 1e12, 2) if kt[3] > 0 else None,
        "_kb": kb,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this pool has one GPU and the path shards by "
              "(slice x subband) with no collectives; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from sfxg import capi, configs, synth

    dev = 0
    torch.cuda.set_device(dev)
    K, W, NSB = args.steps, args.warmup, args.subbands
    wl = Workload(torch, capi, synth, lambda sb, k: configs.config1_jobs(sb, n_integrations=k), K, NSB, dev, 1000)
    print(f"setup done: {len(wl.inputs)} slices in {wl.setup_s:.1f} s", file=sys.stderr, flush=True)
    wall, kt = wl.run(torch, capi, PREC["fp32"], W, args.contexts, dev)
    main_r = summarise(wl, wall, kt)
    print(f"plain FP32 done: {main_r['value']:.2f} station-Gsamples/s", file=sys.stderr, flush=True)
    dom = int(np.argmax(kt))
    dom_name = KERNELS[dom]
    achieved = main_r["_kb"][dom_name] / (kt[dom] * 1e-3) / 1e9
    line = {
        "metric": "station-Gsamples/s correlated (plain FP32 GPU path)",
        "value": main_r["value"],
        "unit": "station-Gsamples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded 2-bit station data generated in device memory, HEADER records from the "
                "input-node delay rule",
        "config": {"workload": "8 stations x 2 pol, 8 subbands x 32 MHz at 64 Ms/s, 2-bit, 4096 channels, "
                               "1 s integrations, 10 s", "stations": 8, "polarisations": 2, "subbands": NSB,
                   "sample_rate": 64_000_000, "channels": 4096, "integration_s": 1.0,
                   "station_samples_per_step": wl.station_samples_per_step},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "kernel_ms_isolated": main_r["kernel_ms_isolated"],
        "kernel_GBs": main_r["kernel_GBs"],
        "setup_s": round(wl.setup_s, 1),
    }
    if not args.main_only:
        wall, kt = wl.run(torch, capi, PREC["split_fp16"], W, args.contexts, dev)
        mixed = summarise(wl, wall, kt)
        mixed.pop("_kb")
        mixed["xstep_mfma_frac"] = round(mixed["xstep_TFLOPs"] / FP16_DENSE_PEAK_TFLOPS, 5) \
            if mixed["xstep_TFLOPs"] else None
        line["mixed_split_fp16"] = mixed
        print(f"mixed done: {mixed['value']:.2f}", file=sys.stderr, flush=True)
        del wl
        torch.cuda.empty_cache()
        w2 = Workload(torch, capi, synth, lambda sb, k: configs.config2_jobs(sb, n_integrations=k), args.c2_steps, 4,
                      dev, 2000)
        print(f"32-station setup done in {w2.setup_s:.1f} s", file=sys.stderr, flush=True)
        c2 = {"workload": "32 stations x 2 pol, 4 subbands x 64 MHz at 128 Ms/s, 2-bit, 16384 channels, "
                          f"0.5 s integrations, {args.c2_steps * 0.5:g} s",
              "station_samples_per_step": w2.station_samples_per_step}
        for name in ("fp32", "split_fp16", "fp16"):
            wall, kt = w2.run(torch, capi, PREC[name], 0, args.contexts, dev)
            r = summarise(w2, wall, kt)
            r.pop("_kb")
            if name != "fp32" and r["xstep_TFLOPs"]:
                r["xstep_mfma_frac"] = round(r["xstep_TFLOPs"] / FP16_DENSE_PEAK_TFLOPS, 5)
            c2[name] = r
            print(f"32-station {name} done: {r['value']:.2f}", file=sys.stderr, flush=True)
        c2["mixed_vs_plain"] = round(c2["split_fp16"]["value"] / c2["fp32"]["value"], 3)
        c2["fp16_vs_plain"] = round(c2["fp16"]["value"] / c2["fp32"]["value"], 3)
        line["config2"] = c2
        del w2
        torch.cuda.empty_cache()
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_usec)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
