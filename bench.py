"""Throughput of the correlator-node data path on one MI355X.

Workload (N=1): 8 stations x 2 circular polarisations, 8 subbands of 32 MHz
sampled at 64 Ms/s, 2-bit, 4096 channels, 1 s integrations, delay offsets
U[-5, 5] ms and rates U[-3e-6, 3e-6]. One step = one 1 s integration of all 8
subbands (8 slices x 16 streams x 64 Msamples = 8.19e9 station-samples);
--steps 10 covers the full 10 s job. The HEADER control records come from the
input-node delay rule on the host; the 2-bit payload is synthesised straight
into device memory, so the timed region starts with the input resident in HBM.

Prints one JSON line (metric, value in station-Gsamples/s, roofline of the
dominant kernel, cpu_baseline of the oracle on a bounded sample).
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
KERNELS = ["stats", "fengine", "channelise", "xstep", "normalise"]


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subbands", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-usec", type=int, default=125000)
    ap.add_argument("--contexts", type=int, default=2, help="contexts (HIP streams) slices alternate over")
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
    """The FP32 oracle (C restatement, one core) on one slice of the same
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
    t_setup = time.perf_counter()
    per_sb = [configs.config1_jobs(sb, n_integrations=K) for sb in range(NSB)]
    inputs = {}  # (k, sb) -> list of (control, tensor, nbytes)
    for sb in range(NSB):
        jobs, raw_start = per_sb[sb]
        for k, job in enumerate(jobs):
            nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
            lst = []
            for i, st in enumerate(job.station_streams):
                ctrl, off, n = synth.header_stream(1 << 40, raw_start, job.delay_models[i], job.start_time,
                                                   nsamp, st.sample_rate, st.bits_per_sample, elided=True)
                buf = torch.empty(n, dtype=torch.uint8, device=f"cuda:{dev}")
                n_int = int(synth.integer_delay(job.delay_models[i], job.start_time, st.sample_rate))
                capi.synthesize_raw(dev, buf.data_ptr(), n, st.bits_per_sample, off * 4, n_int,
                                    common_seed=1000 + 2 * sb + (st.polarisation == "L"),
                                    stream_seed=(sb * 64 + i) * 4096 + k, rho=0.02)
                lst.append((ctrl, buf, n))
            inputs[(k, sb)] = lst
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup
    print(f"setup done: {len(inputs)} slices in {t_setup:.1f} s", file=sys.stderr, flush=True)

    # two contexts on their own HIP streams: the host parses and enqueues the
    # next slice while the device works on the previous one
    ctxs = [capi.Context(dev) for _ in range(max(1, args.contexts))]
    for c in ctxs:
        c.set_options(lut=1)
    counter = [0]

    def run_slice(k, sb):
        ctx = ctxs[counter[0] % len(ctxs)]
        counter[0] += 1
        job = per_sb[sb][0][k]
        ctx.set_parameters(job)
        for i, (ctrl, buf, n) in enumerate(inputs[(k, sb)]):
            ctx.push_stream_elided(i, ctrl, buf.data_ptr(), n)
        ctx.correlate()

    for w in range(W):
        for sb in range(NSB):
            run_slice(w % K, sb)
    for c in ctxs:
        c.synchronize()
        c.kernel_times(reset=True)
    torch.cuda.synchronize()
    print("warmup done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for k in range(K):
        for sb in range(NSB):
            run_slice(k, sb)
    for c in ctxs:
        c.synchronize()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    for c in ctxs:
        c.kernel_times(reset=True)
    print(f"timed region done: {wall:.3f} s", file=sys.stderr, flush=True)
    # kernel durations for the roofline: one step again on one context, each
    # slice waited for, so no kernel overlaps another
    for sb in range(NSB):
        job = per_sb[sb][0][0]
        ctxs[0].set_parameters(job)
        for i, (ctrl, buf, n) in enumerate(inputs[(0, sb)]):
            ctxs[0].push_stream_elided(i, ctrl, buf.data_ptr(), n)
        ctxs[0].correlate()
        ctxs[0].synchronize()
    times = [ctxs[0].kernel_times(reset=True)]

    job = per_sb[0][0][0]
    M = len(job.station_streams)
    nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
    nseg = job.number_ffts_in_integration()
    station_samples = K * NSB * M * nsamp
    value = station_samples / wall / 1e9
    avg = np.mean(np.array(times), axis=0)  # per-slice ms, isolated pass
    B = len(job.baselines())
    nfold = 8
    fused = avg[2] == 0.0
    kb = kernel_bytes(job, M, nseg, nsamp, nfold, B, fused)
    dom = int(np.argmax(avg))
    dom_name = KERNELS[dom]
    achieved = kb[dom_name] / (avg[dom] * 1e-3) / 1e9
    line = {
        "metric": "station-Gsamples/s correlated (plain FP32 GPU path)",
        "value": round(value, 4),
        "unit": "station-Gsamples/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall / K * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seeded 2-bit station data generated in device memory, HEADER records from the "
                "input-node delay rule",
        "config": {"workload": "8 stations x 2 pol, 8 subbands x 32 MHz at 64 Ms/s, 2-bit, 4096 channels, "
                               "1 s integrations, 10 s", "stations": 8, "polarisations": 2, "subbands": NSB,
                   "sample_rate": 64_000_000, "channels": 4096, "integration_s": 1.0,
                   "station_samples_per_step": NSB * M * nsamp},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "kernel_ms_isolated": {n: round(float(a), 3) for n, a in zip(KERNELS, avg)},
        "kernel_GBs": {n: round(kb[n] / (a * 1e-3) / 1e9, 1) for n, a in zip(KERNELS, avg) if a > 0},
        "setup_s": round(t_setup, 1),
    }
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(args.cpu_sample_usec)
    for c in ctxs:
        c.close()
    print(json.dumps(line))


if __name__ == "__main__":
    main()
