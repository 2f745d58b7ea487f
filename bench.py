"""Throughput of the correlator-node data path on one MI355X.

Headline (`value`): the 32-station configuration -- 32 stations x 2 circular
polarisations, 4 subbands of 64 MHz sampled at 128 Ms/s, 2-bit, 16384
channels, 0.5 s integrations, 4 s of data, a common signal (rho 0.01) plus a
second point source (rho 0.005) with its own delays, delay offsets U[-5, 5] ms
and rates U[-3e-6, 3e-6] -- correlated by the batched mixed-precision form:
each step's four subband slices run as one captured HIP graph, with the
X-step on FP16 MFMA with FP32 accumulation. One step is one 0.5 s integration
of the four subbands (4 slices x 64 streams x 64 Msamples = 1.64e10
station-samples); steps cycle over the job's 8 integrations, so --steps 8
times the whole 4 s job. The same line carries the plain FP32 GPU path on
the same workload and the ratio of the two, split FP16 and BF16, each
variant's visibility error on one slice against split FP16, the roofline of
the dominant kernel, the X-step's MFMA fraction and the CPU baseline on a
sample of the same workload.

Extra keys: the 8-station configuration (plain FP32 and split FP16), raw VDIF
ingest, the 48-job mixed batch and pulsar binning.

The HEADER control records come from the input-node delay rule on the host;
the 2-bit payload is synthesised straight into device memory, so the timed
region starts with the input resident in HBM. Prints one JSON line.
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "sfxc-1_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP16_DENSE_PEAK_TFLOPS = 2500.0
# HBM bytes per launch (one chunk of a slice) of a kernel, from the FETCH_SIZE and
# WRITE_SIZE counters scaled by a 1 GiB copy kernel's readings (the
# summaries under profiles/ give the commands); re-measured whenever the
# kernel changes. Key: (kernel, channels, streams, precision).
MEASURED_TRAFFIC = {
    # MFMA channeliser: 4.106e9 read (FP16 series) + 8.198e9 written per
    # chunk of 1953 segments x 64 streams
    ("channelise", 16384, 64, "fp16"): 12304000000,
}
# Wave-level VALU instructions per launch (SQ_INSTS_VALU, same profiles),
# against the wave64 FP32 issue rate: a gfx950 SIMD issues up to two
# independent VALU instructions per quad-cycle (SQ_ACTIVE_INST_VALU2 counts
# those), so 256 CUs x 4 SIMDs x 2 / 4 cycles x 2.4 GHz.
MEASURED_VALU_INSTS = {
    ("channelise", 16384, 64, "fp16"): 1.666e9,
}
VALU_ISSUE_PEAK = 256 * 4 * 2 / 4 * 2.4e9
KERNELS = ["stats", "fengine", "channelise", "xstep", "normalise"]
PREC = {"fp32": 0, "fp16": 1, "bf16": 2, "split_fp16": 3}
C2_SUBBANDS, C2_INTEGRATIONS = 4, 8


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="0.5 s integrations of the 32-station job (8 = all 4 s)")
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--precisions", default="fp16,fp32,split_fp16,bf16",
                    help="32-station legs; the first is the headline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-usec", type=int, default=7812,
                    help="length of each CPU-baseline worker's 32-station slice")
    ap.add_argument("--cpu-cores", type=int, default=0, help="CPU baseline worker processes (default: up to 16)")
    ap.add_argument("--only", choices=("all", "c1", "c2", "c3", "c4"), default="all",
                    help="c2: the headline only; c1, c3, c4: that extra leg only (no headline)")
    ap.add_argument("--no-extras", action="store_true", help="headline only (same as --only c2)")
    ap.add_argument("--c1-steps", type=int, default=2, help="1 s integrations of the 8-station workload")
    ap.add_argument("--c1-contexts", type=int, default=2)
    ap.add_argument("--c3-jobs", type=int, default=48, help="jobs of the mixed batch")
    ap.add_argument("--c3-precisions", default="fp32,fp16")
    ap.add_argument("--c4-steps", type=int, default=1, help="1 s integrations of the pulsar workload")
    ap.add_argument("--c4-precisions", default="fp32,fp16,split_fp16")
    ap.add_argument("--no-peaks", action="store_true", help="skip the achievable-peak microbenchmarks")
    ap.add_argument("--no-autotune", action="store_true",
                    help="default launch shapes instead of the autotuner's (run before each timed pass)")
    ap.add_argument("--no-errors", action="store_true", help="skip the per-variant visibility errors")
    return ap.parse_args()


def kernel_bytes(job, M, nseg, n_samples, nfold, B, fused, spec_bytes=8, ts_bytes=4):
    """Algorithmic HBM bytes per launch of each kernel. fengine: unpack +
    delay correction (+ channeliser when fused: payload in, spectra out).
    Spectra take spec_bytes per bin: 8 (FP32), 4 (FP16 / BF16 pairs); the
    real series between the delay stage and the channeliser ts_bytes per
    sample: 4 (FP32), 2 (FP16, before the MFMA channeliser)."""
    nc = job.fft_size_correlation
    bits = job.station_streams[0].bits_per_sample
    spectra = M * nseg * (nc + 1) * spec_bytes
    return {
        "stats": M * n_samples * bits / 8,
        "fengine": M * n_samples * bits / 8 + (spectra if fused else M * n_samples * ts_bytes),
        "channelise": M * n_samples * ts_bytes + spectra,
        "xstep": spectra + nfold * B * (nc + 1) * 16,
        "normalise": B * (nc + 1) * 8 * 2,
    }


def _c2_cpu_inputs(k, sample_usec):
    """A slice of the 32-station workload (subband k % 4) shortened to
    sample_usec, with random 2-bit payload behind the input-node records."""
    from sfxg import configs, synth

    jobs, raw_start = configs.config2_jobs(k % C2_SUBBANDS, n_integrations=1, t_int_usec=sample_usec)
    job = jobs[0]
    rng = np.random.default_rng(70 + k)
    hdrs = []
    nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
    for i, st in enumerate(job.station_streams):
        ctrl, off, n = synth.header_stream(1 << 40, raw_start, job.delay_models[i], job.start_time, nsamp,
                                           st.sample_rate, st.bits_per_sample, elided=True)
        hdrs.append(synth.join_elided(ctrl, rng.integers(0, 256, n, dtype=np.uint8)))
    return job, hdrs, len(job.station_streams) * nsamp


def _cpu_worker(k, sample_usec, barrier, out):
    """One correlator-node worker on one core: builds its slice, waits for the
    others, then times the FP32 oracle on it."""
    from oracle import oracle as O

    O.set_threads(1)
    job, hdrs, ss = _c2_cpu_inputs(k, sample_usec)
    O.lib("f32")
    barrier.wait()
    t0 = time.perf_counter()
    O.run_slice(job, hdrs, lut=0)
    out.put((k, t0, time.perf_counter(), ss))


def _cpu_run(n, sample_usec):
    import multiprocessing as mp

    ctx = mp.get_context("fork")
    barrier, out = ctx.Barrier(n), ctx.Queue()
    procs = [ctx.Process(target=_cpu_worker, args=(k, sample_usec, barrier, out)) for k in range(n)]
    for p in procs:
        p.start()
    res = [out.get() for _ in procs]
    for p in procs:
        p.join()
        if p.exitcode != 0:
            raise RuntimeError(f"CPU baseline worker failed with exit code {p.exitcode}")
    wall = max(r[2] for r in res) - min(r[1] for r in res)
    return sum(r[3] for r in res), wall


def cpu_cores():
    """Host cores this process may use, at most 16 (a one-GPU host's CPU share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(sample_usec, cores):
    """The FP32 oracle (C restatement, -O2 -ffast-math) as the reference's
    correlator nodes run it: one worker process per host core, each on an
    independent slice of the 32-station workload (subbands 0..3, shortened to
    sample_usec, all 64 streams and 16384 channels), plus one such slice on
    one core alone. Runs before the GPU is initialised (the workers are
    forked)."""
    ss1, wall1 = _cpu_run(1, sample_usec)
    ssn, walln = _cpu_run(cores, sample_usec)
    return {"value": ssn / walln / 1e9, "unit": "station-Gsamples/s", "cores": cores, "kind": "port",
            "single_core_value": ss1 / wall1 / 1e9, "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "fft_backend": "the oracle's own radix-4 Stockham FFT with per-stage twiddle tables (no FFTW or "
                           "MKL headers in this image)",
            "sample": f"{cores} worker processes, each one {sample_usec / 1e6:g} s slice of one subband of the "
                      f"32-station workload (64 streams, 16384 channels, {ss1:.3g} station-samples per worker), "
                      f"FP32 oracle; {walln:.1f} s wall for all, {wall1:.1f} s for one slice on one core"}


def achievable_peaks(torch, dev):
    """Copy bandwidth (read + write bytes), dense FP16 MFMA rate and FP32 VALU
    issue rate (wave-level instructions per second) measured on this GPU by
    the microbenchmarks in libsfxg_peaks.so."""
    import ctypes as C

    path = os.path.join(ROOT, "sfxc-1_amd", "lib", "libsfxg_peaks.so")
    lib = C.CDLL(path)  # raises if the library was not built
    n = 1 << 30
    a = torch.empty(n, dtype=torch.uint8, device=f"cuda:{dev}")
    b = torch.empty_like(a)
    a.fill_(1)
    torch.cuda.synchronize()
    ms = C.c_float()
    rc = lib.sfxg_peak_copy(C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()), C.c_int64(n), 10, C.byref(ms))
    if rc != 0:
        raise RuntimeError(f"sfxg_peak_copy failed: {rc}")
    del a, b
    torch.cuda.empty_cache()
    tf = C.c_double()
    rc = lib.sfxg_peak_mfma_fp16(20000, 8, C.byref(tf))
    if rc != 0:
        raise RuntimeError(f"sfxg_peak_mfma_fp16 failed: {rc}")
    vi = C.c_double()
    rc = lib.sfxg_peak_valu(20000, 32, C.byref(vi))
    if rc != 0:
        raise RuntimeError(f"sfxg_peak_valu failed: {rc}")
    return {"hbm_copy_GBs": round(2 * n / (ms.value * 1e-3) / 1e9, 1), "copy_bytes": n,
            "mfma_fp16_dense_TFLOPs": round(tf.value, 1), "valu_fp32_insts_per_s": float(f"{vi.value:.4g}")}


class Workload:
    """Device-resident input of n_subbands x K slices."""

    def __init__(self, torch, capi, synth, jobs_fn, K, NSB, dev, seed_base, rho=0.02, pulsar_fn=None):
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
                                        stream_seed=(seed_base + sb * 64 + i) * 4096 + k, rho=rho)
                    lst.append((ctrl, buf, n))
                self.inputs[(k, sb)] = lst
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0
        job = self.per_sb[0][0][0]
        self.job = job
        self.pulsar = pulsar_fn(job) if pulsar_fn else None
        self.M = len(job.station_streams)
        self.nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
        self.station_samples_per_step = NSB * self.M * self.nsamp

    def run(self, torch, capi, precision, W, n_ctx, dev, autotune=True):
        ctxs = [capi.Context(dev) for _ in range(max(1, n_ctx))]
        for c in ctxs:
            c.set_options(lut=1, precision=precision)
        counter = [0]
        if autotune:
            c = ctxs[0]
            c.set_parameters(self.per_sb[0][0][0], pulsar=self.pulsar)
            for i, (ctrl, buf, n) in enumerate(self.inputs[(0, 0)]):
                c.push_stream_elided(i, ctrl, buf.data_ptr(), n)
            c.prepare()
            self.tune_report = c.autotune(reps=2)

        def run_slice(k, sb, ctx=None):
            ctx = ctx or ctxs[counter[0] % len(ctxs)]
            counter[0] += 1
            ctx.set_parameters(self.per_sb[sb][0][k], pulsar=self.pulsar)
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
        timed = [c.kernel_times(reset=True) for c in ctxs]
        self.kt_timed = np.mean(timed, axis=0)
        # kernel durations without overlap: one step again on one context,
        # each slice waited for
        for sb in range(self.NSB):
            run_slice(0, sb, ctxs[0])
            ctxs[0].synchronize()
        kt = ctxs[0].kernel_times(reset=True)
        for c in ctxs:
            c.close()
        return wall, kt


class BatchedWorkload:
    """The 32-station job in device memory (every subband and integration),
    correlated step by step: a step's four subband slices are prepared on
    one of two context sets (host work of the next step overlaps the graph
    of the previous one) and run as one captured graph of four branches."""

    def __init__(self, torch, capi, synth, configs, dev, n_int=C2_INTEGRATIONS, nsb=C2_SUBBANDS):
        t0 = time.perf_counter()
        self.per_sb = [configs.config2_jobs(sb, n_integrations=n_int) for sb in range(nsb)]
        rho2, off2, rate2 = configs.second_source(32)
        self.K, self.NSB = n_int, nsb
        self.inputs = {}
        for sb in range(nsb):
            jobs, raw_start = self.per_sb[sb]
            for k, job in enumerate(jobs):
                nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
                tsec = (job.start_time - jobs[0].start_time) / 512e6
                lst = []
                for i, st in enumerate(job.station_streams):
                    ctrl, off, n = synth.header_stream(1 << 40, raw_start, job.delay_models[i], job.start_time,
                                                       nsamp, st.sample_rate, st.bits_per_sample, elided=True)
                    buf = torch.empty(n, dtype=torch.uint8, device=f"cuda:{dev}")
                    n_int = int(synth.integer_delay(job.delay_models[i], job.start_time, st.sample_rate))
                    s = st.station_number
                    n2 = int(round((off2[s] + rate2[s] * tsec) * st.sample_rate))
                    pol = st.polarisation == "L"
                    capi.synthesize_raw(dev, buf.data_ptr(), n, st.bits_per_sample, off * 4, n_int,
                                        common_seed=2000 + 2 * sb + pol, stream_seed=(2000 + sb * 64 + i) * 4096 + k,
                                        rho=0.01, delay2_samples=n2, common2_seed=3000 + 2 * sb + pol, rho2=rho2)
                    lst.append((ctrl, buf, n))
                self.inputs[(k, sb)] = lst
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0
        self.job = self.per_sb[0][0][0]
        self.M = len(self.job.station_streams)
        self.nsamp = self.job.n_ffts_per_integration(0) * self.job.fft_size_delaycor
        self.station_samples_per_step = nsb * self.M * self.nsamp

    def _fill(self, ctx, k, sb):
        ctx.set_parameters(self.per_sb[sb][0][k])
        for i, (ctrl, buf, n) in enumerate(self.inputs[(k, sb)]):
            ctx.push_stream_elided(i, ctrl, buf.data_ptr(), n)

    def run(self, torch, capi, precision, K, W, dev, autotune=True):
        sets = [[capi.Context(dev) for _ in range(self.NSB)] for _ in range(2)]
        for cs in sets:
            for c in cs:
                c.set_options(lut=1, precision=precision)
        batches = [capi.Batch(dev, n_branches=self.NSB) for _ in range(2)]
        if autotune:
            c = sets[0][0]
            self._fill(c, 0, 0)
            c.prepare()
            self.tune_report = c.autotune(reps=2)
        self.host_s = 0.0

        def step(r):
            cs, b = sets[r % 2], batches[r % 2]
            b.synchronize()
            h0 = time.perf_counter()
            b.clear()
            k = r % self.K
            for sb, c in enumerate(cs):
                self._fill(c, k, sb)
                c.prepare()
                b.add(c)
            b.launch()
            self.host_s += time.perf_counter() - h0

        # at least two warm-up steps, so that both context sets hold their
        # buffers before the timed region
        W = max(W, 2)
        for r in range(W):
            step(r)
        for b in batches:
            b.synchronize()
        torch.cuda.synchronize()
        self.host_s = 0.0
        t0 = time.perf_counter()
        for r in range(K):
            step(W + r)
        for b in batches:
            b.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        self.host_ms_per_step = self.host_s / K * 1e3
        for b in batches:
            b.close()
        # kernel durations without overlap: the four slices of one step again,
        # eagerly on one context, each waited for
        c = sets[0][0]
        c.kernel_times(reset=True)
        for sb in range(self.NSB):
            self._fill(c, 0, sb)
            c.correlate()
            c.synchronize()
        kt = c.kernel_times(reset=True)
        self.chunks, self.chunk_segs = c.chunk_plan()
        # one slice's visibilities for the error report, and its read-back time
        self._fill(c, 0, 0)
        c.correlate()
        c.synchronize()
        t1 = time.perf_counter()
        vis = c.stage(capi.STAGE_NORMALISED)
        t2 = time.perf_counter()
        out = c.output()
        t3 = time.perf_counter()
        self.readback = {"normalised_bytes": int(vis.nbytes), "stage_ms": round((t2 - t1) * 1e3, 2),
                         "output_bytes": len(out), "output_ms": round((t3 - t2) * 1e3, 2),
                         "output_GBs": round(len(out) / max(t3 - t2, 1e-9) / 1e9, 2)}
        for cs in sets:
            for c in cs:
                c.close()
        return wall, kt, vis


def errors_vs(vis, ref):
    """Per product (rows) of one slice: relative L2 error over channels and
    the worst channel relative to the product's rms, against ref."""
    l2, worst = [], []
    for k in range(ref.shape[0]):
        r = ref[k].astype(np.complex128)
        g = vis[k].astype(np.complex128)
        rms = np.sqrt(np.mean(np.abs(r) ** 2))
        if rms == 0.0:
            continue
        d = np.abs(g - r)
        l2.append(np.sqrt(np.mean(d ** 2)) / rms)
        worst.append(np.max(d) / rms)
    l2, worst = np.array(l2), np.array(worst)
    return {"relative_l2_max": float(f"{l2.max():.3e}"), "relative_l2_median": float(f"{np.median(l2):.3e}"),
            "worst_channel_max": float(f"{worst.max():.3e}"), "products": int(len(l2))}


class MixedBatch:
    """The mixed batch: each job's 2-bit/1-bit payload for all its slices in
    device memory, one context per job and round parity (two sets, so one
    round's host preparation overlaps the previous round's graph), and per
    round one captured graph of every job's next slice."""

    def __init__(self, torch, capi, synth, configs, dev, n_jobs):
        t0 = time.perf_counter()
        self.jobs = configs.config3_jobs(n_jobs=n_jobs)
        self.inputs = []
        self.station_samples = 0
        for j, d in enumerate(self.jobs):
            per = []
            for k, job in enumerate(d["jobs"]):
                nsamp = job.n_ffts_per_integration(0) * job.fft_size_delaycor
                lst = []
                for i, st in enumerate(job.station_streams):
                    ctrl, off, n = synth.header_stream(1 << 40, d["raw_start"], job.delay_models[i], job.start_time,
                                                       nsamp, st.sample_rate, st.bits_per_sample, elided=True)
                    buf = torch.empty(n, dtype=torch.uint8, device=f"cuda:{dev}")
                    n_int = int(synth.integer_delay(job.delay_models[i], job.start_time, st.sample_rate))
                    capi.synthesize_raw(dev, buf.data_ptr(), n, st.bits_per_sample, off * (8 // st.bits_per_sample),
                                        n_int, common_seed=5000 + j, stream_seed=(5000 + j) * 4096 + i * 64 + k,
                                        rho=d["rho"])
                    lst.append((ctrl, buf, n))
                    self.station_samples += nsamp
                per.append(lst)
            self.inputs.append(per)
        torch.cuda.synchronize()
        self.setup_s = time.perf_counter() - t0
        self.rounds = max(len(d["jobs"]) for d in self.jobs)

    def run_one_graph(self, torch, capi, precision, dev, chunk_mib=128, reps=2):
        """Every slice of every job prepared on a context of its own, all
        captured into one graph (16 branches) and run as one launch; the
        first launch captures and is not timed, then `reps` launches are."""
        ctxs = []
        batch = capi.Batch(dev, n_branches=16)
        try:
            for j, d in enumerate(self.jobs):
                for r, job in enumerate(d["jobs"]):
                    c = capi.Context(dev)
                    ctxs.append(c)
                    c.set_options(lut=1, precision=precision, chunk_mib=chunk_mib)
                    c.set_parameters(job)
                    for i, (ctrl, buf, n) in enumerate(self.inputs[j][r]):
                        c.push_stream_elided(i, ctrl, buf.data_ptr(), n)
                    c.prepare()
                    batch.add(c)
            batch.launch()
            batch.synchronize()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                batch.launch()
            batch.synchronize()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / reps, len(ctxs)
        finally:
            batch.close()
            for c in ctxs:
                c.close()

    def run(self, torch, capi, precision, dev, chunk_mib=512, autotune=True):
        sets = [[capi.Context(dev) for _ in self.jobs] for _ in range(2)]
        for cs in sets:
            for c in cs:
                c.set_options(lut=1, precision=precision, chunk_mib=chunk_mib)
        batches = [capi.Batch(dev, n_branches=4) for _ in range(2)]
        if autotune:
            # one slice of every shape through the autotuner
            seen = set()
            for j, d in enumerate(self.jobs):
                job = d["jobs"][0]
                key = (job.fft_size_correlation, len(job.station_streams))
                if key in seen:
                    continue
                seen.add(key)
                c = sets[0][j]
                c.set_parameters(job)
                for i, (ctrl, buf, n) in enumerate(self.inputs[j][0]):
                    c.push_stream_elided(i, ctrl, buf.data_ptr(), n)
                c.prepare()
                c.autotune(reps=1)
            self.tuned_shapes = len(seen)

        def prepare(r):
            cs, b = sets[r % 2], batches[r % 2]
            b.synchronize()
            b.clear()
            for j, d in enumerate(self.jobs):
                if r < len(d["jobs"]):
                    c = cs[j]
                    c.set_parameters(d["jobs"][r])
                    for i, (ctrl, buf, n) in enumerate(self.inputs[j][r]):
                        c.push_stream_elided(i, ctrl, buf.data_ptr(), n)
                    c.prepare()
                    b.add(c)
            return b

        # a warm-up round on each context set (device buffers allocated at
        # their sizes, graphs captured once), then the timed pass
        for r in range(min(2, self.rounds)):
            prepare(r).launch()
        for b in batches:
            b.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for r in range(self.rounds):
            prepare(r).launch()
        for b in batches:
            b.synchronize()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        for b in batches:
            b.close()
        for cs in sets:
            for c in cs:
                c.close()
        return wall


def vdif_ingest(torch, capi, configs, dev, reps=3):
    """Raw VDIF ingest of one slice of the main workload: every station's
    frames (two threads, R and L, 8000-byte payloads of random 2-bit data)
    resident in HBM; times sfxg_push_vdif for all streams (header words lifted
    on the device, records built on the host, payload gathered on the device),
    then the correlation of that slice."""
    jobs, _ = configs.config1_jobs(0, n_integrations=1)
    job = jobs[0]
    st0 = job.station_streams[0]
    fs, bits, payload, epoch, ref_mjd = st0.sample_rate, st0.bits_per_sample, 8000, 46, 60000
    fb = payload + 32
    frame_samples = payload * 8 // bits
    tps = int(round(512_000_000 / fs))
    frame_ticks = frame_samples * tps
    fps = fs // frame_samples
    margin = 24  # frames before and after the slice: the delays reach 5 ms
    n_frames = (job.stop_time - job.start_time) // frame_ticks + 2 * margin
    epoch_ticks = (59945 - 51544) * 86400 * 512_000_000
    first = (job.start_time - margin * frame_ticks - epoch_ticks) // frame_ticks
    g = first + np.arange(n_frames, dtype=np.int64)
    rng = np.random.default_rng(5)
    stations = sorted({s.station_number for s in job.station_streams})
    bufs, streams = [], {}
    for sn in stations:
        idx = [i for i, s in enumerate(job.station_streams) if s.station_number == sn]
        fr = np.empty((n_frames, len(idx), fb), dtype=np.uint8)
        w = np.zeros((n_frames, len(idx), 8), dtype="<u4")
        for c in range(len(idx)):
            w[:, c, 0] = (g // fps).astype(np.uint32)
            w[:, c, 1] = ((g % fps) | (epoch << 24)).astype(np.uint32)
            w[:, c, 2] = fb // 8
            w[:, c, 3] = sn | ((1 + c) << 16) | ((bits - 1) << 26)
        fr[:, :, :32] = w.view(np.uint8).reshape(n_frames, len(idx), 32)
        fr[:, :, 32:] = rng.integers(0, 256, size=(n_frames, len(idx), payload), dtype=np.uint8)
        t = torch.from_numpy(fr.reshape(-1)).to(f"cuda:{dev}")
        bufs.append(t)
        for c, i in enumerate(idx):
            streams[i] = (len(bufs) - 1, c, [1 + k for k in range(len(idx))])
    torch.cuda.synchronize()
    ctx = capi.Context(dev)
    ctx.set_options(lut=1, precision=0)
    ingest, total = [], []
    for _ in range(reps):
        ctx.set_parameters(job)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(len(job.station_streams)):
            b, c, th = streams[i]
            ctx.push_vdif(i, bufs[b].data_ptr(), bufs[b].numel(), fb, c, th, ref_mjd)
        t1 = time.perf_counter()
        ctx.correlate()
        ctx.synchronize()
        t2 = time.perf_counter()
        ingest.append(t1 - t0)
        total.append(t2 - t0)
    ctx.close()
    frames_bytes = sum(int(b.numel()) for b in bufs)
    used = len(job.station_streams) * (job.n_ffts_per_integration(0) * job.fft_size_delaycor) * bits // 8
    best = min(ingest)
    return {"workload": "one slice of the main workload as 2-thread VDIF per station, 8000-byte payloads",
            "streams": len(job.station_streams), "frames_bytes_in_hbm": frames_bytes,
            "payload_bytes_of_slice": used, "header_bytes_to_host": 16 * frames_bytes // fb,
            "ingest_ms": round(best * 1e3, 2), "ingest_payload_GBs": round(used / best / 1e9, 1),
            "ingest_plus_correlate_ms": round(min(total) * 1e3, 2)}


def summarise(wl, wall, kt, precision="fp32", steps=None):
    """steps: steps in the timed region (default: the workload's own count)."""
    steps = wl.K if steps is None else steps
    job = wl.job
    nseg = job.number_ffts_in_integration()
    B = len(job.baselines())
    nfold = max(1, round(job.integration_time / job.sub_integration_time))
    fused = kt[2] == 0.0
    packed = precision in ("fp16", "bf16")
    # packed spectra at 1024 to 16384 channels: MFMA channeliser fed an FP16 series
    ts_bytes = 2 if packed and not fused and job.fft_size_correlation in (1024, 2048, 4096, 8192, 16384) else 4
    kb = kernel_bytes(job, wl.M, nseg, wl.nsamp, nfold, B, fused, 4 if packed else 8, ts_bytes)
    value = steps * wl.station_samples_per_step / wall / 1e9
    xflops = 8.0 * B * (job.fft_size_correlation + 1) * nseg
    kt_timed = getattr(wl, "kt_timed", None)
    return {
        "value": round(value, 4),
        "ms_per_step": round(wall / steps * 1e3, 3),
        "kernel_ms_isolated": {n: round(float(a), 3) for n, a in zip(KERNELS, kt)},
        "kernel_ms_timed_region": ({n: round(float(a), 3) for n, a in zip(KERNELS, kt_timed)}
                                   if kt_timed is not None else None),
        "kernel_GBs": {n: round(kb[n] / (a * 1e-3) / 1e9, 1) for n, a in zip(KERNELS, kt) if a > 0},
        "xstep_TFLOPs": round(xflops / (kt[3] * 1e-3) / 1e12, 2) if kt[3] > 0 else None,
        "_kb": kb,
        "_kt": kt,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this pool has one GPU and the path shards by "
              "(slice x subband) with no collectives; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    headline = args.only in ("all", "c2")
    extras = set() if args.no_extras or args.only == "c2" else (
        {"c1", "c3", "c4"} if args.only == "all" else {args.only})
    cpu = None
    if headline and not args.no_cpu_baseline:
        # before the GPU is touched: the workers are forked
        cpu = cpu_baseline(args.cpu_sample_usec, args.cpu_cores or cpu_cores())
        print(f"CPU baseline done: {cpu['value']:.4f} station-Gsamples/s on {cpu['cores']} cores",
              file=sys.stderr, flush=True)
    import torch
    from sfxg import capi, configs, synth

    dev = 0
    torch.cuda.set_device(dev)
    peaks = None if args.no_peaks else achievable_peaks(torch, dev)
    if peaks:
        print(f"peaks: {peaks}", file=sys.stderr, flush=True)
    K, W = args.steps, args.warmup
    line = {}
    if headline:
        wl = BatchedWorkload(torch, capi, synth, configs, dev)
        print(f"32-station setup done: {len(wl.inputs)} slices in {wl.setup_s:.1f} s", file=sys.stderr, flush=True)
        legs, vis = {}, {}
        precs = args.precisions.split(",")
        for name in precs:
            wall, kt, v = wl.run(torch, capi, PREC[name], K, W, dev, autotune=not args.no_autotune)
            r = summarise(wl, wall, kt, name, steps=K)
            r["chunks_per_slice"], r["chunk_segments"] = wl.chunks, wl.chunk_segs
            r["readback"] = wl.readback
            r["host_prepare_ms_per_step"] = round(wl.host_ms_per_step, 2)
            if name != "fp32":
                add_mfma_frac(r, peaks)
            legs[name] = r
            vis[name] = v
            print(f"32-station {name} done: {r['value']:.2f} station-Gsamples/s", file=sys.stderr, flush=True)
        head = precs[0]
        h = legs[head]
        kt = h["_kt"]
        dom = int(np.argmax(kt))
        dom_name = KERNELS[dom]
        achieved = h["_kb"][dom_name] / (kt[dom] * 1e-3) / 1e9
        # the delay, channeliser and X-step kernels run once per chunk of the
        # slice; stats and normalise once per slice
        # this leg's chunk plan (packed FP16 / BF16 spectra take half the
        # bytes of FP32 ones, so fewer chunks)
        chunks, chunk_segs = h["chunks_per_slice"], h["chunk_segments"]
        per = chunks if dom_name in ("fengine", "channelise", "xstep") else 1
        roof = {"bound": "hbm", "kernel": dom_name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "bytes_per_launch": int(h["_kb"][dom_name] / per),
                "launch": (f"one chunk of {chunk_segs} segments x 64 streams ({chunks} per 0.5 s slice)"
                           if per > 1 else "one slice (64 streams x 0.5 s)"),
                "launch_ms": round(float(kt[dom]) / per, 3)}
        if peaks:
            roof["frac_of_achievable"] = round(achieved / peaks["hbm_copy_GBs"], 4)
        key = (dom_name, 16384, wl.M, head)
        if key in MEASURED_TRAFFIC:
            roof["traffic"] = MEASURED_TRAFFIC[key]
            roof["traffic_source"] = "rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated"
        if key in MEASURED_VALU_INSTS:
            rate = MEASURED_VALU_INSTS[key] / (roof["launch_ms"] * 1e-3)
            roof["valu"] = {"insts_per_launch": MEASURED_VALU_INSTS[key], "achieved_insts_per_s": round(rate, -8),
                            "peak_insts_per_s": VALU_ISSUE_PEAK, "frac": round(rate / VALU_ISSUE_PEAK, 4),
                            "source": "rocprofv3 SQ_INSTS_VALU"}
            if peaks:
                roof["valu"]["achievable_insts_per_s"] = peaks["valu_fp32_insts_per_s"]
                roof["valu"]["frac_of_achievable"] = round(rate / peaks["valu_fp32_insts_per_s"], 4)
        xs = {"kernel": "xstep", "TFLOPs": h["xstep_TFLOPs"], "mfma_frac": h.get("xstep_mfma_frac"),
              "mfma_frac_of_achievable": h.get("xstep_mfma_frac_of_achievable"),
              "GBs": h["kernel_GBs"].get("xstep"), "launch_ms": round(h["kernel_ms_isolated"]["xstep"] / chunks, 3),
              "launches_per_slice": chunks}
        errs = None
        if not args.no_errors and "split_fp16" in vis:
            errs = {"reference": "split FP16 of the same slice (itself within 1e-5 relative L2 of the FP64 oracle "
                                 "in the GPU tests); one slice: 2080 products x 16385 channels, 3905 segments"}
            for name, v in vis.items():
                if name != "split_fp16":
                    errs[name] = errors_vs(v[0], vis["split_fp16"][0])
        line.update({
            "metric": "station-Gsamples/s correlated (mixed vs FP32 GPU), X-step MFMA %, HBM GB/s",
            "value": h["value"],
            "unit": "station-Gsamples/s",
            "n_gpus": 1,
            "steps": K,
            "warmup": W,
            "ms_per_step": h["ms_per_step"],
            "higher_is_better": True,
            "scaling": "weak",
            "vs_baseline": None,
            "dtype": head,
            "dtype_detail": f"{head} X-step operands on MFMA with FP32 accumulation; FP32 delay correction "
                            "(its real series stored as FP16 for packed FP16 / BF16 spectra, FP32 otherwise); "
                            "channeliser transform on FP16 MFMA operands with FP32 accumulation (packed FP16 / BF16 "
                            "spectra; FP32 transform for plain and split FP16); FP64 delay model and phase arguments",
            "data": "synthetic: seeded 2-bit station data generated in device memory (common signal rho 0.01 and a "
                    "second source rho 0.005 with its own delays), HEADER records from the input-node delay rule",
            "config": {"workload": "32 stations x 2 pol, 4 subbands x 64 MHz at 128 Ms/s, 2-bit, 16384 channels, "
                                   "0.5 s integrations, 4 s; each step's 4 subband slices as one captured graph",
                       "stations": 32, "polarisations": 2, "subbands": wl.NSB, "sample_rate": 128_000_000,
                       "channels": 16384, "integration_s": 0.5, "integrations_in_job": wl.K,
                       "station_samples_per_step": wl.station_samples_per_step,
                       "job_station_samples": wl.K * wl.station_samples_per_step},
            "roofline": roof,
            "xstep": xs,
            "mixed_vs_plain": round(h["value"] / legs["fp32"]["value"], 3) if "fp32" in legs else None,
            "setup_s": round(wl.setup_s, 1),
        })
        for name, r in legs.items():
            r.pop("_kb")
            r.pop("_kt")
            if name != head:
                line[name] = r
                if "fp32" in legs and name != "fp32":
                    r["vs_plain"] = round(r["value"] / legs["fp32"]["value"], 3)
        line["kernel_ms_isolated"] = h["kernel_ms_isolated"]
        line["kernel_GBs"] = h["kernel_GBs"]
        line["readback"] = h["readback"]
        line["host_prepare_ms_per_step"] = h["host_prepare_ms_per_step"]
        if errs:
            line["visibility_errors"] = errs
        del wl, vis
        torch.cuda.empty_cache()
    if "c1" in extras:
        wl = Workload(torch, capi, synth, lambda sb, k: configs.config1_jobs(sb, n_integrations=k), args.c1_steps, 8,
                      dev, 1000)
        print(f"8-station setup done: {len(wl.inputs)} slices in {wl.setup_s:.1f} s", file=sys.stderr, flush=True)
        c1 = {"workload": "8 stations x 2 pol, 8 subbands x 32 MHz at 64 Ms/s, 2-bit, 4096 channels, 1 s "
                          f"integrations, {args.c1_steps} s", "station_samples_per_step": wl.station_samples_per_step}
        for name in ("fp32", "fp16", "split_fp16"):
            wall, kt = wl.run(torch, capi, PREC[name], 1, args.c1_contexts, dev, autotune=not args.no_autotune)
            r = summarise(wl, wall, kt, name)
            r.pop("_kb")
            r.pop("_kt")
            if name != "fp32":
                add_mfma_frac(r, peaks)
            c1[name] = r
            print(f"8-station {name} done: {r['value']:.2f}", file=sys.stderr, flush=True)
        c1["mixed_vs_plain"] = round(c1["fp16"]["value"] / c1["fp32"]["value"], 3)
        c1["split_fp16_vs_plain"] = round(c1["split_fp16"]["value"] / c1["fp32"]["value"], 3)
        line["config1"] = c1
        del wl
        torch.cuda.empty_cache()
        line["vdif_ingest"] = vdif_ingest(torch, capi, configs, dev)
        print(f"VDIF ingest done: {line['vdif_ingest']['ingest_ms']} ms per slice", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if "c3" in extras:
        mb = MixedBatch(torch, capi, synth, configs, dev, args.c3_jobs)
        print(f"mixed batch setup done in {mb.setup_s:.1f} s", file=sys.stderr, flush=True)
        from collections import Counter
        c3 = {"workload": f"{args.c3_jobs} seeded jobs, 2 s each: channels from {{256, 1024, 4096, 16384, 32768}}, "
                          "4..24 stations, 1- or 2-bit, 8/16/32/64 MHz, 0.25/0.5/1/2 s integrations; per round "
                          "every job's next slice as one captured HIP graph (value), and all slices as one graph "
                          "launch (one_graph)",
              "station_samples": mb.station_samples, "rounds": mb.rounds,
              "slices": sum(len(d["jobs"]) for d in mb.jobs),
              "channels": dict(Counter(str(d["nchan"]) for d in mb.jobs)),
              "bits": dict(Counter(str(d["bits"]) for d in mb.jobs))}
        for name in args.c3_precisions.split(","):
            wall = mb.run(torch, capi, PREC[name], dev, autotune=not args.no_autotune)
            c3[name] = {"value": round(mb.station_samples / wall / 1e9, 4), "wall_s": round(wall, 3)}
            print(f"mixed batch {name} done: {c3[name]['value']:.2f}", file=sys.stderr, flush=True)
            torch.cuda.empty_cache()
            try:
                wall1, n_ctx = mb.run_one_graph(torch, capi, PREC[name], dev)
            except capi.SfxgError as e:  # e.g. device memory for every slice's buffers at once
                c3[name]["one_graph"] = {"error": str(e)[:200]}
                print(f"mixed batch {name} one graph failed: {e}", file=sys.stderr, flush=True)
                continue
            c3[name]["one_graph"] = {"value": round(mb.station_samples / wall1 / 1e9, 4), "wall_s": round(wall1, 3),
                                     "slices": n_ctx, "launches": 1,
                                     "note": "every slice prepared beforehand on its own context; one graph "
                                             "launch runs all of them"}
            print(f"mixed batch {name} one graph: {c3[name]['one_graph']['value']:.2f}", file=sys.stderr, flush=True)
        if "fp32" in c3 and "fp16" in c3:
            c3["fp16_vs_plain"] = round(c3["fp16"]["value"] / c3["fp32"]["value"], 3)
        line["config3_batch"] = c3
        del mb
        torch.cuda.empty_cache()
    if "c4" in extras:
        w4 = Workload(torch, capi, synth, lambda sb, k: configs.config4_jobs(sb, n_integrations=k)[:2],
                      args.c4_steps, 8, dev, 4000, rho=0.005, pulsar_fn=configs.pulsar4)
        print(f"pulsar setup done in {w4.setup_s:.1f} s", file=sys.stderr, flush=True)
        c4 = {"workload": "12 stations x 2 pol, 8 subbands x 16 MHz at 32 Ms/s, 2-bit, 1024 channels, 1 s "
                          f"integrations, {args.c4_steps} s, pulsar binning: 64 bins over the gate [0.25, 0.75) "
                          "+ off-gate bin, P = 0.0893 s, DM 68 (device-synthesised continuum: the binning cost "
                          "does not depend on the data)",
              "station_samples_per_step": w4.station_samples_per_step}
        for name in args.c4_precisions.split(","):
            wall, kt = w4.run(torch, capi, PREC[name], 0, 2, dev, autotune=not args.no_autotune)
            r = summarise(w4, wall, kt, name)
            r.pop("_kb")
            r.pop("_kt")
            if name != "fp32":
                add_mfma_frac(r, peaks)
            c4[name] = r
            print(f"pulsar {name} done: {r['value']:.2f}", file=sys.stderr, flush=True)
        for name in ("fp16", "split_fp16"):
            if name in c4 and "fp32" in c4:
                c4[f"{name}_vs_plain"] = round(c4[name]["value"] / c4["fp32"]["value"], 3)
        line["config4_pulsar"] = c4
        del w4
        torch.cuda.empty_cache()
    if peaks:
        line["achievable_peaks"] = peaks
    if not args.no_autotune:
        table = capi.tune_export().strip().split("\n")
        main_keys = [capi.tune_key(16384, 64, p, False, False) for p in range(4)]
        line["autotuner"] = {"shapes_tuned": len(table),
                             "headline_shapes": [t for t in table if t.split("|")[0] in main_keys]}
    if cpu:
        line["cpu_baseline"] = cpu
        if "value" in line:
            line["cpu_baseline"]["gpu_over_cpu"] = round(line["value"] / cpu["value"], 1)
    print(json.dumps(line))


def add_mfma_frac(r, peaks):
    if r["xstep_TFLOPs"]:
        r["xstep_mfma_frac"] = round(r["xstep_TFLOPs"] / FP16_DENSE_PEAK_TFLOPS, 5)
        if peaks:
            r["xstep_mfma_frac_of_achievable"] = round(r["xstep_TFLOPs"] / peaks["mfma_fp16_dense_TFLOPs"], 5)


if __name__ == "__main__":
    main()
