"""Benchmark: WHIR prover time (commit + prove) at d=24 over Goldilocks2 on one MI355X.

Workload (the headline configuration): WHIR PCS, Field64_2 (Basefield embedding),
num_variables = 24, starting log inverse rate 2, folding factor 4, Blake3 Merkle
trees, 100-bit security with 20 PoW bits, list decoding (Johnson), 2 evaluation
points.  2^24 base-field coefficients uniform in F_p and points uniform in the
extension, from a seeded generator (synthetic data).  One step = one full
commit + prove with the input vector already resident in HBM.

PoW: at 100-bit security over Goldilocks2 this configuration asks for 42-55 bits
of grinding per sumcheck round, which no machine finishes; the timed prover runs
the same configuration with every PoW threshold at u64::MAX (the reference's
disable_pow), and the grinding kernel is measured separately ("pow": hashes/s of
the persistent kernel and the projected time of the literal PoW).

Prints one JSON line.  --gpus N with N > 1 is refused: this build measures one GPU.
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
# Per-launch HBM bytes (FETCH_SIZE + WRITE_SIZE, copy-calibrated) from rocprofv3 --pmc
# passes over this workload at d=24; matched to the dominant kernel by name.
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_d24.json")
# SQ_INSTS_VALU per launch of each kernel of the same d=24 command (and of the PoW
# grinds), from a rocprofv3 --pmc pass: the instruction count is fixed per launch, so
# instructions / the live launch time here is the achieved VALU issue rate.
VALU_FILE = os.path.join(ROOT, "profiles", "valu_d24.json")
POW_KERNELS = {"blake3": "k_pow<4>", "sha3": "k_pow<3>"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--num-variables", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-proving measurement")
    ap.add_argument("--no-other-configs", action="store_true", help="skip the d=20 Goldilocks2 and d=26 "
                    "Goldilocks3 prover timings")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving measurement")
    ap.add_argument("--no-pow", action="store_true", help="skip the PoW grinding-rate measurement")
    ap.add_argument("--no-pcie", action="store_true", help="skip the host-memory (PCIe-inclusive) proof timing")
    ap.add_argument("--workers", type=int, default=8, help="executor worker threads (streams) for serving")
    ap.add_argument("--batch-workers", type=int, default=12,
                    help="executor worker threads for the batched measurement (12 measured ~10%% faster than "
                         "8 for the 64-instance batch on MI355X; serving keeps 8, which gave the lower p99)")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-qps", type=float, default=0.0, help="arrival rate; 0 = 70%% of measured capacity")
    ap.add_argument("--inputs", choices=("uniform", "cli"), default="uniform",
                    help="uniform: seeded uniform coefficients and points; cli: the reference CLI's inputs "
                         "(v[i] = i, points (x, ..., x) for x = 0, 1, ..., src/bin/main.rs:150,173-175)")
    ap.add_argument("--linear-constraints", type=int, default=0,
                    help="Covector constraints (0, 1, ..., n-1) before the points, as the reference CLI's "
                         "--linear-constraints (src/bin/main.rs:163-170)")
    ap.add_argument("--ntt-autotune", type=int, default=1, choices=(0, 1),
                    help="tune each NTT shape's schedule during the warmup steps (1, default) or use the "
                         "built-in schedules (0)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run one device-to-device copy of this many bytes (a known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def _note(msg):
    """Progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _proof_digest(t):
    import hashlib
    narg, hints = t.proof()
    return hashlib.sha256(len(narg).to_bytes(8, "little") + narg + hints).digest()


def uniform_field(rng, n):
    p = np.uint64(0xFFFFFFFF00000001)
    v = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    bad = v >= p
    while bad.any():
        v[bad] = rng.integers(0, 2**64, size=int(bad.sum()), dtype=np.uint64)
        bad = v >= p
    return v


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1",
                          "n_gpus_requested": args.gpus}))
        sys.exit(2)

    import whir_amd as W
    from whir_amd import prover as GP

    nv, D = args.num_variables, 2
    params = GP.ProtocolParameters(degree=D, security_level=100, pow_bits=20, initial_folding_factor=4,
                                   folding_factor=4, unique_decoding=False, starting_log_inv_rate=2,
                                   hash_engine=W.HASH_BLAKE3)
    cfg = GP.Config(1 << nv, params).disable_pow()
    desc = cfg.describe()
    rng = np.random.default_rng(20241202)
    n = 1 << nv
    if args.inputs == "cli":
        vec = np.arange(n, dtype=np.uint64)
        points = []
        for x in range(2):
            pt = np.zeros(nv * D, dtype=np.uint64)
            pt[0::D] = x
            points.append(pt)
    else:
        vec = uniform_field(rng, n)
        points = [uniform_field(rng, nv * D) for _ in range(2)]
    dvec = W.DeviceBuffer.from_array(vec)
    # Covector constraints (0, 1, ..., n-1) embedded in the extension; their claimed
    # values are dot products computed on the GPU.
    cov_host = None
    covs = []
    cov_evals = []
    if args.linear_constraints:
        cov_host = np.zeros(n * D, dtype=np.uint64)
        cov_host[0::D] = np.arange(n, dtype=np.uint64)
        lifted = np.zeros(n * D, dtype=np.uint64)
        lifted[0::D] = vec
        ce = W.dot(D, cov_host, lifted)
        del lifted
        for _ in range(args.linear_constraints):
            covs.append(W.DeviceBuffer.from_array(cov_host))
            cov_evals.append(ce)
    evals = cov_evals + [GP.mle_evaluate(D, dvec, nv, pt) for pt in points]
    forms = covs + points
    flat_pts = np.concatenate(points)
    flat_ev = np.concatenate(evals)
    import hashlib
    pid = hashlib.sha3_512(json.dumps(desc, sort_keys=True).encode()).digest()
    sid = hashlib.sha3_256(b"whir_amd bench").digest()

    if args.calibrate_copy:
        src = W.DeviceBuffer(args.calibrate_copy)
        dst = W.DeviceBuffer(args.calibrate_copy)
        W._call("whir_memcpy_d2d", dst.ptr, src.ptr, args.calibrate_copy)
        W.synchronize()
        del src, dst

    def step():
        t = GP.Transcript(pid, sid)
        w = GP.commit(cfg, t, dvec)
        if covs:
            GP.prove_forms(cfg, t, [dvec], [w], forms, np.array(evals, dtype=np.uint64).reshape(len(forms), 1, D))
        else:
            GP.prove(cfg, t, dvec, w, flat_pts, flat_ev)
        return t

    W.ntt_autotune(bool(args.ntt_autotune))
    for _ in range(max(args.warmup, 1 if args.ntt_autotune else 0)):
        step()
    # Shapes tuned so far keep their schedules; later sections (other configurations,
    # batching, serving) use the built-in schedules for new shapes, so no tuning
    # time lands inside their measurements.
    W.ntt_autotune(False)
    W.synchronize()
    t0 = time.perf_counter()
    last = None
    for _ in range(args.steps):
        last = step()
    W.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps

    # Per-kernel times: the same steps again with a HIP event pair around every
    # launch, recorded on the stream the kernel runs on (kept out of the timed
    # region above, since the events add host work between launches).
    W.profile_enable(True)
    W.profile_reset()
    for _ in range(args.steps):
        step()
    W.synchronize()
    prof = W.profile_report()
    W.profile_enable(False)

    # Dominant kernel and its roofline (algorithmic bytes per launch / average launch time).
    dom_name, dom = max(prof.items(), key=lambda kv: kv[1]["total_ms"])
    roofline = None
    if dom["launches"]:
        avg_ms = dom["total_ms"] / dom["launches"]
        bytes_per_launch = dom["bytes"] / dom["launches"]
        gbs = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
        traffic = None
        if nv == 24 and os.path.exists(TRAFFIC_FILE):
            entry = json.load(open(TRAFFIC_FILE))["kernels"].get(dom_name)
            if entry:
                traffic = int(entry["traffic_bytes_per_launch"])
        roofline = {"kernel": dom_name, "bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "avg_launch_ms": round(avg_ms, 4), "algorithmic_bytes_per_launch": int(bytes_per_launch)}
        if nv == 24 and os.path.exists(VALU_FILE) and avg_ms > 0:
            vj = json.load(open(VALU_FILE))
            entry = vj["kernels"].get(dom_name)
            if entry:
                rate = entry["valu_per_launch"] / (avg_ms * 1e-3) / 1e12
                roofline["valu"] = {"instructions_per_launch": int(entry["valu_per_launch"]),
                                    "achieved": round(rate, 4), "peak": round(vj["peak_tinstr_per_s"], 4),
                                    "unit": "T wave-instr/s", "frac": round(rate / vj["peak_tinstr_per_s"], 4)}

    # Host-buffer boundary: a caller that hands over the coefficients in host memory
    # adds their upload to every proof.  Reported beside value, never as it.
    # (a) whir_commit_host from pinned memory: the upload goes a chunk of messages at a
    #     time on a copy stream, each chunk's encode overlapping the later copies;
    # (b) for reference, a plain upload from pageable memory before each proof.
    pcie = None
    if not covs and not args.no_pcie:
        pinned = W.PinnedArray(vec.size)
        pinned.array[:] = vec
        dhost = W.DeviceBuffer(vec.nbytes)

        def step_host():
            t = GP.Transcript(pid, sid)
            w = GP.commit_host(cfg, t, pinned, dhost)
            GP.prove(cfg, t, dhost, w, flat_pts, flat_ev)
            return t

        W.ntt_autotune(bool(args.ntt_autotune))  # the chunked encode's shapes are new
        for _ in range(2):
            step_host()
        W.ntt_autotune(False)
        W.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            th = step_host()
        W.synchronize()
        ms_host = (time.perf_counter() - t0) * 1e3 / args.steps
        same_host = th.proof() == last.proof()
        t0 = time.perf_counter()
        for _ in range(3):
            W.memcpy_h2d(dhost, pinned.array)
        h2d_pinned = (time.perf_counter() - t0) * 1e3 / 3
        h2d = []
        for _ in range(3):
            t0 = time.perf_counter()
            W.memcpy_h2d(dvec, vec)
            h2d.append(time.perf_counter() - t0)
        h2d_ms = min(h2d) * 1e3
        pcie = {"ms_per_proof_from_host": round(ms_host, 3), "bytes": int(vec.nbytes),
                "h2d_pinned_ms": round(h2d_pinned, 3), "h2d_pinned_gb_s": round(vec.nbytes / h2d_pinned / 1e6, 1),
                "identical_proof": same_host,
                "pageable_sequential": {"h2d_ms": round(h2d_ms, 3), "ms_per_proof_with_h2d": round(ms + h2d_ms, 3)},
                "note": "whir_commit_host + whir_prove_dev with the 2^24 coefficients in pinned host memory: "
                        "uploaded four messages at a time on a second stream, each chunk's Reed-Solomon encode "
                        "starting when its copy has landed; pageable_sequential = a plain pageable upload "
                        "followed by the device-resident proof"}
        del pinned, dhost

    verified = None
    if not args.no_verify:
        from oracle import whir_protocol as WP
        ocfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(
            security_level=100, pow_bits=20, initial_folding_factor=4, folding_factor=4, unique_decoding=False,
            starting_log_inv_rate=2, hash_engine=W.HASH_BLAKE3), D)
        ocfg.disable_pow()
        narg, hints = last.proof()
        pts_list = [WP.Covector(cov_host) for _ in covs] + [[p[i * D:(i + 1) * D] for i in range(nv)]
                                                            for p in points]
        try:
            WP.verify(WP.TranscriptReader(pid, sid, narg, hints), ocfg, [e for e in evals], pts_list)
            verified = True
        except Exception as e:  # reported, never hidden
            verified = f"rejected: {e}"

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(nv, D, vec, points, evals, pid, sid, last.proof(),
                           [cov_host for _ in covs])
    pow_rates = None if args.no_pow else bench_pow(W, GP, nv)
    if pow_rates and os.path.exists(VALU_FILE):
        # Achieved VALU issue rate of the grinding kernels: their instructions per hash
        # (profile: VALU per launch over the nonces of that launch) x the live hash rate.
        vj = json.load(open(VALU_FILE))
        for name, kern in POW_KERNELS.items():
            entry = vj["kernels"].get(kern)
            if entry and vj.get("pow_nonces"):
                per_hash = entry["valu_per_launch"] * entry["launches"] / vj["pow_nonces"]
                rate = per_hash * pow_rates[f"{name}_ghashes_per_s"] * 1e9 / 1e12
                pow_rates[f"{name}_valu"] = {"wave_instructions_per_hash": round(per_hash, 3),
                                             "achieved": round(rate, 4), "peak": round(vj["peak_tinstr_per_s"], 4),
                                             "unit": "T wave-instr/s", "frac": round(rate / vj["peak_tinstr_per_s"], 4)}
    _note("other configurations")
    others = None if args.no_other_configs else bench_other_configs(W, GP)
    _note("batched proving")
    batched = None if args.no_batch else bench_batched(args.batch_workers)
    _note("serving")
    serving = None if args.no_serve else bench_serving(args.workers, args.serve_seconds, args.serve_qps)

    kernels = {k: {"launches": v["launches"], "ms_per_step": round(v["total_ms"] / args.steps, 4)}
               for k, v in sorted(prof.items(), key=lambda kv: -kv[1]["total_ms"])}
    line = {
        "metric": "prover ms at d=24 Goldilocks2 (commit + prove, bit-exact)",
        "value": round(ms, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks, exact)",
        "data": (f"synthetic: 2^{nv} uniform Field64 coefficients, 2 uniform Field64_2 points (seed 20241202)"
                 if args.inputs == "uniform" else
                 f"the reference CLI's inputs: v[i] = i (2^{nv} coefficients), points (0,..,0) and (1,..,1)")
                + (f", {args.linear_constraints} covector constraint(s) (0, 1, ..., n-1)" if args.linear_constraints
                   else ""),
        "config": {"workload": f"WHIR PCS Goldilocks2 d={nv} rate 1/4 fold 4 Blake3 100-bit list decoding, "
                               "2 evaluation points, PoW thresholds disabled", "num_variables": nv,
                   "log_inv_rate": 2, "folding_factor": 4, "hash": "Blake3", "evaluations": 2,
                   "linear_constraints": args.linear_constraints, "inputs": args.inputs,
                   "parallelism": "single GPU"},
        "proofs_per_s": round(1e3 / ms, 3),
        "proof_bytes": len(last.proof()[0]) + len(last.proof()[1]),
        "verified_by_oracle_verifier": verified,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "pow": pow_rates,
        "pcie_inclusive": pcie,
        "other_configs": others,
        "batched": batched,
        "serving": serving,
        "kernels": kernels,
        "ntt_schedules": [{k: e[k] for k in ("degree", "log_n", "log_c", "messages", "log_c1", "pass1", "pass2",
                                             "tuned_ms", "default_ms")} for e in W.ntt_schedules()],
    }
    print(json.dumps(line))


def bench_other_configs(W, GP, steps=3):
    """Prover ms of the two other single-instance configurations: d=20 Goldilocks2
    rate 1/2 Blake3, 1 evaluation point (PoW disabled as for the headline: its fold
    PoW is 31-41 bits per round), and d=26 Goldilocks3 rate 1/2 SHA3 with the
    configuration's own PoW (query-phase grinding, 16-20 bits), PCS with one point."""
    import hashlib
    out = {}
    rng = np.random.default_rng(7)
    for name, d, deg, lir, eng, e, dis in (("goldilocks2_d20_blake3", 20, 2, 1, W.HASH_BLAKE3, 1, True),
                                           ("goldilocks3_d26_sha3_pow", 26, 3, 1, W.HASH_SHA3, 1, False)):
        p = GP.ProtocolParameters(degree=deg, security_level=100, pow_bits=20, initial_folding_factor=4,
                                  folding_factor=4, unique_decoding=False, starting_log_inv_rate=lir,
                                  hash_engine=eng)
        cfg = GP.Config(1 << d, p)
        if dis:
            cfg.disable_pow()
        dv = W.DeviceBuffer.from_array(uniform_field(rng, 1 << d))
        pts = uniform_field(rng, e * d * deg)
        ev = np.concatenate([GP.mle_evaluate(deg, dv, d, pts[k * d * deg:(k + 1) * d * deg]) for k in range(e)])
        pid = hashlib.sha3_512(json.dumps(cfg.describe(), sort_keys=True).encode()).digest()
        times = []
        for k in range(steps + 1):
            t = GP.Transcript(pid, hashlib.sha3_256(f"{name} {k}".encode()).digest())
            W.synchronize()
            t0 = time.perf_counter()
            w = GP.commit(cfg, t, dv)
            GP.prove(cfg, t, dv, w, pts, ev)
            W.synchronize()
            times.append((time.perf_counter() - t0) * 1e3)
        out[name] = {"prover_ms": round(float(np.mean(times[1:])), 3), "pow": "disabled" if dis else "as configured",
                     "proof_bytes": sum(len(x) for x in t.proof())}
        del dv
    return out


def bench_pow(W, GP, nv):
    """Grinding rate of the persistent PoW kernel (2^33 nonces against an
    unreachable threshold, so every nonce is hashed), and the expected time of the
    literal PoW of this configuration (every threshold as derived, not disabled):
    sum over PoW steps of 2^64 / (threshold + 1) hashes."""
    out = {}
    ch = bytes(range(32))
    n = 1 << 33
    for name, eng in (("blake3", W.HASH_BLAKE3), ("sha3", W.HASH_SHA3)):
        W.pow_grind_bounded(eng, ch, 0, 1 << 28)  # warm-up
        t0 = time.perf_counter()
        W.pow_grind_bounded(eng, ch, 0, n)
        out[f"{name}_ghashes_per_s"] = round(n / (time.perf_counter() - t0) / 1e9, 2)
    p = GP.ProtocolParameters(degree=2, security_level=100, pow_bits=20, initial_folding_factor=4,
                              folding_factor=4, unique_decoding=False, starting_log_inv_rate=2,
                              hash_engine=W.HASH_BLAKE3)
    desc = GP.Config(1 << nv, p).describe()
    steps = [desc["initial_sumcheck"][2]] * desc["initial_sumcheck"][1]
    for r in desc["rounds"]:
        steps += [r["sumcheck"][2]] * r["sumcheck"][1] + [r["pow"]]
    steps += [desc["final_sumcheck"][2]] * desc["final_sumcheck"][1] + [desc["final_pow"]]
    hashes = sum(2.0 ** 64 / (t + 1) for t in steps if t < 2 ** 64 - 1)
    out["literal_config_expected_hashes"] = float(f"{hashes:.3e}")
    out["literal_config_projected_days"] = round(hashes / (out["blake3_ghashes_per_s"] * 1e9) / 86400, 1)
    out["note"] = ("the timed prover disables PoW (thresholds u64::MAX): this configuration asks for "
                   "42-55 bits per sumcheck round; the projection is the literal PoW at the measured rate")
    return out


class _Instance:
    """One PCS instance: device vector, evaluation points and claimed values."""

    _pids = {}

    def __init__(self, cfg, dv, pts, evals, tag):
        import hashlib
        self.cfg, self.dv, self.pts, self.evals, self.tag = cfg, dv, pts, evals, tag
        # the protocol id depends on the configuration only: derived once per configuration
        if cfg.h not in _Instance._pids:
            _Instance._pids[cfg.h] = hashlib.sha3_512(json.dumps(cfg.describe(), sort_keys=True).encode()).digest()
        self.pid = _Instance._pids[cfg.h]
        self.sid = hashlib.sha3_256(f"{tag} 0".encode()).digest()

    def transcript(self, GP):
        return GP.Transcript(self.pid, self.sid)


def _config(GP, W, d, lir, cache):
    key = (d, lir)
    if key not in cache:
        p = GP.ProtocolParameters(degree=2, security_level=100, pow_bits=20, initial_folding_factor=4,
                                  folding_factor=4, unique_decoding=False, starting_log_inv_rate=lir,
                                  hash_engine=W.HASH_BLAKE3)
        cache[key] = GP.Config(1 << d, p).disable_pow()
    return cache[key]


def _points(GP, rng, dv, d, e):
    pts = uniform_field(rng, e * d * 2)
    evals = np.concatenate([GP.mle_evaluate(2, dv, d, pts[k * d * 2:(k + 1) * d * 2]) for k in range(e)])
    return pts, evals


def bench_batched(workers: int):
    """64 independent PCS instances (Goldilocks2, Blake3, folding 4, rate 1/2,
    num_variables ~ U{16..22}, 1-4 evaluation points, seeded), proved one after the
    other on one stream (the plain path) and through the executor (`workers` proofs
    in flight, one stream each).  Also the commit phase alone both ways."""
    import whir_amd as W
    from whir_amd import prover as GP
    rng = np.random.default_rng(64)
    cache = {}
    insts = []
    for k in range(64):
        d = int(rng.integers(16, 23))
        e = int(rng.integers(1, 5))
        dv = W.DeviceBuffer.from_array(uniform_field(rng, 1 << d))
        pts, evals = _points(GP, rng, dv, d, e)
        insts.append(_Instance(_config(GP, W, d, 1, cache), dv, pts, evals, f"batched {k}"))

    def plain(commit_only):
        W.synchronize()
        t0 = time.perf_counter()
        out = []
        for inst in insts:
            t = inst.transcript(GP)
            w = GP.commit(inst.cfg, t, inst.dv)
            if not commit_only:
                GP.prove(inst.cfg, t, inst.dv, w, inst.pts, inst.evals)
            else:
                W.synchronize()
            out.append(t)
        W.synchronize()
        return time.perf_counter() - t0, out

    # One executor for every timed pass: its workers' streams, buffer pools and pinned
    # mailboxes are set up by the warm-up pass, not inside a timed one.
    ex = GP.Executor(workers=workers)

    def batched(commit_only):
        kind = GP.Executor.COMMIT_ONLY if commit_only else GP.Executor.COMMIT_AND_PROVE
        t0 = time.perf_counter()
        jobs = [(ex.submit(i.cfg, t, i.dv, i.pts, i.evals, kind=kind), t)
                for i, t in ((i, i.transcript(GP)) for i in insts)]
        for jid, _ in jobs:
            ex.wait(jid)
        return time.perf_counter() - t0, [t for _, t in jobs]

    def fused(commit_only):
        # every commit in shared launches (a HIP graph replayed per batch), then the
        # proofs as prove-only jobs on the executor's streams
        W.synchronize()
        t0 = time.perf_counter()
        ts = [i.transcript(GP) for i in insts]
        ws = GP.commit_many([i.cfg for i in insts], ts, [i.dv for i in insts])
        if not commit_only:
            jobs = [ex.submit_prove(i.cfg, t, i.dv, w, i.pts, i.evals) for i, t, w in zip(insts, ts, ws)]
            for jid in jobs:
                ex.wait(jid)
        W.synchronize()
        return time.perf_counter() - t0, ts

    def median_of(fn, reps=5):
        runs = [fn() for _ in range(reps)]
        runs.sort(key=lambda r: r[0])
        return runs[len(runs) // 2]

    try:
        plain(False)  # warm-up: twiddle tables and buffer pools for every size
        batched(False)
        batched(True)
        fused(False)  # records the batch layout's graph
        tp, ref = median_of(lambda: plain(False))
        tb, got = median_of(lambda: batched(False))
        tf, got_f = median_of(lambda: fused(False))
        tpc, _ = median_of(lambda: plain(True))
        tbc, _ = median_of(lambda: batched(True))
        tfc, got_fc = median_of(lambda: fused(True))
    finally:
        ex.close()
    identical = all(a.proof() == b.proof() for a, b in zip(ref, got))
    identical_f = all(a.proof() == b.proof() for a, b in zip(ref, got_f))
    sizes = sorted(int(i.cfg.num_variables) for i in insts)
    return {"instances": len(insts), "workers": workers,
            "num_variables": f"U{{16..22}} (seed 64), median {sizes[len(sizes) // 2]}",
            "plain_proofs_per_s": round(len(insts) / tp, 1),
            "batched_proofs_per_s": round(len(insts) / min(tb, tf), 1),
            "proofs_speedup": round(tp / min(tb, tf), 2),
            "plain_commits_per_s": round(len(insts) / tpc, 1),
            "batched_commits_per_s": round(len(insts) / tfc, 1),
            "commits_speedup": round(tpc / tfc, 2),
            "fused": {"commits_per_s": round(len(insts) / tfc, 1), "proofs_per_s": round(len(insts) / tf, 1),
                      "identical_to_plain": identical_f},
            "streams": {"commits_per_s": round(len(insts) / tbc, 1), "proofs_per_s": round(len(insts) / tb, 1),
                        "identical_to_plain": identical},
            "identical_to_plain": identical and identical_f,
            "method": "fused: whir_commit_many_dev commits all 64 instances in shared launches per shape group "
                      "(a HIP graph recorded on the warm-up pass and replayed), then their proofs run as "
                      "prove-only jobs on the executor's streams; streams: the executor runs commit + prove per "
                      "instance, one HIP stream per worker; plain = the same proofs one after another on one "
                      "stream; batched figures are the faster of fused and streams (commits: fused); each "
                      "figure is the median of 5 passes over the 64 instances after a warm-up pass"}


def bench_serving(workers: int, seconds: float, qps: float):
    """Poisson arrivals of prove requests (Goldilocks2, Blake3, num_variables
    log-uniform in [14, 24], 1-4 evaluations, rate 1/2 or 1/4) into the executor,
    admitted by HBM budget; latency = finish - arrival.  With qps = 0 the rate is
    70% of the throughput measured on a back-to-back burst of the same mix."""
    import threading
    import whir_amd as W
    from whir_amd import prover as GP
    rng = np.random.default_rng(5)
    cache = {}
    vecs = {d: W.DeviceBuffer.from_array(uniform_field(rng, 1 << d)) for d in range(14, 25)}
    pools = {d: [_points(GP, rng, vecs[d], d, 1) for _ in range(4)] for d in range(14, 25)}

    def request(k, r):
        d = int(np.clip(np.floor(np.exp(r.uniform(np.log(14), np.log(25)))), 14, 24))
        e = int(r.integers(1, 5))
        lir = int(r.integers(1, 3))
        sel = [pools[d][j] for j in r.permutation(4)[:e]]
        pts = np.concatenate([p for p, _ in sel])
        ev = np.concatenate([v for _, v in sel])
        return _Instance(_config(GP, W, d, lir, cache), vecs[d], pts, ev, f"serve {k}")

    def run(reqs, arrivals):
        ex = GP.Executor(workers=workers)
        lat, digests = [], []
        jobs = []
        cv = threading.Condition()

        def waiter():
            k = 0
            while True:
                with cv:
                    while k >= len(jobs):
                        cv.wait()
                    jid, t = jobs[k]
                if jid is None:
                    return
                times, _ = ex.wait(jid)
                lat.append(times[2] - times[0])
                digests.append(_proof_digest(t))
                jobs[k] = (0, None)  # the transcript is no longer needed
                k += 1

        th = threading.Thread(target=waiter)
        th.start()
        try:
            t0 = ex.clock()
            next_note = 10.0
            for inst, at in zip(reqs, arrivals):
                delay = at - (ex.clock() - t0)
                if delay > 0:
                    time.sleep(delay)
                if at >= next_note:
                    _note(f"serving: {at:.0f} s, {len(lat)} proofs done")
                    next_note += 10.0
                t = inst.transcript(GP)
                jid = ex.submit(inst.cfg, t, inst.dv, inst.pts, inst.evals)
                with cv:
                    jobs.append((jid, t))
                    cv.notify()
            with cv:
                jobs.append((None, None))
                cv.notify()
            th.join()
            span = ex.clock() - t0
        finally:
            ex.close()
        return lat, span, digests

    r = np.random.default_rng(6)
    burst = [request(k, r) for k in range(4 * workers)]
    run(burst, [0.0] * len(burst))  # warm-up: tables and pools for every size
    lat_b, span_b, _ = run(burst, [0.0] * len(burst))
    capacity = len(burst) / span_b
    rate = qps if qps > 0 else 0.7 * capacity
    gaps = r.exponential(1.0 / rate, size=int(rate * seconds * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < seconds]
    reqs = [request(k, r) for k in range(len(arrivals))]
    lat, span, digests = run(reqs, arrivals)
    lat_ms = np.sort(np.array(lat) * 1e3)
    # every request proved again alone (one after another on this thread's stream):
    # the same proof, compared by digest
    same, t0 = 0, time.perf_counter()
    for k, inst in enumerate(reqs):
        t = inst.transcript(GP)
        w = GP.commit(inst.cfg, t, inst.dv)
        GP.prove(inst.cfg, t, inst.dv, w, inst.pts, inst.evals)
        same += _proof_digest(t) == digests[k]
        if k % 5000 == 4999:
            _note(f"serving check: {k + 1}/{len(reqs)} requests proved alone ({time.perf_counter() - t0:.0f} s)")
    return {"seconds": seconds, "arrival_qps": round(rate, 1), "burst_capacity_proofs_per_s": round(capacity, 1),
            "requests": len(reqs), "throughput_proofs_per_s": round(len(reqs) / span, 1),
            "p50_ms": round(float(np.percentile(lat_ms, 50)), 3), "p99_ms": round(float(np.percentile(lat_ms, 99)), 3),
            "mean_ms": round(float(lat_ms.mean()), 3), "workers": workers,
            "mix": "num_variables log-uniform [14,24], 1-4 evaluations, log_inv_rate 1 or 2, Poisson arrivals",
            "identical_alone": same == len(reqs), "identical_alone_count": f"{same}/{len(reqs)}",
            "check": "every served proof compared (SHA-256 of narg and hints) with the same request proved alone"}


def cpu_baseline(nv, D, vec, points, evals, pid, sid, proof, covectors=()):
    """The oracle prover (oracle/whir_protocol.py over oracle/whir_oracle.c, its
    large loops split over all host threads OpenMP gives it) on the same instance
    as the timed GPU proof: same configuration, vector, points and transcript
    ids.  One commit + prove, timed; its proof bytes are compared with the GPU's."""
    from oracle import whir_protocol as WP
    from oracle import oracle as O
    cores = O.threads()
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, initial_folding_factor=4, folding_factor=4,
                              unique_decoding=False, starting_log_inv_rate=2, hash_engine=O.ENGINE_BLAKE3)
    cfg = WP.WhirConfig.new(1 << nv, p, D)
    cfg.disable_pow()
    pts = [WP.Covector(c) for c in covectors] + [[pt[i * D:(i + 1) * D] for i in range(nv)] for pt in points]
    evs = [np.asarray(e, dtype=np.uint64) for e in evals]
    t0 = time.perf_counter()
    t = WP.Transcript(pid, sid)
    w = WP.commit(t, cfg, vec)
    t1 = time.perf_counter()
    WP.prove(t, cfg, vec, w, pts, evs)
    t2 = time.perf_counter()
    same = (bytes(t.narg), bytes(t.hints)) == proof
    return {"value": round((t2 - t0) * 1e3, 1), "unit": "ms", "cores": cores, "kind": "port",
            "sample": f"restatement, not reference (the Rust reference cannot be built here): "
                      f"one oracle commit + prove of the timed instance itself (d={nv}, Goldilocks{D}, "
                      f"Blake3, rate 1/4, folding 4, PoW off), OpenMP over {cores} host threads: "
                      f"commit {(t1 - t0):.2f} s, prove {(t2 - t1):.2f} s",
            "proof_identical_to_gpu": bool(same)}


if __name__ == "__main__":
    main()
