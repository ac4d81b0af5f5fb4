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
disable_pow), and the grinding kernel is measured separately (pow_mhs).

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


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--num-variables", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-verify", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run one device-to-device copy of this many bytes (a known-size "
                         "reference for the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


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
    vec = uniform_field(rng, 1 << nv)
    points = [uniform_field(rng, nv * D) for _ in range(2)]
    dvec = W.DeviceBuffer.from_array(vec)
    evals = [GP.mle_evaluate(D, dvec, nv, pt) for pt in points]
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
        GP.prove(cfg, t, dvec, w, flat_pts, flat_ev)
        return t

    for _ in range(args.warmup):
        step()
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

    verified = None
    if not args.no_verify:
        from oracle import whir_protocol as WP
        ocfg = WP.WhirConfig.new(1 << nv, WP.ProtocolParameters(
            security_level=100, pow_bits=20, initial_folding_factor=4, folding_factor=4, unique_decoding=False,
            starting_log_inv_rate=2, hash_engine=W.HASH_BLAKE3), D)
        ocfg.disable_pow()
        narg, hints = last.proof()
        pts_list = [[p[i * D:(i + 1) * D] for i in range(nv)] for p in points]
        try:
            WP.verify(WP.TranscriptReader(pid, sid, narg, hints), ocfg, [e for e in evals], pts_list)
            verified = True
        except Exception as e:  # reported, never hidden
            verified = f"rejected: {e}"

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(nv)

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
        "data": "synthetic: 2^24 uniform Field64 coefficients, 2 uniform Field64_2 points (seed 20241202)",
        "config": {"workload": f"WHIR PCS Goldilocks2 d={nv} rate 1/4 fold 4 Blake3 100-bit list decoding, "
                               "2 evaluation points, PoW thresholds disabled", "num_variables": nv,
                   "log_inv_rate": 2, "folding_factor": 4, "hash": "Blake3", "evaluations": 2,
                   "parallelism": "single GPU"},
        "proofs_per_s": round(1e3 / ms, 3),
        "proof_bytes": len(last.proof()[0]) + len(last.proof()[1]),
        "verified_by_oracle_verifier": verified,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "kernels": kernels,
    }
    print(json.dumps(line))


def cpu_baseline(nv_target: int):
    """The oracle prover (oracle/whir_protocol.py over oracle/whir_oracle.c, one host
    thread) on a bounded sample of the same configuration, scaled to d = nv_target."""
    from oracle import whir_protocol as WP
    from oracle import oracle as O
    nv = 18
    D = 2
    p = WP.ProtocolParameters(security_level=100, pow_bits=20, initial_folding_factor=4, folding_factor=4,
                              unique_decoding=False, starting_log_inv_rate=2, hash_engine=O.ENGINE_BLAKE3)
    cfg = WP.WhirConfig.new(1 << nv, p, D)
    cfg.disable_pow()
    rng = np.random.default_rng(1)
    vec = uniform_field(rng, 1 << nv)
    pts = [[uniform_field(rng, D) for _ in range(nv)] for _ in range(2)]
    evals = [WP.mle_evaluate_base(D, vec, pt) for pt in pts]
    t0 = time.perf_counter()
    reps = 0
    while True:
        t = WP.Transcript(b"\0" * 64, b"\0" * 32)
        w = WP.commit(t, cfg, vec)
        WP.prove(t, cfg, vec, w, pts, evals)
        reps += 1
        if time.perf_counter() - t0 > 10.0:
            break
    ms_sample = (time.perf_counter() - t0) * 1e3 / reps
    scale = (1 << (nv_target - nv)) * nv_target / nv
    return {"value": round(ms_sample * scale, 1), "unit": "ms", "cores": 1, "kind": "port",
            "sample": f"oracle prover (single thread, obvious __int128 arithmetic) at d={nv}, same "
                      f"parameters, {reps} proofs in {ms_sample * reps / 1e3:.1f} s = {ms_sample:.1f} ms each; "
                      f"scaled by 2^{nv_target - nv} * {nv_target}/{nv} (n log n) to d={nv_target}"}


if __name__ == "__main__":
    main()
