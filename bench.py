#!/usr/bin/env python3
"""Benchmark of the vkzg hot path on one MI355X.

Headline: computeProofsBatch (FK20, 128 cell proofs per blob) on a 256-blob
batch with the blobs already resident in HBM; one step = one 256-blob call
through vkzg_amd_compute_proofs_device.  Prints ONE JSON line.

Also measured (in "secondary"): recoverCellsBatch on 64 blobs with exactly 64
of 128 cells present, the plain one-call-per-blob proof form, and the
per-stage HIP-event times of the timed steps.  The CPU baseline is the
oracle (a C port of the reference's algorithm) on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import random
import sys
import time

R = 0x73EDA753299D7D483339D80809A1D80553BDA402FFFE5BFEFFFFFFFF00000001
PEAK_INT = 256 * 2.4e9 * 64  # CU x clock x lanes: lane-ops/s (v_mad_u64_u32 issues at full rate)
FP_M = 300                    # limb-MACs per Fp multiplication (12-limb CIOS)
FR_M = 136                    # limb-MACs per Fr multiplication (8-limb CIOS)
W_PROOF = 1.734e9             # limb-MACs per blob of the reference's FK20 as written
METRIC = "proof blobs/s @256-blob batch; recovery blobs/s; p50/p99 serve latency"
HERE = os.path.dirname(os.path.abspath(__file__))


def seeded_blobs(n, seed):
    rng = random.Random(seed)
    return b"".join(rng.randrange(R).to_bytes(32, "big") for _ in range(4096 * n))


def stage_work(info, nb, items_per_output):
    """Algorithmic limb-MACs per launch of each proof stage (formula counts)."""
    W = info["windows"]
    madd, add, dbl_j, add_j = 10, 14, 7, 16  # Fp multiplications per operation
    return {
        "fk20_scalars": nb * (12 * 2048 + 4096 + 64 * 448 + 8192 + 4096) * FR_M,
        "phase_a": nb * 128 * (64 * W * madd + 15 * add) * FP_M,
        "ladder": nb * 128 * 248 * dbl_j * FP_M,
        "fold": nb * 64 * 32 * 2 * add_j * FP_M,
        "phase_b": nb * 128 * items_per_output * madd * FP_M,
        "bucket_reduce": nb * 128 * (2 * 128 + 8) * add * FP_M,
    }


def cpu_baseline(sample_blobs, seed):
    """The oracle's FK20 (a port of the reference algorithm) on a few blobs."""
    threads = min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import oracle as O
    setup = open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "vkzg_amd", "data",
                              "g1_monomial_mainnet.bin"), "rb").read()
    S = O.Setup(setup)
    blobs = seeded_blobs(sample_blobs, seed + 1)
    t0 = time.perf_counter()
    rc, _ = S.compute_proofs(blobs, sample_blobs)
    dt = time.perf_counter() - t0
    assert rc == 0
    return {"value": sample_blobs / dt, "unit": "blobs/s", "cores": threads, "kind": "port",
            "sample": "%d blobs of computeProofsBatch through oracle/oracle.c (C port of the "
                      "reference's FK20: bucket phase A on the 2^(8d) position table, G1 FFT "
                      "phase B), %.1f s; setup build excluded; the reference itself needs "
                      "Vulkan/lavapipe, absent here" % (sample_blobs, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--batch", type=int, default=256)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-sample-blobs", type=int, default=8)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-secondary", action="store_true")
    ap.add_argument("--traffic-file", default=os.path.join(HERE, "profiles", "phase_a_traffic.json"),
                    help="per-launch HBM traffic measured by rocprofv3 --pmc (FETCH_SIZE, WRITE_SIZE)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (counter calibration)")
    a = ap.parse_args()
    if a.calibrate_copy:
        from vkzg_amd import hip
        n = a.calibrate_copy << 20
        src, dst = hip.DeviceBuffer(n), hip.DeviceBuffer(n)
        src.zero()
        for _ in range(3):
            hip.check(hip.lib().hipMemcpy(dst.ptr, src.ptr, n, hip.D2D))
        print(json.dumps({"calibrate_copy_bytes": n, "copies": 3}))
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: this build runs the whole batch on a single "
              "MI355X and has no multi-GPU path (blobs shard with no exchange step).",
              file=sys.stderr)
        sys.exit(2)

    import vkzg_amd
    from vkzg_amd import hip

    nb = a.batch
    blobs = seeded_blobs(nb, a.seed)
    prover = vkzg_amd.Prover(max_batch_size=max(nb, 64))
    info = prover.info()
    d_blobs = hip.DeviceBuffer.from_bytes(blobs)
    d_proofs = hip.DeviceBuffer(nb * 6144)
    stream = hip.Stream()

    for _ in range(a.warmup):
        prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, nb, stream.handle)
    prover.set_stage_timing(True)
    stage_tot = {}
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, nb, stream.handle)
        for k, v in prover.last_stage_times().items():
            stage_tot[k] = stage_tot.get(k, 0.0) + v
    stream.synchronize()
    elapsed = time.perf_counter() - t0
    prover.set_stage_timing(False)
    value = nb * a.steps / elapsed
    stage_ms = {k: v / a.steps for k, v in stage_tot.items()}

    items_per_output = 1048.0  # phase-B items per output (66 split-circulant scalars)
    work = stage_work(info, nb, items_per_output)
    frac = {k: (work[k] / (stage_ms[k] * 1e-3)) / PEAK_INT for k in work if stage_ms.get(k)}
    dom = max(stage_ms, key=stage_ms.get)
    dom_work = work.get(dom)
    traffic = None
    if os.path.exists(a.traffic_file):
        t = json.load(open(a.traffic_file))
        if t.get("kernel") == dom and t.get("blobs") == nb:
            traffic = t["fetch_bytes"] + t["write_bytes"]
    roofline = {
        "bound": "int-valu",
        "kernel": dom,
        "achieved": (dom_work / (stage_ms[dom] * 1e-3)) / 1e12 if dom_work else None,
        "peak": PEAK_INT / 1e12,
        "unit": "T limb-MAC/s",
        "frac": frac.get(dom),
        "traffic": traffic,
    }

    secondary = {"stage_ms": stage_ms, "stage_frac_of_int_peak": frac,
                 "pipeline_frac_of_int_peak_vs_reference_work": value * W_PROOF / PEAK_INT,
                 "phase_a_window_bits": info["window_bits"],
                 "table_gib": info["table_bytes"] / 2**30, "setup_ms": info["setup_ms"]}
    if not a.no_secondary:
        # recoverCellsBatch, 64 blobs, exactly 64 of 128 cells present
        rng = random.Random(a.seed + 2)
        nr = 64
        from vkzg_amd import BYTES_PER_CELL
        cells = rng.randbytes(nr * 128 * BYTES_PER_CELL)
        masks = []
        for _ in range(nr):
            keep = set(rng.sample(range(128), 64))
            masks.append(bytes(1 if k in keep else 0 for k in range(128)))
        present = b"".join(masks)
        d_cells = hip.DeviceBuffer.from_bytes(cells)
        d_present = hip.DeviceBuffer.from_bytes(present)
        d_out = hip.DeviceBuffer(len(cells))
        for _ in range(2):
            prover.recover_cells_device(d_out.ptr, d_cells.ptr, d_present.ptr, present, nr,
                                        stream.handle)
        reps = 20
        t0 = time.perf_counter()
        for _ in range(reps):
            prover.recover_cells_device(d_out.ptr, d_cells.ptr, d_present.ptr, present, nr,
                                        stream.handle)
        rec_dt = (time.perf_counter() - t0) / reps
        secondary["recovery_64_blobs_per_s"] = nr / rec_dt
        secondary["recovery_64_ms"] = rec_dt * 1e3
        prover.set_stage_timing(True)
        prover.recover_cells_device(d_out.ptr, d_cells.ptr, d_present.ptr, present, nr,
                                    stream.handle)
        secondary["recovery_stage_ms"] = prover.last_stage_times()
        prover.set_stage_timing(False)
        # uneven recovery: call sizes U[1,128], per-blob present count U[64,127]
        calls = [rng.randint(1, 128) for _ in range(12)]
        maxn = max(calls)
        ucells = rng.randbytes(maxn * 128 * BYTES_PER_CELL)
        d_ucells = hip.DeviceBuffer.from_bytes(ucells)
        d_uout = hip.DeviceBuffer(len(ucells))
        masks = []
        for _ in range(maxn):
            keep = set(rng.sample(range(128), rng.randint(64, 127)))
            masks.append(bytes(1 if k in keep else 0 for k in range(128)))
        umask = b"".join(masks)
        d_umask = hip.DeviceBuffer.from_bytes(umask)
        prover.recover_cells_device(d_uout.ptr, d_ucells.ptr, d_umask.ptr, umask, maxn, stream.handle)
        t0 = time.perf_counter()
        for n in calls:
            prover.recover_cells_device(d_uout.ptr, d_ucells.ptr, d_umask.ptr, umask[:128 * n], n,
                                        stream.handle)
        secondary["uneven_recovery_blobs_per_s"] = sum(calls) / (time.perf_counter() - t0)
        secondary["uneven_recovery_calls"] = calls
        # plain form: one blob per call through the host C ABI
        one = blobs[:131072]
        prover.compute_cell_kzg_proofs_batch(one, 1)
        calls = 8
        t0 = time.perf_counter()
        for _ in range(calls):
            prover.compute_cell_kzg_proofs_batch(one, 1)
        secondary["plain_per_call_1_blob_blobs_per_s"] = calls / (time.perf_counter() - t0)

    cpu = None
    if not a.no_cpu_baseline:
        cpu = cpu_baseline(a.cpu_sample_blobs, a.seed)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "blobs/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": elapsed / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {"workload": "computeProofsBatch, %d blobs per call, blobs resident in HBM" % nb,
                   "global_batch": nb, "blob": "4096 seeded uniform Fr elements in [0,r)",
                   "setup": "mainnet G1 monomial (throughput does not depend on the setup)",
                   "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
