#!/usr/bin/env python3
"""Benchmark of the vkzg hot path on one MI355X.

Headline: computeProofsBatch (FK20, 128 cell proofs per blob) on a 256-blob
batch with the blobs already resident in HBM; one step = one 256-blob call
through vkzg_amd_compute_proofs_device, which replays the call's captured HIP
graph (the library's default path).  Prints ONE JSON line.

Also measured (in "secondary"): recoverCellsBatch on 64 blobs with exactly 64
of 128 cells present, the plain one-call-per-blob proof form, the per-stage
HIP-event times of a second pass over the same steps (kernels launched one by
one between events; its own rate is stage_timed_blobs_per_s), batch
cell-proof verification and
commitments on a seeded setup, and the request server.  The CPU baseline is the
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
# Algorithmic work of one blob's recovery (lean form: 6 NTT_8192, pointwise
# products, batch-inverted division, vanishing polynomial, I/O), in limb-MACs.
W_REC = 6.38e7
LAD_POS = 16                  # ladder positions 2^(8d) u, d < 16 (8-bit digits of the balanced GLV halves of the phase-B scalars)
DIGIT_BITS = 8                # doublings between ladder positions
BUCKETS = 128                 # phase-B buckets per output (digit magnitudes 1..128)
METRIC = "proof blobs/s @256-blob batch; recovery blobs/s; p50/p99 serve latency"
HERE = os.path.dirname(os.path.abspath(__file__))


def seeded_blobs(n, seed):
    rng = random.Random(seed)
    return b"".join(rng.randrange(R).to_bytes(32, "big") for _ in range(4096 * n))


def stage_work(info, nb, items_per_output):
    """Algorithmic limb-MACs per launch of each proof stage (formula counts)."""
    W = info["windows"]
    madd, add, dbl_j = 10, 14, 7  # Fp multiplications per operation
    return {
        "fk20_scalars": nb * (12 * 2048 + 4096 + 64 * 448 + 8192 + 4096) * FR_M,
        # 8 lanes per MSM, 8 terms each, then a 3-level lane tree (7 adds)
        "phase_a": nb * 128 * (64 * W * madd + 7 * add) * FP_M,
        # 128 chains of 8 (LAD_POS - 1) doublings, after the split-circulant
        # folds of u (3 additions for each of 64 chains, 1 for the other 64)
        "ladder": nb * (128 * (LAD_POS - 1) * DIGIT_BITS * dbl_j + 256 * add) * FP_M,
        "phase_b": nb * 128 * items_per_output * madd * FP_M,
        "bucket_reduce": nb * 128 * (2 * BUCKETS + 8) * add * FP_M,
    }


def host_abi_sweep(prover, blobs, nmax, seed, reps=10):
    """The reference's own convention (bench/bench.cpp:38-60 and its recovery
    sweep): the plain C ABI on host buffers, host<->device copies included,
    1, 2, 4, ... nmax blobs per call, one warm-up call then `reps` timed calls,
    best and mean.  Recovery keeps the first 64 of 128 cells of each blob."""
    import ctypes
    import vkzg_amd
    L = vkzg_amd._lib.load()
    h = prover.handle
    src = ctypes.create_string_buffer(blobs, len(blobs))
    out = ctypes.create_string_buffer(nmax * 6144)
    rng = random.Random(seed)
    cells = ctypes.create_string_buffer(rng.randbytes(nmax * 262144), nmax * 262144)
    present = ctypes.create_string_buffer((bytes([1]) * 64 + bytes(64)) * nmax, nmax * 128)
    rec_out = ctypes.create_string_buffer(nmax * 262144)

    def sweep(call):
        res = {}
        n = 1
        while n <= nmax:
            assert call(n) == 0
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                rc = call(n)
                times.append(time.perf_counter() - t0)
                assert rc == 0
            res[str(n)] = {"best_ms": min(times) * 1e3, "mean_ms": sum(times) / reps * 1e3,
                           "blobs_per_s_best": n / min(times)}
            n *= 2
        return res

    return {
        "convention": "plain C ABI on host buffers (copies included), 1 warm-up + %d reps, best and mean" % reps,
        "proofs": sweep(lambda n: L.vkzg_compute_proofs_batch(h, out, src, n)),
        "recovery_first_64_present": sweep(lambda n: L.vkzg_recover_cells_batch(h, rec_out, cells, present, n)),
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


def verify_bench(seed, nb=64, reps=5):
    """verify_cell_kzg_proof_batch and blob_to_kzg_commitment through the host C
    ABI (host copies and the SHA-256 challenge included) on the seeded setup of
    tests/golden/seeded_setup.bin, whose G2 side makes every verdict real: the
    proofs and commitments are the device's own, and every timed call must
    return valid."""
    import ctypes
    import vkzg_amd
    from vkzg_amd import load
    data = open(os.path.join(HERE, "tests", "golden", "seeded_setup.bin"), "rb").read()
    vp = vkzg_amd.Prover(max_batch_size=nb, setup_g1_monomial=data[:4096 * 48])
    vp.load_g2_setup(data[4096 * 48:])
    blobs = seeded_blobs(nb, seed)
    cells, proofs = vp.compute_cells_and_kzg_proofs_batch(blobs, nb)
    coms = vp.blob_to_kzg_commitments(blobs, nb)
    L = load()
    out = {}
    t0 = time.perf_counter()
    for _ in range(3):
        assert vp.blob_to_kzg_commitments(blobs, nb) == coms
    out["commitments_%d_blobs_per_s" % nb] = 3 * nb / (time.perf_counter() - t0)

    def timed(ncells, picks):
        c = ctypes.create_string_buffer(b"".join(coms[48 * b:48 * (b + 1)] for b, _ in picks), 48 * ncells)
        idx = (ctypes.c_uint64 * ncells)(*[i for _, i in picks])
        ce = ctypes.create_string_buffer(b"".join(cells[(128 * b + i) * 2048:(128 * b + i + 1) * 2048]
                                                  for b, i in picks), 2048 * ncells)
        pr = ctypes.create_string_buffer(b"".join(proofs[(128 * b + i) * 48:(128 * b + i + 1) * 48]
                                                  for b, i in picks), 48 * ncells)
        ok = ctypes.c_int(0)

        def call():
            rc = L.vkzg_amd_verify_cell_kzg_proof_batch(vp.handle, ctypes.byref(ok), c, idx, ce, pr, ncells)
            assert rc == 0 and ok.value == 1
        call()
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        dt = (time.perf_counter() - t0) / reps
        t0 = time.perf_counter()
        vkzg_amd.verify_challenge(c.raw, list(idx), ce.raw, pr.raw)
        return dt, time.perf_counter() - t0

    all_picks = [(b, i) for b in range(nb) for i in range(128)]
    dt, hash_dt = timed(len(all_picks), all_picks)
    out["verify_%d_cells_per_s" % len(all_picks)] = len(all_picks) / dt
    out["verify_%d_cells_ms" % len(all_picks)] = dt * 1e3
    out["verify_%d_cells_challenge_ms" % len(all_picks)] = hash_dt * 1e3
    for ncells in (128, 1):
        dt, hash_dt = timed(ncells, all_picks[:ncells])
        out["verify_%d_cells_ms" % ncells] = dt * 1e3
    out["verify_setup"] = ("seeded tau (tests/golden/seeded_setup.bin), %d blobs' cells, proofs and "
                           "commitments computed on the device; every timed call returned valid" % nb)
    vp.close()
    return out


def serve_bench(prover, rng, seconds, rate, probe_rate, slots):
    """Poisson request stream through vkzg_amd.Server (70/30 proof/recovery mix) plus
    6-blob proof probes; then the same requests replayed one call each
    through the synchronous C ABI (the plain per-call form)."""
    import ctypes
    import threading
    import vkzg_amd
    from vkzg_amd import BYTES_PER_BLOB, BYTES_PER_CELL

    cells_per_blob = 128 * BYTES_PER_CELL
    blob_pool = ctypes.create_string_buffer(seeded_blobs(48, rng.randrange(1 << 30)), 48 * BYTES_PER_BLOB)
    raw = bytearray(rng.randbytes(64 * cells_per_blob))
    raw[0::32] = bytes(b & 0x3F for b in raw[0::32])  # canonical field elements (< r)
    cell_pool = ctypes.create_string_buffer(bytes(raw), 64 * cells_per_blob)
    # schedule
    reqs, t = [], 0.0
    while t < seconds:
        t += rng.expovariate(rate)
        if rng.random() < 0.7:
            reqs.append([t, "p", rng.randint(1, 48), None])
        else:
            k = rng.randint(1, 64)
            masks = []
            for _ in range(k):
                keep = set(rng.sample(range(128), rng.randint(64, 127)))
                masks.append(bytes(1 if q in keep else 0 for q in range(128)))
            reqs.append([t, "r", k, ctypes.create_string_buffer(b"".join(masks), 128 * k)])
    t = 0.0
    while t < seconds:
        t += rng.expovariate(probe_rate)
        reqs.append([t, "probe", 6, None])
    reqs.sort(key=lambda r: r[0])
    n_out = 128
    out_bufs = [ctypes.create_string_buffer(64 * cells_per_blob) for _ in range(n_out)]

    def submit(srv, i, r):
        out = ctypes.addressof(out_bufs[i % n_out])
        if r[1] == "r":
            return srv.submit_recovery_raw(out, ctypes.addressof(cell_pool), ctypes.addressof(r[3]), r[2])
        return srv.submit_proofs_raw(out, ctypes.addressof(blob_pool), r[2])

    def run(srv, paced):
        tickets = [None] * len(reqs)
        infos = [None] * len(reqs)
        ready = threading.Semaphore(0)

        def waiter():
            for i in range(len(reqs)):
                ready.acquire()
                rc, info = srv.wait_raw(tickets[i])
                assert rc == 0, rc
                infos[i] = info

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            if paced:
                delay = r[0] - (time.perf_counter() - t0)
                if delay > 0:
                    time.sleep(delay)
            while i >= n_out and infos[i - n_out] is None:  # output buffer reuse
                time.sleep(0.0002)
            tickets[i] = submit(srv, i, r)
            ready.release()
        th.join()
        first = min(x["submit_ns"] for x in infos)
        last = max(x["done_ns"] for x in infos)
        return infos, (last - first) * 1e-9

    blobs = sum(r[2] for r in reqs)
    with vkzg_amd.Server(prover, slots=slots) as srv:
        run(srv, False)  # warm-up
        infos, wall = run(srv, True)
        sat_infos, sat_wall = run(srv, False)
        st = srv.stats()
    lat = sorted((x["done_ns"] - x["submit_ns"]) * 1e-6 for x in infos)
    probe = sorted((x["done_ns"] - x["submit_ns"]) * 1e-6 for x, r in zip(infos, reqs) if r[1] == "probe")
    pct = lambda v, q: v[min(len(v) - 1, int(q * len(v)))]
    # Plain per-call forms: the same requests, one synchronous call each.
    def per_call():
        t0 = time.perf_counter()
        for i, r in enumerate(reqs):
            out = out_bufs[i % n_out]
            if r[1] == "r":
                rc = vkzg_amd.load().vkzg_recover_cells_batch(prover.handle, out, cell_pool, r[3], r[2])
            else:
                rc = vkzg_amd.load().vkzg_compute_proofs_batch(prover.handle, out, blob_pool, r[2])
            assert rc == 0
        return time.perf_counter() - t0
    # (1) the step-before-batching form: one fixed launch shape for every
    # size and kernels launched one by one (no per-size tuning, no graphs)
    saved = {n: prover.get_tuning(n) for n in (1, 4, 16, 64, 256)}
    fixed = dict(phase_a_groups=1, phase_a_blocks_per_cu=2, normalize_chunk=32, splits=1,
                 recovery_waves=1, ladder_splits=1, reduce_lanes=8)
    for n in saved:
        prover.set_tuning(n, **fixed)
    prover.set_graphs(False)
    plain_wall = per_call()
    prover.set_graphs(True)
    for n, t in saved.items():
        prover.set_tuning(n, **t)
    # (2) the same per-request calls with this build's tuned, graph-replayed pipeline
    tuned_wall = per_call()
    return {
        "offered_requests_per_s": rate + probe_rate, "requests": len(reqs), "blobs": blobs,
        "served_blobs_per_s": blobs / wall, "p50_ms": pct(lat, 0.50), "p99_ms": pct(lat, 0.99),
        "probe_6blob_proof_p50_ms": pct(probe, 0.50), "probe_6blob_proof_p99_ms": pct(probe, 0.99),
        "saturated_served_blobs_per_s": blobs / sat_wall,
        "plain_per_call_blobs_per_s": blobs / plain_wall,
        "plain_per_call_form": "each request one synchronous C-ABI call, one fixed launch shape, "
                               "kernels launched one by one (no per-size tuning, no graphs)",
        "served_over_plain": (blobs / sat_wall) / (blobs / plain_wall),
        "tuned_per_call_blobs_per_s": blobs / tuned_wall,
        "served_over_tuned_per_call": (blobs / sat_wall) / (blobs / tuned_wall),
        "server": st,
        "mix": "70%% proofs k~U[1,48], 30%% recovery k~U[1,64] with 64-127 of 128 cells present, "
               "plus 6-blob proof probes at %g/s; Poisson arrivals at %g requests/s"
               % (probe_rate, rate),
    }


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
    ap.add_argument("--no-autotune", action="store_true", help="keep the default tunings")
    ap.add_argument("--no-verify", action="store_true", help="skip the verification/commitment leg")
    ap.add_argument("--serve-seconds", type=float, default=4.0)
    ap.add_argument("--serve-rate", type=float, default=150.0, help="mix requests per second")
    ap.add_argument("--probe-rate", type=float, default=25.0, help="6-blob proof probes per second")
    ap.add_argument("--serve-slots", type=int, default=2)
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

    tuning = None
    if not a.no_autotune:
        autotune_ms = prover.autotune()
        tuning = {"autotune_ms": autotune_ms,
                  "classes": {str(n): prover.get_tuning(n) for n in (1, 4, 16, 64, 256)}}
    for _ in range(a.warmup):
        prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, nb, stream.handle)
    # The timed steps: the library's default path, one captured HIP graph
    # replayed per call (the synchronous entry points do the same).
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, nb, stream.handle)
    stream.synchronize()
    elapsed = time.perf_counter() - t0
    value = nb * a.steps / elapsed

    # The same steps again with per-stage HIP events (kernels launched one by
    # one between the events): the stage breakdown and the roofline.
    prover.set_stage_timing(True)
    stage_tot = {}
    stream.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        prover.compute_proofs_device(d_proofs.ptr, d_blobs.ptr, nb, stream.handle)
        for k, v in prover.last_stage_times().items():
            stage_tot[k] = stage_tot.get(k, 0.0) + v
    stream.synchronize()
    stage_timed_value = nb * a.steps / (time.perf_counter() - t0)
    prover.set_stage_timing(False)
    stage_ms = {k: v / a.steps for k, v in stage_tot.items()}

    items_per_output = info["phase_b_items"] / 128.0  # one blob's items over its 128 outputs
    work = stage_work(info, nb, items_per_output)
    frac = {k: (work[k] / (stage_ms[k] * 1e-3)) / PEAK_INT for k in work if stage_ms.get(k)}
    dom = max(stage_ms, key=stage_ms.get)
    dom_work = work.get(dom)
    traffic = None
    if os.path.exists(a.traffic_file):
        t = json.load(open(a.traffic_file))
        if (t.get("kernel") == dom and t.get("blobs") == nb and t.get("window_bits") == info["window_bits"]
                and t.get("windows") == info["windows"]):
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

    # ceiling of this build's Fp multiplication: the chip's rate of dependent
    # products at 2 waves per SIMD (the hot kernels' occupancy) and at 8
    fp_rates = {str(w): prover.debug_fp_mul_rate(w)[0] for w in (1, 2, 4, 8)}
    one_wave_cycles = prover.debug_fp_mul_rate(1)[1]
    phase_a_mults = nb * 128 * (64 * info["windows"] * 10 + 15 * 14)
    fp_ceiling = {"fp_mul_per_s_by_waves_per_simd": fp_rates, "one_wave_cycles_per_fp_mul": one_wave_cycles,
                  "phase_a_fp_mul_per_s": phase_a_mults / (stage_ms["phase_a"] * 1e-3),
                  "phase_a_over_2_wave_ceiling": phase_a_mults / (stage_ms["phase_a"] * 1e-3) / fp_rates["2"]}
    # the measured issue rate of v_mad_u64_u32 (8 independent chains per lane)
    # beside the nominal peak the roofline divides by
    mad_rates = {str(w): prover.debug_mad_rate(w)[0] for w in (1, 2, 4, 8)}
    clock_mhz = prover.debug_mad_rate(1, 10)[1]
    measured_peak = max(mad_rates.values())
    roofline["measured_peak"] = measured_peak / 1e12
    roofline["frac_of_measured_peak"] = (dom_work / (stage_ms[dom] * 1e-3)) / measured_peak if dom_work else None
    mad_issue = {"mad_u64_u32_per_s_by_waves_per_simd": mad_rates, "nominal_clock_mhz": clock_mhz,
                 "measured_over_nominal_peak": measured_peak / PEAK_INT}
    secondary = {"stage_timed_blobs_per_s": stage_timed_value, "tuning": tuning, "fp_mul_ceiling": fp_ceiling,
                 "mad_issue_rate": mad_issue,
                 "stage_ms": stage_ms, "stage_frac_of_int_peak": frac,
                 "phase_a_window_bits": info["window_bits"],
                 "table_gib": info["table_bytes"] / 2**30, "setup_ms": info["setup_ms"]}
    if not a.no_secondary:
        # recoverCellsBatch, 64 blobs, exactly 64 of 128 cells present.  The
        # cells are the extension of the seeded blobs (this library's cells
        # and proofs call); the missing cells hold random bytes, and every
        # recovered blob is compared with its original cells afterwards (the
        # encode -> erase -> decode round trip).
        rng = random.Random(a.seed + 2)
        nr = 64
        from vkzg_amd import BYTES_PER_CELL
        cell_bytes = 128 * BYTES_PER_CELL
        ncells = min(nb, 128)
        full = b"".join(prover.compute_cells_and_kzg_proofs_batch(blobs[i * 131072:(i + 64) * 131072],
                                                                  min(64, ncells - i))[0]
                        for i in range(0, ncells, 64))

        def erase(b, keep):
            c = bytearray(rng.randbytes(cell_bytes))
            for k in keep:
                c[2048 * k:2048 * (k + 1)] = full[b * cell_bytes + 2048 * k:b * cell_bytes + 2048 * (k + 1)]
            return bytes(c), bytes(1 if k in keep else 0 for k in range(128))

        pairs = [erase(b % ncells, set(rng.sample(range(128), 64))) for b in range(nr)]
        cells = b"".join(p_[0] for p_ in pairs)
        present = b"".join(p_[1] for p_ in pairs)
        want = b"".join(full[(b % ncells) * cell_bytes:((b % ncells) + 1) * cell_bytes] for b in range(nr))
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
        if d_out.to_bytes() != want:
            raise SystemExit("recovery of the 64-blob workload returned wrong cells")
        secondary["recovery_64_blobs_per_s"] = nr / rec_dt
        secondary["recovery_64_frac_of_int_peak"] = nr / rec_dt * W_REC / PEAK_INT
        secondary["recovery_64_ms"] = rec_dt * 1e3
        secondary["recovery_checked"] = "every blob of the 64-blob and the last uneven call equals its original cells"
        prover.set_stage_timing(True)
        prover.recover_cells_device(d_out.ptr, d_cells.ptr, d_present.ptr, present, nr,
                                    stream.handle)
        rec_stages = prover.last_stage_times()
        secondary["recovery_stage_ms"] = rec_stages
        prover.set_stage_timing(False)
        # Each of the five recovery kernels reads 256 KiB and writes 256 KiB
        # per blob (8192 values of 32 bytes, or the 2048-byte cells of the
        # first and last), plus the 128-byte mask of the first and third.
        pass_bytes = nr * 2 * 8192 * 32
        gbps = {k: pass_bytes / (v * 1e-3) / 1e9 for k, v in rec_stages.items() if v > 0}
        secondary["recovery_pass_hbm"] = {
            "bytes_per_pass": pass_bytes, "gb_per_s": gbps,
            "mean_gb_per_s": sum(gbps.values()) / len(gbps) if gbps else None,
            "frac_of_8_tb_per_s": (sum(gbps.values()) / len(gbps)) / 8000 if gbps else None}
        # uneven recovery: call sizes U[1,128], per-blob present count U[64,127]
        calls = [rng.randint(1, 128) for _ in range(12)]
        maxn = max(calls)
        pairs = [erase(b % ncells, set(rng.sample(range(128), rng.randint(64, 127)))) for b in range(maxn)]
        ucells = b"".join(p_[0] for p_ in pairs)
        umask = b"".join(p_[1] for p_ in pairs)
        d_ucells = hip.DeviceBuffer.from_bytes(ucells)
        d_uout = hip.DeviceBuffer(len(ucells))
        d_umask = hip.DeviceBuffer.from_bytes(umask)
        for n in calls:  # first use of each size captures its graph
            prover.recover_cells_device(d_uout.ptr, d_ucells.ptr, d_umask.ptr, umask[:128 * n], n,
                                        stream.handle)
        t0 = time.perf_counter()
        for n in calls:
            prover.recover_cells_device(d_uout.ptr, d_ucells.ptr, d_umask.ptr, umask[:128 * n], n,
                                        stream.handle)
        secondary["uneven_recovery_blobs_per_s"] = sum(calls) / (time.perf_counter() - t0)
        secondary["uneven_recovery_frac_of_int_peak"] = secondary["uneven_recovery_blobs_per_s"] * W_REC / PEAK_INT
        last = calls[-1]
        if d_uout.to_bytes(last * cell_bytes) != b"".join(
                full[(b % ncells) * cell_bytes:((b % ncells) + 1) * cell_bytes] for b in range(last)):
            raise SystemExit("uneven recovery returned wrong cells")
        secondary["uneven_recovery_calls"] = calls
        # plain form: one blob per call through the host C ABI
        one = blobs[:131072]
        prover.compute_cell_kzg_proofs_batch(one, 1)
        calls = 8
        t0 = time.perf_counter()
        for _ in range(calls):
            prover.compute_cell_kzg_proofs_batch(one, 1)
        secondary["plain_per_call_1_blob_blobs_per_s"] = calls / (time.perf_counter() - t0)
        # latency of one synchronous C-ABI call per batch size (host copies included)
        per_call = {}
        for n in (1, 4, 16, 64):
            chunk = blobs[:131072 * n]
            prover.compute_cell_kzg_proofs_batch(chunk, n)
            t0 = time.perf_counter()
            for _ in range(4):
                prover.compute_cell_kzg_proofs_batch(chunk, n)
            per_call[str(n)] = (time.perf_counter() - t0) / 4 * 1e3
        secondary["per_call_ms"] = per_call
        # compute_cells_and_kzg_proofs and recover_cells_and_kzg_proofs through
        # the host C ABI on preallocated host buffers (host<->device copies
        # included, Python's buffer copies not), best of 5 calls after one
        # warm-up
        import ctypes
        L = vkzg_amd._lib.load()
        nc = min(nb, 64)
        c_blobs = ctypes.create_string_buffer(blobs[:131072 * nc], 131072 * nc)
        c_cells = ctypes.create_string_buffer(262144 * nc)
        c_proofs = ctypes.create_string_buffer(6144 * nc)
        c_rec = ctypes.create_string_buffer(262144 * nc)
        keep = bytearray(128 * nc)
        for b in range(nc):
            for c in rng.sample(range(128), 64):
                keep[128 * b + c] = 1
        c_keep = ctypes.create_string_buffer(bytes(keep), 128 * nc)

        def best_of(call, reps=5):
            assert call() == 0
            times = []
            for _ in range(reps):
                t0 = time.perf_counter()
                assert call() == 0
                times.append(time.perf_counter() - t0)
            return min(times)

        t = best_of(lambda: L.vkzg_amd_compute_cells_and_proofs(prover.handle, c_cells, c_proofs, c_blobs, nc))
        secondary["cells_and_proofs_host_%d_blobs_per_s" % nc] = nc / t
        t = best_of(lambda: L.vkzg_amd_recover_cells_and_proofs(prover.handle, c_rec, c_proofs, c_cells,
                                                                c_keep, nc))
        secondary["recover_cells_and_proofs_host_%d_blobs_per_s" % nc] = nc / t

    if not a.no_secondary:
        secondary["host_abi_sweep"] = host_abi_sweep(prover, blobs, nb, a.seed + 5)

    if not a.no_secondary and not a.no_verify:
        secondary["verify"] = verify_bench(a.seed + 4)

    if not a.no_secondary and a.serve_seconds > 0:
        secondary["serve"] = serve_bench(prover, random.Random(a.seed + 3), a.serve_seconds,
                                         a.serve_rate, a.probe_rate, a.serve_slots)

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
