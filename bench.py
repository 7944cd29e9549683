#!/usr/bin/env python3
# SPDX-License-Identifier: GPL-3.0-or-later
"""Benchmark: batched PM7 SCF single points on one MI355X.

Workload: 256 closed-shell H/C/N/O molecules of uneven size (3-60 atoms, heavy-atom count uniform
in 1-25, seeded random walk, seed 3). One step = one batched solve of all 256 molecules, inputs
uploaded and integrals built inside the step (a step is the whole request as a caller sees it).

Prints one JSON line: solves/s (value), per-step latency p50/p99, the one-molecule-at-a-time FP64
GPU path for comparison, parity of a sample against the CPU FP64 oracle, the roofline of the
dominant kernel (HIP events on that kernel's own stream) and the CPU baseline (the oracle timed on
this host's cores on a bounded sample).
"""

from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "pm7-rs_amd"))
# Batches in flight need their streams on separate hardware queues (HIP's default is 4 per process,
# fewer than two lanes' streams); read when the HIP runtime starts, so set before torch loads it.
os.environ["GPU_MAX_HW_QUEUES"] = "16"

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)


SCF_KERNELS = ["pm7::k_scf<256, 64, true, true>", "pm7::k_scf<512, 120, false, true>",
               "pm7::k_scf<512, 196, true, false>", "pm7::k_scf<512, 1024, false, false>"]
TRAFFIC_FILES = {"fp64": ROOT / "profiles" / "hbm_traffic_organic256_fp64.json",
                 "mixed": ROOT / "profiles" / "hbm_traffic_organic256_mixed.json",
                 "response": ROOT / "profiles" / "hbm_traffic_response_c32h66.json",
                 "large": ROOT / "profiles" / "hbm_traffic_large_water500.json"}
# The closed-shell response path of an s/p molecule, in launch order.
RESPONSE_KERNELS = ["pm7::k_resp_wall", "pm7::k_resp_pack", "pm7::k_resp_gemm", "pm7::k_resp_pairs<false>"]


def measured_traffic(kernel, mode="fp64"):
    """HBM bytes per launch of `kernel` from the FETCH_SIZE / WRITE_SIZE counter runs (scaled by a
    plain copy's readings) kept beside the profiles, or None when that kernel was not profiled."""
    try:
        per = json.loads(TRAFFIC_FILES[mode].read_text())["per_launch"]
    except (OSError, ValueError, KeyError):
        return None
    rec = per.get(kernel)
    return None if rec is None else round(rec["bytes"])


def pct(xs, q):
    s = sorted(xs)
    if not s:
        return None
    k = (len(s) - 1) * q
    lo = int(k)
    hi = min(lo + 1, len(s) - 1)
    return s[lo] + (s[hi] - s[lo]) * (k - lo)


def size_class(n):
    """The megakernel size class of a basis of n functions (<= 64, <= 120, <= 196, larger)."""
    return 0 if n <= 64 else (1 if n <= 120 else (2 if n <= 196 else 3))


def scf_flops(results, classes):
    """Algorithmic FP64 flops per size class: per SCF iteration 2n^3 (commutator FP),
    2 n^2 n_occ (density), 9 n^3 (dense symmetric eigensolve with vectors, Golub & Van Loan's
    symmetric-QR count), doubled for UHF."""
    out = [0.0, 0.0, 0.0, 0.0]
    for r in results:
        n = r.density.shape[0]
        if n == 0:
            continue
        c = size_class(n)
        spins = 2 if r.unrestricted else 1
        per = 11.0 * n ** 3 + 2.0 * n * n * r.n_occ
        out[c] += spins * per * r.iterations
    return out


N_CU = 256  # MI355X compute units; one SCF workgroup owns one molecule on one CU
HBM_PEAK_GBS = 8000.0


def npair_of_atoms(atoms):
    """Packed orbital pairs per atom (1 for H, 10 for an sp atom, 45 for an spd atom)."""
    spd = {"S", "P", "Cl", "Si", "Al", "Br", "I"}
    return [1 if a[0] == "H" else (45 if a[0] in spd else 10) for a in atoms]


def phase_roofline(atoms_list, results):
    """Per-phase rates of the SCF megakernel from each molecule's own device clocks, for the slowest
    molecule (the step's critical path) and summed over the batch, against the chip's peaks and
    against one CU's share of them (a workgroup owns a molecule):
      Fock build: B_fock = 8 sum_pairs npair_a npair_b (w) + 24 nao^2 (P, H read, F written) bytes
                  per iteration (the FP32 phase reads the same FP64 words);
      commutator: 2 nao^3 flop per iteration (E = FP - (FP)^T), timed with the energy dots;
      density:    2 nao^2 n_occ flop per iteration, doubled for UHF;
      eigensolve: time and sweeps only."""
    rows = []
    for atoms, r in zip(atoms_list, results):
        n = r.density.shape[0]
        npair = npair_of_atoms(atoms)
        tot = sum(npair)
        w_words = (tot * tot - sum(x * x for x in npair)) / 2.0
        it = r.iterations
        spins = 2 if r.unrestricted else 1
        rows.append({"bytes_fock": (8.0 * w_words + 24.0 * n * n) * it,
                     "flop_comm": spins * 2.0 * n ** 3 * it,
                     "flop_dens": spins * 2.0 * n * n * r.n_occ * it,
                     "ms": r.device_phase_ms, "n": n})
    def rates(sel):
        fock_s = sum(x["ms"][0] for x in sel) * 1e-3
        comm_s = sum(x["ms"][1] for x in sel) * 1e-3
        dens_s = sum(x["ms"][4] for x in sel) * 1e-3
        gbs = sum(x["bytes_fock"] for x in sel) / fock_s / 1e9 if fock_s else 0.0
        tf_c = sum(x["flop_comm"] for x in sel) / comm_s / 1e12 if comm_s else 0.0
        tf_d = sum(x["flop_dens"] for x in sel) / dens_s / 1e12 if dens_s else 0.0
        cu = len(sel)
        return {"fock_gbs": round(gbs, 2), "fock_frac_of_hbm": round(gbs / HBM_PEAK_GBS, 6),
                "fock_frac_of_cu_share": round(gbs / (HBM_PEAK_GBS / N_CU) / cu, 4),
                "commutator_tflops": round(tf_c, 5), "commutator_frac_of_cu_share": round(tf_c / (FP64_PEAK_TFLOPS / N_CU) / cu, 4),
                "density_tflops": round(tf_d, 5), "density_frac_of_cu_share": round(tf_d / (FP64_PEAK_TFLOPS / N_CU) / cu, 4),
                "eigensolve_ms": round(sum(x["ms"][3] for x in sel), 3)}
    slow = max(rows, key=lambda x: x["ms"][5])
    return {"slowest_molecule": dict(nao=slow["n"], **rates([slow])), "batch": rates(rows),
            "note": "per-CU fractions divide by the number of molecules (one CU each) times one CU's share "
                    "(8 TB/s / 256 of HBM, 78.6 TFLOP/s / 256 of FP64 MFMA)"}


def slowest_info(results):
    """Basis size, iterations and Jacobi cost of the molecule with the longest device SCF time."""
    r = max(results, key=lambda x: x.device_phase_ms[5])
    n = r.density.shape[0]
    n2 = n + (n & 1)
    rounds = r.jacobi_sweeps * max(1, n2 - 1)
    return {"nao": n, "iterations": r.iterations, "jacobi_sweeps": r.jacobi_sweeps,
            "us_per_jacobi_round": round(r.eig_phase_ms[1] * 1e3 / rounds, 3) if rounds else None}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--molecules", type=int, default=256)
    ap.add_argument("--single", type=int, default=32, help="molecules timed one at a time")
    ap.add_argument("--cpu-sample", type=int, default=256, help="molecules timed on one CPU thread")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (engine defaults)")
    ap.add_argument("--serving", type=int, default=2000, help="requests in the serving stream (0: skip)")
    ap.add_argument("--large-waters", type=int, default=500,
                    help="water cluster of the large-system run (0: skip); 500 waters = 3000 AOs")
    ap.add_argument("--large-only", action="store_true", help="only the large-system run")
    ap.add_argument("--large-warmup", type=int, default=1, help="untimed smaller large-path runs first (0: none)")
    ap.add_argument("--large-max-scf", type=int, default=200,
                    help="iteration cap of the large-system runs (a few for counter profiles)")
    ap.add_argument("--response-only", action="store_true", help="only the response-Fock run")
    ap.add_argument("--serving-only", action="store_true", help="only the serving run")
    ap.add_argument("--d-batch", type=int, default=64,
                    help="S/P/Cl molecules (MNDO/d pairs, <= 196 AOs) batched in FP64 and mixed (0: skip)")
    ap.add_argument("--gradient", type=int, default=1, help="time batched nuclear gradients (0: skip)")
    ap.add_argument("--response", type=int, default=1, help="time the CPHF response Fock (0: skip)")
    ap.add_argument("--dandc", type=int, default=1, help="divide-and-conquer scaling series (0: skip)")
    ap.add_argument("--serve-lanes", type=int, default=8)
    ap.add_argument("--serve-batch", type=int, default=64)
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run a plain device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: every box here has a single MI355X; the batched path "
              "shards by molecule with no collective, so N GPUs = N independent copies.", file=sys.stderr)
        return 2

    import numpy as np
    import torch

    if args.copy_probe_mb > 0:
        # A plain copy of known size: its FETCH_SIZE / WRITE_SIZE readings per dispatch give the
        # factor between those counters and bytes moved.
        x = torch.ones(args.copy_probe_mb << 17, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": x.numel() * 8, "dispatches": 3}), flush=True)
        return 0

    # Measured device-copy bandwidth (read + write bytes over the copy time), the HBM peak on this box.
    def copy_bandwidth_gbs(mb=1024, reps=5):
        x = torch.ones(mb << 17, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        y.copy_(x)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            y.copy_(x)
        e1.record()
        torch.cuda.synchronize()
        gbs = 2.0 * x.numel() * 8 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
        del x, y
        torch.cuda.empty_cache()
        return round(gbs, 1)

    def dgemm_tflops(n=8192, reps=4):
        # The matrix-core FP64 rate this box reaches through the vendor BLAS (torch -> rocBLAS/
        # hipBLASLt dgemm), the measured ceiling beside the published 78.6 TFLOP/s.
        a = torch.randn(n, n, dtype=torch.float64, device="cuda")
        b = torch.randn(n, n, dtype=torch.float64, device="cuda")
        c = a @ b
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            c = a @ b
        e1.record()
        torch.cuda.synchronize()
        tf = 2.0 * n ** 3 * reps / (e0.elapsed_time(e1) * 1e-3) / 1e12
        del a, b, c
        torch.cuda.empty_cache()
        return round(tf, 2)

    import pm7amd
    from pm7amd import workloads as W

    def large_system(eng):
        """One large molecule (water cluster, seed 4) in FP64 and mixed precision on the grid-wide
        path; the FP64 dense eigensolve is reported on its own."""
        atoms = W.water_cluster(args.large_waters, 4)
        mol = pm7amd.Molecule.from_symbols_angstrom(atoms)
        out = {"workload": f"({args.large_waters} H2O) cluster, 1.0 g/cm3, O-O >= 2.6 A, seed 4",
               "atoms": len(atoms), "nao": 6 * args.large_waters}
        ref = None
        # warm-up: a 40-water cluster (240 AOs, the same path) in each precision loads every kernel
        # and library code object the timed runs use (lazy loading would otherwise land in the
        # first run's phases)
        warm = pm7amd.Molecule.from_symbols_angstrom(W.water_cluster(40, 4))
        for precision in ("fp64", "mixed") if args.large_warmup else ():
            eng.run(warm, pm7amd.Pm7Options(precision=precision), raise_on_error=False)
        # "mixed" hands over to FP64 at a density change of 1e-5 (the configuration's); the engine's
        # default hand-over (1e-3) is run beside it
        for precision, switch in (("fp64", 1e-3), ("mixed", 1e-5), ("mixed_default_switch", 1e-3)):
            s0 = time.perf_counter()
            r = eng.run(mol, pm7amd.Pm7Options(precision=precision.split("_")[0], mixed_switch=switch,
                                               max_scf=args.large_max_scf), raise_on_error=False)
            wall = time.perf_counter() - s0
            ph = dict(zip(["fock", "energy_commutator", "diis", "eigensolve", "density", "total"],
                          [round(x, 2) for x in r.device_phase_ms]))
            occ = r.mo_coeff[:, : r.n_occ]
            rec = {"status": r.status, "converged": r.converged, "iterations": r.iterations,
                   "fp32_iterations": r.fp32_iterations, "wall_s": round(wall, 3),
                   "solves_per_s": round(1.0 / wall, 4), "phase_ms": ph,
                   "eigensolve_ms_per_iteration": round(r.device_phase_ms[3] / max(1, r.iterations), 2),
                   "heat_of_formation_kcal": r.heat_of_formation_kcal, "total_ev": r.total_ev,
                   "trace_P_minus_electrons": float(np.trace(r.density) - 2 * r.n_occ),
                   "max_abs_2CCt_minus_P": float(np.max(np.abs(2.0 * occ @ occ.T - r.density)))}
            # Fock build F = H + G(P) against the HBM roofline (grid-wide kernels on this path):
            # algorithmic bytes per iteration = 8 sum_pairs npair_a npair_b (w) + 8 nao^2 (P read)
            # + 8 nao^2 (H read) + 8 nao^2 (F write); w and P are read in FP64 in both precisions.
            npair = np.array([1 if a[0] == "H" else (45 if a[0] in ("S", "P", "Cl") else 10) for a in atoms], float)
            w_words = (npair.sum() ** 2 - (npair ** 2).sum()) / 2.0
            nao = 6 * args.large_waters
            b_iter = 8.0 * w_words + 24.0 * nao * nao
            fock_s = r.device_phase_ms[0] * 1e-3
            gbs = b_iter * r.iterations / fock_s / 1e9 if fock_s > 0 else 0.0
            rec["fock_gp"] = {"bytes_per_iteration": b_iter, "ms_per_iteration": round(r.device_phase_ms[0] / max(1, r.iterations), 3),
                              "achieved_gbs": round(gbs, 1), "peak_gbs": 8000.0, "frac": round(gbs / 8000.0, 4)}
            # measured HBM bytes of one FP64 and one FP32 Fock build (rocprofv3 FETCH_SIZE/WRITE_SIZE
            # of the streamed kernels, committed profile of this workload)
            try:
                fb = json.loads(TRAFFIC_FILES["large"].read_text())["fock_build"]
                if args.large_waters == 500:
                    rec["fock_gp"]["traffic_fp64"] = fb["double"]["bytes"]
                    rec["fock_gp"]["traffic_fp32"] = fb["float"]["bytes"]
                    rec["fock_gp"]["traffic_ratio_fp64"] = round(fb["double"]["bytes"] / b_iter, 3)
            except (OSError, KeyError, ValueError):
                pass
            if precision == "fp64":
                ref = r
            else:
                rec["mixed_switch"] = switch
                rec["err_vs_fp64_ev_reported"] = r.err_vs_fp64_ev
                rec["err_vs_fp64_density_reported"] = r.err_vs_fp64_density
                rec["abs_dHf_vs_fp64_kcal"] = abs(r.heat_of_formation_kcal - ref.heat_of_formation_kcal)
                rec["abs_dE_vs_fp64_ev"] = abs(r.total_ev - ref.total_ev)
                rec["max_abs_dP_vs_fp64"] = float(np.max(np.abs(r.density - ref.density)))
            out[precision] = rec
            print(f"[bench] large {precision}: {rec['wall_s']} s, {rec['iterations']} iterations",
                  file=sys.stderr, flush=True)
        return out

    def response_fock(eng):
        """Response Fock (the CPHF kernel of the Hessian): 3N symmetric densities of the C32 alkane
        (98 atoms, 194 AOs, the size of the 102-atom Hessian molecule pm7-rs profiles) against its
        one integral set, one call. Algorithmic HBM bytes per call: the pair integrals and H once,
        each density read and its G written once. kernel_ms spans the response kernels (HIP events
        on their stream); api_ms_per_call adds the integrals, the copies over PCIe and the change of
        layout."""
        ratoms = W.alkane_chain(32)
        rmol = pm7amd.Molecule.from_symbols_angstrom(ratoms)
        rn = 4 * 32 + 66
        rk = 3 * len(ratoms)
        rng = np.random.default_rng(17)
        ra = rng.normal(size=(rk, rn, rn)) * 0.05
        rxs = ra + ra.transpose(0, 2, 1)
        eng.response_fock(rmol, rxs)
        reps = max(2, min(args.steps, 5))
        kms, t = [], time.perf_counter()
        for _ in range(reps):
            eng.response_fock(rmol, rxs)
            kms.append(eng.last_timings_ms()[1])
        dt = (time.perf_counter() - t) / reps
        kernel_ms = float(np.mean(kms))
        n_orb = [4 if a[0] != "H" else 1 for a in ratoms]
        npair = {1: 1, 4: 10}
        w_words = sum(npair[n_orb[u]] * npair[n_orb[v]] for u in range(len(n_orb)) for v in range(u + 1, len(n_orb)))
        alg_bytes = 8.0 * (w_words + rn * rn + 2 * rk * rn * rn)
        traffic = [measured_traffic(k, "response") for k in RESPONSE_KERNELS]
        out = {"workload": f"C32H66 alkane (98 atoms, {rn} AOs), {rk} symmetric response densities (3N), one call",
               "densities_per_s": round(rk / (kernel_ms * 1e-3), 1), "kernel_ms": round(kernel_ms, 4),
               "api_ms_per_call": round(dt * 1e3, 3),
               "kernels": RESPONSE_KERNELS,
               "roofline": {"bound": "hbm", "achieved": round(alg_bytes / (kernel_ms * 1e-3) / 1e9, 1),
                            "peak": 8000.0, "unit": "GB/s",
                            "frac": round(alg_bytes / (kernel_ms * 1e-3) / 8.0e12, 5),
                            "traffic": None if None in traffic else int(sum(traffic)),
                            "algorithmic_bytes": int(alg_bytes),
                            "integral_words": int(w_words)}}
        if not args.no_cpu:
            # the oracle's Fock build minus H, one thread, on the first 24 densities
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_binding as ob

            ob.hcore(rmol.z, rmol.xyz)
            t = time.perf_counter()
            for _ in range(8):
                ob.hcore(rmol.z, rmol.xyz)
            t_int = (time.perf_counter() - t) / 8
            h = ob.hcore(rmol.z, rmol.xyz)
            t = time.perf_counter()
            for x in rxs[:24]:
                ob.fock(rmol.z, rmol.xyz, x, 0.5 * x) - h
            per = (time.perf_counter() - t) / 24
            # each oracle call also rebuilds the pair integrals (one H_core build, reported beside
            # the rate); the reference keeps them between calls
            out["cpu_baseline"] = {"densities_per_s": round(1.0 / per, 1), "cores": 1, "kind": "port",
                                   "ms_per_density": round(per * 1e3, 3),
                                   "integral_build_ms": round(t_int * 1e3, 3),
                                   "sample": "24 of the densities through the CPU oracle's Fock build minus H; "
                                             "each call includes one integral build (integral_build_ms)"}
        print(f"[bench] response fock: {out}", file=sys.stderr, flush=True)
        return out

    # Serving: the continuous-batching server's FP64 capacity is measured first (every request
    # submitted at once), then Poisson arrivals at 75 % of it for FP64 and for mixed precision;
    # latency is submit -> completion per request.
    def serving_run(eng):
        import threading
        stream = W.serving_stream(args.serving, seed=5, rate_per_s=1.0, frac_sp=0.7, frac_d=0.2)
        mols_s = [(pm7amd.Molecule.from_symbols_angstrom(a, c, m), c, m) for _, a, c, m, _ in stream]
        kinds = [k for *_, k in stream]
        gaps = [t for t, *_ in stream]  # arrival times at rate 1/s; scaled below
        reqs = [(0.0, mol, c, m) for mol, c, m in mols_s]

        def serve(precision):
            lat, bad = [], 0
            by_kind, kind_of, got = {}, {}, []
            with pm7amd.Server(eng, lanes=args.serve_lanes, max_batch=args.serve_batch) as srv:
                t0 = time.perf_counter()

                def feed():
                    for k, (t, mol, c, m) in enumerate(reqs):
                        dt = t - (time.perf_counter() - t0)
                        if dt > 0:
                            time.sleep(dt)
                        kind_of[srv.submit(mol, pm7amd.Pm7Options(charge=c, multiplicity=m, precision=precision))] = kinds[k]

                th = threading.Thread(target=feed)
                th.start()
                done = 0
                while done < len(reqs):
                    out = srv.poll(-1, timeout_ms=60000)
                    if out is None:
                        raise SystemExit("serving stalled")
                    done += 1
                    lat.append(out[2])
                    got.append((out[0], out[2]))
                    bad += out[1].status != 0
                elapsed = time.perf_counter() - t0
                th.join()
            for rid, ms in got:
                by_kind.setdefault(kind_of.get(rid, "?"), []).append(ms)
            return {"solves_per_s": round(len(reqs) / elapsed, 3), "failed": bad,
                    "latency_ms": {"p50": round(pct(lat, 0.5), 3), "p99": round(pct(lat, 0.99), 3)},
                    "latency_ms_by_kind": {k: {"n": len(v), "p50": round(pct(v, 0.5), 3), "p99": round(pct(v, 0.99), 3)}
                                           for k, v in sorted(by_kind.items())}}

        serve("fp64")  # warm the server lanes' buffers
        capacity = serve("fp64")  # all requests at t = 0
        rate = 0.75 * capacity["solves_per_s"]
        reqs = [(t / rate, mol, c, m) for t, (mol, c, m) in zip(gaps, mols_s)]
        serving = {"requests": len(reqs), "capacity_fp64_solves_per_s": capacity["solves_per_s"],
                   "rate_per_s": round(rate, 1), "lanes": args.serve_lanes,
                   "max_batch": args.serve_batch,
                   "mix": "70% closed-shell H/C/N/O (3-60 atoms), 20% S/P/Cl molecules of 12-200 atoms "
                          "(MNDO/d pairs), 10% UHF doublets (one H removed)",
                   "fp64": serve("fp64"), "mixed": serve("mixed")}
        print(f"[bench] serving: {serving['fp64']} / {serving['mixed']}", file=sys.stderr, flush=True)
        return serving

    if args.serving_only:
        eng = pm7amd.Engine(0)
        print(json.dumps({"serving": serving_run(eng)}), flush=True)
        eng.close()
        return 0

    if args.response_only:
        eng = pm7amd.Engine(0)
        print(json.dumps({"response_fock": response_fock(eng)}), flush=True)
        eng.close()
        return 0

    if args.large_only:
        eng = pm7amd.Engine(0)
        print(json.dumps({"large_system": large_system(eng)}), flush=True)
        eng.close()
        return 0

    t0 = time.time()
    batch = W.organic_batch(args.molecules, 3)
    mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in batch]
    print(f"[bench] workload built: {len(mols)} molecules in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    eng = pm7amd.Engine(0)
    # The engine's autotuner picks the launch shapes for this workload before anything is timed:
    # per size class block Jacobi or scalar rounds, the reduced-precision sweep cap, and the number
    # of batches in flight (highest throughput whose median step latency stays within 1.5x of one
    # batch at a time). Every measurement below runs with its choice.
    tune = None
    if not args.no_tune:
        tune = eng.autotune(mols, pm7amd.Pm7Options(precision="mixed"), steps=3, latency_budget=1.5)
        tune = {k: (round(v, 3) if isinstance(v, float) else v) for k, v in tune.items()}
        print(f"[bench] autotune: {tune}", file=sys.stderr, flush=True)
    in_flight = tune["batches_in_flight"] if tune else 2

    def run_mode(precision):
        opts = pm7amd.Pm7Options(precision=precision)
        for _ in range(args.warmup):
            eng.run_batch(mols, opts)
        torch.cuda.synchronize()
        lat, kern = [], []
        res = None
        start = time.perf_counter()
        for _ in range(args.steps):
            s = time.perf_counter()
            res = eng.run_batch(mols, opts)
            lat.append((time.perf_counter() - s) * 1e3)
            kern.append(eng.last_timings_ms())
        torch.cuda.synchronize()
        total = time.perf_counter() - start
        return res, lat, kern, total

    def run_pipelined(precision, in_flight=in_flight):
        """The same steps issued back to back through pm7_batch_submit / pm7_batch_wait, `in_flight`
        batches on the device at once (PM7_AMD_ASYNC_LANES): each step is still one batched solve of
        the 256 molecules, its latency is submit -> results on the host; the device span runs from
        the first timed step's integrals kernel to the last one's output gather."""
        opts = pm7amd.Pm7Options(precision=precision)
        steps = args.steps

        def run(n, record):
            inflight, out, res = [], [], None
            for _ in range(n):
                inflight.append((eng.submit_batch(mols, opts), time.perf_counter()))
                if len(inflight) >= in_flight:
                    t, t0 = inflight.pop(0)
                    res = eng.wait_batch(t)
                    tm = eng.last_timings_ms()
                    out.append(((time.perf_counter() - t0) * 1e3, tm[12], tm[13]))
            for t, t0 in inflight:
                res = eng.wait_batch(t)
                tm = eng.last_timings_ms()
                out.append(((time.perf_counter() - t0) * 1e3, tm[12], tm[13]))
            return res, out

        run(max(args.warmup, in_flight), False)
        torch.cuda.synchronize()
        start = time.perf_counter()
        res, rec = run(steps, True)
        torch.cuda.synchronize()
        wall = time.perf_counter() - start
        span_ms = max(r[2] for r in rec) - min(r[1] for r in rec)
        lat = [r[0] for r in rec]
        return res, {"steps": steps, "in_flight": in_flight,
                     "solves_per_s": round(args.molecules * steps / (span_ms * 1e-3), 3),
                     "ms_per_step": round(span_ms / steps, 3),
                     "end_to_end_solves_per_s": round(args.molecules * steps / wall, 3),
                     "latency_ms": {"p50": round(pct(lat, 0.5), 3), "p99": round(pct(lat, 0.99), 3)}}

    # `value` is device-resident throughput: per step, the HIP-event time from the integrals
    # kernel to the end of the output gather (inputs are a few KB of coordinates; results stay in
    # HBM until the host asks). The whole-call rate through the Python mirror, host <-> device
    # copies and result assembly included, is reported beside it as `api`.
    def dev_ms(kern):
        return [k[0] + k[1] + k[2] for k in kern]

    res64, api64, kern64, total64 = run_mode("fp64")
    lat64 = dev_ms(kern64)
    print(f"[bench] fp64 batch: {np.mean(lat64):.1f} ms/step on the device, {np.mean(api64):.1f} ms/call",
          file=sys.stderr, flush=True)
    bad = [r.status for r in res64 if r.status != 0]
    solves64 = args.molecules * args.steps / (sum(lat64) * 1e-3)
    api_solves64 = args.molecules * args.steps / total64
    resmx, apimx, kernmx, totalmx = run_mode("mixed")
    latmx = dev_ms(kernmx)
    print(f"[bench] mixed batch: {np.mean(latmx):.1f} ms/step on the device, {np.mean(apimx):.1f} ms/call",
          file=sys.stderr, flush=True)
    badmx = [r.status for r in resmx if r.status != 0]
    solvesmx = args.molecules * args.steps / (sum(latmx) * 1e-3)
    api_solvesmx = args.molecules * args.steps / totalmx
    dhf_mx = max(abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) for a, b in zip(resmx, res64))
    de_mx = max(abs(a.total_ev - b.total_ev) for a, b in zip(resmx, res64))
    dp_mx = max(float(np.max(np.abs(a.density - b.density))) for a, b in zip(resmx, res64))
    # The FP64 fixed point: FP64 converged to p_tol 1e-10; a default FP64 run stops up to ~1e-6 from
    # it, and the mixed path, which reaches it along another route, is compared with the point itself.
    res_fix = eng.run_batch(mols, pm7amd.Pm7Options(e_tol=1e-12, p_tol=1e-10, max_scf=500))
    dp_fix_mx = max(float(np.max(np.abs(a.density - b.density))) for a, b in zip(resmx, res_fix))
    dp_fix_64 = max(float(np.max(np.abs(a.density - b.density))) for a, b in zip(res64, res_fix))
    err_rep = max(r.err_vs_fp64_density for r in resmx)
    mixed_ok = dhf_mx <= 1e-4 and not badmx

    # The same steps pipelined: two batches in flight through the asynchronous entry points. Their
    # results must be the synchronous ones bit for bit.
    def identical(a, b):
        return all(x.total_ev == y.total_ev and np.array_equal(x.density, y.density) for x, y in zip(a, b))

    pres64, pipe64 = run_pipelined("fp64")
    presmx, pipemx = run_pipelined("mixed")
    pipe64["bitwise_equal_to_one_at_a_time"] = identical(pres64, res64)
    pipemx["bitwise_equal_to_one_at_a_time"] = identical(presmx, resmx)
    # CU-busy fraction with steps in flight: the molecules' own device time per step (each owns one
    # CU while its SCF runs) over the CU-time of a step of the pipelined span.
    for pr, pp in ((pres64, pipe64), (presmx, pipemx)):
        pp["cu_busy_frac"] = round(sum(r.device_phase_ms[5] for r in pr) / (N_CU * pp["ms_per_step"]), 4)
    print(f"[bench] pipelined: fp64 {pipe64}, mixed {pipemx}", file=sys.stderr, flush=True)

    serving = serving_run(eng) if args.serving > 0 else None

    large = large_system(eng) if args.large_waters > 0 else None
    copy_gbs = copy_bandwidth_gbs()
    dgemm_tf = dgemm_tflops()

    # A batch of molecules holding S, P or Cl: every pair through the MNDO/d pair kernel, the
    # 45-word spd one-centre blocks in the Fock build; the megakernel classes only (<= 196 AOs).
    d_batch = None
    if args.d_batch > 0:
        rng_d = W.Rng(11)
        d_atoms = [W.hetero_chain(rng_d, rng_d.randint(4, 30), ("S", "P", "Cl")[rng_d.randint(0, 2)])
                   for _ in range(args.d_batch)]
        d_mols = [pm7amd.Molecule.from_symbols_angstrom(a) for a in d_atoms]
        d_out = {"molecules": len(d_mols), "atoms": [min(map(len, d_atoms)), max(map(len, d_atoms))],
                 "workload": "hetero_chain seed 11: alkanes of 4-30 C with one S (thioether), PH or Cl"}
        d_res = {}
        for precision in ("fp64", "mixed"):
            o = pm7amd.Pm7Options(precision=precision)
            eng.run_batch(d_mols, o)
            t = time.perf_counter()
            reps = max(1, args.steps)
            dev = 0.0
            for _ in range(reps):
                d_res[precision] = eng.run_batch(d_mols, o)
                tm = eng.last_timings_ms()
                dev += (tm[0] + tm[1] + tm[2]) * 1e-3
            dt = (time.perf_counter() - t) / reps
            dev /= reps
            d_out[precision] = {"solves_per_s": round(len(d_mols) / dev, 3), "ms_per_batch": round(dev * 1e3, 3),
                                "api_ms_per_call": round(dt * 1e3, 3),
                                "failed": sum(r.status != 0 for r in d_res[precision])}
        d_out["mixed"]["max_abs_dHf_vs_fp64_kcal"] = max(
            abs(a.heat_of_formation_kcal - b.heat_of_formation_kcal) for a, b in zip(d_res["mixed"], d_res["fp64"]))
        d_batch = d_out
        print(f"[bench] d batch: {d_out['fp64']} / {d_out['mixed']}", file=sys.stderr, flush=True)

    # Nuclear gradients of the same 256-molecule batch: one SCF per molecule, then the pair and atom
    # gradient kernels on the density left in HBM and the corrections' gradient on the host. Timed
    # as whole calls through the Python mirror, beside the plain batched SCF call.
    gradient = None
    if args.gradient:
        gradient = {"workload": "the 256-molecule organic batch", "timing": "whole call through the Python mirror"}
        for precision, api_scf in (("fp64", api64), ("mixed", apimx)):
            o = pm7amd.Pm7Options(precision=precision)
            eng.gradient_batch(mols, o)
            reps = max(1, min(args.steps, 3))
            t = time.perf_counter()
            for _ in range(reps):
                g_res, g_vals = eng.gradient_batch(mols, o)
            dt = (time.perf_counter() - t) / reps
            gradient[precision] = {
                "gradients_per_s": round(len(mols) / dt, 3), "ms_per_call": round(dt * 1e3, 3),
                "scf_only_ms_per_call": round(float(np.mean(api_scf)), 3),
                "failed": sum(r.status != 0 for r in g_res),
                "max_abs_gradient_ev_bohr": round(max(float(np.max(np.abs(g))) for g in g_vals), 6)}
        print(f"[bench] gradient: {gradient}", file=sys.stderr, flush=True)

    response = response_fock(eng) if args.response else None

    # Divide and conquer: the reference's scaling record (docs/divide_and_conquer.md:85-92: flat
    # alkanes, buffer 15 Bohr, core_size 8, D&C p_tol 1e-7; exact SCF e_tol 1e-10, p_tol 1e-8;
    # 16 cores, single wall-clock runs) on the GPU, the CPU oracle's D&C beside it on the small sizes.
    dandc = None
    if args.dandc:
        from pm7amd.dandc import DandcOptions, run_dandc

        ref_s = {40: 2.08, 80: 5.07, 160: 9.53, 320: 19.32, 640: 39.51}
        ref_exact_s = {40: 0.54, 80: 2.48, 160: 15.42}
        ref_de = {40: 0.0019, 80: 0.0043, 160: 0.0089}
        dopt = DandcOptions(buffer=15.0, core_size=8, p_tol=1e-7, max_scf=400)
        xopt = pm7amd.Pm7Options(e_tol=1e-10, p_tol=1e-8, max_scf=400)
        gpu_e = {}
        rows = []
        for nc in (40, 80, 160, 320, 640):
            atoms = W.flat_alkane(nc)
            dmol = pm7amd.Molecule.from_symbols_angstrom(atoms)
            if nc == 40:
                run_dandc(dmol, dandc=dopt, engine=eng)
            t = time.perf_counter()
            r = run_dandc(dmol, dandc=dopt, engine=eng)
            wall = time.perf_counter() - t
            gpu_e[nc] = r.electronic_ev
            row = {"carbons": nc, "atoms": len(atoms), "wall_s": round(wall, 4), "device_ms": round(r.device_ms, 2),
                   "iterations": r.iterations, "ms_per_iteration": round(r.device_ms / r.iterations, 3),
                   "subsystems": r.subsystems, "largest_subsystem": r.largest_subsystem,
                   "largest_subsystem_nao": r.largest_subsystem_nao, "converged": r.converged,
                   "reference_s_16_cores": ref_s[nc], "speedup_vs_reference": round(ref_s[nc] / wall, 1)}
            if nc <= 160:
                t = time.perf_counter()
                ex = eng.run(dmol, xopt)
                row["exact_wall_s"] = round(time.perf_counter() - t, 4)
                row["reference_exact_s_16_cores"] = ref_exact_s[nc]
                row["dE_vs_exact_ev"] = round(abs(r.electronic_ev - ex.electronic_ev), 6)
                row["reference_dE_vs_exact_ev"] = ref_de[nc]
                row["stored_fraction"] = round(r.density.stored_elements() / ex.density.shape[0] ** 2, 4)
            rows.append(row)
        xs = np.log([r_["atoms"] for r_ in rows])
        dandc = {"workload": "flat all-anti alkanes C40-C640 (122-1922 atoms), buffer 15 Bohr, core_size 8, "
                             "p_tol 1e-7 (the reference's scaling record)",
                 "series": rows,
                 "log_log_slope_wall": round(float(np.polyfit(xs, np.log([r_["wall_s"] for r_ in rows]), 1)[0]), 3),
                 "log_log_slope_per_iteration": round(float(np.polyfit(
                     xs, np.log([r_["ms_per_iteration"] for r_ in rows]), 1)[0]), 3)}
        if not args.no_cpu:
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_binding as ob

            cpu_rows = []
            for nc in (40, 80):
                z1, x1 = ob.atoms_to_arrays(W.flat_alkane(nc))
                t = time.perf_counter()
                rr, *_ = ob.run_dandc(z1, x1, buffer=15.0, core_size=8, p_tol=1e-7, max_scf=400,
                                      threads=min(16, os.cpu_count() or 1))
                cpu_rows.append({"carbons": nc, "wall_s": round(time.perf_counter() - t, 3),
                                 "iterations": rr.iterations,
                                 "abs_dE_vs_gpu_ev": abs(rr.electronic_ev - gpu_e[nc])})
            dandc["cpu_baseline"] = {"kind": "port", "cores": min(16, os.cpu_count() or 1), "series": cpu_rows,
                                     "sample": "the CPU oracle's divide and conquer on C40 and C80, subsystems in "
                                               "parallel (one thread each)"}
        print(f"[bench] dandc: {dandc}", file=sys.stderr, flush=True)

    # One molecule at a time (the plain FP64 GPU path the batched path is compared with).
    single = mols[: args.single]
    for m in single[:2]:
        eng.run(m)
    s = time.perf_counter()
    single_dev = 0.0
    for m in single:
        eng.run(m)
        t = eng.last_timings_ms()
        single_dev += (t[0] + t[1] + t[2]) * 1e-3
    single_api_rate = len(single) / (time.perf_counter() - s) if single else 0.0
    single_rate = len(single) / single_dev if single_dev > 0 else 0.0

    # Roofline of the dominant kernel: the SCF megakernel size class with the largest time, its
    # algorithmic flops (11 n^3 + 2 n^2 n_occ per molecule-iteration: the Roothaan step of the
    # reference, whatever route the kernel takes) over its launch time (HIP events on its stream).
    def roofline_of(res, kern, mode):
        cls_ms = [float(np.mean([k[4 + c] for k in kern])) for c in range(4)]
        dom = int(np.argmax(cls_ms))
        flops = scf_flops(res, None)
        achieved_tf = flops[dom] / (cls_ms[dom] * 1e-3) / 1e12 if cls_ms[dom] > 0 else 0.0
        kname = SCF_KERNELS[dom]
        cnt = sum(1 for r in res if size_class(r.density.shape[0]) == dom)
        busy = sum(r.device_phase_ms[5] for r in res) / (N_CU * max(cls_ms))
        return {"bound": "mfma", "kernel": kname, "mode": mode,
                "achieved": round(achieved_tf, 6), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved_tf / FP64_PEAK_TFLOPS, 8), "traffic": measured_traffic(kname, mode),
                "avg_launch_ms": round(cls_ms[dom], 4), "class_ms": [round(x, 4) for x in cls_ms],
                "molecules_in_launch": cnt,
                "frac_of_cu_share": round(achieved_tf / (FP64_PEAK_TFLOPS / N_CU) / max(1, cnt), 6),
                "cu_busy_frac": round(busy, 4),
                "measured_dgemm_peak": dgemm_tf, "frac_of_measured_dgemm": round(achieved_tf / dgemm_tf, 8)}

    roofline_mx = roofline_of(resmx, kernmx, "mixed")
    roofline_64 = roofline_of(res64, kern64, "fp64")
    roofline = roofline_mx if mixed_ok else roofline_64

    # Configurations 1 and 2 one molecule at a time (FP64): water and the jittered all-trans C32H66
    # (98 atoms, 194 AOs), device time and whole-call time, beside the CPU oracle on one thread and
    # on the GPU's 16-CPU share of the host.
    configs12 = {}
    for name, atoms in (("water", W.water()), ("C32H66", W.alkane_chain(32, 2))):
        m1 = pm7amd.Molecule.from_symbols_angstrom(atoms)
        eng.run(m1)
        reps = 20 if name == "water" else 5
        dev, t = 0.0, time.perf_counter()
        for _ in range(reps):
            r1 = eng.run(m1)
            tt = eng.last_timings_ms()
            dev += (tt[0] + tt[1] + tt[2]) * 1e-3
        api = (time.perf_counter() - t) / reps
        rec = {"nao": r1.density.shape[0], "iterations": r1.iterations, "device_ms": round(dev / reps * 1e3, 4),
               "api_ms": round(api * 1e3, 4), "solves_per_s_device": round(reps / dev, 3),
               "total_ev": r1.total_ev}
        if not args.no_cpu:
            sys.path.insert(0, str(ROOT / "tests"))
            import oracle_binding as ob
            z1, x1 = ob.atoms_to_arrays(atoms)
            for label, thr in (("cpu_1_thread", 1), ("cpu_16_threads", 16)):
                ob.run(z1, x1, threads=thr)
                t = time.perf_counter()
                nrep = 0
                while nrep == 0 or time.perf_counter() - t < 1.0:
                    ref1, *_ = ob.run(z1, x1, threads=thr)
                    nrep += 1
                rec[label] = {"ms": round((time.perf_counter() - t) / nrep * 1e3, 3), "cores": thr}
            rec["abs_dE_vs_oracle_ev"] = abs(ref1.total_ev - r1.total_ev)
        configs12[name] = rec
    print(f"[bench] configs 1-2: {configs12}", file=sys.stderr, flush=True)

    cpu = None
    parity = None
    if not args.no_cpu:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_binding as ob

        from concurrent.futures import ThreadPoolExecutor

        # The CPU oracle over the whole batch, molecules in parallel (one host thread each, as
        # rayon callers of run_pm7 do), then one thread alone on a sample of it.
        # 16 = this GPU's share of the host (the pool gives one GPU 16 CPUs, whatever os.cpu_count()
        # reports); os.cpu_count() is printed beside it
        threads = min(16, os.cpu_count() or 1)
        arrays = [ob.atoms_to_arrays(a) for a in batch]

        def solve(i):
            ref, dens, *_ = ob.run(arrays[i][0], arrays[i][1], threads=1)
            return (abs(ref.total_ev - res64[i].total_ev), abs(ref.hof_kcal - res64[i].heat_of_formation_kcal),
                    float(np.max(np.abs(dens - res64[i].density))))

        # Whole passes over the batch until at least 2 s have passed (tens of core-seconds).
        passes = 0
        s = time.perf_counter()
        with ThreadPoolExecutor(threads) as pool:
            while passes == 0 or time.perf_counter() - s < 2.0:
                diffs = list(pool.map(solve, range(len(batch))))
                passes += 1
        cpu_t = (time.perf_counter() - s) / passes
        n1 = min(args.cpu_sample, len(batch))
        s = time.perf_counter()
        for i in range(n1):
            ob.run(arrays[i][0], arrays[i][1], threads=1)
        t1 = time.perf_counter() - s
        cpu = {"value": round(len(batch) / cpu_t, 3), "unit": "solves/s", "cores": threads, "kind": "port",
               "host_cpus_reported": os.cpu_count(),
               "sample": f"the 256-molecule batch on the CPU FP64 oracle, {threads} molecules at a time "
                         f"(one thread each), {passes} passes, {cpu_t:.2f}s per pass; one thread alone on the first {n1}: "
                         f"{n1 / t1:.3f} solves/s ({t1:.1f}s)",
               "value_1_thread": round(n1 / t1, 3)}
        parity = {"sample": len(batch), "max_abs_dE_ev": max(d[0] for d in diffs),
                  "max_abs_dHf_kcal": max(d[1] for d in diffs), "max_abs_dP": max(d[2] for d in diffs)}

    head_mixed = mixed_ok and pipemx["bitwise_equal_to_one_at_a_time"]
    head = pipemx if head_mixed else pipe64
    line = {
        "metric": "batched SCF solves/s + p99 latency, mixed vs FP64 GPU at |ΔHf|≤1e-4 kcal/mol",
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": head["steps"],
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head_mixed else "f64",
        "data": "synthetic",
        "config": {"workload": "256 closed-shell H/C/N/O molecules, 3-60 atoms, seeded random walk (seed 3), "
                               f"one batched solve per step, steps pipelined {head['in_flight']} in flight",
                   "n_molecules": args.molecules,
                   "precision": "mixed" if head_mixed else "fp64", "steps_in_flight": head["in_flight"]},
        "autotune": tune,
        "latency_ms": head["latency_ms"],
        "value_note": "device span from the first timed step's integrals kernel to the last step's output gather "
                      "(HIP events), steps pipelined through pm7_batch_submit/wait; latency = submit -> results on "
                      "the host per step; end_to_end_solves_per_s (host clock, PCIe and result assembly included) "
                      "beside it",
        "pipelined": {"fp64": pipe64, "mixed": pipemx},
        "api": {"note": "whole run_batch call through the Python mirror: uploads, result downloads and "
                        "host assembly included",
                "fp64": {"solves_per_s": round(api_solves64, 3),
                         "latency_ms": {"p50": round(pct(api64, 0.5), 3), "p99": round(pct(api64, 0.99), 3)}},
                "mixed": {"solves_per_s": round(api_solvesmx, 3),
                          "latency_ms": {"p50": round(pct(apimx, 0.5), 3), "p99": round(pct(apimx, 0.99), 3)}},
                "single_molecule_fp64_solves_per_s": round(single_api_rate, 3)},
        "fp64": {"note": "one batch at a time (pm7_run_batch), device time per step",
                 "solves_per_s": round(solves64, 3), "failed": len(bad),
                 "latency_ms": {"p50": round(pct(lat64, 0.5), 3), "p99": round(pct(lat64, 0.99), 3)},
                 "scf_kernel_ms": round(float(np.mean([k[1] for k in kern64])), 4)},
        "mixed": {"note": "one batch at a time (pm7_run_batch), device time per step",
                  "solves_per_s": round(solvesmx, 3), "failed": len(badmx),
                  "latency_ms": {"p50": round(pct(latmx, 0.5), 3), "p99": round(pct(latmx, 0.99), 3)},
                  "scf_kernel_ms": round(float(np.mean([k[1] for k in kernmx])), 4),
                  "speedup_vs_fp64": round(solvesmx / solves64, 4),
                  "max_abs_dHf_vs_fp64_kcal": dhf_mx, "max_abs_dE_vs_fp64_ev": de_mx, "max_abs_dP_vs_fp64": dp_mx,
                  "max_abs_dP_vs_fp64_fixed_point": dp_fix_mx, "fp64_default_max_abs_dP_vs_fixed_point": dp_fix_64,
                  "max_err_vs_fp64_density_reported": err_rep,
                  "phase_roofline": phase_roofline(batch, resmx),
                  "within_1e-4_kcal": mixed_ok,
                  "fp32_iterations_mean": round(float(np.mean([r.fp32_iterations for r in resmx])), 2),
                  "slowest_molecule": slowest_info(resmx),
                  "phase_ms_slowest": dict(zip(["fock", "energy_commutator", "diis", "eigensolve", "density", "total"],
                                               [round(x, 3) for x in max(resmx, key=lambda r: r.device_phase_ms[5]).device_phase_ms])),
                  "iterations_mean": round(float(np.mean([r.iterations for r in resmx])), 2)},
        "serving": serving,
        "large_system": large,
        "hbm_copy_probe_gbs": copy_gbs,
        "dgemm_probe_tflops": dgemm_tf,
        "d_batch": d_batch,
        "gradient": gradient,
        "response_fock": response,
        "dandc": dandc,
        "single_molecule_fp64": {"solves_per_s": round(single_rate, 3), "molecules": len(single),
                                 "batched_speedup": round(head["solves_per_s"] / single_rate, 3) if single_rate else None},
        "kernel_ms": {"integrals": round(float(np.mean([k[0] for k in kern64])), 4),
                      "scf": round(float(np.mean([k[1] for k in kern64])), 4),
                      "gather": round(float(np.mean([k[2] for k in kern64])), 4),
                      "call_wall": round(float(np.mean([k[3] for k in kern64])), 4),
                      "host_prepare": round(float(np.mean([k[8] for k in kern64])), 4),
                      "host_launch": round(float(np.mean([k[9] for k in kern64])), 4),
                      "device_wait": round(float(np.mean([k[10] for k in kern64])), 4),
                      "host_results": round(float(np.mean([k[11] for k in kern64])), 4)},
        "iterations": {"mean": round(float(np.mean([r.iterations for r in res64])), 2),
                       "max": int(max(r.iterations for r in res64))},
        "phase_ms_slowest": dict(zip(["fock", "energy_commutator", "diis", "eigensolve", "density", "total"],
                                     [round(x, 3) for x in max(res64, key=lambda r: r.device_phase_ms[5]).device_phase_ms])),
        "phase_ms_sum": dict(zip(["fock", "energy_commutator", "diis", "eigensolve", "density", "total"],
                                 [round(sum(r.device_phase_ms[k] for r in res64), 2) for k in range(6)])),
        "eig_split_ms_slowest": dict(zip(["warm_start_in", "sweeps", "sort_out"],
                                         [round(x, 3) for x in max(res64, key=lambda r: r.device_phase_ms[5]).eig_phase_ms])),
        "slowest_molecule": slowest_info(res64),
        "jacobi_sweeps_per_solve": round(float(np.mean([r.jacobi_sweeps / max(1, r.iterations) for r in res64])), 2),
        "parity_vs_oracle": parity,
        "phase_roofline_fp64": phase_roofline(batch, res64),
        "roofline_fp64": roofline_64,
        "configs_1_2": configs12,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    eng.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
