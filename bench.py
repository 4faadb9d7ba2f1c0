"""Benchmark: close-evaluation target-panel pairs/sec on one MI355X.

Workload (the order-8 stellarator case): the 5-period surface
R = 1 + 0.3 cos th + 0.06 cos(th - 5 ph), Z = 0.3 sin th - 0.06 sin(th - 5 ph),
48 x 120 charts -> 14,250 order-8 triangles (513,000 nodes); 200,000 targets at
signed distance h 10^U, U ~ U[-6,-1], on both sides; near list by the
reference's ball rule; density from 8 point charges 0.4 outside the surface.

A step is one pass of the close-evaluation path over the resident batch:
pair labelling, direct S/D on the refinement levels, RRQ for the rest, the
naive-block subtraction and the fixed-order scatter into the potential (the
reference's "rrq+fill" region).  Geometry, targets, near list and density are
in HBM before timing starts.  Three paths are timed on the same workload:

  plain     FP64 throughout, every pair labelled by the reference's ball rule;
  selector  the learned level selector (qotential_amd/selector_model.npz) with
            the ball rule below confidence 0.9, FP64;
  split     the ball rule, with the direct weights of levels 1..--split-levels
            from the FP32 matrix cores (x - y and (x - y).n in FP64).

`value` is the path named by --path: plain by default, the path that meets the
metric's condition (GRF error within 1e-9 of plain FP64 -- it is plain FP64).
Every path's Green's identity error (S[du/dn] - D[u] = u inside, 0 outside, on
every target) and its potential's difference from the plain path are reported.
Beside them (not part of `value`) the plain path runs on the configuration's
"radius 2h" near list: every target within twice the panel's largest corner
distance of the panel's node mean.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
# the reference's own single-thread rate on its order-4 fixture run: 137,259
# close pairs in 0.574 s of "rrq+fill" (web/stellarator/fixtures/native/order4-direct.json:20)
REFERENCE_ORDER4_PAIRS_PER_S = 137259 / 0.574


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--mp", type=int, default=48)
    ap.add_argument("--np", type=int, default=120)
    ap.add_argument("--order", type=int, default=8)
    ap.add_argument("--targets", type=int, default=200000)
    ap.add_argument("--path", choices=("plain", "selector", "split"), default="plain")
    ap.add_argument("--no-compare", action="store_true", help="time the --path path only")
    ap.add_argument("--split-levels", type=int, default=2,
                    help="split path: levels 1..N get their direct weights from the FP32 matrix cores")
    ap.add_argument("--threshold", type=float, default=0.9, help="selector path: confidence threshold")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-pairs", type=int, default=300000)
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the timed steps, one device-to-device copy of this many MiB (a plain copy "
                         "kernel of known size that scales the FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def level_flops(p, counts):
    """algorithmic FLOPs of the direct level kernels: per pair served at level l,
    nl(l) (19 + 4h) (entries with one rsqrt, two contractions of h columns)"""
    from qotential_amd.engine import order_info
    info = order_info(p)
    h = info["h"]
    return [float(int(counts[l]) * int(info["nl"][l]) * (19 + 4 * h)) for l in range(4)]


def kernel_flops(p, npan, counts, nrrq, nrrq_panels, split_levels):
    """Algorithmic FLOPs per launch of the kernels that have a fixed count (from
    the formulas, not from executed instructions).  counts[l] = pairs served by
    level l+1, nrrq = RRQ pairs, nrrq_panels = panels with an RRQ pair."""
    from qotential_amd.engine import order_info
    info = order_info(p)
    h, nq, ncoef = info["h"], info["nq"], info["ncoef"]
    nb = 3 * nq
    nc1, nc2 = p * (p + 1) // 2, (p - 1) * p // 2
    ntot = int(sum(info["nl"]))
    lf = level_flops(p, counts)
    ns = max(0, min(4, split_levels))
    out = {
        # 9 sequential FMAs per (level node, panel node); matrix cores up to order 12
        "levels": ("level_geom_mfma_kernel", 18.0 * h * ntot * npan, PEAK_FP64_TFLOPS),
        # real Gauss-Jordan 2h^3; quaternion Gauss-Jordan 32h^3, C = B Fn^T 2h^3, As rows 8h^3
        "rrq_solve": ("rrq_real_inverse_kernel+rrq_quat_inverse_kernel", 44.0 * h ** 3 * nrrq_panels,
                      PEAK_FP64_TFLOPS),
        # raw moments 4 nb (3 nc1 + 9 nc2) (twelve real node weights against the boundary
        # harmonics of degree < p, resp. < p - 1), their combination into the boundary
        # moments ~200 ncoef, degree recursion ~200 ncoef
        "rrq_moments": ("rrq_target_kernel", float(nrrq) * (4 * nb * (3 * nc1 + 9 * nc2) + 400 * ncoef),
                        PEAK_FP64_TFLOPS),
        # W = G V: 2 (2h)(5h) per RRQ pair
        "rrq_weights": ("rrq_weights_kernel", float(nrrq) * 20 * h * h, PEAK_FP64_TFLOPS),
    }
    # the direct stage: levels 1..ns on FP32 matrix cores, the rest on FP64
    if ns > 0:
        out["direct_split"] = ("direct_split_kernel", sum(lf[:ns]), PEAK_FP32_TFLOPS)
    if ns < 4:
        out["direct"] = ("direct_mfma_kernel", sum(lf[ns:]), PEAK_FP64_TFLOPS)
    return out


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_p8_48x120.json")


def measured_traffic(kernel, a):
    """HBM bytes per launch of `kernel` ('+'-joined names are summed) from the
    FETCH_SIZE / WRITE_SIZE counter runs of this workload and path (scaled by a
    plain copy of known size), or None when no such record exists."""
    try:
        with open(TRAFFIC_FILE) as f:
            rec = json.load(f)
    except OSError:
        return None
    if rec.get("workload") != [a.mp, a.np, a.order, a.targets]:
        return None
    total = 0.0
    for name in kernel.split("+"):
        k = rec["kernels"].get(name)
        if k is None:
            return None
        total += k["bytes"]
    return total


def cpu_info():
    info = {"nproc": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Core(s) per socket", "Thread(s) per core"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _panel_sample(offs, idx, sample_pairs, seed):
    npan = len(offs) - 1
    order = np.random.default_rng(seed).permutation(npan)
    chosen, tot = [], 0
    for k in order:
        c = int(offs[k + 1] - offs[k])
        if c == 0:
            continue
        chosen.append(k)
        tot += c
        if tot >= sample_pairs:
            break
    chosen = np.sort(np.array(chosen))
    cnt = np.zeros(npan, dtype=np.int64)
    cnt[chosen] = np.diff(offs)[chosen]
    sub_offs = np.concatenate([[0], np.cumsum(cnt)])
    sub_idx = np.concatenate([idx[offs[k]:offs[k + 1]] for k in chosen])
    return sub_offs, sub_idx, tot, len(chosen)


def cpu_baseline(geo, tx, offs, idx, ub, ubn, sample_pairs):
    """The oracle (CPU FP64 restatement, OpenMP dynamic,4 over panels as the
    reference's GRF loop) on a bounded random sample of panels of the same
    workload: one thread and every host thread the process may use."""
    from oracle import oracle as O
    O.build()
    O.set_side_tol(1e-10)
    info = cpu_info()
    nthr = max(1, min(info["nproc"], 64))
    res = {}
    for threads, pairs in ((1, sample_pairs // 2), (nthr, sample_pairs * min(nthr, 16) // 4)):
        so, si, tot, npn = _panel_sample(offs, idx, pairs, 7 + threads)
        t0 = time.perf_counter()
        O.close_apply(geo["p"], geo, tx, so, si, ub, ubn, exterior=True, nthreads=threads)
        dt = time.perf_counter() - t0
        res[threads] = (tot / dt, tot, npn, dt)
    v1, vn = res[1], res[nthr]
    return dict(value=float(vn[0]), unit="pairs/s", cores=nthr, kind="port",
                value_1thread=float(v1[0]), cpu=info,
                sample=f"{vn[1]} pairs on {vn[2]} random panels of the same workload, oracle FP64 with {nthr} "
                       f"OpenMP threads (dynamic,4 over panels) in {vn[3]:.1f} s; 1 thread: {v1[1]} pairs on "
                       f"{v1[2]} panels in {v1[3]:.1f} s")


def cpu_reference_configs():
    """The oracle at one thread on (a) the CPU reference configuration (16 x 40
    charts, order 4, 4,000 interior targets at h 10^U, U ~ U[-6,-1], point-charge
    density, GRF check) and (b) the reference's order-4 fixture run (built-in
    surface 12 x 36, every node a target, 137,259 pairs), beside the
    reference's own single-thread rate on (b)."""
    from oracle import oracle as O
    from qotential_amd import workload as W
    O.build()
    out = {}
    geo = W.geometry(16, 40, 4, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, 4000, -6.0, -1.0, sides=-1, seed=1)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    offs, idx = O.nearlist(geo["h"], geo["sx"], geo["sw"], tx, nthreads=8)
    u = O.far_direct(tx, geo["sx"], geo["snx"], geo["sw"], ub, ubn, nthreads=8)
    O.set_side_tol(1e-10)
    t0 = time.perf_counter()
    O.close_apply(4, geo, tx, offs, idx, ub, ubn, exterior=True, u=u, nthreads=1)
    dt = time.perf_counter() - t0
    uex, _ = W.charge_field(tx, pos, q)
    out["config1_16x40_p4_4000_interior"] = dict(
        pairs=int(offs[-1]), seconds=dt, pairs_per_s=offs[-1] / dt, threads=1,
        grf_rel_err=float(np.abs(u - uex).max() / np.abs(uex).max()))
    geo4 = W.geometry(12, 36, 4)
    ub, gu = W.harmonic_exp(geo4["sx"])
    ubn = (geo4["snx"] * gu).sum(1)
    offs, idx = O.nearlist(10, geo4["sx"], geo4["sw"], geo4["sx"], nthreads=8)
    O.set_side_tol(1e300)
    t0 = time.perf_counter()
    O.close_apply(4, geo4, geo4["sx"], offs, idx, ub, ubn, exterior=True, nthreads=1)
    dt = time.perf_counter() - t0
    O.set_side_tol(1e-10)
    out["order4_fixture_12x36"] = dict(pairs=int(offs[-1]), seconds=dt, pairs_per_s=offs[-1] / dt, threads=1,
                                       reference_pairs_per_s=REFERENCE_ORDER4_PAIRS_PER_S)
    return out


def timed_steps(eng, steps, warmup):
    for _ in range(warmup):
        eng.eval(exterior=True, side_tol=1e-10)
        eng.apply_resident()
    eng.sync()
    stage_tot = None
    t0 = time.perf_counter()
    for _ in range(steps):
        eng.eval(exterior=True, side_tol=1e-10)
        eng.apply_resident()
        s = eng.stage_ms()
        stage_tot = s if stage_tot is None else {k: stage_tot[k] + s[k] for k in s}
    eng.sync()
    elapsed = time.perf_counter() - t0
    return elapsed, {k: v / steps for k, v in stage_tot.items()}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {a.gpus})",
              file=sys.stderr)
        sys.exit(2)
    from qotential_amd import workload as W
    from qotential_amd.engine import Engine
    from qotential_amd.selector import MODEL_FILE, Selector

    t_setup = time.perf_counter()
    geo = W.geometry(a.mp, a.np, a.order, modes=W.VMEC_BENCH)
    tx, tri, dist = W.targets(geo, a.targets, -6.0, -1.0, sides=0, seed=1)
    pos, q = W.point_charges(geo, nchg=8, offset=0.4, seed=2)
    ub, gu = W.charge_field(geo["sx"], pos, q)
    ubn = (geo["snx"] * gu).sum(1)
    uex, _ = W.charge_field(tx, pos, q)
    uref = np.where(dist < 0.0, uex, 0.0)
    scale = np.abs(uex).max()
    print(f"[bench] geometry {geo['ntri']} triangles p={a.order}, {a.targets} targets "
          f"({time.perf_counter() - t_setup:.1f}s)", flush=True)

    sel = Selector.load(MODEL_FILE)
    runs = {}
    paths = [a.path] if a.no_compare else ["plain", "selector", "split"]
    if "plain" not in paths:
        paths = ["plain"] + paths
    for path in paths:
        eng = Engine(geo)
        eng.set_targets(tx)
        t0 = time.perf_counter()
        nnz = eng.nearlist()
        t_near = time.perf_counter() - t0
        far = eng.far(ub, ubn)
        if path == "selector":
            eng.set_selector(sel, a.threshold)
        elif path == "split":
            eng.set_precision(a.split_levels)
        eng.eval(exterior=True, side_tol=1e-10)
        u0, _ = eng.apply(ub, ubn, far)  # uploads ub, ubn, u (= far) and checks the path once
        lab = eng.weights()["label"]
        counts = np.bincount(lab, minlength=6)
        elapsed, stage = timed_steps(eng, a.steps, a.warmup)
        u, _ = eng.apply(ub, ubn, far)
        offs, idx = eng.pairs()
        csum = np.concatenate([[0], np.cumsum(lab == 5)])
        rrq_panels = int(((csum[offs[1:]] - csum[offs[:-1]]) > 0).sum())
        runs[path] = dict(eng=eng, nnz=nnz, t_near=t_near, counts=counts, elapsed=elapsed, stage=stage, u=u,
                          u0=u0, rrq_panels=rrq_panels, offs=offs, idx=idx)
        ms = 1e3 * elapsed / a.steps
        print(f"[bench] {path}: near list {nnz} pairs ({t_near * 1e3:.1f} ms incl. host sync); served by own nodes, "
              f"L1..L4, RRQ: {counts.tolist()}; {a.steps} steps {ms:.2f} ms/step, {nnz * a.steps / elapsed:.4g} "
              f"pairs/s", flush=True)
        if path != a.path:
            eng.close()
            r = runs[path]
            r["eng"] = None

    # the configuration's "radius 2h" near list (h: the panel's largest corner
    # distance), timed beside the reference ball rule; value stays on the ball rule
    near2h = None
    if not a.no_compare:
        eng = Engine(geo)
        eng.set_targets(tx)
        eng.set_near_radius(2.0)
        t0 = time.perf_counter()
        nnz2 = eng.nearlist()
        t_near2 = time.perf_counter() - t0
        far2 = eng.far(ub, ubn)
        eng.eval(exterior=True, side_tol=1e-10)
        eng.apply(ub, ubn, far2)
        elapsed2, _ = timed_steps(eng, a.steps, a.warmup)
        u2, _ = eng.apply(ub, ubn, far2)
        eng.close()
        near2h = {"pairs": int(nnz2), "near_list_ms": 1e3 * t_near2, "ms_per_step": 1e3 * elapsed2 / a.steps,
                  "pairs_per_s": nnz2 * a.steps / elapsed2, "grf_rel_err": float(np.abs(u2 - uref).max() / scale)}
        print(f"[bench] plain, radius-2h near list: {nnz2} pairs ({nnz2 / runs['plain']['nnz']:.2f}x the ball "
              f"rule, {1e3 * t_near2:.1f} ms incl. host sync); {a.steps} steps {near2h['ms_per_step']:.2f} ms/step, "
              f"{near2h['pairs_per_s']:.4g} pairs/s; GRF rel err {near2h['grf_rel_err']:.3e}", flush=True)

    nnz = runs["plain"]["nnz"]
    grf = {k: float(np.abs(r["u"] - uref).max() / scale) for k, r in runs.items()}
    diff = {k: float(np.abs(r["u"] - runs["plain"]["u"]).max() / scale) for k, r in runs.items()}
    repeat = {k: float(np.abs(r["u"] - r["u0"]).max() / np.abs(r["u0"]).max()) for k, r in runs.items()}
    value = {k: nnz * a.steps / r["elapsed"] for k, r in runs.items()}
    for k in runs:
        print(f"[bench] {k}: GRF rel err {grf[k]:.3e}; vs plain max |du| / max |u| = {diff[k]:.2e}; "
              f"pairs/s {value[k]:.4g} ({value[k] / value['plain']:.3f} of plain)", flush=True)

    if a.calib_copy_mib > 0:
        # plain copy kernel of known size for the HBM-traffic counter calibration (not timed)
        from qotential_amd._lib import lib, check
        import ctypes as C
        cms = C.c_double(0.0)
        check(lib().qol_hbm_copy(C.c_int64(a.calib_copy_mib << 20), C.byref(cms)), "qol_hbm_copy")
        gbs = 2 * (a.calib_copy_mib << 20) / (cms.value * 1e-3) / 1e9
        print(f"[bench] calibration copy: {a.calib_copy_mib} MiB read, {a.calib_copy_mib} MiB written "
              f"({cms.value:.3f} ms, {gbs:.0f} GB/s)", flush=True)

    # roofline of the dominant kernel (most algorithmic work) of the measured path
    r = runs[a.path]
    ns = a.split_levels if a.path == "split" else 0
    kf = kernel_flops(a.order, geo["ntri"], r["counts"][1:5], int(r["counts"][5]), r["rrq_panels"], ns)
    stage_of = {"levels": "levels", "rrq_solve": "rrq_solve", "rrq_moments": "rrq_moments",
                "rrq_weights": "rrq_weights", "direct_split": "direct", "direct": "direct"}
    kernels = {}
    for st, (kname, fl, peak) in kf.items():
        ms = r["stage"][stage_of[st]]
        if st.startswith("direct") and ns > 0 and ns < 4:
            # both direct kernels share the stage's events: split the time by FLOPs at each peak
            tot = sum(kf[s][1] / kf[s][2] for s in ("direct_split", "direct"))
            ms = ms * (fl / peak) / tot
        kernels[kname] = {"ms": ms, "gflop": fl / 1e9, "tflops": fl / (ms * 1e-3) / 1e12, "peak": peak}
    dom = max(kernels, key=lambda k: kernels[k]["gflop"])
    kd = kernels[dom]
    roof = {"kernel": dom, "bound": "mfma", "achieved": kd["tflops"], "peak": kd["peak"], "unit": "TFLOP/s",
            "frac": kd["tflops"] / kd["peak"], "traffic": measured_traffic(dom, a),
            "flops_per_launch": kd["gflop"] * 1e9, "ms_per_launch": kd["ms"]}

    base = None
    if not a.no_cpu_baseline:
        p0 = runs["plain"]
        base = cpu_baseline(geo, tx, p0["offs"], p0["idx"], ub, ubn, a.cpu_sample_pairs)
        print(f"[bench] cpu baseline {base['value']:.4g} pairs/s with {base['cores']} threads, "
              f"{base['value_1thread']:.4g} with 1 ({base['cpu'].get('Model name', '?')})", flush=True)
        base["reference_configs"] = cpu_reference_configs()
        for k, v in base["reference_configs"].items():
            print(f"[bench] cpu {k}: {v['pairs']} pairs, {v['pairs_per_s']:.4g} pairs/s (1 thread)", flush=True)

    runs[a.path]["eng"].close()
    line = {
        "metric": "close-eval target-panel pairs/sec at GRF rel err <=1e-9 vs plain FP64 on GPU",
        "value": value[a.path], "unit": "pairs/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": 1e3 * runs[a.path]["elapsed"] / a.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": f"f64 (levels 1..{a.split_levels} direct weights f32)" if a.path == "split" else "f64",
        "data": "synthetic",
        "config": {"workload": f"{a.path} RRQ close evaluation, stellarator {a.mp}x{a.np} charts, order {a.order}, "
                               f"{geo['ntri']} triangles, {a.targets} near-surface targets both sides "
                               f"U[-6,-1], reference ball near list",
                   "path": a.path, "pairs": int(nnz),
                   "served_own_L1_L4_RRQ": {k: r_["counts"].tolist() for k, r_ in runs.items()},
                   "pairs_per_s": value, "ms_per_step": {k: 1e3 * r_["elapsed"] / a.steps for k, r_ in runs.items()},
                   "selector": {"model": os.path.relpath(str(MODEL_FILE), ROOT), "threshold": a.threshold},
                   "split_levels": a.split_levels,
                   "stage_ms": {k: r_["stage"] for k, r_ in runs.items()}, "kernels": kernels,
                   "grf_rel_err": grf, "rel_err_vs_plain_fp64": diff[a.path], "paths_rel_err_vs_plain_fp64": diff,
                   "repeat_rel_diff": repeat, "plain_near_radius_2h": near2h},
        "roofline": roof,
        "cpu_baseline": base,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
