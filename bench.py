"""Benchmark of the RPD hot path on one MI355X.

Default line (``--workload headline``): learned+batched against plain on
the clustered-sphere stress batch -- 128 independent instances of
50k-500k tets (jittered Kuhn grids) with 5k-40k spheres from a
32-component Gaussian mixture, 25.4 M tets in all, resident in HBM as one
batch.  A step is one rpd_run over the whole batch.  ``value`` is the
learned+batched rate (cells/s): candidates proposed by the MLP (trained
beforehand on the GPU, by the library's own kernels, on 16 other instances
of the same generator), certified, the missed cells re-clipped, the output
checked bit-identical to the plain batched step's.  ``plain`` holds the
plain path batched and run one instance at a time, ``fallback_pct`` the
re-run tets.  Sub-keys: ``cube64`` (one 64^3 instance, 1.57 M tets, plain
vs learned) and ``config3`` (256 instances of 2k-200k tets over 8
sphere-update iterations).  ``cpu_baseline``: the reference algorithm on
the host cores, at one thread and at all threads, on a sample of the same
batch.  Prints one JSON line.

roofline: the dominant kernel of the learned+batched step, its average
launch duration measured with HIP events on the context stream over all of
its launches in a step.  For the clip kernel (FP64 VALU work) achieved =
the algorithmic FP64 flops it executes (45 per 4x4 conflict determinant, 14
per bisector, 59 per output vertex, counted by the kernel in an extra
untimed run) over that time; the HBM bytes per launch it moves come from
rocprofv3 FETCH_SIZE / WRITE_SIZE (profiles/).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
METRIC = ("RPD cells/sec (tets clipped) learned+batched vs plain on 1 "
          "MI355X; fallback %")
# small-tier capacities (vertices, planes, edges) by rpd_run_cfg.small_caps
CAPS_NAMES = {0: "Caps<32, 16, 48>", 1: "Caps<48, 24, 72>",
              2: "Caps<16, 12, 32>", 3: "Caps<24, 14, 40>"}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cells-per-wave", type=int, default=0,
                    help="cells per wavefront of the clip kernel; 0 = the "
                         "autotuner's choice on the instance (untimed)")
    ap.add_argument("--small-caps", type=int, default=0,
                    help="small-tier capacities with --cells-per-wave > 0")
    ap.add_argument("--workload", default="headline",
                    choices=["headline", "cube10", "cube64", "batched",
                             "gmm", "heldout10"])
    ap.add_argument("--no-learned", action="store_true",
                    help="time only the plain path")
    ap.add_argument("--batch-instances", type=int, default=256)
    ap.add_argument("--batch-iterations", type=int, default=8)
    ap.add_argument("--no-cube64", action="store_true",
                    help="headline: skip the cube64 sub-key")
    ap.add_argument("--no-config3", action="store_true",
                    help="headline: skip the config3 sub-key")
    ap.add_argument("--learned-only", action="store_true",
                    help="gmm: time only the learned+batched step (the "
                         "command the rocprofv3 summaries are taken of)")
    ap.add_argument("--batch-model", default="cube10",
                    choices=["same", "cube10"],
                    help="batched: train the proposal model on other "
                         "instances of the batch generator, or on the 10^3 "
                         "Kuhn grid generator (seeds 0-63)")
    ap.add_argument("--cpu-sample-tets", type=int, default=2000000,
                    help="tets of the CPU baseline's sample at all threads "
                         "(1/16 of it at one thread)")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the captured-graph leg of the batched modes")
    ap.add_argument("--no-work-count", action="store_true",
                    help="skip the extra untimed run that counts FP64 work "
                         "(keeps profiler averages to the timed launches)")
    ap.add_argument("--traffic-json", default=os.path.join(
        ROOT, "profiles", "clip_traffic_cube64.json"),
        help="cube64: per-launch HBM traffic of the clip kernel measured by "
             "rocprofv3 --pmc (FETCH_SIZE / WRITE_SIZE, copy-calibrated)")
    ap.add_argument("--traffic-json-gmm", default=os.path.join(
        ROOT, "profiles", "clip_traffic_gmm128.json"),
        help="config 5: the same for the learned step's dominant kernel")
    ap.add_argument("--calib-copy-gib", type=float, default=0.0,
                    help="also run one device copy of this many GiB (the "
                         "plain copy kernel that calibrates PMC counters)")
    return ap.parse_args()


def clip_bytes(st, inst):
    """Compulsory HBM bytes of one small-tier clip launch, each counted once:
    reads of the candidate records (24 B each), the tet vertex ids and
    packed adjacency counts (16 + 16 B per tet), the vertices (24 B), the
    sites (32 B), the neighbour-list offsets (8 B) and ids (4 B); writes of
    one 4 B header per candidate and of the emitted cells' payload words
    (4 B), volume + centroid (32 B) and facet areas (8 B)."""
    reads = (st["n_candidates"] * 24 + inst.n_tet * 32
             + inst.verts.shape[0] * 24 + inst.n_site * (32 + 8)
             + inst.knn_ids.size * 4)
    writes = (st["n_candidates"] * 4 + st["n_words"] * 4
              + st["n_cells"] * 32 + st["n_areas"] * 8)
    return reads + writes


def clip_flops(w):
    """Algorithmic FP64 flops of one clip launch from its executed-work
    counts: 45 per 4x4 conflict determinant, 14 per bisector plane, 59 per
    emitted vertex position (four 3x3 determinants and three divisions)."""
    return 45 * w[0] + 14 * w[1] + 59 * w[2]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def _cpu_rate(insts, n_tets, threads, rng):
    """Cells/s of the reference algorithm on n_tets random tets spread over
    insts in proportion to their sizes: dense relation over every site, then
    the clip of each related cell by its full neighbour list (best of 2)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    sizes = np.array([i.n_tet for i in insts], np.float64)
    per = np.maximum(1, np.round(n_tets * sizes / sizes.sum())).astype(int)
    t_rel = t_clip = 0.0
    cells = tets_done = 0
    for inst, k in zip(insts, per):
        tets = np.sort(rng.choice(inst.n_tet, min(int(k), inst.n_tet),
                                  replace=False)).astype(np.int32)
        oracle.run(inst, tets=tets[:64], mode=1, n_threads=threads)  # warm
        t0 = time.perf_counter()
        rel = oracle.run(inst, tets=tets, mode=1, n_threads=threads)
        t_rel += time.perf_counter() - t0
        best = None
        for _ in range(2):
            t0 = time.perf_counter()
            r = oracle.clip(inst, tets, rel.rel_off, rel.rel_ids, mode=1,
                            n_threads=threads)
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        t_clip += best
        cells += r.cells.size
        tets_done += tets.size
    return cells / (t_rel + t_clip), cells, tets_done, t_rel, t_clip


def cpu_baseline(insts, sample_tets, threads, label):
    """The reference algorithm on the host cores (the reference has no CPU
    RPD: src/CPU/voronoi_fallback.cpp:21-22 is an empty stub, so the
    oracle's restatement in reference arithmetic -- float planes, FP64
    determinant, convex_cell.cu:116-774; dense relation, voronoi.cu:154-193
    -- is the baseline), multithreaded over tets, on a random sample of
    the benchmark's own instances; once at ``threads`` threads and once at
    one thread on 1/16 of the sample."""
    import numpy as np
    rng = np.random.default_rng(12345)
    v, n, nt, tr, tc = _cpu_rate(insts, sample_tets, threads, rng)
    v1, n1, nt1, tr1, tc1 = _cpu_rate(insts, max(1, sample_tets // 16), 1,
                                      rng)
    return {"value": v, "unit": "cells/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(), "host_threads": os.cpu_count(),
            "one_thread": {"value": v1, "cores": 1,
                           "sample": f"{nt1} tets ({n1} cells): relation "
                                     f"{tr1:.2f} s + clip {tc1:.2f} s"},
            "sample": f"{nt} random tets of {label} ({n} cells): dense "
                      f"relation over all sites of each instance {tr:.2f} s "
                      f"+ clip by full neighbour lists in reference float "
                      f"arithmetic {tc:.2f} s (best of 2), {threads} "
                      f"threads"}


def same_result(a, b):
    """Bit-identical cell streams (records, payload words, geometry,
    areas)."""
    import numpy as np
    for f in ("words", "geom", "areas"):
        if not np.array_equal(getattr(a, f), getattr(b, f)):
            return False
    for k in ("tet", "seed", "nb_v", "nb_p", "nb_e", "status", "word_off",
              "area_off"):
        if not np.array_equal(a.cells[k], b.cells[k]):
            return False
    return True


def train_model(instances, epochs, batch, recall=0.999):
    """The proposal model trained on the GPU (own kernels) on the plain
    path's cells of ``instances`` (other seeds of the workload's
    generator), threshold at ``recall`` of the training proposals; returns
    (params, logit threshold, seconds, loss per epoch)."""
    from libmat_amd import learn
    from libmat_amd.rpd import Context
    t0 = time.perf_counter()
    with Context(0) as tc:
        params, loss, thr = learn.train_proposals(
            tc, instances, epochs=epochs, batch=batch, recall=recall)
    return params, thr, time.perf_counter() - t0, [float(x) for x in loss]


def time_batch(ctx, sets, warmup, **kw):
    """Runs of one resident batch: site set ``sets[i]`` for step i (None =
    keep the resident sites).  Returns per-step device ms, cells, learned
    counters summed, and the H2D seconds of the site updates; the averages
    of the per-kernel times are left in time_batch.kernel."""
    if sets[0] is not None:
        ctx.update_sites_batch(sets[0])
    for _ in range(warmup):
        ctx.run(**kw)
    ms, cells, lrn, t_sites = [], [], [0, 0, 0, 0], 0.0
    kern = []
    for it, st_set in enumerate(sets):
        if it and st_set is not None:
            t1 = time.perf_counter()
            ctx.update_sites_batch(st_set)
            t_sites += time.perf_counter() - t1
        st = ctx.run(**kw)
        ms.append(st["ms_total"])
        cells.append(st["n_cells"])
        kern.append(st)
        lrn = [x + y for x, y in zip(lrn, st["learned"])]
    n = len(kern)

    def avg(f):
        return sum(f(k) for k in kern) / n
    if kw.get("learned"):
        names = ["features_phase", "mlp", "proposal_count", "proposal_fill",
                 "plane_masks"]
    else:
        names = ["site_boxes", "binning", "relation_count", "relation_fill",
                 "plane_masks"]
    km = {nm: avg(lambda k, i=i: k["ms_kernel"][i])
          for i, nm in enumerate(names)}
    km.update({
        "clip_first_round": avg(lambda k: k["ms_kernel"][5]),
        "clip_small_tier_all_rounds": avg(lambda k: k["ms_clip_small_sum"]),
        "clip_small_tier_launches": avg(lambda k: k["n_clip_small"]),
        "host_reads": avg(lambda k: k["n_host_reads"]),
        "capacity_and_exact_reruns": avg(lambda k: k["ms_phase"][2]),
        "compaction": avg(lambda k: k["ms_phase"][3])})
    if kw.get("learned"):
        km["features_kernel"] = avg(lambda k: k["ms_features_kernel"])
        km["certificate_closure_fallback"] = avg(lambda k: k["ms_phase"][4])
    time_batch.kernel = {k: round(float(v), 4) for k, v in km.items()}
    return ms, cells, lrn, t_sites


def feature_bytes(insts, n_knn):
    """Compulsory HBM bytes of one proposal/feature launch: per tet its
    vertex ids (16 B) and instance id (4 B) read, its 32 proposals (128 B),
    48 bf16 features (96 B) and owner mask (4 B) written; per vertex its
    coordinates (24 B) and owner (4 B); per site 32 B; per neighbour-list
    entry its id and the neighbour's sphere (36 B)."""
    n_tet = sum(i.n_tet for i in insts)
    n_vert = sum(i.n_vert for i in insts)
    n_site = sum(i.n_site for i in insts)
    return n_tet * (16 + 4 + 128 + 96 + 4) + n_vert * 28 + n_site * 32 + \
        n_knn * 36


def roofline_of_step(kernel_ms, work, n_launch, insts, G, caps, traffic_json):
    """Roofline of the dominant kernel of a learned step: the small-tier
    clip kernel (FP64 VALU work: achieved = algorithmic flops over its
    launches' summed time) or the proposal/feature kernel (compulsory HBM
    bytes over its time), whichever takes longer."""
    clip_ms = kernel_ms["clip_small_tier_all_rounds"]
    feat_ms = kernel_ms.get("features_kernel", 0.0)
    name = f"k_clip<{G}, {CAPS_NAMES[caps]}, false, true>"
    traffic = None
    if traffic_json and os.path.exists(traffic_json):
        with open(traffic_json) as f:
            tj = json.load(f)
        if tj.get("kernel", "").replace("rpd::", "").replace(
                "void ", "").startswith(name if clip_ms >= feat_ms else
                                        "k_knn_features"):
            traffic = tj.get("bytes_per_launch")
    if clip_ms >= feat_ms:
        flops = clip_flops(work)
        achieved = flops / (clip_ms * 1e-3) / 1e12
        return {"bound": "valu", "achieved": achieved,
                "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic,
                "kernel": name, "launches_per_step": n_launch,
                "kernel_ms": clip_ms / max(1, n_launch),
                "flops_per_step": flops, "n_det": work[0],
                "n_bisector": work[1], "n_vout": work[2],
                "note": "FP64 VALU-issue-bound (SQ counters in profiles/); "
                        "achieved = executed algorithmic FP64 flops of the "
                        "step's launches of this kernel / their summed "
                        "device time; kernel_ms = average per launch"}
    n_knn = sum(int(i.knn_ids.size) for i in insts)
    b = feature_bytes(insts, n_knn)
    achieved = b / (feat_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS,
            "unit": "GB/s", "frac": achieved / PEAK_HBM_GBS,
            "traffic": traffic, "kernel": "k_knn_features",
            "launches_per_step": 1, "kernel_ms": feat_ms, "alg_bytes": b}


def bench_batched(a, emit=True):
    """Many independent instances in one resident batch.

    ``batched`` (config 3): 256 instances (tets log-uniform in [2k, 200k]
    as jittered Kuhn grids, spheres log-uniform in [100, 20k],
    r ~ U[0.02, 0.08]) over sphere-update iterations with centre jitter
    N(0, 0.002^2); the learned model is trained on the 10^3 Kuhn grid
    generator, seeds 0-63 (or, with --batch-model same, on 32 other
    instances of the batch generator).  ``gmm`` (config 5): 128 instances
    of 50k-500k tets with 5k-40k clustered spheres (32-component Gaussian
    mixture), ``--steps`` runs of the same batch; the model is trained on
    16 instances of the same generator with another seed.  A step's time
    is the device time of rpd_run (HIP events); per-iteration neighbour
    lists are built beforehand (device builder, rpd_power_neighbors_gpu)
    and their H2D copy is reported apart.  Three paths on the same inputs:
    plain batched, learned+batched (checked bit-identical to plain on the
    last step) and the instances one at a time through the plain path."""
    import numpy as np
    from libmat_amd import prep
    from libmat_amd.rpd import Context

    def progress(b):
        if b % 16 == 15:
            print(f"[{a.workload}] {b + 1} instances prepared",
                  file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    if a.workload == "batched":
        w = prep.batch_workload(a.batch_instances, 0, progress=progress)
        sets = [[(i.sites, i.flags, i.knn_off, i.knn_ids)
                 for i in w.instances]]
        # the same iterations with every sphere a site (hidden ones
        # unflagged), so that each instance keeps its site count: the
        # captured-graph leg replays one graph over them
        sets_all = [[prep.sites_both(x)[1] for x in w.spheres]]
        cur = list(w.spheres)
        # the neighbour lists of each sphere-update iteration come from the
        # device builder (identical to the host builder's)
        nb_ctx = Context(0)
        nb_host = 0

        def sites(x):
            nonlocal nb_host
            try:
                return prep.sites_both(x, ctx=nb_ctx)
            except RuntimeError:  # a cell over the device builder's caps
                nb_host += 1
                return prep.sites_both(x)
        for it in range(1, a.batch_iterations):
            cur = [prep.jitter_spheres(x, 0.002, 7919 * it + b)
                   for b, x in enumerate(cur)]
            both = [sites(x) for x in cur]
            sets.append([x[0] for x in both])
            sets_all.append([x[1] for x in both])
            print(f"[batched] neighbour lists of iteration {it} ready",
                  file=sys.stderr, flush=True)
        nb_ctx.close()
        if nb_host:
            print(f"[batched] {nb_host} sphere sets over the device "
                  f"neighbour builder's capacities: host builder used",
                  file=sys.stderr, flush=True)
        desc = (f"{a.batch_instances} instances, tets log-U[2k,200k] "
                f"(jittered Kuhn grids), spheres log-U[100,20k] "
                f"r~U[0.02,0.08], {a.batch_iterations} iterations of centre "
                f"jitter N(0,0.002^2)")
    else:
        n_inst = a.batch_instances if a.batch_instances != 256 else 128
        w = prep.gmm_workload(n_inst, 0, progress=progress)
        sets = [[(i.sites, i.flags, i.knn_off, i.knn_ids)
                 for i in w.instances]] + [None] * (a.steps - 1)
        sets_all = sets
        desc = (f"{n_inst} instances, tets log-U[50k,500k] (jittered Kuhn "
                f"grids), spheres log-U[5k,40k] from a 32-component "
                f"Gaussian mixture (sigma~U[0.02,0.1]), r~lognormal(-3.5,"
                f"0.6) clipped to the boundary distance")
    print(f"[{a.workload}] host preparation done", file=sys.stderr, flush=True)
    t_prep = time.perf_counter() - t0
    n_tet = sum(i.n_tet for i in w.instances)
    ctx = Context(0)
    ctx.upload_batch(w.instances)
    G, caps, tune_ms = ctx.tune()
    batched = None
    if not a.learned_only:
        ms, cells, _, t_sites = time_batch(ctx, sets, a.warmup,
                                           cells_per_wave=0)
        batched = {"value": sum(cells) / (sum(ms) * 1e-3), "unit": "cells/s",
                   "ms_per_step": sum(ms) / len(ms), "step_ms": ms,
                   "step_cells": cells, "kernel_ms": time_batch.kernel}
        print(f"[{a.workload}] plain batched timed", file=sys.stderr,
              flush=True)
    graph = None
    if not a.learned_only and not a.no_graph:
        graph = bench_graph(ctx, sets_all, a.warmup, G, caps)
        print(f"[{a.workload}] captured graph timed", file=sys.stderr,
              flush=True)
    learned = roofline = None
    if not a.no_learned:
        if a.workload == "batched" and a.batch_model == "cube10":
            train = [prep.cube10_instance(sd) for sd in range(64)]
            params, thr, t_train, loss = train_model(train, 6, 1024)
            model = ("trained on the 10^3 Kuhn grid generator, seeds 0-63 "
                     "(6 epochs)")
        elif a.workload == "batched":
            train = prep.batch_workload(32, 1).instances
            params, thr, t_train, loss = train_model(train, 3, 2048)
            model = ("trained on 32 instances of the same generator, seed 1 "
                     "(3 epochs)")
        else:
            train = prep.gmm_workload(16, 1).instances
            params, thr, t_train, loss = train_model(train, 3, 2048)
            model = ("trained on 16 instances of the same generator, seed 1 "
                     "(3 epochs)")
        del train
        print(f"[{a.workload}] model trained", file=sys.stderr, flush=True)
        ctx.mlp_set(params)
        lkw = dict(cells_per_wave=0, learned=1, logit_threshold=thr)
        lms, lcells, lrn, _ = time_batch(ctx, sets, a.warmup, **lkw)
        lkern = time_batch.kernel
        identical = None
        if not a.learned_only:
            # the last step's cells, learned against plain, bit for bit
            ref = ctx.download()
            ctx.run(cells_per_wave=0)
            identical = same_result(ctx.download(), ref) and lcells == cells
            del ref
        # executed FP64 work of the learned step's clip launches (an extra,
        # untimed run with the kernels' work counters)
        if not a.no_work_count:
            stw = ctx.run(count_work=True, **lkw)
            roofline = roofline_of_step(
                lkern, stw["work"][0], stw["n_clip_small"], w.instances, G,
                caps, a.traffic_json_gmm if a.workload == "gmm" else None)
        learned = {
            "value": sum(lcells) / (sum(lms) * 1e-3), "unit": "cells/s",
            "ms_per_step": sum(lms) / len(lms), "step_ms": lms,
            "identical_to_plain": identical,
            "fallback_pct": 100.0 * lrn[1] / (n_tet * len(sets)),
            "fallback_tets": lrn[1], "full_relation_tets": lrn[3],
            "first_round_candidates": lrn[0], "added_candidates": lrn[2],
            "model": model, "logit_threshold": thr, "train_loss": loss,
            "train_s": t_train, "kernel_ms": lkern}
        print(f"[{a.workload}] learned+batched timed", file=sys.stderr,
              flush=True)
        if not a.learned_only and not a.no_graph:
            # the learned step captured: certificate, closure rounds and
            # output order with their counts on the device
            learned["graph"] = bench_graph(ctx, sets_all, a.warmup, G, caps,
                                           learned=1, logit_threshold=thr)
            print(f"[{a.workload}] captured learned graph timed",
                  file=sys.stderr, flush=True)
    ctx.close()
    if a.calib_copy_gib > 0:  # the copy kernel that calibrates PMC counters
        with Context(0) as cc:
            cc.calib_copy(int(a.calib_copy_gib * (1 << 30)))
    one = None
    if not a.learned_only:
        # the same instances, one plain run each (first step's spheres)
        c1 = Context(0)
        p_ms, p_cells = 0.0, 0
        for inst in w.instances:
            c1.upload(inst)
            c1.run()
            st = c1.run()
            p_ms += st["ms_total"]
            p_cells += st["n_cells"]
        c1.close()
        one = {"value": p_cells / (p_ms * 1e-3), "unit": "cells/s",
               "ms_total": p_ms, "cells": p_cells}
    top = learned if learned is not None else batched
    out = {
        "metric": METRIC, "value": top["value"], "unit": "cells/s",
        "n_gpus": 1, "steps": len(sets), "warmup": a.warmup,
        "ms_per_step": top["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": desc, "n_instances": len(w.instances),
                   "n_tet": n_tet,
                   "n_site": sum(i.n_site for i in w.instances),
                   "path": ("learned+batched" if learned is not None
                            else "plain batched"),
                   "tuned_cells_per_wave": G, "tuned_small_caps": caps},
        "plain": {"batched": batched, "one_at_a_time": one,
                  "graph": graph},
        "learned": learned,
        "learned_vs_plain_batched": (
            learned["value"] / batched["value"]
            if learned is not None and batched is not None else None),
        "learned_vs_plain_one_at_a_time": (
            learned["value"] / one["value"]
            if learned is not None and one is not None else None),
        "fallback_pct": learned["fallback_pct"] if learned else 0.0,
        "identical_to_plain": learned["identical_to_plain"] if learned
        else None,
        "roofline": roofline,
        "tets_per_s": n_tet * 1e3 / top["ms_per_step"],
        "tune_ms": {f"caps{c}_G{g}": v for (c, g), v in tune_ms.items()},
        "host_prep_s": t_prep,
    }
    out["_instances"] = w.instances
    if emit:
        print(json.dumps({k: v for k, v in out.items()
                          if not k.startswith("_")}))
    return out


def bench_graph(ctx, sets, warmup, G, caps, **path):
    """The batched step (plain, or learned with path = learned=1, ...) as
    one replay of a captured hipGraph (rpd_graph_build / rpd_graph_run)
    against eager rpd_run on the same site sets (every sphere a site, so
    instances keep their site counts across iterations); both timed with
    HIP events around the device work.  A replay whose counts outgrow the
    capture's capacities, or a changed batch, is captured again (counted in
    "recaptured")."""
    kw = dict(cells_per_wave=G, small_caps=caps, **path)
    ems, ecells, _, _ = time_batch(ctx, sets, warmup, **kw)
    if sets[0] is not None:
        ctx.update_sites_batch(sets[0])
    t0 = time.perf_counter()
    ctx.graph_build(headroom=0.1, **kw)
    t_build = time.perf_counter() - t0
    for _ in range(warmup):
        ctx.graph_run()
    gms, gcells, replayed, wall = [], [], 0, []
    for it, st_set in enumerate(sets):
        if it and st_set is not None:
            ctx.update_sites_batch(st_set)
        t1 = time.perf_counter()
        st, rep = ctx.graph_run()
        wall.append((time.perf_counter() - t1) * 1e3)
        gms.append(st["ms_total"])
        gcells.append(st["n_cells"])
        replayed += rep
    ref = ctx.download()
    ctx.run(**kw)
    identical = same_result(ctx.download(), ref) and gcells == ecells
    ctx.graph_destroy()
    return {"value": sum(gcells) / (sum(gms) * 1e-3), "unit": "cells/s",
            "ms_per_step": sum(gms) / len(gms), "step_ms": gms,
            "host_wall_ms_per_step": sum(wall) / len(wall),
            "eager_ms_per_step": sum(ems) / len(ems),
            "eager_value": sum(ecells) / (sum(ems) * 1e-3),
            "replayed": replayed, "recaptured": len(sets) - replayed,
            "build_s": t_build, "identical_to_eager": identical,
            "sites": "every sphere a site (hidden ones unflagged)"}


def bench_heldout10(a):
    """The learned-candidate protocol on the 10^3 Kuhn grid generator: the
    model is trained on seeds 0-63 and evaluated on held-out seeds
    1000-1015, resident as one batch.  Reports the proposals' precision
    and recall at the chosen threshold, the fallback rate, and cells/s of
    the plain, learned and top-k (first k proposals by power distance, no
    model) paths on the same batch, each checked bit-identical to plain."""
    import numpy as np
    from libmat_amd import learn, prep
    from libmat_amd.rpd import Context
    train = [prep.cube10_instance(sd) for sd in range(64)]
    params, thr, t_train, loss = train_model(train, 6, 1024)
    test = [prep.cube10_instance(sd) for sd in range(1000, 1016)]
    n_tet = sum(i.n_tet for i in test)
    ctx = Context(0)
    ctx.upload_batch(test)
    ctx.mlp_set(params)
    ctx.run()
    ref = ctx.download()
    ctx.features(download=False)
    Y, M = ctx.labels()
    prec, rec = learn.proposal_quality(ctx.scores(), Y, M, thr)
    paths = {"plain": {}, "learned": dict(learned=1, logit_threshold=thr),
             "topk8": dict(learned=2, proposal_k=8)}
    res = {}
    for name, kw in paths.items():
        ms, cells, lrn, _ = time_batch(ctx, [None] * a.steps, a.warmup,
                                       cells_per_wave=0, **kw)
        same = same_result(ctx.download(), ref)
        res[name] = {"value": sum(cells) / (sum(ms) * 1e-3),
                     "ms_per_step": sum(ms) / len(ms),
                     "identical_to_plain": same,
                     "fallback_pct": 100.0 * lrn[1] / (n_tet * a.steps),
                     "full_relation_tets": lrn[3] // a.steps,
                     "first_round_candidates": lrn[0] // a.steps}
    ctx.close()
    out = {
        "metric": METRIC, "value": res["learned"]["value"], "unit": "cells/s",
        "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
        "ms_per_step": res["learned"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": "10^3 Kuhn grid, 500 spheres r~U[0.02,0.08]: "
                   "model trained on seeds 0-63, evaluated on the batch of "
                   "held-out seeds 1000-1015", "n_tet": n_tet,
                   "path": "learned"},
        "paths": res,
        "learned_vs_plain": res["learned"]["value"] / res["plain"]["value"],
        "learned_vs_topk8": res["learned"]["value"] / res["topk8"]["value"],
        "precision": prec, "recall": rec, "logit_threshold": thr,
        "train_loss": loss, "train_s": t_train,
        "fallback_pct": res["learned"]["fallback_pct"],
    }
    print(json.dumps(out))


def bench_cube(a, emit=True):
    """One instance: the 64^3 jittered Kuhn grid (1,572,864 tets, 50,000
    spheres under the boundary radius rule) or the 10^3 grid; plain path
    against the learned path (model trained on two other seeds of the same
    generator), 20 timed steps each."""
    import numpy as np
    from libmat_amd import prep
    from libmat_amd.rpd import Context
    t0 = time.perf_counter()
    inst = (prep.cube64_instance(0) if a.workload == "cube64"
            else prep.cube10_instance(0))
    t_prep = time.perf_counter() - t0
    ctx = Context(0)
    t0 = time.perf_counter()
    ctx.upload(inst)
    t_upload = time.perf_counter() - t0
    tune_ms = None
    if a.cells_per_wave == 0:
        G, caps, tune_ms = ctx.tune()
    else:
        G, caps = a.cells_per_wave, a.small_caps
    shape = dict(cells_per_wave=G, small_caps=caps)
    for _ in range(a.warmup):
        ctx.run(**shape)
    ctx._L.rpd_sync(ctx._h)
    stats = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        stats.append(ctx.run(**shape))
    ctx._L.rpd_sync(ctx._h)
    dt = time.perf_counter() - t0
    st = stats[-1]
    ms_step = 1e3 * dt / a.steps
    # one more, untimed run that counts the executed FP64 work
    work = None
    active_frac = None
    if not a.no_work_count:
        stw = ctx.run(count_work=True, **shape)
        work = stw["work"]
        active_frac = stw["n_active_planes"] / max(1, stw["n_all_planes"])
    n_cells = st["n_cells"]
    value = n_cells * a.steps / dt
    ph = np.mean(np.array([s["ms_phase"] for s in stats]), axis=0)
    kn = np.mean(np.array([s["ms_kernel"] for s in stats]), axis=0)
    clip_ms = float(kn[5])
    alg = clip_bytes(st, inst)
    achieved = alg / (clip_ms * 1e-3) / 1e9
    fp64 = None
    if work is not None:
        flops = clip_flops(work[0])
        achieved_fp64 = flops / (clip_ms * 1e-3) / 1e12
        fp64 = {"achieved": achieved_fp64, "peak": PEAK_FP64_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved_fp64 / PEAK_FP64_TFLOPS,
                "flops": flops, "n_det": work[0][0],
                "n_bisector": work[0][1], "n_vout": work[0][2],
                "n_exact_det_all_tiers": sum(w[3] for w in work)}
    clip_kernel = f"k_clip<{G}, {CAPS_NAMES[caps]}, false, false>"
    traffic = None
    if os.path.exists(a.traffic_json):
        with open(a.traffic_json) as f:
            tj = json.load(f)
        # only a measurement of this very kernel counts
        if tj.get("kernel", "").replace("rpd::", "").replace(
                "void ", "").startswith(clip_kernel):
            traffic = tj.get("bytes_per_launch")
    # ---- the rows after the clip on the plain run's cells: per-sphere
    # adjacency and components, per-site centroids (host wall time of the
    # calls: device work plus their few count read-backs; nothing copied
    # back to the host)
    import ctypes
    from libmat_amd.rpd import AdjSizes
    ctx.run(**shape)
    after = {}
    for name, call in (
            ("adjacency_ms", lambda: ctx._L.rpd_adjacency(
                ctx._h, ctypes.byref(AdjSizes()))),
            ("centroids_ms", lambda: ctx._L.rpd_site_centroids(
                ctx._h, None, ctypes.byref(ctypes.c_void_p())))):
        assert call() == 0
        t1 = time.perf_counter()
        for _ in range(3):
            assert call() == 0
        after[name] = (time.perf_counter() - t1) * 1e3 / 3
    z = AdjSizes()
    ctx._L.rpd_adjacency(ctx._h, ctypes.byref(z))
    after.update({"facets": z.n_facets, "edges": z.n_edges,
                  "cell_pairs": z.n_cell_pairs})
    # ---- learned path on the same resident instance (model trained on
    # other seeds of the same generator, untimed)
    learned = None
    if not a.no_learned:
        t0 = time.perf_counter()
        if a.workload == "cube64":
            train = [prep.cube64_instance(sd) for sd in (1, 2)]
            params, thr, t_train, loss = train_model(train, 3, 2048)
            model = "trained on the 64^3 generator, seeds 1-2 (3 epochs)"
        else:
            train = [prep.cube10_instance(sd) for sd in range(1, 65)]
            params, thr, t_train, loss = train_model(train, 6, 1024)
            model = "trained on the 10^3 generator, seeds 1-64 (6 epochs)"
        t_learn_prep = time.perf_counter() - t0
        ctx.mlp_set(params)
        kw = dict(shape, learned=1,
                  logit_threshold=thr)
        ctx.run(**shape)
        ref_res = ctx.download()
        lst = ctx.run(**kw)
        identical = same_result(ctx.download(), ref_res)
        del ref_res
        for _ in range(a.warmup):
            ctx.run(**kw)
        ctx._L.rpd_sync(ctx._h)
        lstats = []
        t0 = time.perf_counter()
        for _ in range(a.steps):
            lstats.append(ctx.run(**kw))
        ctx._L.rpd_sync(ctx._h)
        ldt = time.perf_counter() - t0
        lk = np.mean(np.array([x["ms_kernel"] for x in lstats]), axis=0)
        lp = np.mean(np.array([x["ms_phase"] for x in lstats]), axis=0)
        lst = lstats[-1]
        learned = {
            "value": lst["n_cells"] * a.steps / ldt,
            "ms_per_step": 1e3 * ldt / a.steps,
            "identical_to_plain": identical,
            "fallback_pct": 100.0 * lst["learned"][1] / inst.n_tet,
            "fallback_tets": lst["learned"][1],
            "full_relation_tets": lst["learned"][3],
            "first_round_candidates": lst["learned"][0],
            "added_candidates": lst["learned"][2],
            "model": model, "logit_threshold": thr, "train_loss": loss,
            "train_s": t_train, "train_prep_s": t_learn_prep,
            "host_reads": lst["n_host_reads"],
            "phase_ms": {"candidates": lp[0], "clip": lp[1], "rerun": lp[2],
                         "compaction": lp[3], "certificate": lp[4]},
            "kernel_ms": {"features": lk[0], "mlp": lk[1],
                          "proposal_count": lk[2], "proposal_fill": lk[3],
                          "plane_masks": lk[4], "clip": lk[5],
                          "compaction": lk[6], "certificate": lk[7]},
            "clip_alg_gbs": clip_bytes(lst, inst) / (lk[5] * 1e-3) / 1e9,
        }
        if not a.no_graph:  # the learned step captured (one host read)
            learned["graph"] = bench_graph(ctx, [None] * a.steps, a.warmup,
                                           G, caps, learned=1,
                                           logit_threshold=thr)
    # the regular-triangulation neighbours of the same sphere set: the host
    # builder against the device builder (same exact predicates)
    xyzw = (prep.spheres(50000, 1, 0) if a.workload == "cube64"
            else prep.spheres(500, 0, 0, 0.02, 0.08))
    xyzw[:, 3] **= 2
    t0 = time.perf_counter()
    v_h, o_h, i_h = prep.power_neighbors(xyzw, prep.UNIT_BOX)
    t_host = time.perf_counter() - t0
    neighbours = {"n_spheres": int(xyzw.shape[0]), "host_s": t_host,
                  "host_threads": os.cpu_count()}
    try:
        ctx.power_neighbors(xyzw, prep.UNIT_BOX)  # warm-up
        t0 = time.perf_counter()
        v_d, o_d, i_d, nb_ms = ctx.power_neighbors(xyzw, prep.UNIT_BOX)
        t_dev = time.perf_counter() - t0
        neighbours.update({
            "valid": int((v_d > 0).sum()), "pairs": int(i_d.size),
            "device_clip_ms": nb_ms, "device_total_s": t_dev,
            "identical": bool(np.array_equal(v_h > 0, v_d > 0)
                              and np.array_equal(o_h, o_d)
                              and np.array_equal(i_h, i_d))})
    except RuntimeError as e:  # a cell over the builder's capacities
        neighbours["error"] = str(e)
    # host-buffer boundary: upload + one run + download of the cell stream
    t0 = time.perf_counter()
    ctx.upload(inst)
    ctx.run(**shape)
    res = ctx.download()
    t_pcie = time.perf_counter() - t0
    del res
    if a.calib_copy_gib > 0:
        ctx.calib_copy(int(a.calib_copy_gib * (1 << 30)))
    plain = {"value": value, "ms_per_step": ms_step,
             "host_reads": st["n_host_reads"]}
    if learned is not None:
        value, ms_step = learned["value"], learned["ms_per_step"]
    out = {
        "metric": METRIC,
        "value": value, "unit": "cells/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": ms_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic",
        "config": {"workload": (
            "unit cube 64^3 jittered Kuhn grid, "
            f"{inst.n_tet} tets; 50,000 spheres, r = 0.9*dist*U[0.5,1], "
            f"{inst.n_site} non-hidden sites" if a.workload == "cube64"
            else "unit cube 10^3 Kuhn grid, 500 spheres r ~ U[0.02,0.08]"),
            "n_tet": inst.n_tet, "n_site": inst.n_site,
            "cells_per_wave": G, "small_caps": caps,
            "shape": "autotuned" if tune_ms else "given",
            "path": "learned" if learned is not None else "plain"},
        "plain": plain,
        "learned": learned,
        "learned_vs_plain": (learned["value"] / plain["value"]
                             if learned is not None else None),
        "roofline_hbm": {"achieved": achieved,
                         "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": achieved / PEAK_HBM_GBS, "traffic": traffic,
                         "kernel": clip_kernel,
                         "kernel_ms": clip_ms, "alg_bytes": alg,
                         "measured_on": "the plain step's clip launch"},
        "roofline_fp64_valu": fp64,
        "tune_ms": ({f"caps{c}_G{g}": v for (c, g), v in tune_ms.items()}
                    if tune_ms else None),
        "tets_per_s": inst.n_tet * 1e3 / ms_step,
        "fallback_pct": learned["fallback_pct"] if learned else 0.0,
        "exact_rerun_cells": st["n_rerun_exact"],
        "capacity_rerun_cells": st["n_rerun_large"],
        "n_cells": n_cells, "n_candidates": st["n_candidates"],
        "n_words": st["n_words"], "n_areas": st["n_areas"],
        "active_plane_frac": active_frac,
        "phase_ms": {"candidates": ph[0], "clip": ph[1], "rerun": ph[2],
                     "compaction": ph[3]},
        "kernel_ms": {"site_boxes": kn[0], "binning": kn[1],
                      "relation_count": kn[2], "relation_fill": kn[3],
                      "plane_masks": kn[4], "clip": kn[5],
                      "compaction": kn[6]},
        "host_prep_s": t_prep, "upload_s": t_upload,
        "neighbour_builder": neighbours,
        "after_clip": after,
        "pcie_inclusive": {"value": n_cells / t_pcie, "unit": "cells/s",
                           "s": t_pcie, "note": "upload of host buffers + "
                           "one run + download of cells, words, geometry "
                           "and areas; not the metric"},
    }
    ctx.close()
    if not a.no_cpu and emit:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        out["cpu_baseline"] = cpu_baseline([inst], a.cpu_sample_tets,
                                           threads, "the same instance")
    if emit:
        print(json.dumps(out))
    return out


def headline(a):
    """Config 5 learned+batched (the metric's configuration), with the
    cube64 single instance and config 3 (8 sphere-update iterations) as
    sub-keys and the CPU baseline on a sample of config 5."""
    ns = vars(a)
    out = bench_batched(argparse.Namespace(**{**ns, "workload": "gmm"}),
                        emit=False)
    insts = out.pop("_instances")
    out["config"]["workload"] = ("config 5 (learned+batched under stress): "
                                 + out["config"]["workload"])
    if not a.no_cube64:
        c = bench_cube(argparse.Namespace(**{**ns, "workload": "cube64",
                                             "steps": 20}), emit=False)
        out["cube64"] = {k: c[k] for k in (
            "value", "ms_per_step", "plain", "learned", "learned_vs_plain",
            "fallback_pct", "roofline_hbm", "roofline_fp64_valu", "n_cells",
            "n_candidates", "kernel_ms", "phase_ms", "neighbour_builder",
            "after_clip", "pcie_inclusive")}
        out["cube64"]["config"] = c["config"]
        print("[headline] cube64 done", file=sys.stderr, flush=True)
    if not a.no_config3:
        b = bench_batched(argparse.Namespace(**{
            **ns, "workload": "batched", "batch_iterations": 8}), emit=False)
        b.pop("_instances")
        out["config3"] = {k: b[k] for k in (
            "value", "ms_per_step", "config", "plain", "learned",
            "learned_vs_plain_batched", "learned_vs_plain_one_at_a_time",
            "fallback_pct", "identical_to_plain", "roofline")}
        print("[headline] config 3 done", file=sys.stderr, flush=True)
    if not a.no_cpu:
        threads = a.cpu_threads or min(16, os.cpu_count() or 1)
        # every 8th instance of the batch, the sample spread over them
        out["cpu_baseline"] = cpu_baseline(
            insts[::8], a.cpu_sample_tets, threads,
            f"every 8th instance of the config-5 batch "
            f"({len(insts[::8])} instances)")
    print(json.dumps(out))


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "bench.py runs on one GPU only "
                          "(multi-GPU sharding is a design note)",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    if a.workload == "headline":
        return headline(a)
    if a.workload in ("batched", "gmm"):
        return bench_batched(a)
    if a.workload == "heldout10":
        return bench_heldout10(a)
    return bench_cube(a)


if __name__ == "__main__":
    main()
