"""TopoSZp benchmark on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--bs 64] [--shape 512 512 512]

One step = the whole topology pipeline over one 512^3 float32 field resident in HBM:
critical points + per-bin sort positions + sort-position stream + topology stream
(compress), then the topology stream -> field + types, the sort-position stream
-> sort positions and the topology restoration of the field (decompress, as the
reference's decompression timing includes its post-process).  value = 4 N bytes of original data per step / step
time (the reference's GB/s convention, example/test_cuda_compress.cu:95-96).

Also reported: per-stage kernel times (HIP events on the launch stream), the roofline
of the dominant kernel, plain SZp throughput on the same field, the threadblock codec
(float and double data), the double random-access codec and the double 3D topology
path on the same field, a batch of 256 mixed-size fields through the batched plain
codec against a loop of one-field calls, and the reference's own OpenMP CPU path timed
on this host on a bounded slab of the same field.
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "toposzp-cuda_amd"))

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


CPU_SCRIPT = r"""
import ctypes as C, json, os, sys, time
import numpy as np
vp, sz, i32, f32 = C.c_void_p, C.c_size_t, C.c_int, C.c_float
R = C.CDLL(sys.argv[1])
libc = C.CDLL("libc.so.6"); libc.free.argtypes = [vp]
d1, d2, d3, eb, bs = int(sys.argv[3]), int(sys.argv[4]), int(sys.argv[5]), float(sys.argv[6]), int(sys.argv[7])
data = np.fromfile(sys.argv[2], dtype=np.float32)
n = data.size
P = lambda a: a.ctypes.data_as(vp)
R.szp_find_critical_points_3d.argtypes = [vp, C.POINTER(sz), i32, i32, i32, f32]; R.szp_find_critical_points_3d.restype = vp
R.szp_sort_critical_points_3d.argtypes = [vp, sz, vp, i32, i32]
R.szp_compress_sort_positions_3d.argtypes = [vp, sz, C.POINTER(sz), i32]; R.szp_compress_sort_positions_3d.restype = vp
R.szp_float_compress_topology_3d.argtypes = [vp, C.POINTER(sz), f32, sz, i32, vp, i32, i32, i32, i32]
R.szp_float_compress_topology_3d.restype = vp
R.szp_float_decompress_openmp_threadblock_randomaccess_topology_preserved.argtypes = [C.POINTER(vp), sz, f32, i32, vp, C.POINTER(vp)]
R.szp_decompress_sort_positions.argtypes = [vp, sz, i32]; R.szp_decompress_sort_positions.restype = vp
t0 = time.perf_counter()
cnt = sz()
cps = R.szp_find_critical_points_3d(P(data), C.byref(cnt), d1, d2, d3, eb)
R.szp_sort_critical_points_3d(cps, cnt.value, P(data), d2, d3)
spn = sz()
sps = R.szp_compress_sort_positions_3d(cps, cnt.value, C.byref(spn), bs)
tn = sz()
topo = R.szp_float_compress_topology_3d(P(data), C.byref(tn), eb, n, bs, cps, cnt.value, d1, d2, d3)
t1 = time.perf_counter()
out, fn = vp(), vp()
R.szp_float_decompress_openmp_threadblock_randomaccess_topology_preserved(C.byref(out), n, eb, bs, topo, C.byref(fn))
ext = C.c_uint64.from_address(sps).value
spd = R.szp_decompress_sort_positions(sps + 8, ext, bs)
t2 = time.perf_counter()
if len(sys.argv) > 8:  # the streams and the decoded values, for the checks against the GPU
    pre = sys.argv[8]
    open(pre + ".topo", "wb").write(C.string_at(topo, tn.value))
    open(pre + ".sp", "wb").write(C.string_at(sps, spn.value))
    np.ctypeslib.as_array(C.cast(out, C.POINTER(C.c_float)), (n,)).tofile(pre + ".dec")
print(json.dumps({"compress_s": t1 - t0, "decompress_s": t2 - t1, "n": n, "cps": cnt.value,
                  "topo_bytes": tn.value, "sp_bytes": spn.value}))
"""


def cpu_baseline(field_np, eb, bs, threads, slab, dump=None):
    """The reference's OpenMP CPU path (oracle/_ref/libszp_ref.so, built from the
    reference sources) on a bounded slab, in a child process; with dump, a file prefix
    for its streams and decoded values."""
    import numpy as np

    lib = os.path.join(ROOT, "oracle", "_ref", "libszp_ref.so")
    if not os.path.exists(lib):
        return None, "oracle/_ref/libszp_ref.so not built"
    sub = np.ascontiguousarray(field_np[:slab])
    path = os.path.join("/tmp", f"tszp_cpu_slab_{os.getpid()}.f32")
    sub.tofile(path)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        r = subprocess.run([sys.executable, "-c", CPU_SCRIPT, lib, path, *map(str, sub.shape), repr(eb),
                            str(bs)] + ([dump] if dump else []), capture_output=True, text=True, timeout=300,
                           env=env)
    finally:
        os.unlink(path)
    if r.returncode != 0:
        return None, f"reference CPU run failed rc={r.returncode}: {r.stderr[-300:]}"
    res = json.loads(r.stdout.strip().splitlines()[-1])
    res["shape"] = list(sub.shape)
    return res, None


def physical_cores():
    """(physical cores of the host from /proc/cpuinfo, CPUs this process may run on)."""
    pairs = set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            if line.startswith("physical id"):
                phys = line.split(":")[1].strip()
            elif line.startswith("core id"):
                core = line.split(":")[1].strip()
            elif not line.strip() and core is not None:
                pairs.add((phys, core))
                phys = core = None
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return (len(pairs) or None), allowed


def driver_2d(field_np, eb, bs, threads):
    """The reference's own 2D CPU drivers (oracle/_ref/ref_topo_{compress,decompress},
    built from example/testfloat_{compress,decompress}_fastmode1_topology.cc) on a 2D
    field: their compression window (CP find, sort, sort-position and topology
    compression) and decompression window (decoders and the driver's post-process), as
    they print them."""
    exe_c = os.path.join(ROOT, "oracle", "_ref", "ref_topo_compress")
    exe_d = os.path.join(ROOT, "oracle", "_ref", "ref_topo_decompress")
    if not (os.path.exists(exe_c) and os.path.exists(exe_d)):
        return None
    rows, cols = field_np.shape
    path = os.path.join("/tmp", f"tszp_cpu_2d_{os.getpid()}.f32")
    field_np.astype("<f4").tofile(path)
    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    out = {}
    try:
        rc = subprocess.run([exe_c, path, str(bs), repr(eb), str(rows), str(cols)], capture_output=True, text=True,
                            timeout=300, env=env, cwd="/tmp")
        rd = subprocess.run([exe_d, path + ".SZp", str(rows * cols), str(bs), repr(eb), str(rows), str(cols)],
                            capture_output=True, text=True, timeout=300, env=env, cwd="/tmp")
        for line in (rc.stdout + rd.stdout).splitlines():
            if line.startswith("compression time="):
                out["compress_s"] = float(line.split("=")[1].split(",")[0])
            elif line.startswith("decompression time="):
                out["decompress_s"] = float(line.split("=")[1])
            elif line.startswith("Max absolute error"):
                out["max_err_over_eb"] = float(line.split("=")[1]) / eb
            elif line.startswith("compressionRatio"):
                out["ratio"] = float(line.split("=")[1])
    finally:
        for suf in ("", ".SZp", ".SZp.out"):
            if os.path.exists(path + suf):
                os.unlink(path + suf)
    if "compress_s" not in out or "decompress_s" not in out:
        return None
    out["gbs"] = 4.0 * rows * cols / (out["compress_s"] + out["decompress_s"]) / 1e9
    return out


def codec_kernel(mode, bs, ndim, fast):
    """Name of the compression kernel the library launches for this block size: the
    16-byte path for block sizes 32 and 64 (classification also needs the fastest axis a
    multiple of 4), else the generic kernel with its lane geometry."""
    if bs in (32, 64) and (mode != 0 or fast % 4 == 0):
        return f"k_compress_v4<{mode}, {bs // 4}, {ndim if mode == 0 else 2}>"
    if bs <= 16:
        s, e = 1, 1
        while s < bs:
            s <<= 1
    elif bs <= 64:
        s, e = 1, 4
        while 4 * s < bs:
            s <<= 1
    else:
        s, e = 64, 1
        while 64 * e < bs:
            e <<= 1
    return f"k_compress<{mode}, {s}, {e}>"


def pmc_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed FETCH_SIZE / WRITE_SIZE summary
    (already scaled by the copy-kernel calibration there), averaged over its dispatches;
    None when the summary does not list it."""
    path = os.path.join(ROOT, "profiles", "topo_512cube_bs64_pmc_traffic.md")
    if not os.path.exists(path):
        return None
    vals = []
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and kernel in cells[0] and cells[4].isdigit():
            vals.append(int(cells[4]))
    return sum(vals) / len(vals) if vals else None


def elem_side(tz, torch, f, shape, eb, bs, steps):
    """Side figures of the element-partitioned codecs and the double 3D topology path on
    the same field (as float and as double data): whole-call GB/s of input and the HBM
    fraction of the compression / decompression kernels from the library's event timers
    (algorithmic bytes = input + stream)."""
    n = f.numel()
    f64 = f.double()
    out = {}
    cases = [("threadblock_f32", f, False), ("threadblock_f64", f64, False), ("plain_f64", f64, True)]
    for name, x, ra in cases:
        ds = x.element_size()
        sbuf, blk = tz.elem_buffers(n, bs, 1, ds, ra, x.device)
        dec = torch.empty_like(x)

        def run():
            if ra:
                nb, _, _ = tz.plain_compress_f64(x, eb, bs, 1, out=sbuf, blk=blk)
                tz.plain_decompress_f64(sbuf, blk, n, eb, bs, 1, out=dec)
            else:
                nb, _, _ = tz.threadblock_compress(x, eb, bs, 1, out=sbuf, blk=blk)
                tz.threadblock_decompress(sbuf, blk, n, eb, bs, 1, out=dec)
            return nb
        nb = run()
        torch.cuda.synchronize()
        err = float((dec - x).abs().max()) / eb
        tz.prof_reset()
        tz.prof_enable(True)
        t0 = time.perf_counter()
        for _ in range(steps):
            run()
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / steps
        tz.prof_enable(False)
        pr = tz.prof_read()
        cms = pr["elem"][0] / max(1, pr["elem"][1])
        dms = pr["elem_decode"][0] / max(1, pr["elem_decode"][1])
        out[name] = {"gbs": ds * n / el / 1e9, "ratio": ds * n / nb, "max_err_over_eb": err,
                     "compress_kernel_ms": cms, "decompress_kernel_ms": dms,
                     "compress_hbm_frac": (ds * n + nb) / (cms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "decompress_hbm_frac": (ds * n + nb) / (dms * 1e-3) / 1e9 / HBM_PEAK_GBS}
        log(f"{name}: compress+decompress {out[name]['gbs']:.1f} GB/s, ratio {out[name]['ratio']:.2f}, "
            f"kernels {cms:.3f} / {dms:.3f} ms ({out[name]['compress_hbm_frac']:.2f} / "
            f"{out[name]['decompress_hbm_frac']:.2f} of HBM peak), max err {err:.4f} eb")
        # the same stream without its side-car: unit offsets parsed on the device
        run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            pblk, pT = tz.elem_parse(sbuf, n, bs, ds, ra, length=nb)
        torch.cuda.synchronize()
        out[name]["parse_ms"] = 1e3 * (time.perf_counter() - t0) / steps
        out[name]["parse_matches_side_car"] = bool(pT == 1 and torch.equal(pblk, blk))
        log(f"{name}: device parse of the stream without its side-car {out[name]['parse_ms']:.3f} ms, "
            f"offsets equal the side-car: {out[name]['parse_matches_side_car']}")
        del sbuf, blk, dec
    # double 3D topology: critical-point records of the field, then the whole call
    types, _ = tz.classify(f.view(shape))
    idx = torch.nonzero(types.view(-1)).view(-1)
    plane = shape[1] * shape[2]
    rec = torch.stack([idx // plane, (idx % plane) // shape[2], idx % shape[2],
                       types.view(-1)[idx].long(), torch.zeros_like(idx), torch.zeros_like(idx)], 1)
    rec = rec.to(torch.int32).contiguous()
    x = f64.view(shape)
    nb, _, _, _, moved = tz.topo_compress_f64_3d(x, eb, bs, rec)
    torch.cuda.synchronize()
    tz.prof_reset()
    tz.prof_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        nb, _, _, _, moved = tz.topo_compress_f64_3d(x, eb, bs, rec)
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / steps
    tz.prof_enable(False)
    pr = tz.prof_read()
    out["topo_f64_3d"] = {"gbs": 8 * n / el / 1e9, "ms": 1e3 * el, "ratio": 8 * n / nb, "saddles_moved": moved,
                          "cp_count": int(idx.numel()),
                          "quantize_saddle_fix_ms": pr["saddle"][0] / max(1, pr["saddle"][1]),
                          "codec_ms": pr["codec_topo"][0] / max(1, pr["codec_topo"][1])}
    log(f"topo_f64_3d: {out['topo_f64_3d']['ms']:.3f} ms ({out['topo_f64_3d']['gbs']:.1f} GB/s of doubles), "
        f"quantize + saddle fix {out['topo_f64_3d']['quantize_saddle_fix_ms']:.3f} ms, "
        f"codec {out['topo_f64_3d']['codec_ms']:.3f} ms, {moved} saddles moved of {idx.numel()} CPs")
    return out


def batch_side(tz, torch, seed, reps):
    """Mixed-size workload: 256 float32 fields, half 2D (rows and cols log-uniform in
    [64, 2048], blobs of sigma 20-200 + 0.01 noise), half 3D (each dim log-uniform in
    [32, 256], blobs of sigma 5-60 + sine waves + 1e-3 noise), absolute error bound
    log-uniform in [1e-4, 1e-2] per field, block sizes from the library's autotuner.
    Batched compress + decompress (one launch per block-size class and stage) against a
    loop of the one-field device calls on the same fields and block sizes; GB/s of input,
    latency percentiles over the timed repetitions."""
    import numpy as np
    rng = np.random.default_rng(seed)
    fields, ebs = [], []
    for i in range(256):
        if i % 2 == 0:
            shape = [int(round(v)) for v in np.exp(rng.uniform(np.log(64), np.log(2048), 2))]
            f = tz.gen_blobs(shape, seed + i, 64, 20.0, 200.0, 0.01)
        else:
            shape = [int(round(v)) for v in np.exp(rng.uniform(np.log(32), np.log(256), 3))]
            f = tz.gen_blobs(shape, seed + i, 128, 5.0, 60.0, 1e-3, waves=True)
        fields.append(f.view(-1))
        ebs.append(float(np.exp(rng.uniform(np.log(1e-4), np.log(1e-2)))))
    total = sum(f.numel() for f in fields)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bss = tz.batch_tune(fields, ebs)
    tune_ms = 1e3 * (time.perf_counter() - t0)
    sizes, streams, blks = tz.plain_compress_batch(fields, ebs, bss)
    outs = tz.plain_decompress_batch(streams, blks, [f.numel() for f in fields], ebs, bss)
    lstreams = [torch.empty_like(s) for s in streams]
    lblks = [torch.empty_like(b) for b in blks]
    louts = [torch.empty_like(o) for o in outs]

    def loop():
        for i, f in enumerate(fields):
            tz.plain_compress(f, ebs[i], bss[i], 1, out=lstreams[i], blk=lblks[i])
            tz.plain_decompress(lstreams[i], lblks[i], f.numel(), ebs[i], bss[i], 1, out=louts[i])
    loop()
    torch.cuda.synchronize()
    same = all(torch.equal(streams[i][:sizes[i]], lstreams[i][:sizes[i]]) and torch.equal(outs[i], louts[i])
               for i in range(len(fields)))
    err = max(float((outs[i] - fields[i]).abs().max()) / ebs[i] for i in range(len(fields)))
    tz.prof_reset()
    tz.prof_enable(True)
    bt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        tz.plain_compress_batch(fields, ebs, bss, streams, blks)
        tz.plain_decompress_batch(streams, blks, [f.numel() for f in fields], ebs, bss, outs)
        torch.cuda.synchronize()
        bt.append(time.perf_counter() - t0)
    tz.prof_enable(False)
    pr = tz.prof_read()
    lt = []
    for _ in range(max(2, reps // 4)):
        t0 = time.perf_counter()
        loop()
        torch.cuda.synchronize()
        lt.append(time.perf_counter() - t0)
    # the same batch as a device-resident plan, compress + decompress captured in a graph
    pstreams = [torch.empty_like(x) for x in streams]
    pblks = [torch.empty_like(x) for x in blks]
    pouts = [torch.empty_like(x) for x in outs]
    plan = tz.BatchPlan(fields, ebs, bss, pstreams, pblks, pouts)
    plan.compress()
    plan.decompress()
    plan.results()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        plan.compress()
        plan.decompress()
    graph.replay()
    torch.cuda.synchronize()
    gt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        graph.replay()
        torch.cuda.synchronize()
        gt.append(time.perf_counter() - t0)
    psizes = plan.results()
    gsame = psizes == sizes and all(torch.equal(pstreams[i][:sizes[i]], streams[i][:sizes[i]]) and
                                    torch.equal(pouts[i], outs[i]) for i in range(len(fields)))
    plan.close()
    del graph
    cbytes = sum(sizes)
    bt, lt, gt = np.array(bt), np.array(lt), np.array(gt)
    ck = pr["batch"][0] / max(1, reps)
    dk = pr["batch_decode"][0] / max(1, reps)
    res = {"fields": len(fields), "elements": total, "ratio": 4 * total / cbytes,
           "block_sizes": {str(b): bss.count(b) for b in sorted(set(bss))}, "tune_ms": tune_ms,
           "batch_gbs": 4 * total / float(np.median(bt)) / 1e9,
           "batch_p50_ms": 1e3 * float(np.percentile(bt, 50)), "batch_p99_ms": 1e3 * float(np.percentile(bt, 99)),
           "loop_gbs": 4 * total / float(np.median(lt)) / 1e9,
           "loop_p50_ms": 1e3 * float(np.percentile(lt, 50)), "loop_p99_ms": 1e3 * float(np.percentile(lt, 99)),
           "compress_kernels_ms": ck, "decompress_kernels_ms": dk,
           "compress_hbm_frac": (4 * total + cbytes) / (ck * 1e-3) / 1e9 / HBM_PEAK_GBS if ck else None,
           "decompress_hbm_frac": (4 * total + cbytes) / (dk * 1e-3) / 1e9 / HBM_PEAK_GBS if dk else None,
           "identical_to_loop": bool(same), "max_err_over_eb": err,
           "graph_gbs": 4 * total / float(np.median(gt)) / 1e9, "graph_p50_ms": 1e3 * float(np.percentile(gt, 50)),
           "graph_p99_ms": 1e3 * float(np.percentile(gt, 99)), "graph_identical": bool(gsame)}
    res["speedup_vs_loop"] = res["batch_gbs"] / res["loop_gbs"]
    res["_pool"] = {"fields": fields, "ebs": ebs, "bss": bss}
    log(f"batch of {len(fields)} fields ({total / 1e6:.1f} M elements, ratio {res['ratio']:.2f}, block sizes "
        f"{res['block_sizes']}): batched {res['batch_gbs']:.1f} GB/s (p50 {res['batch_p50_ms']:.2f} ms, p99 "
        f"{res['batch_p99_ms']:.2f} ms; kernels {ck:.3f} + {dk:.3f} ms), loop of one-field calls "
        f"{res['loop_gbs']:.1f} GB/s (p50 {res['loop_p50_ms']:.2f} ms) -> {res['speedup_vs_loop']:.1f}x, "
        f"identical {same}, max err {err:.4f} eb, tune {tune_ms:.2f} ms; plan replayed as a graph "
        f"{res['graph_gbs']:.1f} GB/s (p50 {res['graph_p50_ms']:.2f} ms), identical {gsame}")
    return res


def topo_batch_side(tz, torch, seed, reps):
    """Mixed-size workload through the batched topology pipeline (tszp_topo_batch_*):
    compress + decompress + restoration of all 256 fields per step, one launch per stage
    and block-size class, against a loop of the one-field calls (tszp_topo_compress,
    tszp_topo_decompress, tszp_sortpos_decompress, tszp_topo_restore) on the same fields
    and block sizes.  GB/s of input per step; latency percentiles over the repetitions;
    the compress + decompress part also replayed from a captured graph."""
    import numpy as np
    import workloads
    wl = workloads.mixed_fields(seed)
    ts = [f for f, _ in wl]
    ebs = [eb for _, eb in wl]
    total = sum(t.numel() for t in ts)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    batch = tz.TopoBatch(ts, ebs, None)
    create_ms = 1e3 * (time.perf_counter() - t0)
    bss = batch.bss

    def step():
        batch.compress()
        batch.decompress()
        batch.restore()
        return batch.results()
    res = step()
    mism = sum(r.mismatches() for _, r in res)
    maxdev = max(r.max_dev_over_eb for _, r in res)
    tz.prof_reset()
    tz.prof_enable(True)
    bt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        step()
        bt.append(time.perf_counter() - t0)
    tz.prof_enable(False)
    pr = tz.prof_read()
    # compress + decompress from a graph
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.stream(s):
        with torch.cuda.graph(g, stream=s):
            batch.compress()
            batch.decompress()
    torch.cuda.current_stream().wait_stream(s)
    g.replay()
    torch.cuda.synchronize()
    gt = []
    for _ in range(reps):
        t0 = time.perf_counter()
        g.replay()
        torch.cuda.synchronize()
        gt.append(time.perf_counter() - t0)
    gres = batch.results()
    gsame = all(a[0].topo_bytes == b[0].topo_bytes and a[0].sortpos_bytes == b[0].sortpos_bytes
                for a, b in zip(gres, res))
    del g
    # the loop of one-field calls
    bufs = [tz.TopoBuffers(f.numel(), b, 1, f.device) for f, b in zip(ts, bss)]
    outs = [torch.empty(f.numel(), dtype=torch.float32, device=f.device) for f in ts]

    def loop():
        infos = []
        for i, f in enumerate(ts):
            info, _ = tz.topo_compress(f, ebs[i], bss[i], T=1, buf=bufs[i])
            out, types = tz.topo_decompress(bufs[i], f.numel(), ebs[i], bss[i], T=1)
            sp = tz.sortpos_decompress(bufs[i], info.ext_count, bss[i], T=1) if info.ext_count else None
            r = tz.topo_restore(out, types, tuple(f.shape), ebs[i], sp=sp, n_ext=info.ext_count, out=outs[i])
            infos.append((info, r))
        return infos
    linfo = loop()
    torch.cuda.synchronize()
    same = all(a[0].topo_bytes == b[0].topo_bytes and a[0].sortpos_bytes == b[0].sortpos_bytes and
               torch.equal(bufs[i].topo[:a[0].topo_bytes], batch.streams[i][0][:a[0].topo_bytes]) and
               torch.equal(bufs[i].sp[:a[0].sortpos_bytes], batch.streams[i][2][:a[0].sortpos_bytes])
               for i, (a, b) in enumerate(zip(res, linfo)))
    lt = []
    for _ in range(max(2, reps // 4)):
        t0 = time.perf_counter()
        loop()
        torch.cuda.synchronize()
        lt.append(time.perf_counter() - t0)
    batch.close()
    del bufs, outs
    cbytes = sum(r[0].topo_bytes + r[0].sortpos_bytes for r in res)
    bt, lt, gt = np.array(bt), np.array(lt), np.array(gt)
    per = lambda k: pr[k][0] / max(1, reps)
    out = {"fields": len(ts), "elements": total, "ratio": 4 * total / cbytes,
           "block_sizes": {str(b): bss.count(b) for b in sorted(set(bss))}, "create_ms": create_ms,
           "batch_gbs": 4 * total / float(np.median(bt)) / 1e9,
           "batch_p50_ms": 1e3 * float(np.percentile(bt, 50)), "batch_p99_ms": 1e3 * float(np.percentile(bt, 99)),
           "loop_gbs": 4 * total / float(np.median(lt)) / 1e9,
           "loop_p50_ms": 1e3 * float(np.percentile(lt, 50)), "loop_p99_ms": 1e3 * float(np.percentile(lt, 99)),
           "compress_ms": per("topo_batch"), "sort_positions_ms": per("topo_batch_sort"),
           "decompress_ms": per("topo_batch_decode"), "restore_ms": per("topo_batch_restore"),
           "graph_compress_decompress_p50_ms": 1e3 * float(np.percentile(gt, 50)),
           "graph_compress_decompress_gbs": 4 * total / float(np.median(gt)) / 1e9, "graph_identical": bool(gsame),
           "identical_to_loop": bool(same), "mismatches": int(mism), "max_dev_over_eb": float(maxdev)}
    out["speedup_vs_loop"] = out["batch_gbs"] / out["loop_gbs"]
    log(f"topology batch of {len(ts)} fields ({total / 1e6:.1f} M elements, ratio {out['ratio']:.2f}, block sizes "
        f"{out['block_sizes']}): compress+decompress+restore {out['batch_gbs']:.1f} GB/s (p50 {out['batch_p50_ms']:.2f}"
        f" ms, p99 {out['batch_p99_ms']:.2f} ms; stages {out['compress_ms']:.2f} (sort {out['sort_positions_ms']:.2f})"
        f" + {out['decompress_ms']:.2f} + {out['restore_ms']:.2f} ms), loop of one-field calls {out['loop_gbs']:.1f} "
        f"GB/s (p50 {out['loop_p50_ms']:.1f} ms) -> {out['speedup_vs_loop']:.1f}x, identical {same}, mismatches "
        f"{mism}; compress+decompress graph {out['graph_compress_decompress_p50_ms']:.2f} ms")
    return out


def serve_side(tz, torch, pool, big, seed, nreq, rate):
    """Served stream: nreq requests with Poisson arrivals at `rate` per second, 70 %
    compressions of a pool field (the 256 mixed-size fields, or with probability 2 % the
    512^3 field) and 30 % decompressions of a field compressed earlier, submitted to
    the request loop (admission by a 200 GB HBM budget, continuous batching) and,
    separately, run one request at a time with the one-field calls at the same arrival
    times.  Latency = submission (or arrival) to completion."""
    import threading
    import numpy as np
    rng = np.random.default_rng(seed)
    fields = pool["fields"] + [big["field"]]
    ebs = pool["ebs"] + [big["eb"]]
    bss = pool["bss"] + [big["bs"]]
    nbig = len(fields) - 1
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    kinds, which = [], []
    for _ in range(nreq):
        if rng.random() < 0.7 or not kinds:
            kinds.append(0)
            which.append(nbig if rng.random() < 0.02 else int(rng.integers(0, nbig)))
        else:
            kinds.append(1)
            done = [w for k, w in zip(kinds, which) if k == 0]
            which.append(done[int(rng.integers(0, len(done)))])
    outs = [torch.empty_like(f) for f in fields]
    nbytes = sum(4 * fields[w].numel() for w in which)
    budget = 200 << 30
    try:
        srv = tz.Server(budget, max_batch=256)
    except tz.TszpError:
        budget = 64 << 30
        srv = tz.Server(budget, max_batch=256)
    lock = threading.Lock()
    latest = {}                 # field -> its latest finished compress ticket
    tickets = [None] * nreq
    ready = threading.Semaphore(0)
    lat = np.zeros(nreq)

    def waiter():
        for i in range(nreq):
            ready.acquire()
            info = srv.wait(tickets[i])
            lat[i] = info.latency_us
            with lock:
                if kinds[i] == 0:
                    old = latest.get(which[i])
                    latest[which[i]] = tickets[i]
                    if old is not None:
                        srv.release(old)
                else:
                    srv.release(tickets[i])

    torch.cuda.synchronize()
    wt = threading.Thread(target=waiter)
    wt.start()
    t0 = time.perf_counter()
    for i in range(nreq):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        with lock:
            if kinds[i] == 0:
                tickets[i] = srv.compress(fields[which[i]], ebs[which[i]], bss[which[i]])
            else:
                src = latest.get(which[i])
                if src is None:  # its compression has not finished yet: compress again
                    kinds[i] = 0
                    tickets[i] = srv.compress(fields[which[i]], ebs[which[i]], bss[which[i]])
                else:
                    tickets[i] = srv.decompress(src, outs[which[i]])
        ready.release()
    wt.join()
    served_s = time.perf_counter() - t0
    st = srv.stats()
    with lock:
        for t in latest.values():
            srv.release(t)
    srv.close()

    # the same requests one at a time with the one-field calls
    streams = {}
    lat1 = np.zeros(nreq)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    free_at = 0.0
    for i in range(nreq):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        w = which[i]
        if kinds[i] == 0 or w not in streams:
            nb, sbuf, blk = tz.plain_compress(fields[w], ebs[w], bss[w], 1)
            streams[w] = (sbuf, blk)
        else:
            sbuf, blk = streams[w]
            tz.plain_decompress(sbuf, blk, fields[w].numel(), ebs[w], bss[w], 1, out=outs[w])
        torch.cuda.synchronize()
        lat1[i] = 1e6 * (time.perf_counter() - t0 - arrivals[i])
    loop_s = time.perf_counter() - t0
    res = {"requests": nreq, "rate_per_s": rate, "compress_share": float(np.mean(np.array(kinds) == 0)),
           "big_fields": int(sum(1 for k, w in zip(kinds, which) if w == nbig)), "input_gb": nbytes / 1e9,
           "hbm_budget_gb": budget / 2 ** 30, "hbm_peak_gb": st["hbm_peak"] / 2 ** 30, "batches": st["batches"],
           "served_gbs": nbytes / served_s / 1e9, "served_p50_ms": float(np.percentile(lat, 50)) / 1e3,
           "served_p99_ms": float(np.percentile(lat, 99)) / 1e3,
           "loop_gbs": nbytes / loop_s / 1e9, "loop_p50_ms": float(np.percentile(lat1, 50)) / 1e3,
           "loop_p99_ms": float(np.percentile(lat1, 99)) / 1e3}
    log(f"served {nreq} requests at {rate:.0f}/s ({nbytes / 1e9:.1f} GB, {st['batches']} batches, HBM peak "
        f"{res['hbm_peak_gb']:.2f} GiB): {res['served_gbs']:.1f} GB/s, p50 {res['served_p50_ms']:.2f} ms, "
        f"p99 {res['served_p99_ms']:.2f} ms; one request at a time: {res['loop_gbs']:.1f} GB/s, "
        f"p50 {res['loop_p50_ms']:.2f} ms, p99 {res['loop_p99_ms']:.2f} ms")
    return res


def dropin_side(tz, f, eb, bs):
    """The host-pointer drop-in chain on the 512^3 field, timed like the reference's
    drivers time their library calls (host buffers in and out, so PCIe transfers and the
    host-side walks are inside): szp_cuda_find_critical_points_3d ->
    szp_cuda_sort_critical_points_3d -> szp_cuda_compress_sort_positions_3d ->
    szp_cuda_float_compress_topology_3d, then
    szp_cuda_float_decompress_randomaccess_topology_preserved (values and int types) and
    szp_cuda_decompress_sort_positions.  Library-allocated results are freed untouched."""
    import ctypes as C
    import numpy as np
    L = tz.lib()
    libc = C.CDLL("libc.so.6")
    libc.free.argtypes = [C.c_void_p]
    x = np.ascontiguousarray(f.cpu().numpy())
    n = x.size
    d1, d2, d3 = x.shape
    P = x.ctypes.data_as(C.c_void_p)
    res = {}
    for rep in range(2):  # the first run warms the library's device workspace
        t0 = time.perf_counter()
        cnt = C.c_size_t()
        cps = L.szp_cuda_find_critical_points_3d(P, C.byref(cnt), d1, d2, d3, eb)
        L.szp_cuda_sort_critical_points_3d(C.c_void_p(cps), cnt.value, P, d2, d3)
        spn = C.c_size_t()
        sps = L.szp_cuda_compress_sort_positions_3d(C.c_void_p(cps), cnt.value, C.byref(spn), bs)
        tn = C.c_size_t()
        topo = L.szp_cuda_float_compress_topology_3d(P, C.byref(tn), eb, n, bs, C.c_void_p(cps), cnt.value, d1, d2,
                                                      d3)
        t1 = time.perf_counter()
        ext = int(np.count_nonzero(np.isin(np.ctypeslib.as_array(
            C.cast(cps, C.POINTER(C.c_int32)), (cnt.value * 6,))[3::6], (1, 2))))
        t2 = time.perf_counter()
        dp, fp = C.c_void_p(), C.c_void_p()
        L.szp_cuda_float_decompress_randomaccess_topology_preserved(C.byref(dp), n, eb, bs, C.c_void_p(topo),
                                                                    C.byref(fp))
        spd = L.szp_cuda_decompress_sort_positions(C.c_void_p(sps), ext, bs)  # [u64 0][blocks]
        t3 = time.perf_counter()
        ok = bool(dp.value and fp.value and (spd or ext == 0))
        for q in (cps, sps, topo, dp.value, fp.value, spd):
            if q:
                libc.free(C.c_void_p(q))
        res = {"compress_s": t1 - t0, "decompress_s": t3 - t2, "ok": ok, "cp_count": cnt.value,
               "stream_bytes": tn.value + spn.value}
    res["gbs"] = 4.0 * n / (res["compress_s"] + res["decompress_s"]) / 1e9
    log(f"host-pointer drop-in chain on {tuple(x.shape)}: compress {res['compress_s'] * 1e3:.1f} ms, decompress "
        f"{res['decompress_s'] * 1e3:.1f} ms -> {res['gbs']:.2f} GB/s (PCIe and host walks included)")
    return res


def serve_topo_side(tz, torch, big, seed, nreq):
    """Served topology stream: nreq requests, 70 % compressions of a field of the
    mixed-size workload (or with probability 2 % the 512^3 field) and 30 % decompressions
    with the restoration of a field compressed earlier, through the topology request loop
    (two lanes, a 200 GB budget).  Capacity: all requests submitted at once; latency:
    Poisson arrivals at 70 % of that capacity; beside them the same requests one at a
    time with the one-field calls (tszp_topo_compress; tszp_topo_decompress +
    tszp_sortpos_decompress + tszp_topo_restore) at the same arrival times.  GB/s of
    field bytes (input of a compression, output of a decompression)."""
    import threading
    import numpy as np
    import workloads
    rng = np.random.default_rng(seed + 7)
    wl = workloads.mixed_fields(seed)
    fields = [f for f, _ in wl] + [big["field"]]
    ebs = [eb for _, eb in wl] + [big["eb"]]
    bss = [0] * len(wl) + [big["bs"]]
    nbig = len(fields) - 1
    kinds, which = [], []
    for _ in range(nreq):
        if rng.random() < 0.7 or not kinds:
            kinds.append(0)
            which.append(nbig if rng.random() < 0.02 else int(rng.integers(0, nbig)))
        else:
            kinds.append(1)
            comp = [w for k, w in zip(kinds, which) if k == 0]
            which.append(comp[int(rng.integers(0, len(comp)))])
    nbytes = sum(4 * fields[w].numel() for w in which)
    outs = [torch.empty(f.numel(), dtype=torch.float32, device=f.device) for f in fields]
    budget = 200 << 30
    maxe = max(f.numel() for f in fields) + (1 << 24)
    try:
        srv = tz.TopoServer(budget, max_batch=256, max_batch_elements=maxe, lanes=2)
    except tz.TszpError:
        budget = 96 << 30
        srv = tz.TopoServer(budget, max_batch=256, max_batch_elements=maxe, lanes=2)

    def run(arrivals):
        lock = threading.Lock()
        latest = {}
        tickets = [None] * nreq
        kinds_run = list(kinds)
        ready = threading.Semaphore(0)
        lat = np.zeros(nreq)

        def waiter():
            for i in range(nreq):
                ready.acquire()
                info = srv.wait(tickets[i])
                lat[i] = info.latency_us
                with lock:
                    if kinds_run[i] == 0:
                        old = latest.get(which[i])
                        latest[which[i]] = tickets[i]
                        if old is not None:
                            srv.release(old)
                    else:
                        srv.release(tickets[i])

        torch.cuda.synchronize()
        wt = threading.Thread(target=waiter)
        wt.start()
        t0 = time.perf_counter()
        for i in range(nreq):
            if arrivals is not None:
                while time.perf_counter() - t0 < arrivals[i]:
                    pass
            w = which[i]
            with lock:
                src = latest.get(w) if kinds_run[i] == 1 else None
                if kinds_run[i] == 1 and src is None:
                    kinds_run[i] = 0  # its compression has not finished yet: compress again
                if kinds_run[i] == 0:
                    tickets[i] = srv.compress(fields[w], ebs[w], bss[w])
                else:
                    tickets[i] = srv.decompress(src, outs[w])
            ready.release()
        wt.join()
        el = time.perf_counter() - t0
        with lock:
            for t in latest.values():
                srv.release(t)
        return el, lat

    run(None)  # warm-up: the lanes' workspaces grow to their working size
    cap_s, cap_lat = run(None)
    cap_gbs = nbytes / cap_s / 1e9
    rate = 0.7 * nreq / cap_s
    arrivals = np.cumsum(rng.exponential(1.0 / rate, nreq))
    srv_s, lat = run(arrivals)
    st = srv.stats()
    srv.close()
    # the same requests one at a time with the one-field calls
    # the block sizes the request loop's tuner picks, for the one-field calls
    tab = tz._batch_table([x.view(-1) for x in fields[:nbig]], ebs[:nbig], [0] * nbig)
    tz._check(tz.lib().tszp_batch_tune(tab, nbig, tz._stream()))
    tuned = [tab[i].block_size for i in range(nbig)] + [big["bs"]]
    bufs, infos, bsel = {}, {}, {}
    lat1 = np.zeros(nreq)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(nreq):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        w = which[i]
        f = fields[w]
        if kinds[i] == 0 or w not in infos:
            bs = tuned[w]
            if w not in bufs:
                bufs[w] = tz.TopoBuffers(f.numel(), bs, 1, f.device)
            infos[w], _ = tz.topo_compress(f, ebs[w], bs, T=1, buf=bufs[w])
            bsel[w] = bs
        else:
            bs, info = bsel[w], infos[w]
            out, types = tz.topo_decompress(bufs[w], f.numel(), ebs[w], bs, T=1)
            sp = tz.sortpos_decompress(bufs[w], info.ext_count, bs, T=1) if info.ext_count else None
            tz.topo_restore(out, types, tuple(f.shape), ebs[w], sp=sp, n_ext=info.ext_count, out=outs[w])
        torch.cuda.synchronize()
        lat1[i] = 1e6 * (time.perf_counter() - t0 - arrivals[i])
    loop_s = time.perf_counter() - t0
    del bufs
    res = {"requests": nreq, "compress_share": float(np.mean(np.array(kinds) == 0)),
           "big_fields": int(sum(1 for w in which if w == nbig)), "input_gb": nbytes / 1e9,
           "hbm_budget_gb": budget / 2 ** 30, "hbm_peak_gb": st["hbm_peak"] / 2 ** 30, "lanes": 2,
           "capacity_gbs": cap_gbs, "capacity_requests_per_s": nreq / cap_s,
           "rate_per_s": rate, "served_gbs": nbytes / srv_s / 1e9,
           "served_p50_ms": float(np.percentile(lat, 50)) / 1e3, "served_p99_ms": float(np.percentile(lat, 99)) / 1e3,
           "loop_gbs": nbytes / loop_s / 1e9, "loop_p50_ms": float(np.percentile(lat1, 50)) / 1e3,
           "loop_p99_ms": float(np.percentile(lat1, 99)) / 1e3}
    res["capacity_vs_loop"] = res["capacity_gbs"] / res["loop_gbs"]
    log(f"served topology stream of {nreq} requests ({nbytes / 1e9:.1f} GB, {res['big_fields']} 512^3 fields): "
        f"capacity {cap_gbs:.1f} GB/s ({nreq / cap_s:.0f} requests/s); at {rate:.0f}/s: {res['served_gbs']:.1f} GB/s, "
        f"p50 {res['served_p50_ms']:.2f} ms, p99 {res['served_p99_ms']:.2f} ms; one request at a time with the "
        f"one-field calls: {res['loop_gbs']:.1f} GB/s, p50 {res['loop_p50_ms']:.1f} ms, p99 {res['loop_p99_ms']:.1f} ms "
        f"-> capacity {res['capacity_vs_loop']:.1f}x")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--bs", type=int, default=64)
    ap.add_argument("--rel", type=float, default=1e-3)
    ap.add_argument("--shape", type=int, nargs=3, default=[512, 512, 512])
    ap.add_argument("--seed", type=int, default=20251203)
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = min(cores, 16)")
    ap.add_argument("--cpu-slab", type=int, default=64, help="planes of the field for the CPU baseline")
    ap.add_argument("--cpu-slab-t1", type=int, default=8, help="planes for the one-thread CPU baseline")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-restore", action="store_true", help="leave out the topology restoration")
    ap.add_argument("--no-elem", action="store_true", help="leave out the threadblock / double side figures")
    ap.add_argument("--no-batch", action="store_true", help="leave out the mixed-size batch side figures")
    ap.add_argument("--no-serve", action="store_true", help="leave out the served-stream side figure")
    ap.add_argument("--no-other-bs", action="store_true", help="leave out the step at the other block size")
    ap.add_argument("--serve-requests", type=int, default=20000)
    ap.add_argument("--serve-rate", type=float, default=40000.0, help="Poisson arrival rate per second")
    ap.add_argument("--serve-topo-requests", type=int, default=2000)
    ap.add_argument("--profile-pass", action="store_true",
                    help="one step, then a 512 MiB elementwise copy that calibrates HBM counters; for rocprofv3 --pmc")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}): sharding a field over GPUs "
              f"(plane slabs with one-plane halos, a scan of per-shard byte counts, a bin-partitioned "
              f"exchange for the sort positions) is designed but not built or measured.", file=sys.stderr)
        sys.exit(2)

    import torch
    import toposzp_hip as tz

    assert torch.cuda.is_available(), "no GPU"
    dev = torch.device("cuda:0")
    shape = tuple(args.shape)
    n = shape[0] * shape[1] * shape[2]
    bs = args.bs
    log(f"generating fbm field {shape}")
    f = tz.gen_fbm(shape, seed=args.seed, octaves=6, noise=0.05, device=dev)
    eb, rng = tz.rel_to_abs(f, args.rel)
    log(f"eb={eb:.6g} (rel {args.rel} of range {rng:.6g}), bs={bs}")
    buf = tz.TopoBuffers(n, bs, 1, dev)

    last = {}
    restored = torch.empty(n, dtype=torch.float32, device=dev)

    def step():
        info, _ = tz.topo_compress(f, eb, bs, T=1, buf=buf)
        out, types = tz.topo_decompress(buf, n, eb, bs, T=1)
        sp = tz.sortpos_decompress(buf, info.ext_count, bs, T=1) if info.ext_count else None
        if not args.no_restore:
            last["restore"] = tz.topo_restore(out, types, shape, eb, sp=sp, n_ext=info.ext_count, out=restored)
        return info

    if args.profile_pass:
        info = step()
        x = torch.empty_like(f)
        torch.mul(f, 1.0, out=x)  # elementwise: reads and writes 4N bytes
        torch.cuda.synchronize()
        log(f"profile pass: cps={info.cp_count} topo={info.topo_bytes} sp={info.sortpos_bytes}")
        if not args.no_elem:  # one call of each side codec, for per-dispatch counters
            elem_side(tz, torch, f, shape, eb, bs, 1)
        return

    for i in range(args.warmup):
        info = step()
        log(f"warmup {i + 1}/{args.warmup}: cps={info.cp_count} ext={info.ext_count} "
            f"topo={info.topo_bytes} sp={info.sortpos_bytes}")
    torch.cuda.synchronize()
    tz.prof_reset()
    tz.prof_enable(True)
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = step()
    torch.cuda.synchronize()
    elapsed = time.perf_counter() - t0
    tz.prof_enable(False)
    prof = tz.prof_read()
    prof_restore = last.get("restore")  # with the solve's tile chain (computed while profiling)
    # a second, unprofiled timing of the same steps (event records add a little work)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        info = step()
    torch.cuda.synchronize()
    elapsed = min(elapsed, time.perf_counter() - t0)
    ms_step = 1e3 * elapsed / args.steps
    gbs = 4.0 * n / (elapsed / args.steps) / 1e9
    log(f"topology step {ms_step:.3f} ms -> {gbs:.1f} GB/s")

    # stage breakdown per step
    stages = {k: (v[0] / args.steps, v[1] / args.steps) for k, v in prof.items()}
    comp_payload = info.topo_bytes - 8
    sp_payload = max(0, info.sortpos_bytes - 16)
    # Single kernels timed per launch by the library, with their algorithmic bytes per
    # launch (the minimum HBM traffic of what each launch must read and write).
    ndim = len(shape)
    # The kernels of the step that the library times one by one, with their algorithmic
    # bytes per launch (the others, each under 0.35 ms per launch at 512^3, are in the
    # kernel-trace summary); None: no byte model (the window solve is a chain of
    # dependent tiles, bound by latency).
    kernels = {
        # read the field, write the topology payload (the 8 B/CP sort keys it also emits
        # are an intermediate of this design and not counted)
        "codec_topo": (codec_kernel(0, bs, ndim, shape[-1]), 4 * n + comp_payload),
        # bucket the critical points' value keys by bin: read and write 4 B each
        "scatter": ("k_bin_scatter", 8 * info.cp_count),
        # read each critical point's key and payload, write each extremum's sort position
        "query": ("k_bin_query", 8 * info.cp_count + 4 * info.ext_count),
        # read the sort positions, write their stream
        "codec_int": (codec_kernel(3, bs, 2, 4), 4 * info.ext_count + sp_payload),
        # the exact window solve of the restoration
        "solve": (f"k_pp_solve_flow<{ndim}, tszp::Geo1>", None),
    }

    def hbm_roofline(key):
        per_ms = prof[key][0] / max(1, prof[key][1])
        nbytes = kernels[key][1]
        ach = nbytes / (per_ms * 1e-3) / 1e9 if per_ms > 0 else None
        return {"bound": "hbm", "kernel": kernels[key][0], "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (ach / HBM_PEAK_GBS) if ach else None, "traffic": pmc_traffic(kernels[key][0]),
                "algorithmic_bytes_per_launch": nbytes, "avg_launch_ms": per_ms,
                "launches_per_step": prof[key][1] / args.steps}
    dom = max(kernels, key=lambda k: stages.get(k, (0, 0))[0])
    hbm_keys = [k for k in kernels if kernels[k][1] is not None]
    dom_hbm = max(hbm_keys, key=lambda k: stages.get(k, (0, 0))[0])
    if kernels[dom][1] is None:
        r = prof_restore
        per_ms = prof[dom][0] / max(1, prof[dom][1])
        chain = r.solve_chain if r is not None else 0
        roofline = {"bound": "latency", "kernel": kernels[dom][0], "achieved": None, "peak": None, "unit": None,
                    "frac": None, "traffic": None, "avg_launch_ms": per_ms, "launches_per_step": prof[dom][1] / args.steps,
                    "tiles": r.solve_tiles if r is not None else None, "dependent_chain_tiles": chain,
                    "ms_per_chain_tile": per_ms / chain if chain else None,
                    "hbm_kernel": hbm_roofline(dom_hbm)}
    else:
        roofline = hbm_roofline(dom)
    e2e_bytes = (4 * n + info.topo_bytes + info.sortpos_bytes) + (info.topo_bytes + info.sortpos_bytes + 5 * n
                                                                   + 4 * info.ext_count)
    e2e_frac = e2e_bytes / (elapsed / args.steps) / 1e9 / HBM_PEAK_GBS

    # plain SZp on the same field
    nbytes, pout, pblk = tz.plain_compress(f, eb, bs)
    pdec = torch.empty(n, dtype=torch.float32, device=dev)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        nbytes, _, _ = tz.plain_compress(f, eb, bs, out=pout, blk=pblk)
        tz.plain_decompress(pout, pblk, n, eb, bs, out=pdec)
    torch.cuda.synchronize()
    plain_gbs = 4.0 * n / ((time.perf_counter() - t0) / args.steps) / 1e9
    log(f"plain SZp compress+decompress {plain_gbs:.1f} GB/s (ratio {4 * n / nbytes:.2f})")

    # side figure: the same pipeline step at the workload's other block size
    other_bs = {}
    for obs in (64, 128):
        if obs == bs or args.no_other_bs:
            continue
        obuf = tz.TopoBuffers(n, obs, 1, dev)

        def ostep():
            oinfo, _ = tz.topo_compress(f, eb, obs, T=1, buf=obuf)
            oout, otypes = tz.topo_decompress(obuf, n, eb, obs, T=1)
            osp = tz.sortpos_decompress(obuf, oinfo.ext_count, obs, T=1) if oinfo.ext_count else None
            r = tz.topo_restore(oout, otypes, shape, eb, sp=osp, n_ext=oinfo.ext_count, out=restored)
            return oinfo, r
        ostep()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            oinfo, r = ostep()
        torch.cuda.synchronize()
        oms = 1e3 * (time.perf_counter() - t0) / 3
        other_bs[str(obs)] = {"ms_per_step": oms, "gbs": 4.0 * n / (oms * 1e-3) / 1e9,
                              "compression_ratio": 4 * n / (oinfo.topo_bytes + oinfo.sortpos_bytes),
                              "restore_mismatches": r.mismatches()}
        log(f"block size {obs}: step {oms:.3f} ms -> {other_bs[str(obs)]['gbs']:.1f} GB/s, "
            f"ratio {other_bs[str(obs)]['compression_ratio']:.3f}, restore mismatches {r.mismatches()}")
        del obuf

    # side figures: threadblock / double random-access codecs, double 3D topology
    elem = None if args.no_elem else elem_side(tz, torch, f, shape, eb, bs, 3)

    # side figure: 256 mixed-size fields through the batched topology pipeline
    topo_batch = None if args.no_batch else topo_batch_side(tz, torch, args.seed, 8)

    # side figure: 256 mixed-size fields, batched against a loop of one-field calls
    batch = None if args.no_batch else batch_side(tz, torch, args.seed, 12)
    served = None
    if batch is not None:
        pool = batch.pop("_pool")
        if not args.no_serve:
            served = serve_side(tz, torch, pool, {"field": f.view(-1), "eb": eb, "bs": bs}, args.seed,
                                args.serve_requests, args.serve_rate)
        del pool

    # side figure: the host-pointer drop-in chain on the same field (PCIe included)
    dropin = None if args.no_elem else dropin_side(tz, f, eb, bs)

    # side figure: the served topology stream (compress / decompress + restore requests)
    served_topo = None
    if not (args.no_batch or args.no_serve):
        served_topo = serve_topo_side(tz, torch, {"field": f, "eb": eb, "bs": bs}, args.seed,
                                      args.serve_topo_requests)

    # side figure: topology quality of the restored field against the original
    # (the CUDA driver's report), one GPU pass
    tz.quality_metrics(f, restored, shape, eb)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    quality = tz.quality_metrics(f, restored, shape, eb)
    quality_ms = 1e3 * (time.perf_counter() - t0)
    log(f"quality metrics {quality_ms:.3f} ms: preserved {quality['preserved']} of {quality['orig_cp'][0]}, "
        f"lost {quality['lost'][0]}, changed {quality['changed'][0]}, new {quality['new_cps']}, "
        f"max err {quality['max_err'] / eb:.3f} eb, order violations {quality['violations'][0]} of {quality['pairs'][0]}")

    # side figure: block offsets of the topology stream parsed on the device, as for a
    # stream read from a file without its side-car
    info, _ = tz.topo_compress(f, eb, bs, T=1, buf=buf)
    tz.parse_blocks(buf.topo[:info.topo_bytes], n, bs, "topo")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(3):
        tz.parse_blocks(buf.topo[:info.topo_bytes], n, bs, "topo")
    torch.cuda.synchronize()
    parse_ms = 1e3 * (time.perf_counter() - t0) / 3
    log(f"device block-offset parse of the topology stream {parse_ms:.3f} ms")

    cpu = None
    if not args.no_cpu:
        import numpy as np
        phys, allowed = physical_cores()
        threads = args.cpu_threads or min(allowed, 16)
        log(f"reference CPU baseline on a {args.cpu_slab}-plane slab with {threads} threads "
            f"(host: {phys} physical cores, {allowed} CPUs allowed)")
        field_np = f.cpu().numpy()
        pre = os.path.join("/tmp", f"tszp_cpu_dump_{os.getpid()}")
        res, err = cpu_baseline(field_np, eb, bs, threads, args.cpu_slab, dump=pre)
        checks = {}
        if res:
            t = res["compress_s"] + res["decompress_s"]
            cpu = {"value": 4.0 * res["n"] / t / 1e9, "unit": "GB/s", "cores": threads, "kind": "reference",
                   "sample": f"{res['shape']} slab of the same field, reference OpenMP compress "
                             f"({res['compress_s']:.3f} s) + decompress ({res['decompress_s']:.3f} s) "
                             f"library calls (no 3D CPU post-process exists in the reference)",
                   "host_physical_cores": phys, "host_cpus_allowed": allowed}
            # the reference's own stream, decoded by the GPU drop-in, against the reference's decode
            try:
                topo = np.fromfile(pre + ".topo", dtype=np.uint8)
                want = np.fromfile(pre + ".dec", dtype=np.float32)
                t0 = time.perf_counter()
                got, _ = tz.decompress_topology(topo, res["n"], eb, bs)  # host pointers: the drop-in
                checks["dropin_decode_s"] = time.perf_counter() - t0
                checks["decode_matches_reference"] = bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
            finally:
                for suf in (".topo", ".sp", ".dec"):
                    if os.path.exists(pre + suf):
                        os.unlink(pre + suf)
        else:
            cpu = {"value": None, "unit": "GB/s", "cores": threads, "kind": "reference", "sample": err}
        # one thread on a smaller slab, and its bytes against the native compressor's
        log(f"reference CPU path with 1 thread on a {args.cpu_slab_t1}-plane slab")
        res1, err1 = cpu_baseline(field_np, eb, bs, 1, args.cpu_slab_t1, dump=pre)
        if res1:
            t = res1["compress_s"] + res1["decompress_s"]
            cpu["t1"] = {"value": 4.0 * res1["n"] / t / 1e9, "unit": "GB/s", "cores": 1, "shape": res1["shape"],
                         "compress_s": res1["compress_s"], "decompress_s": res1["decompress_s"]}
            try:
                slab = f[:args.cpu_slab_t1].contiguous()
                sinfo, sbuf = tz.topo_compress(slab, eb, bs, T=1)
                ref_topo = np.fromfile(pre + ".topo", dtype=np.uint8)
                ref_sp = np.fromfile(pre + ".sp", dtype=np.uint8)
                checks["t1_streams_match_reference"] = bool(
                    np.array_equal(sbuf.topo[:sinfo.topo_bytes].cpu().numpy(), ref_topo) and
                    np.array_equal(sbuf.sp[:sinfo.sortpos_bytes].cpu().numpy(), ref_sp))
                del sbuf
            finally:
                for suf in (".topo", ".sp", ".dec"):
                    if os.path.exists(pre + suf):
                        os.unlink(pre + suf)
        else:
            cpu["t1"] = {"value": None, "sample": err1}
        # the reference's 2D drivers, post-process included, on the 1800 x 3600 blob plane
        plane = tz.gen_blobs([1800, 3600], args.seed, 64, 20.0, 200.0, 0.01, device=dev).cpu().numpy()
        drv = {}
        for thr in sorted({threads, 1}, reverse=True):
            log(f"reference 2D drivers on the 1800x3600 plane with {thr} thread(s)")
            drv[str(thr)] = driver_2d(plane, 1e-3, 64, thr)
        cpu["drivers_2d"] = {"field": "1800x3600 blobs + noise, abs eb 1e-3, bs 64", "by_threads": drv}
        cpu["checks"] = checks
        log(f"cpu baseline: {cpu}")

    line = {
        "metric": "topo compress+decompress GB/s & %HBM peak (512^3 f32); batched p50/p99 vs plain",
        "value": gbs,
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (device-generated fBm: 6 octaves value noise + 0.05 N(0,1))",
        "config": {"workload": f"topology compress+decompress {shape[0]}x{shape[1]}x{shape[2]} f32, "
                               f"rel {args.rel}, blockSize {bs}",
                   "shape": list(shape), "block_size": bs, "error_bound": f"rel:{args.rel}", "abs_eb": eb,
                   "nchunks": 1, "cp_count": info.cp_count, "ext_count": info.ext_count,
                   "compression_ratio": 4 * n / (info.topo_bytes + info.sortpos_bytes)},
        "roofline": roofline,
        "end_to_end_hbm_frac": e2e_frac,
        "stage_ms_per_step": {k: v[0] for k, v in stages.items()},
        "plain_gbs": plain_gbs,
        "other_block_sizes": other_bs,
        "topo_stream_parse_ms": parse_ms,
        "elem_codecs": elem,
        "batched_topo": topo_batch,
        "batched_plain": batch,
        "served_plain": served,
        "served_topo": served_topo,
        "dropin_host_pointers": dropin,
        "quality_ms": quality_ms,
        "quality": {k: quality[k] for k in ("orig_cp", "preserved", "lost", "changed", "new_cps", "max_err",
                                            "over_1eb", "pairs", "violations", "saddles_above",
                                            "saddles_kept_above")},
        "restore": ({"missing": last["restore"].missing, "spurious": last["restore"].spurious,
                     "changed": last["restore"].changed, "moves": last["restore"].moves,
                     "windows": last["restore"].windows, "rounds": last["restore"].rounds,
                     "max_dev_over_eb": last["restore"].max_dev_over_eb}
                    if "restore" in last else None),
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
