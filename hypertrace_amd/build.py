"""Build steps: the HIP runtime library (hipcc, gfx950) and the ahead-of-time
compiled render programs of the standard workloads (hipRTC into the
code-object cache, so a run on a fresh machine compiles none of them)."""
from __future__ import annotations

import os
import subprocess
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = "/opt/rocm/bin/hipcc"
LIB = os.path.join(PKG, "lib", "libhypertrace_hip.so")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # No packed FP32 instructions in any HIP code of this project.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    return r


def build_library() -> str:
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    tmp = LIB + ".tmp"
    run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(PKG, "csrc", "runtime.hip"), os.path.join(PKG, "csrc", "cache.hip"),
         os.path.join(PKG, "csrc", "batch.hip"), os.path.join(PKG, "csrc", "wavefront.hip"),
         "-L/opt/rocm/lib", "-lhiprtc",
         "-Wl,-rpath,/opt/rocm/lib", "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def standard_scenes():
    """Scene instances whose programs are compiled ahead of time (the program
    depends on the scene type only)."""
    from . import scenes

    return [
        scenes.euclidean_spheres().scene,
        scenes.euclidean_spheres(unit_spheres=True).scene,
        scenes.hyperbolic_cells().scene,
        scenes.reference_euclidean_example().scene,
        scenes.reference_hyperbolic_example().scene,
        scenes.sdf_shapes().scene,
        scenes.sdf_shapes(96, 72, 1, shapes=48, extent=4.0, camera_z=12.0).scene,
    ]


def _standard_jobs():
    """(scene, config, fp options, None) of the standard programs: per
    standard scene the FP32 / FP64 / fast / element-by-element programs,
    Russian roulette for the Euclidean ones, the sphere-tree option, the
    wider FP32 hit window, and the render tuner's candidates for the cell
    frame and (with roulette) the 1024-shape SDF scene."""
    from .processing import FAST_FP, NO_CULL_FP, program_options
    from .tuning import CANDIDATES, candidate_options
    from .types import X32_F32, X32_F64

    rr = program_options(rr_depth=3)
    tree = program_options() + " -DHT_CULL_TREE=1"
    jobs = []
    for k, scene in enumerate(standard_scenes()):
        extra = [(X32_F32, rr), (X32_F64, rr)] if scene.geo == "Eu" else []
        if k in (2, 5):  # the cell scene and the 1024-shape scene: the sphere-tree option
            extra.append((X32_F32, tree))
        if k == 0:  # the Euclidean sphere scene: the wider FP32 hit window, both scans
            extra += [(X32_F32, program_options(eps32=1e-5)), (X32_F32, program_options(eps32=1e-5, no_cull=True))]
        if k == 2:  # the cell scene: the render tuner's candidates
            extra += [(X32_F32, candidate_options(c)) for c in CANDIDATES]
            extra.append((X32_F32, program_options(scalar_bounds=True)))
        if k == 5:  # the 1024-shape scene: the tuner's candidates with Russian roulette
            extra += [(X32_F32, candidate_options(c, rr)) for c in CANDIDATES]
        for cfg, fp in [(X32_F32, None), (X32_F64, None), (X32_F32, FAST_FP), (X32_F32, NO_CULL_FP)] + extra:
            jobs.append((scene, cfg, fp, None))
    return jobs


def precompile_batch_programs(verbose: bool = False) -> None:
    """The batch programs of the autotuner's candidates, and the per-job
    programs (each batch scene type with HOPS = 4..12) that render jobs one
    at a time for comparison."""
    from . import scenes
    from .batch import BLOCKS, OPTIONS, precompile_batch
    from .objects import SceneImpl
    from .processing import precompile
    from .types import X32_F32

    t = time.time()
    for opts, blocks in OPTIONS:
        for block in BLOCKS:
            if block in blocks:
                precompile_batch(X32_F32, block, fp_options=opts)
    rng = __import__("numpy").random.default_rng(0)
    protos = [scenes.batch_euclidean_scene(rng, 8), scenes.batch_hyperbolic_scene(rng, 9)]
    for s in protos:
        for hops in range(4, scenes.BATCH_MAX_HOPS + 1):
            precompile(SceneImpl(s.geo, s.view, s.object, s.background, hops), X32_F32)
    if verbose:
        print(f"  batch and per-job programs in {time.time() - t:.1f}s")


def _suite_jobs():
    """(scene, config, fp options, root source or None[, root name]) of every program the
    GPU tests (tests/test_gpu_*.py, read for their workload and option
    tables) and bench.py's legs load."""
    import importlib.util

    from . import codegen, scenes
    from .processing import FAST_FP, NO_CULL_FP, program_options
    from .tuning import CANDIDATES, candidate_options
    from .types import X32_F32, X32_F64

    tests_dir = os.path.join(ROOT, "tests")

    def load(name):
        path = os.path.join(tests_dir, name + ".py")
        if not os.path.exists(path):
            return None
        spec = importlib.util.spec_from_file_location("_suite_" + name, path)
        mod = importlib.util.module_from_spec(spec)
        spec.loader.exec_module(mod)
        return mod

    jobs = _standard_jobs()  # (scene, config, fp options, root source or None[, root name])
    par = load("test_gpu_parity")
    if par is not None:
        for make in par.WORKLOADS.values():
            s = make().scene
            for fp in list(par.CULLED_VARIANTS.values()) + [None, NO_CULL_FP, FAST_FP, program_options(rr_depth=3)]:
                jobs.append((s, X32_F32, fp, None))
            jobs.append((s, X32_F64, None, None))
        jobs += [(par.WORKLOADS[name]().scene, X32_F32, program_options(hy_distance_fp64=True), None)
                 for name in par.HY_DIST_FP64_CASES]
        rr = program_options(rr_depth=3)
        jobs += [(par.WORKLOADS[name]().scene, X32_F32, fp, None)
                 for name in ("sdf_1024", "eu_ball") for fp in [rr] + [rr + o for o in par.HELPER_RR_OPTIONS.values()]]
        cap = par.CAPACITY_BASE().scene
        jobs += [(cap, X32_F32, par.CULLED_VARIANTS[v], None) for v in par.CAPACITY_VARIANTS]
        jobs.append((cap, X32_F32, NO_CULL_FP, None))
        eu = scenes.euclidean_spheres(96, 72).scene
        jobs += [(scenes.euclidean_spheres(w, h).scene, X32_F32, par.CULLED_VARIANTS[v], None)
                 for (w, h) in ((1, 1), (64, 48)) for v in ("pixel_queue", "sb_q_prio", "gq4_q_prio_w6", "gq4_ho_q_prio",
                                                           "gq4_ho_hl", "gq4_ho_q_s16", "gq4_ho_q_hl_s16",
                                                           "q_hl", "gq4_ho_q_hl_s32_sp64", "sp64", "q_hl_sp4",
                                                           "gq4_ho_hl_s32_sp8_w3", "gq4_ho_q_hl_s32_sp1",
                                                           "q_hl_sp1", "sp1_w3", "gq4_ho_q_hl_s32_sp1_pc",
                                                           "q_sp1_pc")]
        jobs += [(eu, X32_F32, program_options(eps32=1e-5), None),
                 (eu, X32_F32, program_options(eps32=1e-5, no_cull=True), None)]
    bt = load("test_gpu_batch")
    if bt is not None:
        from .batch import batch_root
        from .objects import SceneImpl

        programs = {(block, opts) for block, _, opts in bt.IDENTITY_CASES}
        programs |= {(256, opts) for opts in bt.GRAPH_OPTS} | {(128, None), (256, None)}
        programs |= {(block, opts) for _, opts, block in bt.EDGE_SETS}
        jobs += [(None, X32_F32, opts, batch_root(X32_F32, block), "batch.hip")
                 for block, opts in sorted(programs, key=str)]
        for prims, _, _ in bt.EDGE_SETS:
            for spec in bt.edge_specs(prims):  # the single-job programs the jobs are compared with
                sc = spec.scene
                jobs.append((SceneImpl(sc.geo, sc.view, sc.object, sc.background, spec.depth), X32_F32, None, None))
    ch = load("test_gpu_cache")
    if ch is not None:
        cells_type = scenes.hyperbolic_cells(64, 48).scene  # the program depends on the scene type only
        jobs += [(cells_type, X32_F32, opts, None) for opts in getattr(ch, "CACHE_PROGRAMS", [])]
    wf = load("test_gpu_wavefront")
    if wf is not None:
        for make, opts in wf.CASES.values():
            s = make()[0]
            jobs.append((s, X32_F32, opts, None))
            jobs.append((s, X32_F32, opts, codegen.root_source(s, X32_F32, wavefront=True)))
        s = scenes.hyperbolic_cells(80, 60).scene
        jobs += [(s, X32_F64, None, None), (s, X32_F64, None, codegen.root_source(s, X32_F64, wavefront=True))]
    pl = load("plugin_scenes")
    if pl is not None:  # the plugin scene and its built-in twin (test_gpu_plugins.py)
        jobs += [(s, cfg, None, None) for s in pl.scene_pair() for cfg in (X32_F32, X32_F64)]
    # bench.py legs
    cells = scenes.hyperbolic_cells().scene
    wf_src = codegen.root_source(cells, X32_F32, wavefront=True)
    jobs.append((cells, X32_F32, None, wf_src))
    jobs += [(cells, X32_F32, candidate_options(c), wf_src) for c in CANDIDATES]  # the wavefront leg's program
    for wl in (scenes.hyperbolic_cells(), scenes.hyperbolic_cells(1920, 1080, 1, primitives=256)):
        jobs += [(wl.scene, X32_F32, candidate_options(c), None) for c in CANDIDATES]
        jobs += [(wl.scene, X32_F32, NO_CULL_FP, None), (wl.scene, X32_F64, None, None)]
    rr = program_options(rr_depth=3)
    sdf = scenes.sdf_shapes().scene
    jobs += [(sdf, X32_F32, candidate_options(c, rr), None) for c in CANDIDATES]
    jobs.append((sdf, X32_F64, rr, None))
    return jobs


def _suite_shard(k: int, n: int) -> int:
    from .processing import precompile

    compiled = 0
    for job in _suite_jobs()[k::n]:
        scene, cfg, fp, source = job[:4]
        name = job[4] if len(job) > 4 else "render.hip"
        compiled += not precompile(scene, cfg, name=name, fp_options=fp, source=source)
    return compiled


def precompile_suite_programs(verbose: bool = False, workers: int = 8) -> None:
    """The standard programs and those the GPU tests and bench.py's legs
    load, so that a GPU run compiles nothing: the render tuner's candidates
    for the cell frames and the SDF scene, the plain and FP64 programs of
    each leg, the wavefront forms, and every scene x option pair of the
    parity tables. Cached programs cost only their hash; missing ones
    compile in `workers` processes (each writes the cache through its own
    temporary files), the longest jobs spread over all of them."""
    t0 = time.time()
    n = len(_suite_jobs())
    procs = [subprocess.Popen([sys.executable, "-c",
                               f"import sys; sys.path.insert(0, {ROOT!r}); from hypertrace_amd import build; "
                               f"print(build._suite_shard({k}, {workers}))"],
                              stdout=subprocess.PIPE, text=True) for k in range(workers)]
    compiled = 0
    for p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            raise RuntimeError(f"precompiling the test and bench programs failed (shard exit {p.returncode})")
        compiled += int(out.strip().splitlines()[-1])
    if verbose:
        print(f"  test and bench programs: {n} ({compiled} compiled) in {time.time() - t0:.1f}s")


CACHE_DIR = os.path.join(PKG, "_cache")


def prune_cache(used_since: float, verbose: bool = False) -> int:
    """Removes code objects no build step used since `used_since` (a cache
    hit refreshes the entry's modification time): programs of earlier
    device sources, which would only add to what travels with the tree."""
    removed = 0
    if not os.path.isdir(CACHE_DIR):
        return 0
    for name in os.listdir(CACHE_DIR):
        path = os.path.join(CACHE_DIR, name)
        if name.endswith(".hsaco") and os.path.getmtime(path) < used_since:
            os.remove(path)
            removed += 1
    if verbose:
        print(f"  removed {removed} unused code objects from the cache")
    return removed


def build_all(verbose: bool = True) -> None:
    t = time.time()
    build_library()
    if verbose:
        print(f"built {LIB} in {time.time() - t:.1f}s")
    start = time.time() - 1.0
    precompile_batch_programs(verbose)
    precompile_suite_programs(verbose)
    prune_cache(start, verbose)


if __name__ == "__main__":
    build_all()
