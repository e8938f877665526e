"""Build steps: the HIP runtime library (hipcc, gfx950) and the ahead-of-time
compiled render programs of the standard workloads (hipRTC into the
code-object cache, so a fresh GPU box does not compile them again)."""
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


def precompile_programs(verbose: bool = False) -> None:
    from .processing import FAST_FP, NO_CULL_FP, precompile, program_options
    from .types import X32_F32, X32_F64

    from .tuning import CANDIDATES, candidate_options

    rr = program_options(rr_depth=3)
    tree = program_options() + " -DHT_CULL_TREE=1"
    scenes_ = standard_scenes()
    for k, scene in enumerate(scenes_):
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
            t = time.time()
            hit = precompile(scene, cfg, fp_options=fp)
            if verbose:
                print(f"  program {scene.geo} f64={cfg.double_support} fp={fp or 'strict'}: "
                      f"{'cached' if hit else 'compiled'} in {time.time() - t:.1f}s")


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


def build_all(verbose: bool = True) -> None:
    t = time.time()
    build_library()
    if verbose:
        print(f"built {LIB} in {time.time() - t:.1f}s")
    precompile_programs(verbose)
    precompile_batch_programs(verbose)


if __name__ == "__main__":
    build_all()
