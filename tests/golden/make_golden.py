#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ from the reference's own CPU
path tracer, compiled from its sources by oracle/Makefile into
oracle/_ref/libshiro_ref.so (this needs the reference checkout, so it runs
only where that exists; the fixtures it writes are committed data).

Fixtures (numpy .npz, float32/uint32 arrays):
  qmc.npz       LowDiscrepancySample over pixels x samples x dims and
                RadicalInverse for every prime base (OpenQmcSampler.cpp:43-71)
  bsdf.npz      EvaluateBsdf / EvaluateBsdfPdf on random tuples over varied
                materials, lobe weights and diffuse colours
                (CpuPathTracer.cpp:1232-1403), GGX/Fresnel scalars
  env.npz       BuildImportanceTable rows/conditional CDFs, Sample and
                Pdf/Evaluate of a 64x32 procedural lat-long map
                (EnvironmentMap.cpp:109-236)
  env_angular.npz  Sample / Pdf / Evaluate of a 48x48 angular (light-probe)
                map, which has no importance table (EnvironmentMap.cpp:109-183,
                247-264)
  img_*.npz     CpuPathTracer::RenderSampleBatch frames of small c1/c2/c3
                scenes (brute-force intersection), of a c1 scene with emissive
                area lights and of a c2 scene under a rotated angular dome,
                plus a two-batch FrameAccumulator result (the second batch
                runs with the path-guide state the first one left)
Each image fixture stores a hash of the scene arrays it was rendered from,
so the test detects a change of the workload generator.

    python tests/golden/make_golden.py [section ...]

Sections: qmc bsdf env env_angular images features accumulated, or
img:<name> for one image fixture; all sections when none is named.
`features` renders the scenes of golden_cases.feature_workloads (opacity,
delta glass and mirrors, thin-walled panes, absorbing media, normal
overrides).
"""
import ctypes as C
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from oracle.bindings import ReferenceScene, eval_bsdf, reference_available, reference_fns  # noqa: E402
from shiro_amd import abi  # noqa: E402
from shiro_amd.scene import DomeLight, Scene  # noqa: E402
from shiro_amd.workloads import procedural_dome  # noqa: E402
import golden_cases as cases  # noqa: E402


def scene_hash(scene):
    h = hashlib.sha256()
    for m in scene.meshes:
        h.update(np.ascontiguousarray(m.positions, np.float32).tobytes())
        h.update(np.ascontiguousarray(m.normals, np.float32).tobytes())
        h.update(np.ascontiguousarray(m.indices, np.uint32).tobytes())
    for mat in scene.materials:
        h.update(bytes(mat.to_c()))
    for d in scene.domeLights:
        if d.pixels is not None:
            h.update(np.ascontiguousarray(d.pixels, np.float32).tobytes())
        if d.layout != abi.ENV_AUTOMATIC:
            h.update(np.array([d.layout, *d.right, *d.up, *d.forward], np.float32).tobytes())
    return h.hexdigest()


def main(sections):
    if not reference_available():
        sys.exit("oracle/_ref/libshiro_ref.so is not built (needs the reference checkout)")
    f = reference_fns()

    def want(name):
        return not sections or name in sections

    # --- sampler
    if want("qmc"):
        write_qmc(f)
    if want("bsdf"):
        write_bsdf(f)
    if want("env"):
        px_env = procedural_dome(64, 32, 12345, hot_spots=2, peak=500.0, spot_width_deg=6.0)
        write_env("env.npz", Scene(domeLights=[DomeLight(pixels=px_env)]))
    if want("env_angular"):
        write_env("env_angular.npz", Scene(domeLights=[cases.angular_dome()]))
    for name, wl in cases.image_workloads().items():
        if want("images") or want(f"img:{name}"):
            render_image(name, wl)
    for name, wl in cases.feature_workloads().items():
        if want("features") or want(f"img:{name}"):
            render_image(name, wl)
    if want("accumulated"):
        name, wl, batches = cases.accumulated_workload()
        fb = ReferenceScene(wl.scene).render_accumulated(wl.camera, wl.settings, batches)
        np.savez_compressed(os.path.join(HERE, f"img_{name}.npz"), beauty=fb.beauty, albedo=fb.albedo,
                            normal=fb.normal, depth=fb.depth, scene_hash=scene_hash(wl.scene),
                            batches=np.array(batches, np.uint32))
    print("done")


def render_image(name, wl):
    print("rendering", name, wl.settings.width, wl.settings.height, wl.settings.samplesPerPixel, flush=True)
    fb = ReferenceScene(wl.scene).render(wl.camera, wl.settings)
    np.savez_compressed(os.path.join(HERE, f"img_{name}.npz"), beauty=fb.beauty, albedo=fb.albedo,
                        normal=fb.normal, depth=fb.depth, scene_hash=scene_hash(wl.scene))


def write_qmc(f):
    px, py, ss, dims = cases.qmc_grid()
    vals = np.zeros((len(py), len(px), len(ss), len(dims)), np.float32)
    for j, y in enumerate(py):
        for i, x in enumerate(px):
            for k, s in enumerate(ss):
                for d, dim in enumerate(dims):
                    vals[j, i, k, d] = f.qmc(int(x), int(y), int(s), int(dim))
    bases, idx = cases.radical_inverse_cases()
    ri = np.array([[f.radical_inverse(int(b), int(n)) for n in idx] for b in bases], np.float32)
    np.savez_compressed(os.path.join(HERE, "qmc.npz"), px=px, py=py, s=ss, dims=dims, values=vals,
                        bases=bases, index=idx, radical_inverse=ri)


def write_bsdf(f):
    mats = cases.bsdf_materials()
    mi, n, cn, v, l = cases.bsdf_tuples(len(mats))
    val, pdf = eval_bsdf(f, mats, mi, n, cn, v, l)
    w4 = np.zeros((len(mats), 4), np.float32)
    dc = np.zeros((len(mats), 3), np.float32)
    for i, m in enumerate(mats):
        c = m.to_c()
        f.lobe_weights(C.byref(c), abi.fp(w4[i]), abi.fp(dc[i]))
    nh, rough, cosi, etai, etat = cases.scalar_cases()
    ggx = np.array([f.ggx_d(float(a), float(r)) for a, r in zip(nh, rough)], np.float32)
    g2 = np.array([f.ggx_g2(float(a), float(b), float(r)) for a, b, r in zip(nh, cosi, rough)], np.float32)
    fres = np.array([f.fresnel_dielectric(float(c), float(a), float(b)) for c, a, b in zip(cosi, etai, etat)],
                    np.float32)
    np.savez_compressed(os.path.join(HERE, "bsdf.npz"), mat_index=mi, n=n, cn=cn, v=v, l=l, value=val, pdf=pdf,
                        weights=w4, diffuse=dc, nh=nh, rough=rough, cosi=cosi, etai=etai, etat=etat, ggx_d=ggx,
                        ggx_g2=g2, fresnel=fres)


def write_env(fname, sc):
    d = sc.domeLights[0]
    r = ReferenceScene(sc)
    if d.layout == abi.ENV_ANGULAR:  # no importance table (EnvironmentMap.cpp:189-191)
        row, cond = np.zeros(0, np.float32), np.zeros((0, 0), np.float32)
    else:
        row, cond = r.env_tables()
    u, dirs = cases.env_queries()
    sd, sp, sc_ = r.env_sample(u)
    ep, ec = r.env_eval(dirs)
    np.savez_compressed(os.path.join(HERE, fname), pixels=d.pixels, layout=np.uint32(d.layout), row_cdf=row,
                        cond_cdf=cond, u=u, sample_dir=sd, sample_pdf=sp, sample_rgb=sc_, dirs=dirs, eval_pdf=ep,
                        eval_rgb=ec)


if __name__ == "__main__":
    main(sys.argv[1:])
