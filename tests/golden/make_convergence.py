#!/usr/bin/env python3
"""Writes tests/golden/conv_c2.npz: the CPU oracle's 16,384-sample frame of
the c2 mini scene (64x36, c2's geometry, materials and a 256x128 dome; depth
8), samples [2^20, 2^20 + 16384) of every pixel, the reference image of the
convergence test (tests/test_gpu_transport.py::test_convergence_rate).  It
runs the oracle, not the reference build, so it needs only oracle/liboracle.so:

    python tests/golden/make_convergence.py
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from helpers import c2_mini  # noqa: E402
from oracle.bindings import OracleScene  # noqa: E402
from shiro_amd.scene import RenderSettings  # noqa: E402
import make_golden  # noqa: E402

SPP, START = 16384, 1 << 20


def convergence_workload():
    return c2_mini(width=64, spp=16)


def main():
    wl = convergence_workload()
    st = RenderSettings(**{**wl.settings.__dict__, "samplesPerPixel": SPP})
    t0 = time.perf_counter()
    fb = OracleScene(wl.scene, bvh=True).render(wl.camera, st, START, SPP, threads=os.cpu_count() or 1)
    print(f"{st.width}x{st.height} x {SPP} spp in {time.perf_counter() - t0:.1f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "conv_c2.npz"), beauty=fb.beauty, spp=np.uint32(SPP),
                        sample_start=np.uint32(START), scene_hash=make_golden.scene_hash(wl.scene))


if __name__ == "__main__":
    main()
