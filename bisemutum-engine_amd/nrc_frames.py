"""Workload for profiling the radiance-cache form: the 1920x1080 atrium (262,144
triangles, depth 8) rendered by bi_pt_render_nrc for 8 frames at 1 spp (online
training) and 4 frames at 16 spp.  Run from the repository root, for example under
rocprofv3 --kernel-trace --stats -- python bisemutum-engine_amd/nrc_frames.py
"""
import os
import sys

sys.path[:0] = [os.path.dirname(os.path.abspath(__file__))]

import torch  # noqa: E402

import bi_pt  # noqa: E402
import scenes  # noqa: E402


def main():
    cfg = scenes.CONFIGS["c1"]
    cam = cfg["camera"]()
    pt = bi_pt.PathTracer()
    pt.set_scene(cfg["scene"]())
    nrc = bi_pt.Nrc(pt, train_samples=65536, train_steps=4, cache_bounce=2, seed=1)
    for k in range(8):
        nrc.render(cam, spp=1, max_bounces=cfg["max_bounces"], seed=100 + k)
    for k in range(4):
        nrc.render(cam, spp=16, max_bounces=cfg["max_bounces"], seed=200 + k)
    torch.cuda.synchronize()
    print("cache frames done", nrc.stats(), flush=True)
    nrc.close()
    pt.close()


if __name__ == "__main__":
    main()
