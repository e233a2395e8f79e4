"""The batch workload at full size against the oracle: 64 independent scenes (config0
generator, log-normal 5k-500k triangles, resolutions 256x256 / 512x512 / 768x432 /
1024x1024), each resolution class autotuned on its largest scene, rendered as one hipGraph
launch (lucid_batch_render) and replayed three times. Every frame of the batch is checked
against the oracle's frame of the same scene: LucidInfo counters and per-bin arrays, per-bin
lists (multisets) and per-pixel fragment counts bit-exact, colour within 1/255."""

import numpy as np
import pytest

from lucid_amd import scenes

from test_gpu_parity import _check

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c3_batch():
    import torch
    from lucid_amd.batch import LucidBatch
    from lucid_amd.renderer import GpuScene, LucidRenderer
    scs = scenes.config3()
    print(f"\n[case] C3 batch of {len(scs)} scenes, "
          f"{sum(s.num_quads for s in scs)} triangles", flush=True)
    rs = [LucidRenderer((s.width, s.height), max_quads=s.num_quads) for s in scs]
    gs = [GpuScene(s) for s in scs]
    outs = [r.alloc_outputs(rgba8=True, rgba32f=False, counts=True) for r in rs]
    tuned = {}
    for res in sorted({(s.width, s.height) for s in scs}):
        idx = max((i for i, s in enumerate(scs) if (s.width, s.height) == res),
                  key=lambda i: scs[i].num_quads)
        tuned[res] = rs[idx].autotune(gs[idx], outs[idx], iters=1)
        print(f"[c3] tuning {res}: {tuned[res]}", flush=True)
    for s, r in zip(scs, rs):
        r.set_tuning(**tuned[(s.width, s.height)])
    batch = LucidBatch(rs)
    first = None
    for rep in range(3):
        for o in outs:  # eager work between graph replays
            for v in o.values():
                if v is not None:
                    v.zero_()
        torch.cuda.synchronize()
        batch.render(gs, outs)
        torch.cuda.synchronize()
        got = [(o["image"].cpu().numpy(), o["frag_counts"].cpu().numpy()) for o in outs]
        if first is None:
            first = got
        else:
            for i, (a, b) in enumerate(zip(first, got)):
                assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), (rep, i)
        print(f"[c3] replay {rep} done", flush=True)
    frames = []
    for r, (img, cnt) in zip(rs, first):
        q, t = r.read_bin_lists()
        frames.append({"image": img, "image_f": None,
                       "frag_counts": cnt.astype(np.uint32), "info": r.read_info(),
                       "verify": r.verify_info(), "bin_quads": q, "bin_tris": t})
    del batch, outs, gs, rs
    torch.cuda.empty_cache()
    return scs, frames


@pytest.mark.parametrize("res", scenes.BATCH_RESOLUTIONS)
def test_c3_batch_frames_match_oracle(oracle_mod, c3_batch, res):
    scs, frames = c3_batch
    idx = [i for i, s in enumerate(scs) if (s.width, s.height) == res]
    for i in idx:
        sc = scs[i]
        ref = oracle_mod.render(sc, want_records=False)
        _check(sc, frames[i], ref, oracle_mod)
        print(f"[c3] frame {i} {sc.name}: {ref['num_fragments']} fragments ok", flush=True)
