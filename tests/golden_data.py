"""Loads the committed golden fixtures (tests/golden/) and rebuilds the
surfaces they were made from."""
import json
import os

import numpy as np

from jsi_cad_next_versions_amd import surfaces as S

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    with open(os.path.join(GOLDEN, "golden.json")) as f:
        gold = json.load(f)
    arrays = dict(np.load(os.path.join(GOLDEN, "eval_f32.npz"), allow_pickle=False))
    return gold, arrays


def bits_to_f32(b):
    return np.array([b], np.uint32).view(np.float32)[0]


_surf = None


def surfaces():
    """name -> Surface for every surface the fixtures use that is available here.
    The degree-11 CAD faces (inter_case1/2) come from the reference library
    build (oracle/_ref) when it exists."""
    global _surf
    if _surf is not None:
        return _surf
    wa, wb = S.wave_pair_small()
    ha, hb = S.hard_wave_pair()
    sa, sb = S.saddle_wave_pair()
    out = {"wave16_A": wa, "wave16_B": wb, "hardwave_A": ha, "hardwave_B": hb, "saddle": sa, "waveplane": sb}
    z = np.load(os.path.join(GOLDEN, "sample_surfaces.npz"), allow_pickle=False)
    for c in (3, 4):
        for o in (1, 2):
            n = f"inter_case{c}_obj{o}"
            out[n] = S.Surface(z[n + "_ivec"], z[n + "_data"], n)
    try:
        import oracle
        if oracle.ref_available():
            r = oracle.RefV5()
            for c in (1, 2):
                for o in (1, 2):
                    n = f"inter_case{c}_obj{o}"
                    out[n] = r.surface(n)
    except Exception:
        pass
    _surf = out
    return out
