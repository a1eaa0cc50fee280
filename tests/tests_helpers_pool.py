"""Oracle layers in worker processes (spawned, so that the GPU context of the test
process is not copied) for the full-size parity tests."""
import multiprocessing as mp
import os

_ORC = {}


def _init(kind):
    from clblm_amd import workloads as W
    from oracle.py import Oracle
    lines = W.profile_lines(200000, 500.0, 2500.0, seed=1001) if kind == "profile" else W.large_scene_lines()
    _ORC["o"] = Oracle(lines)


def _layer(args):
    lay, v1, v2, dv, cont = args
    o = _ORC["o"]
    o.jrad = 1
    od, _ = o.odlay(lay, v1, v2, dv, continuum=cont)
    return od


def oracle_layers(kind, jobs, workers=None):
    """The oracle's OD of each (layer, v1, v2, dv, continuum) job, each from JRAD 1, in
    a pool of spawned workers holding the line list of `kind` ("profile" or "large")."""
    n = workers or max(1, min(8, (os.cpu_count() or 2) // 2))
    with mp.get_context("spawn").Pool(n, initializer=_init, initargs=(kind,)) as pool:
        return pool.map(_layer, jobs, chunksize=1)
