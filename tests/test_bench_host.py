"""bench.py's contract on the host: more than one GPU is refused with a message, and
the algorithmic flop counts it reports the roofline with follow their closed forms."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 2
    assert "--gpus 2 is not supported" in out.stderr
    assert out.stdout == ""


def test_roofline_flop_counts():
    import bench
    from exgame_amd import scenes
    sc = scenes.jointed_chains(n_chains=3, m=8, seed=0)
    ns = bench.ns_flops_per_iteration(sc)
    # a chain of 8: each block row meets itself and its neighbours, 8 + 2·7 = 22 blocks;
    # sparse A·X = 2·36·48·22, dense X·(2I − R) = 2·48³
    assert np.all(ns == 2 * 36 * 48 * 22 + 2 * 48 ** 3)
    assert np.all(bench.matvec_flops(sc) == 2 * 48 ** 2)
