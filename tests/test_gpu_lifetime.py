"""Handle lifetimes on the GPU: a context closes what lives in it before it
goes, and a process that ends with live contexts, matrices, vectors and
captured CG graphs exits cleanly (the exit hook frees them while the HIP
runtime is still up)."""
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_context_close_closes_dependents():
    import minife_amd as m
    ctx = m.Context(0)
    _, A, b = m.build_problem(6, ctx=ctx)
    x = m.Vector(A.nrows, ctx)
    res = m.cg_solve(A, b, x, max_iter=20)
    assert res.num_iters > 0
    ctx.close()
    assert A.h is None and b.h is None and x.h is None and ctx.h is None
    A.close()  # already closed: no call into the library
    ctx.close()


EXIT_SCRIPT = r"""
import sys
sys.path.insert(0, sys.argv[1])
import minife_amd as m
_, A, b = m.build_problem(12)
x = m.Vector(A.nrows)
r = m.cg_solve(A, b, x, max_iter=60)
print("iters", r.num_iters, flush=True)
ctx2 = m.Context(0)
_, A2, b2 = m.build_problem(8, ctx=ctx2)
x2 = m.Vector(A2.nrows, ctx2)
r2 = m.cg_solve(A2, b2, x2, max_iter=30)
print("iters", r2.num_iters, flush=True)
# no close(): the exit hook frees every handle
"""


def test_process_exits_cleanly_with_live_handles():
    r = subprocess.run([sys.executable, "-c", EXIT_SCRIPT, ROOT], capture_output=True, text=True,
                       timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    its = [int(l.split()[1]) for l in r.stdout.splitlines() if l.startswith("iters")]
    assert len(its) == 2 and all(i > 0 for i in its)
    assert np.all(np.array(its) <= 60)
