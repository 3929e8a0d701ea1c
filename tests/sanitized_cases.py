"""Oracle runs for the sanitizer test (tests/test_sanitized.py): run as a
script in a child process whose oracle library is the AddressSanitizer /
UBSan build (DESALM_ORACLE_LIB) with the sanitizer runtime preloaded.
Covers the paths with face-dependent extents: inflow (btype 4) and
characteristic-outlet (btype 5) faces in every position, both case hooks
with both u_tau solves, the actuator line with all features together, and
long i-lines.  Prints one line per case."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
for p in (HERE, os.path.join(ROOT, "des-alm_amd", "python"), os.path.join(ROOT, "oracle")):
    sys.path.insert(0, p)

import numpy as np  # noqa: E402

import pyoracle as po  # noqa: E402
from desalm import cases  # noqa: E402


def run(o, steps=2):
    for _ in range(steps):
        o.mg_driver()
    assert np.all(np.isfinite(o.get_state()[0]))


def main():
    print("oracle library", po.LIB_PATH, flush=True)
    for ng, des in ((3, True), (2, False)):
        for bt in ((4, 3, 1, 1, 2, 2), (4, 5, 1, 1, 2, 2), (5, 4, 1, 1, 2, 2), (3, 4, 1, 4, 2, 4),
                   (5, 5, 5, 1, 2, 5), (4, 5, 1, 1, 6, 6)):
            c = cases.wavy(17, 17, 17, ng=ng, itm=2, des=des, btype=bt, iters=(1, 2, 3))
            run(po.Oracle(c.params).load_case(c))
            print("btype", ng, des, bt, "ok", flush=True)
    for hooks in (1, 2, 3):
        for newt in (False, True):
            c = cases.wavy(17, 17, 21, ng=2, itm=2, des=True, btype=(3, 3, 1, 1, 2, 2), iters=(1, 2, 3), seed=5,
                           amp=0.05)
            run(po.Oracle(c.params).load_case(c).set_hooks(hooks, wf_newton=newt))
            print("hooks", hooks, newt, "ok", flush=True)
    from golden.make_golden import build
    c = build("wavy", dict(ni=21, nj=17, nk=17, ng=2, itm=2, amp=0.02, des=True, iters=(1, 2, 3),
                           btype=(4, 5, 1, 1, 2, 2), alm=dict(i_rot=9, center=[0.5, 0.5], radius=0.3, hub=0.06)))
    run(po.Oracle(c.params).load_case(c).set_hooks(3))
    print("actuator line, inflow, outlet, hooks ok", flush=True)
    for dims in ((70, 7, 7), (129, 5, 5)):
        c = cases.wavy(*dims, ng=1, itm=2, des=True, btype=(4, 5, 1, 1, 2, 2), iters=(1,))
        run(po.Oracle(c.params).load_case(c), steps=1)
        print("long i", dims, "ok", flush=True)
    print("ALL OK", flush=True)


if __name__ == "__main__":
    main()
