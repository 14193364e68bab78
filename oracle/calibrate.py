"""CPU calibration of the oracle (test infrastructure, not the product).

The oracle's explicit loop rebuilds every family and element each step, with one LU per bond and
a binary search per scattered entry, as the reference does (pdsolve.cpp:5819-5840, :5842-6090).
This script times it on one thread, on two of the reference's own example meshes, and prints the
ratio to the reference binary's measured seconds per step on the same meshes:
  * examples/box (hex8, m = 4, 102,500 eq; from tests/golden/box): 26.5 s/step;
  * examples/plateCrack (quad4 with a crack, 14,026 eq): 0.290 s/step.
The reference binary was built with g++ -O2, MKL 2021.4 sequential and MPICH 3.3.2 and run as one
MPI rank on an 8-core Xeon at 2.70 GHz, lumped mass, no failure criterion.
The plateCrack mesh is read from the reference checkout, so that case runs only where the
checkout is present:

    python oracle/calibrate.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PLATE_CRACK = "/root/reference/examples/plateCrack/platCrack_NewMarkOFF_FC2.in"
REF_S_PER_STEP = {"box": 26.5, "plateCrack": 0.290}


def time_steps(oracle, mesh, steps):
    """Seconds per explicit step on one thread: (start-up + steps) - (start-up alone), over steps."""
    oracle.lib().om_set_threads(1)
    mesh = dict(mesh)
    mesh["dt"] = 1e-9  # the timing does not depend on dt; keeps the state small
    M = oracle.Model(mesh)
    t0 = time.time()
    M.explicit_run(0)
    t_start = time.time() - t0
    M = oracle.Model(mesh)
    t0 = time.time()
    M.explicit_run(steps)
    t_all = time.time() - t0
    return (t_all - t_start) / steps, M.neq


def main():
    sys.path.insert(0, HERE)
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import oracle
    import make_box_fixture
    oracle.build()
    out = {}
    box, _ = make_box_fixture.load_box()
    s, neq = time_steps(oracle, box, 2)
    out["box"] = {"neq": neq, "oracle_s_per_step": s, "reference_s_per_step": REF_S_PER_STEP["box"],
                  "reference_over_oracle": REF_S_PER_STEP["box"] / s}
    if os.path.exists(PLATE_CRACK):
        pc = oracle.read_case(PLATE_CRACK)
        s, neq = time_steps(oracle, pc, 20)
        out["plateCrack"] = {"neq": neq, "oracle_s_per_step": s,
                             "reference_s_per_step": REF_S_PER_STEP["plateCrack"],
                             "reference_over_oracle": REF_S_PER_STEP["plateCrack"] / s}
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
