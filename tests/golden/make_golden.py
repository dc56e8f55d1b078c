"""Write the golden fixtures of this directory from the reference's own code.

Runs the two drivers that oracle/Makefile builds from the reference headers where they
lie (`make -C oracle ref`), in this container only:
  oracle/_ref/refgraph   -> ref_graph.json   (reference graph.hh: log p, d/dtheta, d2/dtheta2
                                               of the mass and angular densities)
  oracle/_ref/refrandom  -> ref_random.json  (reference random.hh: Splitmix64, Xoshiro128++,
                                               jump, float-rounded uniforms, Poisson)
The fixtures are data (inputs and expected outputs) and are committed; the GPU runs and
the CPU tests read only these JSON files.
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref"])
    for exe, out in (("refgraph", "ref_graph.json"), ("refrandom", "ref_random.json")):
        txt = subprocess.check_output([os.path.join(ROOT, "oracle", "_ref", exe)]).decode()
        json.loads(txt)  # validate
        with open(os.path.join(HERE, out), "w") as f:
            f.write(txt)
        print("wrote", out, len(txt), "bytes")


if __name__ == "__main__":
    sys.exit(main())
