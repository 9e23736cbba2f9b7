"""Write the golden fixtures of the 500-atom workload (cfg0 geometry, seed 0).

The reference ships no test vectors, so these fixtures hold this repository's own FP64
oracle results on the seeded synthetic globule: the geometry (so a change of the generator
is caught), and per model the energy, forces, Born radii and integrals.  They pin the
oracle against regressions (tests/test_golden.py, CPU) and give the GPU path a fixed
target that needs no oracle run (tests/test_gpu_golden.py).

    python3 tests/golden/make_golden.py        # rewrites tests/golden/globule500_seed0.npz
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from openmmcustomgbintegralforce_amd import _native as N  # noqa: E402
from openmmcustomgbintegralforce_amd import synthetic as S  # noqa: E402

OUT = os.path.join(HERE, "globule500_seed0.npz")

# model name -> ABI parameters (defaults of the ABI, quadrature of the workload configs)
MODELS = {
    "gbsw": dict(model="gbsw", n_lebedev=38),
    "gbmv2": dict(model="gbmv", integral=2, n_lebedev=50),
    "gbmv1": dict(model="gbmv", integral=1, n_lebedev=26),
}


def params(name: str):
    m = MODELS[name]
    if m["model"] == "gbsw":
        p = N.default_params(N.MODEL_GBSW)
    else:
        p = N.default_params(N.MODEL_GBMV, m["integral"])
    p.n_lebedev = m["n_lebedev"]
    return p


def main() -> int:
    s = S.globule(500, 0)
    out = {"pos": s.pos, "charge": s.charge, "radius": s.radius}
    for name in MODELS:
        ref = O.compute(O.params_from_abi(params(name), nthreads=8), s.pos, s.charge, s.radius)
        out[f"{name}_energy"] = np.array([ref.energy])
        out[f"{name}_forces"] = ref.forces
        out[f"{name}_born"] = ref.born
        out[f"{name}_integrals"] = ref.integrals
        print(f"{name}: E = {ref.energy:.10f} kJ/mol", flush=True)
    out["meta"] = np.frombuffer(json.dumps({"models": MODELS, "generator": "synthetic.globule(500, 0)"}).encode(),
                                dtype=np.uint8)
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
