"""Write tests/golden/fixtures.json.

Inputs: the two fixtures of the reference's check scripts (restated in
tests/cases.py from examples/check_forward_operator.py:108-133 and
examples/check_operator.py:67-94) and the CPU-runnable reference-size case.
Expected outputs: the scalar pure-Python FP64 restatements in oracle/
(ray_integral_py, bp_scalar_py), which share no code with the C oracle or the
HIP library.  Run from the repository root:  python tests/golden/make_golden.py
"""

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "tests")]

import numpy as np  # noqa: E402

import cases  # noqa: E402
import oracle  # noqa: E402


def main():
    vol, geom, up = cases.fp_case()
    vol_np = vol.contiguous().numpy()
    geom_np = [g.numpy() for g in geom]
    fp_out = oracle.ray_integral_py(vol_np, geom_np, cases.FP_PIXEL, cases.FP_U, cases.FP_V,
                                    cases.FP_STEP, cases.FP_VOXEL, cases.FP_OFFSET)
    proj, bgeom, bup = cases.bp_case()
    proj_np = proj.contiguous().numpy()
    bgeom_np = [g.numpy() for g in bgeom]
    bp_out = oracle.bp_scalar_py(proj_np, bgeom_np, cases.BP_PIXEL, cases.BP_SHAPE,
                                 cases.BP_VOXEL, cases.BP_OFFSET)
    bp_loss = oracle.bp_scalar_py(proj_np, bgeom_np, cases.BP_PIXEL, cases.BP_SHAPE,
                                  cases.BP_VOXEL, cases.BP_OFFSET, weights=bup.contiguous().numpy())
    data = {
        "fp_case": {
            "volume": vol_np.tolist(), "geometry": [g.tolist() for g in geom_np],
            "upstream": up.contiguous().numpy().tolist(),
            "pixel_size": cases.FP_PIXEL, "U": cases.FP_U, "V": cases.FP_V,
            "step_size": cases.FP_STEP, "voxel_size": cases.FP_VOXEL, "offset": list(cases.FP_OFFSET),
            "expected_projection_f64": fp_out.tolist(),
        },
        "bp_case": {
            "projections": proj_np.tolist(), "geometry": [g.tolist() for g in bgeom_np],
            "upstream": bup.contiguous().numpy().tolist(),
            "pixel_size": cases.BP_PIXEL, "volume_shape": list(cases.BP_SHAPE),
            "voxel_size": cases.BP_VOXEL, "offset": list(cases.BP_OFFSET),
            "expected_volume_f64": bp_out.tolist(), "expected_weighted_loss_f64": bp_loss,
        },
        "c1": {"seed": cases.C1_SEED, **cases.C1,
               "volume_sha_sum": float(np.float64(cases.c1_volume()).sum())},
    }
    with open(os.path.join(HERE, "fixtures.json"), "w") as fh:
        json.dump(data, fh, indent=1)
    print("wrote fixtures.json")


if __name__ == "__main__":
    main()
