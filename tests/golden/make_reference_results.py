"""Extract the reference's saved FP32-GPU validation values into a small JSON fixture.

Reads public/validation/results.json of the pocketwind reference (data, not code) and keeps
the rows this build checks against: the laminar sphere Re 100 cases (Cd, voxel frontal area,
blockage, steps) and the verification rows. Run once in the build container:
    python tests/golden/make_reference_results.py /root/reference/public/validation/results.json
"""
import json
import sys

KEEP = ["id", "tier", "cells", "operator", "walls", "storage", "les", "floor", "bodyLengthCells",
        "cellSizeM", "uLu", "re", "steps", "flowThroughSteps", "cd", "cdVoxel", "voxelAreaM2",
        "referenceAreaM2", "blockage", "relStdPct"]


def main(path):
    data = json.load(open(path))
    out = {
        "source": "public/validation/results.json",
        "commit": data.get("commit"),
        "machine": data.get("machine"),
        "laminar": [{k: row.get(k) for k in KEEP} for row in data["laminar"]],
        "verification": data.get("verification"),
    }
    json.dump(out, open(__file__.replace("make_reference_results.py", "reference_results.json"), "w"),
              indent=1)


if __name__ == "__main__":
    main(sys.argv[1])
