"""Extract the reference's saved FP32-GPU validation values into a small JSON fixture.

Reads public/validation/results.json of the pocketwind reference (data, not code) and keeps
the rows this build checks against: every drag row (laminar, validation, confinement and
sensitivity sections, without the browser's adapter and user-agent records), the convergence
and grid-check results, the planned skips and the verification rows. Run once in the build
container:
    python tests/golden/make_reference_results.py /root/reference/public/validation/results.json
"""
import json
import sys



DROP = ("adapter", "userAgent")


def main(path):
    data = json.load(open(path))
    strip = lambda rows: [{k: v for k, v in row.items() if k not in DROP} for row in rows]
    out = {
        "source": "public/validation/results.json",
        "schemaVersion": data.get("schemaVersion"),
        "commit": data.get("commit"),
        "machine": data.get("machine"),
        "factors": data.get("factors"),
        "laminar": strip(data["laminar"]),
        "validation": strip(data["validation"]),
        "confinement": strip(data["confinement"]),
        "sensitivity": strip(data["sensitivity"]),
        "convergence": data.get("convergence"),
        "gridCheck": strip(data.get("gridCheck", [])),
        "skipped": data.get("skipped"),
        "verification": data.get("verification"),
    }
    json.dump(out, open(__file__.replace("make_reference_results.py", "reference_results.json"), "w"),
              indent=1)

if __name__ == "__main__":
    main(sys.argv[1])
