"""Extract the basis-set data this package needs from the reference's BSE JSON files.

The reference reads Basis Set Exchange JSON (schema 0.1) from ``basis_sets/<name>.json``
(reference ``src/Basis.cpp:89-171``).  Only the elements used by the workloads here
(H, C, N, O) are kept, and only the fields the loader reads: ``angular_momentum``,
``exponents`` and ``coefficients`` (kept as the original strings, so the parsed doubles are
the same as the reference's ``from_chars`` results).

``cc-pvdz-expanded`` is derived from ``cc-pvdz``: every general-contraction row becomes a
shell entry of its own, and primitives with a zero coefficient in that row are dropped.
The reference reader keeps only ``coefficients[i]`` for ``i < len(angular_momentum)``
(``src/Basis.cpp:125-127,145-146``), so a general contraction written as one entry loses
its extra rows; written one row per entry, all rows survive.

This is the provenance script of the committed data in ``hartree-fock_amd/hfx/basis_data/``;
the package never runs it.  Run once, in a tree where ``/root/reference`` exists::

    python tests/golden/extract_basis_data.py
"""
import json
import os
import sys

REF = "/root/reference/basis_sets"
HERE = os.path.join(os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))),
                    "hartree-fock_amd", "hfx", "basis_data")
ELEMENTS = ["1", "6", "7", "8"]
FILES = {
    "sto-3g": "sto-3g.json",
    "6-31g": "6-31g.json",
    "6-31g_st__st_": "6-31g_st__st_.json",
    "cc-pvdz": "cc-pvdz.json",
}


def _trim(src):
    out = {"elements": {}}
    for z in ELEMENTS:
        if z not in src["elements"]:
            continue
        shells = []
        for sh in src["elements"][z]["electron_shells"]:
            shells.append({
                "angular_momentum": list(sh["angular_momentum"]),
                "exponents": list(sh["exponents"]),
                "coefficients": [list(r) for r in sh["coefficients"]],
            })
        out["elements"][z] = {"electron_shells": shells}
    return out


def _expand(trimmed):
    out = {"elements": {}}
    for z, el in trimmed["elements"].items():
        shells = []
        for sh in el["electron_shells"]:
            am = sh["angular_momentum"]
            rows = sh["coefficients"]
            for r, row in enumerate(rows):
                l = am[r] if r < len(am) else am[-1]
                keep = [k for k, c in enumerate(row) if float(c) != 0.0]
                shells.append({
                    "angular_momentum": [l],
                    "exponents": [sh["exponents"][k] for k in keep],
                    "coefficients": [[row[k] for k in keep]],
                })
        out["elements"][z] = {"electron_shells": shells}
    return out


def main():
    if not os.path.isdir(REF):
        print("reference basis directory not found:", REF)
        return 1
    for name, fn in FILES.items():
        with open(os.path.join(REF, fn)) as f:
            trimmed = _trim(json.load(f))
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(trimmed, f, indent=1)
        if name == "cc-pvdz":
            with open(os.path.join(HERE, "cc-pvdz-expanded.json"), "w") as f:
                json.dump(_expand(trimmed), f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
