"""Convert the reference's plain-text basis files into this package's JSON form.

The reference keeps its basis sets in ``basis/<name>.dat`` (format read by
src/molecule.cpp:152-235: a ``Z nshell`` header, then per shell a
``S|P|D nprim`` line followed by ``exponent coefficient`` lines, ``#`` lines
being comments).  The numbers are the published STO-nG / Pople basis sets.
Copies of those files ship in ``dftcxx_amd/data/basis`` (read by the library's
input reader, dg_input_read); this recipe reads them and writes
``dftcxx_amd/data/basis/<name>.json`` as::

    {"name": ..., "elements": {"<Z>": [["S", [[alpha, c], ...]], ...]}}

keeping the shell order of the file, which fixes the CGF order (Q8 of the
reference: element in ascending Z, then atom, then shell).

Run:  python dftcxx_amd/data/make_basis.py [basis directory, default dftcxx_amd/data/basis]
"""
import json
import os
import sys


def parse_dat(path):
    """Parse a reference ``.dat`` basis file into {Z: [(type, [(a, c), ...]), ...]}."""
    with open(path) as f:
        lines = [ln for ln in f.read().splitlines() if ln and not ln.startswith("#")]
    out = {}
    i = 0
    while i < len(lines):
        z, nshell = (int(t) for t in lines[i].split()[:2])
        i += 1
        shells = []
        for _ in range(nshell):
            kind, nprim = lines[i].split()[:2]
            nprim = int(nprim)
            i += 1
            prims = []
            for _ in range(nprim):
                a, c = (float(t) for t in lines[i].split()[:2])
                prims.append([a, c])
                i += 1
            shells.append([kind[0], prims])
        out[z] = shells
    return out


def main():
    here = os.path.dirname(os.path.abspath(__file__))
    src = sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "basis")
    os.makedirs(os.path.join(here, "basis"), exist_ok=True)
    for name in ("sto3g", "sto6g", "p321", "p631"):
        data = parse_dat(os.path.join(src, name + ".dat"))
        doc = {"name": name,
               "elements": {str(z): data[z] for z in sorted(data) if z <= 18}}
        with open(os.path.join(here, "basis", name + ".json"), "w") as f:
            json.dump(doc, f, indent=None, separators=(",", ":"))
            f.write("\n")


if __name__ == "__main__":
    main()
