#!/usr/bin/env python3
"""Generate ``lewiner_tables.h`` -- the Lewiner et al. (2003) marching-cubes look-up tables.

The reference reads these tables from MarchingCubes.jl 0.1.11 (an un-vendored dependency,
``Manifest.toml:948-952``; call sites ``src/marching_cubes/marching_cubes.jl:442-448, 540-650`` and
``src/marching_cubes/moments.jl:340``).  That package's tables are Lewiner's ``LookUpTable.h`` with the
tiling edge codes shifted to 1-based.  The same ``LookUpTable.h`` values ship, as base64 ``int8``
text, in scikit-image's ``_marching_cubes_lewiner_luts.py`` (BSD-3-Clause); this script decodes that
*text* with ``base64``/``numpy`` (nothing is imported from scikit-image) and writes a C header.

Everything is kept 0-based, exactly as ``LookUpTable.h`` spells it:

* ``LW_CASES[lut][0..1]`` = (Lewiner case 0..14, configuration), ``lut`` = the 8-bit corner mask;
* tiling codes 0..11 are cube edges, 12 is the interior vertex;
* test tables hold signed face codes (+-1..6), the interior sign (+-7) and 0-based reference edges.

The dispatch in the kernels adds nothing to these numbers; the reference's ``case + 1`` / ``cfg + 1``
1-based indexing is undone at the call sites instead.

Usage::

    python gen_lewiner_tables.py [path/to/_marching_cubes_lewiner_luts.py] > lewiner_tables.h
"""
import base64
import re
import sys

import numpy as np

DEFAULT_SRC = "/opt/conda/lib/python3.9/site-packages/skimage/measure/_marching_cubes_lewiner_luts.py"

# The tables the reference's dispatch reads, in LookUpTable.h order.
WANTED = [
    "CASES", "TILING1", "TILING2", "TILING3_1", "TILING3_2", "TILING4_1", "TILING4_2",
    "TILING5", "TILING6_1_1", "TILING6_1_2", "TILING6_2", "TILING7_1", "TILING7_2", "TILING7_3",
    "TILING7_4_1", "TILING7_4_2", "TILING8", "TILING9", "TILING10_1_1", "TILING10_1_1_",
    "TILING10_1_2", "TILING10_2", "TILING10_2_", "TILING11", "TILING12_1_1", "TILING12_1_1_",
    "TILING12_1_2", "TILING12_2", "TILING12_2_", "TILING13_1", "TILING13_1_", "TILING13_2",
    "TILING13_2_", "TILING13_3", "TILING13_3_", "TILING13_4", "TILING13_5_1", "TILING13_5_2",
    "TILING14", "TEST3", "TEST4", "TEST6", "TEST7", "TEST10", "TEST12", "TEST13", "SUBCONFIG13",
]


def decode(text):
    pat = re.compile(r'^([A-Z0-9_]+)\s*=\s*\(([0-9, ]*)\),\s*"""(.*?)"""', re.S | re.M)
    out = {}
    for m in pat.finditer(text):
        shape = tuple(int(s) for s in m.group(2).replace(" ", "").split(",") if s)
        raw = base64.b64decode("".join(m.group(3).split()))
        out[m.group(1)] = np.frombuffer(raw, dtype=np.int8).reshape(shape)
    return out


def c_array(name, a):
    dims = "".join(f"[{s}]" for s in a.shape)
    flat = a.reshape(-1)
    body = []
    per = 24
    for i in range(0, flat.size, per):
        body.append("    " + ", ".join(str(int(v)) for v in flat[i:i + per]) + ",")
    return f"CC_LUT_QUAL signed char LW_{name}{dims} = {{\n" + "\n".join(body) + "\n};\n"


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_SRC
    with open(path) as f:
        tabs = decode(f.read())
    missing = [w for w in WANTED if w not in tabs]
    if missing:
        raise SystemExit(f"missing tables: {missing}")
    # Sanity pins that hold for LookUpTable.h (and for the reference's own golden cube, lut 60 ->
    # case 10, configuration 5, 0-based; test/test_mc_moments.jl:486-494).
    assert tuple(tabs["CASES"][60]) == (10, 5)
    assert tabs["CASES"][0][0] == 0 and tabs["CASES"][255][0] == 0
    assert (tabs["TEST13"][:, 5] > 0).all() and (tabs["TEST13"][:, 6] > 0).all()

    print("/* Lewiner et al. (2003) marching-cubes look-up tables, 0-based as in LookUpTable.h.")
    print(" *")
    print(" * GENERATED by gen_lewiner_tables.py from the base64 text of scikit-image's")
    print(" * _marching_cubes_lewiner_luts.py -- do not edit by hand.")
    print(" *")
    print(" * Table data: T. Lewiner, H. Lopes, A. W. Vieira, G. Tavares, \"Efficient implementation of")
    print(" * Marching Cubes' cases with topological guarantees\", J. Graphics Tools 8(2), 2003.")
    print(" * Transcribed text: scikit-image, Copyright (C) 2019, the scikit-image team, BSD-3-Clause:")
    print(" *   Redistribution and use in source and binary forms, with or without modification, are")
    print(" *   permitted provided that the following conditions are met: 1. Redistributions of source")
    print(" *   code must retain the above copyright notice, this list of conditions and the following")
    print(" *   disclaimer. 2. Redistributions in binary form must reproduce the above copyright notice,")
    print(" *   this list of conditions and the following disclaimer in the documentation and/or other")
    print(" *   materials provided with the distribution. 3. Neither the name of skimage nor the names")
    print(" *   of its contributors may be used to endorse or promote products derived from this")
    print(" *   software without specific prior written permission.")
    print(" *   THIS SOFTWARE IS PROVIDED BY THE AUTHOR ``AS IS'' AND ANY EXPRESS OR IMPLIED WARRANTIES")
    print(" *   ARE DISCLAIMED. IN NO EVENT SHALL THE AUTHOR BE LIABLE FOR ANY DAMAGES ARISING IN ANY")
    print(" *   WAY OUT OF THE USE OF THIS SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.")
    print(" *")
    print(" * The includer defines CC_LUT_QUAL (e.g. `static const` on the host, `__constant__` on the")
    print(" * device) before including this file.")
    print(" */")
    print("#ifndef CC_LUT_QUAL")
    print("#define CC_LUT_QUAL static const")
    print("#endif")
    print()
    for w in WANTED:
        print(c_array(w, tabs[w]))


if __name__ == "__main__":
    main()
