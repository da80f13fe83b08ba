"""Write tests/golden/sw_layout.json: sizeof and offsetof of every member of
SW2's `map` (sppde/sppde.h:46-67) and `sw` (sw.h:22-95) structs, measured by
compiling a small C program against SW2's own headers where the reference
sources are (REF, default /root/reference).

    python tests/golden/make_layout.py [REF]

tests/test_abi.py compiles the same program against include/sw2_hip.h and
compares, so an object built against SW2's headers reads the library's
structs at the right offsets.
"""
import json
import os
import subprocess
import sys
import tempfile

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sw_layout.json")

MAP_FIELDS = ["sh", "nd", "l0", "l1", "gl0", "gl1", "al0", "al1", "pco", "s", "S", "gpc", "pc", "nb_e", "nb_w", "nb_n",
              "nb_s", "nb_t", "nb_b", "bc0", "bc1", "per"]
SW_FIELDS = ["sch", "polar", "N", "loadFrom", "saveEvery", "IteInfo", "prn", "t0", "t1", "Dt", "Courant_max", "nx", "ny",
             "nz", "lon0", "lat0", "lon1", "lat1", "m", "M", "Omega", "rE", "rP", "epsilon2", "gE", "depth", "sigma_dis",
             "tracer_dis", "beta_sponge", "sigma_sponge", "tau_sponge", "windx", "windy", "u", "v", "eta", "hB", "tracer",
             "perturbs", "xe", "xc", "xv", "yn", "yc", "yv", "Dxe", "Dxc", "Dxv", "Dyn", "Dyc", "Dyv", "ge", "gn", "gc",
             "gv", "fe", "fn", "fc", "fv", "nu2", "nu4", "nu6", "reftype", "refname", "reftable", "reftable_nr",
             "reftable_nc", "geoeq", "num_Gaussians", "Gaussians", "num_vortices", "vortices"]


def program(includes):
    """C source printing the layout as JSON, after the given #include lines."""
    lines = ["#include <stdio.h>", "#include <stdlib.h>", "#include <math.h>", "#include <stddef.h>"] + includes
    lines += ["int main(void) {", '    printf("{\\"map\\": {\\"sizeof\\": %zu", sizeof(map));']
    lines += [f'    printf(", \\"{f}\\": %zu", offsetof(map, {f}));' for f in MAP_FIELDS]
    lines += ['    printf("}, \\"sw\\": {\\"sizeof\\": %zu", sizeof(sw));']
    lines += [f'    printf(", \\"{f}\\": %zu", offsetof(sw, {f}));' for f in SW_FIELDS]
    lines += ['    printf("}}\\n");', "    return 0;", "}"]
    return "\n".join(lines) + "\n"


def measure(includes, incdirs):
    """Compile and run the layout program; returns the parsed layout."""
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as f:
            f.write(program(includes))
        subprocess.run(["gcc", "-std=gnu99", "-o", exe, src] + [f"-I{i}" for i in incdirs], check=True)
        return json.loads(subprocess.run([exe], check=True, capture_output=True, text=True).stdout)


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    lay = measure(['#include "sppde.h"', '#include "sw.h"'], [ref, os.path.join(ref, "sppde")])
    lay["_about"] = {"what": "sizeof/offsetof of SW2's map and sw, gcc on x86-64", "script": "tests/golden/make_layout.py"}
    with open(OUT, "w") as f:
        json.dump(lay, f, indent=1)
    print(f"map {lay['map']['sizeof']} bytes, sw {lay['sw']['sizeof']} bytes -> {OUT}")


if __name__ == "__main__":
    main()
