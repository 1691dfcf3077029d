"""Builds a CPU emulation of the HIP sources under AddressSanitizer
(test infrastructure): rewrites each `kernel<<<grid, block, shmem, stream>>>(args)`
launch into `::emu::launch(grid, block, kernel, args)` and compiles with g++.

Usage: python tests/emu/build_emu.py OUT SOURCE.hip [SOURCE.hip ...]
  OUT ending in .so builds a shared library; otherwise an executable that
  links emu_main.cpp and the oracle sources.
"""
import os
import re
import subprocess
import sys
import tempfile
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE.parents[1] / "del3d_amd" / "csrc"


def split_top(s):
    parts, depth, cur = [], 0, ""
    for ch in s:
        if ch in "([<{":
            depth += 1
        elif ch in ")]>}":
            depth -= 1
        if ch == "," and depth == 0:
            parts.append(cur.strip())
            cur = ""
        else:
            cur += ch
    parts.append(cur.strip())
    return parts


def rewrite(src: str) -> str:
    out, i = [], 0
    pat = re.compile(r"([A-Za-z_][\w:]*(?:<[^<>()]*>)?)\s*<<<(.*?)>>>\s*\(", re.S)
    while True:
        m = pat.search(src, i)
        if not m:
            out.append(src[i:])
            break
        out.append(src[i:m.start()])
        cfg = split_top(m.group(2))
        out.append(f"::emu::launch(dim3({cfg[0]}), dim3({cfg[1]}), {m.group(1)}, ")
        i = m.end()
    text = "".join(out)
    return text.replace(", )", ")")


def main():
    out = sys.argv[1]
    with tempfile.TemporaryDirectory() as td:
        srcs = []
        for s in sys.argv[2:]:
            p = Path(s)
            t = Path(td) / (p.stem + ".cpp")
            t.write_text(f'#include "{CSRC}/common.h"\n' + rewrite(p.read_text()).replace(
                '#include "common.h"', "").replace('#include "pred_core.h"', f'#include "{CSRC}/pred_core.h"').replace(
                '#include "pred2_core.h"', f'#include "{CSRC}/pred2_core.h"').replace(
                '#include "circ_core.h"', f'#include "{CSRC}/circ_core.h"').replace(
                '#include "predicates.h"', f'#include "{CSRC}/predicates.h"'))
            if p.name == "predicates.hip":
                # two sources define ScalarQuery; keep the translation units apart
                t.write_text(t.read_text())
            srcs.append(str(t))
        srcs.append(str(HERE / "emu_driver.cpp"))
        # EMU_SANITIZER=none builds without a sanitizer (the fiber mode, whose
        # switched stacks a sanitizer's stack poisoning does not follow);
        # EMU_OPT sets the optimisation level (default -O1)
        san = os.environ.get("EMU_SANITIZER", "address")
        opt = os.environ.get("EMU_OPT", "-O1")
        cmd = ["g++", "-std=c++20", "-g", opt, "-fPIC", "-pthread", "-ffp-contract=off",
               "-fno-omit-frame-pointer", f"-I{HERE}", f"-I{CSRC}", "-o", out] + srcs
        if san != "none":
            cmd.insert(6, f"-fsanitize={san}")
        if out.endswith(".so"):
            cmd.insert(1, "-shared")
        else:
            oracle = HERE.parents[1] / "oracle"
            objs = []
            for c in ("spatial_sort.c", "pointsets.c", "predicates.c", "exact_int.c"):
                o = Path(td) / (c + ".o")
                subprocess.run(["gcc", "-O2", "-std=gnu11", "-ffp-contract=off", "-c", str(oracle / c),
                                "-o", str(o)], check=True)
                objs.append(str(o))
            cmd += [str(HERE / "emu_main.cpp")] + objs + ["-lm"]
        subprocess.run(cmd, check=True)


if __name__ == "__main__":
    main()
