"""Mutation tests of the walk checks: copies of oracle/mb_oracle.c with one
piece of physics.wgsl's neighbour walk removed or altered are compiled on the
CPU, and tests/walk_checks.py, run against each copy in a subprocess, must fail
(a check that passes a broken walk pins nothing)."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SRC = ROOT / "oracle" / "mb_oracle.c"

MUTANTS = {
    "top_row_removed": ("    if (cz < G - 1u) handle_collision(x, a, cx, k + G);\n", ""),
    "bottom_row_removed": ("    if (cz > 0u) handle_collision(x, a, cx, k - G);\n", ""),
    "middle_right_cell_removed": (
        "            if (st != MBO_INVALID) last = rd_close(x, k + 1u);",
        "            (void)st;"),
    "middle_left_cell_removed": (
        "            if (st != MBO_INVALID) first = st;\n        }\n        if (cx < G - 1u) {",
        "            (void)st;\n        }\n        if (cx < G - 1u) {"),
    "row_right_cell_removed": (
        "    if (cx < G - 1u) {\n        uint32_t st = rd_start(x, base + 1u);",
        "    if (0) {\n        uint32_t st = rd_start(x, base + 1u);"),
    "row_left_cell_removed": (
        "    if (cx > 0u) {\n        uint32_t st = rd_start(x, base - 1u);",
        "    if (0) {\n        uint32_t st = rd_start(x, base - 1u);"),
    "left_edge_guard_dropped": (
        "        if (cx > 0u) {\n            uint32_t st = rd_start(x, k - 1u);",
        "        if (1) {\n            uint32_t st = rd_start(x, k - 1u);"),
    "right_edge_guard_dropped": (
        "        if (cx < G - 1u) {\n            uint32_t st = rd_start(x, k + 1u);",
        "        if (1) {\n            uint32_t st = rd_start(x, k + 1u);"),
    "cell_clamped": (
        "    *cx = f32_to_u32_sat(x + c->half_map_size) / c->grid_size;\n",
        "    *cx = f32_to_u32_sat(x + c->half_map_size) / c->grid_size;\n"
        "    if (*cx > c->grids - 1u) *cx = c->grids - 1u;\n"),
    "row_first_nonempty_rule_broken": (
        "            if (first == MBO_INVALID) first = st;\n            last = rd_close(x, base);",
        "            first = st;\n            last = rd_close(x, base);"),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_mutant_fails_walk_checks(name, tmp_path):
    old, new = MUTANTS[name]
    src = SRC.read_text()
    assert src.count(old) == 1, f"mutation site of {name} not found once in mb_oracle.c"
    mut = tmp_path / "mb_oracle_mut.c"
    mut.write_text(src.replace(old, new))
    lib = tmp_path / "libmb_oracle_mut.so"
    subprocess.run(["gcc", "-O2", "-fPIC", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
                    "-fopenmp", "-shared", "-I", str(SRC.parent), "-o", str(lib), str(mut), "-lm"],
                   check=True)
    env = dict(os.environ, MB_ORACLE_LIB=str(lib), OMP_NUM_THREADS="2")
    r = subprocess.run([sys.executable, str(ROOT / "tests" / "walk_checks.py")], env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, f"mutant {name} passed every walk check:\n{r.stdout}\n{r.stderr}"
    assert "FAIL" in r.stdout
