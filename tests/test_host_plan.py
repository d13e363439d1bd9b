"""Host-side logic of the library, without a GPU: the capacity order of a
level, its slot and tile layout, the records' checks while they are planned,
the rand() stream the sampled shuffles replay (against glibc's own rand()),
and the staging copy of a level's tables.  tests/cpp/plan_layout_main.cpp
holds the checks; it is compiled from the library's own source by
__graft_entry__.build()."""
import os
import subprocess

EXE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "hbt_plan_test")


def test_level_planning_on_the_host():
    assert os.path.exists(EXE), f"{EXE} is missing: build it with __graft_entry__.build()"
    r = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
