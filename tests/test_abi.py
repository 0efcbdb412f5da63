"""The C ABI library loads and exports every entry point include/graphit_amd.h
declares; host-side logic that needs no GPU. CPU only."""
import ctypes as C
import subprocess

import pytest

import easy_graphit_amd as gi


def test_library_loads_and_exports_header_symbols():
    L = gi.lib()
    syms = gi.header_symbols()
    assert len(syms) > 40
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    out = subprocess.run(["nm", "-D", "--defined-only", gi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    assert set(syms) <= exported


def test_version_and_device_count_without_gpu_do_not_fail():
    assert gi.lib().gi_version() >= 100
    assert gi.device_count() >= 0


def test_schedule_defaults_mirror_simple_gpu_schedule():
    s = gi.Schedule()
    # schedule.h:87-97: PUSH, VERTEX_BASED, UNBLOCKED, SPARSE, BOOLMAP pull rep
    assert s.direction == gi.PUSH and s.load_balancing == gi.LB_VERTEX_BASED
    assert s.edge_blocking == gi.UNBLOCKED and s.frontier_creation == gi.FC_SPARSE
    assert s.pull_frontier_rep == gi.PULL_BOOLMAP and s.block_size == -1
    assert s.cta_size == 512  # schedule.cpp:5


def test_null_arguments_report_errors_not_crash():
    L = gi.lib()
    assert L.gi_graph_num_vertices(None, None) == -1
    assert b"null" in L.gi_last_error()
    assert L.gi_frontier_size(None, None, None) == -1
    assert L.gi_graph_destroy(None) == 0


def test_bad_schedule_rejected_before_any_device_work():
    L = gi.lib()
    s = gi.Schedule(cta_size=100)
    h = gi.HybridSchedule(s, gi.Schedule())
    # graph handle is checked first: NULL graph -> INVALID
    assert L.gi_bfs(None, 0, C.byref(h), None, None, None, None) == -1
