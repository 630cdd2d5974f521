"""The C ABI boundary: the library loads, exports every function that
include/vo_api.h declares, and the ctypes mirror has the C struct layout.
No device calls (runs without a GPU)."""
import ctypes as C
import os
import re
import subprocess
import tempfile

import pytest

import velodb_amd
from velodb_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vo_api.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"VO_EXPORT\s+[\w\s\*]*?\b(vo_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = velodb_amd.load_library()
    names = _declared()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert missing == []


def test_ctypes_prototypes_cover_header():
    assert sorted(n for n, _, _ in _abi.PROTOTYPES) == _declared()


def test_no_torch_types_in_header():
    txt = open(HEADER).read()
    assert "torch" not in txt and "at::" not in txt and "c10" not in txt


STRUCTS = ["vo_sort_args", "vo_sort_kp_args", "vo_permute_args", "vo_scatter_args", "vo_gather_args",
           "vo_compare_eq_args", "vo_compose_args", "vo_enrich_args", "vo_compact_args", "vo_shuffle_args",
           "vo_filter_args", "vo_join_args", "vo_batch_item", "vo_sort_batch_args", "vo_command", "vo_config",
           "vo_trace", "vo_predicate_term", "vo_predicate_args", "vo_server_config", "vo_query",
           "vo_query_result", "vo_run_stats", "vo_compact_batch_args", "vo_join_pairs_args",
           "vo_gather_oblivious_args"]
PY = {"vo_sort_args": _abi.SortArgs, "vo_sort_kp_args": _abi.SortKpArgs, "vo_permute_args": _abi.PermuteArgs,
      "vo_scatter_args": _abi.ScatterArgs, "vo_gather_args": _abi.GatherArgs,
      "vo_compare_eq_args": _abi.CompareEqArgs, "vo_compose_args": _abi.ComposeArgs,
      "vo_enrich_args": _abi.EnrichArgs, "vo_compact_args": _abi.CompactArgs, "vo_shuffle_args": _abi.ShuffleArgs,
      "vo_filter_args": _abi.FilterArgs, "vo_join_args": _abi.JoinArgs, "vo_batch_item": _abi.BatchItem,
      "vo_sort_batch_args": _abi.SortBatchArgs, "vo_command": _abi.Command, "vo_config": _abi.Config,
      "vo_trace": _abi.Trace, "vo_predicate_term": _abi.PredicateTerm, "vo_predicate_args": _abi.PredicateArgs,
      "vo_server_config": _abi.ServerConfig, "vo_query": _abi.Query, "vo_query_result": _abi.QueryResult,
      "vo_run_stats": _abi.RunStats, "vo_compact_batch_args": _abi.CompactBatchArgs,
      "vo_join_pairs_args": _abi.JoinPairsArgs, "vo_gather_oblivious_args": _abi.GatherObliviousArgs}


def test_struct_layout_matches_c():
    src = '#include <stdio.h>\n#include "vo_api.h"\nint main(void){\n'
    for s in STRUCTS:
        src += f'printf("%s %zu\\n", "{s}", sizeof({s}));\n'
    src += 'printf("args_offset %zu\\n", __builtin_offsetof(vo_command, args));\nreturn 0;}\n'
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "p.c"), "w") as f:
            f.write(src)
        exe = os.path.join(d, "p")
        subprocess.check_call(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), os.path.join(d, "p.c"),
                               "-o", exe])
        out = dict(line.split() for line in subprocess.check_output([exe]).decode().split("\n") if line)
    for s in STRUCTS:
        assert int(out[s]) == C.sizeof(PY[s]), s
    assert int(out["args_offset"]) == _abi.Command.args.offset


def test_opcode_values_mirror_reference():
    # include/cuda/commands.hpp:16-31
    assert (_abi.OP_NOP, _abi.OP_SCATTER, _abi.OP_GATHER, _abi.OP_SORT, _abi.OP_PERMUTE) == (0, 1, 2, 3, 4)
    assert (_abi.OP_COMPARE_EQ, _abi.OP_COMPOSE_POSITION_MAP, _abi.OP_ENRICH_POSITIONS) == (10, 11, 12)
    assert _abi.OP_TERMINATE == 0xFFFFFFFF


def test_status_strings_without_device():
    lib = velodb_amd.load_library()
    assert lib.vo_version() == 1
    assert lib.vo_status_string(0) == b"ok"
    assert lib.vo_status_string(3) == b"timeout"


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(ImportError):
        velodb_amd.load_library(str(tmp_path / "absent.so"))


def test_cpp_operator_header_compiles():
    """The header-only C++ operator layer builds against the C ABI (syntax
    and templates; nothing runs without a GPU)."""
    import subprocess
    src = os.path.join(ROOT, "include", "velodb_amd", "operators.hpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
                        "-I/opt/rocm/include", "-I" + os.path.join(ROOT, "include"), "-x", "c++", src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
