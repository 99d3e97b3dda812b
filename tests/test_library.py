"""The C-ABI library loads and exports every declared entry point, and the
command line rejects bad input with the reference's exit codes (CPU only;
no compute call is made)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INCLUDE = os.path.join(ROOT, "include")


def declared_functions():
    names = set()
    for h in sorted(os.listdir(INCLUDE)):
        if not h.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"static inline[^{]*\{.*?\n\}", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b([a-z_][a-z0-9_]*)\s*\([^;{]*\)\s*;", text, flags=re.M):
            if m.group(0).lstrip().startswith("typedef"):
                continue
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_reference_phase_api():
    names = declared_functions()
    for ref in ("perform_edge_pull_phase_pr", "perform_edge_push_phase_pr", "perform_vertex_phase_pr",
                "perform_edge_pull_phase_cc", "perform_edge_push_phase_cc", "perform_vertex_phase_cc",
                "perform_edge_pull_phase_bfs", "perform_edge_push_phase_bfs", "perform_vertex_phase_bfs",
                "phase_op_reset_global_accum", "phase_op_write_global_accum_to_buf",
                "phase_op_combine_global_var_from_buf", "edge_pull_op_merge_with_merge_buffer",
                "graph_data_read_from_file", "graph_data_allocate_merge_buffers",
                "graph_data_write_ranks_to_file"):
        assert ref in names


def test_library_exports_every_declared_symbol(gz):
    lib = ctypes.CDLL(gz.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    for var in ("gz_graph", "gz_phase_config", "gz_merge_buffer", "gz_merge_buffer_count"):
        assert hasattr(lib, var)


def exe(gz):
    return os.path.join(os.path.dirname(gz.LIB_PATH), "grazelle")


@pytest.mark.parametrize("args,code", [
    (["-x"], 1), (["-ii", "g"], 1), (["-N", "0", "-i", "g"], 2), (["-N", "abc", "-i", "g"], 2),
    (["-s", "0", "-i", "g"], 2), (["-u", "3", "-i", "g"], 2), (["-a", "sssp", "-i", "g"], 2),
    (["-d", "1.5", "-i", "g"], 2), (["-i"], 3), (["-N"], 3), (["-p", "1", "-i", "g"], 4), ([], 5),
    (["-N", "3"], 5), (["-E", "-X", "-i", "g"], 6), (["-e", "learned", "-i", "g"], 6),
    (["-L", "x", "-i", "g"], 2), (["-L", "0", "-i", "g"], 2), (["-L", "24", "-i", "g"], 2),
    (["-L", "20,4", "-i", "g"], 2), (["-L", "20,2,6", "-i", "g"], 2), (["-L", "20,", "-i", "g"], 2),
    (["-L", "20,1,1,1", "-i", "g"], 2), (["-L", "-i", "g"], 3), (["-L", "20", "-T", "-i", "g"], 6),
    (["-L", "20", "-M", "-i", "g"], 6),
])
def test_cli_error_codes(gz, args, code):
    r = subprocess.run([exe(gz)] + args, capture_output=True, text=True, timeout=60)
    assert r.returncode == code, r.stdout + r.stderr
    assert "Try `" in r.stdout


def test_cli_missing_input_file(gz, tmp_path):
    r = subprocess.run([exe(gz), "-i", str(tmp_path / "none")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255
    assert "unable to read file" in r.stderr
    # a present -pull with a truncated header is unreadable too
    (tmp_path / "t-pull").write_bytes(b"\x01\x00")
    r = subprocess.run([exe(gz), "-i", str(tmp_path / "t")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255


def test_cli_help_and_version(gz):
    r = subprocess.run([exe(gz), "-h"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage:" in r.stdout
    r = subprocess.run([exe(gz), "-V"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "grazelle" in r.stdout
