"""The C-ABI library: loads without a GPU and exports every entry point that
include/swf.h declares; the Python mirror's structures match the header."""
import ctypes
import os
import re
import subprocess

import sweeping_fd_weno_amd as S

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "swf.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(swf(?:1d|7)?_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_boundary():
    names = declared()
    for n in ("swf_create", "swf_destroy", "swf_upload", "swf_download", "swf_step", "swf_advance", "swf_rhs",
              "swf_alpha_max", "swf_check_positivity", "swf_sweep", "swf_bc", "swf_stats", "swf_track",
              "swf_weno_points", "swf_time_continue", "swf_strerror", "swf1d_create", "swf1d_destroy",
              "swf1d_upload", "swf1d_download", "swf1d_advance", "swf1d_stats", "swf1d_track", "swf7_create",
              "swf7_destroy", "swf7_upload", "swf7_download", "swf7_bc", "swf7_advance", "swf7_stats", "swf7_track"):
        assert n in names


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(S.library_path)
    for n in declared():
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", S.library_path], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    assert set(declared()) <= exported


def test_config_struct_layout_matches_header():
    # swf_config: 2 int32, 9 doubles ... as declared; the ctypes mirror must have
    # the same size and field order as the C struct (checked by offsets)
    names = [f[0] for f in S.SwfConfig._fields_]
    text = open(HEADER).read()
    body = text[text.index("typedef struct swf_config"):text.index("} swf_config;")]
    decl = re.findall(r"(?:int32_t|double)\s+([^;]+);", body)
    flat = [v.strip() for d in decl for v in d.split(",")]
    assert names == flat


def test_line_config_struct_layout_matches_header():
    names = [f[0] for f in S.Swf1dConfig._fields_]
    text = open(HEADER).read()
    body = text[text.index("typedef struct swf1d_config"):text.index("} swf1d_config;")]
    decl = re.findall(r"(?:int32_t|double)\s+([^;]+);", body)
    flat = [v.strip() for d in decl for v in d.split(",")]
    assert names == flat


def test_detonation_config_struct_layout_matches_header():
    from sweeping_fd_weno_amd.detonation import Swf7Config
    names = [f[0] for f in Swf7Config._fields_]
    text = open(HEADER).read()
    body = text[text.index("typedef struct swf7_config"):text.index("} swf7_config;")]
    decl = re.findall(r"(?:int32_t|double)\s+([^;]+);", body)
    flat = [v.strip() for d in decl for v in d.split(",")]
    assert names == flat


def test_host_helpers_without_device():
    lib = S.lib()
    assert S.time_continue(0.0, 1.0) and not S.time_continue(1.0, 1.0)
    assert lib.swf_strerror(1).decode().startswith("ERROR: P(u_bar)")
