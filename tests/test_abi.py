"""C-ABI checks that need no GPU: the library loads, exports every symbol that
include/mmcsf.h declares, and parses options with the reference's grammar
(parse_cmd_options, util.h:3341-3442)."""
import ctypes as C
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mmcsf.h")
LIB = os.path.join(ROOT, "mm-csf_amd", "libmmcsf.so")
CLI = os.path.join(ROOT, "mm-csf_amd", "bin", "mttkrp")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mmcsf_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("mmcsf_format_build", "mmcsf_mttkrp", "mmcsf_mttkrp_allmode", "mmcsf_format_export",
                 "mmcsf_options_parse", "mmcsf_tensor_generate", "mmcsf_last_error"):
        assert must in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "libmmcsf.so is not built"
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\b(mmcsf_[a-z0-9_]+)\b", out))
    assert set(declared_functions()) <= exported


def test_oracle_library_loads():
    import oracle
    L = oracle.lib()
    for n in ("orc_build_csf", "orc_mttkrp_coo_f32", "orc_mttkrp_coo_f64", "orc_create_hyb"):
        assert hasattr(L, n)


def test_options_defaults():
    import mmcsf
    o = mmcsf.default_options()
    # Options member initialisers, util.h:189-211
    assert (o.R, o.mode, o.impType, o.warpPerSlice, o.nTile, o.gridSize, o.TBsize, o.MIfbTh,
            o.fiberPerWarp) == (32, 0, 1, 4, 1, 512, 128, 1, 1)
    assert o.fbrThreashold == 99999999 and o.nBin == 10 and o.cpdIters == 10
    assert o.verbose == 0 and o.correctness == 0 and o.natOrdering == 0
    assert o.m0 == b"012" and o.m1 == b"120" and o.m2 == b"201"


def test_options_parse_reference_flags():
    import mmcsf
    o = mmcsf.parse_options(["-i", "x.tns", "-m", "1", "-R", "64", "-t", "12", "-f", "128", "-w", "1",
                             "-b", "256", "-s", "2", "-h", "3", "-g", "99", "-v", "1", "-c", "1",
                             "-o", "out.txt", "-l", "2", "-d", "1"])
    assert o.inFileName == b"x.tns" and o.outFileName == b"out.txt"
    assert (o.mode, o.R, o.impType, o.fbrThreashold, o.warpPerSlice, o.TBsize, o.fiberPerWarp,
            o.MIfbTh, o.gridSize, o.nTile) == (1, 64, 12, 128, 1, 256, 2, 3, 99, 2)
    assert o.verbose == 1 and o.correctness == 1 and o.factorDtype == 1
    # -v / -c take 1 as true and anything else as false
    o = mmcsf.parse_options(["-v", "2", "-c", "0"])
    assert o.verbose == 0 and o.correctness == 0
    # atoi semantics: a negative value wraps like the reference's size_t fields
    o = mmcsf.parse_options(["-f", "-1"])
    assert o.fbrThreashold == 2**64 - 1


def test_options_parse_stops_at_first_non_flag():
    import mmcsf
    o = mmcsf.parse_options(["-R", "8", "positional", "-R", "16"])
    assert o.R == 8


@pytest.mark.parametrize("argv", [["-Q", "1"], ["-R"], ["-i", "a", "-m"]])
def test_options_parse_errors(argv):
    import mmcsf
    with pytest.raises(mmcsf.MMCSFError) as e:
        mmcsf.parse_options(argv)
    assert e.value.code == -1


def test_cli_error_paths_without_gpu(tmp_path):
    # argument errors are reported before any device work, with a non-zero exit
    assert os.path.exists(CLI)
    r = subprocess.run([CLI, "-Q", "1"], capture_output=True, text=True)
    assert r.returncode != 0 and "unknown option" in r.stderr
    r = subprocess.run([CLI, "-R", "8"], capture_output=True, text=True)
    assert r.returncode != 0 and "Provide input file path" in r.stdout
    r = subprocess.run([CLI, "--help"], capture_output=True, text=True)
    assert r.returncode == 0 and "-t implementation type" in r.stdout
