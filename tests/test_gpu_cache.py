"""Binary cache of built formats (mmcsf_format_save / mmcsf_format_load):
two builds of one tensor write byte-identical files, and a reload writes the
same bytes again (every array of every tree round-trips bit for bit);
MTTKRP from a reloaded format against float64 (1e-5 relative, float32
factors); the error paths; the CLI's -k / -i; the nonzeros a format holds
(mmcsf_format_to_tensor); and the reference's launch-shape limits on -w/-s/-b."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from test_gpu_builder import random_tensor

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "mm-csf_amd", "bin", "mttkrp")

FORMATS = [(3, {}), (8, {}), (9, {}), (10, {}), (12, {}), (12, {"mode_trees": False}), (12, {"mixed": True}),
           (13, {}), (14, {})]


def rel(a, b):
    b = np.asarray(b, dtype=np.float64)
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / np.linalg.norm(b))


@pytest.mark.parametrize("dims,skew", [([500, 200, 80], [1.2, 1.2, None]), ([60, 30, 45, 12], [1.1, None, 1.1, None])])
@pytest.mark.parametrize("fmt_type,kw", FORMATS)
def test_cache_round_trip_bit_exact(cuda, tmp_path, dims, skew, fmt_type, kw):
    import torch
    import mmcsf
    rng = np.random.default_rng(fmt_type + len(dims))
    inds, vals = random_tensor(rng, dims, 20000, skew)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    mode = 1
    a, b, c = tmp_path / "a.fmt", tmp_path / "b.fmt", tmp_path / "c.fmt"
    f1 = mmcsf.Format(t, fmt_type, mode=mode, fbr_threshold=16, **kw)
    f1.save(a)
    mmcsf.Format(t, fmt_type, mode=mode, fbr_threshold=16, **kw).save(b)
    assert a.read_bytes() == b.read_bytes(), "two builds of one tensor differ"
    assert mmcsf.lib().mmcsf_format_is_cache(str(a).encode()) == 1
    f3 = mmcsf.Format.load(a)
    f3.save(c)
    assert c.read_bytes() == a.read_bytes(), "reload does not round-trip"
    i1, i3 = f1.info(), f3.info()
    assert i1 == i3
    # the reloaded format computes
    R = 32
    U = [rng.random((d, R)).astype(np.float32) for d in dims]
    Ud = [torch.from_numpy(u).to(cuda) for u in U]
    modes = range(len(dims)) if fmt_type in (9, 12, 13, 14) else [i1["modeOrder"][0]]
    for m in modes:
        got = mmcsf.mttkrp(f3, m, Ud).cpu().numpy()
        assert rel(got, O.mttkrp_f64_mt(inds, vals, m, U)) < 1e-5, f"mode {m}"
    # the nonzeros a format holds are the tensor's
    back = f3.to_tensor()
    _, bi, bv = back.to_numpy()
    want = sorted(zip(*[a.tolist() for a in inds], vals.tolist()))
    assert sorted(zip(*[a.tolist() for a in bi], bv.tolist())) == want


def test_cache_errors(cuda, tmp_path):
    import mmcsf
    rng = np.random.default_rng(3)
    dims = [50, 40, 30]
    inds, vals = random_tensor(rng, dims, 3000)
    f = mmcsf.Format(mmcsf.Tensor.from_numpy(dims, inds, vals), 12)
    p = tmp_path / "x.fmt"
    f.save(p)
    data = p.read_bytes()
    (tmp_path / "cut.fmt").write_bytes(data[:-100])
    with pytest.raises(mmcsf.MMCSFError) as e:
        mmcsf.Format.load(tmp_path / "cut.fmt")
    assert e.value.code == -3
    bad = bytearray(data)
    bad[8] = 99  # version
    (tmp_path / "ver.fmt").write_bytes(bytes(bad))
    with pytest.raises(mmcsf.MMCSFError, match="version"):
        mmcsf.Format.load(tmp_path / "ver.fmt")
    (tmp_path / "x.tns").write_text("3\n2 2 2\n1 1 1 1.0\n")
    with pytest.raises(mmcsf.MMCSFError, match="not a format cache"):
        mmcsf.Format.load(tmp_path / "x.tns")
    assert mmcsf.lib().mmcsf_format_is_cache(str(tmp_path / "x.tns").encode()) == 0
    with pytest.raises(mmcsf.MMCSFError):
        mmcsf.Format.load(tmp_path / "missing.fmt")


def test_cli_writes_and_reads_the_cache(cuda, golden_dir, tmp_path):
    fmt = tmp_path / "toy.fmt"
    r = subprocess.run([CLI, "-i", f"{golden_dir}/toy.tns", "-t", "12", "-R", "4", "-k", str(fmt)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert fmt.exists()
    r = subprocess.run([CLI, "-i", str(fmt), "-t", "12", "-R", "4", "-c", "1"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Total GPU time:" in r.stdout and "Correctness pass!" in r.stdout
    r = subprocess.run([CLI, "-i", str(fmt), "-t", "8", "-R", "4"], capture_output=True, text=True, timeout=120)
    assert r.returncode != 0 and "-t 12" in r.stderr


def test_launch_shape_limits_of_the_reference(cuda):
    import mmcsf
    rng = np.random.default_rng(4)
    dims = [50, 40, 30]
    inds, vals = random_tensor(rng, dims, 2000)
    t = mmcsf.Tensor.from_numpy(dims, inds, vals)
    o = mmcsf.default_options()
    o.warpPerSlice = 17  # 17 * 32 > 512
    with pytest.raises(mmcsf.MMCSFError, match="BLOCKSIZE is smaller than work per slice"):
        mmcsf.Format(t, 8, opts=o)
    o = mmcsf.default_options()
    o.warpPerSlice = 5  # > -b 128 / 32
    with pytest.raises(mmcsf.MMCSFError, match=r"cannot be higher than threadblock size"):
        mmcsf.Format(t, 12, opts=o)
    o = mmcsf.default_options()
    o.fiberPerWarp = 8
    with pytest.raises(mmcsf.MMCSFError, match=r"fbrPerWarp \(-s\)"):
        mmcsf.Format(t, 14, opts=o)
    o = mmcsf.default_options()
    o.warpPerSlice, o.TBsize = 8, 256  # allowed with a larger -b; the launch is the same
    mmcsf.Format(t, 12, opts=o)
