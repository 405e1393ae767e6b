"""The reference application's problems and its command line.

Generator (ElectroMag/src/Electromag utils.h:29-111, built-in glibc rand()):
* rand() equals the host glibc's for many seeds, including seeds >= 2^31
  (stored as negative words), and the published first values after srand(1);
* charges and grid seeds are bit-identical to the reference's own
  InitializePointChargeArray / InitializeFieldLineArray (golden vectors
  written by the compiled reference; live where the reference is present).

Command line (electromag-with-cuda_amd/bin/electromag, ElectroMag.cpp flags):
* CPU: --gpudisable --bogo --save writes data.txt in the reference's format,
  text-identical to the oracle's rows and within FP32 reordering of the
  reference's own rows for the same preset; --cpuprecision=double matches the
  FP64 restatement; --postrundebug answers queries and ends at end of input;
  without a device the GPU mode fails with the library's status (no CPU
  substitute).
* GPU: --bogo --cpu --save --autoregress: GPU rows identical to the device
  path's REF rule, CPU rows to the oracle, and no line flagged; the default
  preset traced on the GPU is no further from FP64 than the reference's rows.
"""
import ctypes
import ctypes.util
import os
import subprocess

import numpy as np
import pytest

from emtrace import legacy, rules

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# EMTRACE_CLI: another build of the same command line (the sanitizer build)
CLI = os.environ.get("EMTRACE_CLI") or os.path.join(ROOT, "electromag-with-cuda_amd", "bin",
                                                    "electromag")
GOLD = os.path.join(ROOT, "tests", "golden")


def _glibc():
    try:
        libc = ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")
        libc.gnu_get_libc_version.restype = ctypes.c_char_p
        libc.gnu_get_libc_version()
        return libc
    except (OSError, AttributeError):
        return None


SEEDS = [0, 1, 2, 12345, 2**31 - 1, 2**31, 2**31 + 3, 3000000000, 2**32 - 1]


@pytest.mark.skipif(_glibc() is None, reason="host C library is not glibc")
def test_rand_matches_host_glibc():
    libc = _glibc()
    libc.srand.argtypes = [ctypes.c_uint]
    for s in SEEDS:
        libc.srand(s)
        want = np.array([libc.rand() for _ in range(5000)], np.int32)
        assert np.array_equal(legacy.rand(s, 5000), want), s


def test_rand_published_values():
    # glibc's first rand() values after srand(1) (and srand(0) == srand(1))
    first = [1804289383, 846930886, 1681692777, 1714636915, 1957747793]
    assert legacy.rand(1, 5).tolist() == first
    assert legacy.rand(0, 5).tolist() == first
    r = legacy.rand(7, 100000)
    assert r.min() >= 0 and r.max() <= 2**31 - 1


def test_init_matches_reference_golden():
    g = np.load(os.path.join(GOLD, "legacy_init.npz"))
    w = legacy.problem(1, 1, 1, 1024, 2)
    assert np.array_equal(w.charges.view(np.uint32), g["charges_p1024"].view(np.uint32))
    for key in g.files:
        if not key.startswith("grid_"):
            continue
        wdt, hgt, dep = (int(v) for v in key[5:].split("x"))
        got = legacy.problem(wdt, hgt, dep, 0, 2).seeds
        assert np.array_equal(got.view(np.uint32), g[key].view(np.uint32)), key


def test_init_matches_live_reference(orc):
    if orc.ref_lib() is None:
        pytest.skip("reference build absent (golden vectors cover this)")
    for p in (1, 3, 333, 5120):
        assert np.array_equal(legacy.problem(1, 1, 1, p, 2).charges.view(np.uint32),
                              orc.ref_init_charges(p).view(np.uint32)), p
    for shape in ((1, 1, 1), (3, 2, 5), (64, 64, 1), (256, 112, 1)):
        got = legacy.problem(*shape, 0, 2).seeds
        assert np.array_equal(got.view(np.uint32), orc.ref_init_grid(*shape).view(np.uint32))


def test_presets_match_the_reference_table():
    # ElectroMag.cpp:43-58 {nx, ny, nz, pStatic, pDynamic, len}
    assert legacy.PRESETS["default"] == (128, 128, 1, 1024, 2500)
    assert legacy.PRESETS["bogo"] == (16, 16, 1, 50, 500)
    assert legacy.PRESETS["fuckingInsane"] == (1024, 1024, 1, 5120, 10000)
    w = legacy.preset("micro")
    assert (w.p, w.n, w.len) == (1000, 256, 1000)


def _run(args, cwd, stdin=None, timeout=300):
    if not os.path.exists(CLI):
        pytest.fail(f"{CLI} not built (make -C electromag-with-cuda_amd)")
    return subprocess.run([CLI] + args, cwd=cwd, input=stdin, capture_output=True, text=True,
                          timeout=timeout)


def _expected_text(tag, rows, fmt="%g"):
    _, length, n = rows.shape
    return [f" {tag} [%d][%d] x: {fmt} y: {fmt} z: {fmt}" % (l, s, rows[0, s, l], rows[1, s, l],
                                                            rows[2, s, l])
            for l in range(n) for s in range(length)]


def _split(lines, tag):
    return [s for s in lines if s.startswith(f" {tag} ")]


def test_cli_cpu_bogo_save(orc, tmp_path):
    r = _run(["--gpudisable", "--bogo", "--save", "--nodisp"], tmp_path)
    assert r.returncode == 0, r.stderr
    assert "CPU kernel execution time" in r.stdout and " Save procedure complete" in r.stdout
    text = open(tmp_path / "data.txt").read().splitlines()
    w = legacy.preset("bogo")
    rc, rows = orc.calc_lines(w.charges, w.seeds, w.len, rules.ref(1.0))
    assert rc == 0
    assert text == _expected_text("CPUL", rows)
    # the reference's own rows for this preset (golden subset)
    g = np.load(os.path.join(GOLD, "legacy_bogo_ref.npz"))
    rs = int(g["row_stride"])
    sub = rows[:, ::rs, :][:, :, g["lines"]].astype(np.float64)
    assert np.abs(sub - g["rows"]).max() <= 2e-6 * np.abs(g["rows"]).max()


def test_cli_cpu_double_precision(orc, tmp_path):
    r = _run(["--gpudisable", "--bogo", "--save", "--nodisp", "--cpuprecision=double"], tmp_path)
    assert r.returncode == 0, r.stderr
    text = open(tmp_path / "data.txt").read().splitlines()
    w = legacy.preset("bogo")
    rc, rows = orc.calc_lines(w.charges.astype(np.float64), w.seeds.astype(np.float64), w.len,
                              rules.ref(1.0), dtype=np.float64)
    assert rc == 0
    assert text == _expected_text("CPUL", rows)


def test_cli_postrundebug_and_flags(tmp_path):
    r = _run(["--gpudisable", "--bogo", "--nodisp", "--postrundebug", "--nonsense"], tmp_path,
             stdin="0 0\n5 499\n256 0\nfoo\n3 1\n")
    assert r.returncode == 0, r.stderr
    out = r.stdout.splitlines()
    assert " Ignoring unknown argument: --nonsense" in out
    assert any(s.startswith(" CPUL [0][0] x: -79.9999 y: -79.9999 z: -4.9999") for s in out)
    assert any(s.startswith(" CPUL [5][499] x: ") for s in out)
    assert any(s.startswith(" Out of range") for s in out)
    assert any(s.startswith(" CPUL [3][1] x: ") for s in out)


def test_cli_seeded_random_modes_repeat(tmp_path):
    a = _run(["--gpudisable", "--bogo", "--nodisp", "--randseed", "--randfieldinit", "--seed=42",
              "--postrundebug"], tmp_path, stdin="7 100\n")
    b = _run(["--gpudisable", "--bogo", "--nodisp", "--randseed", "--randfieldinit", "--seed=42",
              "--postrundebug"], tmp_path, stdin="7 100\n")
    assert a.returncode == 0 and b.returncode == 0
    pa = [s for s in a.stdout.splitlines() if s.startswith(" CPUL")]
    assert pa and pa == [s for s in b.stdout.splitlines() if s.startswith(" CPUL")]
    assert " Charge seed:\t42" in a.stdout and " Line seed:\t43" in a.stdout


def test_cli_gpu_mode_without_device_fails_loudly(tmp_path):
    import emtrace
    if emtrace.lib().em_device_count() > 0:
        pytest.skip("a device is visible")
    r = _run(["--bogo", "--nodisp"], tmp_path)
    assert r.returncode == 19  # EM_ERR_NO_DEVICE: no CPU substitute
    assert "GPU tracer failed" in r.stderr
    assert not (tmp_path / "data.txt").exists()


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
@pytest.mark.gpu
def test_cli_gpu_bogo_compare_save_regress(gpu, orc, tmp_path):
    import torch
    r = _run(["--bogo", "--cpu", "--save", "--autoregress", "--nodisp"], tmp_path)
    print(r.stdout[-600:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "GPU kernel execution time" in r.stdout and "Realistic speedup" in r.stdout
    lines = open(tmp_path / "data.txt").read().splitlines()
    w = legacy.preset("bogo")
    dev = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(), w.len,
                    rules.ref(1.0)).cpu().numpy()
    assert _split(lines, "GPUL") == _expected_text("GPUL", dev)
    rc, rows = orc.calc_lines(w.charges, w.seeds, w.len, rules.ref(1.0))
    assert _split(lines, "CPUL") == _expected_text("CPUL", rows)
    # interleaving: for each (line, step) the CPU point, then the GPU point
    assert lines[0].startswith(" CPUL [0][0]") and lines[1].startswith(" GPUL [0][0]")
    reg = open(tmp_path / "regression.txt").read()
    assert reg == "", reg[:500]


@pytest.mark.gpu
def test_legacy_default_preset_vs_reference_golden(gpu, orc):
    """The reference's default problem (1,024 charges, 128 x 128 grid, 2,500
    rows): all 16,384 lines traced on the GPU; the 64-line golden subset is
    no further from FP64 than the reference's own FP32 rows."""
    import torch
    w = legacy.preset("default")
    dev = gpu.trace(torch.from_numpy(w.charges).cuda(), torch.from_numpy(w.seeds).cuda(), w.len,
                    rules.ref(1.0)).cpu().numpy()
    assert np.isfinite(dev).all()
    g = np.load(os.path.join(GOLD, "legacy_default_ref.npz"))
    rs, idx = int(g["row_stride"]), g["lines"]
    rc, r64 = orc.calc_lines(w.charges.astype(np.float64), w.seeds[idx].astype(np.float64), w.len,
                             rules.ref(1.0), dtype=np.float64)
    r64 = r64[:, ::rs, :]
    e_gpu = np.abs(dev[:, ::rs, :][:, :, idx] - r64).max(axis=(0, 1))
    e_ref = np.abs(g["rows"] - r64).max(axis=(0, 1))
    print(f"[legacy default] vs FP64: gpu median {np.median(e_gpu):.3e} p90 "
          f"{np.quantile(e_gpu, 0.9):.3e} max {e_gpu.max():.3e}; reference median "
          f"{np.median(e_ref):.3e} p90 {np.quantile(e_ref, 0.9):.3e} max {e_ref.max():.3e}",
          flush=True)
    assert np.median(e_gpu) <= 2 * np.median(e_ref)
    assert np.quantile(e_gpu, 0.9) <= 2 * np.quantile(e_ref, 0.9)


@pytest.mark.gpu
def test_cli_gpu_default_preset(gpu, tmp_path):
    r = _run(["--nodisp"], tmp_path, timeout=600)
    print(r.stdout[-800:], flush=True)
    assert r.returncode == 0, r.stderr
    assert "Kernel execution time" in r.stdout and "GPU kernel execution time" in r.stdout
