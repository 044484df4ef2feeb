"""The host code under AddressSanitizer and UBSan (CPU only): the front end's
parsers on the benchmark inputs and the makeLinear topology input (--dry-run,
no device), and the oracle's core, TILD and FTS pins with the sanitized
oracle build loaded into a Python with libasan preloaded."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MFT_ASAN = os.path.join(ROOT, "matilda.ft_amd", "bin", "mft_asan")
ORC_ASAN = os.path.join(ROOT, "oracle", "liboracle64_asan.so")
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
BAD = ("ERROR: AddressSanitizer", "runtime error:")


@pytest.mark.skipif(not os.path.exists(MFT_ASAN), reason="sanitized front end not built")
@pytest.mark.parametrize("name", ["c1_diblock.in", "c2_coacervate.in", "c3_fts_diblock.in", "makelinear"])
def test_front_end_parsers_are_clean_under_asan_and_ubsan(tmp_path, name):
    if name == "makelinear":
        from test_frontend import MAKELINEAR_INPUT
        path = tmp_path / "input"
        path.write_text(MAKELINEAR_INPUT)
        args = ["-in", str(path), "--dry-run", "--dump-topology", str(tmp_path / "topo.txt")]
    else:
        args = ["-in", os.path.join(ROOT, "tests", "inputs", name), "--dry-run"]
    r = subprocess.run([MFT_ASAN] + args, cwd=tmp_path, capture_output=True, text=True, timeout=300, env=ENV)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert not any(b in out for b in BAD), out[-3000:]


def _libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    p = r.stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(not os.path.exists(ORC_ASAN) or _libasan() is None, reason="sanitized oracle not built")
def test_oracle_pins_are_clean_under_asan_and_ubsan():
    env = dict(ENV, ORC_SANITIZE="_asan", LD_PRELOAD=_libasan())
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_core.py"),
                        os.path.join(ROOT, "tests", "test_oracle_tild.py"),
                        os.path.join(ROOT, "tests", "test_oracle_fts.py"), "-k",
                        "not gjf_harmonic and not langevin and not dpd"],
                       cwd=ROOT, capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-3000:]
    assert not any(b in out for b in BAD), out[-3000:]
