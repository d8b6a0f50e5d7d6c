"""The host-side C++ (the XML reader and Param of the C++ host) and the CPU
oracle built with AddressSanitizer and UndefinedBehaviorSanitizer (g++ only,
no GPU code) and run by tests/sanitize_driver.cc: a valid parameter file,
malformed ones (truncated, an unbalanced tag, an attribute without a value,
an unsupported model, an empty file, an unterminated comment) that must be
refused with std::runtime_error, LOS passes with random grids, a mask and
an imported CRE grid, and the generators on power-of-two, mixed-radix and
prime grids.  Any sanitizer report fails the run."""
import os
import subprocess

import pytest

from hammurabix_amd import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_and_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_driver"
    r = subprocess.run(
        ["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-fopenmp",
         "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
         "-I", os.path.join(ROOT, "hammurabix_amd", "host"),
         "-I", os.path.join(ROOT, "hammurabix_amd", "csrc"), "-I", os.path.join(ROOT, "oracle"),
         os.path.join(ROOT, "tests", "sanitize_driver.cc"),
         os.path.join(ROOT, "hammurabix_amd", "host", "hamx_param.cc"),
         os.path.join(ROOT, "hammurabix_amd", "host", "hamx_xml.cc"),
         os.path.join(ROOT, "oracle", "hamx_oracle.cc"),
         os.path.join(ROOT, "oracle", "hamx_oracle_rnd.cc"), "-o", str(exe)],
        capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower():
        pytest.skip("the compiler has no AddressSanitizer runtime")
    assert r.returncode == 0, r.stderr
    valid = tmp_path / "valid.xml"
    valid.write_text(configs.random_forward(freqs=[23.0, 1.4]))
    t = configs.plain_forward()
    bad = {
        "truncated": t[:len(t) // 2],
        "unbalanced": t.replace("</observable>", ""),
        "attribute": t.replace('<b0 value="6.0"/>', '<b0 value="6.0"'),
        "model": t.replace('type="lsa"', 'type="bogus"'),
        "empty": "",
        "comment": t.replace("<shell>", "<shell><!-- unterminated comment"),
    }
    paths = []
    for name, text in bad.items():
        p = tmp_path / f"{name}.xml"
        p.write_text(text)
        paths.append(str(p))
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(exe), str(valid), *paths], capture_output=True, text=True,
                       timeout=600, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr
    assert r.stdout.count("malformed") == len(bad)
    assert "sanitized run ok" in r.stdout
