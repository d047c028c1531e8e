"""The parser of the integer settings read from the environment when a domain
or batch is made (AMCA_DOMAIN_GRAPH_STEPS, AMCA_REMELT_GRAPH_STEPS,
AMCA_BATCH_GRAPHS): unset, valid, and every malformed or out-of-range form."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("env") / "env_check")
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", exe,
                    os.path.join(ROOT, "tests", "env_check.cpp")], check=True)
    return exe


@pytest.mark.parametrize("value,code,parsed", [
    (None, 0, -12345),   # unset: nothing stored
    ("0", 1, 0),
    ("16", 1, 16),
    ("1024", 1, 1024),
    ("1025", -1, -12345),
    ("-1", -1, -12345),
    ("", -1, -12345),
    ("sixteen", -1, -12345),
    ("16x", -1, -12345),
    (" 16", 1, 16),      # strtol skips leading blanks
])
def test_env_int(checker, value, code, parsed):
    env = {k: v for k, v in os.environ.items() if k != "AMCA_TEST_SETTING"}
    if value is not None:
        env["AMCA_TEST_SETTING"] = value
    out = subprocess.run([checker, "AMCA_TEST_SETTING"], env=env, capture_output=True, text=True,
                         check=True).stdout.split()
    assert [int(out[0]), int(out[1])] == [code, parsed]
