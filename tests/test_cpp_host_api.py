"""The C++ host API (include/effsc/effsc.hpp) through its test program."""
import json
import os
import subprocess

import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_api_test")


def run(mode):
    if not os.path.exists(BIN):
        pytest.fail("tests/cpp/host_api_test is not built (make -C tests/cpp)")
    r = subprocess.run([BIN, mode], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return [json.loads(line) for line in r.stdout.splitlines() if line.startswith("{")]


def test_cpp_transcript_and_verifier_errors():
    out = {d["name"]: d for d in run("cpu")}
    t = O.Transcript.shake(b"effsc-hip/cpp")
    t.send([1])
    t.send([2])
    assert out["transcript"]["challenges"] == [t.challenge(1)[0], t.challenge(1)[0]]


@pytest.mark.gpu
def test_cpp_gpu_provers_match_oracle():
    out = {d["name"]: d for d in run("gpu")}
    v, n = 10, 1 << 10
    import brute as B
    a = [B.uniform(0xC0FFEE01, i) for i in range(n)]
    b = [B.uniform(0xC0FFEE02, i) for i in range(n)]
    rp, ch, fin = O.Prover.inner_product(a, b).sumcheck(v, O.Transcript.shake(b"effsc-hip/cpp"))
    assert out["ip"]["round_polys"] == [[x[0] for x in r] for r in rp]
    assert out["ip"]["challenges"] == [c[0] for c in ch]
    assert out["ip"]["final"] == fin[0]

    # the LSB inner product printed by the program, against the oracle's LSB restatement
    rl, cl, fl = O.Prover.inner_product_lsb(a, b).sumcheck(v, O.Transcript.shake(b"effsc-hip/cpp"))
    assert out["ip_lsb"]["round_polys"] == [[x[0] for x in r] for r in rl]
    assert out["ip_lsb"]["final"] == fl[0]
