"""Checkpoint container (particlebase_amd/cpp/include/Serialization.hh): the
self-describing binary that carries the reference's archive field set
(writeSerialization / readSerialization, CUDAenv.hh:268-285).  A host
round trip of every element type, and the errors for a wrong type, a
missing field, a foreign file and a truncated file.  Host only (g++)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def test_archive_round_trip(tmp_path):
    exe = tmp_path / "archive_test"
    subprocess.run(["g++", "-O1", "-std=c++17", "-o", str(exe),
                    os.path.join(HERE, "cpp", "archive_test.cpp")], check=True)
    p = subprocess.run([str(exe), str(tmp_path / "a.ck")], capture_output=True, text=True)
    assert p.returncode == 0, p.stdout + p.stderr
    assert p.stdout.strip().endswith("OK")
