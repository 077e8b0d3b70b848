"""Regenerate philox_rocrand.json from rocRAND's host-callable Philox
(compiles philox_rocrand.cpp for the host with hipcc and runs it)."""
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))

with tempfile.TemporaryDirectory() as d:
    exe = os.path.join(d, "philox_rocrand")
    subprocess.run(["/opt/rocm/bin/hipcc", "-O1", "-o", exe, os.path.join(HERE, "philox_rocrand.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
with open(os.path.join(HERE, "philox_rocrand.json"), "w") as f:
    f.write(out)
print(out.count("ctr"), "vectors written")
