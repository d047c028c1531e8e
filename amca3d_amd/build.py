"""In-tree build of libamca3d.so and the amca3d driver (hipcc for gfx950, g++ host)."""
import os
import subprocess

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG_DIR, "csrc")


def build(jobs=8, verbose=False):
    cmd = ["make", "-C", CSRC, f"-j{jobs}"]
    if not verbose:
        cmd.insert(1, "-s")
    env = dict(os.environ)
    env.setdefault("ROCM", "/opt/rocm")
    r = subprocess.run(cmd, env=env, capture_output=not verbose, text=True)
    if r.returncode != 0:
        raise RuntimeError("libamca3d build failed:\n" + (r.stdout or "") + (r.stderr or ""))


if __name__ == "__main__":
    build(verbose=True)
