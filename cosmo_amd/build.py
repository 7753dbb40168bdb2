"""Builds libcosmo_hip.so (gfx950) and the command-line tools in-tree."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, force=True):
    csrc = os.path.join(HERE, "csrc")
    if force:
        subprocess.run(["make", "-s", "-C", csrc, "clean"], check=True)
    subprocess.run(["make", "-s", "-j%d" % jobs, "-C", csrc], check=True)
    cli = os.path.join(csrc, "cli")
    if os.path.isdir(cli):
        subprocess.run(["make", "-s", "-j%d" % jobs, "-C", cli], check=True)
    return os.path.join(HERE, "libcosmo_hip.so")


if __name__ == "__main__":
    print(build())
