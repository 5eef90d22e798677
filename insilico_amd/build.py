"""Build step: compiles the gfx950 HIP library in tree (make -C csrc)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, force=True):
    args = ["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    return os.path.join(HERE, "libinsilico_hip.so")


if __name__ == "__main__":
    print(build())
