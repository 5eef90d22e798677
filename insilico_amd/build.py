"""Build step: compiles the gfx950 HIP library in tree (make -C csrc) and the
example C++ drivers linked against it (make -C apps)."""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))


def build(jobs=8, force=True):
    args = ["make", "-C", os.path.join(HERE, "csrc"), "-j%d" % jobs]
    if force:
        args.append("-B")
    subprocess.run(args, check=True)
    apps = ["make", "-s", "-C", os.path.join(HERE, "apps")]
    if force:
        apps.append("-B")
    subprocess.run(apps, check=True)
    return os.path.join(HERE, "libinsilico_hip.so")


if __name__ == "__main__":
    print(build())
