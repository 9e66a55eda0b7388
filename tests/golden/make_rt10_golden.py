"""Writes rt10_cornell_900x650.ppm.zlib: the image file the reference's
patch-projection program (experiments/rt10.cpp, built unchanged from its
source by oracle/ref/Makefile into oracle/_ref/rt10) writes for its own
scene, zlib-compressed (1,755,015 bytes of P6 -> about 4 KB).  The tests
compare the GPU renderer's image, written with the reference's 8-bit rule,
with it byte for byte.  Run in a container that has /root/reference."""
import hashlib
import json
import os
import subprocess
import tempfile
import zlib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle", "ref")])
    exe = os.path.join(ROOT, "oracle", "_ref", "rt10")
    with tempfile.TemporaryDirectory() as d:
        subprocess.check_call([exe], cwd=d)
        data = open(os.path.join(d, "output_rt10.ppm"), "rb").read()
    open(os.path.join(HERE, "rt10_cornell_900x650.ppm.zlib"), "wb").write(zlib.compress(data, 9))
    meta = {"source": "experiments/rt10.cpp main() (scene rt10.cpp:837-922), built with g++ -O2 -std=c++17",
            "bytes": len(data), "sha256": hashlib.sha256(data).hexdigest()}
    json.dump(meta, open(os.path.join(HERE, "rt10_cornell_900x650.json"), "w"), indent=1)
    print(meta)


if __name__ == "__main__":
    main()
