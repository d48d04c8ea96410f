"""Write the golden fixtures used by the parity tests.

Reads image files that the reference repository ships as test data
(data/blobs.tif, data/mini_blobs*.tif, data/mini_blobs_measurements_clij.csv and
demo/segmentation/BBBC022/IXMtest_A02_s9.tif) and stores them as small .npz
arrays next to this script, so that the tests never need the reference tree at
run time.  The expected arrays of the reference's own unit tests are transcribed
by hand into reference_tests.json (data, not code).

Run once in a container that has the reference checkout:
    python tests/golden/make_fixtures.py /root/reference
"""
import sys
from pathlib import Path

import numpy as np
from PIL import Image

HERE = Path(__file__).resolve().parent


def read_tif(path: Path, frame: int = 0) -> np.ndarray:
    im = Image.open(path)
    im.seek(frame)
    a = np.array(im)
    return a.astype(a.dtype.newbyteorder("=")) if a.dtype.byteorder not in ("=", "|") else a


def main(ref_root: str) -> None:
    ref = Path(ref_root)
    blobs = read_tif(ref / "data" / "blobs.tif")
    mini = read_tif(ref / "data" / "mini_blobs.tif")
    mini_labels = read_tif(ref / "data" / "mini_blobs_otsu_labels_excluded_edges_imagej.tif").astype(np.uint16)
    csv = np.loadtxt(ref / "data" / "mini_blobs_measurements_clij.csv", delimiter=",", skiprows=1)
    bbbc = read_tif(ref / "demo" / "segmentation" / "BBBC022" / "IXMtest_A02_s9.tif", 0).astype(np.uint16)
    np.savez_compressed(HERE / "blobs.npz", image=blobs)
    np.savez_compressed(HERE / "mini_blobs.npz", image=mini, labels=mini_labels, clij_table=csv[:, 1:])
    # the demo notebook works on frame 0, crop [0:200, 200:400]
    np.savez_compressed(HERE / "bbbc022_crop.npz", image=bbbc[0:200, 200:400])
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
