"""Writes tests/golden/manual12_band.npz: a band of the user manual's section-12 workload
(img_0000.tif / img_0570.tif, 16-bit, ROI (309.5, 69.5)-(759.5, 1145.5), ws 20, step 10,
search radius 60), cropped so that 31 of its 107 subset rows (grid rows 371..681) and all
42 subset columns keep their positions and windows: image columns 235..822 and rows
300..775 (1-based).  Run in the container that holds the reference:
python tests/golden/make_manual12_band.py"""
import os
import sys

import numpy as np
from PIL import Image

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "manual12_band.npz")
X0, X1, Y0, Y1 = 235, 822, 300, 775   # 1-based, inclusive


def main():
    imgs = [np.asarray(Image.open(os.path.join(REF, n))) for n in ("img_0000.tif", "img_0570.tif")]
    crop = [np.ascontiguousarray(a[Y0 - 1:Y1, X0 - 1:X1].astype(np.uint16)) for a in imgs]
    # the full workload's grid [310, 760, 70, 1146] (first row 71 after the search's
    # trimming) in the crop, grid rows 371..681
    grid = np.array([310 - (X0 - 1), 760 - (X0 - 1), 371 - (Y0 - 1), 681 - (Y0 - 1)], np.int32)
    np.savez_compressed(OUT, f=crop[0], g=crop[1], grid=grid, offset=np.array([X0 - 1, Y0 - 1]),
                        ws=20, step=10, s=60)
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes), crop {crop[0].shape}, grid {grid}")


if __name__ == "__main__":
    sys.exit(main())
