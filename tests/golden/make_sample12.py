"""Writes tests/golden/sample12.npz from the reference's stored quadtree result
(results_ForQuadtree_Sample12/results_oht_cfrp_11_ws32_st16.mat, read as data with
scipy.io.loadmat) and its image pair (Images_Quadtree_demo/Images_Sample12/
oht_cfrp_00.tiff, oht_cfrp_11.tiff): the images as row-major [y][x] uint8 arrays (the
transpose of ReadImage.m's x-first layout), the mask likewise, the quadtree DICmesh
(coordinates, 8-column elements, Dirichlet and hole-edge node lists, 1-based as stored)
and the solved ResultDisp.U and ResultDefGrad.F.  Run in the container that holds the
reference: python tests/golden/make_sample12.py"""
import os
import sys

import numpy as np
import scipy.io as sio
from PIL import Image

REF = "/root/reference"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sample12.npz")


def main():
    d = sio.loadmat(os.path.join(REF, "results_ForQuadtree_Sample12",
                                 "results_oht_cfrp_11_ws32_st16.mat"),
                    squeeze_me=True, struct_as_record=False)
    para, mesh = d["DICpara"], d["DICmesh"]
    imgs = []
    for name in ("oht_cfrp_00.tiff", "oht_cfrp_11.tiff"):
        im = Image.open(os.path.join(REF, "Images_Quadtree_demo", "Images_Sample12", name))
        a = np.asarray(im)
        if a.ndim == 3:
            a = a[..., 0]
        imgs.append(np.ascontiguousarray(a.astype(np.uint8)))
    ny, nx = imgs[0].shape
    assert tuple(para.ImgSize) == (nx, ny), (para.ImgSize, imgs[0].shape)
    mask = np.ascontiguousarray(np.asarray(para.ImgRefMask, np.uint8).T)
    np.savez_compressed(
        OUT, f=imgs[0], g=imgs[1], mask=mask,
        coordinatesFEM=np.asarray(mesh.coordinatesFEM, np.float64),
        elementsFEM=np.asarray(mesh.elementsFEM, np.int32),
        dirichlet=np.asarray(mesh.dirichlet, np.int32),
        markCoordHoleEdge=np.asarray(mesh.markCoordHoleEdge, np.int32),
        U=np.asarray(d["ResultDisp"].U, np.float64), F=np.asarray(d["ResultDefGrad"].F, np.float64),
        winsize=int(para.winsize), winstepsize=int(para.winstepsize),
        gridx=np.asarray(para.gridxyROIRange.gridx, np.int32),
        gridy=np.asarray(para.gridxyROIRange.gridy, np.int32))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes)")


if __name__ == "__main__":
    sys.exit(main())
