"""The MATLAB binding (2d_aldic_amd/matlab): the MEX gateway covers every export of
include/aldic.h, compiles (syntax only: MATLAB is not installed here, so the mx*/mex*
functions it calls are declared by a prototype-only header written for this check), and
the .m wrappers keep the reference functions' exact signatures."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MDIR = os.path.join(ROOT, "2d_aldic_amd", "matlab")
GATEWAY = os.path.join(MDIR, "aldic_mex.cpp")

# exports the gateway manages itself or that take device pointers (documented in its header)
NOT_COMMANDS = {"aldic_prep_noalloc"}

# the reference's first lines (func/*.m:1, func_quadtree/*.m:1), continuation included
SIGNATURES = {
    "IntegerSearch": "function [DICpara,x0,y0,u,v,cc]= IntegerSearch(ImgRef,ImgDef,file_name,DICpara)",
    "LocalICGN": "function [U,F,HtempPar,LocalTime,ConvItPerEle,LocalICGNBadPtNum] = LocalICGN(U0,coordinatesFEM,...",
    "Subpb1": "function [U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1(UOld,FOld,udual,vdual,coordinatesFEM,...",
    "Subpb2": "function [Uhat] = Subpb2(DICmesh,beta,mu,U,F,udual,vdual,alpha,GaussPtOrder)",
    "funGlobal_NodalStrainAvg": "function [StrainNodalPt] = funGlobal_NodalStrainAvg(coordinatesFEM,elementsFEM,U,GaussPtOrder)",
    "Subpb1Quadtree": "function [U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1Quadtree(UOld,FOld,udual,vdual,coordinatesFEM,...",
    "Subpb2Quadtree": "function [Uhat] = Subpb2Quadtree(DICmesh,GaussPtOrder,beta,mu,U,F,udual,vdual, ...",
    "funGlobalNodalStrainQuadtree": "function [StrainNodalPt,StrainGaussPt,CoordGaussPt] = funGlobalNodalStrainQuadtree(...",
}

MEX_DECLS = r"""
#pragma once
#include <cstddef>
#include <cstdint>
typedef size_t mwSize;
typedef bool mxLogical;
struct mxArray;
enum mxClassID { mxUNKNOWN_CLASS, mxCELL_CLASS, mxSTRUCT_CLASS, mxLOGICAL_CLASS, mxCHAR_CLASS,
  mxVOID_CLASS, mxDOUBLE_CLASS, mxSINGLE_CLASS, mxINT8_CLASS, mxUINT8_CLASS, mxINT16_CLASS,
  mxUINT16_CLASS, mxINT32_CLASS, mxUINT32_CLASS };
enum mxComplexity { mxREAL, mxCOMPLEX };
extern "C" {
void mexErrMsgIdAndTxt(const char *id, const char *fmt, ...);
int mexAtExit(void (*f)(void));
bool mxIsChar(const mxArray *);
bool mxIsDouble(const mxArray *);
bool mxIsComplex(const mxArray *);
bool mxIsLogical(const mxArray *);
int mxGetString(const mxArray *, char *, mwSize);
size_t mxGetM(const mxArray *);
size_t mxGetN(const mxArray *);
size_t mxGetNumberOfElements(const mxArray *);
mwSize mxGetNumberOfDimensions(const mxArray *);
const mwSize *mxGetDimensions(const mxArray *);
double *mxGetDoubles(const mxArray *);
void *mxGetData(const mxArray *);
mxLogical *mxGetLogicals(const mxArray *);
double mxGetScalar(const mxArray *);
mxClassID mxGetClassID(const mxArray *);
mxArray *mxCreateDoubleMatrix(mwSize, mwSize, mxComplexity);
mxArray *mxCreateNumericMatrix(mwSize, mwSize, mxClassID, mxComplexity);
mxArray *mxCreateNumericArray(mwSize, const mwSize *, mxClassID, mxComplexity);
mxArray *mxCreateDoubleScalar(double);
mxArray *mxDuplicateArray(const mxArray *);
}
"""


def _exports():
    txt = open(os.path.join(ROOT, "include", "aldic.h")).read()
    return set(re.findall(r"^\s*(?:int|void|const char \*)\s*(aldic_\w+)\s*\(", txt, re.M))


def test_gateway_covers_every_export():
    src = open(GATEWAY).read()
    called = set(re.findall(r"\b(aldic_\w+)\s*\(", src))
    missing = sorted(_exports() - called - NOT_COMMANDS)
    assert not missing, f"exports without a gateway command: {missing}"
    for name in NOT_COMMANDS:
        assert name in src  # named in the gateway's header as not reachable from MATLAB
    cmds = re.findall(r'cmd == "(\w+)"', src)
    assert len(cmds) == len(set(cmds)) and len(cmds) >= 25
    assert src.count("{") == src.count("}")


def test_gateway_compiles():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    import tempfile
    tmp = tempfile.mkdtemp(prefix="mexcheck_")
    try:
        with open(os.path.join(tmp, "mex.h"), "w") as fh:
            fh.write(MEX_DECLS)
        r = subprocess.run([gxx, "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                            "-I", os.path.join(ROOT, "include"), "-I", tmp, GATEWAY],
                           capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def test_wrappers_keep_reference_signatures():
    for name, sig in SIGNATURES.items():
        path = os.path.join(MDIR, name + ".m")
        assert os.path.exists(path), path
        first = open(path).read().split("\n", 1)[0].strip()
        assert first == sig, (name, first)
    for fn in os.listdir(MDIR):
        if fn.endswith(".m"):
            body = open(os.path.join(MDIR, fn)).read()
            used = set(re.findall(r"aldic_mex\('(\w+)'", body))
            cmds = set(re.findall(r'cmd == "(\w+)"', open(GATEWAY).read()))
            assert used <= cmds, (fn, used - cmds)
