"""Input formats of the reference's examples.

The sphere->cube examples (Cpp/examples/sphere2cube/run_sphere2cube.m:1-7
and the CUDA/ and MATLAB/ copies) load two MATLAB files: `sphereData.mat`
with `iniNdeMat` (D x N double) and `elmVtxMat` (V x E uint16, 1-based), and
`cubeData.mat` with the target varifold `tgtCenPosMat` (D x Et),
`tgtUniDirMat` (D x Et), `tgtElmVolVec` (1 x Et) and the target mesh
`tgtNdeMat`, `tgtElmVtxMat`.  `load_example` reads such files (scipy's MAT
reader: plain arrays, nothing in the file is executed) into this package's
problem dictionaries: (N, D) point arrays, 0-based int32 (E, V) tables.
"""
import numpy as np

__all__ = ["load_mat", "problem_from_arrays", "load_example"]


def load_mat(path):
    """Variables of a MAT file as numpy arrays (MATLAB metadata dropped)."""
    import scipy.io as sio
    d = sio.loadmat(path)
    return {k: np.asarray(v) for k, v in d.items() if not k.startswith("__")}


def problem_from_arrays(arrs):
    """Problem dictionary from the example variables (MATLAB layouts)."""
    q0 = np.ascontiguousarray(arrs["iniNdeMat"].T, dtype=np.float64)
    elm = np.ascontiguousarray(arrs["elmVtxMat"].T.astype(np.int64) - 1, dtype=np.int32)
    if elm.min() < 0 or elm.max() >= len(q0):
        raise ValueError("elmVtxMat must be 1-based and index iniNdeMat's columns")
    prob = dict(q0=q0, elm=elm,
                tgt_cen=np.ascontiguousarray(arrs["tgtCenPosMat"].T, dtype=np.float64),
                tgt_dir=np.ascontiguousarray(arrs["tgtUniDirMat"].T, dtype=np.float64),
                tgt_vol=np.ascontiguousarray(np.ravel(arrs["tgtElmVolVec"]), dtype=np.float64))
    if "tgtNdeMat" in arrs and "tgtElmVtxMat" in arrs:
        prob["tgt_nde"] = np.ascontiguousarray(arrs["tgtNdeMat"].T, dtype=np.float64)
        prob["tgt_elm"] = np.ascontiguousarray(arrs["tgtElmVtxMat"].T.astype(np.int64) - 1, dtype=np.int32)
    return prob


def load_example(source_mat, target_mat):
    """Problem dictionary from a source and a target MAT file."""
    arrs = load_mat(source_mat)
    arrs.update(load_mat(target_mat))
    return problem_from_arrays(arrs)
