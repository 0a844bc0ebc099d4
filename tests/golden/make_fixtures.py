"""Writes the golden fixtures under tests/golden/ from data the SLEPc source
tree holds (expected test outputs and a PETSc-binary matrix file), plus
independent library results (scipy ARPACK/LAPACK) for the workloads.

Run in a container that has the reference tree mounted:
    python tests/golden/make_fixtures.py /root/reference
The fixtures are committed; tests never read the reference tree.
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def floats(text):
    return [float(x) for x in re.findall(r"-?\d+\.\d*(?:e[-+]\d+)?|-?\d+\.", text)]


def parse_mat_block(lines, start, nrows):
    rows = []
    for ln in lines[start:start + nrows]:
        rows.append([float(x) for x in ln.split()])
    return rows


def bv_test1(ref):
    """src/sys/classes/bv/tests/output/test1_1_bv_type-mat.out (test1.c:15-181)."""
    path = os.path.join(ref, "src/sys/classes/bv/tests/output/test1_1_bv_type-mat.out")
    lines = open(path).read().splitlines()
    out = {}

    def block_after(marker, title, nrows):
        i = next(k for k, ln in enumerate(lines) if marker in ln)
        j = next(k for k in range(i, len(lines)) if title in lines[k])
        # skip "type:" line
        j += 2 if "Mat Object" in lines[j] else 1
        while not re.match(r"^-?\d", lines[j].strip()):
            j += 1
        return parse_mat_block(lines, j, nrows)

    out["X0"] = block_after("BV Object:X", "Mat_X", 10)
    out["Y0"] = block_after("BV Object:Y", "Mat_Y", 10)
    out["Q"] = block_after("Mat Object: Q", "seqdense", 5)
    out["Y_after_mult"] = block_after("After BVMult -", "Mat_Y", 10)
    out["Y_after_multvec"] = block_after("After BVMultVec", "Mat_Y", 10)
    out["M_after_dot"] = block_after("After BVDot -", "seqdense", 3)
    i = next(k for k, ln in enumerate(lines) if "After BVDotVec" in ln)
    out["z_after_dotvec"] = [float(lines[i + 3 + q]) for q in range(5)]
    out["X_after_multinplace_scale"] = block_after("After BVMultInPlace", "Mat_X", 10)
    txt = "\n".join(lines)
    out["norm2_X0"] = float(re.search(r"2-Norm of X\[0\] = ([\d.]+)", txt).group(1))
    out["normF_X"] = float(re.search(r"Frobenius Norm of X = ([\d.]+)", txt).group(1))
    return out


def bv_test2(ref):
    """src/sys/classes/bv/tests/output/test2_1.out: ||ones - X X^T ones|| with X orthonormal."""
    txt = open(os.path.join(ref, "src/sys/classes/bv/tests/output/test2_1.out")).read()
    return {"norm_ones_after": float(re.search(r"X: ([\d.]+)", txt).group(1)),
            "n": 20, "k": 8}


def ds_test3(ref):
    """src/sys/classes/ds/tests/output/test3_1.out: DSHEP compact, n=9 l=2 k=5."""
    txt = open(os.path.join(ref, "src/sys/classes/ds/tests/output/test3_1.out")).read()
    part = txt.split("Computed eigenvalues =")[1]
    return {"n": 9, "l": 2, "k": 5, "eigs": floats(part)[:9]}


def terse_values(ref, rel):
    txt = open(os.path.join(ref, rel)).read()
    line = txt.split("required tolerance:")[1].strip().splitlines()[0]
    return [float(x) for x in line.replace(",", " ").split()]


def petsc_binary(path):
    """PETSc binary AIJ: big-endian int32 header (classid 1211216, M, N, nz),
    row lengths, column indices, then big-endian float64 values."""
    b = open(path, "rb").read()
    hdr = np.frombuffer(b[:16], ">i4")
    assert hdr[0] == 1211216
    M, N, nz = int(hdr[1]), int(hdr[2]), int(hdr[3])
    off = 16
    rl = np.frombuffer(b[off:off + 4 * M], ">i4").astype(np.int32)
    off += 4 * M
    ci = np.frombuffer(b[off:off + 4 * nz], ">i4").astype(np.int32)
    off += 4 * nz
    v = np.frombuffer(b[off:off + 8 * nz], ">f8").astype(np.float64)
    rp = np.concatenate([[0], np.cumsum(rl)]).astype(np.int32)
    return M, rp, ci, v


def main(ref):
    g = {
        "bv_test1": bv_test1(ref),
        "bv_test2": bv_test2(ref),
        "ds_test3": ds_test3(ref),
        "ex2_1": {"n": 72, "nev": 4, "ncv": 20, "values": terse_values(ref, "src/eps/tutorials/output/ex2_1.out")},
        "ex5_1": {"m": 15, "nev": 4, "values": terse_values(ref, "src/eps/tutorials/output/ex5_1.out")},
        "ex4_1": {"nev": 4, "its": 5, "values": terse_values(ref, "src/eps/tutorials/output/ex4_1.out")},
        "test27_1": {"n": 10, "nev": 5, "values": terse_values(ref, "src/eps/tests/output/test27_1.out")},
    }
    with open(os.path.join(HERE, "reference_outputs.json"), "w") as f:
        json.dump(g, f, indent=1)
    M, rp, ci, v = petsc_binary(os.path.join(ref, "share/slepc/datafiles/matrices/rdb200.petsc"))
    np.savez(os.path.join(HERE, "rdb200.npz"), n=M, rowptr=rp, col=ci, val=v)

    # Independent library results (scipy): the 16 smallest eigenvalues of the
    # finite-volume operator L = sigma I - M at 48x48 for two parameter draws.
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    from oracle import oracle as O
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    fv = []
    for sample in (0, 1):
        th = O.fv_theta(5, sample)
        A = O.fv2d(48, 48, th)
        S = sp.csr_matrix((A.val, A.col, A.rowptr), shape=(A.n, A.n))
        w = np.linalg.eigvalsh(S.toarray())
        fv.append({"nx": 48, "ny": 48, "theta": th.tolist(), "sigma": A.sigma,
                   "largest16": sorted(w.tolist())[::-1][:16]})
    # Convection-diffusion at 12^3: nonsymmetric, real spectrum; dense LAPACK eigvals.
    C = O.convdiff3d(12)
    S = sp.csr_matrix((C.val, C.col, C.rowptr), shape=(C.n, C.n)).toarray()
    w = np.linalg.eigvals(S)
    lib = {"fv2d": fv, "convdiff12": {"largest_real20": sorted(w.real.tolist())[::-1][:20],
                                      "max_abs_imag": float(np.abs(w.imag).max())}}
    # 27-point stencil at 10^3 with the U[0,1) diagonal: symmetric, dense LAPACK
    T = O.stencil27(10, seed=1)
    S = sp.csr_matrix((T.val, T.col, T.rowptr), shape=(T.n, T.n)).toarray()
    w = np.linalg.eigvalsh(S)
    lib["stencil27_10"] = {"largest_mag32": sorted(w.tolist(), key=abs)[::-1][:32]}
    with open(os.path.join(HERE, "library_results.json"), "w") as f:
        json.dump(lib, f, indent=1)
    print("fixtures written")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
