"""Write the committed golden fixtures for the linear n=20 workload.

The Femeko reference cannot run in this environment (no Julia runtime, empty
Eigen submodule), so the vectors come from the project's FP64 oracle
(oracle/) and scipy, cross-checked against each other here:
  - CSC pattern (1-based) and values of A(mu) from the oracle's restatement of
    sparse(rowIDs, colIDs, Ak .* mu); checked against scipy's COO->CSC
  - phi from scipy's direct solver (SuperLU) on A[free,free], checked against
    the oracle's Jacobi-PCG at rtol 1e-14
  - |H| per element and the energy from the oracle field recovery
Run: python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as sla

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "femeko.jl_amd")]

from femeko_hip import workloads as W  # noqa: E402
from oracle import oracle as O  # noqa: E402


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    pr = W.linear(20)
    m = pr.mesh
    Ak = O.localStiffnessMatrix(m)
    A = O.sparse_from_local(m, Ak, pr.mu)
    # scipy cross-check of pattern and values
    I = np.concatenate([m.t[n // 4] - 1 for n in range(16)])
    J = np.concatenate([m.t[n % 4] - 1 for n in range(16)])
    V = np.concatenate([Ak[n] * pr.mu for n in range(16)])
    S = sp.coo_matrix((V, (I, J)), shape=(m.nv, m.nv)).tocsc()
    S.sort_indices()
    assert np.array_equal(S.indptr + 1, A.colptr) and np.array_equal(S.indices + 1, A.rowval)
    assert np.max(np.abs(S.data - A.nzval)) <= 1e-14 * np.abs(S.data).max()
    # direct solve
    Sr = S.tocsr()
    free = pr.mask == 0
    b = -Sr[free][:, ~free] @ pr.g[~free]
    phi = pr.g.copy()
    phi[free] = sla.spsolve(Sr[free][:, free].tocsc(), b)
    phi_cg, it, rel, s = O.pcg_dirichlet(A, pr.mask, pr.g, 1e-14)
    assert s == 0 and np.linalg.norm(phi_cg - phi) / np.linalg.norm(phi) < 1e-12
    _it10 = O.pcg_dirichlet(A, pr.mask, pr.g, 1e-10)[1]
    H, Hn = O.field(m, phi)
    E = O.energy(m, Hn, pr.mu)
    np.savez_compressed(
        os.path.join(HERE, "linear20.npz"),
        sha_p=sha(m.p), sha_t=sha(m.t), sha_mu=sha(pr.mu), sha_mask=sha(pr.mask), sha_g=sha(pr.g),
        sha_colptr=sha(A.colptr), sha_rowval=sha(A.rowval), nnz=A.nnz,
        nzval=A.nzval, phi=phi, Hn=Hn, H_sample=H[:, ::997], energy=E, pcg_iters_1e10=_it10,
    )
    print("wrote linear20.npz: nnz", A.nnz, "pcg iterations at 1e-10:", _it10, "energy", E)


if __name__ == "__main__":
    main()
