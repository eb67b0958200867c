"""CPU oracle for the metriko hot path: TEST INFRASTRUCTURE ONLY.

This package is a plain FP64 restatement of komietty/metriko's numerical core
(halfedge mesh, Knoeppel smoothest 4-RoSy field, seam/cut/comb, IGM setup,
Poisson KKT, Gauss-Newton/LM passes and iterative rounding, the local-injectivity
barrier LM, and the QEx / QGP consumers of the uv) written with numpy/scipy.  It is the checker for the HIP implementation in ``metriko_amd``:

* only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
  ``bench.py`` may import it, and only to check or to time a CPU baseline;
* the product path (``metriko_amd``, ``include/``) never imports, links or
  executes anything from here.

Every function cites the reference file:line it follows (paths relative to
the reference's ``include/metriko/core``).  Where the reference delegates to a
third-party library that is absent here (libigl v2.5.0, Eigen 3.4), the
library's published algorithm is restated and named in the docstring.

Parity status of each function (what pins it) is listed in ``PINNING`` below
and checked by ``tests/test_oracle_*.py``.
"""

PINNING = {
    # function                      : how it is pinned
    "edge_topology": "internal consistency (twin/next/prev, EF orientation); "
                     "libigl ids themselves: parity unpinned (libigl absent)",
    "Hmesh.geometry": "Gauss-Bonnet, exterior angle sums, brute-force numpy recompute",
    "connection_laplacian": "Hermitian, |offdiag|=1, planar closed form, dense eigh",
    "solve_smallest_eig": "residual vs Rayleigh quotient, scipy eigsh(sigma=0), "
                          "dense power iteration on tiny meshes",
    "field_roots": "closed form",
    "matching/singular": "Poincare-Hopf (sum = N*chi), rotation invariance",
    "compute_seam": "cut-graph invariants (disk topology, singulars on seam); hand-worked "
                    "octahedron cuts; MeshCutter == 2-core of the networkx BFS dual co-tree; "
                    "Dijkstra == smallest (distance, id) target with smallest-id predecessors "
                    "(tests/test_seam_pins.py)",
    "cut/comb/setup": "transition consistency, constraint rows vanish on x0",
    "reduce_to_linearly_independent": "COLAMD (oracle/colamd.py) == SuperLU's COLAMD through "
                                      "scipy splu's column permutation on random matrices, dense "
                                      "rows under SuperLU's threshold rule, hand-worked orderings; "
                                      "SparseQR rank decisions == brute-force rank growth and the "
                                      "|beta| threshold; Eigen's 0.5 dense fractions stated from "
                                      "Eigen 3.4's Eigen_Colamd.h (tests/test_colamd.py)",
    "fullpiv_kernel": "pivot sequence and factors == LAPACK dgetc2 (complete pivoting) on "
                      "tie-free matrices; hand-worked kernels with forced ties; reduced basis "
                      "(identity on the free columns, direct solve on the pivot columns); "
                      "C @ K == 0, rank (tests/test_fullpiv.py)",
    "kkt": "scipy spsolve on the assembled KKT; dense lstsq on tiny meshes",
    "lm/rounding": "integrality, seamlessness, literal dense normal-equation LM on tiny meshes",
    "injective (barrier, damped LM)": "barrier closed form and finite differences, Rosenbrock "
                                      "minimiser, Woodbury vs direct solve, equal to the plain "
                                      "rounding where the barrier stays inactive",
    "qex": "torus known answers (s^2 n m quads), Euler characteristic 2 on spheres, all ports "
           "connected, transition closed form on the torus seams",
    "motorcycle/tmesh": "one crash and one split per curve, T-quads going round four sides, "
                        "Euler characteristic 2",
    "quantization": "opposite T-quad sides equal, validation, weight closed form, heap order "
                    "vs the comparator; generator rows by the SparseQR<COLAMD> rule above; "
                    "std::sort ties: parity unpinned",
    "poisson/sizing": "see tests/test_batch_oracle.py and tests/test_sizing_oracle.py",
}
