"""CPU oracle for the assembly + equilibrated-flux estimator + MINRES hot path.

TEST INFRASTRUCTURE ONLY.  This package is a plain FP64 numpy/scipy restatement
of the reference algorithms (Rsmeets99/DoubleAdapParabolicFEM).  Only the tests,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker.  The product path
(``doubleadapparabolicfem_amd``) never imports it and fails loudly when its HIP
library is missing.

Each function cites the reference file:line it restates.  How each part is
pinned (closed forms, published anchors, invariants) is recorded in the module
docstrings and in the tests that exercise it:

* quadrature        -- Gauss-Legendre exactness, Duffy exactness
* basis             -- Kronecker delta at nodes, partition of unity
* mesh / spaces     -- published X/Y DOF counts
* assembly          -- Kronecker closed forms of the element matrices, symmetry
* estimator         -- published eta^2, lambda_Y^2, y_estimator^2 anchors
                       (6 significant digits) for 1+1D and 2+1D uniform runs
* minres            -- true residual + agreement with a sparse direct solve
"""
