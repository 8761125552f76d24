"""CPU oracle for the MLS-MPM growth step and its adjoint.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package (``morphompm_amd``)
imports, links or runs anything from this folder; only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` use it,
and only as the checker or the timed host baseline.

Modules
-------
``ref_loop``  FP64 numpy restatement that keeps the reference's per-particle x
              27-node loop structure (python/morphompm/{config,state,
              constitutive,transfer,integrate,diff}.py).  Slow; small cases only.
``ref_vec``   FP64 numpy restatement vectorised over particles, with the
              extensions the workloads need (non-cubic grids, growth schedules,
              Herschel-Bulkley parameter cotangents).  Gated against
              ``ref_loop`` and used at larger sizes and as the CPU baseline.
``scenes``    Seeded synthetic particle sets (free swell, bilayer, batches).
``_ref/``     (git-ignored) CPU build of the reference's own C++ forward
              solver, made by ``oracle/Makefile`` from /root/reference.

Pinning (what checks each function, beside the restatement itself) is listed
in the module docstrings and exercised in ``tests/test_oracle_*.py``.
"""
