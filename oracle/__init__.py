"""Oracle: CPU restatement of Renormalizer's TDVP-PS / DMRG local update.

TEST INFRASTRUCTURE ONLY.  Nothing in the ``renormalizer_amd`` package may
import this package; only ``tests/``, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of ``bench.py`` use it, and only as the checker (or as
the timed CPU baseline), never as part of the product path.

Modules
-------
tdvp    numpy complex128 restatement of the reference's hot path
        (contract_one_site, hop_expr, expm_krylov, svd_qn, the one-site
        TDVP-PS sweep, normalisation and observables), each function citing
        the reference file:line it follows.
dense   independent dense (Kronecker-product) Hamiltonians and exact
        propagation for small spin-boson / Holstein models; used to pin the
        MPO construction and the TDVP-PS sweep without the reference.

Pinning status of each function is listed in its docstring ("pinned by: ...").
"""
