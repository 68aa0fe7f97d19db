"""Device-resident environment cache (reference mps/lib.py:12-118).

Same contract as the reference ``Environ``: ``idx`` is the position of the
last site folded in, ``("L", -1)`` and ``("R", N)`` are the (1,1,1)
sentinels, ``GetLR`` with ``method="System"`` extends an environment by one
site and stores it.  Unlike the reference, environments never leave the
device (the reference copies every one to the host, lib.py:114-118).
Environment dtype is complex128 in the FP64 mode and complex64 in the mixed
mode; the contraction itself is ``native.env_update`` (contract_one_site).
"""
import torch

from renormalizer_amd import native


def contract_one_site(environ, ms, mo, domain, ms_conj=None, prec=native.PREC_FP64):
    """reference lib.py:169-250 for 3-leg sites, on the device."""
    return native.env_update(environ, ms, mo, domain, ms_conj=ms_conj, prec=prec)


class Environ:
    def __init__(self, mps, mpo_tensors, domain=None, mps_conj=None, prec=native.PREC_FP64,
                 build=True):
        self.prec = prec
        cdt, _ = native.prec_dtypes(prec)
        dev = mps[0].device
        self.sentinel = torch.ones((1, 1, 1), dtype=cdt, device=dev)
        self._virtual_disk = {}
        self.nsite = len(mps)
        self.write("L", -1, self.sentinel)
        self.write("R", self.nsite, self.sentinel)
        if build:
            self._construct(mps, mpo_tensors, domain, mps_conj)

    def _construct(self, mps, mpo_tensors, domain, mps_conj):
        assert domain in ("L", "R", None)
        if domain is None:
            self._construct(mps, mpo_tensors, "L", mps_conj)
            self._construct(mps, mpo_tensors, "R", mps_conj)
            return
        n = len(mps)
        rng = range(0, n - 1) if domain == "L" else range(n - 1, 0, -1)
        t = self.sentinel
        for idx in rng:
            t = contract_one_site(t, mps[idx], mpo_tensors[idx], domain,
                                  None if mps_conj is None else mps_conj[idx], self.prec)
            self.write(domain, idx, t)

    def GetLR(self, domain, siteidx, mps, mpo_tensors, itensor=None, method="Scratch", mps_conj=None):
        assert domain in ("L", "R")
        assert method in ("Enviro", "System", "Scratch")
        if siteidx not in range(len(mps)):
            return self.sentinel
        cj = (lambda i: None) if mps_conj is None else (lambda i: mps_conj[i])
        if method == "Scratch":
            itensor = self.sentinel
            sites = range(siteidx + 1) if domain == "L" else range(len(mps) - 1, siteidx - 1, -1)
            for i in sites:
                itensor = contract_one_site(itensor, mps[i], mpo_tensors[i], domain, cj(i), self.prec)
        elif method == "Enviro":
            itensor = self.read(domain, siteidx)
        else:
            if itensor is None:
                itensor = self.read(domain, siteidx + (-1 if domain == "L" else 1))
            itensor = contract_one_site(itensor, mps[siteidx], mpo_tensors[siteidx], domain,
                                        cj(siteidx), self.prec)
            self.write(domain, siteidx, itensor)
        return itensor

    def write(self, domain, siteidx, tensor):
        self._virtual_disk[(domain, siteidx)] = tensor

    def read(self, domain, siteidx):
        return self._virtual_disk[(domain, siteidx)]
