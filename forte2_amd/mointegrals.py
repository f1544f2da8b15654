"""MO-basis integrals for correlated methods on top of the device builder:
the drop-in for forte2's RestrictedMOIntegrals
(forte2/jkbuilder/mointegrals.py:15-86).

E = Vnn (+ the energy of the doubly occupied core), H = C^T (h + 2 J_core -
K_core) C over the active orbitals, V = <pq|rs> over the active orbitals
(physicist's convention, optionally antisymmetrized).  J/K of the core come
from the builder's build_JK, V from its two_electron_integrals_block, so all
heavy work runs in the device kernels.  numpy in, numpy out; torch tensors
on the device stay there.
"""

import dataclasses

import numpy as np
import torch


@dataclasses.dataclass
class RestrictedMOIntegrals:
    system: object
    C: object
    orbitals: list
    core_orbitals: list = dataclasses.field(default_factory=list)
    antisymmetrize: bool = False

    def __post_init__(self):
        fb = self.system.fock_builder
        to_np = not isinstance(self.C, torch.Tensor)
        dev = fb.device
        Call = torch.as_tensor(np.asarray(self.C, dtype=np.float64) if to_np else self.C,
                               dtype=torch.float64).to(dev)
        orbs = list(self.orbitals)
        core = list(self.core_orbitals)
        self.norb = len(orbs)
        C = Call[:, orbs].contiguous()
        H_ao = torch.as_tensor(self.system.ints_hcore(), dtype=torch.float64).to(dev)
        E = float(self.system.nuclear_repulsion)
        H = C.T @ H_ao @ C
        if core:
            Cc = Call[:, core].contiguous()
            J, K = fb.build_JK([Cc])
            G = 2.0 * J[0] - K[0]
            E += 2.0 * float(torch.sum(Cc * (H_ao @ Cc)))
            E += float(torch.sum(Cc * (G @ Cc)))
            H = H + C.T @ G @ C
        V = fb.two_electron_integrals_block(C)
        if self.antisymmetrize:
            V = V - V.transpose(2, 3)
        self.E = E
        self.H = H.cpu().numpy() if to_np else H
        self.V = V.contiguous().cpu().numpy() if to_np else V.contiguous()
