"""shci_amd: the SHCI variational hot path (Hamiltonian build, Davidson H.v,
mixed precision, RDM adjoint) for MI355X (gfx950), behind a C-ABI.

Submodules:
  hamiltonian -- Hamiltonian / SparseMatrix / Davidson mirror of the reference surface
  host        -- CPU integrals + heat-bath selection (libshci_host.so)
  integrals   -- synthetic benchmark integrals, FCIDUMP writer, packing helpers
  autograd    -- torch.autograd.Function E(h, v) with the RDM backward
"""
__all__ = ["hamiltonian", "host", "integrals", "autograd"]
