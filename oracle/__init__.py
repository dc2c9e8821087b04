"""CPU oracle of the CheMPS2 two-site DMRG hot path (test infrastructure only)."""
