"""Backend registry, grid and sponge PML."""
