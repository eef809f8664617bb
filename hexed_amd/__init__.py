"""MI355X (gfx950) implementation of hexed's DG spatial discretization and explicit time step."""
