"""MI355X-native 2D nonlocal finite-element hot path (neighbours, assembly, SpMV, CG)."""
