"""CPU oracle of the voxel pipeline: test infrastructure only (see oracle.py)."""
