"""Learned pressure guess: a U-Net on the project's MFMA kernels, trained on
Jacobi-generated pairs, applied on a coarse grid; plain Jacobi finishes."""
