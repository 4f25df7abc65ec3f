"""Bindings of the HIP kernels (libbh_hip.so)."""
