"""Substrate generators (geometry/substrate.py of the reference)."""
